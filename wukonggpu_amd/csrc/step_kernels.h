// Pattern-step kernels: fused probe + scan (+ filter compaction), merge-path
// partition, load-balanced expand, single-key probe and list copy.
//
// One known_to_unknown step (RCache::known_to_unknown, core/rcache.hpp:
// 232-414) runs as
//   k_clear -> k_probe_scan (K2U) -> [total read back] -> k_partition -> k_expand
// replacing the reference's keygen, probe, edge-list fetch, Thrust scan and
// one-thread-per-row update (core/gpu_hash.cu:145-316, 527-610, 685-694).
// known_to_known / known_to_const (core/gpu_hash.cu:320-466, 613-682) run as
//   k_clear -> k_probe_scan (K2K|K2C)
// which probes, tests membership, scans the flags and compacts the kept rows
// in one pass. index_to_unknown / const_to_unknown are k_probe_one + k_copy.
//
// The per-tile bodies (probe_scan_tile, expand_tile) are shared with the
// batched kernels, which run many independent steps in one launch.
#pragma once

#include "device.h"

namespace wkg {
namespace {

enum StepMode : int { kModeK2U = 0, kModeK2K = 1, kModeK2C = 2 };

__global__ void k_clear(uint64_t *a, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    a[i] = 0;
}

// One step's view for a probe/scan tile.
struct StepParams {
  const int32_t *T;   // input table (null for a single-key step)
  uint64_t R;         // input rows (1 for a single-key step)
  uint32_t c;         // input columns
  uint32_t sc, ec;    // start column; K2K end column
  uint32_t pid, dir;
  uint32_t target;    // K2C constant
  uint32_t target_ok; // K2C constant fits u32 (otherwise nothing matches)
  int mode;           // kMode*
  uint64_t key;       // != 0: single-key step (index / const start), R == 1
  uint2 *rowinfo;     // K2U: (offset in predicate segment, degree) per row
  uint64_t *excl;     // K2U: exclusive output-row offset per row
  int32_t *O;         // filters: compacted output, capacity R x c
  uint64_t *out_rows; // total output rows (written by the step's last tile)
  uint64_t *examined; // filters: edges examined (atomic add)
  uint64_t *status;   // look-back words of the step's tiles
  uint64_t ntiles;    // tiles of this step
};

// Membership of `target` in a group's edge list (known_to_known/const test,
// core/engine.hpp:179-184,211-216). SORTED lists use an 8-lane 9-ary search:
// each lane reads one pivot, a ballot counts the pivots below the target and
// the range shrinks ~9x per round; the last <= 8 entries are compared at once.
// Unsorted lists are scanned 8 entries at a time with early exit, like the
// reference's linear scan. Both report the reference's work count: index of
// the first match + 1, or the degree when absent.
template <bool SORTED, bool FRAMED>
__device__ __forceinline__ void group_member(const GraphView &g, uint64_t estart, uint64_t efb, uint32_t off,
                                             uint32_t deg, uint32_t target, bool valid, bool &found,
                                             uint32_t &examined) {
  const uint32_t lane = lane_id(), l8 = lane & 7, gbase = lane & ~7u;
  bool act = valid && deg > 0;
  found = false;
  uint32_t idx = 0;
  if (SORTED) {
    uint32_t lo = 0, hi = deg;
    while (true) {
      const bool need = act && (hi - lo > 8);
      if (!wave_ballot(need)) break;
      bool lt = false;
      const uint32_t span = hi - lo;
      if (need) {
        const uint32_t pos = lo + (uint32_t)(((uint64_t)span * (l8 + 1)) / 9);
        lt = edge_at<FRAMED>(g, estart, efb, (uint64_t)off + pos) < target;
      }
      const uint32_t m = __builtin_popcount((uint32_t)(wave_ballot(need && lt) >> gbase) & 0xffu);
      if (need) {
        const uint32_t nlo = m == 0 ? lo : lo + (uint32_t)(((uint64_t)span * m) / 9) + 1;
        const uint32_t nhi = m == 8 ? hi : lo + (uint32_t)(((uint64_t)span * (m + 1)) / 9) + 1;
        lo = nlo;
        hi = nhi;
      }
    }
    bool eq = false;
    if (act && l8 < hi - lo) eq = edge_at<FRAMED>(g, estart, efb, (uint64_t)off + lo + l8) == target;
    const uint32_t gb = (uint32_t)(wave_ballot(eq) >> gbase) & 0xffu;
    if (act && gb) {
      found = true;
      idx = lo + (uint32_t)__builtin_ctz(gb);
    }
  } else {
    uint32_t j = 0;
    while (true) {
      const bool go = act && !found && j < deg;
      if (!wave_ballot(go)) break;
      bool eq = false;
      if (go && j + l8 < deg) eq = edge_at<FRAMED>(g, estart, efb, (uint64_t)off + j + l8) == target;
      const uint32_t gb = (uint32_t)(wave_ballot(eq) >> gbase) & 0xffu;
      if (go) {
        if (gb) {
          found = true;
          idx = j + (uint32_t)__builtin_ctz(gb);
        } else {
          j += 8;
        }
      }
    }
  }
  examined = found ? idx + 1 : deg;
}

// One tile (kTileRows rows) of a step. 32 groups of 8 lanes probe 8 rows each
// (8 bucket reads in flight per lane), then the per-row counts (degree, or
// kept flag) are scanned in the block and across the step's tiles with
// decoupled look-back. `lt` is the tile's index within its step.
template <bool SORTED, bool FRAMED>
__device__ __forceinline__ void probe_scan_tile(const GraphView &g, const StepParams &p, uint64_t lt) {
  __shared__ uint32_t s_cnt[kTileRows];
  __shared__ uint32_t s_aux[kTileRows];  // K2U: segment offset; filters: examined
  __shared__ uint32_t s_sel[kTileRows];  // filters: kept row ids in order
  __shared__ uint64_t s_tmp[kBlock / kWave];
  __shared__ uint64_t s_excl;
  constexpr int U = kTileRows / (kBlock / 8);  // rows per group
  const uint32_t tid = threadIdx.x, grp = tid >> 3, l8 = tid & 7;
  const uint64_t base = lt * kTileRows;

  uint64_t key[U];
  uint32_t pid[U], deg[U], off[U];
#pragma unroll
  for (int u = 0; u < U; u++) {
    const uint64_t row = base + grp + (uint64_t)u * (kBlock / 8);
    key[u] = 0;
    pid[u] = p.pid;
    if (row < p.R) {
      if (p.key) {
        key[u] = p.key;
      } else {
        const int32_t cell = p.T[row * p.c + p.sc];
        if (cell >= 0) key[u] = make_key((uint32_t)cell, p.pid, p.dir);
      }
    }
  }
  group_probe<U, FRAMED>(g, key, pid, deg, off);

  if (p.mode == kModeK2U) {
#pragma unroll
    for (int u = 0; u < U; u++)
      if (l8 == 0) {
        s_cnt[grp + u * (kBlock / 8)] = deg[u];
        s_aux[grp + u * (kBlock / 8)] = off[u];
      }
  } else {
    const DevMeta m = g.meta[p.pid < g.num_metas ? p.pid : 0];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const uint64_t row = base + grp + (uint64_t)u * (kBlock / 8);
      uint32_t target = p.target;
      bool valid = p.target_ok != 0;
      if (p.mode == kModeK2K) {
        int32_t cell = 0;
        if (row < p.R) cell = p.T[row * p.c + p.ec];
        target = (uint32_t)cell;
        valid = cell >= 0;
      }
      bool found;
      uint32_t examined;
      group_member<SORTED, FRAMED>(g, m.edge_start, m.eframe_base, off[u], deg[u], target, valid && row < p.R,
                                   found, examined);
      if (l8 == 0) {
        s_cnt[grp + u * (kBlock / 8)] = found ? 1u : 0u;
        s_aux[grp + u * (kBlock / 8)] = examined;
      }
    }
  }
  __syncthreads();

  const uint32_t cnt = s_cnt[tid];
  uint64_t total;
  const uint64_t lex = block_excl_scan(cnt, s_tmp, total);
  if (tid == 0) {
    s_excl = lookback(p.status, lt, total);
    if (lt == p.ntiles - 1) *p.out_rows = s_excl + total;
  }
  __syncthreads();
  const uint64_t row = base + tid;
  if (p.mode == kModeK2U) {
    if (row < p.R) {
      p.rowinfo[row] = make_uint2(s_aux[tid], cnt);
      p.excl[row] = s_excl + lex;
    }
  } else {
    uint64_t ex_total;
    (void)block_excl_scan(s_aux[tid], s_tmp, ex_total);
    if (tid == 0 && ex_total) atomicAdd((unsigned long long *)p.examined, (unsigned long long)ex_total);
    if (cnt) s_sel[lex] = tid;
    __syncthreads();
    // kept rows of this tile are contiguous in O from row s_excl on
    const uint32_t c = p.c, nsel = (uint32_t)total;
    const uint32_t rpp = kBlock / c;
    if (tid < rpp * c) {
      const uint32_t col = tid % c;
      for (uint32_t r = tid / c; r < nsel; r += rpp) {
        const uint64_t src = base + s_sel[r];
        p.O[(s_excl + r) * c + col] = p.T[src * c + col];
      }
    }
  }
}

template <bool SORTED, bool FRAMED>
__global__ __launch_bounds__(kBlock) void k_probe_scan(GraphView g, StepParams p, uint32_t *ticket) {
  __shared__ uint64_t s_tile;
  if (threadIdx.x == 0) s_tile = atomicAdd(ticket, 1u);
  __syncthreads();
  probe_scan_tile<SORTED, FRAMED>(g, p, s_tile);
}

// Merge-path partition of the (row starts, output rows) sequence: part[k] =
// number of rows whose start item lies before diagonal k * kExpandNV. Row i's
// start item sits at i + excl[i].
__device__ __forceinline__ uint64_t merge_split(const uint64_t *excl, uint64_t R, uint64_t E, uint64_t k) {
  const uint64_t diag = k * (uint64_t)kExpandNV < R + E ? k * (uint64_t)kExpandNV : R + E;
  uint64_t lo = 0, hi = R;  // first i with i + excl[i] >= diag (R if none)
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (mid + excl[mid] < diag) lo = mid + 1; else hi = mid;
  }
  return lo;
}

__global__ void k_partition(const uint64_t *excl, uint64_t R, uint64_t E, uint64_t ntiles, uint64_t *part) {
  const uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (k <= ntiles) part[k] = merge_split(excl, R, E, k);
}

struct ExpandParams {
  uint32_t pid;
  const int32_t *T;
  uint64_t R;
  uint32_t c;
  const uint2 *rowinfo;
  const uint64_t *excl;
  const uint64_t *part;  // ntiles + 1 split points
  uint64_t E;
  int32_t *O;  // E x (c+1)
};

// Load-balanced expand (known_to_unknown result update): a tile owns
// kExpandNV items of the merged sequence of row starts and output rows, so a
// hub row's outputs spread over many workgroups and runs of empty rows cost
// one item each. Output row o of input row i is (row i, edge k) with
// k = o - excl[i]; rows come out in input order and edges in list order
// (core/engine.hpp:142-150). Output cells are written with consecutive lanes
// on consecutive ints. `tile` is the tile index within the step.
template <bool FRAMED>
__device__ __forceinline__ void expand_tile(const GraphView &g, const ExpandParams &a, uint64_t tile) {
  __shared__ int32_t s_ex[kExpandNV + 1];
  __shared__ uint32_t s_off[kExpandNV + 1];
  __shared__ uint16_t s_orow[kExpandNV];
  const uint32_t tid = threadIdx.x;
  const uint64_t d0 = tile * (uint64_t)kExpandNV;
  const uint64_t dend = a.R + a.E;
  const uint64_t d1 = d0 + kExpandNV < dend ? d0 + kExpandNV : dend;
  const uint64_t i0 = a.part[tile], i1 = a.part[tile + 1];
  const uint64_t j0 = d0 - i0, j1 = d1 - i1;
  const uint32_t nr = (uint32_t)(i1 - i0 + 1);  // local row r <-> input row i0 - 1 + r
  for (uint32_t r = tid; r < nr; r += kBlock) {
    if (r == 0 && i0 == 0) {
      s_ex[0] = 0;
      s_off[0] = 0;
    } else {
      const uint64_t i = i0 + r - 1;
      s_ex[r] = (int32_t)((int64_t)a.excl[i] - (int64_t)j0);
      s_off[r] = a.rowinfo[i].x;
    }
  }
  __syncthreads();
  // local merged position of local row r's start item: (r - 1) + s_ex[r]
  const int32_t span = (int32_t)(d1 - d0);
  const int32_t x0 = (int32_t)(tid * kExpandVT);
  if (x0 < span) {
    uint32_t lo = 1, hi = nr;  // rows r >= 1 whose start lies before x0
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if ((int32_t)(mid - 1) + s_ex[mid] < x0) lo = mid + 1; else hi = mid;
    }
    uint32_t cur = lo - 1;
#pragma unroll
    for (int v = 0; v < kExpandVT; v++) {
      const int32_t x = x0 + v;
      if (x >= span) break;
      if (cur + 1 < nr && (int32_t)cur + s_ex[cur + 1] == x) {
        cur++;
      } else {
        s_orow[x - (int32_t)cur] = (uint16_t)cur;
      }
    }
  }
  __syncthreads();
  const uint32_t nout = (uint32_t)(j1 - j0);
  const uint32_t c = a.c, nc = c + 1;
  const uint32_t rpp = kBlock / nc;
  if (tid < rpp * nc) {
    const uint32_t col = tid % nc;
    const DevMeta m = g.meta[a.pid];
    for (uint32_t r = tid / nc; r < nout; r += rpp) {
      const uint32_t lr = s_orow[r];
      const uint64_t i = i0 + lr - 1;
      int32_t v;
      if (col < c) {
        v = a.T[i * c + col];
      } else {
        const uint32_t k = (uint32_t)((int32_t)r - s_ex[lr]);
        v = (int32_t)edge_at<FRAMED>(g, m.edge_start, m.eframe_base, (uint64_t)s_off[lr] + k);
      }
      a.O[(j0 + r) * nc + col] = v;
    }
  }
  __syncthreads();  // LDS is reused by the caller's next tile
}

template <bool FRAMED>
__global__ __launch_bounds__(kBlock) void k_expand(GraphView g, ExpandParams a) {
  expand_tile<FRAMED>(g, a, blockIdx.x);
}

// Single-key probe for index_to_unknown / const_to_unknown: out[0] = degree,
// out[1] = offset in the predicate segment. One wave.
template <bool FRAMED>
__global__ __launch_bounds__(kWave) void k_probe_one(GraphView g, uint64_t key, uint32_t pid, uint64_t *out) {
  uint64_t k[1] = {threadIdx.x < 8 ? key : 0};
  uint32_t p[1] = {pid}, deg[1], off[1];
  group_probe<1, FRAMED>(g, k, p, deg, off);
  if (threadIdx.x == 0) {
    out[0] = deg[0];
    out[1] = off[0];
  }
}

// out[j] = list[j] for j < n, the list starting at segment offset `off` of pid.
template <bool FRAMED>
__global__ __launch_bounds__(kBlock) void k_copy_list(GraphView g, uint32_t pid, uint64_t off, uint64_t n,
                                                      int32_t *out) {
  const DevMeta m = g.meta[pid];
  for (uint64_t j = blockIdx.x * (uint64_t)kBlock + threadIdx.x; j < n; j += (uint64_t)gridDim.x * kBlock)
    out[j] = (int32_t)edge_at<FRAMED>(g, m.edge_start, m.eframe_base, off + j);
}

}  // namespace
}  // namespace wkg
