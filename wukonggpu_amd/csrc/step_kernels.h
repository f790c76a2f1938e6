// Pattern-step kernels: probe, tile scan, finalize (compaction of the rows
// that produce output), merge-path partition, load-balanced expand,
// single-key probe and list copy.
//
// One known_to_unknown step (RCache::known_to_unknown, core/rcache.hpp:
// 232-414) runs as
//   k_probe -> k_scan_tiles -> k_finalize -> [totals read back] -> k_partition -> k_expand
// replacing the reference's keygen, probe, edge-list fetch, Thrust scan and
// one-thread-per-row update (core/gpu_hash.cu:145-316, 527-610, 685-694).
// known_to_known / known_to_const (core/gpu_hash.cu:320-466, 613-682) run as
//   k_probe -> k_scan_tiles -> k_finalize (stable compaction of kept rows)
// index_to_unknown / const_to_unknown are k_probe_copy (+ k_copy_list for a
// list above kListCap entries).
//
// The probe writes, per tile, the list of rows that have edges (or that a
// filter keeps) and the tile's sums; the tile scan turns the sums into
// prefixes; the finalize pass moves each tile's list to its global place.
// Everything after the probe touches only rows that produce output. (This is
// reduce-then-scan: a single-pass decoupled look-back serialises on
// device-scope loads that cross the XCDs' separate L2s.)
//
// The per-tile bodies are shared with the batched kernels, which run many
// independent steps in one launch.
#pragma once

#include "../../include/wukong_gpu.h"
#include "device.h"

namespace wkg {
namespace {

enum StepMode : int { kModeK2U = 0, kModeK2K = 1, kModeK2C = 2 };


// One step's view for the probe / scan / finalize passes.
struct StepParams {
  const int32_t *T;    // input table (null for a single-key step)
  uint64_t R;          // input rows (1 for a single-key step); < 2^32
  uint32_t c;          // input columns
  uint32_t sc, ec;     // start column; K2K end column
  uint32_t pid, dir;
  uint32_t target;     // K2C constant
  uint32_t target_ok;  // K2C constant fits u32 (otherwise nothing matches)
  int mode;            // kMode*
  uint64_t key;        // != 0: single-key step (index / const start), R == 1
  uint4 *hits;         // per tile, one entry per tile row: (row, offset, degree, 0) of rows with output
  uint2 *cinfo;        // K2U, compacted rows: (row, offset)
  uint64_t *cexcl;     // K2U, compacted rows: exclusive output-row offset
  int32_t *O;          // filters: compacted output, capacity R x c
  uint64_t *tile_sum;  // per tile: output rows; after the tile scan, the tile's prefix
  uint64_t *tile_nz;   // per tile: rows with output; after the tile scan, the prefix
  uint64_t *tile_ex;   // per tile: edges examined (filters) / keys probed (K2U)
  uint64_t *totals;    // [0] output rows, [1] edges examined (filters) or keys probed (K2U),
                       // [2] rows with output
  uint64_t ntiles;     // tiles of this step
  // two-level tile scan (null when one block scans all tiles): per chunk of
  // kScanChunk tiles, the chunk's prefix of output rows / rows with output
  uint64_t *chunk_sum, *chunk_nz, *chunk_ex;
  // per-step path: the totals are also published to the host by the kernel
  // that computes them (null: not published)
  Signal sg = {nullptr, 0};
  uint64_t tile0 = 0;  // k_probe / k_finalize: the launch's first tile (the per-step path's launches are split)
  // filters: each row's edges examined count `T[row * c + wc]` times (a
  // row standing for that many rows of another table)
  uint32_t weighted = 0, wc = 0;
};

constexpr int kScanChunk = 8 * kBlock;  // tiles per block in the two-level tile scan

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

// Probe pass over one tile (TR rows; `lt` is the tile's index within its
// step). Wave w owns the tile's rows [w * TR/4, (w+1) * TR/4) and probes
// them 64 consecutive rows per round, carrying the last key across rounds.
// Writes the tile's list of rows with output and its sums.
template <int TR, bool SORTED, bool FRAMED>
__device__ __forceinline__ void probe_tile(const GraphView &g, const StepParams &p, uint64_t lt) {
  constexpr int kRounds = TR / kBlock;  // one row per lane per round
  constexpr int kRowsPerThread = TR / kBlock;
  static_assert(TR % kBlock == 0, "tile rows must be a multiple of the block");
  __shared__ uint32_t s_cnt[TR];  // degree, or kept flag
  __shared__ uint32_t s_aux[TR];  // segment offset, or edges examined
  __shared__ uint64_t s_tmp[kBlock / kWave];
  __shared__ uint32_t s_probed[kBlock / kWave];
  __shared__ WaveProbeLds s_wp[kBlock / kWave];
  const uint32_t tid = threadIdx.x, lane = lane_id(), gbase = lane & ~7u, l8 = lane & 7;
  const uint32_t wave = tid / kWave;
  const uint64_t base = lt * TR;
  const DevMeta m = g.meta[p.pid < g.num_metas ? p.pid : 0];

  // all rounds' table reads are issued up front
  constexpr uint32_t kWaveRows = TR / (kBlock / kWave);
  const auto T = gmem(p.T);
  int32_t cells[kRounds], ends[kRounds];
#pragma unroll
  for (int rd = 0; rd < kRounds; rd++) {
    const uint64_t row = base + wave * kWaveRows + rd * kWave + lane;
    const bool in = row < p.R && !p.key;
    cells[rd] = in ? T[row * p.c + p.sc] : -1;
    ends[rd] = in && p.mode == kModeK2K ? T[row * p.c + p.ec] : 0;
  }
  ProbeCarry carry;
  uint32_t probed = 0;
  for (int rd = 0; rd < kRounds; rd++) {
    const uint32_t local = wave * kWaveRows + rd * kWave + lane;
    const uint64_t row = base + local;
    uint64_t key = 0;
    const int32_t end_cell = ends[rd];
    if (row < p.R) {
      if (p.key) key = p.key;
      else if (cells[rd] >= 0) key = make_key((uint32_t)cells[rd], p.pid, p.dir);
    }
    uint32_t deg, off;
    probed += wave_probe<FRAMED>(g, m, key, s_wp[wave], deg, off, carry);
    if (p.mode == kModeK2U) {
      s_cnt[local] = deg;
      s_aux[local] = off;
    } else {
      // the 8 lanes of a group test the group's 8 rows one after another
      uint32_t my_found = 0, my_ex = 0;
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const uint32_t src = gbase + u;
        const uint32_t du = shfl32(deg, src), ou = shfl32(off, src);
        uint32_t target = p.target;
        bool valid = p.target_ok != 0;
        if (p.mode == kModeK2K) {
          const int32_t ec = (int32_t)shfl32((uint32_t)end_cell, src);
          target = (uint32_t)ec;
          valid = ec >= 0;
        }
        bool found;
        uint32_t examined;
        group_member<SORTED, FRAMED>(g, m.edge_start, m.eframe_base, ou, du, target, valid, found, examined);
        if (l8 == (uint32_t)u) {
          my_found = found ? 1u : 0u;
          my_ex = examined;
        }
      }
      s_cnt[local] = my_found;
      s_aux[local] = my_ex;
    }
  }
  if (lane == 0) s_probed[wave] = probed;
  __syncthreads();
  // thread t owns the tile's rows 8t .. 8t+7
  uint64_t sum = 0, ex = 0;
  uint32_t nz = 0;
#pragma unroll
  for (int k = 0; k < kRowsPerThread; k++) {
    const uint32_t c = s_cnt[tid * kRowsPerThread + k];
    sum += c;
    nz += c ? 1u : 0u;
    uint64_t w = 1;
    if (p.weighted) {
      const uint64_t row = base + tid * kRowsPerThread + k;
      w = row < p.R ? (uint32_t)T[row * p.c + p.wc] : 0u;
    }
    ex += (uint64_t)s_aux[tid * kRowsPerThread + k] * w;
  }
  uint64_t nz_total, sum_total, ex_total = 0;
  uint32_t pos = (uint32_t)block_excl_scan(nz, s_tmp, nz_total);
  (void)block_excl_scan(sum, s_tmp, sum_total);
  if (p.mode != kModeK2U) (void)block_excl_scan(ex, s_tmp, ex_total);
  uint4 *hits = p.hits + lt * TR;
#pragma unroll
  for (int k = 0; k < kRowsPerThread; k++) {
    const uint32_t local = tid * kRowsPerThread + k;
    const uint32_t c = s_cnt[local];
    if (c) gstore(hits + pos++, make_uint4((uint32_t)(base + local), p.mode == kModeK2U ? s_aux[local] : 0u, c, 0u));
  }
  if (tid == 0) {
    gmem(p.tile_sum)[lt] = sum_total;
    gmem(p.tile_nz)[lt] = nz_total;
    uint32_t keys = 0;
#pragma unroll
    for (int w = 0; w < kBlock / kWave; w++) keys += s_probed[w];
    gmem(p.tile_ex)[lt] = p.mode == kModeK2U ? (uint64_t)keys : ex_total;
  }
}

// At most 128 VGPRs, so four waves per SIMD hide the probe's load chains.
template <int TR, bool SORTED, bool FRAMED>
__global__ __launch_bounds__(kBlock, 4) void k_probe(GraphView g, StepParams p) {
  probe_tile<TR, SORTED, FRAMED>(g, p, p.tile0 + blockIdx.x);
}

// In-place exclusive scan of a[0, n) by one block; returns the total.
__device__ __forceinline__ uint64_t block_scan_array(uint64_t *a_, uint64_t n, uint64_t *s_tmp) {
  const auto a = gmem(a_);
  const uint64_t per = (n + kBlock - 1) / kBlock;
  const uint64_t lo = threadIdx.x * per < n ? threadIdx.x * per : n;
  const uint64_t hi = lo + per < n ? lo + per : n;
  uint64_t s = 0;
  uint64_t i = lo;
  for (; i + 8 <= hi; i += 8) {  // 8 independent loads in flight
    uint64_t v[8];
#pragma unroll
    for (int k = 0; k < 8; k++) v[k] = a[i + k];
#pragma unroll
    for (int k = 0; k < 8; k++) s += v[k];
  }
  for (; i < hi; i++) s += a[i];
  uint64_t total;
  uint64_t run = block_excl_scan(s, s_tmp, total);
  for (i = lo; i < hi; i++) {
    const uint64_t v = a[i];
    a[i] = run;
    run += v;
  }
  return total;
}

// Tile scan of one step (one block): tile_sum and tile_nz become prefixes;
// writes the step's totals.
__device__ __forceinline__ void scan_step_tiles(const StepParams &p) {
  __shared__ uint64_t s_tmp[kBlock / kWave];
  const uint64_t rows = block_scan_array(p.tile_sum, p.ntiles, s_tmp);
  const uint64_t nz = block_scan_array(p.tile_nz, p.ntiles, s_tmp);
  uint64_t ex = 0, e = 0;
  for (uint64_t i = threadIdx.x; i < p.ntiles; i += kBlock) e += p.tile_ex[i];
  (void)block_excl_scan(e, s_tmp, ex);
  if (threadIdx.x == 0) {
    p.totals[0] = rows;
    p.totals[1] = ex;
    p.totals[2] = nz;
    if (p.sg.h) {
      const uint64_t v[3] = {rows, ex, nz};
      publish(p.sg, v, 3);
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_scan_tiles(StepParams p) { scan_step_tiles(p); }

// Two-level tile scan, level 1: chunk k (kScanChunk tiles) is scanned in
// place (chunk-local prefixes) and its sums stored. With a single chunk the
// step's totals and zero chunk prefixes are written directly.
__device__ __forceinline__ void scan_chunk(const StepParams &p, uint64_t chunk, uint64_t nchunks) {
  __shared__ uint64_t s_tmp[kBlock / kWave];
  constexpr int PT = kScanChunk / kBlock;
  const uint64_t lo = chunk * (uint64_t)kScanChunk + threadIdx.x * PT;
  const auto tile_sum = gmem(p.tile_sum), tile_nz = gmem(p.tile_nz);
  uint64_t vs[PT], vn[PT], s = 0, nzs = 0, e = 0;
#pragma unroll
  for (int k = 0; k < PT; k++) {
    const bool in = lo + k < p.ntiles;
    vs[k] = in ? tile_sum[lo + k] : 0;
    vn[k] = in ? tile_nz[lo + k] : 0;
    if (in) e += gmem(p.tile_ex)[lo + k];
    s += vs[k];
    nzs += vn[k];
  }
  uint64_t ts, tn, te;
  uint64_t rs = block_excl_scan(s, s_tmp, ts);
  uint64_t rn = block_excl_scan(nzs, s_tmp, tn);
  (void)block_excl_scan(e, s_tmp, te);
#pragma unroll
  for (int k = 0; k < PT; k++)
    if (lo + k < p.ntiles) {
      tile_sum[lo + k] = rs;
      tile_nz[lo + k] = rn;
      rs += vs[k];
      rn += vn[k];
    }
  if (threadIdx.x == 0) {
    if (nchunks == 1) {
      gmem(p.totals)[0] = ts;
      gmem(p.totals)[1] = te;
      gmem(p.totals)[2] = tn;
      gmem(p.chunk_sum)[0] = 0;
      gmem(p.chunk_nz)[0] = 0;
      if (p.sg.h) {
        const uint64_t v[3] = {ts, te, tn};
        publish(p.sg, v, 3);
      }
    } else {
      gmem(p.chunk_sum)[chunk] = ts;
      gmem(p.chunk_nz)[chunk] = tn;
      gmem(p.chunk_ex)[chunk] = te;
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_scan_chunks(StepParams p) { scan_chunk(p, blockIdx.x, gridDim.x); }

// Level 2 (one block): chunk sums become chunk prefixes; writes the totals.
__device__ __forceinline__ void scan_top(const StepParams &p, uint64_t nchunks) {
  __shared__ uint64_t s_tmp[kBlock / kWave];
  const uint64_t rows = block_scan_array(p.chunk_sum, nchunks, s_tmp);
  const uint64_t nz = block_scan_array(p.chunk_nz, nchunks, s_tmp);
  uint64_t e = 0, ex = 0;
  for (uint64_t i = threadIdx.x; i < nchunks; i += kBlock) e += gmem(p.chunk_ex)[i];
  (void)block_excl_scan(e, s_tmp, ex);
  if (threadIdx.x == 0) {
    gmem(p.totals)[0] = rows;
    gmem(p.totals)[1] = ex;
    gmem(p.totals)[2] = nz;
    if (p.sg.h) {
      const uint64_t v[3] = {rows, ex, nz};
      publish(p.sg, v, 3);
    }
  }
}

__global__ __launch_bounds__(kBlock) void k_scan_top(StepParams p, uint64_t nchunks) { scan_top(p, nchunks); }

// Finalize pass over one tile of TR rows, by one wave: K2U moves the tile's
// rows with edges to the compacted row list with their exclusive output
// offsets; the filters copy the kept rows, in order, to the compacted output.
// Waves work on their tiles independently (no block barrier), so tiles with
// few rows cost one short chain of loads each.
template <int TR>
__device__ __forceinline__ void finalize_tile(const StepParams &p, uint64_t lt) {
  const uint32_t lane = lane_id();
  auto nz_prefix = [&](uint64_t t) -> uint64_t {
    if (t >= p.ntiles) return gmem(p.totals)[2];
    return gmem(p.tile_nz)[t] + (p.chunk_nz ? gmem(p.chunk_nz)[t / kScanChunk] : 0);
  };
  const uint64_t nz0 = nz_prefix(lt);
  const uint32_t nnz = (uint32_t)(nz_prefix(lt + 1) - nz0);
  if (nnz == 0) return;
  const uint64_t out0 = gmem(p.tile_sum)[lt] + (p.chunk_sum ? gmem(p.chunk_sum)[lt / kScanChunk] : 0);
  const uint4 *hits = p.hits + lt * TR;
  const auto T = gmem(p.T);
  const auto O = gmem(p.O);
  if (p.mode == kModeK2U) {
    uint64_t carry = 0;
    for (uint32_t k0 = 0; k0 < nnz; k0 += kWave) {
      const uint32_t j = k0 + lane;
      const uint4 h = j < nnz ? gload(hits + j) : make_uint4(0, 0, 0, 0);
      const uint64_t inc = wave_incl_scan(h.z);
      if (j < nnz) {
        gstore(p.cinfo + nz0 + j, make_uint2(h.x, h.y));
        gmem(p.cexcl)[nz0 + j] = out0 + carry + inc - h.z;
      }
      carry += shfl64(inc, kWave - 1);
    }
  } else {
    // kept rows k0.. of this tile go to output rows out0 + k0 ..
    const uint32_t c = p.c;
    for (uint32_t k0 = 0; k0 < nnz; k0 += kWave) {
      const uint32_t nsel = nnz - k0 < (uint32_t)kWave ? nnz - k0 : (uint32_t)kWave;
      const uint32_t myrow = lane < nsel ? gload(hits + k0 + lane).x : 0u;
      if (c <= (uint32_t)kWave) {  // 64 / c rows per pass, one column per lane
        const uint32_t rpp = kWave / c, col = lane % c, r0 = lane / c;
        for (uint32_t it = 0; it * rpp < nsel; it++) {
          const uint32_t r = it * rpp + r0;
          const uint32_t src = shfl32(myrow, r < (uint32_t)kWave ? r : 0u);
          if (r0 < rpp && r < nsel) O[(out0 + k0 + r) * c + col] = T[(uint64_t)src * c + col];
        }
      } else {
        for (uint32_t r = 0; r < nsel; r++) {
          const uint32_t src = shfl32(myrow, r);
          for (uint32_t col = lane; col < c; col += kWave) O[(out0 + k0 + r) * c + col] = T[(uint64_t)src * c + col];
        }
      }
    }
  }
}

// One wave per tile: tile blockIdx.x * 4 + wave.
template <int TR>
__global__ __launch_bounds__(kBlock) void k_finalize(StepParams p) {
  const uint64_t t = p.tile0 + blockIdx.x * (uint64_t)(kBlock / kWave) + threadIdx.x / kWave;
  if (t < p.ntiles) finalize_tile<TR>(p, t);
}

// Merge-path partition of the (row starts, output rows) sequence over the
// compacted rows: part[k] = number of rows whose start item lies before
// diagonal k * NV. Row i's start item sits at i + cexcl[i].
template <int NV>
__device__ __forceinline__ uint64_t merge_split(const uint64_t *cexcl, uint64_t R, uint64_t E, uint64_t k) {
  const uint64_t diag = k * (uint64_t)NV < R + E ? k * (uint64_t)NV : R + E;
  uint64_t lo = 0, hi = R;  // first i with i + cexcl[i] >= diag (R if none)
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (mid + cexcl[mid] < diag) lo = mid + 1; else hi = mid;
  }
  return lo;
}

__global__ void k_partition(const uint64_t *cexcl, uint64_t R, uint64_t E, uint64_t ntiles, uint64_t *part) {
  const uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (k <= ntiles) part[k] = merge_split<kExpandNV>(cexcl, R, E, k);
}

struct ExpandParams {
  uint32_t pid;
  const int32_t *T;      // input table
  uint32_t c;            // input columns
  uint64_t R;            // compacted rows (rows with at least one edge)
  const uint2 *cinfo;    // compacted rows: (input row, segment offset)
  const uint64_t *cexcl; // compacted rows: exclusive output-row offset
  const uint64_t *part;  // ntiles + 1 split points
  uint64_t E;
  int32_t *O;  // E x (c+1)
  uint64_t tile0;  // k_expand: the launch's first tile
};

// Load-balanced expand (known_to_unknown result update): a tile owns
// NV = kBlock * VT items of the merged sequence of row starts and output rows, so a
// hub row's outputs spread over many workgroups. Output row o of compacted
// row i is (input row, edge k) with k = o - cexcl[i]; rows come out in input
// order and edges in list order (core/engine.hpp:142-150).
// Phase 1 stages the tile's rows in LDS; phase 2 walks the merge path and
// issues each thread's (up to VT) edge loads together, parking the
// edges in LDS; phase 3 writes the tile's contiguous output region with
// aligned 16-byte stores. `tile` is the tile index within the step.
template <int VT, bool FRAMED>
__device__ __forceinline__ void expand_tile(const GraphView &g, const ExpandParams &a, uint64_t tile) {
  constexpr int NV = kBlock * VT;
  static_assert(NV < 65536, "local output rows are kept in 16 bits");
  __shared__ int32_t s_ex[NV + 1];
  __shared__ uint32_t s_off[NV + 1];
  __shared__ uint32_t s_row[NV + 1];
  __shared__ uint16_t s_orow[NV];
  __shared__ uint32_t s_edge[NV];
  const uint32_t tid = threadIdx.x;
  const uint64_t d0 = tile * (uint64_t)NV;
  const uint64_t dend = a.R + a.E;
  const uint64_t d1 = d0 + NV < dend ? d0 + NV : dend;
  const uint64_t i0 = gmem(a.part)[tile], i1 = gmem(a.part)[tile + 1];
  const uint64_t j0 = d0 - i0, j1 = d1 - i1;
  const uint32_t nr = (uint32_t)(i1 - i0 + 1);  // local row r <-> compacted row i0 - 1 + r
  for (uint32_t r = tid; r < nr; r += kBlock) {
    if (r == 0 && i0 == 0) {
      s_ex[0] = 0;
      s_off[0] = 0;
      s_row[0] = 0;
    } else {
      const uint64_t i = i0 + r - 1;
      const uint2 ci = gload(a.cinfo + i);
      s_ex[r] = (int32_t)((int64_t)gmem(a.cexcl)[i] - (int64_t)j0);
      s_row[r] = ci.x;
      s_off[r] = ci.y;
    }
  }
  __syncthreads();
  // local merged position of local row r's start item: (r - 1) + s_ex[r]
  const int32_t span = (int32_t)(d1 - d0);
  const int32_t x0 = (int32_t)(tid * VT);
  const DevMeta m = g.meta[a.pid];
  if (x0 < span) {
    uint32_t lo = 1, hi = nr;  // rows r >= 1 whose start lies before x0
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if ((int32_t)(mid - 1) + s_ex[mid] < x0) lo = mid + 1; else hi = mid;
    }
    uint32_t cur = lo - 1;
    int32_t oidx[VT];
    uint32_t epos[VT];
#pragma unroll
    for (int v = 0; v < VT; v++) {
      const int32_t x = x0 + v;
      oidx[v] = -1;
      epos[v] = 0;
      if (x < span) {
        if (cur + 1 < nr && (int32_t)cur + s_ex[cur + 1] == x) {
          cur++;
        } else {
          const int32_t o = x - (int32_t)cur;
          oidx[v] = o;
          epos[v] = s_off[cur] + (uint32_t)(o - s_ex[cur]);
          s_orow[o] = (uint16_t)cur;
        }
      }
    }
    uint32_t ev[VT];
#pragma unroll
    for (int v = 0; v < VT; v++)
      ev[v] = oidx[v] >= 0 ? edge_at<FRAMED>(g, m.edge_start, m.eframe_base, epos[v]) : 0u;
#pragma unroll
    for (int v = 0; v < VT; v++)
      if (oidx[v] >= 0) s_edge[oidx[v]] = ev[v];
  }
  __syncthreads();
  // the tile's output is the contiguous int range [j0 * nc, j1 * nc)
  const uint32_t nout = (uint32_t)(j1 - j0);
  const uint32_t c = a.c, nc = c + 1;
  const uint32_t nint = nout * nc;
  const uint64_t g0 = j0 * nc;
  auto cell = [&](uint32_t r, uint32_t col) -> int32_t {
    return col < c ? gmem(a.T)[(uint64_t)s_row[s_orow[r]] * c + col] : (int32_t)s_edge[r];
  };
  // streaming (nontemporal) stores: the table is far larger than the caches
  typedef int32_t v4i __attribute__((ext_vector_type(4)));
  const uint32_t head = (uint32_t)((4 - (g0 & 3)) & 3) < nint ? (uint32_t)((4 - (g0 & 3)) & 3) : nint;
  const auto O = gmem(a.O);
  if (tid < head) __builtin_nontemporal_store(cell(tid / nc, tid % nc), &O[g0 + tid]);
  const uint32_t nquad = (nint - head) / 4;
  WKG_GLOBAL v4i *oq = (WKG_GLOBAL v4i *)(O + g0 + head);
  for (uint32_t q = tid; q < nquad; q += kBlock) {
    const uint32_t e = head + 4 * q;
    uint32_t r = e / nc, col = e - r * nc;
    v4i v;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      v[k] = cell(r, col);
      if (++col == nc) {
        col = 0;
        r++;
      }
    }
    __builtin_nontemporal_store(v, &oq[q]);
  }
  const uint32_t tail0 = head + 4 * nquad;
  if (tid < nint - tail0) {
    const uint32_t e = tail0 + tid;
    __builtin_nontemporal_store(cell(e / nc, e % nc), &O[g0 + e]);
  }
  __syncthreads();  // LDS is reused by the caller's next tile
}

template <bool FRAMED>
__global__ __launch_bounds__(kBlock) void k_expand(GraphView g, ExpandParams a) {
  expand_tile<kExpandVT, FRAMED>(g, a, a.tile0 + blockIdx.x);
}

// ---- a small step in one launch --------------------------------------------
// A known-start step over at most kSmallRows input rows runs in one
// workgroup: the probe tiles (kBlock rows each), the tile scan, the
// finalize and, for known_to_unknown with at most `cap` output rows, the
// merge-path split points and the expand into x.O (cap rows reserved) --
// the multi-launch path's own tile bodies, separated by workgroup barriers
// (after which the block sees its own global writes), so the tables are the
// same. A bigger output leaves the compacted rows for the multi-launch
// expand. Saves four launches on the per-step path's small steps, whose
// time is launch latency.
constexpr uint64_t kSmallRows = 2 * kBlock;
constexpr uint64_t kSmallCap = 4 * kExpandNV;  // output rows expanded in the same launch

template <bool SORTED, bool FRAMED>
__global__ __launch_bounds__(kBlock) void k_small_step(GraphView g, StepParams p, ExpandParams x, uint64_t cap,
                                                       uint64_t *part, Signal sg) {
  for (uint64_t lt = 0; lt < p.ntiles; lt++) {
    probe_tile<kBlock, SORTED, FRAMED>(g, p, lt);
    __syncthreads();  // the tile's LDS is reused by the next
  }
  scan_chunk(p, 0, 1);
  __syncthreads();
  for (uint64_t lt = threadIdx.x / kWave; lt < p.ntiles; lt += kBlock / kWave) finalize_tile<kBlock>(p, lt);
  __syncthreads();
  const uint64_t E = p.totals[0], nz = p.totals[2];
  if (p.mode == kModeK2U && E != 0 && E <= cap) {
    const uint64_t nx = (nz + E + kExpandNV - 1) / kExpandNV;
    for (uint64_t k = threadIdx.x; k <= nx; k += kBlock) part[k] = merge_split<kExpandNV>(p.cexcl, nz, E, k);
    __syncthreads();
    ExpandParams xx = x;
    xx.R = nz;
    xx.E = E;
    xx.part = part;
    for (uint64_t t = 0; t < nx; t++) expand_tile<kExpandVT, FRAMED>(g, xx, t);
  }
  if (threadIdx.x == 0) publish(sg, p.totals, 3);  // (thread 0 wrote the totals)
}

// ---- a whole small query in one workgroup ----
// Every step of a query whose tables stay small runs in one 1024-thread
// workgroup with every table in LDS (two ping-pong tables of kSQCells
// cells), so the only global accesses are the graph's buckets and lists and
// the final copy of the result to the host:
//  - an index / constant start: wave 0 probes the key, the workgroup copies
//    its list into the table;
//  - a known-start step: wave w probes rows [64 w r, 64 (w + 1) r) of the
//    input (r = rounds of 64 rows, at most 4), filters test membership with
//    the 8-lane group test, then a workgroup scan of the degrees (kept
//    flags) gives every row its first output row; known_to_unknown writes
//    output row k = (the last row whose first output row is <= k, its list
//    entry k - that first row), filters copy the kept rows in order. Output
//    order is the reference's (core/engine.hpp:126-194).
// A table past the caps ends the query early with status 1 (the caller then
// runs it step by step), so the result is always the per-step path's.
// Two kernels run this body: k_small_query, one launch per query, and
// k_sq_server, a persistent server that takes queries from a ring in pinned
// host memory as they arrive.
constexpr int kSQThreads = 1024;
constexpr int kSQWaves = kSQThreads / kWave;
constexpr uint32_t kSQSteps = 8;        // patterns
constexpr uint32_t kSQRows = 4096;      // input rows of a known-start step
constexpr uint32_t kSQCells = 8192;     // cells of a table (32 KB)
constexpr int kSQHalf = 4;  // list entries / output rows a thread gathers at once
static_assert(kSQCells * 4 <= kStageBytes, "a small query's result fits the staging area");
struct SmallQuery {
  wkg_plan_step steps[kSQSteps];
  DevMeta meta[kSQSteps];  // each step's predicate metadata (passed in, not loaded)
  uint32_t nsteps;
  uint32_t pad[15];
};
static_assert(sizeof(SmallQuery) % 64 == 0, "whole cache lines");

struct SQShared {
  WaveProbeLds wp[kSQWaves];
  uint32_t deg[kSQRows];  // degree (known_to_unknown) or kept flag (filters)
  uint32_t off[kSQRows];  // list offset (known_to_unknown)
  uint32_t pos[kSQRows];  // first output row
  alignas(16) int32_t tab[2][kSQCells];
  uint64_t tmp[2][kSQWaves];
  SmallQuery q;           // the query (copied in)
};

// Exclusive u64 scan over the kSQThreads threads; `total` gets the sum.
__device__ __forceinline__ uint64_t sq_block_scan(uint64_t v, uint64_t *tmp, uint64_t &total) {
  const uint32_t lane = lane_id(), wave = threadIdx.x / kWave;
  const uint64_t inc = wave_incl_scan(v);
  if (lane == kWave - 1) tmp[wave] = inc;
  __syncthreads();
  uint64_t pre = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kSQWaves; w++) {
    const uint64_t t = tmp[w];
    if ((uint32_t)w < wave) pre += t;
    tot += t;
  }
  __syncthreads();
  total = tot;
  return pre + inc - v;
}

__device__ __forceinline__ uint32_t sq_uniform(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }
__device__ __forceinline__ uint64_t sq_uniform(uint64_t v) {
  return ((uint64_t)sq_uniform((uint32_t)(v >> 32)) << 32) | sq_uniform((uint32_t)v);
}

// Runs sh.q; wave 0 writes the result table to dst (rows x cols cells, 16-
// byte aligned) unless the status is 1. res = {status, rows, cols,
// expanded, the step that declined}. Every thread of the workgroup calls
// it; wave 0 then releases its stores to the host.
template <bool SORTED, bool FRAMED>
__device__ __forceinline__ void sq_run(const GraphView &g, SQShared &sh, int32_t *dst, uint64_t (&res)[5]) {
  const SmallQuery &q = sh.q;
  const uint32_t tid = threadIdx.x, lane = lane_id(), wave = tid / kWave, l8 = lane & 7, gbase = lane & ~7u;
  uint64_t expanded = 0, status = 0;
  uint32_t rows = 0, cols = 0, cur = 0, l = 0;
  const uint32_t nsteps = sq_uniform((uint32_t)q.nsteps);
  for (; l < nsteps; l++) {
    // the step and its metadata are the same in every lane: into scalar registers
    wkg_plan_step d;
    d.kind = (int32_t)sq_uniform((uint32_t)q.steps[l].kind);
    d.pid = (int32_t)sq_uniform((uint32_t)q.steps[l].pid);
    d.dir = (int32_t)sq_uniform((uint32_t)q.steps[l].dir);
    d.start_col = (int32_t)sq_uniform((uint32_t)q.steps[l].start_col);
    d.end_col = (int32_t)sq_uniform((uint32_t)q.steps[l].end_col);
    d.vid = (int64_t)sq_uniform((uint64_t)q.steps[l].vid);
    DevMeta m;
    m.main_start = sq_uniform(q.meta[l].main_start);
    m.psz = sq_uniform(q.meta[l].psz);
    m.magic = sq_uniform(q.meta[l].magic);
    m.indrct_start = sq_uniform(q.meta[l].indrct_start);
    m.edge_start = sq_uniform(q.meta[l].edge_start);
    m.edge_len = 0;
    m.vframe_base = sq_uniform(q.meta[l].vframe_base);
    m.eframe_base = sq_uniform(q.meta[l].eframe_base);
    const int32_t *in = sh.tab[cur];
    int32_t *out = sh.tab[cur ^ 1];
    if (d.kind == WKG_STEP_I2U || d.kind == WKG_STEP_C2U) {  // one key: its list is the table
      const uint64_t key = d.kind == WKG_STEP_I2U ? make_key(0, (uint64_t)d.pid, (uint64_t)d.dir)
                                                  : make_key((uint64_t)d.vid, (uint64_t)d.pid, (uint64_t)d.dir);
      if (wave == 0) {
        uint32_t deg, off;
        ProbeCarry carry;
        wave_probe<FRAMED>(g, m, lane == 0 ? key : 0, sh.wp[0], deg, off, carry);
        if (lane == 0) {
          sh.deg[0] = deg;
          sh.off[0] = off;
        }
      }
      __syncthreads();
      const uint32_t deg = sh.deg[0], off = sh.off[0];
      expanded += deg;
      if (deg > kSQCells) {
        status = 1;
        break;
      }
#pragma unroll 1
      for (uint32_t j0 = 0; j0 < deg; j0 += kSQHalf * kSQThreads) {
        uint32_t v[kSQHalf];  // a thread's list loads in flight together
#pragma unroll
        for (int k = 0; k < kSQHalf; k++) {
          const uint32_t j = j0 + tid + k * kSQThreads;
          if (j < deg) v[k] = edge_at<FRAMED>(g, m.edge_start, m.eframe_base, (uint64_t)off + j);
        }
#pragma unroll
        for (int k = 0; k < kSQHalf; k++) {
          const uint32_t j = j0 + tid + k * kSQThreads;
          if (j < deg) out[j] = (int32_t)v[k];
        }
      }
      rows = deg;
      cols = 1;
    } else {
      if (rows > kSQRows) {
        status = 1;
        break;
      }
      const int mode = d.kind == WKG_STEP_K2U ? kModeK2U : d.kind == WKG_STEP_K2K ? kModeK2K : kModeK2C;
      const uint32_t sc = (uint32_t)d.start_col, ec = mode == kModeK2K ? (uint32_t)d.end_col : 0u;
      const bool target_ok = mode == kModeK2C && d.vid >= 0 && d.vid <= 0xffffffffLL;
      const uint32_t rounds = (rows + kSQThreads - 1) / kSQThreads;  // 64-row rounds of a wave
      ProbeCarry carry;
      uint64_t ex = 0;  // filters: edges examined by this lane's rows
      for (uint32_t rd = 0; rd < rounds; rd++) {
        const uint32_t row = (wave * rounds + rd) * kWave + lane;
        uint64_t key = 0;
        int32_t end_cell = -1;
        if (row < rows) {
          const int32_t cell = in[row * cols + sc];
          if (cell >= 0) key = make_key((uint32_t)cell, (uint64_t)d.pid, (uint64_t)d.dir);
          if (mode == kModeK2K) end_cell = in[row * cols + ec];
        }
        uint32_t deg, off;
        wave_probe<FRAMED>(g, m, key, sh.wp[wave], deg, off, carry);
        if (mode == kModeK2U) {
          if (row < rows) {
            sh.deg[row] = deg;
            sh.off[row] = off;
          }
        } else {
          // the 8 lanes of a group test the group's 8 rows one after another
          uint32_t my_found = 0, my_ex = 0;
#pragma unroll 1
          for (int u = 0; u < 8; u++) {
            const uint32_t src = gbase + u;
            const uint32_t du = shfl32(deg, src), ou = shfl32(off, src);
            uint32_t target = (uint32_t)d.vid;
            bool valid = target_ok;
            if (mode == kModeK2K) {
              const int32_t e = (int32_t)shfl32((uint32_t)end_cell, src);
              target = (uint32_t)e;
              valid = e >= 0;
            }
            bool found;
            uint32_t examined;
            group_member<SORTED, FRAMED>(g, m.edge_start, m.eframe_base, ou, du, target, valid, found, examined);
            if (l8 == (uint32_t)u) {
              my_found = found ? 1u : 0u;
              my_ex = examined;
            }
          }
          if (row < rows) sh.deg[row] = my_found;
          ex += my_ex;
        }
      }
      __syncthreads();
      // thread t owns rows [t r, t r + r) of the scan
      uint64_t cnt = 0;
      for (uint32_t k = 0; k < rounds; k++) {
        const uint32_t row = tid * rounds + k;
        if (row < rows) cnt += sh.deg[row];
      }
      uint64_t total, ex_total = 0;
      uint64_t pre = sq_block_scan(cnt, sh.tmp[0], total);
      if (mode != kModeK2U) (void)sq_block_scan(ex, sh.tmp[1], ex_total);
      for (uint32_t k = 0; k < rounds; k++) {
        const uint32_t row = tid * rounds + k;
        if (row < rows) {
          sh.pos[row] = (uint32_t)pre;  // (a total past the caps declines below before any use)
          pre += sh.deg[row];
        }
      }
      __syncthreads();
      if (mode == kModeK2U) {
        expanded += total;
        const uint32_t nc = cols + 1;
        if (total * nc > kSQCells) {
          status = 1;
          break;
        }
        const uint32_t E = (uint32_t)total;  // <= kSQCells
#pragma unroll 1
        for (uint32_t r0 = 0; r0 < E; r0 += kSQHalf * kSQThreads) {  // kSQHalf output rows a thread at once
          uint32_t par[kSQHalf], v[kSQHalf];
#pragma unroll
          for (int k = 0; k < kSQHalf; k++) {  // the last row whose first output row is <= r
            const uint32_t r = r0 + tid + k * kSQThreads;
            uint32_t lo = 0;
#pragma unroll
            for (uint32_t step = kSQRows / 2; step; step >>= 1)
              if (lo + step < rows && sh.pos[lo + step] <= r) lo += step;
            par[k] = lo;
          }
#pragma unroll
          for (int k = 0; k < kSQHalf; k++) {  // the gathered edges in flight together
            const uint32_t r = r0 + tid + k * kSQThreads;
            if (r < E) v[k] = edge_at<FRAMED>(g, m.edge_start, m.eframe_base, (uint64_t)sh.off[par[k]] + (r - sh.pos[par[k]]));
          }
#pragma unroll
          for (int k = 0; k < kSQHalf; k++) {
            const uint32_t r = r0 + tid + k * kSQThreads;
            if (r < E) {
              int32_t *o = out + r * nc;
              for (uint32_t j = 0; j < cols; j++) o[j] = in[par[k] * cols + j];
              o[cols] = (int32_t)v[k];
            }
          }
        }
        rows = E;
        cols = nc;
      } else {
        expanded += ex_total;
        for (uint32_t row = tid; row < rows; row += kSQThreads)
          if (sh.deg[row])
            for (uint32_t j = 0; j < cols; j++) out[sh.pos[row] * cols + j] = in[row * cols + j];
        rows = (uint32_t)total;
      }
    }
    __syncthreads();  // the output is complete before it is the next step's input
    cur ^= 1;
  }
  // the result to the host: wave 0 alone (one wave's system-scope release
  // then covers every store), 16 bytes a lane
  if (status == 0 && wave == 0) {
    const uint32_t cells = rows * cols, nq = cells / 4;
    const uint4 *src = (const uint4 *)sh.tab[cur];
    for (uint32_t i = lane; i < nq; i += kWave) ((uint4 *)dst)[i] = src[i];
    if (lane < cells - 4 * nq) dst[4 * nq + lane] = sh.tab[cur][4 * nq + lane];
  }
  res[0] = status;
  res[1] = rows;
  res[2] = cols;
  res[3] = expanded;
  res[4] = l;
}

// One query, one launch: the result into the pinned staging area, then
// {status, rows, cols, expanded} and the flag.
template <bool SORTED, bool FRAMED>
__global__ __launch_bounds__(kSQThreads) void k_small_query(GraphView g, SmallQuery q, int32_t *stage, Signal sg) {
  __shared__ SQShared sh;
  if (threadIdx.x == 0) sh.q = q;
  __syncthreads();
  uint64_t res[5];
  sq_run<SORTED, FRAMED>(g, sh, stage, res);
  if (threadIdx.x < kWave) {
    __threadfence_system();
    if (threadIdx.x == 0) publish(sg, res, 5);
  }
}

// ---- the persistent small-query server ----
// A ring of slots in pinned host memory: the host writes a query into slot
// t mod nslots, then publishes t + 1 as the ring's tail. Each workgroup of
// k_sq_server claims the next unclaimed ticket (a compare-and-swap on a
// device-memory head counter, so every ticket is taken once, in order),
// copies the query into LDS, runs it, writes the result table and
// {status, rows, cols, expanded, device ticks} into the slot, then t + 1 as
// the slot's done word (system-scope release). A workgroup exits when no
// ticket arrived for idle_ticks, when it has run for life_ticks, or when
// the host sets stop; every wait is bounded by those clocks (the device's
// constant wall clock), so the grid always drains. Queries arriving while
// others run are admitted as soon as a workgroup is free.
struct SQResp {
  uint64_t status, rows, cols, expanded, ticks;
  uint64_t step;  // status 1: the step that outgrew the limits
  uint64_t pad;
  uint64_t done;  // ticket + 1, written last
};
struct SQSlot {
  SmallQuery q;
  SQResp resp;
  int32_t cells[kSQCells];
};
struct SQBox {
  const uint64_t *tail;  // pinned: tickets issued (host-written)
  const uint64_t *stop;  // pinned: nonzero asks the workgroups to exit
  uint64_t *head;        // device memory: tickets claimed
  SQSlot *slots;         // pinned ring (device address)
  uint64_t nslots;       // a power of two
  uint64_t idle_ticks, life_ticks;
};
constexpr uint64_t kSQNoClaim = ~uint64_t(0);

template <bool SORTED>
__global__ __launch_bounds__(kSQThreads) void k_sq_server(GraphView g, SQBox box) {
  __shared__ SQShared sh;
  __shared__ uint64_t s_claim;
  const uint64_t t_start = wall_clock64();
  uint64_t t_idle = t_start;  // (thread 0) since when the ring is idle
  // (thread 0) the head as last seen: a stale value only fails one
  // compare-and-swap, which returns the current one
  uint64_t h = threadIdx.x == 0 ? __hip_atomic_load(box.head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
  for (;;) {
    if (threadIdx.x == 0) {
      uint64_t claim = kSQNoClaim;
      for (;;) {
        const uint64_t now = wall_clock64();
        if (now - t_start > box.life_ticks || now - t_idle > box.idle_ticks) break;
        if (__hip_atomic_load(box.stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM)) break;
        const uint64_t tail = __hip_atomic_load(box.tail, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
        if (h < tail) {
          if (__hip_atomic_compare_exchange_strong(box.head, &h, h + 1, __ATOMIC_ACQUIRE, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT)) {
            claim = h++;
            break;
          }
          continue;  // another workgroup took it (h is now the current head)
        }
        __builtin_amdgcn_s_sleep(4);
      }
      s_claim = claim;
    }
    __syncthreads();
    const uint64_t claim = s_claim;
    if (claim == kSQNoClaim) break;
    const uint64_t t0 = wall_clock64();
    SQSlot *slot = box.slots + (claim & (box.nslots - 1));
    // the query into LDS, word by word past every cache
    const uint32_t *src = (const uint32_t *)&slot->q;
    uint32_t *qw = (uint32_t *)&sh.q;
    for (uint32_t i = threadIdx.x; i < sizeof(SmallQuery) / 4; i += kSQThreads)
      qw[i] = __hip_atomic_load(src + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    __syncthreads();
    uint64_t res[5];
    sq_run<SORTED, false>(g, sh, slot->cells, res);
    if (threadIdx.x < kWave) __threadfence_system();  // (wave 0 wrote the cells)
    if (threadIdx.x == 0) {
      slot->resp.status = res[0];
      slot->resp.rows = res[1];
      slot->resp.cols = res[2];
      slot->resp.expanded = res[3];
      slot->resp.step = res[4];
      slot->resp.ticks = wall_clock64() - t0;
      __hip_atomic_store(&slot->resp.done, claim + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
      t_idle = wall_clock64();
    }
    __syncthreads();  // s_claim and sh are reused
  }
}

// A small table to the host: copies n cells into the pinned staging area,
// then publishes (every wave's stores complete at system scope first).
__global__ __launch_bounds__(kBlock) void k_stage(const int32_t *src, uint64_t n, int32_t *dst, Signal sg) {
  for (uint64_t i = threadIdx.x; i < n; i += kBlock) dst[i] = src[i];
  __threadfence_system();
  __syncthreads();
  if (threadIdx.x == 0) publish(sg, nullptr, 0);
}

// Index / const start in one launch: wave 0 probes the key; when the list
// has at most `cap` entries the workgroup copies it to out; then the degree
// and offset are published.
constexpr uint64_t kListCap = 8192;
template <bool FRAMED>
__global__ __launch_bounds__(kBlock) void k_probe_copy(GraphView g, uint64_t key, uint32_t pid, uint64_t cap,
                                                       int32_t *out, Signal sg) {
  __shared__ WaveProbeLds s_wp;
  __shared__ uint32_t s_deg, s_off;
  const DevMeta m = g.meta[pid];
  if (threadIdx.x < kWave) {
    uint32_t deg, off;
    ProbeCarry carry;
    wave_probe<FRAMED>(g, m, threadIdx.x == 0 ? key : 0, s_wp, deg, off, carry);
    if (threadIdx.x == 0) {
      s_deg = deg;
      s_off = off;
    }
  }
  __syncthreads();
  const uint64_t deg = s_deg, off = s_off;
  if (deg <= cap)
    for (uint64_t j = threadIdx.x; j < deg; j += kBlock) out[j] = (int32_t)edge_at<FRAMED>(g, m.edge_start, m.eframe_base, off + j);
  if (threadIdx.x == 0) {
    const uint64_t v[2] = {deg, off};
    publish(sg, v, 2);
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

// ---- known_to_unknown on an inherited column (probe once per parent row) ----
// A known_to_unknown output repeats compacted input row i (parent row
// info_p[i].x) in its rows [excl_p[i], excl_p[i+1]). A later step whose key
// column is one of the parent's gathers that key once per compacted parent
// row (k_gather_col), probes those keys as a one-column table, and spreads
// each kept parent row (m, with d list entries) back to its f child rows:
// child k of it gets output rows [out_pre[m] + k d, out_pre[m] + (k+1) d).

// K[i] = T[info[i].x * c + sc]
__global__ __launch_bounds__(kBlock) void k_gather_col(const int32_t *T, uint32_t c, uint32_t sc, const uint2 *info,
                                                       uint64_t n, int32_t *K) {
  for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock)
    K[i] = T[(uint64_t)info[i].x * c + sc];
}

// The kept parent rows: m < nk is key-table row info2[m].x (a compacted
// parent row i) with list length d = excl2[m+1] - excl2[m] (the last: dtot -
// excl2[m]); its child rows number f = excl_p[i+1] - excl_p[i] (the last
// compacted row's: rows - excl_p[i]).
struct KeptRows {
  const uint2 *info2;
  const uint64_t *excl2;
  uint64_t nk, dtot;
  const uint64_t *excl_p;
  uint64_t np, rows;
  __device__ __forceinline__ uint64_t deg(uint64_t m) const { return (m + 1 < nk ? excl2[m + 1] : dtot) - excl2[m]; }
  __device__ __forceinline__ uint64_t fan(uint32_t i) const {
    return (i + 1 < np ? excl_p[i + 1] : rows) - excl_p[i];
  }
};

// cnt[m] = f, out[m] = f d
__global__ __launch_bounds__(kBlock) void k_lift(KeptRows k, uint64_t *cnt, uint64_t *out) {
  for (uint64_t m = blockIdx.x * (uint64_t)kBlock + threadIdx.x; m < k.nk; m += (uint64_t)gridDim.x * kBlock) {
    const uint64_t f = k.fan(k.info2[m].x);
    cnt[m] = f;
    out[m] = f * k.deg(m);
  }
}

// Kept child row j (of kc): its parent m is the last with cnt_pre(m) <= j
// (cnt_pre strictly ascends: every compacted parent row has a child). The
// prefixes are the two-level tile scan's: chunk-local value + chunk prefix.
__global__ __launch_bounds__(kBlock) void k_spread(KeptRows k, const uint64_t *cnt_loc, const uint64_t *cnt_chunk,
                                                   const uint64_t *out_loc, const uint64_t *out_chunk, uint64_t kc,
                                                   uint2 *cinfo, uint64_t *cexcl) {
  auto cnt_pre = [&](uint64_t m) { return cnt_loc[m] + cnt_chunk[m / kScanChunk]; };
  for (uint64_t j = blockIdx.x * (uint64_t)kBlock + threadIdx.x; j < kc; j += (uint64_t)gridDim.x * kBlock) {
    uint64_t lo = 0, hi = k.nk;  // cnt_pre(lo) <= j < cnt_pre(hi)
    while (hi - lo > 1) {
      const uint64_t mid = (lo + hi) >> 1;
      if (cnt_pre(mid) <= j) lo = mid;
      else hi = mid;
    }
    const uint64_t kk = j - cnt_pre(lo);
    const uint2 pi = k.info2[lo];
    cinfo[j] = make_uint2((uint32_t)(k.excl_p[pi.x] + kk), pi.y);
    cexcl[j] = out_loc[lo] + out_chunk[lo / kScanChunk] + kk * k.deg(lo);
  }
}

// ---- filters on a star arm's output (decide once per parent group) ----
// An output A of the inherited-column known_to_unknown above consists of
// groups: group m < ng is kept key row m, i.e. grandparent row ginfo[m].x
// whose list (segment offset ginfo[m].y, d_m = gs[m+1] - gs[m] entries)
// was joined to each of its f_m child rows. Its rows are
// [go[m], go[m] + f_m d_m), child by child, so row r of group m holds list
// entry (r - go[m]) mod d_m in its last column and the grandparent row's
// cells in its first columns. A known_to_known / known_to_const step from
// that last column (with a grandparent column as the end) gives the same
// answer for all f_m rows that hold one entry: it is decided once per
// (group, entry) -- the d_m rows of a table S -- and each decision is
// counted f_m times.

// Group m's first child row of the output, its grandparent row and list.
__global__ __launch_bounds__(kBlock) void k_keep_groups(KeptRows k, const uint2 *pinfo, const uint64_t *out_loc,
                                                        const uint64_t *out_chunk, uint2 *ginfo, uint64_t *gs,
                                                        uint64_t *go) {
  for (uint64_t m = blockIdx.x * (uint64_t)kBlock + threadIdx.x; m < k.nk; m += (uint64_t)gridDim.x * kBlock) {
    const uint2 i2 = k.info2[m];
    ginfo[m] = make_uint2(pinfo[i2.x].x, i2.y);
    gs[m] = k.excl2[m];
    go[m] = out_loc[m] + out_chunk[m / kScanChunk];
  }
}

// The group table: G[m] = (end cell of the grandparent row (0 without an
// end column), f_m), the list of G row m at segment offset ginfo[m].y.
__global__ __launch_bounds__(kBlock) void k_group_table(const int32_t *P, uint32_t pc, int32_t ec, const uint2 *ginfo,
                                                        const uint64_t *gs, const uint64_t *go, uint64_t ng,
                                                        uint64_t dtot, uint64_t rows, int32_t *G, uint2 *gci) {
  for (uint64_t m = blockIdx.x * (uint64_t)kBlock + threadIdx.x; m < ng; m += (uint64_t)gridDim.x * kBlock) {
    const uint2 gi = ginfo[m];
    const uint64_t d = (m + 1 < ng ? gs[m + 1] : dtot) - gs[m];
    const uint64_t f = ((m + 1 < ng ? go[m + 1] : rows) - go[m]) / d;
    G[2 * m] = ec >= 0 ? P[(uint64_t)gi.x * pc + (uint32_t)ec] : 0;
    G[2 * m + 1] = (int32_t)(uint32_t)f;
    gci[m] = make_uint2((uint32_t)m, gi.y);
  }
}

// D[row] = 1 for every row of a filter's probe pass that was kept: tile lt's
// kept rows are hits[lt * tr + i], i below its count (tile prefixes of the
// two-level scan).
__global__ __launch_bounds__(kBlock) void k_hits_flags(const uint4 *hits, const uint64_t *tile_nz,
                                                       const uint64_t *chunk_nz, uint64_t ntiles, uint64_t nz_total,
                                                       uint32_t tr, uint8_t *D) {
  auto pre = [&](uint64_t t) -> uint64_t {
    return t >= ntiles ? nz_total : tile_nz[t] + chunk_nz[t / kScanChunk];
  };
  for (uint64_t lt = blockIdx.x; lt < ntiles; lt += gridDim.x) {
    const uint64_t a = pre(lt), n = pre(lt + 1) - a;
    for (uint64_t i = threadIdx.x; i < n; i += kBlock) D[hits[lt * tr + i].x] = 1;
  }
}

__global__ __launch_bounds__(kBlock) void k_zero_bytes(uint8_t *D, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) D[i] = 0;
}

// Largest i in [lo, hi] with a[i] <= x (a ascending, a[lo] <= x).
__device__ __forceinline__ uint64_t last_le(const uint64_t *a, uint64_t lo, uint64_t hi, uint64_t x) {
  while (lo < hi) {
    const uint64_t mid = (lo + hi + 1) >> 1;
    if (a[mid] <= x) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Filter pass over one TR-row tile of a group-structured table: row r of
// group m is kept iff D[gs[m] + (r - go[m]) mod d_m]. Writes the tile's
// kept rows and sums exactly as the probe pass of a filter does (no edges
// examined: they were counted on S), so the tile scan and the compaction
// that follow are the filter's own.
template <int TR>
__global__ __launch_bounds__(kBlock) void k_flag_groups(StepParams p, const uint64_t *go, const uint64_t *gs,
                                                        uint64_t ng, uint64_t dtot, const uint8_t *D) {
  constexpr int kRowsPerThread = TR / kBlock;
  static_assert(TR % kBlock == 0, "tile rows must be a multiple of the block");
  __shared__ uint64_t s_tmp[kBlock / kWave];
  __shared__ uint64_t s_m[2];
  const uint64_t lt = p.tile0 + blockIdx.x;
  const uint32_t tid = threadIdx.x;
  const uint64_t base = lt * TR;
  if (tid == 0) {  // the groups this tile's rows lie in: [m0, m1]
    const uint64_t last = (base + TR < p.R ? base + TR : p.R) - 1;
    const uint64_t m0 = last_le(go, 0, ng - 1, base);
    s_m[0] = m0;
    s_m[1] = last_le(go, m0, ng - 1, last);
  }
  __syncthreads();
  const uint64_t m0 = s_m[0], m1 = s_m[1];
  const uint64_t r0 = base + (uint64_t)tid * kRowsPerThread;
  uint32_t keep = 0, nz = 0;  // bit k: row r0 + k kept
  if (r0 < p.R) {
    uint64_t m = last_le(go, m0, m1, r0);
#pragma unroll
    for (int k = 0; k < kRowsPerThread; k++) {
      const uint64_t r = r0 + k;
      if (r < p.R) {
        while (m < m1 && go[m + 1] <= r) m++;  // every group has rows: at most one step
        const uint64_t d = (m + 1 < ng ? gs[m + 1] : dtot) - gs[m];
        const uint32_t e = (uint32_t)(r - go[m]) % (uint32_t)d;
        if (D[gs[m] + e]) {
          keep |= 1u << k;
          nz++;
        }
      }
    }
  }
  uint64_t nz_total;
  uint32_t pos = (uint32_t)block_excl_scan(nz, s_tmp, nz_total);
  uint4 *hits = p.hits + lt * TR;
#pragma unroll
  for (int k = 0; k < kRowsPerThread; k++)
    if (keep >> k & 1u) hits[pos++] = make_uint4((uint32_t)(r0 + k), 0u, 1u, 0u);
  if (tid == 0) {
    p.tile_sum[lt] = nz_total;
    p.tile_nz[lt] = nz_total;
    p.tile_ex[lt] = 0;
  }
}


// ---- filters from a fan-out's new column (the parent's list decided per entry once) ----
// A known_to_unknown output T with lineage repeats compacted parent row i
// (parent table row info[i].x, list at segment offset info[i].y) in its rows
// [excl[i], excl[i+1]); row excl[i] + j holds the list's entry j in its last
// column. A known_to_known / known_to_const step from that column (the end
// column, if any, one of the parent's) keeps row excl[i] + j iff the filter
// target of parent i is in the list of (entry j, pid, dir). Parents that
// share a list share its entries' keys, so each distinct list's entries are
// probed once (D: (segment offset, degree) per entry under the filter's
// predicate); each row then only tests its target against that list, and
// the input table's cells are read for kept rows alone.
struct FanoutFilter {
  const uint64_t *excl;   // first child row per compacted parent (ascending)
  const uint2 *info;      // (parent row, list offset) per compacted parent
  uint64_t np;            // compacted parents
  const uint64_t *pbase;  // per compacted parent: its list's first entry in D
  const uint2 *D;         // per distinct-list entry: (segment offset, or the entry if degree 1; degree) of its key
  const int32_t *P;       // parent table
  uint32_t pc;            // parent columns
  int32_t ec;             // known_to_known end column (a parent column); -1: known_to_const
  uint32_t *bits;         // per tile: TR / 32 words, bit k of word w: row 32 w + k kept
};

// Keys of the distinct lists: K[k] = entry k - lbase[L] of list L (segment
// offset ly[L] of pid), L the list holding k.
template <bool FRAMED>
__global__ __launch_bounds__(kBlock) void k_gather_lists(GraphView g, uint32_t pid, const uint64_t *lbase,
                                                         const uint32_t *ly, uint64_t nl, uint64_t total, int32_t *K) {
  const DevMeta m = g.meta[pid];
  for (uint64_t k = blockIdx.x * (uint64_t)kBlock + threadIdx.x; k < total; k += (uint64_t)gridDim.x * kBlock) {
    uint64_t lo = 0, hi = nl - 1;  // last L with lbase[L] <= k
    while (lo < hi) {
      const uint64_t mid = (lo + hi + 1) >> 1;
      if (lbase[mid] <= k) lo = mid;
      else hi = mid - 1;
    }
    K[k] = (int32_t)edge_at<FRAMED>(g, m.edge_start, m.eframe_base, (uint64_t)ly[lo] + (k - lbase[lo]));
  }
}

// D[k] = (0, 0) for every key, then (x, degree) for the keys with edges
// (the compacted rows of a known_to_unknown probe pass over K): x is the
// list's only entry when the degree is 1 (a row then needs no list read),
// else the list's segment offset.
__global__ __launch_bounds__(kBlock) void k_zero_u64(uint64_t *a, uint64_t n) {
  for (uint64_t i = blockIdx.x * (uint64_t)kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) a[i] = 0;
}
__global__ __launch_bounds__(kBlock) void k_key_degrees(GraphView g, uint32_t pid, const uint2 *cinfo,
                                                        const uint64_t *cexcl, uint64_t nz, uint64_t E, uint2 *D) {
  const DevMeta m = g.meta[pid];
  for (uint64_t k = blockIdx.x * (uint64_t)kBlock + threadIdx.x; k < nz; k += (uint64_t)gridDim.x * kBlock) {
    const uint64_t deg = (k + 1 < nz ? cexcl[k + 1] : E) - cexcl[k];
    const uint32_t x = deg == 1 ? edge_at<false>(g, m.edge_start, m.eframe_base, cinfo[k].y) : cinfo[k].y;
    D[cinfo[k].x] = make_uint2(x, (uint32_t)deg);
  }
}

// Filter pass over one TR-row tile of T. Each thread takes RPT consecutive
// rows: its parent (a binary search for the first row, then walking), list
// entry, D entry and target; a list of at most kLaneList entries is scanned
// by the row's own lane (the reference's early-exit scan, so its work count
// is the scan's: first match + 1, or the degree). Longer lists go through
// the probe pass's 8-lane membership test in a second phase, which a block
// skips when it has none. Kept flags become the tile's bit words; the tile's
// sums are kept rows and edges examined.
template <int TR, bool SORTED>
__global__ __launch_bounds__(kBlock) void k_fanout_filter(GraphView g, StepParams p, FanoutFilter f) {
  constexpr int RPT = TR / kBlock;
  constexpr uint32_t kLaneList = 8;
  static_assert(TR % kBlock == 0 && (TR / 32) <= kBlock, "tile rows");
  __shared__ uint32_t s_off[TR], s_deg[TR], s_tgt[TR];
  __shared__ uint8_t s_keep[TR], s_big[TR];
  __shared__ uint64_t s_m[2];
  __shared__ uint32_t s_any;
  const uint64_t lt = p.tile0 + blockIdx.x;
  const uint32_t tid = threadIdx.x, lane = lane_id(), gbase = lane & ~7u, l8 = lane & 7;
  const uint64_t base = lt * TR;
  const DevMeta m = g.meta[p.pid < g.num_metas ? p.pid : 0];
  if (tid == 0) {  // the parents this tile's rows lie in: [i0, i1]
    const uint64_t last = (base + TR < p.R ? base + TR : p.R) - 1;
    const uint64_t i0 = last_le(f.excl, 0, f.np - 1, base);
    s_m[0] = i0;
    s_m[1] = last_le(f.excl, i0, f.np - 1, last);
    s_any = 0;
  }
  __syncthreads();
  uint64_t ex = 0;
  {
    const uint64_t i0 = s_m[0], i1 = s_m[1];
    const uint64_t r0 = base + (uint64_t)tid * RPT;
    uint64_t i = r0 < p.R ? last_le(f.excl, i0, i1, r0) : i0;
    uint32_t off[RPT], deg[RPT], tgt[RPT];
    bool valid[RPT];
#pragma unroll
    for (int k = 0; k < RPT; k++) {  // the D entries and targets, loads issued together
      const uint64_t r = r0 + k;
      off[k] = 0;
      deg[k] = 0;
      tgt[k] = 0;
      valid[k] = false;
      if (r < p.R) {
        while (i < i1 && f.excl[i + 1] <= r) i++;  // every parent has rows: one step at most per row
        const uint2 e = f.D[f.pbase[i] + (r - f.excl[i])];
        off[k] = e.x;
        deg[k] = e.y;
        if (f.ec >= 0) {
          const int32_t cell = f.P[(uint64_t)f.info[i].x * f.pc + (uint32_t)f.ec];
          tgt[k] = (uint32_t)cell;
          valid[k] = cell >= 0;
        } else {
          tgt[k] = p.target;
          valid[k] = p.target_ok != 0;
        }
      }
    }
    bool big_any = false;
#pragma unroll
    for (int k = 0; k < RPT; k++) {
      const uint32_t local = tid * RPT + k;
      uint8_t keep = 0, big = 0;
      if (valid[k] && deg[k] > kLaneList) {
        big = 1;
        s_off[local] = off[k];
        s_deg[local] = deg[k];
        s_tgt[local] = tgt[k];
      } else if (deg[k] == 1) {  // D holds the list's only entry
        keep = valid[k] && off[k] == tgt[k] ? 1 : 0;
        ex += 1;
      } else {
        uint32_t x = deg[k];
        if (valid[k])
          for (uint32_t j = 0; j < deg[k]; j++)
            if (edge_at<false>(g, m.edge_start, m.eframe_base, (uint64_t)off[k] + j) == tgt[k]) {
              keep = 1;
              x = j + 1;
              break;
            }
        ex += x;
      }
      s_keep[local] = keep;
      s_big[local] = big;
      big_any |= big != 0;
    }
    if (big_any) s_any = 1;
  }
  __syncthreads();
  if (s_any) {  // long lists: the 8-lane membership test, 64 rows per wave per round
    constexpr uint32_t kWaveRows = TR / (kBlock / kWave);
    const uint32_t wave = tid / kWave;
    for (int rd = 0; rd < TR / kBlock; rd++) {
      const uint32_t local = wave * kWaveRows + rd * kWave + lane;
      const bool big = s_big[local] != 0;
      const uint32_t off = big ? s_off[local] : 0u, deg = big ? s_deg[local] : 0u, tgt = big ? s_tgt[local] : 0u;
      if (!wave_ballot(big)) continue;
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const uint32_t src = gbase + u;
        const uint32_t du = shfl32(deg, src), ou = shfl32(off, src), tu = shfl32(tgt, src);
        const bool bu = shfl32(big ? 1u : 0u, src) != 0;
        bool found;
        uint32_t examined;
        group_member<SORTED, false>(g, m.edge_start, m.eframe_base, ou, du, tu, bu, found, examined);
        if (l8 == (uint32_t)u && big) {
          s_keep[local] = found ? 1 : 0;
          ex += examined;
        }
      }
    }
  }
  __syncthreads();
  uint32_t word = 0, nz = 0;
  if (tid < TR / 32) {
#pragma unroll
    for (int b = 0; b < 32; b++) word |= (uint32_t)s_keep[tid * 32 + b] << b;
    nz = __builtin_popcount(word);
    f.bits[lt * (TR / 32) + tid] = word;
  }
  uint64_t v[2] = {nz, ex}, tot[2];
  __shared__ uint64_t s_tmp2[2 * (kBlock / kWave)];
  block_excl_scan_n<2>(v, s_tmp2, tot);
  if (tid == 0) {
    p.tile_sum[lt] = tot[0];
    p.tile_nz[lt] = tot[0];
    p.tile_ex[lt] = tot[1];
  }
}

// The same decisions over parent-aligned units: unit (i, jb) is rows
// [excl[i] + jb TR, ...) of parent i, at most TR of them, all with one
// target. Units are ranked in row order (rank = ubase[i] + jb: the order of
// the scan and the compaction) but executed list by list, j-block by
// j-block, the parents sharing a list innermost (block e of list L: j-block
// (e - xbase[L]) / |parents of L|), so the blocks that read one slice of a
// shared list's D entries run together and find it in their L2.
struct FanoutUnits {
  const uint64_t *xbase;  // per distinct list (+1): first execution index
  const uint64_t *poff;   // per distinct list (+1): its parents' first slot in par
  const uint32_t *par;    // compacted parents, grouped by list, row order within a list
  const uint64_t *ubase;  // per compacted parent: rank of its first unit
  uint64_t nl;            // distinct lists
  uint64_t *row0;         // per unit rank: its first row
};

template <int TR, bool SORTED>
__global__ __launch_bounds__(kBlock) void k_fanout_units(GraphView g, StepParams p, FanoutFilter f, FanoutUnits u) {
  constexpr int RPT = TR / kBlock;
  constexpr uint32_t kLaneList = 8;
  static_assert(RPT == 8 && (TR / 32) <= kBlock, "a thread's rows are one byte of kept bits");
  __shared__ uint32_t s_off[TR], s_deg[TR];
  __shared__ uint8_t s_keep[TR];             // long lists' decisions (phase 2)
  __shared__ uint8_t s_bigb[kBlock];         // per thread: which of its rows have long lists
  __shared__ __align__(4) uint8_t s_kb[kBlock];  // per thread: its rows' kept bits
  __shared__ uint32_t s_any;
  const uint32_t tid = threadIdx.x, lane = lane_id(), gbase = lane & ~7u, l8 = lane & 7;
  const uint64_t e = p.tile0 + blockIdx.x;
  // the unit: list L, j-block, parent
  uint64_t L = 0, hi = u.nl - 1;
  while (L < hi) {
    const uint64_t mid = (L + hi + 1) >> 1;
    if (u.xbase[mid] <= e) L = mid;
    else hi = mid - 1;
  }
  const uint64_t npl = u.poff[L + 1] - u.poff[L], e2 = e - u.xbase[L];
  const uint64_t jb = e2 / npl;
  const uint32_t i = u.par[u.poff[L] + e2 % npl];
  const uint64_t first = f.excl[i];
  const uint64_t n = (i + 1 < f.np ? f.excl[i + 1] : p.R) - first;
  const uint64_t j0 = jb * TR;
  const uint32_t nrows = (uint32_t)(n - j0 < (uint64_t)TR ? n - j0 : (uint64_t)TR);
  const uint64_t rank = u.ubase[i] + jb;
  uint32_t tgt;
  bool valid;
  if (f.ec >= 0) {
    const int32_t cell = f.P[(uint64_t)f.info[i].x * f.pc + (uint32_t)f.ec];
    tgt = (uint32_t)cell;
    valid = cell >= 0;
  } else {
    tgt = p.target;
    valid = p.target_ok != 0;
  }
  const DevMeta m = g.meta[p.pid < g.num_metas ? p.pid : 0];
  if (tid == 0) s_any = 0;
  const uint2 *Dp = f.D + f.pbase[i] + j0;
  uint2 d[RPT];
#pragma unroll
  for (int k = 0; k < RPT; k++) {  // the unit's D entries, loads issued together
    const uint32_t local = tid * RPT + k;
    d[k] = local < nrows ? Dp[local] : make_uint2(0u, 0u);
  }
  __syncthreads();
  uint64_t ex = 0;
  uint32_t kb = 0, bigb = 0;  // this thread's rows: kept bits, long-list bits
#pragma unroll
  for (int k = 0; k < RPT; k++) {
    const uint32_t local = tid * RPT + k;
    const uint32_t off = d[k].x, deg = d[k].y;
    if (valid && deg > kLaneList) {
      bigb |= 1u << k;
      s_off[local] = off;
      s_deg[local] = deg;
    } else if (deg == 1) {  // D holds the list's only entry
      kb |= (valid && off == tgt ? 1u : 0u) << k;
      ex += 1;
    } else {
      uint32_t x = deg;
      if (valid)
        for (uint32_t j = 0; j < deg; j++)
          if (edge_at<false>(g, m.edge_start, m.eframe_base, (uint64_t)off + j) == tgt) {
            kb |= 1u << k;
            x = j + 1;
            break;
          }
      ex += x;
    }
  }
  s_bigb[tid] = (uint8_t)bigb;
  if (bigb) s_any = 1;
  __syncthreads();
  if (s_any) {  // long lists: the 8-lane membership test, 64 rows per wave per round
    constexpr uint32_t kWaveRows = TR / (kBlock / kWave);
    const uint32_t wave = tid / kWave;
    for (int rd = 0; rd < TR / kBlock; rd++) {
      const uint32_t local = wave * kWaveRows + rd * kWave + lane;
      const bool big = (s_bigb[local / RPT] >> (local % RPT)) & 1u;
      const uint32_t off = big ? s_off[local] : 0u, deg = big ? s_deg[local] : 0u;
      if (!wave_ballot(big)) continue;
#pragma unroll
      for (int q = 0; q < 8; q++) {
        const uint32_t src = gbase + q;
        const uint32_t du = shfl32(deg, src), ou = shfl32(off, src);
        const bool bu = shfl32(big ? 1u : 0u, src) != 0;
        bool found;
        uint32_t examined;
        group_member<SORTED, false>(g, m.edge_start, m.eframe_base, ou, du, tgt, bu, found, examined);
        if (l8 == (uint32_t)q && big) {
          s_keep[local] = found ? 1 : 0;
          ex += examined;
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < RPT; k++)
      if (bigb >> k & 1u) kb |= (uint32_t)s_keep[tid * RPT + k] << k;
  }
  s_kb[tid] = (uint8_t)kb;
  __syncthreads();
  uint32_t word = 0, nz = 0;
  if (tid < TR / 32) {  // rows 32 tid .. 32 tid + 31: four threads' bytes
    word = ((const uint32_t *)s_kb)[tid];
    nz = __builtin_popcount(word);
    f.bits[rank * (TR / 32) + tid] = word;
  }
  uint64_t v[2] = {nz, ex}, tot[2];
  __shared__ uint64_t s_tmp2[2 * (kBlock / kWave)];
  block_excl_scan_n<2>(v, s_tmp2, tot);
  if (tid == 0) {
    p.tile_sum[rank] = tot[0];
    p.tile_nz[rank] = tot[0];
    p.tile_ex[rank] = tot[1];
    u.row0[rank] = first + j0;
  }
}

// The fan-out filter for long shared lists, a block of list entries at a
// time: a workgroup takes kFBRows entries of one distinct list and up to
// kFBGroup (= 64) of the parents sharing that list, and tests every parent
// against every entry's list, holding the block in registers:
//  - thread t takes entries t, t + 256, ...: their D entries are read once;
//    an entry whose list has one element keeps it in a register (D holds
//    it), so a (parent, entry) test is one compare;
//  - the few entries with longer lists are tested before the parent loop,
//    each by one wave with lane pi testing parent pi (the list's entries
//    broadcast lane by lane, or a binary search per lane for a long sorted
//    list), leaving a 64-bit parent mask per entry;
//  - per parent, each wave's compares are ballots: 64 rows' kept flags at
//    once (kept rows are rare: pass 1 only notes the parents with one in a
//    mask, and pass 2 redoes those parents' ballots);
//  - a workgroup's time is its chain of dependent loads (list, parent slot,
//    parent row, end cell), so the block's D entries are loaded first, from
//    the list's base, and arrive while the parents' targets are fetched.
// The output is per unit (parent x block of kFBRows rows, ranked
// ubase[parent] + block) its kept rows and edges examined, for the tile
// scan, and a list of the kept rows with their unit and index in it, for
// k_compact_kept. Work counts follow the reference's scan: first match + 1,
// or the degree.
constexpr int kFBRows = 2048;
constexpr int kFBPer = kFBRows / kBlock;
constexpr int kFBGroup = 64;
struct FanoutBlocks {
  const uint64_t *xb;     // per distinct list (+1): first work index (blocks x parent groups)
  const uint64_t *poff;   // per distinct list (+1): its parents' first slot in par
  const uint32_t *par;    // compacted parents, grouped by list, row order within a list
  const uint64_t *ubase;  // per compacted parent: rank of its first unit of kFBRows rows
  const uint64_t *ln;     // per distinct list: its entries
  const uint64_t *lbase;  // per distinct list: its first entry in D
  uint64_t nl;
  uint64_t *row0;         // per unit rank: its first row
  // kept rows, unordered: (row, unit rank, index within the unit, 0); a row
  // past the capacity is counted but not stored (the caller then reruns)
  uint4 *kept;
  uint64_t *nkept;
  uint64_t cap;
};

template <bool SORTED>
__global__ __launch_bounds__(kBlock) void k_fanout_blocks(GraphView g, StepParams p, FanoutFilter f, FanoutBlocks u) {
  constexpr int W = kBlock / kWave;
  constexpr uint32_t kNone = 0xffffffffu, kNoTarget = 0xfffffffeu;  // (ids and valid targets are below 2^31)
  static_assert(kFBGroup == kWave, "lane pi tests parent pi");
  __shared__ uint32_t s_boff[kFBRows], s_bdeg[kFBRows];
  __shared__ uint64_t s_bmask[kFBRows];            // per longer list: the parents in it
  __shared__ uint16_t s_pc[kFBGroup][kFBPer][W];   // per parent: kept rows per (k, wave) ballot
  __shared__ uint32_t s_kcnt[kFBGroup];            // per parent: kept rows in the block
  __shared__ uint32_t s_tgt[kFBGroup], s_par[kFBGroup];
  __shared__ uint64_t s_rank[kFBGroup];
  __shared__ uint64_t s_bex[W][kWave], s_tmp[W];
  __shared__ uint32_t s_nb;
  __shared__ unsigned long long s_hit;            // parents with a kept row in the block
  const uint32_t tid = threadIdx.x, lane = lane_id(), wave = tid / kWave;
  const uint64_t e = p.tile0 + blockIdx.x;
  uint64_t L = 0, hi = u.nl - 1;  // the list: the last with xb <= e
  while (L < hi) {
    const uint64_t mid = (L + hi + 1) >> 1;
    if (u.xb[mid] <= e) L = mid;
    else hi = mid - 1;
  }
  const uint64_t cnt = u.poff[L + 1] - u.poff[L], ngr = (cnt + kFBGroup - 1) / kFBGroup;
  const uint64_t e2 = e - u.xb[L], jb = e2 / ngr, grp = e2 % ngr;
  const uint64_t n = u.ln[L], j0 = jb * kFBRows;
  const uint32_t nrows = (uint32_t)(n - j0 < (uint64_t)kFBRows ? n - j0 : (uint64_t)kFBRows);
  const uint32_t npg = (uint32_t)(cnt - grp * kFBGroup < (uint64_t)kFBGroup ? cnt - grp * kFBGroup : kFBGroup);
  // the block's D entries, loaded now and used once the parents' targets
  // (a chain of dependent loads) are in
  const uint2 *Dp = f.D + u.lbase[L] + j0;
  uint2 dv[kFBPer];
#pragma unroll
  for (int k = 0; k < kFBPer; k++) {
    const uint32_t row = tid + k * kBlock;
    dv[k] = row < nrows ? Dp[row] : make_uint2(0u, 0u);
  }
  const DevMeta m = g.meta[p.pid < g.num_metas ? p.pid : 0];
  if (tid == 0) {
    s_nb = 0;
    s_hit = 0;
  }
  for (uint32_t z = tid; z < kFBGroup * kFBPer * W / 2; z += kBlock) ((uint32_t *)&s_pc[0][0][0])[z] = 0;
  if (tid < kFBGroup) {
    uint32_t tg = kNoTarget, i = 0;
    if (tid < npg) {
      i = u.par[u.poff[L] + grp * kFBGroup + tid];
      if (f.ec >= 0) {
        const int32_t cell = f.P[(uint64_t)f.info[i].x * f.pc + (uint32_t)f.ec];
        if (cell >= 0) tg = (uint32_t)cell;
      } else if (p.target_ok) {
        tg = p.target;
      }
      s_rank[tid] = u.ubase[i] + jb;
    }
    s_par[tid] = i;
    s_tgt[tid] = tg;
    s_kcnt[tid] = 0;
  }
  __syncthreads();
  uint32_t c0[kFBPer], bi[kFBPer];  // single entry (or kNone); a longer list's index (or kNone)
  uint64_t ex1 = 0;  // single-entry rows: one edge examined each, for every parent
#pragma unroll
  for (int k = 0; k < kFBPer; k++) {
    const uint2 d = dv[k];
    c0[k] = d.y == 1 ? d.x : kNone;
    ex1 += d.y == 1 ? 1u : 0u;
    bi[k] = kNone;
    if (d.y >= 2) {
      const uint32_t b = atomicAdd(&s_nb, 1u);
      bi[k] = b;
      s_boff[b] = d.x;
      s_bdeg[b] = d.y;
    }
  }
  __syncthreads();
  const uint32_t nb = s_nb;
  {
    // each longer list by one wave, lane pi testing parent pi
    const uint32_t tgt = s_tgt[lane];
    uint64_t exl = 0;
    for (uint32_t b = wave; b < nb; b += W) {
      const uint32_t off = s_boff[b], deg = s_bdeg[b];
      bool fnd = false;
      uint32_t x = deg;
      if (SORTED && deg > kWave) {
        if (tgt != kNoTarget) {
          uint32_t lo = 0, hi2 = deg;
          while (lo < hi2) {
            const uint32_t mid = (lo + hi2) >> 1;
            if (edge_at<false>(g, m.edge_start, m.eframe_base, (uint64_t)off + mid) < tgt) lo = mid + 1;
            else hi2 = mid;
          }
          if (lo < deg && edge_at<false>(g, m.edge_start, m.eframe_base, (uint64_t)off + lo) == tgt) {
            fnd = true;
            x = lo + 1;
          }
        }
      } else {  // wave-uniform: 64 entries at a time, broadcast lane by lane
        for (uint32_t q0 = 0; q0 < deg; q0 += kWave) {
          const uint32_t ev = q0 + lane < deg ? edge_at<false>(g, m.edge_start, m.eframe_base, (uint64_t)off + q0 + lane) : kNone;
          const uint32_t qn = deg - q0 < (uint32_t)kWave ? deg - q0 : (uint32_t)kWave;
          for (uint32_t q = 0; q < qn; q++) {
            const uint32_t v = (uint32_t)__builtin_amdgcn_readlane((int)ev, (int)q);
            if (!fnd && tgt != kNoTarget && v == tgt) {
              fnd = true;
              x = q0 + q + 1;
            }
          }
        }
      }
      const uint64_t mask = wave_ballot(fnd);
      if (lane == 0) s_bmask[b] = mask;
      exl += x;
    }
    s_bex[wave][lane] = exl;
  }
  uint64_t ex1_total;
  (void)block_excl_scan(ex1, s_tmp, ex1_total);  // (its barriers also order the masks above)
  uint64_t mk[kFBPer];
  uint32_t kmulti = 0;  // wave-uniform: bit k set when a row k of the wave has a longer list
#pragma unroll
  for (int k = 0; k < kFBPer; k++) {
    mk[k] = bi[k] != kNone ? s_bmask[bi[k]] : 0ull;
    kmulti |= wave_ballot(bi[k] != kNone) ? 1u << k : 0u;
  }
  const uint32_t tgt_l = s_tgt[lane];  // lane pi: parent pi's target
  auto ballot_k = [&](uint32_t pi, uint32_t tgt, int k) -> uint64_t {
    const uint64_t m1 = __builtin_amdgcn_uicmp(c0[k], tgt, 32);  // (32: equal) 64 rows' compares at once
    if (kmulti >> k & 1u) return m1 | wave_ballot(((mk[k] >> pi) & 1u) != 0);
    return m1;
  };
  // pass 1: every parent against every single-entry row: one compare per
  // 64 rows, only noting parents with a match (rare) in a wave-uniform mask
  uint64_t hit = 0;
  for (uint32_t pi = 0; pi < npg; pi++) {
    const uint32_t tgt = (uint32_t)__builtin_amdgcn_readlane((int)tgt_l, (int)pi);
    uint64_t any = 0;
#pragma unroll
    for (int k = 0; k < kFBPer; k++) any |= __builtin_amdgcn_uicmp(c0[k], tgt, 32);
    hit |= (any ? 1ull : 0ull) << pi;
  }
  if (hit && lane == 0) atomicOr(&s_hit, (unsigned long long)hit);
  // ... and the parents a longer list keeps
  for (uint32_t b = tid; b < nb; b += kBlock)
    if (s_bmask[b]) atomicOr(&s_hit, (unsigned long long)s_bmask[b]);
  __syncthreads();
  // pass 2, parents with kept rows only: the ballots again, counted per
  // (k, wave), then each kept row listed with its index among its parent's
  // kept rows (row order: k, then wave, then lane)
  for (uint64_t hm = s_hit; hm; hm &= hm - 1) {  // block-uniform
    const uint32_t pi = (uint32_t)__builtin_ctzll(hm);
    const uint32_t tgt = (uint32_t)__builtin_amdgcn_readlane((int)tgt_l, (int)pi);
    uint64_t bw[kFBPer];
#pragma unroll
    for (int k = 0; k < kFBPer; k++) {
      bw[k] = ballot_k(pi, tgt, k);
      if (lane == 0) s_pc[pi][k][wave] = (uint16_t)__builtin_popcountll(bw[k]);
    }
    __syncthreads();
    const uint64_t rank = s_rank[pi], first = f.excl[s_par[pi]] + j0;
    uint32_t before = 0;  // kept rows of parent pi at smaller (k, wave)
#pragma unroll
    for (int k = 0; k < kFBPer; k++) {
      for (int w = 0; w < W; w++)
        if (w < (int)wave) before += s_pc[pi][k][w];
      if (bw[k]) {
        uint64_t base = 0;
        if (lane == 0) base = atomicAdd((unsigned long long *)u.nkept, (unsigned long long)__builtin_popcountll(bw[k]));
        base = shfl64(base, 0);
        if ((bw[k] >> lane) & 1u) {
          const uint32_t lr = (uint32_t)__builtin_popcountll(bw[k] & lanes_below());
          const uint64_t slot = base + lr;
          if (slot < u.cap)
            u.kept[slot] = make_uint4((uint32_t)(first + (uint32_t)k * kBlock + wave * kWave + lane), (uint32_t)rank,
                                      before + lr, 0u);
        }
      }
      for (int w = (int)wave; w < W; w++) before += s_pc[pi][k][w];
    }
    if (tid == 0) s_kcnt[pi] = before;  // (wave 0 counted every wave's rows)
  }
  __syncthreads();
  if (tid < npg) {
    uint64_t xt = ex1_total;
#pragma unroll
    for (int w = 0; w < W; w++) xt += s_bex[w][tid];
    const uint64_t rank = s_rank[tid];
    p.tile_sum[rank] = s_kcnt[tid];
    p.tile_nz[rank] = s_kcnt[tid];
    p.tile_ex[rank] = xt;
  }
}

// Compaction of the kept rows k_fanout_blocks listed: one thread per kept
// row, its place = its unit's prefix from the two-level scan + its index in
// the unit; the row is built from its lineage as in k_compact_lineage.
__global__ __launch_bounds__(kBlock) void k_compact_kept(GraphView g, StepParams p, const uint4 *kept, uint64_t n,
                                                         FanoutFilter f, uint64_t edge_start) {
  const uint64_t k = blockIdx.x * (uint64_t)kBlock + threadIdx.x;
  if (k >= n) return;
  const uint4 e = kept[k];
  const uint64_t r = e.x, rank = e.y;
  const uint64_t o = p.tile_nz[rank] + (p.chunk_nz ? p.chunk_nz[rank / kScanChunk] : 0) + e.z;
  uint64_t lo = 0, hi = f.np - 1;
  while (lo < hi) {
    const uint64_t mid = (lo + hi + 1) >> 1;
    if (f.excl[mid] <= r) lo = mid;
    else hi = mid - 1;
  }
  const uint32_t c = p.c, pc = f.pc;
  const uint2 inf = f.info[lo];
  const int32_t *prow = f.P + (uint64_t)inf.x * pc;
  for (uint32_t col = 0; col < pc; col++) p.O[o * c + col] = prow[col];
  p.O[o * c + pc] = (int32_t)g.edges[edge_start + inf.y + (r - f.excl[lo])];
}

// Compaction of a fan-out filter's kept rows: one wave per tile reads the
// tile's TR / 32 bit words and copies each kept row of T to its place in O
// (the tile's prefix from the two-level scan, then bit order).
template <int TR>
__global__ __launch_bounds__(kBlock) void k_compact_bits(StepParams p, const uint32_t *bits, const uint64_t *row0) {
  static_assert(TR / 32 <= kWave, "one word per lane");
  const uint64_t lt = p.tile0 + blockIdx.x * (uint64_t)(kBlock / kWave) + threadIdx.x / kWave;
  if (lt >= p.ntiles) return;
  const uint64_t r_base = row0 ? row0[lt] : lt * TR;  // unit-ranked tiles carry their first row
  const uint32_t lane = lane_id();
  const uint32_t w = lane < TR / 32 ? bits[lt * (TR / 32) + lane] : 0u;
  if (!wave_ballot(w != 0)) return;
  const uint64_t pre = p.tile_nz[lt] + (p.chunk_nz ? p.chunk_nz[lt / kScanChunk] : 0);
  const uint64_t inc = wave_incl_scan((uint64_t)__builtin_popcount(w));
  uint64_t o = pre + inc - (uint64_t)__builtin_popcount(w);
  const uint32_t c = p.c;
  for (uint32_t x = w; x; x &= x - 1) {
    const uint64_t r = r_base + (uint64_t)lane * 32 + (uint32_t)__builtin_ctz(x);
    for (uint32_t col = 0; col < c; col++) p.O[o * c + col] = p.T[r * c + col];
    o++;
  }
}

// The same compaction for a fan-out that was never written: kept row r is
// built from its lineage -- the parent row (the last compacted parent whose
// first child row is <= r, a binary search) and that parent's list entry
// r - its first child row, of the step's predicate (resident graph).
template <int TR>
__global__ __launch_bounds__(kBlock) void k_compact_lineage(GraphView g, StepParams p, const uint32_t *bits,
                                                            const uint64_t *row0, FanoutFilter f,
                                                            uint64_t edge_start) {
  static_assert(TR / 32 <= kWave, "one word per lane");
  const uint64_t lt = p.tile0 + blockIdx.x * (uint64_t)(kBlock / kWave) + threadIdx.x / kWave;
  if (lt >= p.ntiles) return;
  const uint64_t r_base = row0 ? row0[lt] : lt * TR;
  const uint32_t lane = lane_id();
  const uint32_t w = lane < TR / 32 ? bits[lt * (TR / 32) + lane] : 0u;
  if (!wave_ballot(w != 0)) return;
  const uint64_t pre = p.tile_nz[lt] + (p.chunk_nz ? p.chunk_nz[lt / kScanChunk] : 0);
  const uint64_t inc = wave_incl_scan((uint64_t)__builtin_popcount(w));
  uint64_t o = pre + inc - (uint64_t)__builtin_popcount(w);
  const uint32_t c = p.c, pc = f.pc;
  for (uint32_t x = w; x; x &= x - 1) {
    const uint64_t r = r_base + (uint64_t)lane * 32 + (uint32_t)__builtin_ctz(x);
    uint64_t lo = 0, hi = f.np - 1;
    while (lo < hi) {
      const uint64_t mid = (lo + hi + 1) >> 1;
      if (f.excl[mid] <= r) lo = mid;
      else hi = mid - 1;
    }
    const uint2 inf = f.info[lo];
    const int32_t *prow = f.P + (uint64_t)inf.x * pc;
    for (uint32_t col = 0; col < pc; col++) p.O[o * c + col] = prow[col];
    p.O[o * c + pc] = (int32_t)g.edges[edge_start + inf.y + (r - f.excl[lo])];
    o++;
  }
}

}  // namespace
}  // namespace wkg
