// Fork-join split and reply merge of result tables: the one exchange step a
// sharded deployment adds per pattern, on one GPU's share
// (GPU_Engine::generate_sub_query, core/gpu_engine.hpp:293-350;
// RCache::generate_sub_query, core/rcache.hpp:964-1043; its kernels
// core/gpu_hash.cu:698-873; the merge of the sub-replies,
// core/reply_map.hpp:43-80).
//
// Row i goes to sub-table cell[i][col] mod nparts, read as an unsigned 32-bit
// vertex id (the owner rule hash_mod(vid, servers) = vid % servers,
// utils/mymath.hpp:55-59). The sub-tables land back to back in one output
// table. The reference gives each row its slot with an atomicAdd, so the
// order inside a sub-table is whatever order the atomics ran in; here it is
// the input order (a stable partition), one of those orders and the one a
// test can pin.
//
// Two passes over slabs of rows (1024, or 16 per part when there are more
// than 64 parts, so the count array stays under a byte per row), one wave per
// slab:
//   count:   rows per part in the slab -> cnt[part][slab];
//   scan:    exclusive scan of cnt in part-major order (rocPRIM) -> base,
//            so base[part][slab] is where the slab's rows of that part go;
//   scatter: each 64-row chunk ranks its rows inside their parts, then the
//            rows are copied to their places (rows of one part are
//            contiguous there too); with 9..64 parts and rows of 1, 2 or 4
//            cells, a workgroup's four slabs are first sorted by part in LDS
//            and leave as one run per part (k_split_scatter_staged).
// A chunk's rows are grouped by part with one ballot per bit of the part
// number (ceil(log2 parts) ballots, whatever the parts in the chunk); with
// at most 64 parts the per-part counts and offsets live in one register per
// lane (lane q: part q), with more in LDS.
#include <hip/hip_runtime.h>
#include <string.h>

#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <vector>

#include "../../include/wukong_gpu.h"
#include "wkg_internal.h"

using namespace wkg;

namespace {

constexpr int kWaves = 4;  // waves per workgroup
constexpr int kSplitBlock = kWaves * kWave;
constexpr int kMaxParts = 1024;
constexpr int kU = 8;    // 64-row chunks whose row loads a wave issues together (scatter)
constexpr int kUC = 16;  // the same for the count pass's key loads

struct SplitArgs {
  const int32_t *in;
  int32_t *out;
  uint64_t rows;
  uint32_t cols, col, parts;
  uint64_t pmagic;  // fastmod magic of parts
  uint64_t cmagic;  // fastmod magic of cols
  uint32_t nbits;   // bits of the largest part number
  uint64_t slab;    // rows per slab (a multiple of 64)
  uint64_t nslabs;
  uint32_t *cnt;         // [parts][nslabs]
  const uint64_t *base;  // exclusive scan of cnt, same layout
};

// floor(x / d) with m = fastmod_magic(d): the estimate is at most two low.
__device__ __forceinline__ uint32_t fastdiv32(uint32_t x, uint32_t d, uint64_t m) {
  uint64_t q = __umul64hi((uint64_t)x, m);
  uint64_t r = (uint64_t)x - q * d;
  if (r >= d) {
    q++;
    r -= d;
  }
  if (r >= d) q++;
  return (uint32_t)q;
}

__device__ __forceinline__ uint32_t part_of(const SplitArgs &a, uint64_t r) {
  const uint32_t v = (uint32_t)a.in[r * a.cols + a.col];
  return (uint32_t)fastmod(v, a.parts, a.pmagic);
}

// A chunk's rows grouped by part, with one ballot per bit of the part
// number: peers = the valid lanes whose row has this lane's part; own = the
// valid lanes whose row's part is this lane's index (meaningful for lanes
// below the part count when there are at most 64 parts).
struct Groups {
  uint64_t peers, own;
};
__device__ __forceinline__ Groups group_parts(uint32_t p, bool valid, uint32_t lane, uint32_t nbits) {
  const uint64_t v = wave_ballot(valid);
  uint64_t peers = v, own = v;
  for (uint32_t b = 0; b < nbits; b++) {
    const uint64_t bb = wave_ballot(valid && ((p >> b) & 1u));
    peers &= ((p >> b) & 1u) ? bb : ~bb;
    own &= ((lane >> b) & 1u) ? bb : ~bb;
  }
  return {peers, own};
}

// V consecutive cells as one vector access.
template <int V>
struct VecOf {
  typedef int T __attribute__((ext_vector_type(V)));
};

// Per slab: rows of each part. SMALL (at most 64 parts): lane q keeps part
// q's count in a register; otherwise one LDS add per distinct part per
// chunk, by the group's first lane. Dynamic LDS: kWaves * parts counters
// (none when SMALL). The keys of kUC chunks are loaded before any is used,
// so a wave has kUC loads in flight.
template <bool SMALL, int V>
__global__ __launch_bounds__(kSplitBlock) void k_split_count(SplitArgs a) {
  using T = typename VecOf<V>::T;
  extern __shared__ uint64_t lds_count[];
  const uint32_t w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const uint64_t slab = (uint64_t)blockIdx.x * kWaves + w;
  if (slab >= a.nslabs) return;  // the whole wave; no workgroup barrier below
  uint32_t *hw = (uint32_t *)lds_count + (uint64_t)w * a.parts;
  if (!SMALL) {
    for (uint32_t p = lane; p < a.parts; p += kWave) hw[p] = 0;
    wave_lds_sync();
  }
  const uint64_t below = lanes_below();
  uint32_t own_rows = 0;
  const uint64_t r0 = slab * a.slab, r1 = min(r0 + a.slab, a.rows);
  for (uint64_t cb = r0; cb < r1; cb += kUC * kWave) {
    T kv[kUC];  // the V-cell vector holding each row's key cell
#pragma unroll
    for (int u = 0; u < kUC; u++) {
      const uint64_t r = cb + (uint64_t)u * kWave + lane;
      kv[u] = T(0);
      if (r < r1) kv[u] = ((const T *)(a.in + r * a.cols))[a.col / V];
    }
    uint32_t key[kUC];
#pragma unroll
    for (int u = 0; u < kUC; u++) {
      uint32_t k = 0;
#pragma unroll
      for (int c = 0; c < V; c++)
        if ((uint32_t)c == a.col % V) k = (uint32_t)kv[u][c];
      key[u] = k;
    }
#pragma unroll
    for (int u = 0; u < kUC; u++) {
      const uint64_t r = cb + (uint64_t)u * kWave + lane;
      const bool valid = r < r1;
      const uint32_t p = valid ? (uint32_t)fastmod(key[u], a.parts, a.pmagic) : 0u;
      const Groups g = group_parts(p, valid, lane, a.nbits);
      if (SMALL) own_rows += (uint32_t)__builtin_popcountll(g.own);
      else if (valid && (g.peers & below) == 0) atomicAdd(&hw[p], (uint32_t)__builtin_popcountll(g.peers));
    }
  }
  if (SMALL) {
    if (lane < a.parts) a.cnt[(uint64_t)lane * a.nslabs + slab] = own_rows;
  } else {
    wave_lds_sync();
    for (uint32_t p = lane; p < a.parts; p += kWave) a.cnt[(uint64_t)p * a.nslabs + slab] = hw[p];
  }
}

// First row of every part: base[part][0].
__global__ void k_split_heads(const uint64_t *base, uint64_t nslabs, uint32_t parts, uint64_t *heads) {
  const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < parts) heads[p] = base[(uint64_t)p * nslabs];
}

// Per slab: each row's place is its part's running offset plus its rank
// among the chunk's rows of that part. SMALL: lane q holds part q's offset
// in a register; otherwise the offsets are in LDS (kWaves * parts, dynamic).
// ROW > 0 (rows of V * ROW cells): every lane loads its rows once, takes
// the key cell from them and stores them. Otherwise, V = 2 or 4: every lane
// copies its own row in V-cell vectors (the column count is a multiple of
// V); V = 1: the wave copies the chunk's cells together, each lane taking
// its cell's row destination from the row's lane.
template <bool SMALL, int V, int ROW>
__global__ __launch_bounds__(kSplitBlock) void k_split_scatter(SplitArgs a) {
  using T = typename VecOf<V>::T;
  extern __shared__ uint64_t lds_scatter[];
  const int32_t *__restrict__ in = a.in;
  int32_t *__restrict__ out = a.out;
  const uint32_t w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const uint64_t slab = (uint64_t)blockIdx.x * kWaves + w;
  if (slab >= a.nslabs) return;
  uint64_t *ow = lds_scatter + (uint64_t)w * a.parts;
  uint64_t off_reg = 0;  // SMALL: the running offset of part `lane`
  if (SMALL) {
    if (lane < a.parts) off_reg = a.base[(uint64_t)lane * a.nslabs + slab];
  } else {
    for (uint32_t p = lane; p < a.parts; p += kWave) ow[p] = a.base[(uint64_t)p * a.nslabs + slab];
    wave_lds_sync();
  }
  const uint64_t below = lanes_below();
  const uint64_t r0 = slab * a.slab, r1 = min(r0 + a.slab, a.rows);
  for (uint64_t cb = r0; cb < r1; cb += kU * kWave) {
    uint32_t key[kU];
    T rowv[ROW > 0 ? kU * ROW : 1];  // ROW > 0: the rows themselves, key cell included
#pragma unroll
    for (int u = 0; u < kU; u++) {
      const uint64_t r = cb + (uint64_t)u * kWave + lane;
      if (ROW > 0) {
        uint32_t k = 0;
#pragma unroll
        for (int j = 0; j < (ROW > 0 ? ROW : 1); j++) {
          rowv[u * ROW + j] = T(0);
          if (r < r1) rowv[u * ROW + j] = ((const T *)(in + r * a.cols))[j];
#pragma unroll
          for (int c = 0; c < V; c++)
            if ((uint32_t)(j * V + c) == a.col) k = (uint32_t)rowv[u * ROW + j][c];
        }
        key[u] = k;
      } else {
        key[u] = r < r1 ? (uint32_t)in[r * a.cols + a.col] : 0u;
      }
    }
    uint64_t mine[kU];
#pragma unroll
    for (int u = 0; u < kU; u++) {
      const uint64_t r = cb + (uint64_t)u * kWave + lane;
      const bool valid = r < r1;
      const uint32_t p = valid ? (uint32_t)fastmod(key[u], a.parts, a.pmagic) : 0u;
      const Groups g = group_parts(p, valid, lane, a.nbits);
      const uint64_t rank = (uint64_t)__builtin_popcountll(g.peers & below);
      if (SMALL) {
        mine[u] = shfl64(off_reg, p) + rank;
        off_reg += (uint64_t)__builtin_popcountll(g.own);
      } else {
        const uint64_t o = valid ? ow[p] : 0;
        mine[u] = o + rank;
        wave_lds_sync();  // every lane has read its part's offset before it moves
        if (valid && rank == 0) ow[p] = o + (uint64_t)__builtin_popcountll(g.peers);
        wave_lds_sync();
      }
    }
    if (ROW > 0) {
#pragma unroll
      for (int u = 0; u < kU; u++) {
        const uint64_t r = cb + (uint64_t)u * kWave + lane;
        if (r < r1) {
#pragma unroll
          for (int j = 0; j < (ROW > 0 ? ROW : 1); j++) ((T *)(out + mine[u] * a.cols))[j] = rowv[u * ROW + j];
        }
      }
    } else if (V > 1) {
      const uint32_t nv = a.cols / V;
      for (uint32_t j = 0; j < nv; j++) {  // kU rows' vectors in flight at once
        T tmp[kU];
#pragma unroll
        for (int u = 0; u < kU; u++) {
          const uint64_t r = cb + (uint64_t)u * kWave + lane;
          tmp[u] = T(0);
          if (r < r1) tmp[u] = ((const T *)(in + r * a.cols))[j];
        }
#pragma unroll
        for (int u = 0; u < kU; u++) {
          const uint64_t r = cb + (uint64_t)u * kWave + lane;
          if (r < r1) ((T *)(out + mine[u] * a.cols))[j] = tmp[u];
        }
      }
    } else {
#pragma unroll
      for (int u = 0; u < kU; u++) {
        const uint64_t c0 = cb + (uint64_t)u * kWave;
        if (c0 >= r1) break;  // uniform
        const uint32_t n = (uint32_t)min<uint64_t>(kWave, r1 - c0);
        const uint32_t cells = n * a.cols;
        const int32_t *src = in + c0 * a.cols;
        for (uint32_t k0 = 0; k0 < cells; k0 += kWave) {  // uniform trip count: shfl64 needs every lane
          const uint32_t k = k0 + lane;
          const uint32_t row = k < cells ? fastdiv32(k, a.cols, a.cmagic) : 0u;
          const uint64_t d = shfl64(mine[u], row);
          if (k < cells) out[d * a.cols + (k - row * a.cols)] = src[k];
        }
      }
    }
  }
}

// Scatter with the workgroup's rows sorted by part in LDS first (at most 64
// parts, rows of C = 1, 2 or 4 cells): workgroup b takes the kWaves slabs
// kWaves*b .. kWaves*b + kWaves - 1 (kStageRows rows). Each wave ranks its
// slab's rows as the direct scatter does, but puts each row at its place in
// the workgroup's part-sorted block in LDS; then the workgroup copies the
// block out, each part's rows to their contiguous destination, so a part's
// rows from the whole workgroup leave as one run of whole cache lines
// rather than a few rows per chunk (partly written lines otherwise make
// the writes grow with the part count).
constexpr uint32_t kStageRows = kWaves * 1024;
template <int C>
__global__ __launch_bounds__(kSplitBlock) void k_split_scatter_staged(SplitArgs a) {
  using T = typename VecOf<C>::T;
  __shared__ T s_rows[kStageRows];
  __shared__ uint8_t s_part[kStageRows];
  __shared__ uint32_t s_pstart[kWave];
  __shared__ uint32_t s_woff[kWaves][kWave];
  __shared__ uint64_t s_gbase[kWave];
  const uint32_t w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
  const uint64_t s0 = (uint64_t)blockIdx.x * kWaves;  // the workgroup's first slab
  const uint64_t row0 = s0 * a.slab;
  const uint32_t nrows = (uint32_t)min<uint64_t>(kStageRows, a.rows - row0);
  const uint32_t nsl = (uint32_t)min<uint64_t>(kWaves, a.nslabs - s0);
  if (w == 0) {
    // lane q: part q's places for the workgroup's slabs, read as one short
    // run of the part-major scan (its end: the next slab's place, or the
    // next part's first place after the last slab)
    uint64_t b[kWaves + 1] = {0, 0, 0, 0, 0};
    if (lane < a.parts) {
      const uint64_t *bp = a.base + (uint64_t)lane * a.nslabs + s0;
#pragma unroll
      for (uint32_t k = 0; k < kWaves; k++)
        if (k < nsl) b[k] = bp[k];
      b[kWaves] = s0 + nsl < a.nslabs ? bp[nsl] : lane + 1 < a.parts ? a.base[(uint64_t)(lane + 1) * a.nslabs] : a.rows;
    }
    const uint64_t c = b[kWaves] - b[0];  // part q's rows in the workgroup
    const uint64_t incl = wave_incl_scan(c);
    s_pstart[lane] = (uint32_t)(incl - c);
    s_gbase[lane] = b[0];
#pragma unroll
    for (uint32_t k = 0; k < kWaves; k++) s_woff[k][lane] = (uint32_t)(b[k] - b[0]);
  }
  __syncthreads();
  const uint64_t slab = s0 + w;
  if (slab < a.nslabs) {
    const uint64_t below = lanes_below();
    // lane q: where the wave's next row of part q goes in LDS
    uint32_t run = 0;
    if (lane < a.parts) run = s_pstart[lane] + s_woff[w][lane];
    const uint64_t r0 = slab * a.slab, r1 = min(r0 + a.slab, a.rows);
    for (uint64_t cb = r0; cb < r1; cb += kU * kWave) {
      T rowv[kU];
#pragma unroll
      for (int u = 0; u < kU; u++) {
        const uint64_t r = cb + (uint64_t)u * kWave + lane;
        rowv[u] = T(0);
        if (r < r1) rowv[u] = ((const T *)a.in)[r];
      }
#pragma unroll
      for (int u = 0; u < kU; u++) {
        const uint64_t r = cb + (uint64_t)u * kWave + lane;
        const bool valid = r < r1;
        uint32_t key = 0;
#pragma unroll
        for (int c = 0; c < C; c++)
          if ((uint32_t)c == a.col) key = (uint32_t)rowv[u][c];
        const uint32_t p = valid ? (uint32_t)fastmod(key, a.parts, a.pmagic) : 0u;
        const Groups g = group_parts(p, valid, lane, a.nbits);
        const uint32_t at = shfl32(run, p) + (uint32_t)__builtin_popcountll(g.peers & below);
        run += (uint32_t)__builtin_popcountll(g.own);
        if (valid) {
          s_rows[at] = rowv[u];
          s_part[at] = (uint8_t)p;
        }
      }
    }
  }
  __syncthreads();
  for (uint32_t j = threadIdx.x; j < nrows; j += kSplitBlock) {
    const uint32_t p = s_part[j];
    ((T *)a.out)[s_gbase[p] + (j - s_pstart[p])] = s_rows[j];
  }
}

// Key loads as V-cell vectors (the row's vector that holds the key cell):
// rows of 2 or 4k cells are read as whole 8- or 16-byte lane accesses.
template <bool SMALL>
void launch_count(const SplitArgs &a, unsigned blocks, size_t lds, hipStream_t s) {
  if (a.cols % 4 == 0) k_split_count<SMALL, 4><<<blocks, kSplitBlock, lds, s>>>(a);
  else if (a.cols % 2 == 0) k_split_count<SMALL, 2><<<blocks, kSplitBlock, lds, s>>>(a);
  else k_split_count<SMALL, 1><<<blocks, kSplitBlock, lds, s>>>(a);
}

template <bool SMALL>
void launch_scatter(const SplitArgs &a, unsigned blocks, size_t lds, hipStream_t s) {
  switch (a.cols) {  // rows of 1, 2, 4 or 8 cells are held in registers between load and store
    case 1: k_split_scatter<SMALL, 1, 1><<<blocks, kSplitBlock, lds, s>>>(a); return;
    case 2: k_split_scatter<SMALL, 2, 1><<<blocks, kSplitBlock, lds, s>>>(a); return;
    case 4: k_split_scatter<SMALL, 4, 1><<<blocks, kSplitBlock, lds, s>>>(a); return;
    case 8: k_split_scatter<SMALL, 4, 2><<<blocks, kSplitBlock, lds, s>>>(a); return;
    default: break;
  }
  if (a.cols % 4 == 0) k_split_scatter<SMALL, 4, 0><<<blocks, kSplitBlock, lds, s>>>(a);
  else if (a.cols % 2 == 0) k_split_scatter<SMALL, 2, 0><<<blocks, kSplitBlock, lds, s>>>(a);
  else k_split_scatter<SMALL, 1, 0><<<blocks, kSplitBlock, lds, s>>>(a);
}

}  // namespace

extern "C" {

int wkg_table_split(wkg_ctx *ctx, const wkg_table *in, int col, int nparts, void *stream, wkg_table **out,
                    uint64_t *part_rows, uint64_t *part_head) {
  if (!ctx || !in || !out || !part_rows) return WKG_EINVAL;
  *out = nullptr;
  if (nparts < 1 || nparts > kMaxParts)
    return ctx->fail(WKG_EINVAL, "split into %d parts (1..%d)", nparts, kMaxParts);
  if (col < 0 || col >= in->cols)
    return ctx->fail(WKG_EINVAL, "split column %d of a %d-column table", col, in->cols);
  if (in->cols > 65536) return ctx->fail(WKG_EINVAL, "split of a %d-column table (at most 65536)", in->cols);
  if (in->rows * (uint64_t)in->cols >= (uint64_t(1) << 62))
    return ctx->fail(WKG_EINVAL, "table too large to split");
  hipStream_t s = ctx->pick(stream);
  wkg_table *t;
  int rc = ctx->new_table(in->rows, in->cols, s, &t);
  if (rc) return rc;
  std::vector<uint64_t> heads((size_t)nparts, 0);
  if (in->rows) {
    SplitArgs a;
    a.in = in->d;
    a.out = t->d;
    a.rows = in->rows;
    a.cols = (uint32_t)in->cols;
    a.col = (uint32_t)col;
    a.parts = (uint32_t)nparts;
    a.pmagic = fastmod_magic((uint64_t)nparts);
    a.cmagic = fastmod_magic((uint64_t)in->cols);
    a.nbits = 0;
    while ((1u << a.nbits) < (uint32_t)nparts) a.nbits++;
    a.slab = std::max<uint64_t>(1024, 16 * (uint64_t)nparts);
    a.nslabs = (in->rows + a.slab - 1) / a.slab;
    const uint64_t ncnt = a.nslabs * (uint64_t)nparts;
    size_t tmp_bytes = 0;
    hipError_t e = rocprim::exclusive_scan(nullptr, tmp_bytes, (const uint32_t *)nullptr, (uint64_t *)nullptr,
                                           (uint64_t)0, (size_t)ncnt, rocprim::plus<uint64_t>(), s);
    if (e != hipSuccess) {
      ctx->drop_table(t);
      return ctx->fail(WKG_EDEVICE, "split scan sizing failed: %s", hipGetErrorString(e));
    }
    auto up = [](uint64_t b) { return (b + 255) & ~uint64_t(255); };
    const uint64_t b_cnt = up(ncnt * 4), b_base = up(ncnt * 8), b_heads = up((uint64_t)nparts * 8);
    char *scr = (char *)ctx->pool->alloc(b_cnt + b_base + b_heads + up(tmp_bytes), s, wkg_ctx::err_slot());
    if (!scr) {
      ctx->drop_table(t);
      return WKG_ENOMEM;
    }
    a.cnt = (uint32_t *)scr;
    uint64_t *base = (uint64_t *)(scr + b_cnt);
    uint64_t *d_heads = (uint64_t *)(scr + b_cnt + b_base);
    void *tmp = scr + b_cnt + b_base + b_heads;
    a.base = base;
    const uint64_t blocks = (a.nslabs + kWaves - 1) / kWaves;
    const bool small = nparts <= kWave;
    {
      ProfScope ps(ctx, s, "split_count", 4.0 * (double)in->rows);
      if (small) launch_count<true>(a, (unsigned)blocks, 0, s);
      else launch_count<false>(a, (unsigned)blocks, (size_t)kWaves * nparts * 4, s);
    }
    e = hipGetLastError();
    if (e == hipSuccess) {
      ProfScope ps(ctx, s, "split_scan", 12.0 * (double)ncnt);
      e = rocprim::exclusive_scan(tmp, tmp_bytes, (const uint32_t *)a.cnt, base, (uint64_t)0, (size_t)ncnt,
                                  rocprim::plus<uint64_t>(), s);
    }
    if (e == hipSuccess) {
      k_split_heads<<<(unsigned)((nparts + 255) / 256), 256, 0, s>>>(base, a.nslabs, (uint32_t)nparts, d_heads);
      e = hipGetLastError();
    }
    if (e == hipSuccess) {
      ProfScope ps(ctx, s, "split_scatter", 8.0 * (double)in->rows * (double)in->cols);
      // (with at most 8 parts the direct scatter's runs are long already,
      // and it measured faster there)
      if (small && nparts > 8 && a.slab == 1024 && (a.cols == 1 || a.cols == 2 || a.cols == 4)) {
        if (a.cols == 1) k_split_scatter_staged<1><<<(unsigned)blocks, kSplitBlock, 0, s>>>(a);
        else if (a.cols == 2) k_split_scatter_staged<2><<<(unsigned)blocks, kSplitBlock, 0, s>>>(a);
        else k_split_scatter_staged<4><<<(unsigned)blocks, kSplitBlock, 0, s>>>(a);
      } else if (small) {
        launch_scatter<true>(a, (unsigned)blocks, 0, s);
      } else {
        launch_scatter<false>(a, (unsigned)blocks, (size_t)kWaves * nparts * 8, s);
      }
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(heads.data(), d_heads, (size_t)nparts * 8, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    ctx->pool->release(scr, s);
    if (e != hipSuccess) {
      ctx->drop_table(t);
      return ctx->fail(WKG_EDEVICE, "split failed: %s", hipGetErrorString(e));
    }
  }
  for (int p = 0; p < nparts; p++) {
    const uint64_t end = p + 1 < nparts ? heads[(size_t)p + 1] : in->rows;
    part_rows[p] = end - heads[(size_t)p];
    if (part_head) part_head[p] = heads[(size_t)p];
  }
  *out = t;
  return WKG_OK;
}

int wkg_table_concat(wkg_ctx *ctx, int n, const wkg_table *const *parts, void *stream, wkg_table **out) {
  if (!ctx || !out || n < 0 || (n && !parts)) return WKG_EINVAL;
  *out = nullptr;
  int cols = -1;
  uint64_t rows = 0;
  for (int i = 0; i < n; i++) {
    if (!parts[i]) return ctx->fail(WKG_EINVAL, "part %d is null", i);
    if (parts[i]->rows == 0) continue;  // an empty reply carries no cells
    if (cols >= 0 && parts[i]->cols != cols)
      return ctx->fail(WKG_EINVAL, "part %d has %d columns, earlier parts %d", i, parts[i]->cols, cols);
    cols = parts[i]->cols;
    rows += parts[i]->rows;
  }
  if (cols < 0) cols = n ? parts[0]->cols : 0;
  hipStream_t s = ctx->pick(stream);
  wkg_table *t;
  int rc = ctx->new_table(rows, cols, s, &t);
  if (rc) return rc;
  uint64_t at = 0;
  hipError_t e = hipSuccess;
  {
    ProfScope ps(ctx, s, "concat", 8.0 * (double)rows * (double)cols);
    for (int i = 0; i < n && e == hipSuccess; i++) {
      if (!parts[i]->rows || !parts[i]->bytes) continue;
      e = hipMemcpyAsync(t->d + at, parts[i]->d, parts[i]->bytes, hipMemcpyDeviceToDevice, s);
      at += parts[i]->rows * (uint64_t)cols;
    }
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    ctx->drop_table(t);
    return ctx->fail(WKG_EDEVICE, "concat failed: %s", hipGetErrorString(e));
  }
  *out = t;
  return WKG_OK;
}

}  // extern "C"
