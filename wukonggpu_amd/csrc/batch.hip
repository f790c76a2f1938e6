// Fused batched steps: many independent pattern steps (different tables,
// predicates, directions and step kinds) run as one probe launch, one tile
// scan launch, one finalize launch and one partition + expand pair. Each
// workgroup finds its step by binary search over the steps' first tiles.
// This is the primitive under batched and served execution; the reference
// serialises queries on the GPU (core/gpu.hpp:53-61) and launches five
// kernels plus two or three synchronisations per step
// (core/rcache.hpp:232-414).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "../../include/wukong_gpu.h"
#include "step_kernels.h"
#include "wkg_internal.h"

using namespace wkg;

namespace {

struct BatchItem {
  StepParams p;    // probe / scan / finalize view (pointers offset to this step)
  ExpandParams x;  // expand view (filled after the totals are known)
};

// last q in [0, n] with base[q] <= t (base ascending, base[n] = total)
__device__ __forceinline__ uint32_t find_item(const uint64_t *base, uint32_t n, uint64_t t) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (base[mid] <= t) lo = mid; else hi = mid - 1;
  }
  return lo;
}

template <bool SORTED, bool FRAMED>
__global__ __launch_bounds__(kBlock, 3) void k_batch_probe(GraphView g, const BatchItem *items,
                                                        const uint64_t *tile_base, uint32_t n) {
  __shared__ uint32_t s_q;
  if (threadIdx.x == 0) s_q = find_item(tile_base, n, blockIdx.x);
  __syncthreads();
  const uint32_t q = s_q;
  probe_tile<kTileRows, SORTED, FRAMED>(g, items[q].p, blockIdx.x - tile_base[q]);
}

__global__ __launch_bounds__(kBlock) void k_batch_scan(const BatchItem *items) {
  scan_step_tiles(items[blockIdx.x].p);
}

// One wave per tile.
__global__ __launch_bounds__(kBlock) void k_batch_finalize(const BatchItem *items, const uint64_t *tile_base,
                                                           uint32_t n) {
  const uint64_t t = blockIdx.x * (uint64_t)(kBlock / kWave) + threadIdx.x / kWave;
  if (t >= tile_base[n]) return;
  const uint32_t q = find_item(tile_base, n, t);
  finalize_tile<kTileRows>(items[q].p, t - tile_base[q]);
}

__global__ void k_batch_partition(const BatchItem *items, const uint64_t *part_base, uint32_t n, uint64_t *part) {
  const uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
  if (k >= part_base[n]) return;
  const uint32_t q = find_item(part_base, n, k);
  const ExpandParams &x = items[q].x;
  part[k] = merge_split<kExpandNV>(x.cexcl, x.R, x.E, k - part_base[q]);
}

template <bool FRAMED>
__global__ __launch_bounds__(kBlock) void k_batch_expand(GraphView g, const BatchItem *items,
                                                         const uint64_t *xtile_base, uint32_t n) {
  __shared__ uint32_t s_q;
  if (threadIdx.x == 0) s_q = find_item(xtile_base, n, blockIdx.x);
  __syncthreads();
  const uint32_t q = s_q;
  expand_tile<kExpandVT, FRAMED>(g, items[q].x, blockIdx.x - xtile_base[q]);
}

}  // namespace

extern "C" int wkg_step_batch(wkg_ctx *ctx, int n, const wkg_step_desc *steps, void *stream, wkg_table **out,
                              uint64_t *expanded) {
  if (!ctx || n < 0 || (n && (!steps || !out))) return WKG_EINVAL;
  if (n == 0) return WKG_OK;
  if (!ctx->loaded) return ctx->fail(WKG_ESTATE, "no graph loaded");
  hipStream_t s = ctx->pick(stream);
  for (int i = 0; i < n; i++) out[i] = nullptr;

  // ---- validate and lay out -----------------------------------------------------
  std::vector<int> pids;
  std::vector<uint64_t> R(n), rowb(n + 1, 0), tileb(n + 1, 0);
  std::vector<int> mode(n), outc(n);
  for (int i = 0; i < n; i++) {
    const wkg_step_desc &d = steps[i];
    int rc = ctx->check_pattern(d.pid, d.dir);
    if (rc) return ctx->fail(rc, "step %d: %s", i, wkg_ctx::err_slot().c_str());
    switch (d.kind) {
      case WKG_STEP_I2U:
        if (!is_tpid(d.pid)) return ctx->fail(WKG_EINVAL, "step %d: index id %d is not a predicate/type id", i, d.pid);
        R[i] = 1, mode[i] = kModeK2U, outc[i] = 1;
        break;
      case WKG_STEP_C2U:
        if (d.vid < 0 || (uint64_t)d.vid > kMaxVid) return ctx->fail(WKG_EINVAL, "step %d: constant out of range", i);
        R[i] = 1, mode[i] = kModeK2U, outc[i] = 1;
        break;
      case WKG_STEP_K2U:
      case WKG_STEP_K2K:
      case WKG_STEP_K2C: {
        if (!d.in) return ctx->fail(WKG_EINVAL, "step %d: null input table", i);
        const int c = d.in->cols;
        if (c < 1 || d.start_col < 0 || d.start_col >= c)
          return ctx->fail(WKG_EINVAL, "step %d: start column %d outside the %d-column table", i, d.start_col, c);
        if (d.kind == WKG_STEP_K2K && (d.end_col < 0 || d.end_col >= c))
          return ctx->fail(WKG_EINVAL, "step %d: end column %d outside the %d-column table", i, d.end_col, c);
        mode[i] = d.kind == WKG_STEP_K2U ? kModeK2U : d.kind == WKG_STEP_K2K ? kModeK2K : kModeK2C;
        outc[i] = d.kind == WKG_STEP_K2U ? c + 1 : c;
        if (outc[i] > kBlock) return ctx->fail(WKG_ERANGE, "step %d: tables are limited to %d columns", i, kBlock);
        R[i] = d.in->rows;
        if (R[i] >= (1ull << 32)) return ctx->fail(WKG_ERANGE, "step %d: tables are limited to 2^32 - 1 rows", i);
        break;
      }
      default:
        return ctx->fail(WKG_EINVAL, "step %d: unknown kind %d", i, d.kind);
    }
    rowb[i + 1] = rowb[i] + R[i];
    tileb[i + 1] = tileb[i] + (R[i] + kTileRows - 1) / kTileRows;
    if (std::find(pids.begin(), pids.end(), d.pid) == pids.end()) pids.push_back(d.pid);
  }
  for (int p : pids)
    if (int rc = ctx->ensure_pred(p, s, &pids)) return rc;

  const uint64_t NR = rowb[n], NT = tileb[n];
  // scratch: totals[4n] | tile_base[n+1] | tile_sum, tile_nz, tile_ex [NT each] |
  //          hits[NT * kTileRows] (16 B) | cinfo[NR] (8 B) | cexcl[NR] | items
  const uint64_t w_items = (sizeof(BatchItem) * n + 7) / 8;
  const uint64_t hit_words = 2 * NT * kTileRows;
  const uint64_t words = 4 * (uint64_t)n + (n + 1) + 3 * NT + hit_words + 2 * NR + w_items;
  uint64_t *scr = (uint64_t *)ctx->pool->alloc(words * 8, s, wkg_ctx::err_slot());
  if (!scr) return WKG_ENOMEM;
  uint64_t *totals = scr;
  uint64_t *d_tileb = totals + 4 * n;
  uint64_t *tsum = d_tileb + (n + 1);
  uint64_t *tnz = tsum + NT;
  uint64_t *tex = tnz + NT;
  uint4 *hits = (uint4 *)(tex + NT);
  uint2 *cinfo = (uint2 *)((uint64_t *)hits + hit_words);
  uint64_t *cexcl = (uint64_t *)(cinfo + NR);
  BatchItem *d_items = (BatchItem *)(cexcl + NR);
  auto bail = [&](int rc) {
    ctx->pool->release(scr, s);
    for (int i = 0; i < n; i++) {
      ctx->drop_table(out[i]);
      out[i] = nullptr;
    }
    return rc;
  };

  std::vector<BatchItem> items(n);
  for (int i = 0; i < n; i++) {
    const wkg_step_desc &d = steps[i];
    StepParams &p = items[i].p;
    p = StepParams{};
    p.pid = (uint32_t)d.pid;
    p.dir = (uint32_t)d.dir;
    p.mode = mode[i];
    p.R = R[i];
    p.hits = hits + tileb[i] * kTileRows;
    p.cinfo = cinfo + rowb[i];
    p.cexcl = cexcl + rowb[i];
    p.tile_sum = tsum + tileb[i];
    p.tile_nz = tnz + tileb[i];
    p.tile_ex = tex + tileb[i];
    p.totals = totals + 4 * i;
    p.ntiles = tileb[i + 1] - tileb[i];
    if (d.kind == WKG_STEP_I2U || d.kind == WKG_STEP_C2U) {
      p.key = d.kind == WKG_STEP_I2U ? make_key(0, (uint64_t)d.pid, (uint64_t)d.dir)
                                     : make_key((uint64_t)d.vid, (uint64_t)d.pid, (uint64_t)d.dir);
      p.T = nullptr;
      p.c = 0;
    } else {
      p.key = 0;
      p.T = d.in->d;
      p.c = (uint32_t)d.in->cols;
      p.sc = (uint32_t)d.start_col;
      p.ec = (uint32_t)(d.kind == WKG_STEP_K2K ? d.end_col : 0);
      p.target = (uint32_t)d.vid;
      p.target_ok = (d.kind == WKG_STEP_K2C && d.vid >= 0 && d.vid <= 0xffffffffLL) ? 1u : 0u;
      if (mode[i] != kModeK2U) {  // filters write into a table of capacity R x c
        if (int rc = ctx->new_table(R[i], outc[i], s, &out[i])) return bail(rc);
        p.O = out[i]->d;
      }
    }
  }
  hipError_t e = hipMemcpyAsync(d_tileb, tileb.data(), (n + 1) * 8, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(d_items, items.data(), sizeof(BatchItem) * n, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return bail(ctx->fail(WKG_EDEVICE, "descriptor upload failed: %s", hipGetErrorString(e)));
  const GraphView g = ctx->view();
  if (NT) {
    double pb = 0;
    for (int i = 0; i < n; i++) pb += (double)R[i] * (4.0 + (mode[i] == kModeK2K ? 4.0 : 0.0) + 128.0);
    ProfScope ps(ctx, s, "batch_probe", pb);
    if (ctx->framed) {
      if (ctx->sorted) k_batch_probe<true, true><<<(unsigned)NT, kBlock, 0, s>>>(g, d_items, d_tileb, n);
      else k_batch_probe<false, true><<<(unsigned)NT, kBlock, 0, s>>>(g, d_items, d_tileb, n);
    } else {
      if (ctx->sorted) k_batch_probe<true, false><<<(unsigned)NT, kBlock, 0, s>>>(g, d_items, d_tileb, n);
      else k_batch_probe<false, false><<<(unsigned)NT, kBlock, 0, s>>>(g, d_items, d_tileb, n);
    }
  }
  {
    ProfScope ps(ctx, s, "batch_scan", 24.0 * (double)NT);
    k_batch_scan<<<(unsigned)n, kBlock, 0, s>>>(d_items);
  }
  if (NT) {
    ProfScope ps(ctx, s, "batch_finalize", 0.0);
    k_batch_finalize<<<(unsigned)((NT + kBlock / kWave - 1) / (kBlock / kWave)), kBlock, 0, s>>>(d_items, d_tileb, n);
  }
  if ((e = hipGetLastError()) != hipSuccess)
    return bail(ctx->fail(WKG_EDEVICE, "batched probe launch failed: %s", hipGetErrorString(e)));
  std::vector<uint64_t> tot(4 * (size_t)n);
  if (int rc = ctx->read_words(totals, tot.size(), s, tot.data())) return bail(rc);

  // ---- outputs and the expand pass --------------------------------------------------
  std::vector<uint64_t> partb(n + 1, 0), xtileb(n + 1, 0);
  double fin_bytes = 0;
  for (int i = 0; i < n; i++) {
    const uint64_t E = tot[4 * i], nz = tot[4 * i + 2];
    uint64_t nx = 0;
    if (mode[i] == kModeK2U) {
      fin_bytes += 32.0 * nz;
      if (int rc = ctx->new_table(E, outc[i], s, &out[i])) return bail(rc);
      if (E) nx = (nz + E + kExpandNV - 1) / kExpandNV;
      ExpandParams &x = items[i].x;
      x.pid = items[i].p.pid;
      x.T = items[i].p.T;
      x.c = items[i].p.c;
      x.R = nz;
      x.cinfo = items[i].p.cinfo;
      x.cexcl = items[i].p.cexcl;
      x.E = E;
      x.O = out[i]->d;
      if (expanded) expanded[i] = E;
    } else {
      fin_bytes += 16.0 * nz + 8.0 * nz * outc[i];
      out[i]->rows = E;
      out[i]->bytes = E * (uint64_t)outc[i] * 4;
      if (expanded) expanded[i] = tot[4 * i + 1];
    }
    partb[i + 1] = partb[i] + (nx ? nx + 1 : 0);
    xtileb[i + 1] = xtileb[i] + nx;
  }
  ctx->prof_add_bytes("batch_finalize", fin_bytes);
  const uint64_t NX = xtileb[n], NP = partb[n];
  if (NX) {
    uint64_t *xs = (uint64_t *)ctx->pool->alloc((NP + 2 * (n + 1)) * 8, s, wkg_ctx::err_slot());
    if (!xs) return bail(WKG_ENOMEM);
    uint64_t *d_part = xs, *d_partb = xs + NP, *d_xtileb = d_partb + (n + 1);
    for (int i = 0; i < n; i++) items[i].x.part = d_part + partb[i];
    e = hipMemcpyAsync(d_items, items.data(), sizeof(BatchItem) * n, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(d_partb, partb.data(), (n + 1) * 8, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(d_xtileb, xtileb.data(), (n + 1) * 8, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
      {
        ProfScope ps(ctx, s, "batch_partition", 8.0 * (double)NP);
        k_batch_partition<<<(unsigned)((NP + kBlock - 1) / kBlock), kBlock, 0, s>>>(d_items, d_partb, n, d_part);
      }
      double xb = 0;
      for (int i = 0; i < n; i++) {
        const ExpandParams &x = items[i].x;
        if (mode[i] == kModeK2U && x.E)
          xb += 4.0 * x.E + 4.0 * x.E * (x.c + 1) + 4.0 * x.R * x.c + 16.0 * x.R;
      }
      ProfScope ps(ctx, s, "batch_expand", xb);
      if (ctx->framed) k_batch_expand<true><<<(unsigned)NX, kBlock, 0, s>>>(g, d_items, d_xtileb, n);
      else k_batch_expand<false><<<(unsigned)NX, kBlock, 0, s>>>(g, d_items, d_xtileb, n);
      e = hipGetLastError();
    }
    // the host descriptor vectors are read by the async copies above
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    ctx->pool->release(xs, s);
    if (e != hipSuccess) return bail(ctx->fail(WKG_EDEVICE, "batched expand failed: %s", hipGetErrorString(e)));
  }
  ctx->pool->release(scr, s);
  return WKG_OK;
}
