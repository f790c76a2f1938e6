// libwkg_hip.so: device context, HBM pool, gstore residency and the C ABI of
// the pattern-step engine (include/wukong_gpu.h).
//
// The reference's RCache (core/rcache.hpp:114-221) allocates one fixed set of
// 20M-element buffers and one device gcache, and every handler synchronises
// two or three times. Here tables and step scratch come from a stream-aware
// caching pool under an explicit HBM budget, sizes are 64-bit, and a step
// synchronises once (to learn its output size).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/wukong_gpu.h"
#include "step_kernels.h"
#include "wkg_internal.h"

using namespace wkg;

// ---------------------------------------------------------------------------
// error helpers

int wkg_ctx::fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  err_slot() = buf;
  return code;
}

#define HIPCHK(ctx, call)                                                                    \
  do {                                                                                       \
    hipError_t e_ = (call);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      return (ctx)->fail(WKG_EDEVICE, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                         __FILE__, __LINE__);                                                \
  } while (0)

#define LAUNCHCHK(ctx)                                                                          \
  do {                                                                                          \
    hipError_t e_ = hipGetLastError();                                                          \
    if (e_ != hipSuccess)                                                                       \
      return (ctx)->fail(WKG_EDEVICE, "kernel launch failed: %s (%s:%d)", hipGetErrorString(e_), \
                         __FILE__, __LINE__);                                                   \
  } while (0)

// ---------------------------------------------------------------------------
// pool

namespace {
size_t round_size(size_t n) {
  if (n <= 4096) return 4096;
  if (n < (size_t(2) << 20)) {
    size_t p = 4096;
    while (p < n) p <<= 1;
    return p;
  }
  const size_t g = size_t(2) << 20;
  return (n + g - 1) / g * g;
}
}  // namespace

void *Pool::alloc(size_t bytes, hipStream_t s, std::string &err) {
  if (BumpArena *b = thread_arena()) {
    const uint64_t sz = ((uint64_t)(bytes ? bytes : 1) + 255) & ~uint64_t(255);
    if (b->used + sz <= b->cap) {
      void *p = b->base + b->used;
      b->used += sz;
      return p;
    }  // (a query bigger than the region takes pool memory as usual)
  }
  std::lock_guard<std::mutex> lk(mu);
  const size_t sz = round_size(bytes ? bytes : 1);
  // reuse a cached block at most a quarter larger: a small request must not
  // pin a big block, which counts in full against the budget
  for (auto it = cached.lower_bound(sz); it != cached.end() && it->first <= sz + sz / 4; ++it) {
    if (it->second.stream != s) continue;
    void *p = it->second.ptr;
    live[p] = it->first;
    live_bytes += it->first;
    cached_bytes -= it->first;
    cached.erase(it);
    return p;
  }
  if (live_bytes + sz > budget) {
    err = "HBM table budget exhausted: need " + std::to_string(sz) + " bytes, " +
          std::to_string(budget - std::min<uint64_t>(budget, live_bytes)) + " of " + std::to_string(budget) +
          " left";
    return nullptr;
  }
  if (live_bytes + cached_bytes + sz > budget) trim_locked(sz);
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, sz);
  if (e != hipSuccess) {
    trim_locked(0);
    e = hipMalloc(&p, sz);
  }
  if (e != hipSuccess) {
    (void)hipGetLastError();
    err = std::string("hipMalloc failed: ") + hipGetErrorString(e);
    return nullptr;
  }
  live[p] = sz;
  live_bytes += sz;
  mallocs++;
  return p;
}

void Pool::release(void *p, hipStream_t s) {
  if (!p) return;
  if (BumpArena *b = thread_arena())
    if ((char *)p >= b->base && (char *)p < b->base + b->cap) return;
  std::lock_guard<std::mutex> lk(mu);
  auto it = live.find(p);
  if (it == live.end()) return;
  const size_t sz = it->second;
  live.erase(it);
  live_bytes -= sz;
  cached.emplace(sz, Cached{p, s});
  cached_bytes += sz;
}

// Frees cached blocks, largest first, until `need` more bytes fit the
// budget (need 0: all of them). hipFree waits for the device, so the cache
// is only cut as far as needed.
void Pool::trim_locked(uint64_t need) {
  if (cached.empty()) return;
  trims++;
  (void)hipDeviceSynchronize();
  while (!cached.empty() && (need == 0 || live_bytes + cached_bytes + need > budget)) {
    auto it = std::prev(cached.end());
    (void)hipFree(it->second.ptr);
    cached_bytes -= it->first;
    cached.erase(it);
  }
}

void Pool::stats(uint64_t *live, uint64_t *cached_b, uint64_t *nmalloc, uint64_t *ntrim) {
  std::lock_guard<std::mutex> lk(mu);
  if (live) *live = live_bytes;
  if (cached_b) *cached_b = cached_bytes;
  if (nmalloc) *nmalloc = mallocs;
  if (ntrim) *ntrim = trims;
}

void Pool::trim() {
  std::lock_guard<std::mutex> lk(mu);
  trim_locked(0);
}

Pool::~Pool() {
  for (auto &kv : cached) (void)hipFree(kv.second.ptr);
  for (auto &kv : live) (void)hipFree(kv.first);
}

// ---------------------------------------------------------------------------
// context internals

GraphView wkg_ctx::view() const {
  GraphView g;
  g.slots = d_slots;
  g.edges = d_edges;
  g.meta = d_meta;
  g.vframes = d_vframes;
  g.eframes = d_eframes;
  g.vshift = vshift;
  g.eshift = eshift;
  g.num_metas = num_metas;
  return g;
}

// The launch stream of an entry point; also makes the context's device
// current on the calling thread (a context may be driven from several).
hipStream_t wkg_ctx::pick(void *s) const {
  (void)hipSetDevice(device);
  if (s) return (hipStream_t)s;
  const ThreadCtx *t = tctx();
  return t ? t->stream : stream;
}

int wkg_ctx::new_table(uint64_t rows, int cols, hipStream_t s, wkg_table **out) {
  wkg_table *t = new (std::nothrow) wkg_table;
  if (!t) return fail(WKG_ENOMEM, "out of host memory");
  t->rows = rows;
  t->cols = cols;
  t->stream = s;
  t->bytes = rows * (uint64_t)cols * sizeof(int32_t);
  t->d = nullptr;
  if (t->bytes) {
    t->d = (int32_t *)pool->alloc(t->bytes, s, err_slot());
    if (!t->d) {
      delete t;
      return WKG_ENOMEM;
    }
  }
  *out = t;
  return WKG_OK;
}

// The caller's handle goes away: the table's lineage is released at once (it
// only serves steps that read this table), its cells once no child table's
// lineage reads them any more.
void wkg_ctx::drop_table(wkg_table *t) {
  if (!t) return;
  if (t->lin_parent) {
    pool->release(t->lin_alloc, t->stream);
    wkg_table *par = t->lin_parent;
    t->lin_parent = nullptr;
    t->lin_alloc = nullptr;
    unref_table(par);
  }
  if (t->glin_parent) {
    pool->release(t->glin_alloc, t->stream);
    wkg_table *gp = t->glin_parent;
    t->glin_parent = nullptr;
    t->glin_alloc = nullptr;
    unref_table(gp);
  }
  unref_table(t);
}

void wkg_ctx::unref_table(wkg_table *t) {
  if (--t->refs > 0) return;
  if (t->owned) pool->release(t->d, t->stream);
  wkg_table *owner = t->slice_of;
  delete t;
  if (owner) unref_table(owner);
}

// The host learns a step's counts by polling the flag that a kernel queued
// behind the step's work sets in pinned host memory: no copy command and no
// stream synchronisation, whose wake-up latency dominated small steps. After
// 20 ms of polling it waits on the stream instead (which also reports a
// failed kernel whose signal never comes).
int wkg_ctx::wait_signal(const Signal &sg, hipStream_t s) {
  const uint64_t *flag = rb_host() + kRbFlag;
  const auto t0 = std::chrono::steady_clock::now();
  for (uint32_t spins = 0; __atomic_load_n(flag, __ATOMIC_ACQUIRE) != sg.seq; spins++) {
    if ((spins & 1023) == 1023 && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(20)) {
      HIPCHK(this, hipStreamSynchronize(s));
      if (__atomic_load_n(flag, __ATOMIC_ACQUIRE) != sg.seq)
        return fail(WKG_EDEVICE, "read-back signal %llu missing after the stream completed",
                    (unsigned long long)sg.seq);
      break;
    }
  }
  return WKG_OK;
}

// n device words to the host: through the pinned staging area (a one-block
// copy kernel, then the polled flag) when they fit, else a copy command.
int wkg_ctx::read_words(const uint64_t *d, uint64_t n, hipStream_t s, uint64_t *host) {
  if (!n) return WKG_OK;
  if (n * 8 > kStageBytes) {
    HIPCHK(this, hipMemcpyAsync(host, d, n * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(this, hipStreamSynchronize(s));
    return WKG_OK;
  }
  const Signal sg = next_signal();
  k_stage<<<1, kBlock, 0, s>>>((const int32_t *)d, 2 * n, (int32_t *)((char *)rb_dev() + kStageOff), sg);
  LAUNCHCHK(this);
  if (int rc = wait_signal(sg, s)) return rc;
  memcpy(host, (const char *)rb_host() + kStageOff, n * 8);
  return WKG_OK;
}

int wkg_ctx::check_pattern(int pid, int dir) {
  if (!loaded) return fail(WKG_ESTATE, "no graph loaded");
  if (pid < 1 || (uint32_t)pid >= num_metas) return fail(WKG_EINVAL, "predicate id %d outside [1, %u)", pid, num_metas);
  if (dir != 0 && dir != 1) return fail(WKG_EINVAL, "direction %d is not IN(0) or OUT(1)", dir);
  return WKG_OK;
}

// ---------------------------------------------------------------------------
// step implementations (plain path)

int wkg_ctx::step_single_list(uint64_t key, int pid, hipStream_t s, wkg_table **out, uint64_t *expanded) {
  int rc = ensure_pred(pid, s);
  if (rc) return rc;
  // a list of at most kListCap entries is copied by the probing launch
  wkg_table *t;
  if ((rc = new_table(kListCap, 1, s, &t))) return rc;
  GraphView g = view();
  const Signal sg = next_signal();
  {
    ProfScope ps(this, s, "probe_one", 128.0 + 16.0);
    if (framed) k_probe_copy<true><<<1, kBlock, 0, s>>>(g, key, (uint32_t)pid, kListCap, t->d, sg);
    else k_probe_copy<false><<<1, kBlock, 0, s>>>(g, key, (uint32_t)pid, kListCap, t->d, sg);
  }
  hipError_t le = hipGetLastError();
  rc = le == hipSuccess ? wait_signal(sg, s) : fail(WKG_EDEVICE, "probe launch failed: %s", hipGetErrorString(le));
  if (rc) {
    drop_table(t);
    return rc;
  }
  const uint64_t deg = rb_host()[0], off = rb_host()[1];
  if (deg <= kListCap) {
    t->rows = deg;
    t->bytes = deg * 4;
  } else {
    drop_table(t);
    if ((rc = new_table(deg, 1, s, &t))) return rc;
    const uint64_t blocks = std::min<uint64_t>((deg + kBlock - 1) / kBlock, 4096);
    {
      ProfScope ps(this, s, "copy_list", 8.0 * (double)deg);
      if (framed) k_copy_list<true><<<blocks, kBlock, 0, s>>>(g, (uint32_t)pid, off, deg, t->d);
      else k_copy_list<false><<<blocks, kBlock, 0, s>>>(g, (uint32_t)pid, off, deg, t->d);
    }
    LAUNCHCHK(this);
  }
  if (expanded) *expanded = deg;
  *out = t;
  return WKG_OK;
}

// Probe tile rows for an R-row input: the largest tile that still gives
// every CU at least four tiles, down to one 64-row round per wave. Small
// inputs otherwise run on a few blocks whose rounds' bucket walks are one
// long chain of load latencies each.
static int probe_tile_rows(uint64_t R, int ncu) {
  for (int tr = kTileRows; tr > kBlock; tr /= 2)
    if (R >= (uint64_t)tr * 4 * (uint64_t)ncu) return tr;
  return kBlock;
}

// The per-step path splits its probe, finalize and expand passes into
// launches of thread_launch_rows() input rows (expand: that / 64 tiles of
// 512 items): the GPU runs another queue's kernels only between this
// queue's kernels, so one launch over a huge table would hold up every
// concurrent batch (a served light query) for its whole length. The default,
// 2^26 rows, is ~1.4 ms of probing; the served path's heavy worker uses 2^22.
template <int TR>
static void launch_probe_tr(bool framed, bool sorted, const GraphView &g, const StepParams &p0, hipStream_t s) {
  const uint64_t chunk = std::max<uint64_t>(thread_launch_rows() / TR, 4);
  StepParams p = p0;
  for (uint64_t t0 = 0; t0 < p0.ntiles; t0 += chunk) {
    p.tile0 = t0;
    const unsigned grid = (unsigned)std::min<uint64_t>(chunk, p0.ntiles - t0);
    if (framed) {
      if (sorted) k_probe<TR, true, true><<<grid, kBlock, 0, s>>>(g, p);
      else k_probe<TR, false, true><<<grid, kBlock, 0, s>>>(g, p);
    } else {
      if (sorted) k_probe<TR, true, false><<<grid, kBlock, 0, s>>>(g, p);
      else k_probe<TR, false, false><<<grid, kBlock, 0, s>>>(g, p);
    }
  }
}

static void launch_probe(int tr, bool framed, bool sorted, const GraphView &g, const StepParams &p, hipStream_t s) {
  switch (tr) {
    case 256: launch_probe_tr<256>(framed, sorted, g, p, s); break;
    case 512: launch_probe_tr<512>(framed, sorted, g, p, s); break;
    case 1024: launch_probe_tr<1024>(framed, sorted, g, p, s); break;
    default: launch_probe_tr<2048>(framed, sorted, g, p, s); break;
  }
}

static void launch_finalize(int tr, const StepParams &p0, hipStream_t s) {
  const uint64_t chunk = std::max<uint64_t>(thread_launch_rows() / (uint64_t)tr, 4) & ~uint64_t(3);  // tiles, whole blocks of 4 waves
  StepParams p = p0;
  for (uint64_t t0 = 0; t0 < p0.ntiles; t0 += chunk) {
    p.tile0 = t0;
    const uint64_t n = std::min<uint64_t>(chunk, p0.ntiles - t0);
    const unsigned grid = (unsigned)((n + kBlock / kWave - 1) / (kBlock / kWave));
    switch (tr) {
      case 256: k_finalize<256><<<grid, kBlock, 0, s>>>(p); break;
      case 512: k_finalize<512><<<grid, kBlock, 0, s>>>(p); break;
      case 1024: k_finalize<1024><<<grid, kBlock, 0, s>>>(p); break;
      default: k_finalize<2048><<<grid, kBlock, 0, s>>>(p); break;
    }
  }
}

// Scratch words of one step's probe / scan / finalize passes over R rows in
// TR-row tiles: totals[4] | tile_sum | tile_nz | tile_ex | chunk sums x 3 |
// hits[ntiles * TR] (16 B each).
static uint64_t pass_scratch_words(uint64_t R, int TR) {
  const uint64_t ntiles = (R + TR - 1) / TR;
  const uint64_t nchunks = (ntiles + kScanChunk - 1) / kScanChunk;
  return 4 + 3 * ntiles + 3 * nchunks + 2 * ntiles * TR;
}

// Scratch and parameters of one step's probe / scan / finalize passes over
// TR-row tiles, with `extra` more scratch words after them (at p's end: the
// returned pointer is pp.scr + pp's own words). With ext_scr (and, for
// known_to_unknown, ext_lin of 16 B per row) the scratch is the caller's and
// nothing is allocated.
static int pass_setup(wkg_ctx *ctx, int mode, const wkg_table *in, int sc, int pid, int dir, int ec, int64_t cval,
                      int32_t *O, hipStream_t s, int TR, uint64_t extra, ProbePass &pp, StepParams &p,
                      uint64_t *ext_scr = nullptr, void *ext_lin = nullptr) {
  const uint64_t R = in->rows;
  const int c = in->cols;
  const uint64_t ntiles = (R + TR - 1) / TR;
  const uint64_t nchunks = (ntiles + kScanChunk - 1) / kScanChunk;
  if (ext_scr) {
    pp.external = true;
    pp.scr = ext_scr;
  } else {
    pp.scr = (uint64_t *)ctx->pool->alloc((pass_scratch_words(R, TR) + extra) * 8, s, wkg_ctx::err_slot());
    if (!pp.scr) return WKG_ENOMEM;
  }
  if (mode == kModeK2U) {  // cinfo[R] (8 B) | cexcl[R]: may outlive the step as the output's lineage
    pp.lin = ext_scr ? ext_lin : ctx->pool->alloc(R * 16, s, wkg_ctx::err_slot());
    if (!pp.lin) {
      ctx->release_pass(pp, s);
      return WKG_ENOMEM;
    }
    pp.cinfo = (uint2 *)pp.lin;
    pp.cexcl = (uint64_t *)((char *)pp.lin + R * 8);
  }
  uint64_t *scr = pp.scr;
  uint64_t *totals = scr, *tile_sum = scr + 4, *tile_nz = tile_sum + ntiles, *tile_ex = tile_nz + ntiles;
  uint64_t *chunks = tile_ex + ntiles;
  p = StepParams{};
  p.T = in->d;
  p.R = R;
  p.c = (uint32_t)c;
  p.sc = (uint32_t)sc;
  p.ec = (uint32_t)(ec < 0 ? 0 : ec);
  p.pid = (uint32_t)pid;
  p.dir = (uint32_t)dir;
  p.target = (uint32_t)cval;
  p.target_ok = (cval >= 0 && cval <= 0xffffffffLL) ? 1u : 0u;
  p.mode = mode;
  p.key = 0;
  p.hits = (uint4 *)(chunks + 3 * nchunks);
  p.cinfo = pp.cinfo;
  p.cexcl = pp.cexcl;
  p.O = O;
  p.tile_sum = tile_sum;
  p.tile_nz = tile_nz;
  p.tile_ex = tile_ex;
  p.totals = totals;
  p.ntiles = ntiles;
  p.chunk_sum = chunks;
  p.chunk_nz = chunks + nchunks;
  p.chunk_ex = chunks + 2 * nchunks;
  return WKG_OK;
}

// Probe, tile scan and compaction of one step over `in` (the first half of
// every known-start step): for K2U the compacted rows with edges (cinfo) and
// their output offsets (cexcl) in pp.lin; the filters compact into O.
int wkg_ctx::probe_pass(int mode, const wkg_table *in, int sc, int pid, int dir, int ec, int64_t cval, int32_t *O,
                        hipStream_t s, ProbePass &pp, int wcol, uint64_t *ext_scr, void *ext_lin, int tile_rows) {
  const uint64_t R = in->rows;
  const int c = in->cols;
  const int TR = tile_rows ? tile_rows : probe_tile ? probe_tile : probe_tile_rows(R, ncu);
  StepParams p{};
  int rc = pass_setup(this, mode, in, sc, pid, dir, ec, cval, O, s, TR, 0, pp, p, ext_scr, ext_lin);
  if (rc) return rc;
  if (wcol >= 0) {
    p.weighted = 1;
    p.wc = (uint32_t)wcol;
  }
  pp.hits = p.hits;
  pp.tile_nz = p.tile_nz;
  pp.chunk_nz = p.chunk_nz;
  pp.ntiles = p.ntiles;
  pp.tr = TR;
  const uint64_t ntiles = p.ntiles, nchunks = (ntiles + kScanChunk - 1) / kScanChunk;
  uint64_t *totals = p.totals;
  const GraphView g = view();
  {
    // the input table (4 R c: its rows are read whole, as cache lines); K2U
    // adds one 128-byte bucket per key probed and a 16-byte hit per row with
    // edges once the counts are known; the filters are charged one bucket per row
    const double b = 4.0 * (double)R * c + (mode == kModeK2U ? 0.0 : 128.0 * (double)R);
    ProfScope ps(this, s, mode == kModeK2U ? "probe" : "probe_filter", b);
    launch_probe(TR, framed, sorted || mode == kModeK2U, g, p, s);
  }
  // the scan publishes the totals; the host goes on while the finalize runs
  p.sg = next_signal();
  {
    ProfScope ps(this, s, "scan_tiles", 48.0 * (double)ntiles);
    k_scan_chunks<<<(unsigned)nchunks, kBlock, 0, s>>>(p);
    if (nchunks > 1) k_scan_top<<<1, kBlock, 0, s>>>(p, nchunks);
  }
  if (mode == kModeK2U || O) {  // (a filter without O leaves its kept rows in pp.hits)
    ProfScope ps(this, s, mode == kModeK2U ? "finalize" : "compact", 0.0);
    launch_finalize(TR, p, s);
  }
  hipError_t le = hipGetLastError();
  rc = le == hipSuccess ? wait_signal(p.sg, s) : WKG_OK;
  if (le != hipSuccess || rc) {
    release_pass(pp, s);
    return le != hipSuccess ? fail(WKG_EDEVICE, "probe launch failed: %s", hipGetErrorString(le)) : rc;
  }
  memcpy(pp.cnt, rb_host(), 3 * sizeof(uint64_t));
  if (mode == kModeK2U) {
    prof_add_bytes("finalize", 32.0 * pp.cnt[2]);
    prof_add_bytes("probe", 128.0 * (double)pp.cnt[1] + 16.0 * (double)pp.cnt[2]);
    prof_add_bytes("probe_keys", (double)pp.cnt[1]);  // counts, not bytes: keys probed (run leaders)
    prof_add_bytes("probe_rows", (double)R);           // and rows looked up
  } else if (O) {
    prof_add_bytes("compact", 16.0 * pp.cnt[2] + 8.0 * pp.cnt[2] * c);
  }
  return WKG_OK;
}

void wkg_ctx::release_pass(ProbePass &pp, hipStream_t s) {
  if (!pp.external) {
    if (pp.scr) pool->release(pp.scr, s);
    if (pp.lin) pool->release(pp.lin, s);
  }
  pp = ProbePass{};
}

// Merge-path expand of a known_to_unknown step: output row o of compacted
// row i is (input row cinfo[i].x, edge o - cexcl[i]).
int wkg_ctx::expand_pass(const wkg_table *in, int pid, const uint2 *cinfo, const uint64_t *cexcl, uint64_t nz,
                         uint64_t E, hipStream_t s, wkg_table *t, const char *tag, uint64_t *ext_part) {
  if (!E) return WKG_OK;
  const int c = in->cols;
  const uint64_t nx = (nz + E + kExpandNV - 1) / kExpandNV;
  // split points: (nx + 1) words, the caller's (ext_part) or the pool's
  uint64_t *part = ext_part ? ext_part : (uint64_t *)pool->alloc((nx + 1) * 8, s, err_slot());
  if (!part) return WKG_ENOMEM;
  {
    ProfScope ps(this, s, "partition", 8.0 * (double)(nx + 1));
    k_partition<<<(unsigned)((nx + 1 + kBlock - 1) / kBlock), kBlock, 0, s>>>(cexcl, nz, E, nx, part);
  }
  ExpandParams x{};
  x.pid = (uint32_t)pid;
  x.T = in->d;
  x.c = (uint32_t)c;
  x.R = nz;
  x.cinfo = cinfo;
  x.cexcl = cexcl;
  x.part = part;
  x.E = E;
  x.O = t->d;
  {
    // edges gathered (4E) + output rows written (4E(c+1)) + the input rows
    // with edges and their offsets read once (4 nz c + 16 nz)
    const double b = 4.0 * E + 4.0 * E * (c + 1) + 4.0 * nz * c + 16.0 * nz;
    ProfScope ps(this, s, tag, b);
    const GraphView g = view();
    // split launches, as the probe's
    const uint64_t per = std::max<uint64_t>(thread_launch_rows() / 64, 1);
    for (uint64_t t0 = 0; t0 < nx; t0 += per) {
      x.tile0 = t0;
      const unsigned grid = (unsigned)std::min<uint64_t>(per, nx - t0);
      if (framed) k_expand<true><<<grid, kBlock, 0, s>>>(g, x);
      else k_expand<false><<<grid, kBlock, 0, s>>>(g, x);
    }
  }
  hipError_t e2 = hipGetLastError();
  if (!ext_part) pool->release(part, s);
  if (e2 != hipSuccess) return fail(WKG_EDEVICE, "expand launch failed: %s", hipGetErrorString(e2));
  return WKG_OK;
}

// Keeps (cinfo, cexcl) of a known_to_unknown step with its output `t` when
// the step fanned out (at least 4 output rows per input row with edges) and
// both the input and the arrays are small next to the budget: a later
// known_to_unknown step on one of the input's columns then probes once per
// compacted input row (step_k2u_lineage).
void wkg_ctx::keep_lineage(const wkg_table *in, ProbePass &pp, uint64_t nz, wkg_table *t, int pid) {
  if (!lineage || !pp.lin || nz == 0 || t->rows < 4 * nz) return;
  if (!in->owned && !in->slice_of) return;  // a caller's buffer (no-alloc step): nothing keeps it alive
  if (in->bytes + in->rows * 16 > pool->budget / 16) return;
  wkg_table *par = const_cast<wkg_table *>(in);
  par->refs++;
  t->lin_parent = par;
  t->lin_info = pp.cinfo;
  t->lin_excl = pp.cexcl;
  t->lin_n = nz;
  t->lin_pid = pid;
  t->lin_alloc = pp.lin;
  pp.lin = nullptr;
}

// known_to_unknown on column sc of a table whose rows are its parent's
// compacted rows fanned out (t->lin_*), with sc one of the parent's columns:
// all rows from one parent row share the key, so the keys are gathered
// once per compacted parent row and probed as a one-column table. Kept
// parent rows are then spread back to their child rows in order, each with
// the parent's list, which gives the per-row step's output row for row.
int wkg_ctx::step_k2u_lineage(const wkg_table *in, int sc, int pid, int dir, hipStream_t s, wkg_table **out,
                              uint64_t *expanded) {
  const wkg_table *par = in->lin_parent;
  const uint64_t np = in->lin_n;
  const int out_cols = in->cols + 1;
  wkg_table *keys = nullptr;
  int rc = new_table(np, 1, s, &keys);
  if (rc) return rc;
  {
    ProfScope ps(this, s, "lineage", 12.0 * (double)np);
    k_gather_col<<<(unsigned)std::min<uint64_t>((np + kBlock - 1) / kBlock, 65535), kBlock, 0, s>>>(
        par->d, (uint32_t)par->cols, (uint32_t)sc, in->lin_info, np, keys->d);
  }
  ProbePass pp;
  rc = probe_pass(kModeK2U, keys, 0, pid, dir, -1, -1, nullptr, s, pp);
  if (rc) {
    drop_table(keys);
    return rc;
  }
  prof_add_bytes("lineage_steps", 1.0);  // a count
  const uint64_t nk = pp.cnt[2], dtot = pp.cnt[0];  // kept parent rows, their list lengths summed
  uint64_t kc = 0, E = 0;
  uint64_t *lift = nullptr;
  void *lin = nullptr, *glin = nullptr;
  if (nk) {
    // per kept parent row: child rows and output rows, scanned like tile sums
    const uint64_t nchunks = (nk + kScanChunk - 1) / kScanChunk;
    lift = (uint64_t *)pool->alloc((4 + 2 * nk + 3 * nchunks) * 8, s, err_slot());
    if (!lift) {
      release_pass(pp, s);
      drop_table(keys);
      return WKG_ENOMEM;
    }
    const KeptRows kept{pp.cinfo, pp.cexcl, nk, dtot, in->lin_excl, np, in->rows};
    StepParams q{};
    q.totals = lift;
    q.tile_sum = lift + 4;       // output rows -> prefix
    q.tile_nz = q.tile_sum + nk; // child rows -> prefix
    q.tile_ex = q.tile_nz;       // (read before it is rewritten)
    q.ntiles = nk;
    q.chunk_sum = q.tile_nz + nk;
    q.chunk_nz = q.chunk_sum + nchunks;
    q.chunk_ex = q.chunk_nz + nchunks;
    q.sg = next_signal();  // the scan publishes the totals
    {
      ProfScope ps(this, s, "lineage", 48.0 * (double)nk);
      k_lift<<<(unsigned)std::min<uint64_t>((nk + kBlock - 1) / kBlock, 65535), kBlock, 0, s>>>(kept, q.tile_nz,
                                                                                                q.tile_sum);
      k_scan_chunks<<<(unsigned)nchunks, kBlock, 0, s>>>(q);
      if (nchunks > 1) k_scan_top<<<1, kBlock, 0, s>>>(q, nchunks);
    }
    hipError_t le = hipGetLastError();
    uint64_t cnt[3];
    rc = le == hipSuccess ? wait_signal(q.sg, s) : fail(WKG_EDEVICE, "lineage launch failed: %s",
                                                         hipGetErrorString(le));
    if (!rc) memcpy(cnt, rb_host(), sizeof cnt);
    if (rc) {
      pool->release(lift, s);
      release_pass(pp, s);
      drop_table(keys);
      return rc;
    }
    E = cnt[0];
    kc = cnt[2];
    if (table_limit && in->rows >= 2 && E * (uint64_t)out_cols * 4 > table_limit) {
      pool->release(lift, s);
      release_pass(pp, s);
      drop_table(keys);
      return fail(WKG_ENOMEM, "a %llu-row output is above the table limit", (unsigned long long)E);
    }
    // kept child rows: (row, list offset) and first output row
    lin = pool->alloc(kc * 16, s, err_slot());
    if (!lin) {
      pool->release(lift, s);
      release_pass(pp, s);
      drop_table(keys);
      return WKG_ENOMEM;
    }
    {
      ProfScope ps(this, s, "lineage", 28.0 * (double)kc);
      k_spread<<<(unsigned)std::min<uint64_t>((kc + kBlock - 1) / kBlock, 1u << 20), kBlock, 0, s>>>(
          kept, q.tile_nz, q.chunk_nz, q.tile_sum, q.chunk_sum, kc, (uint2 *)lin, (uint64_t *)((char *)lin + kc * 8));
    }
    // the output's groups, when a group's list repeats over 4+ rows on
    // average (filters from the new column then decide once per entry).
    // They keep the grandparent table alive with the output, so its bytes
    // count against the same share of the budget as a kept lineage parent.
    if (E >= 4 * dtot && par->bytes + nk * 24 <= pool->budget / 16) {
      glin = pool->alloc(nk * 24, s, err_slot());
      if (glin) {
        ProfScope ps(this, s, "lineage", 40.0 * (double)nk);
        k_keep_groups<<<(unsigned)std::min<uint64_t>((nk + kBlock - 1) / kBlock, 65535), kBlock, 0, s>>>(
            kept, in->lin_info, q.tile_sum, q.chunk_sum, (uint2 *)glin, (uint64_t *)glin + nk,
            (uint64_t *)glin + 2 * nk);
      }
    }
  }
  if (expanded) *expanded = E;
  wkg_table *t = nullptr;
  rc = new_table(E, out_cols, s, &t);
  if (!rc && E) rc = expand_pass(in, pid, (const uint2 *)lin, (const uint64_t *)((char *)lin + kc * 8), kc, E, s, t);
  if (lift) pool->release(lift, s);
  release_pass(pp, s);
  drop_table(keys);
  if (rc) {
    if (lin) pool->release(lin, s);
    if (glin) pool->release(glin, s);
    drop_table(t);
    return rc;
  }
  if (glin) {
    wkg_table *gp = const_cast<wkg_table *>(par);
    gp->refs++;
    t->glin_parent = gp;
    t->glin_info = (const uint2 *)glin;
    t->glin_s = (const uint64_t *)glin + nk;
    t->glin_o = (const uint64_t *)glin + 2 * nk;
    t->glin_n = nk;
    t->glin_dtot = dtot;
    t->glin_pid = pid;
    t->glin_alloc = glin;
  }
  ProbePass own;  // the spread rows are this output's own lineage
  own.lin = lin;
  own.cinfo = (uint2 *)lin;
  own.cexcl = lin ? (uint64_t *)((char *)lin + kc * 8) : nullptr;
  keep_lineage(in, own, kc, t, pid);
  if (own.lin) pool->release(own.lin, s);
  *out = t;
  return WKG_OK;
}

// known_to_known / known_to_const from the last column of a table with
// group lineage (the end column, if any, one of the grandparent's): each
// group's list entry is decided once, on the table S of (end cell, child
// rows, entry) rows -- the group table expanded by its lists -- with each
// entry's examined edges counted once per child row. The decisions are then
// read back for every row of the input and the kept rows compacted in
// order, as the per-row filter's compaction does, which gives its output
// and its work count.
int wkg_ctx::step_filter_groups(int mode, const wkg_table *in, int pid, int dir, int ec, int64_t cval,
                                hipStream_t s, wkg_table **out, uint64_t *expanded) {
  const wkg_table *gp = in->glin_parent;
  const uint64_t ng = in->glin_n, dtot = in->glin_dtot, R = in->rows;
  const int c = in->cols;
  int rc = ensure_pred(in->glin_pid, s);
  if (rc) return rc;
  prof_add_bytes("lineage_filters", 1.0);  // a count
  wkg_table *G = nullptr, *S = nullptr, *t = nullptr;
  uint2 *gci = nullptr;
  uint8_t *D = nullptr;
  ProbePass sp, pp;
  auto done = [&](int r) {
    release_pass(sp, s);
    release_pass(pp, s);
    if (gci) pool->release(gci, s);
    if (D) pool->release(D, s);
    drop_table(G);
    drop_table(S);
    if (r) drop_table(t);
    return r;
  };
  if ((rc = new_table(ng, 2, s, &G))) return done(rc);
  gci = (uint2 *)pool->alloc(ng * 8, s, err_slot());
  if (!gci) return done(WKG_ENOMEM);
  {
    ProfScope ps(this, s, "lineage", 40.0 * (double)ng);
    k_group_table<<<(unsigned)std::min<uint64_t>((ng + kBlock - 1) / kBlock, 65535), kBlock, 0, s>>>(
        gp->d, (uint32_t)gp->cols, mode == kModeK2K ? ec : -1, in->glin_info, in->glin_s, in->glin_o, ng, dtot, R,
        G->d, gci);
  }
  if ((rc = new_table(dtot, 3, s, &S)) ||
      (rc = expand_pass(G, in->glin_pid, gci, in->glin_s, ng, dtot, s, S, "lineage")))
    return done(rc);
  // decisions per (group, entry): kept S rows, examined edges weighted by column 1
  if ((rc = probe_pass(mode, S, 2, pid, dir, mode == kModeK2K ? 0 : -1, cval, nullptr, s, sp, 1)))
    return done(rc);
  const uint64_t examined = sp.cnt[1];
  D = (uint8_t *)pool->alloc(dtot, s, err_slot());
  if (!D) return done(WKG_ENOMEM);
  {
    ProfScope ps(this, s, "lineage", (double)dtot + 16.0 * (double)sp.cnt[2]);
    k_zero_bytes<<<(unsigned)std::min<uint64_t>((dtot + kBlock - 1) / kBlock, 65535), kBlock, 0, s>>>(D, dtot);
    if (sp.cnt[2])
      k_hits_flags<<<(unsigned)std::min<uint64_t>(sp.ntiles, 65535), kBlock, 0, s>>>(
          sp.hits, sp.tile_nz, sp.chunk_nz, sp.ntiles, sp.cnt[2], (uint32_t)sp.tr, D);
  }
  // the input's rows: kept flags, tile scan, compaction
  constexpr int TR = 2048;
  if ((rc = new_table(R, c, s, &t))) return done(rc);
  StepParams p{};
  if ((rc = pass_setup(this, mode, in, c - 1, pid, dir, ec, cval, t->d, s, TR, 0, pp, p))) return done(rc);
  const uint64_t ntiles = p.ntiles, nchunks = (ntiles + kScanChunk - 1) / kScanChunk;
  {
    ProfScope ps(this, s, "lineage", 1.0 * (double)R);
    const uint64_t per = std::max<uint64_t>(thread_launch_rows() / TR, 1);
    for (uint64_t t0 = 0; t0 < ntiles; t0 += per) {
      p.tile0 = t0;
      k_flag_groups<TR><<<(unsigned)std::min<uint64_t>(per, ntiles - t0), kBlock, 0, s>>>(
          p, in->glin_o, in->glin_s, ng, dtot, D);
    }
    p.tile0 = 0;
  }
  p.sg = next_signal();
  {
    ProfScope ps(this, s, "scan_tiles", 48.0 * (double)ntiles);
    k_scan_chunks<<<(unsigned)nchunks, kBlock, 0, s>>>(p);
    if (nchunks > 1) k_scan_top<<<1, kBlock, 0, s>>>(p, nchunks);
  }
  {
    ProfScope ps(this, s, "compact", 0.0);
    launch_finalize(TR, p, s);
  }
  hipError_t le = hipGetLastError();
  if (le != hipSuccess) return done(fail(WKG_EDEVICE, "filter launch failed: %s", hipGetErrorString(le)));
  if ((rc = wait_signal(p.sg, s))) return done(rc);
  uint64_t cnt[3];
  memcpy(cnt, rb_host(), sizeof cnt);
  prof_add_bytes("compact", 16.0 * cnt[2] + 8.0 * cnt[2] * c);
  t->rows = cnt[0];
  t->bytes = t->rows * (uint64_t)c * 4;
  if (expanded) *expanded = examined;
  *out = t;
  return done(WKG_OK);
}

// known_to_known / known_to_const from the last column of a table with
// lineage (a known_to_unknown output that fanned out; the end column, if
// any, one of the parent's): the distinct lists of its compacted parents are
// found on the host from the lineage (parents sharing a list offset share
// the list), each distinct list's entries are probed once under the
// filter's predicate, and every row then tests its parent's target against
// its entry's list (k_fanout_filter), with the reference's work count. Only
// kept rows are read from the input. Returns kNotTaken (nothing done) when
// the rows do not repeat the distinct entries at least 4 times, or the
// parents are too many to dedup on the host.
int wkg_ctx::step_filter_fanout(int mode, const wkg_table *in, int pid, int dir, int ec, int64_t cval, hipStream_t s,
                                wkg_table **out, uint64_t *expanded) {
  constexpr uint64_t kMaxParents = 1u << 20;
  const uint64_t np = in->lin_n, R = in->rows;
  const int c = in->cols;
  if (np == 0 || np > kMaxParents) return kNotTaken;
  // the compacted parents' lists: (offset, first child row) per parent
  std::vector<uint64_t> h((size_t)np * 2);
  {
    std::vector<uint2> info(np);
    HIPCHK(this, hipMemcpyAsync(info.data(), in->lin_info, np * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(this, hipMemcpyAsync(h.data() + np, in->lin_excl, np * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(this, hipStreamSynchronize(s));
    for (uint64_t i = 0; i < np; i++) h[i] = info[i].y;
  }
  // distinct lists by offset: ly (offset), lbase (first entry in D), ln
  // (length); per parent its list and that list's first D entry
  std::unordered_map<uint64_t, uint64_t> seen;
  seen.reserve(np * 2);
  std::vector<uint64_t> lbase, ln, pbase(np), plist(np);
  std::vector<uint32_t> ly;
  uint64_t total = 0;
  for (uint64_t i = 0; i < np; i++) {
    const uint64_t n = (i + 1 < np ? h[np + i + 1] : R) - h[np + i];
    auto it = seen.find(h[i]);
    if (it == seen.end()) {
      it = seen.emplace(h[i], lbase.size()).first;
      lbase.push_back(total);
      ly.push_back((uint32_t)h[i]);
      ln.push_back(n);
      total += n;
    }
    plist[i] = it->second;
    pbase[i] = lbase[it->second];
  }
  if (R < 4 * total) return kNotTaken;
  const uint64_t nl = lbase.size();
  // parent-aligned units of TRU rows when they average at least 256 rows
  // (long shared lists), else row tiles of TR rows
  constexpr int TR = 1024, TRU = 2048;
  static_assert(TRU == kFBRows, "k_fanout_blocks fills the units' bit words");
  uint64_t nunits = 0;
  std::vector<uint64_t> ubase(np);
  for (uint64_t i = 0; i < np; i++) {
    ubase[i] = nunits;
    nunits += (ln[plist[i]] + TRU - 1) / TRU;
  }
  const bool units = nunits * 256 <= R;
  int rc = ensure_pred(in->lin_pid, s);
  if (rc) return rc;
  prof_add_bytes("fanout_filters", 1.0);  // counts: taken, and by tile kind
  prof_add_bytes(units ? "fanout_units" : "fanout_row_tiles", 1.0);
  prof_add_bytes("fanout_parents", (double)np);  // sizes, summed over the filters
  prof_add_bytes("fanout_lists", (double)nl);
  prof_add_bytes("fanout_list_entries", (double)total);
  const int tr = units ? TRU : TR;
  const uint64_t ntiles = units ? nunits : (R + TR - 1) / TR, nchunks = (ntiles + kScanChunk - 1) / kScanChunk;
  // uploads: pbase | lbase | ly (u32) [| ubase | xbase | poff | par (u32) | ln | xb]
  const uint64_t w_ly = (nl + 1) / 2, w_par = (np + 1) / 2;
  const uint64_t up_words = np + nl + w_ly + (units ? np + 2 * (nl + 1) + w_par + nl + (nl + 1) : 0);
  uint64_t nwork = 0;  // units through k_fanout_blocks: (block of a list) x (group of its parents) items
  double blocks_bytes = 0;  // k_fanout_blocks' reads and writes (see below)
  const uint64_t scan_words = 4 + 3 * ntiles + 3 * nchunks + (units ? ntiles : 0);  // (+ unit first rows)
  const uint64_t bit_words = (ntiles * (uint64_t)(tr / 32) + 1) / 2;
  uint64_t *buf = (uint64_t *)pool->alloc((up_words + scan_words + bit_words + total) * 8, s, err_slot());
  if (!buf) return WKG_ENOMEM;
  std::vector<uint64_t> up(up_words, 0);
  {
    uint64_t o = 0;
    memcpy(up.data() + o, pbase.data(), np * 8);
    o += np;
    memcpy(up.data() + o, lbase.data(), nl * 8);
    o += nl;
    memcpy((uint32_t *)(up.data() + o), ly.data(), nl * 4);
    o += w_ly;
    if (units) {
      memcpy(up.data() + o, ubase.data(), np * 8);
      o += np;
      // parents grouped by list (row order within a list); execution bases
      std::vector<uint64_t> cnt(nl + 1, 0);
      for (uint64_t i = 0; i < np; i++) cnt[plist[i] + 1]++;
      uint64_t *xbase = up.data() + o, *poff = xbase + nl + 1;
      uint32_t *par = (uint32_t *)(poff + nl + 1);
      poff[0] = 0;
      xbase[0] = 0;
      for (uint64_t L = 0; L < nl; L++) {
        poff[L + 1] = poff[L] + cnt[L + 1];
        xbase[L + 1] = xbase[L] + cnt[L + 1] * ((ln[L] + TRU - 1) / TRU);
      }
      std::vector<uint64_t> fill(poff, poff + nl);
      for (uint64_t i = 0; i < np; i++) par[fill[plist[i]]++] = (uint32_t)i;
      o += 2 * (nl + 1) + w_par;  // past xbase, poff and par
      uint64_t *lnd = up.data() + o, *xb = lnd + nl;
      xb[0] = 0;
      for (uint64_t L = 0; L < nl; L++) {
        lnd[L] = ln[L];
        xb[L + 1] = xb[L] + ((ln[L] + kFBRows - 1) / kFBRows) * ((cnt[L + 1] + kFBGroup - 1) / kFBGroup);
      }
      nwork = xb[nl];
      prof_add_bytes("fanout_work_items", (double)nwork);
      // per work item its block's D entries (8 B each) and its parents'
      // cells, ranks and first rows (24 B each); per unit its three counters
      for (uint64_t L = 0; L < nl; L++)
        blocks_bytes += 8.0 * (double)ln[L] * (double)((cnt[L + 1] + kFBGroup - 1) / kFBGroup) + 24.0 * (double)cnt[L + 1] * (double)((ln[L] + kFBRows - 1) / kFBRows);
      blocks_bytes += 24.0 * (double)ntiles;
    }
  }
  uint64_t *d_pbase = buf, *d_lbase = buf + np;
  uint32_t *d_ly = (uint32_t *)(d_lbase + nl);
  uint64_t *d_ubase = buf + np + nl + w_ly, *d_xbase = d_ubase + np, *d_poff = d_xbase + nl + 1;
  uint32_t *d_par = (uint32_t *)(d_poff + nl + 1);
  uint64_t *d_ln = (uint64_t *)d_par + w_par, *d_xb = d_ln + nl;
  uint64_t *scr = buf + up_words;
  uint32_t *bits = (uint32_t *)(scr + scan_words);
  uint2 *D = (uint2 *)(scr + scan_words + bit_words);
  wkg_table *K = nullptr, *t = nullptr;
  ProbePass kp;
  auto done = [&](int r) {
    release_pass(kp, s);
    drop_table(K);
    if (r) drop_table(t);
    pool->release(buf, s);
    return r;
  };
  if (hipMemcpyAsync(buf, up.data(), up_words * 8, hipMemcpyHostToDevice, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return done(fail(WKG_EDEVICE, "fan-out filter upload failed"));
  if ((rc = new_table(total, 1, s, &K))) return done(rc);
  const GraphView g = view();
  {
    ProfScope ps(this, s, "lineage", 8.0 * (double)total + 8.0 * (double)total);
    const unsigned grid = (unsigned)std::min<uint64_t>((total + kBlock - 1) / kBlock, 65535);
    k_gather_lists<false><<<grid, kBlock, 0, s>>>(g, (uint32_t)in->lin_pid, d_lbase, d_ly, nl, total, K->d);
    k_zero_u64<<<grid, kBlock, 0, s>>>((uint64_t *)D, total);
  }
  if ((rc = probe_pass(kModeK2U, K, 0, pid, dir, -1, -1, nullptr, s, kp))) return done(rc);
  if (kp.cnt[2]) {
    ProfScope ps(this, s, "lineage", 24.0 * (double)kp.cnt[2]);
    k_key_degrees<<<(unsigned)std::min<uint64_t>((kp.cnt[2] + kBlock - 1) / kBlock, 65535), kBlock, 0, s>>>(
        g, (uint32_t)pid, kp.cinfo, kp.cexcl, kp.cnt[2], kp.cnt[0], D);
  }
  // the rows: decisions as tile bit words, the tile scan, then the kept rows
  StepParams p{};
  p.T = in->d;
  p.R = R;
  p.c = (uint32_t)c;
  p.pid = (uint32_t)pid;
  p.dir = (uint32_t)dir;
  p.mode = mode;
  p.target = (uint32_t)cval;
  p.target_ok = (cval >= 0 && cval <= 0xffffffffLL) ? 1u : 0u;
  p.totals = scr;
  p.tile_sum = scr + 4;
  p.tile_nz = p.tile_sum + ntiles;
  p.tile_ex = p.tile_nz + ntiles;
  p.chunk_sum = p.tile_ex + ntiles;
  p.chunk_nz = p.chunk_sum + nchunks;
  p.chunk_ex = p.chunk_nz + nchunks;
  p.ntiles = ntiles;
  uint64_t *row0 = units ? p.chunk_ex + nchunks : nullptr;
  FanoutFilter f{};
  f.excl = in->lin_excl;
  f.info = in->lin_info;
  f.np = np;
  f.pbase = d_pbase;
  f.D = D;
  f.P = in->lin_parent->d;
  f.pc = (uint32_t)in->lin_parent->cols;
  f.ec = mode == kModeK2K ? ec : -1;
  f.bits = bits;
  FanoutUnits u{d_xbase, d_poff, d_par, d_ubase, nl, row0};
  // long shared lists: k_fanout_blocks, which lists its kept rows (up to
  // kept_cap; more kept rows than that and the filter reruns on k_fanout_units)
  bool blocks = units && fanout_blocks;
  const uint64_t kept_cap = std::min<uint64_t>(R, fanout_kept_cap);
  uint4 *kept_rows = nullptr;
  if (blocks) {
    kept_rows = (uint4 *)pool->alloc(kept_cap * 16 + 256, s, err_slot());
    if (!kept_rows) return done(WKG_ENOMEM);
  }
  auto done2 = [&](int r) {
    if (kept_rows) pool->release(kept_rows, s);
    return done(r);
  };
  uint64_t *nkept = (uint64_t *)(kept_rows + kept_cap);
  FanoutBlocks fb{d_xb, d_poff, d_par, d_ubase, d_ln, d_lbase, nl, row0, kept_rows, nkept, kept_cap};
  hipError_t le;
  uint64_t kept = 0, examined = 0;
  for (;;) {
    {
      // one parent per workgroup: per row a D entry (8 B, mostly cached) and
      // the list it names; blocks: each D entry once per parent group (the
      // input rows are never read)
      ProfScope ps(this, s, "fanout_filter", blocks ? blocks_bytes : 8.0 * (double)R);
      const uint64_t per = std::max<uint64_t>(thread_launch_rows() / (uint64_t)tr, 1);
      if (blocks) {
        k_zero_u64<<<1, kBlock, 0, s>>>(nkept, 1);
        // work items of kFBRows rows x up to kFBGroup parents: per-launch rows as
        // for units, but at least 8 workgroups per CU (a launch fills the GPU)
        const uint64_t wper = std::max<uint64_t>(per / kFBGroup, 8 * (uint64_t)ncu);
        for (uint64_t w0 = 0; w0 < nwork; w0 += wper) {
          p.tile0 = w0;
          const unsigned grid = (unsigned)std::min<uint64_t>(wper, nwork - w0);
          if (sorted) k_fanout_blocks<true><<<grid, kBlock, 0, s>>>(g, p, f, fb);
          else k_fanout_blocks<false><<<grid, kBlock, 0, s>>>(g, p, f, fb);
        }
      } else {
        for (uint64_t t0 = 0; t0 < ntiles; t0 += per) {
          p.tile0 = t0;
          const unsigned grid = (unsigned)std::min<uint64_t>(per, ntiles - t0);
          if (units) {
            if (sorted) k_fanout_units<TRU, true><<<grid, kBlock, 0, s>>>(g, p, f, u);
            else k_fanout_units<TRU, false><<<grid, kBlock, 0, s>>>(g, p, f, u);
          } else {
            if (sorted) k_fanout_filter<TR, true><<<grid, kBlock, 0, s>>>(g, p, f);
            else k_fanout_filter<TR, false><<<grid, kBlock, 0, s>>>(g, p, f);
          }
        }
      }
      p.tile0 = 0;
    }
    p.sg = next_signal();
    {
      ProfScope ps(this, s, "scan_tiles", 48.0 * (double)ntiles);
      k_scan_chunks<<<(unsigned)nchunks, kBlock, 0, s>>>(p);
      if (nchunks > 1) k_scan_top<<<1, kBlock, 0, s>>>(p, nchunks);
    }
    le = hipGetLastError();
    if (le != hipSuccess) return done2(fail(WKG_EDEVICE, "fan-out filter launch failed: %s", hipGetErrorString(le)));
    if ((rc = wait_signal(p.sg, s))) return done2(rc);
    kept = rb_host()[0];
    examined = rb_host()[1];
    if (!blocks || kept <= kept_cap) break;
    blocks = false;  // more kept rows than the list holds: the bit-word path
    prof_add_bytes("fanout_kept_overflow", 1.0);  // a count
  }
  if ((rc = new_table(kept, c, s, &t))) return done2(rc);
  if (kept) {
    p.O = t->d;
    ProfScope ps(this, s, "compact", 8.0 * (double)kept * c);
    const uint64_t estart = h_meta[in->lin_pid].edge_start;
    if (blocks) {
      k_compact_kept<<<(unsigned)((kept + kBlock - 1) / kBlock), kBlock, 0, s>>>(g, p, kept_rows, kept, f, estart);
    } else {
      const uint64_t per = std::max<uint64_t>(thread_launch_rows() / (uint64_t)tr, 4) & ~uint64_t(3);
      for (uint64_t t0 = 0; t0 < ntiles; t0 += per) {
        p.tile0 = t0;
        const unsigned grid = (unsigned)((std::min<uint64_t>(per, ntiles - t0) + kBlock / kWave - 1) / (kBlock / kWave));
        if (in->d) {
          if (units) k_compact_bits<TRU><<<grid, kBlock, 0, s>>>(p, bits, row0);
          else k_compact_bits<TR><<<grid, kBlock, 0, s>>>(p, bits, nullptr);
        } else {  // a fan-out that was never written: rows from the lineage
          if (units) k_compact_lineage<TRU><<<grid, kBlock, 0, s>>>(g, p, bits, row0, f, estart);
          else k_compact_lineage<TR><<<grid, kBlock, 0, s>>>(g, p, bits, nullptr, f, estart);
        }
      }
    }
    if ((le = hipGetLastError()) != hipSuccess)
      return done2(fail(WKG_EDEVICE, "fan-out compaction launch failed: %s", hipGetErrorString(le)));
  }
  if (kept_rows) {
    pool->release(kept_rows, s);
    kept_rows = nullptr;
  }
  if (expanded) *expanded = examined;
  *out = t;
  t = nullptr;
  return done(WKG_OK);
}

// known_to_unknown on column sc of `in` followed by a filter (fmode) whose
// start column is the new column: when the filter can run on the fan-out's
// lineage (step_filter_fanout), the fan-out is never written -- a table
// with lineage and no cells stands for it -- and *out is the filter's
// output (*fused = 1); the rows and work counts are those of the two steps
// run one after the other. Otherwise *out is the known_to_unknown output
// (*fused = 0) and the caller runs the filter on it. The reference runs the
// two handlers one after the other (core/rcache.hpp:232-414, then 334-466).
int wkg_ctx::step_k2u_filter(const wkg_table *in, int sc, int pid, int dir, int fmode, int fpid, int fdir, int ec,
                             int64_t cval, hipStream_t s, wkg_table **out, int *fused, uint64_t *ex_u,
                             uint64_t *ex_f) {
  *fused = 0;
  *ex_f = 0;
  const bool fits = lineage && !framed && in && in->rows > kSmallRows && in->cols >= 1 && sc >= 0 && sc < in->cols &&
                    !(in->lin_parent && sc < in->lin_parent->cols) && in->cols + 1 <= kBlock &&
                    in->rows < (1ull << 32) && (in->owned || in->slice_of);
  if (!fits || check_pattern(fpid, fdir) != WKG_OK || (fmode == kModeK2K && (ec < 0 || ec >= in->cols)))
    return step_known(kModeK2U, in, sc, pid, dir, -1, -1, s, out, ex_u);
  int rc = check_pattern(pid, dir);
  if (rc) return rc;
  if ((rc = ensure_pred(pid, s))) return rc;
  const int c = in->cols;
  ProbePass pp;
  if ((rc = probe_pass(kModeK2U, in, sc, pid, dir, -1, -1, nullptr, s, pp))) return rc;
  const uint64_t E = pp.cnt[0], nz = pp.cnt[2];
  *ex_u = E;
  if (E >= group_filter_rows && E < (1ull << 32) && nz > 0 && E >= 4 * nz && in->bytes + in->rows * 16 <= pool->budget / 16) {
    // the fan-out as lineage only
    wkg_table *v = new (std::nothrow) wkg_table;
    if (!v) {
      release_pass(pp, s);
      return fail(WKG_ENOMEM, "out of host memory");
    }
    v->d = nullptr;  // (no cells: not owned)
    v->owned = false;
    v->rows = E;
    v->cols = c + 1;
    v->bytes = 0;
    v->stream = s;
    wkg_table *par = const_cast<wkg_table *>(in);
    par->refs++;
    v->lin_parent = par;
    v->lin_info = pp.cinfo;
    v->lin_excl = pp.cexcl;
    v->lin_n = nz;
    v->lin_pid = pid;
    v->lin_alloc = pp.lin;
    pp.lin = nullptr;
    release_pass(pp, s);
    prof_add_bytes("fused_fanouts", 1.0);  // a count
    rc = step_filter_fanout(fmode, v, fpid, fdir, ec, cval, s, out, ex_f);
    if (rc != kNotTaken) {
      drop_table(v);
      if (rc == WKG_OK) *fused = 1;
      return rc;
    }
    // the filter declined: write the fan-out after all
    drop_table(v);
    prof_add_bytes("fused_fanouts", -1.0);
    return step_known(kModeK2U, in, sc, pid, dir, -1, -1, s, out, ex_u);
  }
  const int out_cols = c + 1;
  if (table_limit && in->rows >= 2 && E * (uint64_t)out_cols * 4 > table_limit) {
    release_pass(pp, s);
    return fail(WKG_ENOMEM, "a %llu-row output is above the table limit", (unsigned long long)E);
  }
  wkg_table *t = nullptr;
  if ((rc = new_table(E, out_cols, s, &t)) || (rc = expand_pass(in, pid, pp.cinfo, pp.cexcl, nz, E, s, t))) {
    release_pass(pp, s);
    drop_table(t);
    return rc;
  }
  keep_lineage(in, pp, nz, t, pid);
  release_pass(pp, s);
  *out = t;
  return WKG_OK;
}

int wkg_ctx::step_known(int mode, const wkg_table *in, int sc, int pid, int dir, int ec, int64_t cval,
                        hipStream_t s, wkg_table **out, uint64_t *expanded) {
  int rc = check_pattern(pid, dir);
  if (rc) return rc;
  if (!in) return fail(WKG_EINVAL, "null input table");
  const uint64_t R = in->rows;
  const int c = in->cols;
  if (c < 1 || sc < 0 || sc >= c) return fail(WKG_EINVAL, "start column %d outside the %d-column table", sc, c);
  if (mode == kModeK2K && (ec < 0 || ec >= c))
    return fail(WKG_EINVAL, "end column %d outside the %d-column table", ec, c);
  const int out_cols = mode == kModeK2U ? c + 1 : c;
  if (out_cols > kBlock) return fail(WKG_ERANGE, "tables are limited to %d columns", kBlock);
  if (R >= (1ull << 32)) return fail(WKG_ERANGE, "tables are limited to 2^32 - 1 rows per step");
  if ((rc = ensure_pred(pid, s))) return rc;
  if (R == 0) {  // empty input gives an empty result (core/rcache.hpp:242-250)
    if (expanded) *expanded = 0;
    return new_table(0, out_cols, s, out);
  }
  if (mode == kModeK2U && lineage && in->lin_parent && sc < in->lin_parent->cols)
    return step_k2u_lineage(in, sc, pid, dir, s, out, expanded);
  if (small_steps && R <= kSmallRows) return step_small(mode, in, sc, pid, dir, ec, cval, s, out, expanded);
  if (mode != kModeK2U && lineage && !framed && in->glin_parent && sc == c - 1 &&
      (mode == kModeK2C || ec < in->glin_parent->cols) && R >= group_filter_rows && R >= 4 * in->glin_dtot)
    return step_filter_groups(mode, in, pid, dir, ec, cval, s, out, expanded);
  if (mode != kModeK2U && lineage && !framed && in->lin_parent && sc == c - 1 &&
      (mode == kModeK2C || ec < in->lin_parent->cols) && R >= group_filter_rows) {
    rc = step_filter_fanout(mode, in, pid, dir, ec, cval, s, out, expanded);
    if (rc != kNotTaken) return rc;
  }

  wkg_table *t = nullptr;
  if (mode != kModeK2U && (rc = new_table(R, c, s, &t))) return rc;
  ProbePass pp;
  if ((rc = probe_pass(mode, in, sc, pid, dir, ec, cval, t ? t->d : nullptr, s, pp))) {
    drop_table(t);
    return rc;
  }
  if (mode != kModeK2U) {  // filters: kept rows compacted in place
    t->rows = pp.cnt[0];
    t->bytes = t->rows * (uint64_t)c * 4;
    if (expanded) *expanded = pp.cnt[1];
    release_pass(pp, s);
    *out = t;
    return WKG_OK;
  }
  const uint64_t E = pp.cnt[0], nz = pp.cnt[2];
  if (table_limit && R >= 2 && E * (uint64_t)out_cols * 4 > table_limit) {  // (one row cannot be split)
    release_pass(pp, s);
    return fail(WKG_ENOMEM, "a %llu-row output is above the table limit", (unsigned long long)E);
  }
  if (expanded) *expanded = E;
  if ((rc = new_table(E, out_cols, s, &t)) || (rc = expand_pass(in, pid, pp.cinfo, pp.cexcl, nz, E, s, t))) {
    release_pass(pp, s);
    drop_table(t);
    return rc;
  }
  keep_lineage(in, pp, nz, t, pid);
  release_pass(pp, s);
  *out = t;
  return WKG_OK;
}

// A known-start step over at most kSmallRows rows in one launch
// (k_small_step) and one read-back of its totals. A known_to_unknown output
// of more than kSmallCap rows is expanded by the multi-launch expand from
// the compacted rows the launch left.
int wkg_ctx::step_small(int mode, const wkg_table *in, int sc, int pid, int dir, int ec, int64_t cval,
                        hipStream_t s, wkg_table **out, uint64_t *expanded) {
  const uint64_t R = in->rows;
  const int c = in->cols;
  const int out_cols = mode == kModeK2U ? c + 1 : c;
  wkg_table *t = nullptr;
  int rc = new_table(mode == kModeK2U ? kSmallCap : R, out_cols, s, &t);
  if (rc) return rc;
  ProbePass pp;
  StepParams p{};
  const uint64_t part_words = kSmallCap / kExpandNV + 2;  // split points of the in-launch expand
  if ((rc = pass_setup(this, mode, in, sc, pid, dir, ec, cval, mode == kModeK2U ? nullptr : t->d, s, kBlock,
                       part_words, pp, p))) {
    drop_table(t);
    return rc;
  }
  uint64_t *part = (uint64_t *)(p.hits + p.ntiles * kBlock);
  ExpandParams x{};
  x.pid = (uint32_t)pid;
  x.T = in->d;
  x.c = (uint32_t)c;
  x.cinfo = p.cinfo;
  x.cexcl = p.cexcl;
  x.O = t->d;
  const GraphView g = view();
  const bool srt = sorted || mode == kModeK2U;
  const Signal sg = next_signal();
  {
    ProfScope ps(this, s, "small_step", 4.0 * (double)R * c + 128.0 * (double)R);
    if (framed) {
      if (srt) k_small_step<true, true><<<1, kBlock, 0, s>>>(g, p, x, kSmallCap, part, sg);
      else k_small_step<false, true><<<1, kBlock, 0, s>>>(g, p, x, kSmallCap, part, sg);
    } else {
      if (srt) k_small_step<true, false><<<1, kBlock, 0, s>>>(g, p, x, kSmallCap, part, sg);
      else k_small_step<false, false><<<1, kBlock, 0, s>>>(g, p, x, kSmallCap, part, sg);
    }
  }
  hipError_t le = hipGetLastError();
  rc = le == hipSuccess ? wait_signal(sg, s) : fail(WKG_EDEVICE, "small step launch failed: %s", hipGetErrorString(le));
  if (rc) {
    release_pass(pp, s);
    drop_table(t);
    return rc;
  }
  memcpy(pp.cnt, rb_host(), 3 * sizeof(uint64_t));
  if (mode != kModeK2U) {  // filters: kept rows compacted in place
    t->rows = pp.cnt[0];
    t->bytes = t->rows * (uint64_t)c * 4;
    if (expanded) *expanded = pp.cnt[1];
    release_pass(pp, s);
    *out = t;
    return WKG_OK;
  }
  const uint64_t E = pp.cnt[0], nz = pp.cnt[2];
  if (table_limit && R >= 2 && E * (uint64_t)out_cols * 4 > table_limit) {  // (one row cannot be split)
    release_pass(pp, s);
    drop_table(t);
    return fail(WKG_ENOMEM, "a %llu-row output is above the table limit", (unsigned long long)E);
  }
  if (E <= kSmallCap) {  // expanded in the launch
    t->rows = E;
    t->bytes = E * (uint64_t)out_cols * 4;
  } else {
    drop_table(t);
    t = nullptr;
    if ((rc = new_table(E, out_cols, s, &t)) || (rc = expand_pass(in, pid, pp.cinfo, pp.cexcl, nz, E, s, t))) {
      release_pass(pp, s);
      drop_table(t);
      return rc;
    }
  }
  if (expanded) *expanded = E;
  keep_lineage(in, pp, nz, t, pid);
  release_pass(pp, s);
  *out = t;
  return WKG_OK;
}

int wkg_thread_launch_rows(wkg_ctx *ctx, uint64_t rows) {
  if (!ctx) return WKG_EINVAL;
  if (rows != 0 && (rows < 4096 || rows > (1ull << 32)))
    return ctx->fail(WKG_EINVAL, "launch rows %llu outside [4096, 2^32] (0: the default)", (unsigned long long)rows);
  thread_launch_rows() = rows ? rows : kDefaultLaunchRows;
  return WKG_OK;
}

int wkg_set_small_steps(wkg_ctx *ctx, int on) {
  if (!ctx) return WKG_EINVAL;
  ctx->small_steps = on != 0;
  return WKG_OK;
}

// ---------------------------------------------------------------------------
// profiling and the copy-bandwidth probe

hipEvent_t wkg_ctx::prof_event() {
  {
    std::lock_guard<std::mutex> lk(prof_mu);
    if (!prof_free.empty()) {
      hipEvent_t e = prof_free.back();
      prof_free.pop_back();
      return e;
    }
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

void wkg_ctx::prof_add_bytes(const char *name, double bytes) {
  if (!profiling) return;
  std::lock_guard<std::mutex> lk(prof_mu);
  prof_stats[name].bytes += bytes;
}

void wkg_ctx::prof_collect() {
  std::lock_guard<std::mutex> lk(prof_mu);
  for (ProfRec &r : prof_pending) {
    float ms = 0;
    if (hipEventSynchronize(r.b) == hipSuccess && hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
      KStat &k = prof_stats[r.name];
      k.launches++;
      k.ms += ms;
      k.bytes += r.bytes;
    }
    prof_free.push_back(r.a);
    prof_free.push_back(r.b);
  }
  prof_pending.clear();
}

namespace {
// 16-byte-per-lane streaming copy (nontemporal loads and stores, 4
// workgroups per CU, the fastest of the shapes measured): the reference point
// for achievable HBM bandwidth and for scaling FETCH_SIZE / WRITE_SIZE counters.
__global__ __launch_bounds__(256) void k_copy_calibrate(const uint4 *__restrict__ src, uint4 *__restrict__ dst,
                                                         uint64_t n) {
  typedef int32_t v4i __attribute__((ext_vector_type(4)));
  const v4i *s4 = (const v4i *)src;
  v4i *d4 = (v4i *)dst;
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256ull)
    __builtin_nontemporal_store(__builtin_nontemporal_load(&s4[i]), &d4[i]);
}
// Random 128-byte line reads, 8 lanes per line (the way the probe reads a
// bucket) and 8 independent lines in flight per 8-lane group: the
// random-access roofline for the hash probe. buf holds line_mask + 1 lines;
// out holds one word per thread (keeps the loads alive).
__global__ __launch_bounds__(256) void k_random_lines(const uint4 *__restrict__ buf, uint64_t line_mask,
                                                      uint32_t rounds, uint64_t seed, uint32_t *out) {
  const uint32_t l8 = threadIdx.x & 7;
  const uint64_t grp = (blockIdx.x * 256ull + threadIdx.x) >> 3;
  uint32_t acc = 0;
  for (uint32_t r = 0; r < rounds; r++) {
    uint4 v[8];
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const uint64_t line = mix64(seed + (grp * rounds + r) * 8 + k) & line_mask;
      v[k] = buf[line * 8 + l8];
    }
#pragma unroll
    for (int k = 0; k < 8; k++) acc ^= v[k].x + v[k].w;
  }
  out[blockIdx.x * 256ull + threadIdx.x] = acc;
}
}  // namespace

// ---------------------------------------------------------------------------
// C ABI

extern "C" {

int wkg_profile_enable(wkg_ctx *ctx, int on) {
  if (!ctx) return WKG_EINVAL;
  ctx->prof_collect();
  if (on) {
    std::lock_guard<std::mutex> lk(ctx->prof_mu);
    ctx->prof_stats.clear();
  }
  ctx->profiling = on != 0;
  return WKG_OK;
}

int wkg_profile_read(wkg_ctx *ctx, const char *kernel, uint64_t *launches, double *total_ms, double *alg_bytes) {
  if (!ctx || !kernel) return WKG_EINVAL;
  ctx->prof_collect();
  std::lock_guard<std::mutex> lk(ctx->prof_mu);
  auto it = ctx->prof_stats.find(kernel);
  const wkg_ctx::KStat k = it == ctx->prof_stats.end() ? wkg_ctx::KStat{} : it->second;
  if (launches) *launches = k.launches;
  if (total_ms) *total_ms = k.ms;
  if (alg_bytes) *alg_bytes = k.bytes;
  return WKG_OK;
}

int wkg_bandwidth_probe(wkg_ctx *ctx, uint64_t bytes, int iters, double *copy_gbps) {
  if (!ctx || !copy_gbps || bytes < 16 || iters < 1) return WKG_EINVAL;
  const uint64_t n = bytes / 16;
  void *a = nullptr, *b = nullptr;
  if (hipMalloc(&a, n * 16) != hipSuccess || hipMalloc(&b, n * 16) != hipSuccess) {
    (void)hipGetLastError();
    if (a) (void)hipFree(a);
    return ctx->fail(WKG_ENOMEM, "cannot allocate %llu bytes for the bandwidth probe", (unsigned long long)bytes);
  }
  hipStream_t s = ctx->stream;
  (void)hipMemsetAsync(a, 1, n * 16, s);
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  int dev_cus = 256;
  (void)hipDeviceGetAttribute(&dev_cus, hipDeviceAttributeMultiprocessorCount, ctx->device);
  const unsigned grid = (unsigned)std::min<uint64_t>((n + 255) / 256, (uint64_t)dev_cus * 4);
  k_copy_calibrate<<<grid, 256, 0, s>>>((const uint4 *)a, (uint4 *)b, n);  // warm up
  (void)hipEventRecord(e0, s);
  for (int i = 0; i < iters; i++) {
    ProfScope ps(ctx, s, "copy_calibrate", 2.0 * n * 16);
    k_copy_calibrate<<<grid, 256, 0, s>>>((const uint4 *)a, (uint4 *)b, n);
  }
  (void)hipEventRecord(e1, s);
  hipError_t e = hipEventSynchronize(e1);
  float ms = 0;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(a);
  (void)hipFree(b);
  if (e != hipSuccess) return ctx->fail(WKG_EDEVICE, "bandwidth probe failed: %s", hipGetErrorString(e));
  *copy_gbps = 2.0 * n * 16 * iters / (ms * 1e-3) / 1e9;
  return WKG_OK;
}

int wkg_random_access_probe(wkg_ctx *ctx, uint64_t bytes, int iters, double *lines_per_s) {
  if (!ctx || !lines_per_s || bytes < (1u << 20) || iters < 1) return WKG_EINVAL;
  uint64_t nlines = 1;
  while (nlines * 2 * 128 <= bytes) nlines *= 2;
  int dev_cus = 256;
  (void)hipDeviceGetAttribute(&dev_cus, hipDeviceAttributeMultiprocessorCount, ctx->device);
  const unsigned grid = (unsigned)dev_cus * 8;
  const uint32_t rounds = 16;
  void *buf = nullptr, *out = nullptr;
  if (hipMalloc(&buf, nlines * 128) != hipSuccess || hipMalloc(&out, (uint64_t)grid * 256 * 4) != hipSuccess) {
    (void)hipGetLastError();
    if (buf) (void)hipFree(buf);
    return ctx->fail(WKG_ENOMEM, "cannot allocate %llu bytes for the random-access probe",
                     (unsigned long long)(nlines * 128));
  }
  hipStream_t s = ctx->stream;
  (void)hipMemsetAsync(buf, 1, nlines * 128, s);
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  k_random_lines<<<grid, 256, 0, s>>>((const uint4 *)buf, nlines - 1, rounds, 1, (uint32_t *)out);  // warm up
  (void)hipEventRecord(e0, s);
  for (int i = 0; i < iters; i++) {
    ProfScope ps(ctx, s, "random_lines", (double)grid * 32 * rounds * 8 * 128);
    k_random_lines<<<grid, 256, 0, s>>>((const uint4 *)buf, nlines - 1, rounds, 1000003ull * (i + 2),
                                       (uint32_t *)out);
  }
  (void)hipEventRecord(e1, s);
  hipError_t e = hipEventSynchronize(e1);
  float ms = 0;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(buf);
  (void)hipFree(out);
  if (e != hipSuccess) return ctx->fail(WKG_EDEVICE, "random-access probe failed: %s", hipGetErrorString(e));
  *lines_per_s = (double)grid * 32 * rounds * 8 * iters / (ms * 1e-3);
  return WKG_OK;
}

// Step scratch of wkg_step_k2u_noalloc, one copy per GPU, held until the
// process ends. Host code never uses this name for memory: it is a host-side
// stand-in for the symbol. Each device's address comes from
// hipGetSymbolAddress once per device (noalloc_scratch_of, called by
// wkg_init with that device current) and is what the kernels are given.
static __device__ __align__(256) uint8_t scratch[WKG_NOALLOC_SCRATCH_BYTES];

static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices];  // per device: the address of its copy of `scratch`

// The current device's scratch address (looked up on its first use).
static void *noalloc_scratch_of(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
      (void)hipGetLastError();
      return nullptr;
    }
    g_scratch_ptr[device] = p;
  }
  return g_scratch_ptr[device];
}

int wkg_init(int device, uint64_t table_budget_bytes, wkg_ctx **out) {
  if (!out) return WKG_EINVAL;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return WKG_EDEVICE;
  if (device < 0 || device >= n) return WKG_EINVAL;
  if (hipSetDevice(device) != hipSuccess) return WKG_EDEVICE;
  wkg_ctx *c = new (std::nothrow) wkg_ctx;
  if (!c) return WKG_ENOMEM;
  c->device = device;
  c->budget_request = table_budget_bytes;
  if (hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || c->ncu < 1)
    c->ncu = 256;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipHostMalloc((void **)&c->h_rb, kStageOff + kStageBytes, hipHostMallocCoherent | hipHostMallocMapped) !=
          hipSuccess ||
      hipHostGetDevicePointer((void **)&c->h_rb_dev, c->h_rb, 0) != hipSuccess) {
    if (c->h_rb) (void)hipHostFree(c->h_rb);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return WKG_EDEVICE;
  }
  memset(c->h_rb, 0, kStageOff);  // the signal flag starts below every sequence number
  c->noalloc_scratch = (uint64_t *)noalloc_scratch_of(device);
  if (!c->noalloc_scratch) {
    (void)hipHostFree(c->h_rb);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return WKG_EDEVICE;
  }
  size_t fr = 0, tot = 0;
  (void)hipMemGetInfo(&fr, &tot);
  c->pool = new Pool(table_budget_bytes ? table_budget_bytes : (uint64_t)(fr * 0.75));
  *out = c;
  return WKG_OK;
}

int wkg_destroy(wkg_ctx *ctx) {
  if (!ctx) return WKG_EINVAL;
  (void)hipSetDevice(ctx->device);
  ctx->sq_free();
  (void)hipDeviceSynchronize();
  ctx->free_graph();
  delete ctx->pool;
  if (ctx->h_rb) (void)hipHostFree(ctx->h_rb);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return WKG_OK;
}

const char *wkg_last_error(const wkg_ctx *ctx) { return ctx ? wkg_ctx::err_slot().c_str() : "null context"; }

int wkg_table_from_host(wkg_ctx *ctx, const int32_t *host, uint64_t rows, int cols, void *stream,
                        wkg_table **out) {
  if (!ctx || !out || cols < 0 || (rows && cols && !host)) return WKG_EINVAL;
  hipStream_t s = ctx->pick(stream);
  wkg_table *t;
  int rc = ctx->new_table(rows, cols, s, &t);
  if (rc) return rc;
  if (t->bytes) {
    hipError_t e = hipMemcpyAsync(t->d, host, t->bytes, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      ctx->drop_table(t);
      return ctx->fail(WKG_EDEVICE, "table upload failed: %s", hipGetErrorString(e));
    }
  }
  *out = t;
  return WKG_OK;
}

int wkg_table_to_host(wkg_ctx *ctx, const wkg_table *t, int32_t *host, void *stream) {
  if (!ctx || !t) return WKG_EINVAL;
  if (!t->bytes) return WKG_OK;
  if (!host) return ctx->fail(WKG_EINVAL, "null host buffer");
  hipStream_t s = ctx->pick(stream);
  if (t->bytes <= kStageBytes && t->bytes % 8 == 0)  // through the pinned staging area, polled
    return ctx->read_words((const uint64_t *)t->d, t->bytes / 8, s, (uint64_t *)host);
  if (t->bytes <= kStageBytes) {
    const Signal sg = ctx->next_signal();
    k_stage<<<1, kBlock, 0, s>>>(t->d, t->bytes / 4, (int32_t *)((char *)ctx->rb_dev() + kStageOff), sg);
    LAUNCHCHK(ctx);
    if (int rc = ctx->wait_signal(sg, s)) return rc;
    memcpy(host, (const char *)ctx->rb_host() + kStageOff, t->bytes);
    return WKG_OK;
  }
  HIPCHK(ctx, hipMemcpyAsync(host, t->d, t->bytes, hipMemcpyDeviceToHost, s));
  HIPCHK(ctx, hipStreamSynchronize(s));
  return WKG_OK;
}

int wkg_table_shape(const wkg_table *t, uint64_t *rows, int *cols) {
  if (!t) return WKG_EINVAL;
  if (rows) *rows = t->rows;
  if (cols) *cols = t->cols;
  return WKG_OK;
}

const int32_t *wkg_table_device_ptr(const wkg_table *t) { return t ? t->d : nullptr; }

int wkg_set_table_limit(wkg_ctx *ctx, uint64_t max_table_bytes) {
  if (!ctx) return WKG_EINVAL;
  ctx->table_limit = max_table_bytes;
  return WKG_OK;
}

int wkg_set_probe_tile_rows(wkg_ctx *ctx, int rows) {
  if (!ctx) return WKG_EINVAL;
  if (rows != 0 && rows != 256 && rows != 512 && rows != 1024 && rows != 2048)
    return ctx->fail(WKG_EINVAL, "probe tile rows must be 0, 256, 512, 1024 or 2048 (got %d)", rows);
  ctx->probe_tile = rows;
  return WKG_OK;
}

int wkg_thread_arena_begin(wkg_ctx *ctx, uint64_t bytes) {
  if (!ctx || bytes < 4096) return WKG_EINVAL;
  if (thread_arena()) return ctx->fail(WKG_ESTATE, "this thread already has an arena");
  hipStream_t s = ctx->pick(nullptr);
  char *base = (char *)ctx->pool->alloc(bytes, s, wkg_ctx::err_slot());
  if (!base) return WKG_ENOMEM;
  thread_arena() = new BumpArena{base, bytes, 0};
  return WKG_OK;
}

int wkg_thread_arena_reset(wkg_ctx *ctx) {
  if (!ctx) return WKG_EINVAL;
  BumpArena *b = thread_arena();
  if (!b) return ctx->fail(WKG_ESTATE, "this thread has no arena");
  HIPCHK(ctx, hipStreamSynchronize(ctx->pick(nullptr)));  // (a failed query may have left work queued)
  b->used = 0;
  return WKG_OK;
}

int wkg_thread_arena_end(wkg_ctx *ctx) {
  if (!ctx) return WKG_EINVAL;
  BumpArena *b = thread_arena();
  if (!b) return ctx->fail(WKG_ESTATE, "this thread has no arena");
  thread_arena() = nullptr;
  ctx->pool->release(b->base, ctx->pick(nullptr));
  delete b;
  return WKG_OK;
}

int wkg_pool_stats(wkg_ctx *ctx, uint64_t *live_bytes, uint64_t *cached_bytes, uint64_t *mallocs,
                   uint64_t *trims) {
  if (!ctx) return WKG_EINVAL;
  ctx->pool->stats(live_bytes, cached_bytes, mallocs, trims);
  return WKG_OK;
}

int wkg_set_lineage(wkg_ctx *ctx, int on) {
  if (!ctx) return WKG_EINVAL;
  ctx->lineage = on != 0;
  return WKG_OK;
}

int wkg_set_group_filter_rows(wkg_ctx *ctx, uint64_t rows) {
  if (!ctx) return WKG_EINVAL;
  ctx->group_filter_rows = rows;
  return WKG_OK;
}

int wkg_thread_context_begin(wkg_ctx *ctx) {
  if (!ctx) return WKG_EINVAL;
  ThreadCtx *&t = thread_ctx();
  if (t) return ctx->fail(WKG_ESTATE, "the calling thread already has a thread context");
  ThreadCtx *n = new (std::nothrow) ThreadCtx;
  if (!n) return ctx->fail(WKG_ENOMEM, "out of host memory");
  (void)hipSetDevice(ctx->device);
  if (hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) != hipSuccess ||
      hipHostMalloc((void **)&n->h_rb, kStageOff + kStageBytes, hipHostMallocCoherent | hipHostMallocMapped) !=
          hipSuccess ||
      hipHostGetDevicePointer((void **)&n->h_rb_dev, n->h_rb, 0) != hipSuccess) {
    if (n->h_rb) (void)hipHostFree(n->h_rb);
    if (n->stream) (void)hipStreamDestroy(n->stream);
    delete n;
    return ctx->fail(WKG_EDEVICE, "thread context set-up failed");
  }
  memset(n->h_rb, 0, kStageOff);  // the flag starts below every sequence number
  n->ctx = ctx;
  t = n;
  return WKG_OK;
}

int wkg_thread_context_end(wkg_ctx *ctx) {
  if (!ctx) return WKG_EINVAL;
  ThreadCtx *&t = thread_ctx();
  if (!t || t->ctx != ctx) return ctx->fail(WKG_ESTATE, "the calling thread has no context of its own here");
  (void)hipStreamSynchronize(t->stream);  // (its tables' memory returns to the pool with this stream)
  (void)hipStreamDestroy(t->stream);
  (void)hipHostFree(t->h_rb);
  delete t;
  t = nullptr;
  return WKG_OK;
}

int wkg_set_fanout_blocks(wkg_ctx *ctx, int on, uint64_t kept_cap) {
  if (!ctx) return WKG_EINVAL;
  ctx->fanout_blocks = on != 0;
  ctx->fanout_kept_cap = kept_cap ? kept_cap : (1ull << 22);
  return WKG_OK;
}

int wkg_table_slice(wkg_ctx *ctx, const wkg_table *t, uint64_t row0, uint64_t nrows, wkg_table **out) {
  if (!ctx || !t || !out) return WKG_EINVAL;
  if (row0 > t->rows || nrows > t->rows - row0)
    return ctx->fail(WKG_EINVAL, "rows [%llu, %llu) outside a %llu-row table", (unsigned long long)row0,
                     (unsigned long long)(row0 + nrows), (unsigned long long)t->rows);
  wkg_table *v = new (std::nothrow) wkg_table;
  if (!v) return ctx->fail(WKG_ENOMEM, "out of host memory");
  v->d = t->d ? t->d + row0 * (uint64_t)t->cols : nullptr;
  v->rows = nrows;
  v->cols = t->cols;
  v->bytes = nrows * (uint64_t)t->cols * sizeof(int32_t);
  v->stream = t->stream;
  v->owned = false;
  // the view keeps its table's cells alive (a child step's lineage may read them)
  v->slice_of = const_cast<wkg_table *>(t);
  v->slice_of->refs++;
  *out = v;
  return WKG_OK;
}

int wkg_table_wrap_device(wkg_ctx *ctx, const int32_t *cells, uint64_t rows, int cols, wkg_table **out) {
  if (!ctx || !out) return WKG_EINVAL;
  if (cols < 1 || cols > kBlock) return ctx->fail(WKG_EINVAL, "a table has 1..%d columns (got %d)", kBlock, cols);
  if (rows && !cells) return ctx->fail(WKG_EINVAL, "null cells for a %llu-row table", (unsigned long long)rows);
  wkg_table *v = new (std::nothrow) wkg_table;
  if (!v) return ctx->fail(WKG_ENOMEM, "out of host memory");
  v->d = const_cast<int32_t *>(cells);
  v->rows = rows;
  v->cols = cols;
  v->bytes = rows * (uint64_t)cols * sizeof(int32_t);
  v->stream = ctx->stream;
  v->owned = false;  // (and no owner: steps keep no lineage on it)
  *out = v;
  return WKG_OK;
}

int wkg_table_free(wkg_ctx *ctx, wkg_table *t) {
  if (!ctx) return WKG_EINVAL;
  ctx->drop_table(t);
  return WKG_OK;
}

int wkg_step_i2u(wkg_ctx *ctx, int idx_id, int dir, void *stream, wkg_table **out, uint64_t *expanded) {
  if (!ctx || !out) return WKG_EINVAL;
  int rc = ctx->check_pattern(idx_id, dir);
  if (rc) return rc;
  if (!is_tpid(idx_id)) return ctx->fail(WKG_EINVAL, "index id %d is not a predicate/type id", idx_id);
  return ctx->step_single_list(make_key(0, (uint64_t)idx_id, (uint64_t)dir), idx_id, ctx->pick(stream), out,
                               expanded);
}

int wkg_step_c2u(wkg_ctx *ctx, int64_t vid, int pid, int dir, void *stream, wkg_table **out,
                 uint64_t *expanded) {
  if (!ctx || !out) return WKG_EINVAL;
  int rc = ctx->check_pattern(pid, dir);
  if (rc) return rc;
  if (vid < 0 || (uint64_t)vid > kMaxVid) return ctx->fail(WKG_EINVAL, "constant %lld outside the vertex id range", (long long)vid);
  const uint64_t key = make_key((uint64_t)vid, (uint64_t)pid, (uint64_t)dir);
  return ctx->step_single_list(key, pid, ctx->pick(stream), out, expanded);
}

int wkg_step_k2u(wkg_ctx *ctx, const wkg_table *in, int start_col, int pid, int dir, void *stream,
                 wkg_table **out, uint64_t *expanded) {
  if (!ctx || !out) return WKG_EINVAL;
  return ctx->step_known(kModeK2U, in, start_col, pid, dir, -1, 0, ctx->pick(stream), out, expanded);
}

int wkg_step_k2u_noalloc(wkg_ctx *ctx, const int32_t *in, uint64_t rows, int cols, int start_col, int pid, int dir,
                         int32_t *out, uint64_t out_cap_rows, void *stream, uint64_t *out_rows, uint64_t *expanded) {
  if (!ctx || !out_rows) return WKG_EINVAL;
  *out_rows = 0;
  if (expanded) *expanded = 0;
  int rc = ctx->check_pattern(pid, dir);
  if (rc) return rc;
  if (ctx->framed)
    return ctx->fail(WKG_ESTATE, "the no-alloc step needs a resident graph (framed segments load on demand)");
  if (cols < 1 || start_col < 0 || start_col >= cols)
    return ctx->fail(WKG_EINVAL, "start column %d outside the %d-column table", start_col, cols);
  if (cols + 1 > kBlock) return ctx->fail(WKG_ERANGE, "tables are limited to %d columns", kBlock);
  if (rows >= (1ull << 32)) return ctx->fail(WKG_ERANGE, "tables are limited to 2^32 - 1 rows per step");
  if (rows && !in) return ctx->fail(WKG_EINVAL, "null input table");
  if (!ctx->noalloc_scratch) return ctx->fail(WKG_ESTATE, "no scratch for the no-alloc step on this device");
  if (rows == 0) return WKG_OK;
  hipStream_t s = ctx->pick(stream);
  // scratch: the probe passes' words | compacted rows (16 B per row) | the expand's split points
  auto up = [](uint64_t b) { return (b + 255) & ~uint64_t(255); };
  const int TR = ctx->probe_tile ? ctx->probe_tile : probe_tile_rows(rows, ctx->ncu);
  const uint64_t pass_b = up(pass_scratch_words(rows, TR) * 8), lin_b = up(rows * 16);
  if (pass_b + lin_b + 256 > WKG_NOALLOC_SCRATCH_BYTES)
    return ctx->fail(WKG_ERANGE, "a %llu-row input needs %llu bytes of step scratch, above the %llu of the no-alloc step",
                     (unsigned long long)rows, (unsigned long long)(pass_b + lin_b + 256),
                     (unsigned long long)WKG_NOALLOC_SCRATCH_BYTES);
  char *base = (char *)ctx->noalloc_scratch;
  wkg_table tin{};
  tin.d = const_cast<int32_t *>(in);
  tin.rows = rows;
  tin.cols = cols;
  tin.bytes = rows * (uint64_t)cols * 4;
  tin.stream = s;
  tin.owned = false;
  if ((rc = ctx->ensure_pred(pid, s))) return rc;
  ProbePass pp;
  if ((rc = ctx->probe_pass(kModeK2U, &tin, start_col, pid, dir, -1, 0, nullptr, s, pp, -1, (uint64_t *)base,
                            base + pass_b, TR)))
    return rc;
  const uint64_t E = pp.cnt[0], nz = pp.cnt[2];
  *out_rows = E;
  if (expanded) *expanded = E;
  const uint64_t part_b = ((nz + E + kExpandNV - 1) / kExpandNV + 1) * 8;
  if (E > out_cap_rows || (E && !out) || pass_b + lin_b + part_b > WKG_NOALLOC_SCRATCH_BYTES) {
    ctx->release_pass(pp, s);
    if (E > out_cap_rows)
      return ctx->fail(WKG_ERANGE, "the output has %llu rows; the buffer holds %llu", (unsigned long long)E,
                       (unsigned long long)out_cap_rows);
    if (E && !out) return ctx->fail(WKG_EINVAL, "null output buffer");
    return ctx->fail(WKG_ERANGE, "a %llu-row output needs %llu bytes of step scratch, above the %llu of the "
                     "no-alloc step", (unsigned long long)E, (unsigned long long)(pass_b + lin_b + part_b),
                     (unsigned long long)WKG_NOALLOC_SCRATCH_BYTES);
  }
  wkg_table tout{};
  tout.d = out;
  tout.rows = E;
  tout.cols = cols + 1;
  tout.bytes = E * (uint64_t)(cols + 1) * 4;
  tout.stream = s;
  tout.owned = false;
  rc = ctx->expand_pass(&tin, pid, pp.cinfo, pp.cexcl, nz, E, s, &tout, "expand", (uint64_t *)(base + pass_b + lin_b));
  ctx->release_pass(pp, s);
  return rc;
}

}  // extern "C"

// ---- small queries: the persistent server's host side ----
struct SQServer {
  std::mutex mu;
  hipStream_t stream = nullptr;
  hipEvent_t end_ev = nullptr;  // recorded after the last launch
  bool launched = false;
  char *h = nullptr;            // pinned: tail line, stop line, then the slots
  char *d = nullptr;            // the same block's device address
  uint64_t *d_head = nullptr;   // device memory: tickets claimed
  uint64_t tail = 0;            // tickets issued
  std::vector<uint8_t> busy;    // per slot: issued and not yet collected
  // host clock of the last launch and of the last ticket: within the idle
  // time of the last ticket and the life time of the launch a workgroup
  // surely runs (or is about to start), so no event query is needed
  std::chrono::steady_clock::time_point launch_t, submit_t;
  std::atomic<uint64_t> polls{0};  // unanswered polls (collect without waiting)
  double tick_hz = 1e8;         // the device wall clock
  uint64_t *tail_h() { return (uint64_t *)h; }
  uint64_t *stop_h() { return (uint64_t *)(h + 64); }
  SQSlot *slot_h(uint64_t t) { return (SQSlot *)(h + 128) + (t & (kSQSlots - 1)); }
  static constexpr uint64_t kSQSlots = 64;
};

namespace {

// A server workgroup exits after ctx->sq_idle_s without a query and takes
// no query after ctx->sq_life_s (the next query launches it again). Its
// stream may share a hardware queue with other streams of the process: work
// queued there waits at most that long behind it.

int sq_check_plan(wkg_ctx *ctx, const wkg_plan_step *steps, int n) {
  if (!ctx->loaded) return ctx->fail(WKG_ESTATE, "no graph loaded");
  if (n < 1 || n > (int)kSQSteps) return ctx->fail(WKG_ERANGE, "a small query has 1..%u steps", kSQSteps);
  if (ctx->framed) return ctx->fail(WKG_ERANGE, "small queries run on resident graphs");
  int c = 0;  // the plan's column bookkeeping, as the executor checks it
  for (int l = 0; l < n; l++) {
    const wkg_plan_step &d = steps[l];
    if (int rc = ctx->check_pattern(d.pid, d.dir)) return rc;
    const bool start = d.kind == WKG_STEP_I2U || d.kind == WKG_STEP_C2U;
    if (start != (l == 0)) return ctx->fail(WKG_EINVAL, "step %d: an index or constant start only at step 0", l);
    if (d.kind == WKG_STEP_I2U && !is_tpid(d.pid)) return ctx->fail(WKG_EINVAL, "index id %d is not a predicate/type id", d.pid);
    if (d.kind == WKG_STEP_C2U && (d.vid < 0 || (uint64_t)d.vid > kMaxVid))
      return ctx->fail(WKG_EINVAL, "constant %lld outside the vertex id range", (long long)d.vid);
    if (start) {
      c = 1;
      continue;
    }
    if (d.kind != WKG_STEP_K2U && d.kind != WKG_STEP_K2K && d.kind != WKG_STEP_K2C)
      return ctx->fail(WKG_EINVAL, "step %d: unknown kind %d", l, d.kind);
    if (d.start_col < 0 || d.start_col >= c) return ctx->fail(WKG_EINVAL, "step %d: start column %d outside %d", l, d.start_col, c);
    if (d.kind == WKG_STEP_K2K && (d.end_col < 0 || d.end_col >= c))
      return ctx->fail(WKG_EINVAL, "step %d: end column %d outside %d", l, d.end_col, c);
    if (d.kind == WKG_STEP_K2U) c++;
  }
  return WKG_OK;
}

std::tuple<int, int, int, int64_t> sq_start_key(const wkg_plan_step &d) {
  return {d.kind, d.pid, d.dir, d.kind == WKG_STEP_C2U ? d.vid : 0};
}

// A start list known to outgrow the limits: WKG_ERANGE at once.
int sq_known_big(wkg_ctx *ctx, const wkg_plan_step *steps) {
  std::lock_guard<std::mutex> lk(ctx->sq_big_mu);
  if (!ctx->sq_big_starts.count(sq_start_key(steps[0]))) return WKG_OK;
  ctx->prof_add_bytes("small_query_declined", 1.0);  // a count
  return ctx->fail(WKG_ERANGE, "the start list outgrows the small-query limits (%u cells)", kSQCells);
}

void sq_note_decline(wkg_ctx *ctx, const wkg_plan_step &start, uint64_t step) {
  if (step != 0) return;
  std::lock_guard<std::mutex> lk(ctx->sq_big_mu);
  if (ctx->sq_big_starts.size() >= 65536) ctx->sq_big_starts.clear();  // (bounded)
  ctx->sq_big_starts.insert(sq_start_key(start));
}

void sq_fill(wkg_ctx *ctx, const wkg_plan_step *steps, int n, SmallQuery &q) {
  memset(&q, 0, sizeof q);
  for (int l = 0; l < n; l++) {
    q.steps[l] = steps[l];
    q.meta[l] = ctx->h_meta[steps[l].pid];
  }
  q.nsteps = (uint32_t)n;
}

int sq_init(wkg_ctx *ctx) {
  static std::mutex init_mu;  // (two host threads may make a context's first small query)
  std::lock_guard<std::mutex> lk(init_mu);
  if (ctx->sqs) return WKG_OK;
  SQServer *v = new (std::nothrow) SQServer;
  if (!v) return ctx->fail(WKG_ENOMEM, "small-query server state");
  const size_t bytes = 128 + SQServer::kSQSlots * sizeof(SQSlot);
  int khz = 0;
  bool ok = hipHostMalloc((void **)&v->h, bytes, hipHostMallocCoherent | hipHostMallocMapped) == hipSuccess &&
            hipHostGetDevicePointer((void **)&v->d, v->h, 0) == hipSuccess &&
            hipMalloc((void **)&v->d_head, 256) == hipSuccess && hipMemset(v->d_head, 0, 256) == hipSuccess &&
            hipStreamCreateWithFlags(&v->stream, hipStreamNonBlocking) == hipSuccess &&
            hipEventCreateWithFlags(&v->end_ev, hipEventDisableTiming) == hipSuccess &&
            hipDeviceSynchronize() == hipSuccess;
  if (ok && hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx->device) == hipSuccess && khz > 0)
    v->tick_hz = std::min(1e9, std::max(1e6, khz * 1e3));
  if (!ok) {
    if (v->end_ev) (void)hipEventDestroy(v->end_ev);
    if (v->stream) (void)hipStreamDestroy(v->stream);
    if (v->d_head) (void)hipFree(v->d_head);
    if (v->h) (void)hipHostFree(v->h);
    delete v;
    return ctx->fail(WKG_EDEVICE, "small-query server set-up failed");
  }
  memset(v->h, 0, bytes);
  v->busy.assign(SQServer::kSQSlots, 0);
  // With HIP's default 4 hardware queues, other streams of the process are
  // likely to share the server's: keep its workgroups' life short then.
  if (wkg_hw_queues() < 8) ctx->sq_life_s = std::min(ctx->sq_life_s, 0.002);
  ctx->sqs = v;
  return WKG_OK;
}

// Whether the last launch still runs (v->mu held).
int sq_alive(wkg_ctx *ctx, SQServer *v, bool &alive) {
  alive = false;
  if (!v->launched) return WKG_OK;
  const hipError_t e = hipEventQuery(v->end_ev);
  if (e == hipErrorNotReady) {
    alive = true;
    return WKG_OK;
  }
  v->launched = false;
  if (e != hipSuccess) return ctx->fail(WKG_EDEVICE, "small-query server: %s", hipGetErrorString(e));
  return WKG_OK;
}

// Launches the server unless it runs (v->mu held).
int sq_ensure(wkg_ctx *ctx, SQServer *v) {
  const auto now = std::chrono::steady_clock::now();
  if (v->launched && now - v->submit_t < std::chrono::duration<double>(ctx->sq_idle_s / 2) &&
      now - v->launch_t < std::chrono::duration<double>(ctx->sq_life_s / 2))
    return WKG_OK;
  bool alive;
  if (int rc = sq_alive(ctx, v, alive)) return rc;
  if (alive) return WKG_OK;
  SQBox b;
  b.tail = (const uint64_t *)v->d;
  b.stop = (const uint64_t *)(v->d + 64);
  b.head = v->d_head;
  b.slots = (SQSlot *)(v->d + 128);
  b.nslots = SQServer::kSQSlots;
  b.idle_ticks = (uint64_t)(ctx->sq_idle_s * v->tick_hz);
  b.life_ticks = (uint64_t)(ctx->sq_life_s * v->tick_hz);
  const int nwg = std::max(1, std::min(ctx->sq_workgroups, ctx->ncu / 2));
  if (ctx->sorted) k_sq_server<true><<<nwg, kSQThreads, 0, v->stream>>>(ctx->view(), b);
  else k_sq_server<false><<<nwg, kSQThreads, 0, v->stream>>>(ctx->view(), b);
  const hipError_t le = hipGetLastError();
  if (le != hipSuccess) return ctx->fail(WKG_EDEVICE, "small-query server launch failed: %s", hipGetErrorString(le));
  HIPCHK(ctx, hipEventRecord(v->end_ev, v->stream));
  v->launched = true;
  v->launch_t = std::chrono::steady_clock::now();
  return WKG_OK;
}

int sq_submit(wkg_ctx *ctx, const wkg_plan_step *steps, int n, uint64_t *ticket) {
  if (int rc = sq_init(ctx)) return rc;
  SQServer *v = ctx->sqs;
  std::lock_guard<std::mutex> lk(v->mu);
  const uint64_t t = v->tail;
  if (v->busy[t & (SQServer::kSQSlots - 1)])
    return ctx->fail(WKG_EAGAIN, "all %llu small-query slots are in flight", (unsigned long long)SQServer::kSQSlots);
  SQSlot *sl = v->slot_h(t);
  sq_fill(ctx, steps, n, sl->q);
  v->busy[t & (SQServer::kSQSlots - 1)] = 1;
  __atomic_store_n(v->tail_h(), t + 1, __ATOMIC_RELEASE);  // the slot is written before the tail moves
  v->tail = t + 1;
  *ticket = t;
  const int rc = sq_ensure(ctx, v);
  v->submit_t = std::chrono::steady_clock::now();  // (after: a workgroup then runs until at least idle past it)
  return rc;
}

// Copies ticket t's answer out and frees its slot; WKG_EAGAIN while it is
// not answered (wait = 0).
int sq_collect(wkg_ctx *ctx, uint64_t t, bool wait, int32_t *cells, uint64_t cap, uint64_t *rows, int *cols,
               uint64_t *expanded) {
  SQServer *v = ctx->sqs;
  if (!v) return ctx->fail(WKG_EINVAL, "ticket %llu is not in flight", (unsigned long long)t);
  if (!wait && __atomic_load_n(&v->slot_h(t)->resp.done, __ATOMIC_ACQUIRE) != t + 1 &&
      (v->polls.fetch_add(1, std::memory_order_relaxed) & 63) != 63)
    return WKG_EAGAIN;  // not answered (every 64th such poll also checks that a server runs, below)
  {
    std::lock_guard<std::mutex> lk(v->mu);
    if (t >= v->tail || t + SQServer::kSQSlots < v->tail || !v->busy[t & (SQServer::kSQSlots - 1)])
      return ctx->fail(WKG_EINVAL, "ticket %llu is not in flight", (unsigned long long)t);
  }
  SQSlot *sl = v->slot_h(t);
  const auto t0 = std::chrono::steady_clock::now();
  for (uint32_t spins = 0; __atomic_load_n(&sl->resp.done, __ATOMIC_ACQUIRE) != t + 1; spins++) {
    if (!wait && spins == 0) {
      // not answered yet: make sure a server runs to answer it
      std::lock_guard<std::mutex> lk(v->mu);
      if (__atomic_load_n(&sl->resp.done, __ATOMIC_ACQUIRE) == t + 1) break;
      if (int rc = sq_ensure(ctx, v)) return rc;
      return WKG_EAGAIN;
    }
    if ((spins & 255) == 255) {
      std::lock_guard<std::mutex> lk(v->mu);
      // every workgroup of the last launch may have gone idle just before
      // this ticket arrived: a fresh launch takes it
      if (__atomic_load_n(&sl->resp.done, __ATOMIC_ACQUIRE) == t + 1) break;
      if (int rc = sq_ensure(ctx, v)) return rc;
      if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(10))
        return ctx->fail(WKG_EDEVICE, "small-query ticket %llu not answered in 10 s", (unsigned long long)t);
    }
  }
  const uint64_t status = sl->resp.status, r = sl->resp.rows, c = sl->resp.cols;
  if (expanded) *expanded = sl->resp.expanded;
  if (status != 0) sq_note_decline(ctx, sl->q.steps[0], sl->resp.step);
  if (ctx->profiling) {
    std::lock_guard<std::mutex> lk(ctx->prof_mu);
    wkg_ctx::KStat &k = ctx->prof_stats["small_query"];
    k.launches++;
    k.ms += sl->resp.ticks / v->tick_hz * 1e3;
    ctx->prof_stats[status ? "small_query_declined" : "small_query_taken"].bytes += 1.0;
  }
  int rc = WKG_OK;
  if (status != 0) {
    rc = ctx->fail(WKG_ERANGE, "a table outgrew the small-query limits (%u input rows, %u cells)", kSQRows, kSQCells);
  } else if (r * c > cap) {
    rc = ctx->fail(WKG_ERANGE, "a %llu-cell answer does not fit %llu cells", (unsigned long long)(r * c),
                   (unsigned long long)cap);
  } else {
    *rows = r;
    *cols = (int)c;
    if (r * c) memcpy(cells, sl->cells, r * c * 4);
  }
  std::lock_guard<std::mutex> lk(v->mu);
  v->busy[t & (SQServer::kSQSlots - 1)] = 0;
  return rc;
}

// One launch per query (no server).
int sq_one_launch(wkg_ctx *ctx, const wkg_plan_step *steps, int n, void *stream, int32_t *cells, uint64_t cap,
                  uint64_t *rows, int *cols, uint64_t *expanded) {
  hipStream_t s = ctx->pick(stream);
  SmallQuery q;
  sq_fill(ctx, steps, n, q);
  int32_t *stage = (int32_t *)((char *)ctx->rb_dev() + kStageOff);
  const GraphView g = ctx->view();
  const Signal sg = ctx->next_signal();
  {
    ProfScope ps(ctx, s, "small_query", 0.0);
    if (ctx->sorted) k_small_query<true, false><<<1, kSQThreads, 0, s>>>(g, q, stage, sg);
    else k_small_query<false, false><<<1, kSQThreads, 0, s>>>(g, q, stage, sg);
  }
  hipError_t le = hipGetLastError();
  if (le != hipSuccess) return ctx->fail(WKG_EDEVICE, "small query launch failed: %s", hipGetErrorString(le));
  if (int rc = ctx->wait_signal(sg, s)) return rc;
  const uint64_t *rb = ctx->rb_host();
  const uint64_t status = rb[0], r = rb[1], c = rb[2];
  if (expanded) *expanded = rb[3];
  if (status != 0) sq_note_decline(ctx, steps[0], rb[4]);
  ctx->prof_add_bytes(status ? "small_query_declined" : "small_query_taken", 1.0);  // counts
  if (status != 0)
    return ctx->fail(WKG_ERANGE, "a table outgrew the small-query limits (%u input rows, %u cells)", kSQRows, kSQCells);
  if (r * c > cap)
    return ctx->fail(WKG_ERANGE, "a %llu-cell answer does not fit %llu cells", (unsigned long long)(r * c),
                     (unsigned long long)cap);
  *rows = r;
  *cols = (int)c;
  if (r * c) memcpy(cells, (const char *)rb + kStageOff, r * c * 4);
  return WKG_OK;
}

}  // namespace

void wkg_ctx::sq_stop() {
  SQServer *v = sqs;
  if (!v) return;
  std::lock_guard<std::mutex> lk(v->mu);
  if (v->launched) {
    __atomic_store_n(v->stop_h(), 1, __ATOMIC_RELEASE);
    (void)hipStreamSynchronize(v->stream);
    __atomic_store_n(v->stop_h(), 0, __ATOMIC_RELEASE);
    v->launched = false;
  }
}

void wkg_ctx::sq_free() {
  sq_stop();
  SQServer *v = sqs;
  if (!v) return;
  (void)hipEventDestroy(v->end_ev);
  (void)hipStreamDestroy(v->stream);
  (void)hipFree(v->d_head);
  (void)hipHostFree(v->h);
  delete v;
  sqs = nullptr;
}

extern "C" {

int wkg_query_small(wkg_ctx *ctx, const wkg_plan_step *steps, int n, void *stream, int32_t *cells, uint64_t cap,
                    uint64_t *rows, int *cols, uint64_t *expanded) {
  if (!ctx || !steps || !rows || !cols || (cap && !cells)) return WKG_EINVAL;
  *rows = 0;
  *cols = 0;
  if (expanded) *expanded = 0;
  if (int rc = sq_check_plan(ctx, steps, n)) return rc;
  if (int rc = sq_known_big(ctx, steps)) return rc;
  if (ctx->sq_workgroups == 0) return sq_one_launch(ctx, steps, n, stream, cells, cap, rows, cols, expanded);
  uint64_t t;
  if (int rc = sq_submit(ctx, steps, n, &t)) return rc;
  return sq_collect(ctx, t, true, cells, cap, rows, cols, expanded);
}

int wkg_query_small_submit(wkg_ctx *ctx, const wkg_plan_step *steps, int n, uint64_t *ticket) {
  if (!ctx || !steps || !ticket) return WKG_EINVAL;
  if (int rc = sq_check_plan(ctx, steps, n)) return rc;
  if (int rc = sq_known_big(ctx, steps)) return rc;
  return sq_submit(ctx, steps, n, ticket);
}

int wkg_query_small_collect(wkg_ctx *ctx, uint64_t ticket, int wait, int32_t *cells, uint64_t cap, uint64_t *rows,
                            int *cols, uint64_t *expanded) {
  if (!ctx || !rows || !cols || (cap && !cells)) return WKG_EINVAL;
  *rows = 0;
  *cols = 0;
  if (expanded) *expanded = 0;
  return sq_collect(ctx, ticket, wait != 0, cells, cap, rows, cols, expanded);
}

int wkg_set_small_query_server(wkg_ctx *ctx, int workgroups) {
  if (!ctx || workgroups < 0) return WKG_EINVAL;
  ctx->sq_stop();  // the next query launches the new shape
  ctx->sq_workgroups = workgroups;
  return WKG_OK;
}

int wkg_get_small_query_server_times(const wkg_ctx *ctx, double *idle_s, double *life_s) {
  if (!ctx || !idle_s || !life_s) return WKG_EINVAL;
  *idle_s = ctx->sq_idle_s;
  *life_s = ctx->sq_life_s;
  return WKG_OK;
}

int wkg_hw_queues(void) {
  const char *hq = getenv("GPU_MAX_HW_QUEUES");
  const int n = hq ? atoi(hq) : 0;
  return n > 0 ? n : 4;
}

int wkg_set_small_query_server_times(wkg_ctx *ctx, double idle_s, double life_s) {
  if (!ctx || !(idle_s > 0) || !(life_s >= idle_s) || life_s > 10.0) return WKG_EINVAL;
  ctx->sq_stop();
  ctx->sq_idle_s = idle_s;
  ctx->sq_life_s = life_s;
  return WKG_OK;
}

int wkg_step_k2u_filter(wkg_ctx *ctx, const wkg_table *in, int start_col, int pid, int dir, int filter_kind,
                        int filter_pid, int filter_dir, int end_col, int64_t const_end, void *stream, wkg_table **out,
                        int *fused, uint64_t *expanded_k2u, uint64_t *expanded_filter) {
  if (!ctx || !out || !fused) return WKG_EINVAL;
  *out = nullptr;
  *fused = 0;
  uint64_t eu = 0, ef = 0;
  if (filter_kind != WKG_STEP_K2K && filter_kind != WKG_STEP_K2C)
    return ctx->fail(WKG_EINVAL, "the second step must be known_to_known or known_to_const (kind %d)", filter_kind);
  if (!in) return ctx->fail(WKG_EINVAL, "null input table");
  const int fmode = filter_kind == WKG_STEP_K2K ? kModeK2K : kModeK2C;
  const int rc = ctx->step_k2u_filter(in, start_col, pid, dir, fmode, filter_pid, filter_dir, end_col, const_end,
                                      ctx->pick(stream), out, fused, &eu, &ef);
  if (expanded_k2u) *expanded_k2u = eu;
  if (expanded_filter) *expanded_filter = ef;
  return rc;
}

int wkg_step_k2k(wkg_ctx *ctx, const wkg_table *in, int start_col, int pid, int dir, int end_col,
                 void *stream, wkg_table **out, uint64_t *expanded) {
  if (!ctx || !out) return WKG_EINVAL;
  return ctx->step_known(kModeK2K, in, start_col, pid, dir, end_col, 0, ctx->pick(stream), out, expanded);
}

int wkg_step_k2c(wkg_ctx *ctx, const wkg_table *in, int start_col, int pid, int dir, int64_t const_end,
                 void *stream, wkg_table **out, uint64_t *expanded) {
  if (!ctx || !out) return WKG_EINVAL;
  return ctx->step_known(kModeK2C, in, start_col, pid, dir, -1, const_end, ctx->pick(stream), out, expanded);
}

}  // extern "C"
