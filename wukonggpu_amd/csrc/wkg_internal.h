// Host-side state of libwkg_hip.so (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <list>
#include <map>
#include <mutex>
#include <set>
#include <tuple>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/wukong_gpu.h"
#include "device.h"

// Stream-aware caching allocator for tables and step scratch under an HBM
// budget. A freed block is reused only by later work on the stream that
// freed it (stream order makes that safe without a synchronisation); blocks
// cached for other streams are returned to HIP when the budget needs them.
// A per-thread bump region (wkg_thread_arena_*): while it is set, the
// calling thread's pool allocations are carved from it (no hipMalloc) and its
// releases inside it are no-ops; a reset drops them all at once.
struct BumpArena {
  char *base = nullptr;
  uint64_t cap = 0, used = 0;
};
inline BumpArena *&thread_arena() {
  static thread_local BumpArena *a = nullptr;
  return a;
}
// Input rows per probe / finalize launch of the calling thread's per-step
// calls (expands: rows / 64 tiles per launch); wkg_thread_launch_rows.
constexpr uint64_t kDefaultLaunchRows = 1ull << 26;
inline uint64_t &thread_launch_rows() {
  static thread_local uint64_t r = kDefaultLaunchRows;
  return r;
}

struct Pool {
  struct Cached {
    void *ptr;
    hipStream_t stream;
  };
  explicit Pool(uint64_t b) : budget(b) {}
  ~Pool();
  void *alloc(size_t bytes, hipStream_t s, std::string &err);
  void release(void *p, hipStream_t s);
  void trim();
  void stats(uint64_t *live, uint64_t *cached_b, uint64_t *nmalloc, uint64_t *ntrim);
  uint64_t budget;
  uint64_t live_bytes = 0, cached_bytes = 0, peak_bytes = 0;
  uint64_t mallocs = 0, trims = 0;  // hipMalloc calls; cache drops (each a device synchronisation)

 private:
  void trim_locked(uint64_t need);
  std::mutex mu;
  std::multimap<size_t, Cached> cached;
  std::unordered_map<void *, size_t> live;
};

struct wkg_table {
  int32_t *d;
  uint64_t rows;
  int cols;
  uint64_t bytes;
  hipStream_t stream;  // stream that produced it; its memory returns there
  bool owned = true;   // false: a row slice of another table (wkg_table_slice)
  wkg_table *slice_of = nullptr;  // a slice: the table whose cells it views (held by a reference)
  int refs = 1;        // the caller's handle + child tables whose lineage reads these cells
  // Lineage, set when this table is a known_to_unknown step's output: rows
  // [lin_excl[i], lin_excl[i+1]) (the last up to `rows`) repeat the parent's
  // row lin_info[i].x in their first parent->cols cells, i < lin_n.
  wkg_table *lin_parent = nullptr;
  const uint2 *lin_info = nullptr;
  const uint64_t *lin_excl = nullptr;
  uint64_t lin_n = 0;
  int lin_pid = 0;            // the step's predicate (lin_info[i].y is an offset in its segment)
  void *lin_alloc = nullptr;  // pool memory holding lin_info and lin_excl
  // Group lineage, set when this table is a known_to_unknown output over a
  // column its input inherited (step_k2u_lineage): group m < glin_n is
  // grandparent row glin_info[m].x joined with the list at segment offset
  // glin_info[m].y of predicate glin_pid (glin_s[m+1] - glin_s[m] entries,
  // the last up to glin_dtot), repeated for each child row, in rows
  // [glin_o[m], glin_o[m+1]) (the last up to `rows`).
  wkg_table *glin_parent = nullptr;
  const uint2 *glin_info = nullptr;
  const uint64_t *glin_s = nullptr, *glin_o = nullptr;
  uint64_t glin_n = 0, glin_dtot = 0;
  int glin_pid = 0;
  void *glin_alloc = nullptr;
};

// Returned by a specialised step path that declined the step (nothing was
// done; the general path runs it). Positive, so never a WKG_* status.
constexpr int kNotTaken = 1;

// One step's probe, tile scan and compaction (wkg_ctx::probe_pass).
struct ProbePass {
  uint64_t *scr = nullptr;    // step scratch
  void *lin = nullptr;        // known_to_unknown: cinfo | cexcl (may become the output's lineage)
  uint2 *cinfo = nullptr;     // compacted rows with edges: (row, list offset)
  uint64_t *cexcl = nullptr;  // their first output row
  uint64_t cnt[3] = {0, 0, 0};  // output rows, edges examined or keys probed, rows with output
  // the probe's kept-row lists and their tile prefixes (valid until release)
  const uint4 *hits = nullptr;
  const uint64_t *tile_nz = nullptr, *chunk_nz = nullptr;
  uint64_t ntiles = 0;
  int tr = 0;
  bool external = false;  // scr / lin are the caller's memory, not the pool's
};

// Framed residency (graph larger than its HBM budget): per-predicate
// segments in fixed power-of-two frames with least-recently-used eviction
// (ShardManager, core/shard_manager.hpp:44-557).
struct Residency {
  uint64_t vframe_buckets = 0, eframe_edges = 0;   // frame geometry
  uint64_t n_vframes = 0, n_eframes = 0;           // physical frames in the pools
  std::vector<uint64_t> vfree, efree;              // free physical frames
  std::vector<uint64_t> h_vframes, h_eframes;      // logical -> physical (or ~0)
  std::vector<uint64_t> vlen, elen;                // frames per predicate
  std::vector<uint64_t> used_v, used_e;            // loaded frames per predicate
  std::list<int> lru;                              // resident predicates, most recent first
  std::vector<std::list<int>::iterator> lru_pos;
  std::vector<char> in_lru;
  const uint64_t *h_slots = nullptr;               // the image on the host: a pinned copy, or the caller's
  const uint32_t *h_edges = nullptr;
  char *pinned = nullptr;                          // the pinned copy (owned), when it could be made
  hipStream_t copy_stream = nullptr;               // prefetches
  std::vector<hipEvent_t> ready;                   // per predicate: its prefetch's copies are done
  std::vector<char> pending;                       // per predicate: prefetched, not yet waited for
  uint64_t stat_prefetches = 0;
};

struct SQServer;
struct wkg_ctx;

// A thread's own stream and pinned read-back block on one context
// (wkg_thread_context_begin): that thread's per-step calls use them instead
// of the context's, so several threads can run queries at once.
struct ThreadCtx {
  const wkg_ctx *ctx = nullptr;
  hipStream_t stream = nullptr;
  uint64_t *h_rb = nullptr, *h_rb_dev = nullptr;
  uint64_t seq = 0;
};
inline ThreadCtx *&thread_ctx() {
  static thread_local ThreadCtx *t = nullptr;
  return t;
}

struct wkg_ctx {
  int device = 0;
  int ncu = 256;  // compute units of the device
  hipStream_t stream = nullptr;
  Pool *pool = nullptr;
  uint64_t budget_request = 0;
  uint64_t table_limit = 0;  // largest step output table in bytes (0: no limit)
  int probe_tile = 0;        // rows per probe tile of the per-step calls (0: by input size)
  bool lineage = true;       // probe inherited columns once per parent row
  uint64_t group_filter_rows = 1ull << 22;  // smallest input a filter decides per lineage group
  bool fanout_blocks = true;  // long shared lists: k_fanout_blocks (else k_fanout_units)
  uint64_t fanout_kept_cap = 1ull << 22;  // k_fanout_blocks' kept-row list (more: rerun on k_fanout_units)
  bool small_steps = true;   // steps over at most kSmallRows rows in one launch
  uint64_t *noalloc_scratch = nullptr;  // this device's no-alloc step scratch (wkg_step_k2u_noalloc)
  SQServer *sqs = nullptr;   // the persistent small-query server (created on first use)
  int sq_workgroups = 4;     // its workgroups (0: one launch per small query)
  double sq_idle_s = 200e-6, sq_life_s = 0.02;  // a server workgroup's idle exit and life
  // index / constant starts whose list alone outgrew the small-query limits
  // (the graph does not change while loaded: such a query is declined at once)
  std::set<std::tuple<int, int, int, int64_t>> sq_big_starts;
  std::mutex sq_big_mu;
  uint64_t *h_rb = nullptr;  // pinned read-back block: counters, the signal flag, small-table staging
  uint64_t *h_rb_dev = nullptr;  // the same block's device address
  uint64_t rb_seq = 0;       // the last read-back's signal value
  // The calling thread's last failure message (wkg_last_error): the served
  // path drives one context from two host threads, so it is kept per thread.
  static std::string &err_slot() {
    static thread_local std::string s;
    return s;
  }

  // graph
  bool loaded = false, framed = false, sorted = false;
  uint4 *d_slots = nullptr;
  uint32_t *d_edges = nullptr;
  wkg::DevMeta *d_meta = nullptr;
  uint64_t *d_vframes = nullptr, *d_eframes = nullptr;
  uint32_t vshift = 0, eshift = 0;
  uint32_t num_metas = 0;
  uint64_t num_slots = 0, num_edges = 0, num_buckets_total = 0, num_main = 0;
  std::vector<wkg_pred_meta_abi> metas;
  std::vector<wkg::DevMeta> h_meta;
  Residency *res = nullptr;
  uint64_t resident_bytes = 0, stat_loads = 0, stat_evictions = 0;

  // per-kernel timing with HIP events on the launch stream
  struct ProfRec {
    const char *name;
    hipEvent_t a, b;
    double bytes;
  };
  struct KStat {
    uint64_t launches = 0;
    double ms = 0, bytes = 0;
  };
  bool profiling = false;
  std::mutex prof_mu;
  std::vector<ProfRec> prof_pending;
  std::vector<hipEvent_t> prof_free;
  std::map<std::string, KStat> prof_stats;
  hipEvent_t prof_event();
  void prof_collect();
  void prof_add_bytes(const char *name, double bytes);  // bytes known only after a launch

  int fail(int code, const char *fmt, ...) __attribute__((format(printf, 3, 4)));
  void sq_stop();  // stops the small-query server (it holds the graph view it was launched with)
  void sq_free();
  wkg::GraphView view() const;
  hipStream_t pick(void *s) const;
  int new_table(uint64_t rows, int cols, hipStream_t s, wkg_table **out);
  void drop_table(wkg_table *t);
  void unref_table(wkg_table *t);
  int read_words(const uint64_t *d, uint64_t n, hipStream_t s, uint64_t *host);
  // the calling thread's read-back block (its own, or the context's)
  ThreadCtx *tctx() const {
    ThreadCtx *t = thread_ctx();
    return t && t->ctx == this ? t : nullptr;
  }
  uint64_t *rb_host() const {
    const ThreadCtx *t = tctx();
    return t ? t->h_rb : h_rb;
  }
  uint64_t *rb_dev() const {
    const ThreadCtx *t = tctx();
    return t ? t->h_rb_dev : h_rb_dev;
  }
  wkg::Signal next_signal() {
    ThreadCtx *t = tctx();
    return t ? wkg::Signal{t->h_rb_dev, ++t->seq} : wkg::Signal{h_rb_dev, ++rb_seq};
  }
  int wait_signal(const wkg::Signal &sg, hipStream_t s);
  int check_pattern(int pid, int dir);
  int step_single_list(uint64_t key, int pid, hipStream_t s, wkg_table **out, uint64_t *expanded);
  int step_known(int mode, const wkg_table *in, int sc, int pid, int dir, int ec, int64_t cval, hipStream_t s,
                 wkg_table **out, uint64_t *expanded);
  int step_k2u_filter(const wkg_table *in, int sc, int pid, int dir, int fmode, int fpid, int fdir, int ec,
                      int64_t cval, hipStream_t s, wkg_table **out, int *fused, uint64_t *ex_u, uint64_t *ex_f);
  int probe_pass(int mode, const wkg_table *in, int sc, int pid, int dir, int ec, int64_t cval, int32_t *O,
                 hipStream_t s, ProbePass &pp, int wcol = -1, uint64_t *ext_scr = nullptr, void *ext_lin = nullptr,
                 int tile_rows = 0);
  int step_small(int mode, const wkg_table *in, int sc, int pid, int dir, int ec, int64_t cval, hipStream_t s,
                 wkg_table **out, uint64_t *expanded);
  void release_pass(ProbePass &pp, hipStream_t s);
  int expand_pass(const wkg_table *in, int pid, const uint2 *cinfo, const uint64_t *cexcl, uint64_t nz, uint64_t E,
                  hipStream_t s, wkg_table *t, const char *tag = "expand", uint64_t *ext_part = nullptr);
  void keep_lineage(const wkg_table *in, ProbePass &pp, uint64_t nz, wkg_table *t, int pid);
  int step_filter_fanout(int mode, const wkg_table *in, int pid, int dir, int ec, int64_t cval, hipStream_t s,
                         wkg_table **out, uint64_t *expanded);
  int step_k2u_lineage(const wkg_table *in, int sc, int pid, int dir, hipStream_t s, wkg_table **out,
                       uint64_t *expanded);
  int step_filter_groups(int mode, const wkg_table *in, int pid, int dir, int ec, int64_t cval, hipStream_t s,
                         wkg_table **out, uint64_t *expanded);
  // residency
  int set_image_metas(uint64_t nslots, uint64_t nedges, const wkg_pred_meta_abi *m_in, int nmetas);
  int finish_resident(uint32_t flags, hipStream_t s);
  int ensure_pred(int pid, hipStream_t s, const std::vector<int> *keep = nullptr);
  int load_pred(int pid, hipStream_t s, const std::vector<int> *keep, bool prefetch);
  void free_graph();
};

// Records HIP events around the launches in its scope when profiling is on.
// `bytes` is the algorithmic byte count of the launch (for GB/s figures).
struct ProfScope {
  wkg_ctx *c;
  hipStream_t s;
  const char *name;
  double bytes;
  hipEvent_t a = nullptr;
  ProfScope(wkg_ctx *ctx, hipStream_t st, const char *n, double b) : c(ctx), s(st), name(n), bytes(b) {
    if (c->profiling) {
      a = c->prof_event();
      if (a) (void)hipEventRecord(a, s);
    }
  }
  ~ProfScope() {
    if (!a) return;
    hipEvent_t b = c->prof_event();
    if (!b) return;
    (void)hipEventRecord(b, s);
    std::lock_guard<std::mutex> lk(c->prof_mu);
    c->prof_pending.push_back({name, a, b, bytes});
  }
};
