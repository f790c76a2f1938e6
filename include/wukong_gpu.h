/*
 * wukong_gpu.h -- C ABI of libwkg_hip.so, the MI355X (gfx950) graph-exploration
 * engine for one server's gstore shard.
 *
 * This is the drop-in boundary below the reference's GPU engine. The
 * reference's host side (GPU_Engine, core/gpu_engine.hpp:171-467) drives
 * RCache (core/rcache.hpp:114-962), which launches the kernels of
 * core/gpu_hash.cu through C++ free functions (core/gpu_hash.h:18-146). Here
 * one C call replaces one RCache handler: key-list generation, the hash probe,
 * the edge-list fetch, the prefix sum and the result-table update all run
 * behind it, and the result stays on the device as a table handle until the
 * caller copies it out.
 *
 * Conventions
 *  - Every entry point returns 0 (WKG_OK) or a negative WKG_E* status; none
 *    aborts. wkg_last_error(ctx) describes the calling thread's most recent
 *    failure.
 *  - A result table is row-major int32, rows x cols, like
 *    request_or_reply::result_table (core/query.hpp:69). Row counts are 64-bit.
 *  - Column numbers are 0-based; the reference's variable -k is column k-1
 *    (var2column, core/query.hpp:151-154).
 *  - `stream` is a hipStream_t or NULL (the context's own stream). Calls that
 *    return a row count wait for the count on that stream; the table contents
 *    are ordered on the stream.
 *  - A context is used by one host thread at a time (the reference runs one
 *    GPU engine thread per GPU, core/wukong.cpp:320-325).
 */
#ifndef WUKONG_GPU_H
#define WUKONG_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WKG_OK 0
#define WKG_EINVAL (-1)    /* bad argument or unsupported pattern */
#define WKG_ENOMEM (-2)    /* HBM budget or device allocation exhausted */
#define WKG_EDEVICE (-3)   /* HIP runtime error */
#define WKG_ERANGE (-4)    /* a size exceeds a format limit */
#define WKG_ESTATE (-5)    /* call out of order (e.g. no graph loaded) */
#define WKG_EAGAIN (-6)    /* not finished yet / no free slot: call again */

typedef struct wkg_ctx wkg_ctx;
typedef struct wkg_table wkg_table;

/* pred_meta_t (core/rdf_meta.hpp:31-54), 64 bytes per predicate/type id. */
typedef struct wkg_pred_meta_abi {
  uint64_t partition_sz, main_hdr_start, main_hdr_end;
  uint64_t indrct_hdr_start, indrct_hdr_end, num_buckets;
  uint64_t edge_start, edge_end;
} wkg_pred_meta_abi;

/* ---- context ------------------------------------------------------------- */

/* Opens `device` and reserves nothing yet. table_budget_bytes caps the device
 * memory held by result tables and step scratch (0: 75% of free HBM after the
 * graph is loaded). Replaces the RCache constructor's fixed 20M-element
 * buffers (core/rcache.hpp:145-165). */
int wkg_init(int device, uint64_t table_budget_bytes, wkg_ctx **out);
int wkg_destroy(wkg_ctx *ctx);
const char *wkg_last_error(const wkg_ctx *ctx);

/* ---- graph residency (gstore segment cache) -------------------------------- */

#define WKG_LOAD_SORTED_LISTS 1u  /* caller guarantees every edge list ascends */

/* Uploads a gstore image: num_slots 16-byte {key, ptr} slots (2 u64 each),
 * num_edges u32 edges, num_metas pred_meta_t records. With
 * graph_budget_bytes == 0 or >= the image size, the whole image becomes
 * resident (the default on 288 GB of HBM3E). A smaller budget keeps the image
 * in pinned host memory and loads per-predicate segments into fixed-size
 * frames on demand, evicting least recently used predicates not needed by the
 * current step (ShardManager, core/shard_manager.hpp:44-557). The host arrays
 * must stay valid until wkg_destroy in the framed case (the segments are
 * copied from a pinned copy of them when one can be allocated). Without
 * WKG_LOAD_SORTED_LISTS the lists are checked once on the device; unsorted
 * lists make the known_to_known/known_to_const filters use a linear scan. */
int wkg_load_graph(wkg_ctx *ctx, const uint64_t *slots, uint64_t num_slots,
                   const uint32_t *edges, uint64_t num_edges,
                   const wkg_pred_meta_abi *metas, int num_metas,
                   uint64_t graph_budget_bytes, uint32_t flags);

/* Makes predicate/type `pid` resident (ShardManager::load_predicate,
 * core/shard_manager.hpp:473-555; check_pred_exist :351-359). A no-op when the
 * whole image is resident. When a prefetch of pid is under way, later work on
 * the context's streams waits for it on the device instead. */
int wkg_ensure_predicate(wkg_ctx *ctx, int pid);

/* 1 when pid's whole segment is in frames (or the image is resident), 0 when
 * not (ShardManager::check_pred_exist, core/shard_manager.hpp:351-359). */
int wkg_predicate_resident(const wkg_ctx *ctx, int pid);

/* Starts loading pid's segment on the context's copy stream and returns
 * (the pipelined load of a later pattern's predicate while predicate
 * `in_use` runs, ShardManager::load_predicate(pred, current, ..., async),
 * core/rcache.hpp:273-281, global_gpu_enable_pipeline). Frames are taken
 * now; neither `in_use` (< 0: none) nor the most recently used predicate is
 * evicted for them. The next wkg_ensure_predicate or step on pid waits for
 * the copies on the device. Copies run from a pinned copy of the host image,
 * made at wkg_load_graph when host memory allows (else they read the
 * caller's arrays and the host waits for them). A no-op for a resident image
 * or a predicate already resident or under way. */
int wkg_prefetch_predicate(wkg_ctx *ctx, int pid, int in_use);

/* Loads the listed predicates' segments now, in order (n < 0: every
 * predicate id from 1 up), as the reference does at start-up
 * (global_gpu_preload_predicts, core/rcache.hpp:191-217); later ones may
 * evict earlier ones when they do not all fit. Returns when the copies are
 * done. */
int wkg_preload_predicates(wkg_ctx *ctx, const int *pids, int n);

/* Prefetches started (wkg_prefetch_predicate that copied something). */
int wkg_graph_prefetches(const wkg_ctx *ctx, uint64_t *n);

/* Builds the gstore image on the device from n ID triples (host arrays, as
 * read by wks_triples_read_file) and makes it the context's resident graph:
 * the same image, slot for slot, that wks_build makes with the same
 * main_load (0: 0.75) -- the reference builds on the host (core/gstore.hpp:
 * 600-659, 665-870, 957-978; core/dgraph.hpp:268-330). *build_s (may be
 * NULL) gets the build time, upload included. */
int wkg_build_graph(wkg_ctx *ctx, const uint32_t *s, const uint32_t *p, const uint32_t *o, uint64_t n,
                    double main_load, double *build_s);
/* Sizes of the loaded image, and (resident graphs) a copy of it to the host:
 * num_slots x 2 u64, num_edges u32, num_metas pred metas. */
int wkg_graph_shape(const wkg_ctx *ctx, uint64_t *num_slots, uint64_t *num_edges, int *num_metas);
int wkg_graph_download(wkg_ctx *ctx, uint64_t *slots, uint32_t *edges, wkg_pred_meta_abi *metas);

/* Graph residency statistics: bytes resident, frame loads, evictions. */
int wkg_graph_stats(const wkg_ctx *ctx, uint64_t *resident_bytes,
                    uint64_t *segment_loads, uint64_t *segment_evictions);
/* 1 when the loaded graph is framed (segments loaded on demand under a graph
 * budget), 0 when the whole image is resident, negative on error. */
int wkg_graph_framed(const wkg_ctx *ctx);

/* ---- tables ------------------------------------------------------------------ */

/* Host rows -> device table (the first-step upload the reference requires
 * before a GPU handler, core/rcache.hpp:252-262). */
int wkg_table_from_host(wkg_ctx *ctx, const int32_t *host, uint64_t rows, int cols,
                        void *stream, wkg_table **out);
/* Device table -> host (the last-step D2H copy, core/rcache.hpp:401-412).
 * `host` must hold rows*cols int32. Tables up to 64 KB are copied by a
 * kernel into the context's pinned staging block, whose flag the host polls
 * (no stream synchronisation); larger ones by a copy command. */
int wkg_table_to_host(wkg_ctx *ctx, const wkg_table *t, int32_t *host, void *stream);
int wkg_table_shape(const wkg_table *t, uint64_t *rows, int *cols);
/* Device address of the table's first cell (row-major int32). */
const int32_t *wkg_table_device_ptr(const wkg_table *t);
/* A table over rows x cols device cells the caller owns -- the history
 * buffers the reference receives into (GPU::history_inbuf /
 * load_history_data, core/gpu.hpp:64-92; the GPUDirect receive of
 * core/rdma_transport.hpp:138-150) -- so a step can read a history that is
 * only a device pointer (request_or_reply::gpu_history_ptr,
 * core/query.hpp:63). The view never frees the cells; they must stay valid
 * until the view is freed and every step reading it has completed. */
int wkg_table_wrap_device(wkg_ctx *ctx, const int32_t *cells, uint64_t rows, int cols, wkg_table **out);
/* A known_to_unknown step of two or more input rows whose output table would
 * exceed max_table_bytes fails with WKG_ENOMEM before allocating it (0: no
 * limit, the default). The
 * engine then runs the step on row halves of its input, so a huge table is
 * produced and consumed in chunks of at most this size. */
int wkg_set_table_limit(wkg_ctx *ctx, uint64_t max_table_bytes);
/* Rows per probe tile of the per-step calls: 256, 512, 1024 or 2048, or 0
 * (the default) to choose by input size (the largest tile that leaves at
 * least four tiles per compute unit). A tuning knob with no counterpart in
 * the reference; results do not depend on it. */
int wkg_set_probe_tile_rows(wkg_ctx *ctx, int rows);
/* A known_to_unknown step that fans out (at least 4 output rows per input
 * row with edges) keeps, with its output table, which input row each output
 * row repeats. A later known_to_unknown step whose start column is one of
 * those inherited columns (a star's next arm from the same variable) then
 * probes once per input row instead of once per output row; its output is
 * the same, row for row. On by default; 0 turns it off. No counterpart in the
 * reference, which probes every row (core/gpu_hash.cu:145-316). */
int wkg_set_lineage(wkg_ctx *ctx, int on);
/* With lineage on, the output of such a once-per-input-row step also keeps
 * its groups (an input row's list, joined to each of that row's children).
 * A known_to_known / known_to_const step from that output's last column
 * (end column, if any, one of the input's) of at least `rows` rows (default
 * 2^22; 0: any size), and at least 4 per list entry of its groups, decides
 * each group's list entry once and reads the decision back for every row
 * holding it; its output and work count are the per-row filter's. No
 * counterpart in the reference (core/gpu_hash.cu:320-466 tests every row). */
int wkg_set_group_filter_rows(wkg_ctx *ctx, uint64_t rows);
/* Such a filter over long lists shared by many parents runs a block of a
 * list's entries against 64 of its parents per workgroup and lists the kept
 * rows (at most kept_cap of them, 0: 2^22; with more it reruns with one
 * parent per workgroup). On by default; 0 always runs one parent per
 * workgroup. The output and work count are the same either way. */
int wkg_set_fanout_blocks(wkg_ctx *ctx, int on, uint64_t kept_cap);
/* The calling thread's per-step calls on ctx (stream NULL) get a stream and
 * a pinned read-back block of their own until _end, so several threads can
 * run per-step calls on one context at once (a graph that is resident;
 * framed residency is not thread-safe). Tables it made stay valid after
 * _end. No counterpart in the reference (one engine thread per GPU). */
int wkg_thread_context_begin(wkg_ctx *ctx);
int wkg_thread_context_end(wkg_ctx *ctx);
/* A known-start step over at most 512 input rows runs its probe, scan,
 * compaction and (up to 2048 output rows) expand in one launch instead of
 * five; its output is the same, row for row. On by default; 0 turns it off.
 * A launch-count optimisation with no counterpart in the reference. */
int wkg_set_small_steps(wkg_ctx *ctx, int on);
/* Input rows per launch of the calling thread's per-step probe and
 * compaction passes (the expand: rows / 64 tiles of 512 items per launch);
 * 0 restores the default, 2^26. The GPU runs another queue's kernels only
 * between this queue's kernels, so smaller launches let concurrent work (a
 * served light batch, another process) in sooner, at a few microseconds per
 * extra launch. A scheduling knob with no counterpart in the reference;
 * results do not depend on it. */
int wkg_thread_launch_rows(wkg_ctx *ctx, uint64_t rows);
/* Per-thread arena: until wkg_thread_arena_end, the tables and step scratch
 * that the calling thread's per-step calls allocate are carved from `bytes`
 * of pool memory reserved now, with no hipMalloc; wkg_thread_arena_reset
 * drops them all at once (call it only when every table the thread made is
 * freed and its work has completed, e.g. after a query's last table reached
 * the host). An allocation that does not fit takes pool memory as usual.
 * For a worker thread that runs one query at a time beside other streams. */
int wkg_thread_arena_begin(wkg_ctx *ctx, uint64_t bytes);
int wkg_thread_arena_reset(wkg_ctx *ctx);
int wkg_thread_arena_end(wkg_ctx *ctx);
/* Table/scratch pool: bytes live and cached, hipMalloc calls, and cache drops
 * (each drop synchronises the device and frees every cached block). */
int wkg_pool_stats(wkg_ctx *ctx, uint64_t *live_bytes, uint64_t *cached_bytes, uint64_t *mallocs,
                   uint64_t *trims);
/* A view of rows [row0, row0 + nrows) of t, sharing its memory (for running
 * a step over a table in row chunks). Free the view before t. */
int wkg_table_slice(wkg_ctx *ctx, const wkg_table *t, uint64_t row0, uint64_t nrows, wkg_table **out);
int wkg_table_free(wkg_ctx *ctx, wkg_table *t);

/* Fork-join split (GPU_Engine::generate_sub_query, core/gpu_engine.hpp:293-350,
 * over RCache::generate_sub_query, core/rcache.hpp:964-1043, and the kernels
 * of core/gpu_hash.cu:698-873): row i of `in` goes to sub-table
 * in[i][col] mod nparts, the cell read as an unsigned 32-bit vertex id (the
 * owner rule vid % servers, utils/mymath.hpp:55-59). *out holds the
 * sub-tables back to back: sub-table p is part_rows[p] rows starting at row
 * part_head[p] (part_head may be NULL); wkg_table_slice gives views of them.
 * Inside a sub-table rows keep their input order (the reference's atomic
 * slots allow any order; this is the stable one). 1 <= nparts <= 1024. */
int wkg_table_split(wkg_ctx *ctx, const wkg_table *in, int col, int nparts, void *stream, wkg_table **out,
                    uint64_t *part_rows, uint64_t *part_head);
/* Merge of sub-replies (RMap::put_reply, core/reply_map.hpp:43-80): the
 * parts' rows appended in the order given. Empty parts are skipped; the
 * others must have the same column count. */
int wkg_table_concat(wkg_ctx *ctx, int n, const wkg_table *const *parts, void *stream, wkg_table **out);

/* ---- pattern steps (one RCache handler each) -------------------------------------
 * `expanded` (may be NULL) receives the step's work count: edges emitted for
 * index/const/known_to_unknown, edges examined by the reference's early-exit
 * linear scan for known_to_known/known_to_const. A step returns once its
 * row count is known; its output table is complete in the order of its
 * stream (later work there sees it; other streams must wait for it). */

/* index_to_unknown: rows = the index list of key (0, idx_id, dir).
 * RCache::index_to_unknown, core/rcache.hpp:811-962. */
int wkg_step_i2u(wkg_ctx *ctx, int idx_id, int dir, void *stream, wkg_table **out,
                 uint64_t *expanded);

/* const_to_unknown: rows = the edge list of key (vid, pid, dir). The reference
 * runs this on the host (GPU_Engine::const_to_unknown,
 * core/gpu_engine.hpp:171-189) and cannot continue on the GPU from it
 * (core/rcache.hpp:252-262); here it is a device probe. */
int wkg_step_c2u(wkg_ctx *ctx, int64_t vid, int pid, int dir, void *stream,
                 wkg_table **out, uint64_t *expanded);

/* known_to_unknown: for each row in order, one output row per edge of key
 * (row[start_col], pid, dir), in edge-list order, with the neighbour appended
 * as the last column. RCache::known_to_unknown, core/rcache.hpp:232-414. */
int wkg_step_k2u(wkg_ctx *ctx, const wkg_table *in, int start_col, int pid, int dir,
                 void *stream, wkg_table **out, uint64_t *expanded);

/* known_to_unknown that allocates nothing. No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc). The
 * result is wkg_step_k2u's, row for row. `in` is rows x cols device cells
 * (row-major int32) and `out` a device buffer of out_cap_rows x (cols + 1)
 * cells, both the caller's. The step's scratch (tile sums, compacted rows,
 * split points) is a fixed device array of WKG_NOALLOC_SCRATCH_BYTES per
 * GPU, set up by wkg_init and held until the process ends; an input or
 * output too big for it fails with WKG_ERANGE before any launch or before
 * the expand (about 32 bytes per input row plus 1 byte per 64 output rows).
 * *out_rows gets the output's row count; when it exceeds out_cap_rows the
 * call fails with WKG_ERANGE and writes nothing to `out`. The graph must be
 * resident (WKG_ESTATE when framed). Every call on one GPU shares that GPU's
 * scratch: calls that overlap on one GPU (other streams, other threads,
 * other contexts) race; calls on one stream are ordered. The output is
 * complete in stream order. */
#define WKG_NOALLOC_SCRATCH_BYTES (64ull << 20)
int wkg_step_k2u_noalloc(wkg_ctx *ctx, const int32_t *in, uint64_t rows, int cols, int start_col, int pid, int dir,
                         int32_t *out, uint64_t out_cap_rows, void *stream, uint64_t *out_rows,
                         uint64_t *expanded);

/* known_to_known: keep the rows whose row[end_col] is in the edge list of
 * (row[start_col], pid, dir). RCache::known_to_known, core/rcache.hpp:417-612. */
/* known_to_unknown (as wkg_step_k2u) followed by a filter whose start
 * column is the new column: filter_kind WKG_STEP_K2K with end_col an input
 * column, or WKG_STEP_K2C with const_end, on filter_pid / filter_dir. When
 * the filter can decide the fan-out's rows from its lineage (a parent row
 * and a list entry each) the fan-out is never written: *fused = 1 and *out
 * is the filter's output. Otherwise *fused = 0 and *out is the
 * known_to_unknown output (the caller runs the filter on it). Either way
 * the tables and work counts are those of the two calls made one after the
 * other (core/rcache.hpp:232-414, then the filter's handler); this call has
 * no counterpart in the reference, it fuses two of its handlers. */
int wkg_step_k2u_filter(wkg_ctx *ctx, const wkg_table *in, int start_col, int pid, int dir, int filter_kind,
                        int filter_pid, int filter_dir, int end_col, int64_t const_end, void *stream, wkg_table **out,
                        int *fused, uint64_t *expanded_k2u, uint64_t *expanded_filter);
int wkg_step_k2k(wkg_ctx *ctx, const wkg_table *in, int start_col, int pid, int dir,
                 int end_col, void *stream, wkg_table **out, uint64_t *expanded);

/* known_to_const: keep the rows whose edge list of (row[start_col], pid, dir)
 * holds const_end. RCache::known_to_const, core/rcache.hpp:614-809. */
int wkg_step_k2c(wkg_ctx *ctx, const wkg_table *in, int start_col, int pid, int dir,
                 int64_t const_end, void *stream, wkg_table **out, uint64_t *expanded);

/* ---- batched steps ------------------------------------------------------------------
 * Many independent (table, pattern) steps in one fused launch sequence: one
 * probe/scan pass over the rows of all inputs and one expand/compact pass over
 * all outputs. Each descriptor is one reference handler call. The reference
 * runs one query at a time on the GPU (GPU::query_id CAS, core/gpu.hpp:53-61);
 * this is the primitive under batched and served execution. */
#define WKG_STEP_I2U 0
#define WKG_STEP_C2U 1
#define WKG_STEP_K2U 2
#define WKG_STEP_K2K 3
#define WKG_STEP_K2C 4

typedef struct wkg_step_desc {
  int32_t kind;       /* WKG_STEP_* */
  int32_t pid;        /* predicate (or index id for I2U) */
  int32_t dir;        /* 0 = IN, 1 = OUT */
  int32_t start_col;  /* K2*: column of the start variable */
  int32_t end_col;    /* K2K: column of the end variable */
  int32_t pad;
  int64_t vid;        /* C2U: start constant; K2C: end constant */
  const wkg_table *in; /* K2*: input table */
} wkg_step_desc;

/* Runs n steps; out[i] receives a new table, expanded[i] (may be NULL) its
 * work count. Inputs are not modified or freed. */
int wkg_step_batch(wkg_ctx *ctx, int n, const wkg_step_desc *steps, void *stream,
                   wkg_table **out, uint64_t *expanded);

/* ---- device-resident batched execution -----------------------------------------------
 * Whole queries, not single steps: every query of a batch advances through
 * all of its pattern steps on the device, with no host round trip between
 * steps. A query is a compiled plan, one wkg_plan_step per pattern (the step
 * GPU_Engine::execute_one_step would dispatch, core/gpu_engine.hpp:395-467).
 * Per step level, a single-block plan kernel sizes every query's work and
 * carves its tables and scratch out of an HBM arena; persistent kernels then
 * take the probe, finalize and expand tiles of all queries from one shared
 * work queue (an atomic tile ticket over the concatenated tile space). The
 * whole sequence -- plan upload, every level, the retirement of finished
 * queries' tables and the read-back of the result records -- is captured once in a hipGraph
 * and replayed per batch. Results are the same, cell for cell, as the plain
 * per-step calls give. A query whose tables do not fit the arena ends with
 * status WKG_ENOMEM and can be rerun on the plain path. */
typedef struct wkg_batch wkg_batch;

typedef struct wkg_plan_step {
  int32_t kind;      /* WKG_STEP_* (I2U / C2U only as a query's first step) */
  int32_t pid, dir;  /* predicate (index id for I2U), 0 = IN / 1 = OUT */
  int32_t start_col, end_col;
  int32_t pad;
  int64_t vid;       /* C2U start constant / K2C end constant */
} wkg_plan_step;

/* ---- small queries: one workgroup each, a persistent server ---------------------------
 * Runs every step of a compiled plan (wkg_plan_step, above; at most 8 steps,
 * the first an index or constant start) in one 1024-thread workgroup with
 * the tables in LDS, while every table stays within 4096 input rows per step
 * and 8192 cells; the result is the per-step calls' table, cell for cell.
 * The reference runs a query's steps one handler at a time
 * (GPU_Engine::execute_request, core/gpu_engine.hpp:471-526); for a query
 * this small each step costs launch and read-back latency, not work.
 *
 * By default the queries go to a persistent server: a kernel of a few
 * workgroups (wkg_set_small_query_server) that takes queries from a ring of
 * 64 slots in pinned host memory as they arrive, with no launch per query,
 * and exits by itself after 200 us without a query (the next query launches
 * it again). With 0 workgroups each query is one launch instead.
 *
 * wkg_query_small runs one query and waits: WKG_OK with the table copied
 * into cells (cap cells; WKG_SMALL_QUERY_CELLS always suffices), or
 * WKG_ERANGE when a table outgrew the limits (or the graph is framed):
 * then nothing is returned and the caller runs the query step by step.
 * An index or constant start whose list alone outgrew the limits is
 * remembered while the graph is loaded: the next query from it is declined
 * at once, with no launch.
 * wkg_query_small_submit / _collect are the same, split: submit returns a
 * ticket at once (WKG_EAGAIN when all 64 slots are in flight; WKG_ERANGE for
 * a start remembered as too big), collect
 * returns its answer, or WKG_EAGAIN while it runs when wait is 0. Every
 * submitted ticket must be collected once. A context's calls may come from
 * several host threads. */
#define WKG_SMALL_QUERY_CELLS 8192
int wkg_query_small(wkg_ctx *ctx, const wkg_plan_step *steps, int n, void *stream, int32_t *cells, uint64_t cap,
                    uint64_t *rows, int *cols, uint64_t *expanded);
int wkg_query_small_submit(wkg_ctx *ctx, const wkg_plan_step *steps, int n, uint64_t *ticket);
int wkg_query_small_collect(wkg_ctx *ctx, uint64_t ticket, int wait, int32_t *cells, uint64_t cap, uint64_t *rows,
                            int *cols, uint64_t *expanded);
/* Server workgroups (default 4, at most half the CUs; 0: one launch per
 * query). Stops a running server; the next query starts the new shape. */
int wkg_set_small_query_server(wkg_ctx *ctx, int workgroups);
/* A server workgroup exits after idle_s without a query (default 200 us)
 * and takes no query after life_s (default 20 ms, at most 10 s); the next
 * query launches it again. Work on a stream that shares the server's
 * hardware queue waits behind it up to that long: give the process enough
 * hardware queues for its streams (environment GPU_MAX_HW_QUEUES, read when
 * the HIP runtime starts; HIP's default is 4, and with fewer than 8 the
 * server's first use lowers life_s to 2 ms). */
int wkg_set_small_query_server_times(wkg_ctx *ctx, double idle_s, double life_s);
int wkg_get_small_query_server_times(const wkg_ctx *ctx, double *idle_s, double *life_s);
/* The process's hardware queues per device as HIP reads them
 * (GPU_MAX_HW_QUEUES; 4 when unset). */
int wkg_hw_queues(void);

/* Launch and tile shapes (0 = default). The autotuner below picks them. */
typedef struct wkg_batch_tuning {
  int32_t probe_wg_per_cu;   /* persistent probe / finalize grid, workgroups per CU */
  int32_t expand_wg_per_cu;  /* persistent expand grid */
  int32_t aux_wg_per_cu;     /* tile scan, merge-path partition, gather */
  int32_t probe_tile_rows;   /* rows per probe tile: 1024 or 2048 */
  int32_t expand_tile_items; /* merged items per expand tile: 512, 1024 or 2048 */
  int32_t schedule;          /* probe and expand tiles: 0 drawn from a shared work
                                queue, 1 dealt to the workgroups statically */
} wkg_batch_tuning;

typedef struct wkg_batch_result {
  uint64_t offset;   /* first cell of the query's table in the returned cells */
  uint64_t rows;
  int32_t cols;
  int32_t status;    /* WKG_OK, or WKG_ENOMEM when the arena was full */
  uint64_t expanded; /* work count summed over the steps */
} wkg_batch_result;

/* max_queries and max_patterns bound a batch; arena_bytes of HBM (taken from
 * the table budget) hold the intermediate tables and step scratch. Level l's
 * tables go to one of two stacks by the parity of l, so only a level's
 * inputs, outputs and scratch are live at once; a query's last table is
 * retired to the host as soon as its last step ends. stream may be NULL (a
 * stream of the batch's own). */
int wkg_batch_create(wkg_ctx *ctx, int max_queries, int max_patterns, uint64_t arena_bytes,
                     void *stream, wkg_batch **out);
int wkg_batch_destroy(wkg_batch *b);
int wkg_batch_set_tuning(wkg_batch *b, const wkg_batch_tuning *t);
int wkg_batch_get_tuning(const wkg_batch *b, wkg_batch_tuning *t);
/* Starts nq <= max_queries queries and returns at once; query q's steps are
 * steps[step_off[q] .. step_off[q+1]). The arrays are copied. */
int wkg_batch_launch(wkg_batch *b, int nq, const wkg_plan_step *steps, const int32_t *step_off);
/* 1 while the launched batch runs, 0 once it is done. */
int wkg_batch_busy(wkg_batch *b);
/* Waits for the launched batch. res[nq] gets the records; *cells points at
 * the batch's result buffer (pinned host memory the retire kernels write
 * directly), which holds every result table at its record's offset and stays
 * valid until the next launch on this batch. A query whose table did not fit
 * the buffer ends with WKG_ENOMEM; the buffer grows for the next launch. */
int wkg_batch_collect(wkg_batch *b, wkg_batch_result *res, const int32_t **cells, uint64_t *ncells);
/* launch + collect */
int wkg_batch_run(wkg_batch *b, int nq, const wkg_plan_step *steps, const int32_t *step_off,
                  wkg_batch_result *res, const int32_t **cells, uint64_t *ncells);
/* Sets the result buffer to `bytes` (default 64 MiB; it also grows by itself
 * after a batch whose results overflowed it). */
int wkg_batch_reserve_results(wkg_batch *b, uint64_t bytes);
/* Of the last collected batch: the arena's high-water mark and the result
 * bytes (for admission against a memory budget). */
int wkg_batch_stats(const wkg_batch *b, uint64_t *arena_peak, uint64_t *result_bytes);
/* Autotuner: times `reps` runs of the given batch under each candidate
 * launch/tile shape, keeps the fastest (median) and returns it. */
int wkg_batch_autotune(wkg_batch *b, int nq, const wkg_plan_step *steps, const int32_t *step_off,
                       int reps, wkg_batch_tuning *best, double *best_ms);

/* ---- measurement ---------------------------------------------------------------------
 * The reference has printf timers around asynchronous launches only
 * (core/rcache.hpp:300-385, BREAK_DOWN). Here each kernel launch can be
 * bracketed by HIP events on its own stream. Kernel names: "probe",
 * "probe_filter", "scan_tiles", "finalize", "compact", "partition", "expand",
 * "probe_one", "copy_list", "batch_probe", "batch_scan", "batch_finalize",
 * "batch_partition", "batch_expand", "copy_calibrate".
 * alg_bytes sums each launch's algorithmic byte count. */
int wkg_profile_enable(wkg_ctx *ctx, int on);
int wkg_profile_read(wkg_ctx *ctx, const char *kernel, uint64_t *launches, double *total_ms,
                     double *alg_bytes);
/* Device-to-device copy bandwidth of a plain 16-byte-per-lane copy kernel
 * over `bytes` (read + written bytes per second). */
int wkg_bandwidth_probe(wkg_ctx *ctx, uint64_t bytes, int iters, double *copy_gbps);
/* Random 128-byte line reads over a bytes-sized buffer (8 lanes per line, 8
 * lines in flight per lane group, as the hash probe reads buckets): the
 * random-access roofline of the probe, in lines per second. */
int wkg_random_access_probe(wkg_ctx *ctx, uint64_t bytes, int iters, double *lines_per_s);

#ifdef __cplusplus
}
#endif
#endif
