/*
 * wukong_engine.h -- C ABI of libwkg_engine.so, the host engine above
 * libwkg_hip.so.
 *
 * The engine keeps the reference GPU engine's request model and dispatch:
 * a query is cmd_chains, 4 x int64 per pattern (start, predicate, direction,
 * end) (core/query.hpp:68); variables are -1, -2, ... and variable -k is
 * column k-1 (core/query.hpp:151-154); a variable is known once its column
 * exists (core/query.hpp:141-149); step 0 starts from an index when
 * cmd_chains[0] is a predicate/type id (core/query.hpp:120-139); every other
 * step dispatches on (type(start), type(end)) exactly as
 * GPU_Engine::execute_one_step does (core/gpu_engine.hpp:395-467).
 * Intermediate tables stay on the device between steps; the last step's
 * table is copied to the host (core/rcache.hpp:389-412).
 *
 * Status codes are those of wukong_gpu.h.
 */
#ifndef WUKONG_ENGINE_H
#define WUKONG_ENGINE_H

#include <stdint.h>

#include "wukong_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct wke_engine wke_engine;

/* Result of one query: row-major int32 table (malloc'd; free with wke_free). */
typedef struct wke_result {
  int32_t *table;
  uint64_t rows;
  int32_t cols;
  int32_t status;     /* WKG_OK or the error of the failing step */
  uint64_t expanded;  /* work count summed over the steps */
  double latency_s;   /* submit -> host table */
} wke_result;

int wke_create(wkg_ctx *ctx, wke_engine **out);
int wke_destroy(wke_engine *e);
const char *wke_last_error(const wke_engine *e);
void wke_free(void *p);

/* Plain path: one query, one step at a time (GPU_Engine::execute_request,
 * core/gpu_engine.hpp:471-526, collapsed to one server). A step whose tables
 * do not fit the context's table budget runs on the row halves of its input,
 * each carried depth-first through the rest of the query (HBM-budgeted
 * chunks); the result is the same table an unchunked run gives. */
int wke_run_query(wke_engine *e, const int64_t *cmd, int ncmd, wke_result *res);

/* The plain path first offers an index- or constant-start query of at most
 * 8 patterns to wkg_query_small (all steps in one launch while its tables
 * stay small) and runs it step by step when that declines. On by default; 0
 * turns it off. The result is the same either way. */
int wke_set_small_queries(wke_engine *e, int on);

/* Row-half splits made so far by the plain path's chunking. */
int wke_stats(const wke_engine *e, uint64_t *splits);

/* Plain path over many queries, one after another (the reference runs one
 * query on the GPU at a time, core/gpu.hpp:53-61). */
int wke_run_plain(wke_engine *e, const int64_t *cmd, const int64_t *cmd_off, int nq, wke_result *res);

/* Batched path: the queries are compiled into device plans and run by the
 * device-resident executor (wkg_batch_*, include/wukong_gpu.h): every query
 * advances through all its steps on the device inside one captured launch
 * sequence, max_queries at a time. A query that does not fit the executor's
 * arena is rerun on the plain path, so every result matches wke_run_query. */
int wke_run_batched(wke_engine *e, const int64_t *cmd, const int64_t *cmd_off, int nq, wke_result *res);

/* wke_run_batched with every result table in one buffer: *cells (malloc'd,
 * release with wke_free) holds *ncells cells; query q's table starts at
 * cell_off[q] (res[q].table stays NULL). */
int wke_run_batched_packed(wke_engine *e, const int64_t *cmd, const int64_t *cmd_off, int nq, wke_result *res,
                           uint64_t *cell_off, int32_t **cells, uint64_t *ncells);

/* Step-synchronous batched path: all queries advance one step per round, the
 * current steps of all unfinished queries fused into one wkg_step_batch. */
int wke_run_rounds(wke_engine *e, const int64_t *cmd, const int64_t *cmd_off, int nq, wke_result *res);

/* Executor shape used by wke_run_batched and wke_serve (defaults: 1024
 * queries, 8 patterns, 16 GiB arena per executor). Drops existing executors. */
int wke_configure(wke_engine *e, int max_queries, int max_patterns, uint64_t arena_bytes);

/* Runs the autotuner of the batched executor on these queries (reps timed
 * runs per candidate). The chosen shape stays in effect for the engine's
 * executors: with heavy == 0 for wke_run_batched and the light serving lanes,
 * with heavy != 0 for the heavy serving lane. */
int wke_autotune(wke_engine *e, const int64_t *cmd, const int64_t *cmd_off, int nq, int reps, int heavy,
                 wkg_batch_tuning *best, double *best_ms);
/* Sets the executor shape without autotuning (heavy as for wke_autotune). */
int wke_set_tuning(wke_engine *e, const wkg_batch_tuning *t, int heavy);

/* Where wke_serve runs heavy queries: -1 (default) a worker thread on the
 * per-step path when the graph is resident, else a heavy executor lane; 0 a
 * heavy executor lane; 1 the worker thread (resident graphs only). */
int wke_set_heavy_mode(wke_engine *e, int mode);
/* Worker threads of that heavy lane (default 2, at most 8): each runs heavy
 * queries one at a time on its own stream and read-back block
 * (wkg_thread_context_begin) with its share of the heavy arena, so heavy
 * queries overlap on the device. */
int wke_set_heavy_workers(wke_engine *e, int n);
/* Plain path: a known_to_unknown step whose new column the next pattern
 * filters (known_to_known back to an earlier column, or known_to_const)
 * runs with that filter in one call (wkg_step_k2u_filter), which does not
 * write the fan-out when the filter can run on its lineage. Same tables and
 * work counts; on by default, 0 runs the two steps one after the other. */
int wke_set_fused_filters(wke_engine *e, int on);
/* Plain path on a framed graph: before each step, the first later pattern's
 * predicate that is not resident starts loading on the copy stream
 * (wkg_prefetch_predicate, the current step's predicate kept), as the
 * reference's global_gpu_enable_pipeline does (core/rcache.hpp:273-281). On
 * by default (the reference's shipped config, scripts/config:33); no effect
 * on a resident graph. */
int wke_set_pipeline(wke_engine *e, int on);

/* Light queries of wke_serve / wke_serve_closed that start from an index or
 * a constant and have at most 8 patterns go first to the small-query server
 * (wkg_query_small_submit), which admits each one as soon as one of its
 * `workgroups` workgroups is free, with no batch boundary; one that
 * outgrows its limits goes on to the light lanes. Under load (more than 64
 * light queries in flight) new light queries go to the light lanes, whose
 * batches serve more queries per second. Default 16; 0 sends
 * every light query to the light lanes. The counters of the last serve:
 * queries the server answered and queries it declined. */
int wke_set_serve_small(wke_engine *e, int workgroups);
int wke_serve_small_stats(const wke_engine *e, uint64_t *answered, uint64_t *declined);
/* Counters of the last wke_serve: plain-path fallbacks on the serving
 * thread, queries retried alone, queries run by the worker thread, and the
 * worker's longest backlog. */
int wke_serve_stats(const wke_engine *e, uint64_t *fallbacks, uint64_t *solo, uint64_t *worker_queries,
                    uint64_t *worker_backlog);

/* Served path: an open-loop arrival stream. Query q arrives at arrival_s[q]
 * (seconds from the start, ascending). Arrived queries wait in a light or a
 * heavy queue (heavy[q] != 0; heavy may be NULL). Light queries run on
 * `light_lanes` device executors, each with its own stream and arena; heavy
 * queries run as wke_set_heavy_mode says: on a worker thread, one at a time
 * on the per-step path (whose star arms use lineage), or on one more
 * executor lane. The arenas together are the table budget
 * (table_budget_bytes; half to a heavy lane if there is one, the rest split
 * over the light lanes). Whenever a lane is free, the queries waiting in its queue are
 * admitted as one device batch, as many as its arena holds at the bytes per
 * query the lane has seen (its arena high-water mark over its batch size; a
 * heavy lane starts with one query), so work that arrives while other
 * batches run overlaps them on the device. A query that does not fit its
 * lane's arena beside the others is retried alone, then on the plain path.
 * latency_s in each result is completion - arrival. */
int wke_serve(wke_engine *e, const int64_t *cmd, const int64_t *cmd_off, int nq, const double *arrival_s,
              const uint8_t *heavy, uint64_t table_budget_bytes, int light_lanes, wke_result *res,
              double *wall_s);
/* Closed-loop batch mode, the reference's batch driver (core/proxy.hpp:
 * 287-457): the light queries are issued in order while fewer than
 * light_inflight of them are in flight, the heavy ones while fewer than
 * heavy_inflight are (its heavy in-flight cap), until duration_s has passed;
 * then everything issued completes. Two issue rounds are at least
 * interval_s apart (its send interval, -i; 0: issue whenever a slot frees). Lanes, admission and the heavy worker
 * are wke_serve's. issue_s[q] gets each query's issue time in seconds from
 * the start (-1: not issued before the end, its result zeroed); latency_s is
 * completion - issue (send to reply, as the reference's Logger records it).
 * Queries are issued in order within their class. keep (may be NULL: all)
 * marks the queries whose tables are returned; the others are blind, as the
 * reference's batch requests are (core/proxy.hpp:384): rows, cols, status
 * and expanded only, table NULL. */
int wke_serve_closed(wke_engine *e, const int64_t *cmd, const int64_t *cmd_off, int nq, const uint8_t *heavy,
                     int light_inflight, int heavy_inflight, double duration_s, double interval_s,
                     uint64_t table_budget_bytes, int light_lanes, const uint8_t *keep, wke_result *res,
                     double *issue_s, double *wall_s);

#ifdef __cplusplus
}
#endif
#endif
