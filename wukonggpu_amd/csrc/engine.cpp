// libwkg_engine.so: the host engine above the device C ABI.
//
// Request model and dispatch follow the reference GPU engine:
//   request_or_reply        core/query.hpp:45-175
//   execute_one_step        core/gpu_engine.hpp:395-467 (var_pair dispatch)
//   handlers                core/gpu_engine.hpp:171-291
//   execute_request         core/gpu_engine.hpp:471-526 (one server: the steps
//                           run back to back, tables stay on the device)
// Batched and served execution are this engine's own: the reference runs one
// query at a time on the GPU (core/gpu.hpp:53-61).
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "wukong_engine.h"
#include "wukong_gpu.h"

namespace {

using Clock = std::chrono::steady_clock;

enum VarType { kKnown, kUnknown, kConst };

bool is_tpid(int64_t id) { return id > 1 && id < (1 << 17); }

struct Request {
  const int64_t *cmd = nullptr;
  int ncmd = 0;
  int step = 0;
  int col_num = 0;
  wkg_table *t = nullptr;  // device table between steps (gpu_history_ptr)
  uint64_t expanded = 0;
  int status = WKG_OK;
  bool done = false;

  bool is_finished() const { return step * 4 >= ncmd; }
  static int64_t var2column(int64_t v) { return -v - 1; }
};

// execute_one_step's dispatch (core/gpu_engine.hpp:395-467) for pattern
// `step` of cmd when col_num columns are bound: the step kind, its columns
// and the column count after it.
int dispatch(const int64_t *cmd, int step, int col_num, wkg_step_desc *d, int *new_cols, std::string &err) {
  memset(d, 0, sizeof *d);
  auto fail = [&](const char *m) {
    err = m;
    return WKG_EINVAL;
  };
  auto var_type = [&](int64_t v) {
    if (v >= 0) return kConst;
    if (-v > col_num) return kUnknown;
    return kKnown;
  };
  const int64_t *c = cmd + 4 * step;
  const int64_t start = c[0], pid = c[1], dir = c[2], end = c[3];
  if (dir != 0 && dir != 1) return fail("direction must be IN(0) or OUT(1)");
  if (step == 0 && is_tpid(cmd[0])) {
    if (!(pid == 0 || pid == 1)) return fail("index start needs __PREDICATE__ or rdf:type");
    if (!(col_num == 0 && Request::var2column(end) == 0)) return fail("wrong query plan");
    d->kind = WKG_STEP_I2U;
    d->pid = (int32_t)start;
    d->dir = (int32_t)dir;
    *new_cols = 1;
    return WKG_OK;
  }
  if (pid < 0) return fail("predicate variables are not supported");
  if (pid >= (1 << 17)) return fail("predicate id out of range");
  const VarType ts = var_type(start), te = var_type(end);
  d->pid = (int32_t)pid;
  d->dir = (int32_t)dir;
  if (ts == kConst && te == kUnknown) {
    if (!(col_num == 0 && Request::var2column(end) == 0)) return fail("wrong query plan");
    d->kind = WKG_STEP_C2U;
    d->vid = start;
    *new_cols = 1;
  } else if (ts == kKnown && te == kConst) {
    d->kind = WKG_STEP_K2C;
    d->start_col = (int32_t)Request::var2column(start);
    d->vid = end;
    *new_cols = col_num;
  } else if (ts == kKnown && te == kKnown) {
    d->kind = WKG_STEP_K2K;
    d->start_col = (int32_t)Request::var2column(start);
    d->end_col = (int32_t)Request::var2column(end);
    *new_cols = col_num;
  } else if (ts == kKnown && te == kUnknown) {
    if (col_num != Request::var2column(end)) return fail("wrong query plan");
    d->kind = WKG_STEP_K2U;
    d->start_col = (int32_t)Request::var2column(start);
    *new_cols = col_num + 1;
  } else {
    return fail("unsupported triple pattern");
  }
  return WKG_OK;
}

// A whole query as a device plan: the dispatch of every step, with the
// column count carried from step to step.
int compile_query(const int64_t *cmd, int ncmd, std::vector<wkg_plan_step> &out, std::string &err) {
  out.clear();
  if (!cmd || ncmd <= 0 || ncmd % 4) {
    err = "a query is 4 x int64 per pattern";
    return WKG_EINVAL;
  }
  int cols = 0;
  for (int step = 0; step * 4 < ncmd; step++) {
    wkg_step_desc d;
    int nc;
    if (int rc = dispatch(cmd, step, cols, &d, &nc, err)) return rc;
    wkg_plan_step p{};
    p.kind = d.kind;
    p.pid = d.pid;
    p.dir = d.dir;
    p.start_col = d.start_col;
    p.end_col = d.end_col;
    p.vid = d.vid;
    out.push_back(p);
    cols = nc;
  }
  return WKG_OK;
}

// Copies batch record r (cells of the whole batch) into a wke_result.
// with_table false: rows and cols only (a blind request's reply)
void fill_result(const wkg_batch_result &r, const int32_t *cells, wke_result *out, bool with_table = true) {
  out->status = r.status;
  out->expanded = r.expanded;
  out->rows = r.status == WKG_OK ? r.rows : 0;
  out->cols = r.cols;
  out->table = nullptr;
  const uint64_t n = out->rows * (uint64_t)(r.cols > 0 ? r.cols : 0);
  if (r.status == WKG_OK && n && with_table) {
    out->table = (int32_t *)malloc(n * 4);
    if (!out->table) {
      out->status = WKG_ENOMEM;
      out->rows = 0;
    } else {
      memcpy(out->table, cells + r.offset, n * 4);
    }
  }
}

}  // namespace

struct wke_engine {
  wkg_ctx *ctx = nullptr;
  // The calling thread's last failure message (wke_last_error): the served
  // path runs heavy queries on a second host thread.
  static std::string &err_slot() {
    static thread_local std::string s;
    return s;
  }
  // Held around every per-step (plain) query while a served loop's worker
  // thread may run one: the per-step path keeps one read-back buffer and
  // stream per context.
  std::mutex plain_mu;
  int heavy_mode = -1;  // wke_set_heavy_mode
  int heavy_workers = 2;  // wke_set_heavy_workers: served heavy queries at once (worker threads)
  bool small_queries = true;  // wke_set_small_queries: try wkg_query_small first
  int serve_small = 16;       // wke_set_serve_small: small-query server workgroups while serving (0: none)
  bool fuse_filters = true;   // wke_set_fused_filters: known_to_unknown + filter on its new column in one call
  bool pipeline = true;       // wke_set_pipeline: framed graphs load a later pattern's predicate meanwhile
  uint64_t sv_fallbacks = 0, sv_solo = 0, sv_worker = 0, sv_backlog = 0;  // last wke_serve
  uint64_t sv_small = 0, sv_small_declined = 0;
  // device-resident executor shape (wke_configure) and tuning (wke_autotune)
  int cfg_queries = 1024, cfg_patterns = 8;
  uint64_t cfg_arena = 16ull << 30;
  bool tuned = false, tuned_heavy = false;
  wkg_batch_tuning tuning{}, tuning_heavy{};  // light / default executors, heavy lane
  wkg_batch *exec = nullptr;

  ~wke_engine() { drop_exec(); }
  void drop_exec() {
    if (exec) wkg_batch_destroy(exec);
    exec = nullptr;
  }
  // A new executor with the engine's shape and the tuning of its kind
  // (arena_bytes 0: the default).
  int new_exec(uint64_t arena_bytes, bool heavy, wkg_batch **out) {
    int rc = wkg_batch_create(ctx, cfg_queries, cfg_patterns, arena_bytes ? arena_bytes : cfg_arena, nullptr, out);
    if (rc) return fail_ctx(rc);
    const bool have = heavy ? tuned_heavy : tuned;
    if (have && (rc = wkg_batch_set_tuning(*out, heavy ? &tuning_heavy : &tuning))) {
      wkg_batch_destroy(*out);
      *out = nullptr;
      return fail_ctx(rc);
    }
    return WKG_OK;
  }
  int get_exec(wkg_batch **out) {
    if (!exec)
      if (int rc = new_exec(0, false, &exec)) return rc;
    *out = exec;
    return WKG_OK;
  }
  int fail(int code, const std::string &m) {
    err_slot() = m;
    return code;
  }
  int fail_ctx(int code) {
    err_slot() = wkg_last_error(ctx);
    return code;
  }

  // The step descriptor of r's current step.
  int plan(const Request &r, wkg_step_desc *d, int *new_cols) {
    int rc = dispatch(r.cmd, r.step, r.col_num, d, new_cols, err_slot());
    d->in = r.t;
    return rc;
  }

  // Framed graphs: the first later pattern's predicate that is not resident
  // starts loading on the copy stream while this step (on predicate
  // `in_use`) runs (global_gpu_enable_pipeline, core/rcache.hpp:273-281). A
  // prefetch that fails leaves the load to that step.
  void prefetch_next(const Request &r, int cols, int in_use) {
    if (wkg_graph_framed(ctx) != 1) return;
    std::string why;
    for (int st = r.step + 1; st * 4 < r.ncmd; st++) {
      wkg_step_desc d;
      int nc;
      if (dispatch(r.cmd, st, cols, &d, &nc, why) != WKG_OK) return;
      cols = nc;
      const int res = wkg_predicate_resident(ctx, d.pid);
      if (res == 1) continue;
      if (res == 0) (void)wkg_prefetch_predicate(ctx, d.pid, in_use);
      return;
    }
  }

  int run_one_step(Request &r, void *stream) {
    wkg_step_desc d;
    int nc;
    int rc = plan(r, &d, &nc);
    if (rc) return rc;
    if (pipeline) prefetch_next(r, nc, d.pid);
    wkg_table *o = nullptr;
    uint64_t ex = 0;
    // known_to_unknown whose new column the next pattern filters: one call,
    // and the fan-out it would write is not written when the filter can run
    // on its lineage
    if (fuse_filters && d.kind == WKG_STEP_K2U && (r.step + 1) * 4 < r.ncmd) {
      wkg_step_desc d2;
      int nc2;
      std::string why;
      if (dispatch(r.cmd, r.step + 1, nc, &d2, &nc2, why) == WKG_OK && d2.start_col == nc - 1 &&
          (d2.kind == WKG_STEP_K2C || (d2.kind == WKG_STEP_K2K && d2.end_col < nc - 1))) {
        int fused = 0;
        uint64_t ex2 = 0;
        rc = wkg_step_k2u_filter(ctx, r.t, d.start_col, d.pid, d.dir, d2.kind, d2.pid, d2.dir, d2.end_col, d2.vid,
                                 stream, &o, &fused, &ex, &ex2);
        if (rc) return fail_ctx(rc);
        advance(r, o, nc, ex);
        if (fused) {  // the filter is done too: its output is the table
          r.step++;
          r.col_num = nc2;
          r.expanded += ex2;
        }
        return WKG_OK;
      }
    }
    switch (d.kind) {
      case WKG_STEP_I2U: rc = wkg_step_i2u(ctx, d.pid, d.dir, stream, &o, &ex); break;
      case WKG_STEP_C2U: rc = wkg_step_c2u(ctx, d.vid, d.pid, d.dir, stream, &o, &ex); break;
      case WKG_STEP_K2U: rc = wkg_step_k2u(ctx, r.t, d.start_col, d.pid, d.dir, stream, &o, &ex); break;
      case WKG_STEP_K2K: rc = wkg_step_k2k(ctx, r.t, d.start_col, d.pid, d.dir, d.end_col, stream, &o, &ex); break;
      default: rc = wkg_step_k2c(ctx, r.t, d.start_col, d.pid, d.dir, d.vid, stream, &o, &ex); break;
    }
    if (rc) return fail_ctx(rc);
    advance(r, o, nc, ex);
    return WKG_OK;
  }

  void advance(Request &r, wkg_table *o, int nc, uint64_t ex) {
    if (r.t) wkg_table_free(ctx, r.t);
    r.t = o;
    r.col_num = nc;
    r.step++;
    r.expanded += ex;
  }

  int finish(Request &r, wke_result *res, void *stream) {
    res->status = r.status;
    res->expanded = r.expanded;
    res->table = nullptr;
    res->rows = 0;
    res->cols = r.col_num;
    if (r.status == WKG_OK && r.t) {
      uint64_t rows;
      int cols;
      wkg_table_shape(r.t, &rows, &cols);
      res->rows = rows;
      res->cols = cols;
      if (rows && cols) {
        res->table = (int32_t *)malloc(rows * (uint64_t)cols * 4);
        if (!res->table) {
          res->status = WKG_ENOMEM;
        } else if (int rc = wkg_table_to_host(ctx, r.t, res->table, stream)) {
          free(res->table);
          res->table = nullptr;
          res->status = rc;
          err_slot() = wkg_last_error(ctx);
        }
      }
    }
    if (r.t) wkg_table_free(ctx, r.t);
    r.t = nullptr;
    r.done = true;
    return res->status;
  }

  void drop(Request &r) {
    if (r.t) wkg_table_free(ctx, r.t);
    r.t = nullptr;
  }

  // Host rows of a query's result, appended chunk by chunk.
  struct HostRows {
    int32_t *data = nullptr;
    uint64_t rows = 0, cap_cells = 0;
    int cols = 0;
    ~HostRows() { free(data); }
  };

  std::atomic<uint64_t> splits{0};  // row-half splits made by run_rest (wke_stats)

  // Runs r's remaining steps. A step whose tables do not fit the table
  // budget (WKG_ENOMEM) is retried on the two row halves of its input, each
  // carried depth-first through the rest of the query, so only one chunk's
  // intermediate tables are live at a time. Final rows are appended in input
  // order -- the order an unchunked run gives, since every step keeps its
  // input rows' order -- and work counts add up. Takes ownership of r.t.
  int run_rest(Request &r, HostRows &out, uint64_t &expanded, int depth, void *stream) {
    while (!r.is_finished()) {
      const int rc = run_one_step(r, stream);
      if (rc == WKG_OK) continue;
      uint64_t R = 0;
      int c = 0;
      if (r.t) wkg_table_shape(r.t, &R, &c);
      if (rc != WKG_ENOMEM || !r.t || R < 2 || depth >= 48) {
        drop(r);
        return rc;
      }
      splits++;
      expanded += r.expanded;
      const uint64_t half = R / 2;
      for (int h = 0; h < 2; h++) {
        Request sub = r;
        sub.expanded = 0;
        sub.t = nullptr;
        if (int src = wkg_table_slice(ctx, r.t, h ? half : 0, h ? R - half : half, &sub.t)) {
          drop(r);
          return fail_ctx(src);
        }
        if (int sr = run_rest(sub, out, expanded, depth + 1, stream)) {
          drop(r);
          return sr;
        }
      }
      drop(r);  // the views of both halves are gone
      return WKG_OK;
    }
    expanded += r.expanded;
    uint64_t rows = 0;
    int cols = r.col_num;
    if (r.t) wkg_table_shape(r.t, &rows, &cols);
    out.cols = cols;
    const uint64_t n = rows * (uint64_t)cols;
    if (n) {
      const uint64_t need = out.rows * (uint64_t)cols + n;
      if (need > out.cap_cells) {
        const uint64_t cap = std::max<uint64_t>(need, out.cap_cells + out.cap_cells / 2);
        int32_t *d = (int32_t *)realloc(out.data, cap * 4);
        if (!d) {
          drop(r);
          return fail(WKG_ENOMEM, "host result buffer");
        }
        out.data = d;
        out.cap_cells = cap;
      }
      if (int rc = wkg_table_to_host(ctx, r.t, out.data + out.rows * (uint64_t)cols, stream)) {
        drop(r);
        return fail_ctx(rc);
      }
      out.rows += rows;
    }
    drop(r);
    return WKG_OK;
  }
};

extern "C" {

int wke_create(wkg_ctx *ctx, wke_engine **out) {
  if (!ctx || !out) return WKG_EINVAL;
  wke_engine *e = new wke_engine;
  e->ctx = ctx;
  *out = e;
  return WKG_OK;
}

int wke_destroy(wke_engine *e) {
  delete e;
  return WKG_OK;
}

const char *wke_last_error(const wke_engine *e) { return e ? wke_engine::err_slot().c_str() : "null engine"; }
void wke_free(void *p) { free(p); }

int wke_run_query(wke_engine *e, const int64_t *cmd, int ncmd, wke_result *res) {
  if (!e || !cmd || !res || ncmd <= 0 || ncmd % 4) return WKG_EINVAL;
  memset(res, 0, sizeof *res);
  const auto t0 = Clock::now();
  // a query whose tables stay small: all of its steps in one launch (the
  // library declines with WKG_ERANGE once a table outgrows that; then, and
  // on any other status, the query runs step by step, which also reports
  // the canonical error of a bad plan)
  if (e->small_queries && ncmd / 4 <= 8) {
    std::vector<wkg_plan_step> plan;
    std::string why;
    if (compile_query(cmd, ncmd, plan, why) == WKG_OK &&
        (plan[0].kind == WKG_STEP_I2U || plan[0].kind == WKG_STEP_C2U)) {
      int32_t *cells = (int32_t *)malloc(WKG_SMALL_QUERY_CELLS * 4);
      if (!cells) return e->fail(WKG_ENOMEM, "host result table");
      uint64_t rows = 0, ex = 0;
      int cols = 0;
      if (wkg_query_small(e->ctx, plan.data(), (int)plan.size(), nullptr, cells, WKG_SMALL_QUERY_CELLS, &rows, &cols,
                          &ex) == WKG_OK) {
        const uint64_t n = rows * (uint64_t)cols;
        if (n) {
          int32_t *t = (int32_t *)realloc(cells, n * 4);  // (shrinks in place)
          res->table = t ? t : cells;
        } else {
          free(cells);
        }
        res->rows = rows;
        res->cols = cols;
        res->expanded = ex;
        res->status = WKG_OK;
        res->latency_s = std::chrono::duration<double>(Clock::now() - t0).count();
        return WKG_OK;
      }
      free(cells);
    }
  }
  Request r;
  r.cmd = cmd;
  r.ncmd = ncmd;
  wke_engine::HostRows out;
  uint64_t ex = 0;
  int rc = e->run_rest(r, out, ex, 0, nullptr);
  res->status = rc;
  if (rc == WKG_OK) {
    res->rows = out.rows;
    res->cols = out.cols;
    res->expanded = ex;
    res->table = out.data;  // handed to the caller (wke_free)
    out.data = nullptr;
  }
  res->latency_s = std::chrono::duration<double>(Clock::now() - t0).count();
  return rc;
}

int wke_set_small_queries(wke_engine *e, int on) {
  if (!e) return WKG_EINVAL;
  e->small_queries = on != 0;
  return WKG_OK;
}

int wke_stats(const wke_engine *e, uint64_t *splits) {
  if (!e) return WKG_EINVAL;
  if (splits) *splits = e->splits;
  return WKG_OK;
}

int wke_run_plain(wke_engine *e, const int64_t *cmd, const int64_t *cmd_off, int nq, wke_result *res) {
  if (!e || !cmd || !cmd_off || !res || nq < 0) return WKG_EINVAL;
  int first = WKG_OK;
  for (int q = 0; q < nq; q++) {
    int rc = wke_run_query(e, cmd + cmd_off[q], (int)(cmd_off[q + 1] - cmd_off[q]), &res[q]);
    if (rc && !first) first = rc;
  }
  return first;
}

namespace {

// Runs one round: the current step of every request in `act` as one fused
// batch. Requests whose step cannot be planned fail individually; if the
// batch as a whole fails for lack of memory, its steps are retried one by one.
int batched_round(wke_engine *e, std::vector<Request *> &act, void *stream) {
  std::vector<wkg_step_desc> descs;
  std::vector<int> cols;
  std::vector<Request *> who;
  for (Request *r : act) {
    wkg_step_desc d;
    int nc;
    int rc = e->plan(*r, &d, &nc);
    if (rc) {
      r->status = rc;
      continue;
    }
    descs.push_back(d);
    cols.push_back(nc);
    who.push_back(r);
  }
  if (descs.empty()) return WKG_OK;
  std::vector<wkg_table *> outs(descs.size(), nullptr);
  std::vector<uint64_t> ex(descs.size(), 0);
  int rc = wkg_step_batch(e->ctx, (int)descs.size(), descs.data(), stream, outs.data(), ex.data());
  if (rc == WKG_OK) {
    for (size_t i = 0; i < who.size(); i++) e->advance(*who[i], outs[i], cols[i], ex[i]);
    return WKG_OK;
  }
  if (rc != WKG_ENOMEM) {
    e->err_slot() = wkg_last_error(e->ctx);
    for (Request *r : who) r->status = rc;
    return rc;
  }
  for (Request *r : who) {
    int r2 = e->run_one_step(*r, stream);
    if (r2) r->status = r2;
  }
  return WKG_OK;
}

}  // namespace

int wke_run_rounds(wke_engine *e, const int64_t *cmd, const int64_t *cmd_off, int nq, wke_result *res) {
  if (!e || !cmd || !cmd_off || !res || nq < 0) return WKG_EINVAL;
  const auto t0 = Clock::now();
  std::vector<Request> reqs(nq);
  for (int q = 0; q < nq; q++) {
    memset(&res[q], 0, sizeof res[q]);
    reqs[q].cmd = cmd + cmd_off[q];
    reqs[q].ncmd = (int)(cmd_off[q + 1] - cmd_off[q]);
    if (reqs[q].ncmd <= 0 || reqs[q].ncmd % 4) reqs[q].status = WKG_EINVAL;
  }
  int first = WKG_OK;
  while (true) {
    std::vector<Request *> act;
    for (int q = 0; q < nq; q++) {
      Request &r = reqs[q];
      if (r.done) continue;
      if (r.status != WKG_OK || r.is_finished()) {
        if (r.status != WKG_OK) e->drop(r);
        e->finish(r, &res[q], nullptr);
        res[q].latency_s = std::chrono::duration<double>(Clock::now() - t0).count();
        if (res[q].status && !first) first = res[q].status;
        continue;
      }
      act.push_back(&r);
    }
    if (act.empty()) break;
    batched_round(e, act, nullptr);
  }
  return first;
}

int wke_configure(wke_engine *e, int max_queries, int max_patterns, uint64_t arena_bytes) {
  if (!e || max_queries < 1 || max_patterns < 1 || max_patterns > 64 || arena_bytes < (1u << 20)) return WKG_EINVAL;
  e->drop_exec();
  e->cfg_queries = max_queries;
  e->cfg_patterns = max_patterns;
  e->cfg_arena = arena_bytes;
  return WKG_OK;
}

namespace {

struct Compiled {
  std::vector<wkg_plan_step> steps;
  std::vector<int32_t> off;
  std::vector<int> who;  // query index of each compiled plan
};

// Appends query q's plan to c.
void add_plan(Compiled &c, const std::vector<wkg_plan_step> &plan, int q) {
  if (c.off.empty()) c.off.push_back(0);
  c.steps.insert(c.steps.end(), plan.begin(), plan.end());
  c.off.push_back((int32_t)c.steps.size());
  c.who.push_back(q);
}

}  // namespace

namespace {

// Result sink of the batched path: per-query malloc'd tables (wke_run_batched)
// or one packed buffer (wke_run_batched_packed).
struct Sink {
  bool packed = false;
  uint64_t *cell_off = nullptr;  // packed: first cell of each query's table
  int32_t *cells = nullptr;
  uint64_t ncells = 0, cap = 0;
  bool reserve(uint64_t n) {
    if (ncells + n <= cap) return true;
    const uint64_t c = std::max<uint64_t>(ncells + n, cap + cap / 2);
    int32_t *d = (int32_t *)realloc(cells, c * 4);
    if (!d) return false;
    cells = d;
    cap = c;
    return true;
  }
};

int run_batched_into(wke_engine *e, const int64_t *cmd, const int64_t *cmd_off, int nq, wke_result *res, Sink &sk) {
  const auto t0 = Clock::now();
  auto since = [&] { return std::chrono::duration<double>(Clock::now() - t0).count(); };
  int first = WKG_OK;
  auto note = [&](int rc) {
    if (rc && !first) first = rc;
  };
  std::vector<int> plain;  // queries for the plain path
  std::vector<std::vector<wkg_plan_step>> plans(nq);
  for (int q = 0; q < nq; q++) {
    memset(&res[q], 0, sizeof res[q]);
    if (sk.packed) sk.cell_off[q] = 0;
    const int n = (int)(cmd_off[q + 1] - cmd_off[q]);
    std::string why;
    if (int rc = compile_query(cmd + cmd_off[q], n, plans[q], why)) {
      res[q].status = rc;
      e->err_slot() = why;
      note(rc);
    } else if ((int)plans[q].size() > e->cfg_patterns) {
      plain.push_back(q);
    }
  }
  wkg_batch *b = nullptr;
  if (int rc = e->get_exec(&b)) return rc;
  int q = 0;
  std::vector<wkg_batch_result> rec((size_t)e->cfg_queries);
  while (q < nq) {
    Compiled c;
    for (; q < nq && (int)c.who.size() < e->cfg_queries; q++)
      if (res[q].status == WKG_OK && !plans[q].empty() && (int)plans[q].size() <= e->cfg_patterns)
        add_plan(c, plans[q], q);
    if (c.who.empty()) break;
    const int32_t *cells = nullptr;
    uint64_t ncells = 0;
    int rc = wkg_batch_run(b, (int)c.who.size(), c.steps.data(), c.off.data(), rec.data(), &cells, &ncells);
    if (rc == WKG_ENOMEM) {  // e.g. the batch's predicates do not all fit a framed graph budget at once
      plain.insert(plain.end(), c.who.begin(), c.who.end());
      continue;
    }
    if (rc) return e->fail_ctx(rc);
    const double lat = since();
    uint64_t base = 0;
    if (sk.packed) {  // the batch's result buffer in one piece
      if (!sk.reserve(ncells)) return e->fail(WKG_ENOMEM, "host result buffer");
      base = sk.ncells;
      if (ncells) memcpy(sk.cells + base, cells, ncells * 4);
      sk.ncells += ncells;
    }
    for (size_t i = 0; i < c.who.size(); i++) {
      const int qi = c.who[i];
      if (rec[i].status == WKG_ENOMEM) {
        plain.push_back(qi);
        continue;
      }
      if (sk.packed) {
        res[qi].status = rec[i].status;
        res[qi].expanded = rec[i].expanded;
        res[qi].rows = rec[i].status == WKG_OK ? rec[i].rows : 0;
        res[qi].cols = rec[i].cols;
        sk.cell_off[qi] = base + rec[i].offset;
      } else {
        fill_result(rec[i], cells, &res[qi]);
      }
      res[qi].latency_s = lat;
      note(res[qi].status);
    }
  }
  for (int qi : plain) {  // beyond the executor's shape or its arena
    int rc = wke_run_query(e, cmd + cmd_off[qi], (int)(cmd_off[qi + 1] - cmd_off[qi]), &res[qi]);
    res[qi].latency_s = since();
    note(rc);
    if (sk.packed && rc == WKG_OK) {
      const uint64_t n = res[qi].rows * (uint64_t)res[qi].cols;
      if (!sk.reserve(n)) return e->fail(WKG_ENOMEM, "host result buffer");
      if (n) memcpy(sk.cells + sk.ncells, res[qi].table, n * 4);
      sk.cell_off[qi] = sk.ncells;
      sk.ncells += n;
      free(res[qi].table);
      res[qi].table = nullptr;
    }
  }
  return first;
}

}  // namespace

int wke_run_batched(wke_engine *e, const int64_t *cmd, const int64_t *cmd_off, int nq, wke_result *res) {
  if (!e || !cmd || !cmd_off || !res || nq < 0) return WKG_EINVAL;
  Sink sk;
  return run_batched_into(e, cmd, cmd_off, nq, res, sk);
}

int wke_run_batched_packed(wke_engine *e, const int64_t *cmd, const int64_t *cmd_off, int nq, wke_result *res,
                           uint64_t *cell_off, int32_t **cells, uint64_t *ncells) {
  if (!e || !cmd || !cmd_off || !res || nq < 0 || !cell_off || !cells || !ncells) return WKG_EINVAL;
  Sink sk;
  sk.packed = true;
  sk.cell_off = cell_off;
  int rc = run_batched_into(e, cmd, cmd_off, nq, res, sk);
  *cells = sk.cells;
  *ncells = sk.ncells;
  return rc;
}

int wke_set_heavy_mode(wke_engine *e, int mode) {
  if (!e || mode < -1 || mode > 1) return WKG_EINVAL;
  e->heavy_mode = mode;
  return WKG_OK;
}

int wke_set_heavy_workers(wke_engine *e, int n) {
  if (!e || n < 1 || n > 8) return WKG_EINVAL;
  e->heavy_workers = n;
  return WKG_OK;
}

int wke_set_fused_filters(wke_engine *e, int on) {
  if (!e) return WKG_EINVAL;
  e->fuse_filters = on != 0;
  return WKG_OK;
}

int wke_set_pipeline(wke_engine *e, int on) {
  if (!e) return WKG_EINVAL;
  e->pipeline = on != 0;
  return WKG_OK;
}

int wke_set_serve_small(wke_engine *e, int workgroups) {
  if (!e || workgroups < 0) return WKG_EINVAL;
  e->serve_small = workgroups;
  return WKG_OK;
}

int wke_serve_small_stats(const wke_engine *e, uint64_t *answered, uint64_t *declined) {
  if (!e) return WKG_EINVAL;
  if (answered) *answered = e->sv_small;
  if (declined) *declined = e->sv_small_declined;
  return WKG_OK;
}

int wke_serve_stats(const wke_engine *e, uint64_t *fallbacks, uint64_t *solo, uint64_t *worker_queries,
                    uint64_t *worker_backlog) {
  if (!e) return WKG_EINVAL;
  if (fallbacks) *fallbacks = e->sv_fallbacks;
  if (solo) *solo = e->sv_solo;
  if (worker_queries) *worker_queries = e->sv_worker;
  if (worker_backlog) *worker_backlog = e->sv_backlog;
  return WKG_OK;
}

int wke_set_tuning(wke_engine *e, const wkg_batch_tuning *t, int heavy) {
  if (!e || !t) return WKG_EINVAL;
  if (!heavy && e->exec) {
    if (int rc = wkg_batch_set_tuning(e->exec, t)) return e->fail_ctx(rc);
    if (int rc = wkg_batch_get_tuning(e->exec, &e->tuning)) return e->fail_ctx(rc);
  } else {
    (heavy ? e->tuning_heavy : e->tuning) = *t;
  }
  (heavy ? e->tuned_heavy : e->tuned) = true;
  return WKG_OK;
}

int wke_autotune(wke_engine *e, const int64_t *cmd, const int64_t *cmd_off, int nq, int reps, int heavy,
                 wkg_batch_tuning *best, double *best_ms) {
  if (!e || !cmd || !cmd_off || nq < 1 || reps < 1) return WKG_EINVAL;
  Compiled c;
  std::vector<wkg_plan_step> plan;
  for (int q = 0; q < nq && (int)c.who.size() < e->cfg_queries; q++) {
    std::string why;
    if (compile_query(cmd + cmd_off[q], (int)(cmd_off[q + 1] - cmd_off[q]), plan, why) == WKG_OK &&
        (int)plan.size() <= e->cfg_patterns)
      add_plan(c, plan, q);
  }
  if (c.who.empty()) return e->fail(WKG_EINVAL, "no query of the sample compiles to a device plan");
  wkg_batch *b = nullptr;
  if (heavy) {  // a separate executor, so the default one keeps its captured sequences
    if (int rc = e->new_exec(0, true, &b)) return rc;
  } else if (int rc = e->get_exec(&b)) {
    return rc;
  }
  wkg_batch_tuning t{};
  double ms = 0;
  int rc = wkg_batch_autotune(b, (int)c.who.size(), c.steps.data(), c.off.data(), reps, &t, &ms);
  if (rc) e->fail_ctx(rc);
  if (heavy) wkg_batch_destroy(b);
  if (rc) return rc;
  (heavy ? e->tuning_heavy : e->tuning) = t;
  (heavy ? e->tuned_heavy : e->tuned) = true;
  if (best) *best = t;
  if (best_ms) *best_ms = ms;
  return WKG_OK;
}

namespace {

// One device executor of the served path with the queries it runs. `est`
// is the arena bytes per query seen so far (its high-water mark over the
// batch size); a batch admits at most arena / est queries.
struct Lane {
  wkg_batch *b = nullptr;
  bool heavy = false;
  bool busy = false;
  uint64_t arena = 0;
  double est = 0;
  std::vector<int> run;
};

// The heavy lane on a resident graph: host threads that each run heavy
// queries one at a time on the per-step path (whose star arms use lineage),
// each on a stream and read-back block of its own (wkg_thread_context_*),
// while the light lanes' executors run on theirs.
struct PlainWorker {
  std::mutex mu;
  std::condition_variable cv;
  std::deque<int> in;
  struct Done {
    int q, rc;
    std::string why;
  };
  std::vector<Done> out;
  bool stop = false;
  int ready = 0;
  int pending = 0;  // queued or running (main thread only)
  std::vector<std::thread> th;
};

}  // namespace

namespace {

// Closed-loop admission (the reference's batch driver): per-class in-flight
// caps and a duration instead of arrival times.
struct Closed {
  int cap[2];            // light, heavy
  double duration;       // no issue after this many seconds
  double interval;       // at least this long between two issue rounds (0: none)
  double *issue_s;       // per query: issue time, -1 if never issued
  const uint8_t *keep;   // per query: return the table (NULL: all)
  bool blind(int q) const { return keep && !keep[q]; }
};

// The served loop, open (arrival_s) or closed (cl).
int serve_loop(wke_engine *e, const int64_t *cmd, const int64_t *cmd_off, int nq, const double *arrival_s,
               const Closed *cl, const uint8_t *heavy, uint64_t table_budget_bytes, int light_lanes, wke_result *res,
               double *wall_s) {
  bool any_heavy = false;
  for (int q = 0; heavy && q < nq; q++) any_heavy |= heavy[q] != 0;
  if (table_budget_bytes == 0) table_budget_bytes = 64ull << 30;
  // heavy queries: a worker thread on the per-step path when the graph is
  // resident; with a framed graph (segment loads and evictions are not
  // thread-safe) a heavy executor lane instead
  const bool resident = wkg_graph_framed(e->ctx) == 0;
  const bool worker = any_heavy && resident && e->heavy_mode != 0;
  e->sv_fallbacks = e->sv_solo = e->sv_worker = e->sv_backlog = 0;
  e->sv_small = e->sv_small_declined = 0;
  // light queries whose tables stay small: the persistent small-query
  // server, admitted one by one as they arrive (a query its limits decline
  // goes on to the light lanes)
  const bool small = resident && e->serve_small > 0;
  if (small && wkg_set_small_query_server(e->ctx, e->serve_small) != WKG_OK) return e->fail_ctx(WKG_EINVAL);
  // With few hardware queues the lanes' streams may share the server's: its
  // workgroups then leave after 20 us without a query and within 1 ms, so a
  // lane batch queued behind them waits at most that long.
  double sq_idle = 0, sq_life = 0;
  const bool sq_short = small && wkg_hw_queues() < 8;
  if (sq_short) {
    (void)wkg_get_small_query_server_times(e->ctx, &sq_idle, &sq_life);
    (void)wkg_set_small_query_server_times(e->ctx, std::min(sq_idle, 20e-6), std::min(sq_life, 1e-3));
  }
  const bool heavy_lane = any_heavy && !worker;
  const uint64_t heavy_arena = (heavy_lane || worker) ? table_budget_bytes / 2 : 0;
  const uint64_t light_arena = (table_budget_bytes - heavy_arena) / (uint64_t)light_lanes;
  std::vector<Lane> lanes;
  PlainWorker pw;
  std::deque<int> sq_wait;                          // light queries waiting for a server slot
  std::vector<std::pair<int, uint64_t>> sq_run;     // (query, ticket) on the server
  std::vector<int32_t> sq_buf(WKG_SMALL_QUERY_CELLS);
  const int sq_keep = 4;                            // the context's server shape after serving
  auto release = [&] {
    if (!pw.th.empty()) {
      {
        std::lock_guard<std::mutex> lk(pw.mu);
        pw.stop = true;
      }
      pw.cv.notify_all();
      for (std::thread &t : pw.th) t.join();
      pw.th.clear();
    }
    for (Lane &l : lanes) {
      if (l.busy) {
        std::vector<wkg_batch_result> sink(l.run.size());
        wkg_batch_collect(l.b, sink.data(), nullptr, nullptr);
      }
      wkg_batch_destroy(l.b);
    }
    std::vector<int32_t> sink(WKG_SMALL_QUERY_CELLS);
    for (const auto &qt : sq_run) {  // every ticket is collected
      uint64_t r, x;
      int c;
      (void)wkg_query_small_collect(e->ctx, qt.second, 1, sink.data(), sink.size(), &r, &c, &x);
    }
    sq_run.clear();
    if (small) (void)wkg_set_small_query_server(e->ctx, sq_keep);
    if (sq_short) (void)wkg_set_small_query_server_times(e->ctx, sq_idle, sq_life);
  };
  for (int i = 0; i < light_lanes + (heavy_lane ? 1 : 0); i++) {
    Lane l;
    l.heavy = i == light_lanes;
    l.arena = l.heavy ? heavy_arena : light_arena;
    if (int rc = e->new_exec(l.arena, l.heavy, &l.b)) {
      release();
      return rc;
    }
    lanes.push_back(l);
  }
  std::vector<std::vector<wkg_plan_step>> plans(nq);
  std::vector<char> solo(nq, 0);
  std::deque<int> wait[2];  // light, heavy
  if (worker) {
    // The worker's tables and scratch come from its own arena (the heavy
    // share of the budget), reserved before arrivals start and reset after
    // every query: no hipMalloc or hipFree (which wait on the device) while
    // the light lanes keep it busy.
    // Each of heavy_workers threads has its own share of the arena, stream
    // and read-back block, so heavy queries overlap on the device.
    const int nw = std::max(1, e->heavy_workers);
    for (int w = 0; w < nw; w++)
      pw.th.emplace_back([&, nw] {
        const bool own = wkg_thread_context_begin(e->ctx) == WKG_OK;
        const bool arena = wkg_thread_arena_begin(e->ctx, heavy_arena / (uint64_t)nw) == WKG_OK;
        // short launches, so that light batches get the GPU between them
        (void)wkg_thread_launch_rows(e->ctx, 1ull << 22);
        {
          std::lock_guard<std::mutex> lk(pw.mu);
          pw.ready++;
        }
        pw.cv.notify_all();
        for (;;) {
          int q;
          {
            std::unique_lock<std::mutex> lk(pw.mu);
            pw.cv.wait(lk, [&] { return pw.stop || !pw.in.empty(); });
            if (pw.stop) break;
            q = pw.in.front();
            pw.in.pop_front();
          }
          int rc;
          if (own) {
            rc = wke_run_query(e, cmd + cmd_off[q], (int)(cmd_off[q + 1] - cmd_off[q]), &res[q]);
          } else {  // (no stream of its own: the context's, one query at a time)
            std::lock_guard<std::mutex> pl(e->plain_mu);
            rc = wke_run_query(e, cmd + cmd_off[q], (int)(cmd_off[q + 1] - cmd_off[q]), &res[q]);
          }
          if (arena) (void)wkg_thread_arena_reset(e->ctx);
          std::lock_guard<std::mutex> lk(pw.mu);
          pw.out.push_back({q, rc, rc ? std::string(wke_last_error(e)) : std::string()});
        }
        if (arena) (void)wkg_thread_arena_end(e->ctx);
        (void)wkg_thread_launch_rows(e->ctx, 0);
        if (own) (void)wkg_thread_context_end(e->ctx);
      });
    std::unique_lock<std::mutex> lk(pw.mu);
    pw.cv.wait(lk, [&] { return pw.ready == nw; });
  }
  const auto t0 = Clock::now();
  auto now = [&] { return std::chrono::duration<double>(Clock::now() - t0).count(); };
  int next = 0, completed = 0, first = WKG_OK;
  // closed loop: each class's queries in order, issued below its cap
  std::vector<int> order[2];
  size_t onext[2] = {0, 0};
  int inflight[2] = {0, 0}, issued = 0;
  double next_round = 0;  // closed loop: earliest time of the next issue round
  if (cl) {
    for (int q = 0; q < nq; q++) {
      order[heavy && heavy[q] ? 1 : 0].push_back(q);
      cl->issue_s[q] = -1;
      memset(&res[q], 0, sizeof res[q]);  // what a query never issued returns
    }
  }
  auto cls = [&](int q) { return heavy && heavy[q] ? 1 : 0; };
  auto done = [&](int q, int rc) {
    res[q].latency_s = now() - (cl ? cl->issue_s[q] : arrival_s[q]);
    if (cl && cl->blind(q) && res[q].table) {
      free(res[q].table);
      res[q].table = nullptr;
    }
    if (rc && !first) first = rc;
    completed++;
    inflight[cls(q)]--;
  };
  auto run_plain = [&](int q) {
    e->sv_fallbacks++;
    int rc;
    {
      std::lock_guard<std::mutex> pl(e->plain_mu);
      rc = wke_run_query(e, cmd + cmd_off[q], (int)(cmd_off[q + 1] - cmd_off[q]), &res[q]);
    }
    done(q, rc);
  };
  std::vector<wkg_batch_result> rec((size_t)e->cfg_queries);
  for (;;) {
    if (cl) {  // issued all it will issue, and all of that completed
      const bool more = now() < cl->duration && (onext[0] < order[0].size() || onext[1] < order[1].size());
      if (!more && completed == issued) break;
    } else if (completed >= nq) {
      break;
    }
    bool progressed = false;
    // admit arrivals (open loop) or issue below the caps (closed loop)
    const double t = now();
    int round_issued = 0;
    for (;;) {
      int q = -1;
      if (cl) {
        if (t >= cl->duration || t < next_round) break;
        for (int k = 0; k < 2 && q < 0; k++)
          if (inflight[k] < cl->cap[k] && onext[k] < order[k].size()) q = order[k][onext[k]++];
        if (q < 0) break;
        cl->issue_s[q] = t;
      } else {
        if (next >= nq || arrival_s[next] > t) break;
        q = next++;
      }
      issued++;
      round_issued++;
      inflight[cls(q)]++;
      memset(&res[q], 0, sizeof res[q]);
      std::string why;
      int rc = compile_query(cmd + cmd_off[q], (int)(cmd_off[q + 1] - cmd_off[q]), plans[q], why);
      if (rc) {
        res[q].status = rc;
        e->err_slot() = why;
        done(q, rc);
      } else if ((int)plans[q].size() > e->cfg_patterns) {
        run_plain(q);
      } else if (small && !(any_heavy && heavy[q]) && plans[q].size() <= 8 &&
                 (plans[q][0].kind == WKG_STEP_I2U || plans[q][0].kind == WKG_STEP_C2U) && inflight[0] <= 64) {
        // (under load -- more than 64 light queries in flight -- the light
        // lanes' batches serve more queries per second than the server)
        sq_wait.push_back(q);
      } else if (worker && heavy[q]) {
        {
          std::lock_guard<std::mutex> lk(pw.mu);
          pw.in.push_back(q);
        }
        pw.cv.notify_one();
        pw.pending++;
        e->sv_worker++;
        e->sv_backlog = std::max<uint64_t>(e->sv_backlog, (uint64_t)pw.pending);
      } else {
        wait[any_heavy && heavy[q] ? 1 : 0].push_back(q);
      }
      progressed = true;
    }
    if (cl && round_issued && cl->interval > 0) next_round = t + cl->interval;
    // heavy queries the worker finished
    if (worker) {
      std::vector<PlainWorker::Done> fin;
      {
        std::lock_guard<std::mutex> lk(pw.mu);
        fin.swap(pw.out);
      }
      for (const PlainWorker::Done &d : fin) {
        if (d.rc) e->err_slot() = d.why;
        pw.pending--;
        done(d.q, d.rc);
        progressed = true;
      }
    }
    // the small-query server: answers, then new tickets for free slots
    for (size_t i = 0; i < sq_run.size();) {
      const int q = sq_run[i].first;
      uint64_t r = 0, x = 0;
      int c = 0;
      const int rc = wkg_query_small_collect(e->ctx, sq_run[i].second, 0, sq_buf.data(), sq_buf.size(), &r, &c, &x);
      if (rc == WKG_EAGAIN) {
        i++;
        continue;
      }
      sq_run[i] = sq_run.back();
      sq_run.pop_back();
      progressed = true;
      if (rc == WKG_ERANGE) {  // past the limits: a light lane runs it
        e->sv_small_declined++;
        wait[0].push_back(q);
        continue;
      }
      e->sv_small++;
      res[q].status = rc;
      if (rc == WKG_OK) {
        res[q].rows = r;
        res[q].cols = c;
        res[q].expanded = x;
        const uint64_t n = r * (uint64_t)c;
        if (n && !(cl && cl->blind(q))) {
          res[q].table = (int32_t *)malloc(n * 4);
          if (res[q].table) memcpy(res[q].table, sq_buf.data(), n * 4);
          else res[q].status = WKG_ENOMEM, res[q].rows = 0;
        }
      } else {
        e->err_slot() = wkg_last_error(e->ctx);
      }
      done(q, res[q].status);
    }
    while (!sq_wait.empty()) {
      const int q = sq_wait.front();
      uint64_t ticket;
      const int rc = wkg_query_small_submit(e->ctx, plans[q].data(), (int)plans[q].size(), &ticket);
      if (rc == WKG_EAGAIN) break;  // every slot in flight
      sq_wait.pop_front();
      progressed = true;
      if (rc == WKG_OK) sq_run.emplace_back(q, ticket);
      else wait[0].push_back(q);  // (e.g. a framed graph)
    }
    // retire finished batches
    for (Lane &l : lanes) {
      if (!l.busy || wkg_batch_busy(l.b)) continue;
      const int32_t *cells = nullptr;
      int rc = wkg_batch_collect(l.b, rec.data(), &cells, nullptr);
      l.busy = false;
      progressed = true;
      uint64_t peak = 0;
      if (rc == WKG_OK && wkg_batch_stats(l.b, &peak, nullptr) == WKG_OK && !l.run.empty())
        l.est = std::max(l.est, (double)peak / (double)l.run.size());
      std::deque<int> &w = wait[l.heavy ? 1 : 0];
      for (size_t i = 0; i < l.run.size(); i++) {
        const int q = l.run[i];
        if (rc) {
          res[q].status = rc;
          done(q, rc);
        } else if (rec[i].status == WKG_ENOMEM && l.run.size() > 1) {
          solo[q] = 1;  // retried alone
          e->sv_solo++;
          w.push_front(q);
        } else if (rec[i].status == WKG_ENOMEM) {
          run_plain(q);
        } else {
          fill_result(rec[i], cells, &res[q], !(cl && cl->blind(q)));
          done(q, res[q].status);
        }
      }
      l.run.clear();
    }
    // start every idle lane on what waits in its queue
    for (Lane &l : lanes) {
      std::deque<int> &w = wait[l.heavy ? 1 : 0];
      if (l.busy || w.empty()) continue;
      Compiled c;
      // admission: what the lane's arena holds at the bytes per query seen so
      // far (a heavy lane starts with one query to learn it)
      int cap_n = e->cfg_queries;
      if (l.est > 0) cap_n = (int)std::min<double>(cap_n, std::max(1.0, 0.9 * (double)l.arena / l.est));
      else if (l.heavy) cap_n = 1;
      if (solo[w.front()]) {
        add_plan(c, plans[w.front()], w.front());
        w.pop_front();
      } else {
        while (!w.empty() && !solo[w.front()] && (int)c.who.size() < cap_n) {
          add_plan(c, plans[w.front()], w.front());
          w.pop_front();
        }
      }
      int rc = wkg_batch_launch(l.b, (int)c.who.size(), c.steps.data(), c.off.data());
      if (rc == WKG_ENOMEM) {  // the batch's predicates do not all fit a framed graph budget at once
        for (int q : c.who) run_plain(q);
      } else if (rc) {
        e->err_slot() = wkg_last_error(e->ctx);
        for (int q : c.who) {
          res[q].status = rc;
          done(q, rc);
        }
      } else {
        l.busy = true;
        l.run = c.who;
      }
      progressed = true;
    }
    if (progressed) continue;
    if (!sq_run.empty()) {  // answers come within microseconds: poll, do not sleep
      std::this_thread::yield();
      continue;
    }
    bool any_busy = pw.pending > 0;
    for (Lane &l : lanes) any_busy |= l.busy;
    if (!cl && !any_busy && next < nq) {
      const double dt = arrival_s[next] - now();
      if (dt > 0) std::this_thread::sleep_for(std::chrono::duration<double>(dt));
    } else {
      std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
  }
  if (wall_s) *wall_s = now();
  release();
  return first;
}

}  // namespace

int wke_serve(wke_engine *e, const int64_t *cmd, const int64_t *cmd_off, int nq, const double *arrival_s,
              const uint8_t *heavy, uint64_t table_budget_bytes, int light_lanes, wke_result *res,
              double *wall_s) {
  if (!e || !cmd || !cmd_off || !arrival_s || !res || nq < 0 || light_lanes < 1 || light_lanes > 16)
    return WKG_EINVAL;
  return serve_loop(e, cmd, cmd_off, nq, arrival_s, nullptr, heavy, table_budget_bytes, light_lanes, res, wall_s);
}

int wke_serve_closed(wke_engine *e, const int64_t *cmd, const int64_t *cmd_off, int nq, const uint8_t *heavy,
                     int light_inflight, int heavy_inflight, double duration_s, double interval_s,
                     uint64_t table_budget_bytes, int light_lanes, const uint8_t *keep, wke_result *res,
                     double *issue_s, double *wall_s) {
  if (!e || !cmd || !cmd_off || !res || !issue_s || nq < 0 || light_lanes < 1 || light_lanes > 16 ||
      light_inflight < 0 || heavy_inflight < 0 || duration_s <= 0 || interval_s < 0)
    return WKG_EINVAL;
  Closed cl{{light_inflight, heavy_inflight}, duration_s, interval_s, issue_s, keep};
  return serve_loop(e, cmd, cmd_off, nq, nullptr, &cl, heavy, table_budget_bytes, light_lanes, res, wall_s);
}

}  // extern "C"
