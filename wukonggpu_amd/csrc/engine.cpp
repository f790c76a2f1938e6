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
#include <chrono>
#include <deque>
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
  bool start_from_index() const { return is_tpid(cmd[0]); }
  VarType variable_type(int64_t v) const {
    if (v >= 0) return kConst;
    if (-v > col_num) return kUnknown;
    return kKnown;
  }
  static int64_t var2column(int64_t v) { return -v - 1; }
};

}  // namespace

struct wke_engine {
  wkg_ctx *ctx = nullptr;
  std::string err;
  int fail(int code, const std::string &m) {
    err = m;
    return code;
  }
  int fail_ctx(int code) {
    err = wkg_last_error(ctx);
    return code;
  }

  // The step descriptor of r's current step (execute_one_step's dispatch).
  int plan(const Request &r, wkg_step_desc *d, int *new_cols) {
    memset(d, 0, sizeof *d);
    const int64_t *c = r.cmd + 4 * r.step;
    const int64_t start = c[0], pid = c[1], dir = c[2], end = c[3];
    if (dir != 0 && dir != 1) return fail(WKG_EINVAL, "direction must be IN(0) or OUT(1)");
    if (r.step == 0 && r.start_from_index()) {
      if (!(pid == 0 || pid == 1)) return fail(WKG_EINVAL, "index start needs __PREDICATE__ or rdf:type");
      if (!(r.col_num == 0 && Request::var2column(end) == 0)) return fail(WKG_EINVAL, "wrong query plan");
      d->kind = WKG_STEP_I2U;
      d->pid = (int32_t)start;
      d->dir = (int32_t)dir;
      *new_cols = 1;
      return WKG_OK;
    }
    if (pid < 0) return fail(WKG_EINVAL, "predicate variables are not supported");
    if (pid >= (1 << 17)) return fail(WKG_EINVAL, "predicate id out of range");
    const VarType ts = r.variable_type(start), te = r.variable_type(end);
    d->pid = (int32_t)pid;
    d->dir = (int32_t)dir;
    d->in = r.t;
    if (ts == kConst && te == kUnknown) {
      if (!(r.col_num == 0 && Request::var2column(end) == 0)) return fail(WKG_EINVAL, "wrong query plan");
      d->kind = WKG_STEP_C2U;
      d->vid = start;
      *new_cols = 1;
    } else if (ts == kKnown && te == kConst) {
      d->kind = WKG_STEP_K2C;
      d->start_col = (int32_t)Request::var2column(start);
      d->vid = end;
      *new_cols = r.col_num;
    } else if (ts == kKnown && te == kKnown) {
      d->kind = WKG_STEP_K2K;
      d->start_col = (int32_t)Request::var2column(start);
      d->end_col = (int32_t)Request::var2column(end);
      *new_cols = r.col_num;
    } else if (ts == kKnown && te == kUnknown) {
      if (r.col_num != Request::var2column(end)) return fail(WKG_EINVAL, "wrong query plan");
      d->kind = WKG_STEP_K2U;
      d->start_col = (int32_t)Request::var2column(start);
      *new_cols = r.col_num + 1;
    } else {
      return fail(WKG_EINVAL, "unsupported triple pattern");
    }
    return WKG_OK;
  }

  int run_one_step(Request &r, void *stream) {
    wkg_step_desc d;
    int nc;
    int rc = plan(r, &d, &nc);
    if (rc) return rc;
    wkg_table *o = nullptr;
    uint64_t ex = 0;
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
          err = wkg_last_error(ctx);
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

const char *wke_last_error(const wke_engine *e) { return e ? e->err.c_str() : "null engine"; }
void wke_free(void *p) { free(p); }

int wke_run_query(wke_engine *e, const int64_t *cmd, int ncmd, wke_result *res) {
  if (!e || !cmd || !res || ncmd <= 0 || ncmd % 4) return WKG_EINVAL;
  memset(res, 0, sizeof *res);
  const auto t0 = Clock::now();
  Request r;
  r.cmd = cmd;
  r.ncmd = ncmd;
  while (!r.is_finished()) {
    int rc = e->run_one_step(r, nullptr);
    if (rc) {
      e->drop(r);
      res->status = rc;
      return rc;
    }
  }
  int rc = e->finish(r, res, nullptr);
  res->latency_s = std::chrono::duration<double>(Clock::now() - t0).count();
  return rc;
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
    e->err = wkg_last_error(e->ctx);
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

int wke_run_batched(wke_engine *e, const int64_t *cmd, const int64_t *cmd_off, int nq, wke_result *res) {
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

int wke_serve(wke_engine *e, const int64_t *cmd, const int64_t *cmd_off, int nq, const double *arrival_s,
              const uint8_t *heavy, uint64_t table_budget_bytes, wke_result *res, double *wall_s) {
  if (!e || !cmd || !cmd_off || !arrival_s || !res || nq < 0) return WKG_EINVAL;
  (void)heavy;
  (void)table_budget_bytes;
  std::vector<Request> reqs(nq);
  for (int q = 0; q < nq; q++) {
    memset(&res[q], 0, sizeof res[q]);
    reqs[q].cmd = cmd + cmd_off[q];
    reqs[q].ncmd = (int)(cmd_off[q + 1] - cmd_off[q]);
    if (reqs[q].ncmd <= 0 || reqs[q].ncmd % 4) reqs[q].status = WKG_EINVAL;
  }
  const auto t0 = Clock::now();
  auto now = [&] { return std::chrono::duration<double>(Clock::now() - t0).count(); };
  int next = 0, completed = 0, first = WKG_OK;
  std::vector<Request *> act;
  std::vector<int> act_q;
  while (completed < nq) {
    // admit every arrival so far (continuous batching: they join this round)
    const double t = now();
    while (next < nq && arrival_s[next] <= t) {
      act.push_back(&reqs[next]);
      act_q.push_back(next);
      next++;
    }
    if (act.empty()) {
      const double wait = arrival_s[next] - now();
      if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
      continue;
    }
    std::vector<Request *> run;
    for (Request *r : act)
      if (r->status == WKG_OK && !r->is_finished()) run.push_back(r);
    if (!run.empty()) batched_round(e, run, nullptr);
    // retire finished requests
    std::vector<Request *> keep;
    std::vector<int> keep_q;
    for (size_t i = 0; i < act.size(); i++) {
      Request &r = *act[i];
      if (r.status != WKG_OK || r.is_finished()) {
        if (r.status != WKG_OK) e->drop(r);
        e->finish(r, &res[act_q[i]], nullptr);
        res[act_q[i]].latency_s = now() - arrival_s[act_q[i]];
        if (res[act_q[i]].status && !first) first = res[act_q[i]].status;
        completed++;
      } else {
        keep.push_back(act[i]);
        keep_q.push_back(act_q[i]);
      }
    }
    act.swap(keep);
    act_q.swap(keep_q);
  }
  if (wall_s) *wall_s = now();
  return first;
}

}  // extern "C"
