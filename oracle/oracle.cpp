// ORACLE -- test infrastructure only.
//
// A CPU restatement of the reference engine's pattern-step semantics, used as
// the parity checker for the GPU path and as the timed CPU baseline
// ("port" kind). Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg load this library; the product never does.
//
// It shares no code with the product: key packing, the hash and the bucket
// walk are written out again here from the reference.
//
//   key packing         core/gstore.hpp:102-149 (ikey_t), 151-176 (iptr_t)
//   hash                utils/mymath.hpp:62-71 (hash_u64) via ikey_t::hash
//   bucket_for          core/gstore.hpp:289-293
//   get_vertex_local    core/gstore.hpp:389-409
//   get_edges_local     core/gstore.hpp:432-450
//   get_index_edges     core/gstore.hpp:1136-1139
//   const_to_unknown    core/engine.hpp:96-122
//   known_to_unknown    core/engine.hpp:126-160
//   known_to_known      core/engine.hpp:162-194
//   known_to_const      core/engine.hpp:196-225
//   index_to_unknown    core/engine.hpp:227-255 (tid 0, global_mt_threshold 1:
//                       the full list, as scripts/config:18 sets it)
//   execute_one_step    core/engine.hpp:500-570
//   request state       core/query.hpp:45-175 (variable_type 141-149,
//                       var2column 151-154, start_from_index 120-139)
//
// Deviations, all on paths where the reference has undefined behaviour or
// aborts: table sizes are 64-bit (the reference uses int), unsupported
// patterns return an error code instead of assert(false), and a predicate id
// outside the image or with an empty partition is a miss (the reference
// divides by zero).
//
// "Expanded edges" are counted the way the benchmark defines them: edges
// emitted by index_to_unknown / const_to_unknown / known_to_unknown, and
// edges examined by the early-exit scans of known_to_known / known_to_const.
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <atomic>
#include <chrono>
#include <thread>
#include <vector>

namespace {

struct Meta {  // pred_meta_t, core/rdf_meta.hpp:31-54
  uint64_t partition_sz, main_hdr_start, main_hdr_end, indrct_hdr_start, indrct_hdr_end,
      num_buckets, edge_start, edge_end;
};

struct Image {
  const uint64_t *slots;  // (key, ptr) per slot
  uint64_t num_slots;
  const uint32_t *edges;
  uint64_t num_edges;
  const Meta *metas;
  int num_metas;
};

const int64_t NBITS_DIR = 1, NBITS_IDX = 17;
const int ASSOC = 8;
enum { PREDICATE_ID = 0, TYPE_ID = 1 };

uint64_t ikey(uint64_t vid, uint64_t pid, uint64_t dir) {
  uint64_t r = 0;  // same assembly as ikey_t::hash
  r += vid;
  r <<= NBITS_IDX;
  r += pid;
  r <<= NBITS_DIR;
  r += dir;
  return r;
}

uint64_t hash_u64(uint64_t key) {
  key = (~key) + (key << 21);
  key = key ^ (key >> 24);
  key = (key + (key << 3)) + (key << 8);
  key = key ^ (key >> 14);
  key = (key + (key << 2)) + (key << 4);
  key = key ^ (key >> 28);
  key = key + (key << 31);
  return key;
}

bool is_tpid(int64_t id) { return (id > 1) && (id < (1 << NBITS_IDX)); }

// get_vertex_local + get_edges_local: returns the list and its size (0 = miss)
const uint32_t *get_edges(const Image &g, int64_t vid, int64_t dir, int64_t pid, uint64_t *sz) {
  *sz = 0;
  if (pid < 0 || pid >= g.num_metas || vid < 0 || (dir != 0 && dir != 1)) return nullptr;
  const Meta &m = g.metas[pid];
  if (m.partition_sz == 0) return nullptr;
  const uint64_t key = ikey((uint64_t)vid, (uint64_t)pid, (uint64_t)dir);
  uint64_t bucket_id = m.main_hdr_start + hash_u64(key) % m.partition_sz;
  while (true) {
    for (int i = 0; i < ASSOC; i++) {
      uint64_t slot_id = bucket_id * ASSOC + i;
      if (slot_id >= g.num_slots) return nullptr;  // corrupt image
      const uint64_t k = g.slots[2 * slot_id];
      if (i < ASSOC - 1) {
        if (k == key) {
          const uint64_t ptr = g.slots[2 * slot_id + 1];
          *sz = ptr & ((1ull << 28) - 1);
          return g.edges + (ptr >> 28);
        }
      } else {
        if (k == 0) return nullptr;  // not found
        bucket_id = k >> (NBITS_IDX + NBITS_DIR);
        break;
      }
    }
  }
}

enum var_type { known_var, unknown_var, const_var };

struct Req {  // request_or_reply, the fields the engine uses
  const int64_t *cmd;
  int ncmd;
  int step = 0;
  int col_num = 0;
  std::vector<int> table;
  uint64_t expanded = 0;
  int threads = 1;  // known-start steps over row ranges on this many threads

  bool is_finished() const { return step * 4 >= ncmd; }
  bool start_from_index() const { return is_tpid(cmd[0]); }
  var_type variable_type(int64_t vid) const {
    if (vid >= 0) return const_var;
    if ((-vid) > col_num) return unknown_var;
    return known_var;
  }
  int64_t var2column(int64_t vid) const { return (-vid) - 1; }
  uint64_t rows() const { return col_num == 0 ? 0 : table.size() / col_num; }
  int get(uint64_t r, int64_t c) const { return table[col_num * r + c]; }
  void append_row_to(uint64_t r, std::vector<int> &out) const {
    for (int c = 0; c < col_num; c++) out.push_back(get(r, c));
  }
};

// error codes
const int OK = 0, E_PLAN = -10, E_UNSUPPORTED = -11;

int const_to_unknown(const Image &g, Req &req) {
  int64_t start = req.cmd[req.step * 4], pid = req.cmd[req.step * 4 + 1];
  int64_t d = req.cmd[req.step * 4 + 2], end = req.cmd[req.step * 4 + 3];
  if (!(req.col_num == 0 && req.var2column(end) == 0)) return E_PLAN;
  std::vector<int> upd;
  uint64_t sz = 0;
  const uint32_t *res = get_edges(g, start, d, pid, &sz);
  for (uint64_t k = 0; k < sz; k++) upd.push_back((int)res[k]);
  req.expanded += sz;
  req.table.swap(upd);
  req.col_num = 1;
  req.step++;
  return OK;
}

// The three known-start handlers below process input rows [r0, r1) into
// `upd` (appending) and return the step's work count over those rows; a
// step is the concatenation of its row ranges, in order, since each row's
// output depends on that row alone (run_known).
uint64_t known_to_unknown_rows(const Image &g, const Req &req, uint64_t r0, uint64_t r1,
                               std::vector<int> &upd) {
  int64_t start = req.cmd[req.step * 4], pid = req.cmd[req.step * 4 + 1];
  int64_t d = req.cmd[req.step * 4 + 2];
  const int64_t sc = req.var2column(start);
  uint64_t expanded = 0;
  for (uint64_t i = r0; i < r1; i++) {
    int64_t prev_id = req.get(i, sc);
    uint64_t sz = 0;
    const uint32_t *res = get_edges(g, prev_id, d, pid, &sz);
    for (uint64_t k = 0; k < sz; k++) {
      req.append_row_to(i, upd);
      upd.push_back((int)res[k]);
    }
    expanded += sz;
  }
  return expanded;
}

uint64_t known_to_known_rows(const Image &g, const Req &req, uint64_t r0, uint64_t r1,
                             std::vector<int> &upd) {
  int64_t start = req.cmd[req.step * 4], pid = req.cmd[req.step * 4 + 1];
  int64_t d = req.cmd[req.step * 4 + 2], end = req.cmd[req.step * 4 + 3];
  const int64_t sc = req.var2column(start), ec = req.var2column(end);
  uint64_t expanded = 0;
  for (uint64_t i = r0; i < r1; i++) {
    int64_t prev_id = req.get(i, sc);
    uint64_t sz = 0;
    const uint32_t *res = get_edges(g, prev_id, d, pid, &sz);
    int64_t end2 = req.get(i, ec);
    uint64_t k = 0;
    for (; k < sz; k++) {
      if ((int64_t)res[k] == end2) {
        req.append_row_to(i, upd);
        break;
      }
    }
    expanded += k < sz ? k + 1 : sz;
  }
  return expanded;
}

uint64_t known_to_const_rows(const Image &g, const Req &req, uint64_t r0, uint64_t r1,
                             std::vector<int> &upd) {
  int64_t start = req.cmd[req.step * 4], pid = req.cmd[req.step * 4 + 1];
  int64_t d = req.cmd[req.step * 4 + 2], end = req.cmd[req.step * 4 + 3];
  const int64_t sc = req.var2column(start);
  uint64_t expanded = 0;
  for (uint64_t i = r0; i < r1; i++) {
    int64_t prev_id = req.get(i, sc);
    uint64_t sz = 0;
    const uint32_t *res = get_edges(g, prev_id, d, pid, &sz);
    uint64_t k = 0;
    for (; k < sz; k++) {
      if ((int64_t)res[k] == end) {
        req.append_row_to(i, upd);
        break;
      }
    }
    expanded += k < sz ? k + 1 : sz;
  }
  return expanded;
}

typedef uint64_t (*RowsFn)(const Image &, const Req &, uint64_t, uint64_t, std::vector<int> &);

// Runs one known-start handler over all rows: on one thread in one range, or
// (req.threads > 1 and enough rows) in contiguous row ranges on several
// threads whose outputs are concatenated in range order.
void run_known(const Image &g, Req &req, RowsFn fn, std::vector<int> &upd) {
  const uint64_t n = req.rows();
  const int T = req.threads;
  if (T <= 1 || n < (1u << 16)) {
    req.expanded += fn(g, req, 0, n, upd);
    return;
  }
  const uint64_t parts = (uint64_t)T * 8;
  std::vector<std::vector<int>> outs(parts);
  std::vector<uint64_t> ex(parts, 0);
  std::atomic<uint64_t> next{0};
  auto work = [&]() {
    for (uint64_t c; (c = next.fetch_add(1)) < parts;)
      ex[c] = fn(g, req, n * c / parts, n * (c + 1) / parts, outs[c]);
  };
  std::vector<std::thread> ts;
  for (int i = 1; i < T; i++) ts.emplace_back(work);
  work();
  for (auto &t : ts) t.join();
  uint64_t total = 0;
  for (auto &o : outs) total += o.size();
  upd.reserve(total);
  for (uint64_t c = 0; c < parts; c++) {
    upd.insert(upd.end(), outs[c].begin(), outs[c].end());
    std::vector<int>().swap(outs[c]);
    req.expanded += ex[c];
  }
}

int known_to_unknown(const Image &g, Req &req) {
  int64_t end = req.cmd[req.step * 4 + 3];
  if (req.col_num != req.var2column(end)) return E_PLAN;
  std::vector<int> upd;
  run_known(g, req, known_to_unknown_rows, upd);
  req.col_num = req.col_num + 1;
  req.table.swap(upd);
  req.step++;
  return OK;
}

int known_to_known(const Image &g, Req &req) {
  std::vector<int> upd;
  run_known(g, req, known_to_known_rows, upd);
  req.table.swap(upd);
  req.step++;
  return OK;
}

int known_to_const(const Image &g, Req &req) {
  std::vector<int> upd;
  run_known(g, req, known_to_const_rows, upd);
  req.table.swap(upd);
  req.step++;
  return OK;
}

int index_to_unknown(const Image &g, Req &req) {
  int64_t idx = req.cmd[req.step * 4];
  int64_t d = req.cmd[req.step * 4 + 2], var = req.cmd[req.step * 4 + 3];
  if (!(req.col_num == 0 && req.var2column(var) == 0)) return E_PLAN;
  std::vector<int> upd;
  uint64_t sz = 0;
  const uint32_t *res = get_edges(g, 0, d, idx, &sz);  // get_index_edges_local
  for (uint64_t k = 0; k < sz; k += 1) upd.push_back((int)res[k]);
  req.expanded += sz;
  req.table.swap(upd);
  req.col_num = 1;
  req.step++;
  return OK;
}

// returns 1 if a step ran, 0 if finished, <0 on error
int execute_one_step(const Image &g, Req &req) {
  if (req.is_finished()) return 0;
  if (req.step == 0 && req.start_from_index()) {
    if (!(req.cmd[1] == PREDICATE_ID || req.cmd[1] == TYPE_ID)) return E_PLAN;
    int rc = index_to_unknown(g, req);
    return rc < 0 ? rc : 1;
  }
  int64_t start = req.cmd[req.step * 4], predicate = req.cmd[req.step * 4 + 1];
  int64_t end = req.cmd[req.step * 4 + 3];
  if (predicate < 0) return E_UNSUPPORTED;  // needs VERSATILE
  int rc;
  var_type ts = req.variable_type(start), te = req.variable_type(end);
  if (ts == const_var && te == unknown_var) rc = const_to_unknown(g, req);
  else if (ts == known_var && te == const_var) rc = known_to_const(g, req);
  else if (ts == known_var && te == known_var) rc = known_to_known(g, req);
  else if (ts == known_var && te == unknown_var) rc = known_to_unknown(g, req);
  else rc = E_UNSUPPORTED;
  return rc < 0 ? rc : 1;
}

// Order-sensitive 64-bit digest of a row-major table: fmix64(rows * 256 +
// cols) plus, over cells i = 0, 1, ..., fmix64((i + 1) * golden ^ cell_i)
// (cell as its u32 bit pattern), all mod 2^64. oracle.table_digest computes
// the same for a numpy table.
uint64_t fmix64(uint64_t x) {  // MurmurHash3's 64-bit finalizer
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdull;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ull;
  x ^= x >> 33;
  return x;
}

uint64_t table_digest(const std::vector<int> &t, uint64_t rows, int cols) {
  uint64_t h = fmix64(rows * 256 + (uint64_t)cols);
  for (uint64_t i = 0; i < t.size(); i++)
    h += fmix64((i + 1) * 0x9E3779B97F4A7C15ull ^ (uint64_t)(uint32_t)t[i]);
  return h;
}

Image make_image(const uint64_t *slots, uint64_t num_slots, const uint32_t *edges, uint64_t num_edges,
                 const void *metas, int num_metas) {
  return Image{slots, num_slots, edges, num_edges, (const Meta *)metas, num_metas};
}

}  // namespace

extern "C" {

// Runs steps [0, stop_step) of a query (stop_step <= 0: all), the
// known-start steps on `threads` threads over row ranges (<= 1: one thread;
// the result is the same). On success the result table is returned in *out
// (free with or_free), *rows x *cols.
int or_run_query_mt(const uint64_t *slots, uint64_t num_slots, const uint32_t *edges,
                    uint64_t num_edges, const void *metas, int num_metas, const int64_t *cmd, int ncmd,
                    int stop_step, int threads, int32_t **out, uint64_t *rows, int *cols,
                    uint64_t *expanded, double *seconds) {
  Image g = make_image(slots, num_slots, edges, num_edges, metas, num_metas);
  Req req;
  req.cmd = cmd;
  req.ncmd = ncmd;
  req.threads = threads;
  if (ncmd <= 0 || ncmd % 4) return E_PLAN;
  auto t0 = std::chrono::steady_clock::now();
  while (stop_step <= 0 || req.step < stop_step) {
    int rc = execute_one_step(g, req);
    if (rc < 0) return rc;
    if (rc == 0) break;
  }
  auto t1 = std::chrono::steady_clock::now();
  if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
  *rows = req.rows();
  *cols = req.col_num;
  *expanded = req.expanded;
  *out = nullptr;
  if (!req.table.empty()) {
    *out = (int32_t *)malloc(req.table.size() * sizeof(int32_t));
    if (!*out) return -2;
    memcpy(*out, req.table.data(), req.table.size() * sizeof(int32_t));
  }
  return 0;
}

int or_run_query(const uint64_t *slots, uint64_t num_slots, const uint32_t *edges,
                 uint64_t num_edges, const void *metas, int num_metas, const int64_t *cmd, int ncmd,
                 int stop_step, int32_t **out, uint64_t *rows, int *cols, uint64_t *expanded,
                 double *seconds) {
  return or_run_query_mt(slots, num_slots, edges, num_edges, metas, num_metas, cmd, ncmd, stop_step, 1,
                         out, rows, cols, expanded, seconds);
}

// One step on a given input table (the state a request has before `step`),
// a known-start step on `threads` threads over row ranges (<= 1: one).
int or_run_step(const uint64_t *slots, uint64_t num_slots, const uint32_t *edges,
                uint64_t num_edges, const void *metas, int num_metas, const int64_t *cmd, int ncmd,
                int step, const int32_t *in, uint64_t in_rows, int in_cols, int threads, int32_t **out,
                uint64_t *rows, int *cols, uint64_t *expanded) {
  Image g = make_image(slots, num_slots, edges, num_edges, metas, num_metas);
  Req req;
  req.cmd = cmd;
  req.ncmd = ncmd;
  req.step = step;
  req.threads = threads;
  req.col_num = in_cols;
  if (ncmd <= 0 || ncmd % 4 || step < 0 || step * 4 >= ncmd) return E_PLAN;
  req.table.assign(in, in + in_rows * (uint64_t)in_cols);
  int rc = execute_one_step(g, req);
  if (rc < 0) return rc;
  *rows = req.rows();
  *cols = req.col_num;
  *expanded = req.expanded;
  *out = nullptr;
  if (!req.table.empty()) {
    *out = (int32_t *)malloc(req.table.size() * sizeof(int32_t));
    if (!*out) return -2;
    memcpy(*out, req.table.data(), req.table.size() * sizeof(int32_t));
  }
  return 0;
}

// Throughput baseline: nq queries (query q = cmd[cmd_off[q] .. cmd_off[q+1]))
// run to completion on `threads` threads, each query on one thread like one
// reference Engine thread (core/engine.hpp:659-725). Per-query latencies and
// result sizes are written to lat_s[q], rows[q]; returns wall seconds in *wall.
// digest (may be NULL) gets each result table's table_digest.
int or_run_many(const uint64_t *slots, uint64_t num_slots, const uint32_t *edges,
                uint64_t num_edges, const void *metas, int num_metas, const int64_t *cmd,
                const int64_t *cmd_off, int nq, int threads, double *lat_s, uint64_t *rows,
                uint64_t *expanded, uint64_t *digest, double *wall) {
  Image g = make_image(slots, num_slots, edges, num_edges, metas, num_metas);
  std::atomic<int> next{0}, err{0};
  auto work = [&]() {
    while (true) {
      int q = next.fetch_add(1);
      if (q >= nq) return;
      Req req;
      req.cmd = cmd + cmd_off[q];
      req.ncmd = (int)(cmd_off[q + 1] - cmd_off[q]);
      auto t0 = std::chrono::steady_clock::now();
      int rc = 1;
      while (rc == 1) rc = execute_one_step(g, req);
      auto t1 = std::chrono::steady_clock::now();
      if (rc < 0) err = rc;
      lat_s[q] = std::chrono::duration<double>(t1 - t0).count();
      rows[q] = req.rows();
      expanded[q] = req.expanded;
      if (digest) digest[q] = table_digest(req.table, req.rows(), req.col_num);
    }
  };
  auto t0 = std::chrono::steady_clock::now();
  std::vector<std::thread> ts;
  for (int i = 1; i < threads; i++) ts.emplace_back(work);
  work();
  for (auto &t : ts) t.join();
  *wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return err.load();
}

// Raw lookup, for tests of the image itself.
int or_get_edges(const uint64_t *slots, uint64_t num_slots, const uint32_t *edges,
                 uint64_t num_edges, const void *metas, int num_metas, int64_t vid, int64_t dir,
                 int64_t pid, uint64_t *off, uint64_t *size) {
  Image g = make_image(slots, num_slots, edges, num_edges, metas, num_metas);
  const uint32_t *res = get_edges(g, vid, dir, pid, size);
  *off = res ? (uint64_t)(res - edges) : 0;
  return res ? 1 : 0;
}

uint64_t or_hash_u64(uint64_t x) { return hash_u64(x); }

// inverse_hash_u64, utils/mymath.hpp:73-107: undoes hash_u64 step by step.
uint64_t or_inverse_hash_u64(uint64_t key) {
  uint64_t tmp;
  tmp = key - (key << 31);  // key = key + (key << 31)
  key = key - (tmp << 31);
  tmp = key ^ key >> 28;    // key = key ^ (key >> 28)
  key = key ^ tmp >> 28;
  key *= 14933078535860113213u;  // key *= 21
  tmp = key ^ key >> 14;    // key = key ^ (key >> 14)
  tmp = key ^ tmp >> 14;
  tmp = key ^ tmp >> 14;
  key = key ^ tmp >> 14;
  key *= 15244667743933553977u;  // key *= 265
  tmp = key ^ key >> 24;    // key = key ^ (key >> 24)
  key = key ^ tmp >> 24;
  tmp = ~key;               // key = (~key) + (key << 21)
  tmp = ~(key - (tmp << 21));
  tmp = ~(key - (tmp << 21));
  key = ~(key - (tmp << 21));
  return key;
}

void or_free(void *p) { free(p); }

}  // extern "C"
