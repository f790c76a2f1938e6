// Host gstore: synthetic generator, ID-triple file IO and the image builder.
//
// The builder produces the reference's key/value image (core/gstore.hpp:
// init 600-659, insert_normal 665-833, index maps 453-524 and 835-870,
// finish_pred_metas 957-978) deterministically:
//   * per-predicate lists sorted by (p,s,o) for OUT and (p,o,s) for IN, with
//     duplicates removed (core/dgraph.hpp:320-324);
//   * main partitions proportional to each id's triple count
//     (core/gstore.hpp:618-651);
//   * epochs in the reference's insert order: rdf:type (pid 1) first, then the
//     type indexes, then predicates ascending; within a predicate epoch the
//     (s,p,OUT) keys, the (o,p,IN) keys, then the index keys (0,p,IN) and
//     (0,p,OUT); each epoch's edges and indirect buckets form one contiguous
//     segment;
//   * index lists in ascending vertex order (the reference's order comes from
//     racing TBB maps and is not reproducible; ascending makes it canonical).
// Chain placement follows the cluster-chaining rule of insert_key
// (core/gstore.hpp:296-343): a bucket's keys fill slots 0..6 in insert order,
// the overflow goes to linked indirect buckets; indirect buckets of an epoch
// are numbered in main-bucket order.
#include "gstore_format.h"
#include "gstore_layout.h"
#include "../../include/wukong_gstore.h"

#include <dirent.h>
#include <omp.h>
#include <sys/stat.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <numeric>
#include <string>
#include <vector>

struct wks_triples {
  std::vector<uint32_t> s, p, o;
};

struct wks_gstore {
  uint64_t num_buckets = 0;
  uint64_t num_buckets_ext = 0;
  uint64_t num_keys = 0;
  std::vector<uint64_t> slots;  // 2 words per slot
  std::vector<uint32_t> edges;
  std::vector<wkg_pred_meta> metas;
};

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

constexpr uint64_t kVidBase = uint64_t(1) << wkg::kIdxBits;

// splitmix64 (Vigna): a seedable counter-based generator.
struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  uint64_t below(uint64_t n) {  // uniform in [0, n)
    return (uint64_t)(((unsigned __int128)next() * n) >> 64);
  }
};

uint64_t derive(uint64_t seed, uint64_t a, uint64_t b) {
  Rng r(seed ^ wkg::mix64(a * 0x9E3779B97F4A7C15ull + b + 1));
  r.next();
  return r.next();
}

// Keyed bijection of [0, n): a 4-round Feistel network over the smallest
// even-width power of two >= n, with cycle walking back into range.
struct Perm {
  uint64_t n, half, mask, keys[4];
  Perm(uint64_t n_, uint64_t seed) : n(n_) {
    int bits = 2;
    while ((uint64_t(1) << bits) < n) bits++;
    if (bits & 1) bits++;
    half = bits / 2;
    mask = (uint64_t(1) << half) - 1;
    Rng r(seed);
    for (auto &k : keys) k = r.next();
  }
  uint64_t enc(uint64_t x) const {
    uint64_t l = x >> half, r = x & mask;
    for (int i = 0; i < 4; i++) {
      uint64_t nl = r;
      r = l ^ (wkg::mix64(r ^ keys[i]) & mask);
      l = nl;
    }
    return (l << half) | r;
  }
  uint64_t dec(uint64_t y) const {  // one round of enc undone per step, last key first
    uint64_t l = y >> half, r = y & mask;
    for (int i = 3; i >= 0; i--) {
      uint64_t pr = l;
      l = r ^ (wkg::mix64(l ^ keys[i]) & mask);
      r = pr;
    }
    return (l << half) | r;
  }
  uint64_t operator()(uint64_t x) const {
    uint64_t y = enc(x);
    while (y >= n) y = enc(y);
    return y;
  }
  // inverse of operator(): walking dec back from y passes the out-of-range
  // values the forward walk passed and stops at its in-range start
  uint64_t inverse(uint64_t y) const {
    uint64_t x = dec(y);
    while (x >= n) x = dec(x);
    return x;
  }
};

struct ZipfDeg {
  double C, s;
  uint64_t cap, n;
  uint64_t operator()(uint64_t r) const {  // r >= 1
    double d = std::floor(C * std::pow((double)r, -s));
    uint64_t v = d <= 0 ? 0 : (uint64_t)d;
    if (cap && v > cap) v = cap;
    if (v > n) v = n;
    return v;
  }
  uint64_t max_rank() const {  // last rank with a nonzero degree
    double r = std::pow(C, 1.0 / s);
    uint64_t m = r >= (double)n ? n : (uint64_t)r;
    while (m < n && (*this)(m + 1) > 0) m++;
    while (m > 0 && (*this)(m) == 0) m--;
    return m;
  }
  double total() const {
    uint64_t m = max_rank();
    double sum = 0;
    uint64_t r0 = 1;
    if (cap) {
      double rc = std::pow(C / (double)cap, 1.0 / s);
      uint64_t nc = rc >= (double)m ? m : (uint64_t)rc;
      while (nc > 0 && (*this)(nc) < std::min<uint64_t>(cap, n)) nc--;
      sum += (double)nc * (double)std::min<uint64_t>(cap, n);
      r0 = nc + 1;
    }
    for (uint64_t r = r0; r <= m; r++) sum += (double)(*this)(r);
    return sum;
  }
};

// d distinct values of [0, n), ascending.
void sample_distinct(Rng &rng, uint64_t n, uint64_t d, std::vector<uint32_t> &out) {
  out.clear();
  if (d == 0) return;
  if (d * 2 > n) {
    std::vector<uint32_t> all(n);
    std::iota(all.begin(), all.end(), 0u);
    for (uint64_t i = 0; i < d; i++) std::swap(all[i], all[i + rng.below(n - i)]);
    out.assign(all.begin(), all.begin() + d);
    std::sort(out.begin(), out.end());
    return;
  }
  while (out.size() < d) {
    size_t need = d - out.size();
    for (size_t i = 0; i < need; i++) out.push_back((uint32_t)rng.below(n));
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
  }
}

// A linked list: `out` starts with m distinct values taken from the source
// list `src` (a seeded partial shuffle), then uniform distinct values of
// [0, n) are added until it holds d values; returned ascending.
void sample_linked(Rng &rng, uint64_t n, uint64_t d, uint64_t m, std::vector<uint32_t> src,
                   std::vector<uint32_t> &out) {
  out.clear();
  m = std::min<uint64_t>(std::min<uint64_t>(m, d), src.size());
  for (uint64_t i = 0; i < m; i++) {
    std::swap(src[i], src[i + rng.below(src.size() - i)]);
    out.push_back(src[i]);
  }
  std::sort(out.begin(), out.end());
  if (d * 2 > n) {
    std::vector<char> used(n, 0);
    for (uint32_t v : out) used[v] = 1;
    std::vector<uint32_t> rest;
    rest.reserve(n - out.size());
    for (uint64_t v = 0; v < n; v++)
      if (!used[v]) rest.push_back((uint32_t)v);
    for (uint64_t i = 0; out.size() < d; i++) {
      std::swap(rest[i], rest[i + rng.below(rest.size() - i)]);
      out.push_back(rest[i]);
    }
    std::sort(out.begin(), out.end());
    return;
  }
  while (out.size() < d) {
    size_t need = d - out.size();
    for (size_t i = 0; i < need; i++) out.push_back((uint32_t)rng.below(n));
    std::sort(out.begin(), out.end());
    out.erase(std::unique(out.begin(), out.end()), out.end());
  }
}

}  // namespace

extern "C" {

const char *wks_last_error(void) { return g_err.c_str(); }

int wks_generate(const wks_gen_params *prm, wks_triples **out) {
  if (!prm || !out) return fail(WKS_EINVAL, "null argument");
  *out = nullptr;
  const uint64_t N = prm->num_entities;
  if (N == 0 || kVidBase + N > (uint64_t(1) << 31))
    return fail(WKS_EINVAL, "num_entities must be in [1, 2^31 - 2^17)");
  if (prm->num_preds == 0) return fail(WKS_EINVAL, "num_preds must be >= 1");
  if (2ull + prm->num_preds + prm->num_types > kVidBase)
    return fail(WKS_EINVAL, "too many predicate and type ids for 17 bits");
  if (prm->num_types && !prm->type_sizes) return fail(WKS_EINVAL, "type_sizes is null");
  if (!(prm->zipf_s > 0)) return fail(WKS_EINVAL, "zipf_s must be > 0");
  const uint32_t P = prm->num_preds;
  const uint64_t Tp = prm->num_triples / P;
  const uint64_t cap = prm->degree_cap;
  if (cap && (double)cap * (double)N < (double)Tp)
    return fail(WKS_EINVAL, "degree cap too small for the requested triples");

  ZipfDeg z{0.0, prm->zipf_s, cap, N};
  if (Tp > 0) {
    double lo = 0, hi = (double)Tp;
    z.C = hi;
    while (z.total() < (double)Tp) {
      hi *= 2;
      z.C = hi;
      if (hi > 1e30) return fail(WKS_EINVAL, "cannot reach the requested triple count");
    }
    for (int it = 0; it < 100; it++) {
      double mid = 0.5 * (lo + hi);
      z.C = mid;
      if (z.total() < (double)Tp) lo = mid; else hi = mid;
    }
    z.C = hi;
  }
  const uint64_t rmax = Tp > 0 ? z.max_rank() : 0;
  if (prm->num_linked > P / 2) return fail(WKS_EINVAL, "num_linked must be <= num_preds / 2");
  if (prm->num_linked && !(prm->link_frac >= 0 && prm->link_frac <= 1))
    return fail(WKS_EINVAL, "link_frac must be in [0, 1]");
  if (prm->num_sibling && (!prm->sibling_src || !prm->sibling_dst))
    return fail(WKS_EINVAL, "sibling_src / sibling_dst is null");
  for (uint32_t k = 0; k < prm->num_sibling; k++)
    if (prm->sibling_src[k] < 2 || prm->sibling_src[k] >= 2 + P || prm->sibling_dst[k] < 2 ||
        prm->sibling_dst[k] >= 2 + P)
      return fail(WKS_EINVAL, "sibling plant %u names a predicate outside 2..%u", k, P + 1);

  // degree of rank r in predicate j (hub predicates differ only in rank 1)
  const uint64_t hub_deg = std::min<uint64_t>(prm->hub_degree, N);
  auto degree = [&](uint32_t j, uint64_t r) -> uint64_t {
    uint64_t d = r <= rmax ? z(r) : 0;
    if (r == 1 && j < prm->num_hubs && hub_deg > d) d = hub_deg;
    return d;
  };
  auto top_rank = [&](uint32_t j) -> uint64_t {
    return (j < prm->num_hubs && rmax == 0 && hub_deg > 0) ? 1 : rmax;
  };

  // per-predicate totals
  std::vector<uint64_t> pred_total(P), pred_base(P + 1, 0);
  uint64_t base_total = 0;
  for (uint64_t r = 1; r <= rmax; r++) base_total += z(r);
  for (uint32_t j = 0; j < P; j++) {
    uint64_t t = base_total;
    if (j < prm->num_hubs) {
      uint64_t d1 = rmax >= 1 ? z(1) : 0;
      if (hub_deg > d1) t += hub_deg - d1;
    }
    pred_total[j] = t;
    pred_base[j + 1] = pred_base[j] + t;
  }
  uint64_t type_total = 0;
  for (uint32_t k = 0; k < prm->num_types; k++) {
    if (prm->type_sizes[k] > N) return fail(WKS_EINVAL, "type %u larger than num_entities", k);
    type_total += prm->type_sizes[k];
  }
  // sibling plants: floor(d / 2) triples per subject of the source predicate
  std::vector<uint64_t> sib_base(prm->num_sibling + 1, pred_base[P] + type_total);
  for (uint32_t k = 0; k < prm->num_sibling; k++) {
    const uint32_t j = prm->sibling_src[k] - 2;
    uint64_t c = 0;
    for (uint64_t r = 1; r <= top_rank(j); r++) c += degree(j, r) / 2;
    sib_base[k + 1] = sib_base[k] + c;
  }
  const uint64_t total = sib_base[prm->num_sibling];

  wks_triples *t = new (std::nothrow) wks_triples;
  if (!t) return fail(WKS_ENOMEM, "out of memory");
  try {
    t->s.resize(total);
    t->p.resize(total);
    t->o.resize(total);
  } catch (...) {
    delete t;
    return fail(WKS_ENOMEM, "cannot allocate %llu triples", (unsigned long long)total);
  }

#pragma omp parallel
  {
    std::vector<uint32_t> objs, src_objs;
#pragma omp for schedule(dynamic, 1)
    for (uint32_t j = 0; j < P; j++) {
      Perm subj(N, derive(prm->seed, 1, j));
      // a linked predicate (odd index below 2 * num_linked) reads its source's ranks
      const bool linked = (j & 1) && j < 2 * prm->num_linked;
      Perm src_subj(N, derive(prm->seed, 1, linked ? j - 1 : j));
      std::vector<uint64_t> sib_pos;  // plants this predicate sources
      std::vector<uint32_t> sib_pid;
      for (uint32_t k = 0; k < prm->num_sibling; k++)
        if (prm->sibling_src[k] == 2 + j) {
          sib_pos.push_back(sib_base[k]);
          sib_pid.push_back(prm->sibling_dst[k]);
        }
      uint64_t pos = pred_base[j];
      const uint32_t pid = 2 + j;
      const uint64_t rtop = top_rank(j);
      for (uint64_t r = 1; r <= rtop; r++) {
        const uint64_t d = degree(j, r);
        if (!d) continue;
        const uint64_t sv = subj(r - 1);
        uint64_t d_src = 0, r_src = 0;
        if (linked) {
          r_src = src_subj.inverse(sv) + 1;
          d_src = r_src <= top_rank(j - 1) ? degree(j - 1, r_src) : 0;
        }
        if (d_src) {
          Rng rs(derive(prm->seed, 2 + (j - 1), r_src));
          sample_distinct(rs, N, d_src, src_objs);
          Rng rl(derive(prm->seed, (0x5ull << 40) | j, r));
          const uint64_t m = (uint64_t)std::ceil(prm->link_frac * (double)d);
          sample_linked(rl, N, d, m, src_objs, objs);
        } else {
          Rng rng(derive(prm->seed, 2 + j, r));
          sample_distinct(rng, N, d, objs);
        }
        const uint32_t s = (uint32_t)(kVidBase + sv);
        for (uint32_t ob : objs) {
          t->s[pos] = s;
          t->p[pos] = pid;
          t->o[pos] = (uint32_t)(kVidBase + ob);
          pos++;
        }
        for (size_t k = 0; k < sib_pos.size(); k++)
          for (size_t i = 0; i + 1 < objs.size(); i += 2) {
            const uint64_t q = sib_pos[k]++;
            t->s[q] = (uint32_t)(kVidBase + objs[i]);
            t->p[q] = sib_pid[k];
            t->o[q] = (uint32_t)(kVidBase + objs[i + 1]);
          }
      }
    }
  }
  uint64_t pos = pred_base[P];
  for (uint32_t k = 0; k < prm->num_types; k++) {
    Perm mem(N, derive(prm->seed, 0x7ull << 40, k));
    const uint32_t tid = 2 + P + k;
    for (uint64_t i = 0; i < prm->type_sizes[k]; i++) {
      t->s[pos] = (uint32_t)(kVidBase + mem(i));
      t->p[pos] = wkg::kTypeId;
      t->o[pos] = tid;
      pos++;
    }
  }
  *out = t;
  return WKS_OK;
}

int wks_triples_from_arrays(const uint32_t *s, const uint32_t *p, const uint32_t *o,
                            uint64_t n, wks_triples **out) {
  if (!out || (n && (!s || !p || !o))) return fail(WKS_EINVAL, "null argument");
  wks_triples *t = new (std::nothrow) wks_triples;
  if (!t) return fail(WKS_ENOMEM, "out of memory");
  try {
    t->s.assign(s, s + n);
    t->p.assign(p, p + n);
    t->o.assign(o, o + n);
  } catch (...) {
    delete t;
    return fail(WKS_ENOMEM, "out of memory");
  }
  *out = t;
  return WKS_OK;
}

}  // extern "C"

namespace {
// Appends the "s\tp\to" lines of one ID-format file to t.
int read_triples_into(const char *path, wks_triples *t) {
  FILE *f = fopen(path, "rb");
  if (!f) return fail(WKS_EIO, "cannot open %s: %s", path, strerror(errno));
  std::vector<char> buf(1 << 22);
  uint64_t v[3];
  int field = 0;
  bool in_num = false;
  uint64_t cur = 0, line = 1;
  size_t got;
  int rc = WKS_OK;
  auto end_line = [&]() -> bool {
    if (in_num) { v[field++] = cur; in_num = false; cur = 0; }
    if (field == 0) return true;  // blank line
    if (field != 3) { rc = fail(WKS_EIO, "%s:%llu: expected 3 fields", path, (unsigned long long)line); return false; }
    if (v[0] > 0xffffffffull || v[1] > 0xffffffffull || v[2] > 0xffffffffull) {
      rc = fail(WKS_ERANGE, "%s:%llu: id exceeds 32 bits", path, (unsigned long long)line);
      return false;
    }
    t->s.push_back((uint32_t)v[0]);
    t->p.push_back((uint32_t)v[1]);
    t->o.push_back((uint32_t)v[2]);
    field = 0;
    return true;
  };
  while (rc == WKS_OK && (got = fread(buf.data(), 1, buf.size(), f)) > 0) {
    for (size_t i = 0; i < got; i++) {
      char c = buf[i];
      if (c >= '0' && c <= '9') {
        if (field >= 3) { rc = fail(WKS_EIO, "%s:%llu: too many fields", path, (unsigned long long)line); break; }
        cur = cur * 10 + (uint64_t)(c - '0');
        if (cur > 0xffffffffull * 10) { rc = fail(WKS_ERANGE, "%s:%llu: id too large", path, (unsigned long long)line); break; }
        in_num = true;
      } else if (c == '\n') {
        if (!end_line()) break;
        line++;
      } else if (c == '\t' || c == ' ' || c == '\r') {
        if (in_num) { v[field++] = cur; in_num = false; cur = 0; }
      } else {
        rc = fail(WKS_EIO, "%s:%llu: unexpected character", path, (unsigned long long)line);
        break;
      }
    }
  }
  if (rc == WKS_OK) end_line();
  fclose(f);
  return rc;
}

// The "<string>\t<id>" lines of an ID-mapping file (str_index / str_normal,
// datagen/generate_data.cpp:155-167): calls fn(id) per line.
template <class Fn>
int read_id_map(const std::string &path, Fn fn) {
  FILE *f = fopen(path.c_str(), "rb");
  if (!f) return fail(WKS_EIO, "cannot open %s: %s", path.c_str(), strerror(errno));
  std::vector<char> line(1 << 16);
  uint64_t n = 0;
  int rc = WKS_OK;
  while (fgets(line.data(), (int)line.size(), f)) {
    n++;
    char *tab = strrchr(line.data(), '\t');
    if (!tab) {
      if (line[0] == '\n' || line[0] == '\0') continue;
      rc = fail(WKS_EIO, "%s:%llu: expected <string>\\t<id>", path.c_str(), (unsigned long long)n);
      break;
    }
    char *end = nullptr;
    const unsigned long long id = strtoull(tab + 1, &end, 10);
    if (end == tab + 1 || id > 0xffffffffull) {
      rc = fail(WKS_EIO, "%s:%llu: bad id", path.c_str(), (unsigned long long)n);
      break;
    }
    fn((uint32_t)id);
  }
  fclose(f);
  return rc;
}
}  // namespace

extern "C" {

int wks_triples_read_file(const char *path, wks_triples **out) {
  if (!path || !out) return fail(WKS_EINVAL, "null argument");
  wks_triples *t = new wks_triples;
  int rc = read_triples_into(path, t);
  if (rc != WKS_OK) { delete t; return rc; }
  *out = t;
  return WKS_OK;
}

int wks_dataset_read(const char *dir, wks_triples **out, wks_dataset_info *info) {
  if (!dir || !out) return fail(WKS_EINVAL, "null argument");
  *out = nullptr;
  DIR *d = opendir(dir);
  if (!d) return fail(WKS_EIO, "cannot open directory %s: %s", dir, strerror(errno));
  std::vector<std::string> files;
  bool has_index = false, has_normal = false;
  for (struct dirent *e; (e = readdir(d));) {
    const std::string n = e->d_name;
    if (n.empty() || n[0] == '.') continue;
    if (n.compare(0, 3, "id_") == 0) files.push_back(n);
    if (n == "str_index") has_index = true;
    if (n == "str_normal") has_normal = true;
  }
  closedir(d);
  if (files.empty()) return fail(WKS_EIO, "no id_* files in %s", dir);
  std::sort(files.begin(), files.end());
  const std::string base = std::string(dir) + "/";
  wks_dataset_info inf{};
  inf.files = files.size();
  wks_triples *t = new wks_triples;
  for (const std::string &n : files)
    if (int rc = read_triples_into((base + n).c_str(), t)) {
      delete t;
      return rc;
    }
  inf.triples = t->s.size();
  // index ids (predicates and types): the id space of the pred metas
  std::vector<char> index_id;
  if (has_index) {
    int rc = read_id_map(base + "str_index", [&](uint32_t id) {
      if (id >= wkg::kMaxIdx + 1u) return;
      if (id >= index_id.size()) index_id.resize(id + 1, 0);
      if (id && !index_id[id]) inf.index_ids++;
      index_id[id] = 1;
      inf.max_index_id = std::max<uint64_t>(inf.max_index_id, id);
    });
    if (rc) { delete t; return rc; }
    for (size_t i = 0; i < t->s.size(); i++) {
      const uint32_t pid = t->p[i];
      const bool ok = pid < index_id.size() && index_id[pid] &&
                      (pid != wkg::kTypeId || (t->o[i] < index_id.size() && index_id[t->o[i]]));
      if (!ok) {
        delete t;
        return fail(WKS_EINVAL, "triple %llu names an index id missing from str_index", (unsigned long long)i);
      }
    }
  }
  if (has_normal) {
    int rc = read_id_map(base + "str_normal", [&](uint32_t id) {
      inf.normal_ids++;
      inf.max_normal_id = std::max<uint64_t>(inf.max_normal_id, id);
    });
    if (rc) { delete t; return rc; }
  }
  inf.has_str_index = has_index;
  inf.has_str_normal = has_normal;
  if (info) *info = inf;
  *out = t;
  return WKS_OK;
}

int wks_dataset_write(const wks_triples *t, const char *dir, int files) {
  if (!t || !dir || files < 1) return fail(WKS_EINVAL, "null argument or no files");
  if (mkdir(dir, 0755) != 0 && errno != EEXIST) return fail(WKS_EIO, "cannot create %s: %s", dir, strerror(errno));
  const std::string base = std::string(dir) + "/";
  const size_t n = t->s.size();
  for (int k = 0; k < files; k++) {
    char name[64];
    snprintf(name, sizeof name, "id_%04d.nt", k);
    FILE *f = fopen((base + name).c_str(), "wb");
    if (!f) return fail(WKS_EIO, "cannot open %s%s: %s", base.c_str(), name, strerror(errno));
    for (size_t i = n * k / files; i < n * (k + 1) / files; i++) fprintf(f, "%u\t%u\t%u\n", t->s[i], t->p[i], t->o[i]);
    if (fclose(f) != 0) return fail(WKS_EIO, "write to %s%s failed", base.c_str(), name);
  }
  // ID mappings with made-up names in the reference's layout: __PREDICATE__ 0,
  // rdf:type 1, then every predicate and type id; every normal vertex id
  std::vector<char> idx, normal;
  auto mark = [](std::vector<char> &v, uint32_t id) {
    if (id >= v.size()) v.resize((size_t)id + 1, 0);
    v[id] = 1;
  };
  for (size_t i = 0; i < n; i++) {
    mark(idx, t->p[i]);
    mark(normal, t->s[i]);
    if (t->p[i] == wkg::kTypeId) mark(idx, t->o[i]);
    else mark(normal, t->o[i]);
  }
  FILE *f = fopen((base + "str_index").c_str(), "wb");
  if (!f) return fail(WKS_EIO, "cannot open %sstr_index: %s", base.c_str(), strerror(errno));
  fprintf(f, "__PREDICATE__\t0\n<http://www.w3.org/1999/02/22-rdf-syntax-ns#type>\t1\n");
  for (size_t id = 2; id < idx.size(); id++)
    if (idx[id]) fprintf(f, "<http://wukong.example/index/%zu>\t%zu\n", id, id);
  if (fclose(f) != 0) return fail(WKS_EIO, "write to %sstr_index failed", base.c_str());
  f = fopen((base + "str_normal").c_str(), "wb");
  if (!f) return fail(WKS_EIO, "cannot open %sstr_normal: %s", base.c_str(), strerror(errno));
  for (size_t id = 0; id < normal.size(); id++)
    if (normal[id]) fprintf(f, "<http://wukong.example/vertex/%zu>\t%zu\n", id, id);
  if (fclose(f) != 0) return fail(WKS_EIO, "write to %sstr_normal failed", base.c_str());
  return WKS_OK;
}

int wks_triples_write_file(const wks_triples *t, const char *path) {
  if (!t || !path) return fail(WKS_EINVAL, "null argument");
  FILE *f = fopen(path, "wb");
  if (!f) return fail(WKS_EIO, "cannot open %s: %s", path, strerror(errno));
  for (size_t i = 0; i < t->s.size(); i++)
    fprintf(f, "%u\t%u\t%u\n", t->s[i], t->p[i], t->o[i]);
  if (fclose(f) != 0) return fail(WKS_EIO, "write to %s failed", path);
  return WKS_OK;
}

uint64_t wks_triples_count(const wks_triples *t) { return t ? t->s.size() : 0; }
const uint32_t *wks_triples_s(const wks_triples *t) { return t ? t->s.data() : nullptr; }
const uint32_t *wks_triples_p(const wks_triples *t) { return t ? t->p.data() : nullptr; }
const uint32_t *wks_triples_o(const wks_triples *t) { return t ? t->o.data() : nullptr; }
void wks_triples_free(wks_triples *t) { delete t; }

}  // extern "C"

namespace {

using wkg::Epoch;

struct Grouped {
  // per predicate id: [so_off[p], so_off[p+1]) of (s<<32|o), deduplicated
  std::vector<uint64_t> so, os;
  std::vector<uint64_t> so_off, os_off, so_cnt, os_cnt;
  std::vector<uint64_t> nsubj, nobj;  // distinct subjects / objects per id
};

// Calls fn(key, size, off_in_epoch) for every key of the epoch in insertion
// order, and (if edges != nullptr) writes the epoch's edges.
template <class Fn>
void walk_epoch(const Epoch &e, const Grouped &g, const std::vector<uint64_t> &type_off,
                const std::vector<uint32_t> &type_members, uint32_t *edges, Fn &&fn) {
  using namespace wkg;
  if (e.kind == 0) {
    const uint64_t *a = g.so.data() + g.so_off[kTypeId];
    const uint64_t n = g.so_cnt[kTypeId];
    uint64_t i = 0;
    while (i < n) {
      uint64_t s = a[i] >> 32, j = i;
      while (j < n && (a[j] >> 32) == s) {
        if (edges) edges[e.edge_base + j] = (uint32_t)a[j];
        j++;
      }
      fn(make_key(s, kTypeId, kDirOut), j - i, i);
      i = j;
    }
  } else if (e.kind == 1) {
    const uint64_t b = type_off[e.id], n = type_off[e.id + 1] - b;
    if (edges) std::memcpy(edges + e.edge_base, type_members.data() + b, n * sizeof(uint32_t));
    fn(make_key(0, e.id, kDirIn), n, 0);
  } else {
    const uint32_t p = e.id;
    const uint64_t *a = g.so.data() + g.so_off[p];
    const uint64_t na = g.so_cnt[p];
    const uint64_t *b = g.os.data() + g.os_off[p];
    const uint64_t nb = g.os_cnt[p];
    uint32_t *subj = edges ? edges + e.edge_base + na + nb : nullptr;
    uint32_t *obj = edges ? subj + g.nsubj[p] : nullptr;
    uint64_t i = 0, ns = 0;
    while (i < na) {  // (s,p,OUT) -> objects
      uint64_t s = a[i] >> 32, j = i;
      while (j < na && (a[j] >> 32) == s) {
        if (edges) edges[e.edge_base + j] = (uint32_t)a[j];
        j++;
      }
      fn(make_key(s, p, kDirOut), j - i, i);
      if (subj) subj[ns] = (uint32_t)s;
      ns++;
      i = j;
    }
    i = 0;
    uint64_t no = 0;
    while (i < nb) {  // (o,p,IN) -> subjects
      uint64_t o = b[i] >> 32, j = i;
      while (j < nb && (b[j] >> 32) == o) {
        if (edges) edges[e.edge_base + na + j] = (uint32_t)b[j];
        j++;
      }
      fn(make_key(o, p, kDirIn), j - i, na + i);
      if (obj) obj[no] = (uint32_t)o;
      no++;
      i = j;
    }
    if (na) fn(make_key(0, p, kDirIn), g.nsubj[p], na + nb);
    if (nb) fn(make_key(0, p, kDirOut), g.nobj[p], na + nb + g.nsubj[p]);
  }
}

}  // namespace

extern "C" {

int wks_build(const wks_triples *t, const wks_build_params *bp, wks_gstore **out) {
  using namespace wkg;
  if (!t || !out) return fail(WKS_EINVAL, "null argument");
  *out = nullptr;
  const double load = bp && bp->main_load > 0 ? bp->main_load : 0.75;
  if (load > 1.0) return fail(WKS_EINVAL, "main_load must be in (0, 1]");
  if (bp && bp->num_threads > 0) omp_set_num_threads(bp->num_threads);
  const uint64_t n = t->s.size();

  // ---- validate ids, find the id range --------------------------------------
  uint32_t max_id = 1;
  bool bad = false;
  std::string why;
#pragma omp parallel for reduction(max : max_id)
  for (uint64_t i = 0; i < n; i++) {
    const uint32_t s = t->s[i], p = t->p[i], o = t->o[i];
    bool ok = s >= kVidBase && s < (1u << 31);
    if (p == kTypeId) ok = ok && is_tpid(o);
    else ok = ok && is_tpid(p) && o >= kVidBase && o < (1u << 31);
    if (!ok) {
#pragma omp critical
      {
        if (!bad) {
          char b[160];
          snprintf(b, sizeof b, "invalid triple %llu: (%u, %u, %u)", (unsigned long long)i, s, p, o);
          why = b;
        }
        bad = true;
      }
    }
    uint32_t m = p == kTypeId ? o : p;
    if (m > max_id) max_id = m;
  }
  if (bad) return fail(WKS_EINVAL, "%s", why.c_str());
  const uint32_t num_metas = max_id + 1;

  Grouped g;
  try {
    g.so_off.assign(num_metas + 1, 0);
    g.os_off.assign(num_metas + 1, 0);
    g.so_cnt.assign(num_metas, 0);
    g.os_cnt.assign(num_metas, 0);
    g.nsubj.assign(num_metas, 0);
    g.nobj.assign(num_metas, 0);
  } catch (...) {
    return fail(WKS_ENOMEM, "out of memory");
  }

  // ---- group by predicate (counting sort), then sort+dedup each group --------
  {
    const int nt = omp_get_max_threads();
    std::vector<std::vector<uint64_t>> hs(nt, std::vector<uint64_t>(num_metas, 0)),
        ho(nt, std::vector<uint64_t>(num_metas, 0));
#pragma omp parallel
    {
      const int tid = omp_get_thread_num();
      auto &a = hs[tid];
      auto &b = ho[tid];
#pragma omp for schedule(static)
      for (uint64_t i = 0; i < n; i++) {
        a[t->p[i]]++;
        if (t->p[i] != kTypeId) b[t->p[i]]++;
      }
    }
    // per-thread cursors in (thread-major within predicate) order keeps the
    // scatter deterministic for a static schedule
    std::vector<std::vector<uint64_t>> cs(nt, std::vector<uint64_t>(num_metas)),
        co(nt, std::vector<uint64_t>(num_metas));
    uint64_t ps = 0, po = 0;
    for (uint32_t p = 0; p < num_metas; p++) {
      g.so_off[p] = ps;
      g.os_off[p] = po;
      for (int k = 0; k < nt; k++) {
        cs[k][p] = ps;
        ps += hs[k][p];
        co[k][p] = po;
        po += ho[k][p];
      }
    }
    g.so_off[num_metas] = ps;
    g.os_off[num_metas] = po;
    try {
      g.so.resize(ps);
      g.os.resize(po);
    } catch (...) {
      return fail(WKS_ENOMEM, "cannot allocate grouped triples");
    }
#pragma omp parallel
    {
      const int tid = omp_get_thread_num();
      auto &a = cs[tid];
      auto &b = co[tid];
#pragma omp for schedule(static)
      for (uint64_t i = 0; i < n; i++) {
        const uint64_t s = t->s[i], p = t->p[i], o = t->o[i];
        g.so[a[p]++] = (s << 32) | o;
        if (p != kTypeId) g.os[b[p]++] = (o << 32) | s;
      }
    }
  }
#pragma omp parallel for schedule(dynamic, 1)
  for (uint64_t k = 0; k < 2ull * num_metas; k++) {
    const bool is_so = k < num_metas;
    const uint32_t p = (uint32_t)(is_so ? k : k - num_metas);
    auto &v = is_so ? g.so : g.os;
    const auto &off = is_so ? g.so_off : g.os_off;
    uint64_t *b = v.data() + off[p], *e = v.data() + off[p + 1];
    std::sort(b, e);
    uint64_t *u = std::unique(b, e);
    (is_so ? g.so_cnt : g.os_cnt)[p] = (uint64_t)(u - b);
    uint64_t d = 0;
    for (uint64_t *x = b; x < u; x++)
      if (x == b || (x[0] >> 32) != (x[-1] >> 32)) d++;
    (is_so ? g.nsubj : g.nobj)[p] = d;
  }

  // ---- type index lists: members of t in ascending vertex order -------------
  std::vector<uint64_t> type_cnt(num_metas, 0), type_off(num_metas + 1, 0);
  std::vector<uint32_t> type_members;
  {
    const uint64_t *a = g.so.data() + g.so_off[kTypeId];
    const uint64_t na = g.so_cnt[kTypeId];
    for (uint64_t i = 0; i < na; i++) type_cnt[(uint32_t)a[i]]++;
    for (uint32_t p = 0; p < num_metas; p++) type_off[p + 1] = type_off[p] + type_cnt[p];
    type_members.resize(type_off[num_metas]);
    std::vector<uint64_t> cur(type_off.begin(), type_off.end() - 1);
    for (uint64_t i = 0; i < na; i++) type_members[cur[(uint32_t)a[i]]++] = (uint32_t)(a[i] >> 32);
    for (uint32_t p = 2; p < num_metas; p++)
      if (type_cnt[p] && (g.so_cnt[p] || g.os_cnt[p]))
        return fail(WKS_EINVAL, "id %u is used both as a predicate and as a type", p);
  }

  // ---- partition sizes and epochs (core/gstore.hpp:618-651) -----------------
  Layout lay;
  {
    IdCounts cnt;
    cnt.so_cnt = g.so_cnt;
    cnt.os_cnt = g.os_cnt;
    cnt.nsubj = g.nsubj;
    cnt.nobj = g.nobj;
    cnt.type_cnt = type_cnt;
    plan_layout(cnt, num_metas, load, lay);
  }
  wks_gstore *G = new (std::nothrow) wks_gstore;
  if (!G) return fail(WKS_ENOMEM, "out of memory");
  G->num_buckets = lay.num_buckets;
  G->num_keys = lay.num_keys;
  G->metas = lay.metas;
  std::vector<Epoch> &ep = lay.ep;
  const uint64_t total_edges = lay.total_edges;
  try {
    G->edges.resize(total_edges);
  } catch (...) {
    delete G;
    return fail(WKS_ENOMEM, "cannot allocate %llu edges", (unsigned long long)total_edges);
  }

  // ---- pass 1: write edges, count the overflow buckets of every epoch ------------
#pragma omp parallel
  {
    std::vector<uint32_t> cnt;
#pragma omp for schedule(dynamic, 1)
    for (size_t k = 0; k < ep.size(); k++) {
      Epoch &e = ep[k];
      const wkg_pred_meta &m = G->metas[e.id];
      const uint64_t psz = m.partition_sz, mag = fastmod_magic(psz);
      cnt.assign(psz, 0);
      walk_epoch(e, g, type_off, type_members, G->edges.data(),
                 [&](uint64_t key, uint64_t, uint64_t) { cnt[fastmod(mix64(key), psz, mag)]++; });
      uint64_t ext = 0;
      for (uint64_t b = 0; b < psz; b++)
        ext += chain_buckets(cnt[b]);
      e.ext_count = ext;
    }
  }
  const uint64_t ext_total = place_epochs(lay);
  G->metas = lay.metas;
  G->num_buckets_ext = ext_total;
  const uint64_t nb_all = G->num_buckets + ext_total;
  try {
    G->slots.assign(nb_all * kAssoc * 2, 0);
  } catch (...) {
    delete G;
    return fail(WKS_ENOMEM, "cannot allocate %llu buckets", (unsigned long long)nb_all);
  }

  // ---- pass 2: place keys ---------------------------------------------------------
  std::atomic<bool> too_big{false};
#pragma omp parallel
  {
    std::vector<uint32_t> cur;
    std::vector<uint64_t> chain;
#pragma omp for schedule(dynamic, 1)
    for (size_t k = 0; k < ep.size(); k++) {
      const Epoch &e = ep[k];
      const wkg_pred_meta &m = G->metas[e.id];
      const uint64_t psz = m.partition_sz, mag = fastmod_magic(psz);
      cur.assign(psz, 0);
      walk_epoch(e, g, type_off, type_members, nullptr,
                 [&](uint64_t key, uint64_t, uint64_t) { cur[fastmod(mix64(key), psz, mag)]++; });
      chain.assign(psz, 0);
      uint64_t next = G->num_buckets + e.ext_base;
      for (uint64_t b = 0; b < psz; b++) {
        chain[b] = next;
        next += chain_buckets(cur[b]);
      }
      // link chains: main bucket -> first indirect -> ... -> last (link 0)
      for (uint64_t b = 0; b < psz; b++) {
        const uint64_t nchain = chain_buckets(cur[b]);
        if (!nchain) continue;
        uint64_t prev = m.main_hdr_start + b;
        for (uint64_t j = 0; j < nchain; j++) {
          G->slots[(prev * kAssoc + 7) * 2] = make_key(chain[b] + j, 0, 0);
          prev = chain[b] + j;
        }
      }
      std::fill(cur.begin(), cur.end(), 0);
      walk_epoch(e, g, type_off, type_members, nullptr, [&](uint64_t key, uint64_t size, uint64_t off) {
        const uint64_t b = fastmod(mix64(key), psz, mag);
        const uint32_t r = cur[b]++;
        uint64_t bucket, slot;
        if (r < 7) {
          bucket = m.main_hdr_start + b;
          slot = r;
        } else {
          bucket = chain[b] + (r - 7) / 7;
          slot = (r - 7) % 7;
        }
        if (size > kSizeMask) too_big = true;
        uint64_t *sp = &G->slots[(bucket * kAssoc + slot) * 2];
        sp[0] = key;
        sp[1] = make_ptr(size & kSizeMask, e.edge_base + off);
      });
    }
  }
  if (too_big) {
    delete G;
    return fail(WKS_ERANGE, "an edge list exceeds 2^28 entries");
  }
  if (total_edges >= (uint64_t(1) << kOffBits)) {
    delete G;
    return fail(WKS_ERANGE, "edge region exceeds 2^36 entries");
  }
  *out = G;
  return WKS_OK;
}

void wks_gstore_free(wks_gstore *g) { delete g; }
uint64_t wks_num_slots(const wks_gstore *g) { return g ? g->slots.size() / 2 : 0; }
uint64_t wks_num_buckets(const wks_gstore *g) { return g ? g->num_buckets : 0; }
uint64_t wks_num_buckets_ext(const wks_gstore *g) { return g ? g->num_buckets_ext : 0; }
uint64_t wks_num_keys(const wks_gstore *g) { return g ? g->num_keys : 0; }
const uint64_t *wks_slots(const wks_gstore *g) { return g ? g->slots.data() : nullptr; }
uint64_t wks_num_edges(const wks_gstore *g) { return g ? g->edges.size() : 0; }
const uint32_t *wks_edges(const wks_gstore *g) { return g ? g->edges.data() : nullptr; }
int wks_num_metas(const wks_gstore *g) { return g ? (int)g->metas.size() : 0; }
const void *wks_metas(const wks_gstore *g) { return g ? g->metas.data() : nullptr; }

int wks_get_edges(const wks_gstore *g, uint64_t vid, uint32_t pid, int dir,
                  const uint32_t **edges, uint64_t *size) {
  using namespace wkg;
  if (!g || !edges || !size) return fail(WKS_EINVAL, "null argument");
  *edges = nullptr;
  *size = 0;
  if (pid >= g->metas.size() || vid > kMaxVid || (dir != 0 && dir != 1)) return WKS_OK;
  const wkg_pred_meta &m = g->metas[pid];
  if (!m.partition_sz) return WKS_OK;
  const uint64_t key = make_key(vid, pid, (uint64_t)dir);
  uint64_t b = m.main_hdr_start + mix64(key) % m.partition_sz;
  const uint64_t nb = g->slots.size() / (2 * kAssoc);
  while (b < nb) {
    const uint64_t *sp = &g->slots[b * kAssoc * 2];
    for (int i = 0; i < kAssoc - 1; i++)
      if (sp[2 * i] == key) {
        *size = ptr_size(sp[2 * i + 1]);
        *edges = g->edges.data() + ptr_off(sp[2 * i + 1]);
        return WKS_OK;
      }
    const uint64_t link = sp[2 * (kAssoc - 1)];
    if (!link) return WKS_OK;
    b = key_vid(link);
  }
  return fail(WKS_ERANGE, "corrupt chain");
}

int wks_gstore_save(const wks_gstore *g, const char *path) {
  if (!g || !path) return fail(WKS_EINVAL, "null argument");
  FILE *f = fopen(path, "wb");
  if (!f) return fail(WKS_EIO, "cannot open %s: %s", path, strerror(errno));
  const uint64_t hdr[8] = {0x31474B5753ull /* "SWKG1" */, 1, g->num_buckets, g->num_buckets_ext,
                           g->num_keys, g->edges.size(), g->metas.size(), g->slots.size()};
  bool ok = fwrite(hdr, sizeof hdr, 1, f) == 1 &&
            fwrite(g->metas.data(), sizeof(wkg_pred_meta), g->metas.size(), f) == g->metas.size() &&
            fwrite(g->slots.data(), 8, g->slots.size(), f) == g->slots.size() &&
            fwrite(g->edges.data(), 4, g->edges.size(), f) == g->edges.size();
  ok = (fclose(f) == 0) && ok;
  return ok ? WKS_OK : fail(WKS_EIO, "write to %s failed", path);
}

int wks_gstore_load(const char *path, wks_gstore **out) {
  if (!path || !out) return fail(WKS_EINVAL, "null argument");
  FILE *f = fopen(path, "rb");
  if (!f) return fail(WKS_EIO, "cannot open %s: %s", path, strerror(errno));
  uint64_t hdr[8];
  if (fread(hdr, sizeof hdr, 1, f) != 1 || hdr[0] != 0x31474B5753ull || hdr[1] != 1) {
    fclose(f);
    return fail(WKS_EIO, "%s is not a gstore image", path);
  }
  wks_gstore *g = new (std::nothrow) wks_gstore;
  if (!g) { fclose(f); return fail(WKS_ENOMEM, "out of memory"); }
  g->num_buckets = hdr[2];
  g->num_buckets_ext = hdr[3];
  g->num_keys = hdr[4];
  try {
    g->edges.resize(hdr[5]);
    g->metas.resize(hdr[6]);
    g->slots.resize(hdr[7]);
  } catch (...) {
    fclose(f);
    delete g;
    return fail(WKS_ENOMEM, "cannot allocate image");
  }
  bool ok = fread(g->metas.data(), sizeof(wkg_pred_meta), g->metas.size(), f) == g->metas.size() &&
            fread(g->slots.data(), 8, g->slots.size(), f) == g->slots.size() &&
            fread(g->edges.data(), 4, g->edges.size(), f) == g->edges.size();
  fclose(f);
  if (!ok || g->slots.size() != (g->num_buckets + g->num_buckets_ext) * wkg::kAssoc * 2) {
    delete g;
    return fail(WKS_EIO, "%s is truncated or inconsistent", path);
  }
  *out = g;
  return WKS_OK;
}

}  // extern "C"
