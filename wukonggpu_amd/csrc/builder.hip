// Device gstore builder: the image wks_build makes (gstore.cpp), slot for
// slot, built in HBM from ID triples and installed as the context's resident
// graph (the reference builds on the host with 11 threads and TBB,
// core/gstore.hpp:600-659, 665-870, 957-978; core/dgraph.hpp:268-330).
//
//   1. triples -> (p, s<<32|o) and (p, o<<32|s) pairs, radix-sorted by p then
//      by the pair (two stable passes), duplicates dropped;
//   2. per-id counts -> partitions and epochs on the host (gstore_layout.h,
//      the host builder's own code);
//   3. edges: each id's object lists, subject lists, subject and object
//      index lists, and type member lists, scattered from the sorted pairs;
//   4. keys in the host builder's insertion order (epoch by epoch: runs of a
//      subject, runs of an object, then the index keys), each hashed to its
//      main bucket; a stable sort by bucket gives every key its rank in its
//      bucket by insertion order, which fixes its slot: ranks 0..6 in the
//      main bucket, the rest in the bucket's chain of indirect buckets,
//      numbered per epoch in main-bucket order.
// Sorts and scans are rocPRIM's; the kernels here are the gstore rules.
#include <hip/hip_runtime.h>
#include <string.h>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <chrono>
#include <vector>

#include "../../include/wukong_gpu.h"
#include "gstore_layout.h"
#include "wkg_internal.h"

using namespace wkg;

namespace {

constexpr uint32_t kNone = 0xffffffffu;
constexpr uint64_t kVidBase = uint64_t(1) << kIdxBits;

__device__ __forceinline__ uint64_t gidx() { return blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; }
__device__ __forceinline__ uint64_t gstride() { return (uint64_t)gridDim.x * blockDim.x; }

// Validity (gstore.cpp wks_build: s a vertex; rdf:type objects are type ids,
// other predicates are ids and their objects vertices), the id range, and
// the two pair arrays.
__global__ void k_b_pairs(const uint32_t *s, const uint32_t *p, const uint32_t *o, uint64_t n, uint64_t *so_key,
                          uint32_t *so_pid, uint64_t *os_key, uint32_t *os_pid, unsigned long long *first_bad,
                          uint32_t *max_id) {
  uint32_t mx = 1;
  for (uint64_t i = gidx(); i < n; i += gstride()) {
    const uint32_t si = s[i], pi = p[i], oi = o[i];
    bool ok = si >= kVidBase && si < (1u << 31);
    if (pi == (uint32_t)kTypeId) ok = ok && is_tpid(oi);
    else ok = ok && is_tpid(pi) && oi >= kVidBase && oi < (1u << 31);
    if (!ok) atomicMin(first_bad, (unsigned long long)i);
    const uint32_t m = pi == (uint32_t)kTypeId ? oi : pi;
    mx = m > mx ? m : mx;
    so_key[i] = ((uint64_t)si << 32) | oi;
    so_pid[i] = pi;
    os_key[i] = ((uint64_t)oi << 32) | si;
    os_pid[i] = pi;
  }
  atomicMax(max_id, mx);
}

// keep[i]: first of its (pid, pair) run.
__global__ void k_b_first(const uint64_t *key, const uint32_t *pid, uint64_t n, uint32_t *keep) {
  for (uint64_t i = gidx(); i < n; i += gstride())
    keep[i] = (i == 0 || pid[i] != pid[i - 1] || key[i] != key[i - 1]) ? 1u : 0u;
}

__global__ void k_b_compact(const uint64_t *key, const uint32_t *pid, const uint32_t *keep, const uint32_t *pos,
                            uint64_t n, uint64_t *key_out, uint32_t *pid_out) {
  for (uint64_t i = gidx(); i < n; i += gstride())
    if (keep[i]) {
      key_out[pos[i]] = key[i];
      pid_out[pos[i]] = pid[i];
    }
}

// Over deduplicated pairs: first index of each id (start[pid]), run flags of
// the high word within an id (run[i]), and for rdf:type pairs in (type,
// member) order the first index of each type (tstart[type]).
__global__ void k_b_bounds(const uint64_t *key, const uint32_t *pid, uint64_t n, uint32_t *start, uint32_t *run,
                           uint32_t *tstart) {
  for (uint64_t i = gidx(); i < n; i += gstride()) {
    const bool first_id = i == 0 || pid[i] != pid[i - 1];
    if (first_id) start[pid[i]] = (uint32_t)i;
    const bool first_hi = first_id || (key[i] >> 32) != (key[i - 1] >> 32);
    run[i] = first_hi ? 1u : 0u;
    if (tstart && first_hi && pid[i] == (uint32_t)kTypeId) tstart[(uint32_t)(key[i] >> 32)] = (uint32_t)i;
  }
}

// at[id] = v[start[id]] for ids present (the run prefix at each id's start).
__global__ void k_b_gather(const uint32_t *start, const uint32_t *v, uint32_t nids, uint32_t *at) {
  for (uint64_t i = gidx(); i < nids; i += gstride()) at[i] = start[i] == kNone ? 0u : v[start[i]];
}

// Run starts' positions, in run order (runpos[rank] = i), and the end.
__global__ void k_b_runpos(const uint32_t *run, const uint32_t *rpre, uint64_t n, uint32_t *runpos, uint32_t nruns) {
  for (uint64_t i = gidx(); i < n; i += gstride())
    if (run[i]) runpos[rpre[i]] = (uint32_t)i;
  if (gidx() == 0) runpos[nruns] = (uint32_t)n;
}

// Per id on the device (ids < num_metas).
struct IdInfo {
  uint64_t edge_base;   // the id's epoch's first edge
  uint64_t key_base;    // the id's epoch's first key (insertion order)
  uint64_t main_start;  // first main bucket
  uint64_t psz, magic;  // partition size and its fastmod constant
  uint64_t chain_base;  // num_buckets + the epoch's ext_base, minus the chain prefix at main_start
  uint32_t so_start, os_start, rs_at, ro_at;  // pair and run prefixes at the id's start
  uint32_t na, nb, nsubj, pad;                // so / os pairs, distinct subjects
};

// Edges and keys from the (s, p, o) pairs: rdf:type objects (epoch kind 0)
// or a predicate's objects (kind 2, first segment), its subject index list,
// and the keys (s,p,OUT) of subject runs.
__global__ void k_b_from_so(const uint64_t *key, const uint32_t *pid, const uint32_t *run, const uint32_t *rpre,
                            const uint32_t *runpos, uint64_t n, const IdInfo *ids, uint32_t *edges, uint64_t *kkey,
                            uint64_t *ksize, uint64_t *koff, uint32_t *kid) {
  for (uint64_t i = gidx(); i < n; i += gstride()) {
    const uint32_t p = pid[i];
    const IdInfo &d = ids[p];
    const uint64_t j = i - d.so_start;
    edges[d.edge_base + j] = (uint32_t)key[i];
    if (!run[i]) continue;
    const uint32_t r = rpre[i], k = r - d.rs_at;
    const uint64_t s = key[i] >> 32;
    if (p != (uint32_t)kTypeId) edges[d.edge_base + d.na + d.nb + k] = (uint32_t)s;
    const uint64_t idx = d.key_base + k;
    kkey[idx] = make_key(s, p, kDirOut);
    ksize[idx] = runpos[r + 1] - i;
    koff[idx] = j;
    kid[idx] = p;
  }
}

// From the (o, p, s) pairs: type members (rdf:type pairs, in (type, member)
// order), a predicate's subjects (second segment), its object index list,
// and the keys (o,p,IN) of object runs.
__global__ void k_b_from_os(const uint64_t *key, const uint32_t *pid, const uint32_t *run, const uint32_t *rpre,
                            const uint32_t *runpos, uint64_t n, const IdInfo *ids, const uint32_t *tstart,
                            uint32_t *edges, uint64_t *kkey, uint64_t *ksize, uint64_t *koff, uint32_t *kid) {
  for (uint64_t i = gidx(); i < n; i += gstride()) {
    const uint32_t p = pid[i];
    const uint64_t hi = key[i] >> 32;
    if (p == (uint32_t)kTypeId) {
      edges[ids[hi].edge_base + (i - tstart[hi])] = (uint32_t)key[i];
      continue;
    }
    const IdInfo &d = ids[p];
    const uint64_t j = i - d.os_start;
    edges[d.edge_base + d.na + j] = (uint32_t)key[i];
    if (!run[i]) continue;
    const uint32_t r = rpre[i], k = r - d.ro_at;
    edges[d.edge_base + d.na + d.nb + d.nsubj + k] = (uint32_t)hi;
    const uint64_t idx = d.key_base + d.nsubj + k;
    kkey[idx] = make_key(hi, p, kDirIn);
    ksize[idx] = runpos[r + 1] - i;
    koff[idx] = d.na + j;
    kid[idx] = p;
  }
}

// Keys made on the host (index keys): scattered to their places.
struct HostKey {
  uint64_t idx, key, size, off;
  uint32_t id, pad;
};
__global__ void k_b_host_keys(const HostKey *hk, uint64_t n, uint64_t *kkey, uint64_t *ksize, uint64_t *koff,
                              uint32_t *kid) {
  for (uint64_t i = gidx(); i < n; i += gstride()) {
    const HostKey h = hk[i];
    kkey[h.idx] = h.key;
    ksize[h.idx] = h.size;
    koff[h.idx] = h.off;
    kid[h.idx] = h.id;
  }
}

// Main bucket of every key; the key's own index rides along the sort.
__global__ void k_b_bucket(const uint64_t *kkey, const uint32_t *kid, uint64_t nk, const IdInfo *ids, uint32_t *kb,
                           uint32_t *kidx) {
  for (uint64_t i = gidx(); i < nk; i += gstride()) {
    const IdInfo &d = ids[kid[i]];
    kb[i] = (uint32_t)(d.main_start + (d.psz ? fastmod(mix64(kkey[i]), d.psz, d.magic) : 0));
    kidx[i] = (uint32_t)i;
  }
}

// Over keys sorted by bucket (stable: insertion order within a bucket):
// keys per bucket and each bucket's first position.
__global__ void k_b_bucket_runs(const uint32_t *kb, uint64_t nk, uint32_t *cnt, uint32_t *first) {
  for (uint64_t i = gidx(); i < nk; i += gstride()) {
    if (i == 0 || kb[i] != kb[i - 1]) first[kb[i]] = (uint32_t)i;
    atomicAdd(&cnt[kb[i]], 1u);
  }
}

__global__ void k_b_chains(const uint32_t *cnt, uint64_t nb, uint64_t *nchain) {
  for (uint64_t b = gidx(); b < nb; b += gstride()) nchain[b] = chain_buckets(cnt[b]);
}

// Places every key: rank r in its bucket by insertion order, slot r of the
// main bucket for r < 7, else slot (r-7)%7 of the ((r-7)/7)-th chained
// bucket; rank 0 also writes its bucket's chain links.
__global__ void k_b_place(const uint32_t *kb, const uint32_t *kidx, uint64_t nk, const uint32_t *first,
                          const uint32_t *cnt, const uint64_t *cpre, const uint64_t *kkey, const uint64_t *ksize,
                          const uint64_t *koff, const uint32_t *kid, const IdInfo *ids, uint64_t *slots,
                          uint32_t *too_big) {
  for (uint64_t i = gidx(); i < nk; i += gstride()) {
    const uint64_t b = kb[i];
    const uint32_t r = (uint32_t)(i - first[b]);
    const uint32_t k = kidx[i];
    const IdInfo &d = ids[kid[k]];
    const uint64_t cs = d.chain_base + cpre[b];  // first chained bucket of b
    uint64_t bucket, slot;
    if (r < 7) {
      bucket = b;
      slot = r;
    } else {
      bucket = cs + (r - 7) / 7;
      slot = (r - 7) % 7;
    }
    const uint64_t size = ksize[k];
    if (size > kSizeMask) *too_big = 1u;
    uint64_t *sp = slots + (bucket * kAssoc + slot) * 2;
    sp[0] = kkey[k];
    sp[1] = make_ptr(size & kSizeMask, d.edge_base + koff[k]);
    if (r == 0) {
      const uint64_t nch = chain_buckets(cnt[b]);
      uint64_t prev = b;
      for (uint64_t j = 0; j < nch; j++) {
        slots[(prev * kAssoc + 7) * 2] = make_key(cs + j, 0, 0);
        prev = cs + j;
      }
    }
  }
}

unsigned grid_for(uint64_t n) { return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, 8192)); }

// Device buffers of one build, freed on every exit path.
struct Bufs {
  std::vector<void *> v;
  template <class T>
  hipError_t get(T **p, uint64_t count) {
    *p = nullptr;
    hipError_t e = hipMalloc((void **)p, std::max<uint64_t>(count, 1) * sizeof(T));
    if (e == hipSuccess) v.push_back(*p);
    return e;
  }
  void put(void *p) {
    auto it = std::find(v.begin(), v.end(), p);
    if (it != v.end()) {
      (void)hipFree(p);
      v.erase(it);
    }
  }
  ~Bufs() {
    for (void *p : v) (void)hipFree(p);
  }
};

// Stable sort of (pid, key) pairs by pid then key: LSD, key first.
hipError_t sort_pairs(uint64_t *&key, uint32_t *&pid, uint64_t n, uint32_t max_id, Bufs &B, hipStream_t s) {
  uint64_t *key2;
  uint32_t *pid2;
  hipError_t e = B.get(&key2, n);
  if (e == hipSuccess) e = B.get(&pid2, n);
  if (e != hipSuccess) return e;
  size_t t1 = 0, t2 = 0;
  int pid_bits = 1;
  while ((1u << pid_bits) <= max_id) pid_bits++;
  e = rocprim::radix_sort_pairs(nullptr, t1, key, key2, pid, pid2, n, 0, 63, s);
  if (e == hipSuccess) e = rocprim::radix_sort_pairs(nullptr, t2, pid2, pid, key2, key, n, 0, pid_bits, s);
  if (e != hipSuccess) return e;
  char *tmp;
  if ((e = B.get(&tmp, std::max(t1, t2))) != hipSuccess) return e;
  e = rocprim::radix_sort_pairs(tmp, t1, key, key2, pid, pid2, n, 0, 63, s);
  if (e == hipSuccess) e = rocprim::radix_sort_pairs(tmp, t2, pid2, pid, key2, key, n, 0, pid_bits, s);
  B.put(tmp);
  B.put(key2);
  B.put(pid2);
  return e;
}

hipError_t excl_scan_u32(const uint32_t *in, uint32_t *out, uint64_t n, Bufs &B, hipStream_t s) {
  size_t t = 0;
  hipError_t e = rocprim::exclusive_scan(nullptr, t, in, out, 0u, n, rocprim::plus<uint32_t>(), s);
  char *tmp;
  if (e == hipSuccess) e = B.get(&tmp, t);
  if (e == hipSuccess) e = rocprim::exclusive_scan(tmp, t, in, out, 0u, n, rocprim::plus<uint32_t>(), s);
  if (e == hipSuccess) B.put(tmp);
  return e;
}

hipError_t excl_scan_u64(const uint64_t *in, uint64_t *out, uint64_t n, Bufs &B, hipStream_t s) {
  size_t t = 0;
  hipError_t e = rocprim::exclusive_scan(nullptr, t, in, out, (uint64_t)0, n, rocprim::plus<uint64_t>(), s);
  char *tmp;
  if (e == hipSuccess) e = B.get(&tmp, t);
  if (e == hipSuccess) e = rocprim::exclusive_scan(tmp, t, in, out, (uint64_t)0, n, rocprim::plus<uint64_t>(), s);
  if (e == hipSuccess) B.put(tmp);
  return e;
}

template <class T>
hipError_t down(std::vector<T> &h, const T *d, uint64_t n, hipStream_t s) {
  h.resize(n);
  hipError_t e = hipMemcpyAsync(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost, s);
  return e == hipSuccess ? hipStreamSynchronize(s) : e;
}

// Sorted, deduplicated pairs of one kind and what the layout needs of them.
struct Side {
  uint64_t *key = nullptr;
  uint32_t *pid = nullptr;
  uint32_t *run = nullptr, *rpre = nullptr, *runpos = nullptr;
  uint64_t n = 0;
  uint32_t nruns = 0;
  std::vector<uint32_t> start, rat;  // per id: first pair (kNone if absent), run prefix there
};

// Sort, drop duplicates, find id starts and high-word runs.
hipError_t prepare_side(Side &sd, uint64_t n, uint32_t num_metas, uint32_t *tstart, Bufs &B, hipStream_t s) {
  hipError_t e = sort_pairs(sd.key, sd.pid, n, num_metas - 1, B, s);
  uint32_t *keep, *pos;
  if (e == hipSuccess) e = B.get(&keep, n + 1);
  if (e == hipSuccess) e = B.get(&pos, n + 1);
  if (e != hipSuccess) return e;
  k_b_first<<<grid_for(n), 256, 0, s>>>(sd.key, sd.pid, n, keep);
  if ((e = hipMemsetAsync(keep + n, 0, 4, s)) != hipSuccess) return e;
  if ((e = excl_scan_u32(keep, pos, n + 1, B, s)) != hipSuccess) return e;
  uint32_t m = 0;
  if ((e = hipMemcpyAsync(&m, pos + n, 4, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
  uint64_t *k2;
  uint32_t *p2;
  if ((e = B.get(&k2, m)) != hipSuccess || (e = B.get(&p2, m)) != hipSuccess) return e;
  k_b_compact<<<grid_for(n), 256, 0, s>>>(sd.key, sd.pid, keep, pos, n, k2, p2);
  B.put(sd.key);
  B.put(sd.pid);
  B.put(keep);
  B.put(pos);
  sd.key = k2;
  sd.pid = p2;
  sd.n = m;
  uint32_t *start;
  if ((e = B.get(&start, num_metas)) != hipSuccess) return e;
  if ((e = hipMemsetAsync(start, 0xff, (uint64_t)num_metas * 4, s)) != hipSuccess) return e;
  if ((e = B.get(&sd.run, m + 1)) != hipSuccess || (e = B.get(&sd.rpre, m + 1)) != hipSuccess) return e;
  k_b_bounds<<<grid_for(m), 256, 0, s>>>(sd.key, sd.pid, m, start, sd.run, tstart);
  if ((e = hipMemsetAsync(sd.run + m, 0, 4, s)) != hipSuccess) return e;
  if ((e = excl_scan_u32(sd.run, sd.rpre, m + 1, B, s)) != hipSuccess) return e;
  if ((e = hipMemcpyAsync(&sd.nruns, sd.rpre + m, 4, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
  if ((e = B.get(&sd.runpos, (uint64_t)sd.nruns + 1)) != hipSuccess) return e;
  k_b_runpos<<<grid_for(m), 256, 0, s>>>(sd.run, sd.rpre, m, sd.runpos, sd.nruns);
  uint32_t *rat;
  if ((e = B.get(&rat, num_metas)) != hipSuccess) return e;
  k_b_gather<<<grid_for(num_metas), 256, 0, s>>>(start, sd.rpre, num_metas, rat);
  if ((e = down(sd.start, (const uint32_t *)start, num_metas, s)) != hipSuccess) return e;
  if ((e = down(sd.rat, (const uint32_t *)rat, num_metas, s)) != hipSuccess) return e;
  B.put(start);
  B.put(rat);
  return hipGetLastError();
}

// Per-id pair counts and distinct high words from the id starts (ids appear
// in ascending order) and the run prefix at each start.
void side_counts(const Side &sd, uint32_t num_metas, std::vector<uint64_t> &cnt, std::vector<uint64_t> &distinct) {
  cnt.assign(num_metas, 0);
  distinct.assign(num_metas, 0);
  uint32_t prev = kNone;
  for (uint32_t p = 0; p <= num_metas; p++) {
    if (p < num_metas && sd.start[p] == kNone) continue;
    if (prev != kNone) {
      const uint64_t end = p < num_metas ? sd.start[p] : sd.n;
      const uint64_t rend = p < num_metas ? sd.rat[p] : sd.nruns;
      cnt[prev] = end - sd.start[prev];
      distinct[prev] = rend - sd.rat[prev];
    }
    prev = p;
  }
}

}  // namespace

extern "C" {

int wkg_build_graph(wkg_ctx *ctx, const uint32_t *hs, const uint32_t *hp, const uint32_t *ho, uint64_t n,
                    double main_load, double *build_s) {
  if (!ctx || (n && (!hs || !hp || !ho))) return WKG_EINVAL;
  if (ctx->loaded) return ctx->fail(WKG_ESTATE, "a graph is already loaded");
  if (n == 0) return ctx->fail(WKG_EINVAL, "no triples");
  if (n >= 0xffffffffull) return ctx->fail(WKG_ERANGE, "the device builder takes fewer than 2^32 triples");
  const double load = main_load > 0 ? main_load : 0.75;
  if (load > 1.0) return ctx->fail(WKG_EINVAL, "main_load must be in (0, 1]");
  hipStream_t s = ctx->pick(nullptr);
  const auto t0 = std::chrono::steady_clock::now();
  Bufs B;
  hipError_t e = hipSuccess;
#define BCHK(call)                                                                     \
  do {                                                                                 \
    if ((e = (call)) != hipSuccess) {                                                  \
      (void)hipGetLastError();                                                         \
      return ctx->fail(e == hipErrorOutOfMemory ? WKG_ENOMEM : WKG_EDEVICE,            \
                       "device build: %s at %s:%d", hipGetErrorString(e), __FILE__, __LINE__); \
    }                                                                                  \
  } while (0)

  // ---- 1. pairs, validity, id range --------------------------------------------
  uint32_t *ds, *dp, *dov;
  BCHK(B.get(&ds, n));
  BCHK(B.get(&dp, n));
  BCHK(B.get(&dov, n));
  BCHK(hipMemcpyAsync(ds, hs, n * 4, hipMemcpyHostToDevice, s));
  BCHK(hipMemcpyAsync(dp, hp, n * 4, hipMemcpyHostToDevice, s));
  BCHK(hipMemcpyAsync(dov, ho, n * 4, hipMemcpyHostToDevice, s));
  Side so, os;
  BCHK(B.get(&so.key, n));
  BCHK(B.get(&so.pid, n));
  BCHK(B.get(&os.key, n));
  BCHK(B.get(&os.pid, n));
  unsigned long long *bad;
  uint32_t *maxid;
  BCHK(B.get(&bad, 1));
  BCHK(B.get(&maxid, 1));
  BCHK(hipMemsetAsync(bad, 0xff, 8, s));
  BCHK(hipMemsetAsync(maxid, 0, 4, s));
  k_b_pairs<<<grid_for(n), 256, 0, s>>>(ds, dp, dov, n, so.key, so.pid, os.key, os.pid, bad, maxid);
  BCHK(hipGetLastError());
  unsigned long long first_bad = 0;
  uint32_t max_id = 0;
  BCHK(hipMemcpyAsync(&first_bad, bad, 8, hipMemcpyDeviceToHost, s));
  BCHK(hipMemcpyAsync(&max_id, maxid, 4, hipMemcpyDeviceToHost, s));
  BCHK(hipStreamSynchronize(s));
  if (first_bad != ~0ull) {
    uint32_t t[3] = {0, 0, 0};
    (void)hipMemcpy(&t[0], ds + first_bad, 4, hipMemcpyDeviceToHost);
    (void)hipMemcpy(&t[1], dp + first_bad, 4, hipMemcpyDeviceToHost);
    (void)hipMemcpy(&t[2], dov + first_bad, 4, hipMemcpyDeviceToHost);
    return ctx->fail(WKG_EINVAL, "invalid triple %llu: (%u, %u, %u)", first_bad, t[0], t[1], t[2]);
  }
  B.put(ds);
  B.put(dp);
  B.put(dov);
  const uint32_t num_metas = std::max<uint32_t>(max_id, 1) + 1;

  // ---- 2. sort, dedup, counts, layout ---------------------------------------------
  uint32_t *tstart;
  BCHK(B.get(&tstart, num_metas));
  BCHK(hipMemsetAsync(tstart, 0xff, (uint64_t)num_metas * 4, s));
  BCHK(prepare_side(so, n, num_metas, nullptr, B, s));
  BCHK(prepare_side(os, n, num_metas, tstart, B, s));
  IdCounts c;
  side_counts(so, num_metas, c.so_cnt, c.nsubj);
  side_counts(os, num_metas, c.os_cnt, c.nobj);
  // rdf:type pairs on the (o, p, s) side are type member lists, not a
  // predicate's IN side
  std::vector<uint32_t> h_tstart;
  BCHK(down(h_tstart, (const uint32_t *)tstart, num_metas, s));
  c.type_cnt.assign(num_metas, 0);
  if (os.start[kTypeId] != kNone) {
    const uint64_t seg_end = os.start[kTypeId] + c.os_cnt[kTypeId];
    uint32_t prev = kNone;
    for (uint32_t t = 0; t <= num_metas; t++) {
      if (t < num_metas && h_tstart[t] == kNone) continue;
      if (prev != kNone) c.type_cnt[prev] = (t < num_metas ? h_tstart[t] : seg_end) - h_tstart[prev];
      prev = t;
    }
  }
  c.os_cnt[kTypeId] = 0;
  c.nobj[kTypeId] = 0;
  for (uint32_t p = 2; p < num_metas; p++)
    if (c.type_cnt[p] && (c.so_cnt[p] || c.os_cnt[p]))
      return ctx->fail(WKG_EINVAL, "id %u is used both as a predicate and as a type", p);
  Layout lay;
  plan_layout(c, num_metas, load, lay);
  if (lay.total_edges >= (uint64_t(1) << kOffBits)) return ctx->fail(WKG_ERANGE, "edge region exceeds 2^36 entries");
  if (lay.num_keys >= 0xffffffffull || lay.num_buckets >= 0xffffffffull)
    return ctx->fail(WKG_ERANGE, "the device builder takes fewer than 2^32 keys and buckets");

  // per-id device info (chain bases come after the chains are counted)
  std::vector<IdInfo> hid(num_metas, IdInfo{});
  std::vector<uint64_t> key_base(num_metas, 0);
  {
    uint64_t kb = 0;
    for (const Epoch &ep : lay.ep) {
      key_base[ep.id] = kb;
      hid[ep.id].edge_base = ep.edge_base;
      kb += ep.nkeys;
    }
  }
  for (uint32_t p = 0; p < num_metas; p++) {
    IdInfo &d = hid[p];
    d.key_base = key_base[p];
    d.main_start = lay.metas[p].main_hdr_start;
    d.psz = lay.metas[p].partition_sz;
    d.magic = d.psz ? fastmod_magic(d.psz) : 0;
    d.so_start = so.start[p] == kNone ? 0 : so.start[p];
    d.os_start = os.start[p] == kNone ? 0 : os.start[p];
    d.rs_at = so.rat[p];
    d.ro_at = os.rat[p];
    d.na = (uint32_t)c.so_cnt[p];
    d.nb = (uint32_t)c.os_cnt[p];
    d.nsubj = (uint32_t)c.nsubj[p];
  }
  IdInfo *dids;
  BCHK(B.get(&dids, num_metas));
  BCHK(hipMemcpyAsync(dids, hid.data(), num_metas * sizeof(IdInfo), hipMemcpyHostToDevice, s));

  // ---- 3 + 4a. edges and keys ----------------------------------------------------------
  uint32_t *edges = nullptr;
  BCHK(hipMalloc(&edges, std::max<uint64_t>(lay.total_edges, 1) * 4));
  B.v.push_back(edges);
  const uint64_t nk = lay.num_keys;
  uint64_t *kkey, *ksize, *koff;
  uint32_t *kid;
  BCHK(B.get(&kkey, nk));
  BCHK(B.get(&ksize, nk));
  BCHK(B.get(&koff, nk));
  BCHK(B.get(&kid, nk));
  std::vector<HostKey> hk;
  uint64_t ntypes = 0;
  for (uint32_t t = 0; t < num_metas; t++) ntypes += c.type_cnt[t] ? 1 : 0;
  for (const Epoch &ep : lay.ep) {
    const uint32_t p = ep.id;
    if (ep.kind == 1) {
      hk.push_back(HostKey{key_base[p], make_key(0, p, kDirIn), c.type_cnt[p], 0, p, 0});
    } else if (ep.kind == 2) {
      const uint64_t na = c.so_cnt[p], nb = c.os_cnt[p], base = key_base[p] + c.nsubj[p] + c.nobj[p];
      if (na) hk.push_back(HostKey{base, make_key(0, p, kDirIn), c.nsubj[p], na + nb, p, 0});
      if (nb) hk.push_back(HostKey{base + (na ? 1 : 0), make_key(0, p, kDirOut), c.nobj[p], na + nb + c.nsubj[p], p, 0});
    }
  }
  // every key slot is written exactly once: subject runs, objects runs of
  // predicates (the rdf:type pairs' runs are the types), index keys
  if ((uint64_t)so.nruns + (os.nruns - ntypes) + hk.size() != nk)
    return ctx->fail(WKG_EINVAL, "device build: %llu keys planned, %llu made", (unsigned long long)nk,
                     (unsigned long long)(so.nruns + (os.nruns - ntypes) + hk.size()));
  k_b_from_so<<<grid_for(so.n), 256, 0, s>>>(so.key, so.pid, so.run, so.rpre, so.runpos, so.n, dids, edges, kkey,
                                              ksize, koff, kid);
  k_b_from_os<<<grid_for(os.n), 256, 0, s>>>(os.key, os.pid, os.run, os.rpre, os.runpos, os.n, dids, tstart, edges,
                                              kkey, ksize, koff, kid);
  BCHK(hipGetLastError());
  if (!hk.empty()) {
    HostKey *dhk;
    BCHK(B.get(&dhk, hk.size()));
    BCHK(hipMemcpyAsync(dhk, hk.data(), hk.size() * sizeof(HostKey), hipMemcpyHostToDevice, s));
    k_b_host_keys<<<grid_for(hk.size()), 256, 0, s>>>(dhk, hk.size(), kkey, ksize, koff, kid);
    BCHK(hipGetLastError());
    BCHK(hipStreamSynchronize(s));  // (hk is a host vector)
  }
  // the sorted pairs are no longer needed
  for (Side *sd : {&so, &os}) {
    B.put(sd->key);
    B.put(sd->pid);
    B.put(sd->run);
    B.put(sd->rpre);
    B.put(sd->runpos);
  }

  // ---- 4b. buckets, ranks, chains, slots ----------------------------------------------------
  uint32_t *kb, *kidx, *kb2, *kidx2;
  BCHK(B.get(&kb, nk));
  BCHK(B.get(&kidx, nk));
  BCHK(B.get(&kb2, nk));
  BCHK(B.get(&kidx2, nk));
  k_b_bucket<<<grid_for(nk), 256, 0, s>>>(kkey, kid, nk, dids, kb, kidx);
  BCHK(hipGetLastError());
  {
    int bits = 1;
    while ((1ull << bits) <= lay.num_buckets) bits++;
    size_t t = 0;
    BCHK(rocprim::radix_sort_pairs(nullptr, t, kb, kb2, kidx, kidx2, nk, 0, bits, s));
    char *tmp;
    BCHK(B.get(&tmp, t));
    BCHK(rocprim::radix_sort_pairs(tmp, t, kb, kb2, kidx, kidx2, nk, 0, bits, s));
    B.put(tmp);
  }
  B.put(kb);
  B.put(kidx);
  const uint64_t nbm = lay.num_buckets;
  uint32_t *bcnt, *bfirst;
  uint64_t *nchain, *cpre;
  BCHK(B.get(&bcnt, nbm));
  BCHK(B.get(&bfirst, nbm));
  BCHK(B.get(&nchain, nbm + 1));
  BCHK(B.get(&cpre, nbm + 1));
  BCHK(hipMemsetAsync(bcnt, 0, nbm * 4, s));
  k_b_bucket_runs<<<grid_for(nk), 256, 0, s>>>(kb2, nk, bcnt, bfirst);
  k_b_chains<<<grid_for(nbm), 256, 0, s>>>(bcnt, nbm, nchain);
  BCHK(hipMemsetAsync(nchain + nbm, 0, 8, s));
  BCHK(excl_scan_u64(nchain, cpre, nbm + 1, B, s));
  BCHK(hipGetLastError());
  // each epoch's indirect buckets: the chain prefix over its main range
  std::vector<uint64_t> h_cpre;
  BCHK(down(h_cpre, (const uint64_t *)cpre, nbm + 1, s));
  for (Epoch &ep : lay.ep) {
    const wkg_pred_meta &m = lay.metas[ep.id];
    ep.ext_count = h_cpre[m.main_hdr_end] - h_cpre[m.main_hdr_start];
  }
  const uint64_t ext_total = place_epochs(lay);
  for (const Epoch &ep : lay.ep)
    hid[ep.id].chain_base = lay.num_buckets + ep.ext_base - h_cpre[lay.metas[ep.id].main_hdr_start];
  BCHK(hipMemcpyAsync(dids, hid.data(), num_metas * sizeof(IdInfo), hipMemcpyHostToDevice, s));
  const uint64_t nslots = (nbm + ext_total) * kAssoc;
  if (nslots / kAssoc >= 0xffffffffull) return ctx->fail(WKG_ERANGE, "images are limited to 2^32 - 1 buckets");
  uint64_t *slots = nullptr;
  BCHK(hipMalloc(&slots, nslots * 16));
  B.v.push_back(slots);
  BCHK(hipMemsetAsync(slots, 0, nslots * 16, s));
  uint32_t *too_big;
  BCHK(B.get(&too_big, 1));
  BCHK(hipMemsetAsync(too_big, 0, 4, s));
  k_b_place<<<grid_for(nk), 256, 0, s>>>(kb2, kidx2, nk, bfirst, bcnt, cpre, kkey, ksize, koff, kid, dids, slots,
                                          too_big);
  BCHK(hipGetLastError());
  uint32_t big = 0;
  BCHK(hipMemcpyAsync(&big, too_big, 4, hipMemcpyDeviceToHost, s));
  BCHK(hipStreamSynchronize(s));
  if (big) return ctx->fail(WKG_ERANGE, "an edge list exceeds 2^28 entries");
#undef BCHK

  // ---- install: the image stays where it was built ---------------------------------------
  std::vector<wkg_pred_meta_abi> am(num_metas);
  static_assert(sizeof(wkg_pred_meta_abi) == sizeof(wkg_pred_meta), "pred meta layouts differ");
  memcpy(am.data(), lay.metas.data(), num_metas * sizeof(wkg_pred_meta));
  B.v.erase(std::find(B.v.begin(), B.v.end(), (void *)slots));
  B.v.erase(std::find(B.v.begin(), B.v.end(), (void *)edges));
  for (void *p : B.v) (void)hipFree(p);  // (before the table budget is taken from what is free)
  B.v.clear();
  ctx->d_slots = (uint4 *)slots;
  ctx->d_edges = edges;
  if (int rc = ctx->set_image_metas(nslots, lay.total_edges, am.data(), (int)num_metas)) {
    ctx->free_graph();
    return rc;
  }
  const int rc = ctx->finish_resident(0, s);
  if (build_s) *build_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return rc;
}

int wkg_graph_shape(const wkg_ctx *ctx, uint64_t *num_slots, uint64_t *num_edges, int *num_metas) {
  if (!ctx) return WKG_EINVAL;
  if (!ctx->loaded) return WKG_ESTATE;
  if (num_slots) *num_slots = ctx->num_slots;
  if (num_edges) *num_edges = ctx->num_edges;
  if (num_metas) *num_metas = (int)ctx->num_metas;
  return WKG_OK;
}

int wkg_graph_download(wkg_ctx *ctx, uint64_t *slots, uint32_t *edges, wkg_pred_meta_abi *metas) {
  if (!ctx) return WKG_EINVAL;
  if (!ctx->loaded || ctx->framed) return ctx->fail(WKG_ESTATE, "no resident graph");
  hipError_t e = hipSuccess;
  if (slots) e = hipMemcpy(slots, ctx->d_slots, ctx->num_slots * 16, hipMemcpyDeviceToHost);
  if (e == hipSuccess && edges && ctx->num_edges) e = hipMemcpy(edges, ctx->d_edges, ctx->num_edges * 4, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return ctx->fail(WKG_EDEVICE, "image download failed: %s", hipGetErrorString(e));
  if (metas) memcpy(metas, ctx->metas.data(), ctx->metas.size() * sizeof(wkg_pred_meta_abi));
  return WKG_OK;
}

}  // extern "C"
