// Gstore residency: whole-image upload (the default on 288 GB of HBM3E) or a
// framed per-predicate segment cache when the image exceeds its HBM budget
// (ShardManager, core/shard_manager.hpp:44-557: frame allocation 473-555,
// eviction 130-293, header tables 361-374, segment copies 378-442).
//
// A predicate's segment is its main buckets followed by its indirect
// buckets (core/shard_manager.hpp:378-427) plus its edge range (:429-442).
// Frames are powers of two so the device maps a segment index with a shift
// and a mask.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstring>

#include "../../include/wukong_gpu.h"
#include "device.h"
#include "wkg_internal.h"

using namespace wkg;

namespace {

// Device image checks after a whole-image upload: each stored list lies in
// the image and in its predicate's segment, each chain link points into the
// indirect region and no bucket is linked twice (so no chain can cycle), and
// optionally each list ascends strictly.
// flags: bit 0 list bounds, bit 1 links, bit 2 unsorted list.
__global__ void k_validate(const uint4 *slots, uint64_t num_buckets_total, uint64_t num_main, uint64_t num_edges,
                           const DevMeta *meta, uint32_t num_metas, const uint32_t *edges, uint32_t *linked,
                           int check_sorted, uint32_t *flags) {
  for (uint64_t s = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; s < num_buckets_total * kAssoc;
       s += (uint64_t)gridDim.x * blockDim.x) {
    const uint4 v = slots[s];
    const uint64_t k = ((uint64_t)v.y << 32) | v.x;
    if (k == 0) continue;
    if ((s & 7) == 7) {
      const uint64_t nb = key_vid(k);
      if (nb < num_main || nb >= num_buckets_total) {
        atomicOr(flags, 2u);
      } else if (atomicOr(&linked[nb >> 5], 1u << (nb & 31)) & (1u << (nb & 31))) {
        atomicOr(flags, 2u);
      }
      continue;
    }
    const uint64_t p = ((uint64_t)v.w << 32) | v.z;
    const uint64_t sz = ptr_size(p), off = ptr_off(p);
    const uint64_t pid = key_pid(k);
    if (pid >= num_metas || off + sz > num_edges || off < meta[pid].edge_start ||
        off + sz > meta[pid].edge_start + meta[pid].edge_len) {
      atomicOr(flags, 1u);
      continue;
    }
    if (check_sorted) {
      for (uint64_t j = 1; j < sz; j++)
        if (edges[off + j - 1] >= edges[off + j]) {
          atomicOr(flags, 4u);
          break;
        }
    }
  }
}

int log2_floor(uint64_t x) {
  int r = 0;
  while ((x >> (r + 1)) != 0) r++;
  return r;
}

}  // namespace

void wkg_ctx::free_graph() {
  sq_stop();  // the server reads the graph it was launched with
  {
    std::lock_guard<std::mutex> lk(sq_big_mu);
    sq_big_starts.clear();
  }
  if (d_slots) (void)hipFree(d_slots);
  if (d_edges) (void)hipFree(d_edges);
  if (d_meta) (void)hipFree(d_meta);
  if (d_vframes) (void)hipFree(d_vframes);
  if (d_eframes) (void)hipFree(d_eframes);
  if (res) {
    if (res->copy_stream) {
      (void)hipStreamSynchronize(res->copy_stream);
      (void)hipStreamDestroy(res->copy_stream);
    }
    for (hipEvent_t e : res->ready)
      if (e) (void)hipEventDestroy(e);
    if (res->pinned) (void)hipHostFree(res->pinned);
    delete res;
  }
  d_slots = nullptr;
  d_edges = nullptr;
  d_meta = nullptr;
  d_vframes = d_eframes = nullptr;
  res = nullptr;
  loaded = framed = sorted = false;
  metas.clear();
  h_meta.clear();
  resident_bytes = 0;
}

// Makes `pid` resident for work queued on stream s: waits (on s) for its
// prefetch if one is under way, else loads its missing frames on s.
int wkg_ctx::ensure_pred(int pid, hipStream_t s, const std::vector<int> *keep) {
  if (!framed) return WKG_OK;
  if (pid < 0 || (uint32_t)pid >= num_metas) return fail(WKG_EINVAL, "predicate id %d out of range", pid);
  Residency &r = *res;
  if (r.pending[pid]) {
    // its copies were queued on some stream (a prefetch, or a load for
    // another stream): work on s waits for them until they are known done
    if (hipEventQuery(r.ready[pid]) == hipSuccess) {
      r.pending[pid] = 0;
    } else if (hipStreamWaitEvent(s, r.ready[pid], 0) != hipSuccess) {
      return fail(WKG_EDEVICE, "waiting for predicate %d's segment copies failed", pid);
    }
  }
  const int rc = load_pred(pid, s, keep, false);
  if (rc != WKG_OK) (void)hipStreamSynchronize(s);  // copies queued before the failure are not tracked
  return rc;
}

// Loads missing frames of `pid` with copies queued on stream s, evicting
// least recently used predicates that are neither `pid` nor listed in `keep`.
// A prefetch (on the copy stream) records pid's ready event after its copies
// and leaves the predicate pending; it does not count as a use.
int wkg_ctx::load_pred(int pid, hipStream_t s, const std::vector<int> *keep, bool prefetch) {
  Residency &r = *res;
  auto touch = [&](int p) {
    if (r.in_lru[p]) r.lru.erase(r.lru_pos[p]);
    r.lru.push_front(p);
    r.lru_pos[p] = r.lru.begin();
    r.in_lru[p] = 1;
  };
  if (r.used_v[pid] == r.vlen[pid] && r.used_e[pid] == r.elen[pid]) {
    if (!prefetch) touch(pid);
    return WKG_OK;
  }
  if (r.vlen[pid] > r.n_vframes || r.elen[pid] > r.n_eframes)
    return fail(WKG_ENOMEM, "predicate %d's segment is larger than the graph budget", pid);
  auto kept = [&](int p) {
    if (p == pid) return true;
    if (keep)
      for (int q : *keep)
        if (q == p) return true;
    return false;
  };
  bool synced = false;
  auto evict_until = [&](bool vert, uint64_t need) -> bool {
    std::vector<uint64_t> &fr = vert ? r.vfree : r.efree;
    while (fr.size() < need) {
      int victim = -1;
      for (auto it = r.lru.rbegin(); it != r.lru.rend(); ++it) {
        const int p = *it;
        if (kept(p)) continue;
        if ((vert ? r.used_v[p] : r.used_e[p]) == 0) continue;
        victim = p;
        break;
      }
      if (victim < 0) return false;
      if (!synced) {  // frames may still be read (or prefetched into) by queued work
        (void)hipDeviceSynchronize();
        synced = true;
        std::fill(r.pending.begin(), r.pending.end(), 0);  // every prefetch has landed
      }
      std::vector<uint64_t> &tab = vert ? r.h_vframes : r.h_eframes;
      const uint64_t base = vert ? h_meta[victim].vframe_base : h_meta[victim].eframe_base;
      const uint64_t len = vert ? r.vlen[victim] : r.elen[victim];
      const uint64_t shift = vert ? vshift : eshift;
      for (uint64_t j = 0; j < len; j++)
        if (tab[base + j] != ~0ull) {
          fr.push_back(tab[base + j] >> shift);
          tab[base + j] = ~0ull;
        }
      (vert ? r.used_v : r.used_e)[victim] = 0;
      uint64_t *dtab = vert ? d_vframes : d_eframes;
      // synchronous: the host table changes again on the next call
      if (len && hipMemcpy(dtab + base, tab.data() + base, len * 8, hipMemcpyHostToDevice) != hipSuccess)
        return false;
      stat_evictions++;
      if (r.used_v[victim] == 0 && r.used_e[victim] == 0 && r.in_lru[victim]) {
        r.lru.erase(r.lru_pos[victim]);
        r.in_lru[victim] = 0;
      }
    }
    return true;
  };
  const DevMeta &m = h_meta[pid];
  const wkg_pred_meta_abi &pm = metas[pid];
  // bucket frames
  if (r.used_v[pid] < r.vlen[pid]) {
    if (!evict_until(true, r.vlen[pid] - r.used_v[pid]))
      return fail(WKG_ENOMEM, "cannot free bucket frames for predicate %d", pid);
    const uint64_t F = r.vframe_buckets, L = m.psz + (pm.indrct_hdr_end - pm.indrct_hdr_start);
    for (uint64_t j = 0; j < r.vlen[pid]; j++) {
      uint64_t &ent = r.h_vframes[m.vframe_base + j];
      if (ent != ~0ull) continue;
      const uint64_t fr = r.vfree.back();
      r.vfree.pop_back();
      const uint64_t lo = j * F, hi = std::min(L, lo + F);
      // main part [lo, min(hi, psz)), indirect part [max(lo, psz), hi)
      if (lo < m.psz) {
        const uint64_t e = std::min(hi, m.psz);
        if (hipMemcpyAsync(d_slots + (fr * F) * kAssoc, r.h_slots + (m.main_start + lo) * kAssoc * 2,
                           (e - lo) * kAssoc * 16, hipMemcpyHostToDevice, s) != hipSuccess)
          return fail(WKG_EDEVICE, "segment copy failed");
      }
      if (hi > m.psz) {
        const uint64_t b = std::max(lo, m.psz);
        if (hipMemcpyAsync(d_slots + (fr * F + (b - lo)) * kAssoc,
                           r.h_slots + (pm.indrct_hdr_start + (b - m.psz)) * kAssoc * 2, (hi - b) * kAssoc * 16,
                           hipMemcpyHostToDevice, s) != hipSuccess)
          return fail(WKG_EDEVICE, "segment copy failed");
      }
      ent = fr * F;
      r.used_v[pid]++;
      stat_loads++;
    }
    if (r.vlen[pid] &&
        hipMemcpy(d_vframes + m.vframe_base, r.h_vframes.data() + m.vframe_base, r.vlen[pid] * 8,
                  hipMemcpyHostToDevice) != hipSuccess)
      return fail(WKG_EDEVICE, "frame table update failed");
  }
  // edge frames
  if (r.used_e[pid] < r.elen[pid]) {
    if (!evict_until(false, r.elen[pid] - r.used_e[pid]))
      return fail(WKG_ENOMEM, "cannot free edge frames for predicate %d", pid);
    const uint64_t F = r.eframe_edges, L = m.edge_len;
    for (uint64_t j = 0; j < r.elen[pid]; j++) {
      uint64_t &ent = r.h_eframes[m.eframe_base + j];
      if (ent != ~0ull) continue;
      const uint64_t fr = r.efree.back();
      r.efree.pop_back();
      const uint64_t lo = j * F, hi = std::min(L, lo + F);
      if (hipMemcpyAsync(d_edges + fr * F, r.h_edges + m.edge_start + lo, (hi - lo) * 4, hipMemcpyHostToDevice,
                         s) != hipSuccess)
        return fail(WKG_EDEVICE, "segment copy failed");
      ent = fr * F;
      r.used_e[pid]++;
      stat_loads++;
    }
    if (r.elen[pid] &&
        hipMemcpy(d_eframes + m.eframe_base, r.h_eframes.data() + m.eframe_base, r.elen[pid] * 8,
                  hipMemcpyHostToDevice) != hipSuccess)
      return fail(WKG_EDEVICE, "frame table update failed");
  }
  // The copies read pinned memory, so they may still run when this returns:
  // the ready event lets work on other streams wait for them (ensure_pred).
  if (!r.ready[pid] && hipEventCreateWithFlags(&r.ready[pid], hipEventDisableTiming) != hipSuccess) {
    r.ready[pid] = nullptr;
    (void)hipStreamSynchronize(s);
    return fail(WKG_EDEVICE, "event creation failed");
  }
  if (hipEventRecord(r.ready[pid], s) != hipSuccess) {
    (void)hipStreamSynchronize(s);
    return fail(WKG_EDEVICE, "event record failed");
  }
  r.pending[pid] = 1;
  if (prefetch) r.stat_prefetches++;
  touch(pid);  // (a prefetch too: newest in the LRU order, so the next prefetch does not evict it first)
  return WKG_OK;
}

// Host-side checks for the framed mode (the image stays on the host there):
// per predicate, every link of its main and indirect buckets stays inside its
// own indirect range and no bucket is linked twice; every list lies inside
// its predicate's edge segment.
static int validate_host(wkg_ctx *c, const uint64_t *slots, const uint32_t *edges, bool check_sorted,
                         bool *is_sorted) {
  const uint64_t nb = c->num_buckets_total;
  std::vector<uint8_t> linked(nb, 0);
  bool srt = true;
  for (uint32_t p = 0; p < c->num_metas; p++) {
    const wkg_pred_meta_abi &pm = c->metas[p];
    if (!pm.partition_sz) continue;
    auto scan = [&](uint64_t b) -> int {
      const uint64_t *sp = slots + b * kAssoc * 2;
      for (int i = 0; i < kAssoc - 1; i++) {
        const uint64_t k = sp[2 * i];
        if (!k) continue;
        const uint64_t kp = key_pid(k), sz = ptr_size(sp[2 * i + 1]), off = ptr_off(sp[2 * i + 1]);
        if (kp >= c->num_metas || off < c->metas[kp].edge_start || off + sz > c->metas[kp].edge_end)
          return c->fail(WKG_EINVAL, "corrupt image: list of slot %llu outside its segment",
                         (unsigned long long)(b * kAssoc + i));
        if (check_sorted && srt)
          for (uint64_t j = 1; j < sz; j++)
            if (edges[off + j - 1] >= edges[off + j]) {
              srt = false;
              break;
            }
      }
      const uint64_t link = sp[2 * (kAssoc - 1)];
      if (link) {
        const uint64_t t = key_vid(link);
        if (t < pm.indrct_hdr_start || t >= pm.indrct_hdr_end || linked[t])
          return c->fail(WKG_EINVAL, "corrupt image: bad chain link in bucket %llu", (unsigned long long)b);
        linked[t] = 1;
      }
      return WKG_OK;
    };
    for (uint64_t b = pm.main_hdr_start; b < pm.main_hdr_end; b++)
      if (int rc = scan(b)) return rc;
    for (uint64_t b = pm.indrct_hdr_start; b < pm.indrct_hdr_end; b++)
      if (int rc = scan(b)) return rc;
  }
  *is_sorted = srt;
  return WKG_OK;
}

extern "C" {

// Checks an image's sizes and pred metas and derives the device metas
// (first half of a graph load, shared with the device builder).
int wkg_ctx::set_image_metas(uint64_t nslots, uint64_t nedges, const wkg_pred_meta_abi *m_in, int nmetas) {
  if (nslots == 0 || nslots % kAssoc)
    return fail(WKG_EINVAL, "num_slots must be a positive multiple of %d", kAssoc);
  if (nmetas > (1 << kIdxBits)) return fail(WKG_EINVAL, "too many predicate ids");
  if (nslots / kAssoc > 0xffffffffull)  // the probe addresses buckets with 32 bits
    return fail(WKG_ERANGE, "images are limited to 2^32 - 1 buckets");
  num_slots = nslots;
  num_edges = nedges;
  num_buckets_total = nslots / kAssoc;
  num_metas = (uint32_t)nmetas;
  metas.assign(m_in, m_in + nmetas);
  uint64_t nmain = 0;
  for (int p = 0; p < nmetas; p++) {
    const wkg_pred_meta_abi &m = m_in[p];
    if (!m.partition_sz) continue;
    if (m.main_hdr_end - m.main_hdr_start != m.partition_sz || m.main_hdr_end > num_buckets_total ||
        m.indrct_hdr_start > m.indrct_hdr_end || m.indrct_hdr_end > num_buckets_total ||
        m.edge_start > m.edge_end || m.edge_end > nedges || m.edge_end - m.edge_start >= (1ull << 32)) {
      metas.clear();
      return fail(WKG_EINVAL, "pred_meta %d is inconsistent with the image", p);
    }
    nmain = std::max(nmain, m.main_hdr_end);
  }
  num_main = nmain;
  h_meta.assign(nmetas, DevMeta{});
  for (int p = 0; p < nmetas; p++) {
    const wkg_pred_meta_abi &m = m_in[p];
    DevMeta &d = h_meta[p];
    d.main_start = m.main_hdr_start;
    d.psz = m.partition_sz;
    d.magic = m.partition_sz ? fastmod_magic(m.partition_sz) : 0;
    d.indrct_start = m.indrct_hdr_start;
    d.edge_start = m.edge_start;
    d.edge_len = m.edge_end - m.edge_start;
  }
  return WKG_OK;
}

// Second half of a resident load, with d_slots and d_edges in place: device
// metas, the image checks on the device, and the table budget.
int wkg_ctx::finish_resident(uint32_t flags, hipStream_t s) {
  if (hipMalloc(&d_meta, sizeof(DevMeta) * num_metas) != hipSuccess) {
    (void)hipGetLastError();
    free_graph();
    return fail(WKG_ENOMEM, "cannot allocate the predicate metadata");
  }
  hipError_t e = hipMemcpy(d_meta, h_meta.data(), sizeof(DevMeta) * num_metas, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    free_graph();
    return fail(WKG_EDEVICE, "metadata upload failed: %s", hipGetErrorString(e));
  }
  const uint64_t nwords = (num_buckets_total + 31) / 32;
  uint32_t *tmp = nullptr;
  if (hipMalloc(&tmp, (nwords + 1) * 4) != hipSuccess) {
    free_graph();
    return fail(WKG_ENOMEM, "cannot allocate validation bitmap");
  }
  (void)hipMemsetAsync(tmp, 0, (nwords + 1) * 4, s);
  const int check_sorted = (flags & WKG_LOAD_SORTED_LISTS) ? 0 : 1;
  const uint64_t blocks = std::min<uint64_t>((num_slots + 255) / 256, 1u << 16);
  k_validate<<<(unsigned)blocks, 256, 0, s>>>(d_slots, num_buckets_total, num_main, num_edges, d_meta, num_metas,
                                               d_edges, tmp + 1, check_sorted, tmp);
  uint32_t vflags = 0;
  e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(&vflags, tmp, 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(tmp);
  if (e != hipSuccess) {
    free_graph();
    return fail(WKG_EDEVICE, "image validation failed: %s", hipGetErrorString(e));
  }
  if (vflags & 3u) {
    free_graph();
    return fail(WKG_EINVAL, "corrupt image (%s)", (vflags & 1u) ? "list outside its segment" : "bad chain link");
  }
  sorted = (flags & WKG_LOAD_SORTED_LISTS) || !(vflags & 4u);
  resident_bytes = num_slots * 16 + std::max<uint64_t>(num_edges, 1) * 4;
  loaded = true;
  if (!budget_request) {  // default table budget: 75% of what is left
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) == hipSuccess) pool->budget = (uint64_t)(fr * 0.75);
  }
  return WKG_OK;
}

int wkg_load_graph(wkg_ctx *ctx, const uint64_t *slots, uint64_t num_slots, const uint32_t *edges,
                   uint64_t num_edges, const wkg_pred_meta_abi *metas, int num_metas, uint64_t graph_budget_bytes,
                   uint32_t flags) {
  if (!ctx) return WKG_EINVAL;
  if (ctx->loaded) return ctx->fail(WKG_ESTATE, "a graph is already loaded");
  if (!slots || !metas || num_metas < 2 || (num_edges && !edges))
    return ctx->fail(WKG_EINVAL, "null image array");
  (void)hipSetDevice(ctx->device);
  if (int rc = ctx->set_image_metas(num_slots, num_edges, metas, num_metas)) return rc;
  const uint64_t num_main = ctx->num_main;
  const uint64_t slot_bytes = num_slots * 16, edge_bytes = std::max<uint64_t>(num_edges, 1) * 4;
  const bool framed = graph_budget_bytes && graph_budget_bytes < slot_bytes + edge_bytes;
  hipStream_t s = ctx->stream;

  if (!framed) {
    if (hipMalloc(&ctx->d_slots, slot_bytes) != hipSuccess || hipMalloc(&ctx->d_edges, edge_bytes) != hipSuccess) {
      (void)hipGetLastError();
      ctx->free_graph();
      return ctx->fail(WKG_ENOMEM, "cannot allocate %llu bytes for the resident image",
                       (unsigned long long)(slot_bytes + edge_bytes));
    }
    hipError_t e = hipMemcpy(ctx->d_slots, slots, slot_bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess && num_edges) e = hipMemcpy(ctx->d_edges, edges, num_edges * 4, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      ctx->free_graph();
      return ctx->fail(WKG_EDEVICE, "image upload failed: %s", hipGetErrorString(e));
    }
    return ctx->finish_resident(flags, s);
  } else {
    bool srt = true;
    int rc = validate_host(ctx, slots, edges, !(flags & WKG_LOAD_SORTED_LISTS), &srt);
    if (rc) {
      ctx->free_graph();
      return rc;
    }
    ctx->sorted = srt;
    Residency *r = new Residency;
    ctx->res = r;
    r->h_slots = slots;
    r->h_edges = edges;
    // Segment copies read the image from pinned memory, so that a prefetch
    // on the copy stream runs while the host goes on (the reference pins the
    // host image in place with cudaHostRegister, core/rcache.hpp:187-189;
    // here the image is copied once into memory allocated pinned). Without
    // room for the copy, copies read the caller's arrays.
    if (hipHostMalloc((void **)&r->pinned, slot_bytes + edge_bytes, hipHostMallocDefault) == hipSuccess) {
      memcpy(r->pinned, slots, slot_bytes);
      if (num_edges) memcpy(r->pinned + slot_bytes, edges, num_edges * 4);
      r->h_slots = (const uint64_t *)r->pinned;
      r->h_edges = (const uint32_t *)(r->pinned + slot_bytes);
    } else {
      (void)hipGetLastError();
      r->pinned = nullptr;
    }
    if (hipStreamCreateWithFlags(&r->copy_stream, hipStreamNonBlocking) != hipSuccess) {
      r->copy_stream = nullptr;
      ctx->free_graph();
      return ctx->fail(WKG_EDEVICE, "copy stream creation failed");
    }
    r->ready.assign(num_metas, nullptr);
    r->pending.assign(num_metas, 0);
    // split the budget between bucket and edge frames in proportion to the
    // segments, frames of at most 16 MiB and at least 64 per pool
    uint64_t seg_b = 0, seg_e = 0, max_b = 1, max_e = 1;
    for (int p = 0; p < num_metas; p++) {
      const uint64_t lb = ctx->h_meta[p].psz + (metas[p].indrct_hdr_end - metas[p].indrct_hdr_start);
      seg_b += lb;
      seg_e += ctx->h_meta[p].edge_len;
      max_b = std::max(max_b, lb);
      max_e = std::max(max_e, ctx->h_meta[p].edge_len);
    }
    const double fb = (double)(seg_b * 128) / (double)(seg_b * 128 + seg_e * 4 + 1);
    const uint64_t bud_b = (uint64_t)(graph_budget_bytes * fb), bud_e = graph_budget_bytes - bud_b;
    const int vs = std::max(9, std::min(17, log2_floor(std::max<uint64_t>(bud_b / 128 / 64, 1))));
    const int es = std::max(14, std::min(22, log2_floor(std::max<uint64_t>(bud_e / 4 / 64, 1))));
    ctx->vshift = (uint32_t)vs;
    ctx->eshift = (uint32_t)es;
    r->vframe_buckets = 1ull << vs;
    r->eframe_edges = 1ull << es;
    r->n_vframes = bud_b / (r->vframe_buckets * 128);
    r->n_eframes = bud_e / (r->eframe_edges * 4);
    r->vlen.resize(num_metas);
    r->elen.resize(num_metas);
    r->used_v.assign(num_metas, 0);
    r->used_e.assign(num_metas, 0);
    r->lru_pos.resize(num_metas);
    r->in_lru.assign(num_metas, 0);
    uint64_t vb = 0, eb = 0;
    for (int p = 0; p < num_metas; p++) {
      DevMeta &d = ctx->h_meta[p];
      const uint64_t lb = d.psz + (metas[p].indrct_hdr_end - metas[p].indrct_hdr_start);
      r->vlen[p] = (lb + r->vframe_buckets - 1) / r->vframe_buckets;
      r->elen[p] = (d.edge_len + r->eframe_edges - 1) / r->eframe_edges;
      d.vframe_base = vb;
      d.eframe_base = eb;
      vb += r->vlen[p];
      eb += r->elen[p];
    }
    r->h_vframes.assign(std::max<uint64_t>(vb, 1), ~0ull);
    r->h_eframes.assign(std::max<uint64_t>(eb, 1), ~0ull);
    for (uint64_t f = 0; f < r->n_vframes; f++) r->vfree.push_back(r->n_vframes - 1 - f);
    for (uint64_t f = 0; f < r->n_eframes; f++) r->efree.push_back(r->n_eframes - 1 - f);
    const uint64_t pool_b = std::max<uint64_t>(r->n_vframes, 1) * r->vframe_buckets * 128;
    const uint64_t pool_e = std::max<uint64_t>(r->n_eframes, 1) * r->eframe_edges * 4;
    if (hipMalloc(&ctx->d_slots, pool_b) != hipSuccess || hipMalloc(&ctx->d_edges, pool_e) != hipSuccess ||
        hipMalloc(&ctx->d_meta, sizeof(DevMeta) * num_metas) != hipSuccess ||
        hipMalloc(&ctx->d_vframes, r->h_vframes.size() * 8) != hipSuccess ||
        hipMalloc(&ctx->d_eframes, r->h_eframes.size() * 8) != hipSuccess) {
      (void)hipGetLastError();
      ctx->free_graph();
      return ctx->fail(WKG_ENOMEM, "cannot allocate the frame pools");
    }
    hipError_t e = hipMemcpy(ctx->d_meta, ctx->h_meta.data(), sizeof(DevMeta) * num_metas, hipMemcpyHostToDevice);
    if (e == hipSuccess)
      e = hipMemcpy(ctx->d_vframes, r->h_vframes.data(), r->h_vframes.size() * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess)
      e = hipMemcpy(ctx->d_eframes, r->h_eframes.data(), r->h_eframes.size() * 8, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      ctx->free_graph();
      return ctx->fail(WKG_EDEVICE, "frame table upload failed: %s", hipGetErrorString(e));
    }
    ctx->framed = true;
    ctx->resident_bytes = pool_b + pool_e;
  }
  ctx->loaded = true;
  if (!ctx->budget_request) {  // default table budget: 75% of what is left
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) == hipSuccess) ctx->pool->budget = (uint64_t)(fr * 0.75);
  }
  return WKG_OK;
}

int wkg_ensure_predicate(wkg_ctx *ctx, int pid) {
  if (!ctx) return WKG_EINVAL;
  if (!ctx->loaded) return ctx->fail(WKG_ESTATE, "no graph loaded");
  return ctx->ensure_pred(pid, ctx->stream);
}

int wkg_prefetch_predicate(wkg_ctx *ctx, int pid, int in_use) {
  if (!ctx) return WKG_EINVAL;
  if (!ctx->loaded) return ctx->fail(WKG_ESTATE, "no graph loaded");
  if (!ctx->framed) return WKG_OK;
  if (pid < 0 || (uint32_t)pid >= ctx->num_metas) return ctx->fail(WKG_EINVAL, "predicate id %d out of range", pid);
  Residency &r = *ctx->res;
  if (r.pending[pid]) return WKG_OK;
  // neither the predicate in use nor the most recently used one is a victim
  std::vector<int> keep;
  if (in_use >= 0) keep.push_back(in_use);
  if (!r.lru.empty()) keep.push_back(r.lru.front());
  const int rc = ctx->load_pred(pid, r.copy_stream, &keep, true);
  if (rc != WKG_OK && r.ready[pid]) {
    // a part of the segment may have been queued before the failure: the
    // next use of pid waits for it too
    if (hipEventRecord(r.ready[pid], r.copy_stream) == hipSuccess) r.pending[pid] = 1;
  } else if (rc != WKG_OK) {
    (void)hipStreamSynchronize(r.copy_stream);
  }
  return rc;
}

int wkg_predicate_resident(const wkg_ctx *ctx, int pid) {
  if (!ctx) return WKG_EINVAL;
  if (!ctx->loaded) return WKG_ESTATE;
  if (pid < 0 || (uint32_t)pid >= ctx->num_metas) return WKG_EINVAL;
  if (!ctx->framed) return 1;
  const Residency &r = *ctx->res;
  return r.used_v[pid] == r.vlen[pid] && r.used_e[pid] == r.elen[pid] ? 1 : 0;
}

int wkg_graph_prefetches(const wkg_ctx *ctx, uint64_t *n) {
  if (!ctx || !n) return WKG_EINVAL;
  *n = ctx->res ? ctx->res->stat_prefetches : 0;
  return WKG_OK;
}

int wkg_preload_predicates(wkg_ctx *ctx, const int *pids, int n) {
  if (!ctx) return WKG_EINVAL;
  if (!ctx->loaded) return ctx->fail(WKG_ESTATE, "no graph loaded");
  if (n > 0 && !pids) return ctx->fail(WKG_EINVAL, "null predicate list");
  if (!ctx->framed) return WKG_OK;
  std::vector<int> list;
  if (n < 0)
    for (uint32_t p = 1; p < ctx->num_metas; p++) list.push_back((int)p);
  else
    list.assign(pids, pids + n);
  for (int p : list) {
    if (p < 0 || (uint32_t)p >= ctx->num_metas) return ctx->fail(WKG_EINVAL, "predicate id %d out of range", p);
    if (ctx->h_meta[p].psz == 0 && ctx->h_meta[p].edge_len == 0) continue;  // nothing stored
    if (int rc = ctx->ensure_pred(p, ctx->stream)) return rc;
  }
  return hipStreamSynchronize(ctx->stream) == hipSuccess ? WKG_OK : ctx->fail(WKG_EDEVICE, "preload failed");
}

int wkg_graph_framed(const wkg_ctx *ctx) {
  if (!ctx) return WKG_EINVAL;
  if (!ctx->loaded) return WKG_ESTATE;
  return ctx->framed ? 1 : 0;
}

int wkg_graph_stats(const wkg_ctx *ctx, uint64_t *resident_bytes, uint64_t *segment_loads,
                    uint64_t *segment_evictions) {
  if (!ctx) return WKG_EINVAL;
  if (resident_bytes) *resident_bytes = ctx->resident_bytes;
  if (segment_loads) *segment_loads = ctx->stat_loads;
  if (segment_evictions) *segment_evictions = ctx->stat_evictions;
  return WKG_OK;
}

}  // extern "C"
