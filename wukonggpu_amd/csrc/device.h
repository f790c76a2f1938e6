// Device-side building blocks for the pattern-step kernels (gfx950, wave64).
//
// Data layout in HBM
//   slots   16-byte {key, ptr} slots, 8 per 128-byte bucket (the gstore image
//           as built on the host); read as one uint4 per lane.
//   edges   u32 neighbour ids.
//   meta    DevMeta per predicate/type id (probe constants precomputed).
//   tables  row-major int32, rows x cols.
//
// Resident mode addresses buckets and edges by their global ids. Framed mode
// (graph larger than its HBM budget) keeps each predicate's segment -- main
// buckets followed by its indirect buckets, and its edge range -- in
// fixed-size frames and maps a segment-local index through a frame table,
// the way the reference's map_location_on_shards does (core/gpu_hash.cu:88-92)
// but with power-of-two frames (shift/mask instead of div/mod).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gstore_format.h"

namespace wkg {

constexpr int kWave = 64;
constexpr int kBlock = 256;         // threads per workgroup for every step kernel
constexpr int kTileRows = kBlock;   // probe/scan tile: one row per thread
constexpr int kExpandVT = 8;        // merged items per thread in the expand
constexpr int kExpandNV = kBlock * kExpandVT;
constexpr uint32_t kMaxHops = 1u << 20;  // chain-walk guard (images are validated)

struct DevMeta {
  uint64_t main_start;    // first main bucket, global id
  uint64_t psz;           // partition_sz (0: the id has no keys)
  uint64_t magic;         // fastmod_magic(psz)
  uint64_t indrct_start;  // first indirect bucket, global id
  uint64_t edge_start;    // first edge, global index
  uint64_t edge_len;
  uint64_t vframe_base;   // framed: first frame-table entry of the bucket segment
  uint64_t eframe_base;   // framed: first frame-table entry of the edge segment
};

struct GraphView {
  const uint4 *slots;       // resident: whole image; framed: bucket frame pool
  const uint32_t *edges;    // resident: whole edge array; framed: edge frame pool
  const DevMeta *meta;
  const uint64_t *vframes;  // framed: physical first bucket of each logical frame
  const uint64_t *eframes;  // framed: physical first edge of each logical frame
  uint32_t vshift, eshift;  // log2 of buckets / edges per frame
  uint32_t num_metas;
};

__device__ __forceinline__ uint32_t lane_id() {
  return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}
__device__ __forceinline__ uint64_t wave_ballot(bool p) { return __ballot(p); }
__device__ __forceinline__ uint32_t shfl32(uint32_t v, uint32_t src) {
  return (uint32_t)__builtin_amdgcn_ds_bpermute((int)(src << 2), (int)v);
}
__device__ __forceinline__ uint64_t shfl64(uint64_t v, uint32_t src) {
  uint32_t lo = shfl32((uint32_t)v, src), hi = shfl32((uint32_t)(v >> 32), src);
  return ((uint64_t)hi << 32) | lo;
}

template <bool FRAMED>
__device__ __forceinline__ uint64_t bucket_phys(const GraphView &g, const DevMeta &m, uint64_t local) {
  if (!FRAMED) return local;  // resident: callers pass global ids
  const uint64_t mask = (uint64_t(1) << g.vshift) - 1;
  return g.vframes[m.vframe_base + (local >> g.vshift)] + (local & mask);
}

template <bool FRAMED>
__device__ __forceinline__ uint32_t edge_at(const GraphView &g, uint64_t edge_start, uint64_t eframe_base,
                                            uint64_t local) {
  if (!FRAMED) return g.edges[edge_start + local];
  const uint64_t mask = (uint64_t(1) << g.eshift) - 1;
  return g.edges[g.eframes[eframe_base + (local >> g.eshift)] + (local & mask)];
}

// Cluster-chaining probe, 8 lanes per key: each lane of a group loads one
// 16-byte slot of the bucket (one 128-byte coalesced read per bucket), a
// wave ballot finds the matching slot, and lane 7's link moves the group to
// the next bucket (get_vertex_local, core/gstore.hpp:389-409; device form
// core/gpu_hash.cu:181-234). Every lane of the wave must call this; all 8
// lanes of a group carry the same key. key == 0 means "no key" (a miss).
// Returns, in every lane of the group: deg (0 on a miss) and the list's
// offset relative to its predicate's edge_start.
template <int U, bool FRAMED>
__device__ __forceinline__ void group_probe(const GraphView &g, const uint64_t (&key)[U],
                                            const uint32_t (&pid)[U], uint32_t (&deg)[U],
                                            uint32_t (&off)[U], uint32_t *hops_out = nullptr) {
  const uint32_t lane = lane_id(), l8 = lane & 7, gbase = lane & ~7u;
  uint64_t bkt[U];
  bool act[U];
  uint64_t istart[U], psz[U], vfb[U], estart[U];
  uint32_t hops = 0;
#pragma unroll
  for (int u = 0; u < U; u++) {
    deg[u] = 0;
    off[u] = 0;
    act[u] = key[u] != 0 && pid[u] < g.num_metas;
    bkt[u] = 0;
    if (act[u]) {
      const DevMeta &m = g.meta[pid[u]];
      psz[u] = m.psz;
      istart[u] = m.indrct_start;
      vfb[u] = m.vframe_base;
      estart[u] = m.edge_start;
      if (psz[u] == 0) {
        act[u] = false;
      } else {
        uint64_t local = fastmod(mix64(key[u]), m.psz, m.magic);
        bkt[u] = FRAMED ? bucket_phys<true>(g, m, local) : m.main_start + local;
      }
    }
  }
  while (wave_ballot([&] {
    bool a = false;
#pragma unroll
    for (int u = 0; u < U; u++) a |= act[u];
    return a;
  }())) {
    uint4 s[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      s[u] = make_uint4(0, 0, 0, 0);
      if (act[u]) s[u] = g.slots[bkt[u] * kAssoc + l8];
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const uint64_t k = ((uint64_t)s[u].y << 32) | s[u].x;
      const bool hit = act[u] && l8 < 7 && k == key[u];
      const uint32_t gb = (uint32_t)(wave_ballot(hit) >> gbase) & 0xffu;
      const uint32_t src = gbase + (gb ? (uint32_t)__builtin_ctz(gb) : 7u);
      const uint32_t plo = shfl32(s[u].z, src), phi = shfl32(s[u].w, src);
      const uint32_t klo = shfl32(s[u].x, src), khi = shfl32(s[u].y, src);
      if (act[u]) {
        const uint64_t v = ((uint64_t)phi << 32) | plo;
        if (gb) {
          deg[u] = (uint32_t)ptr_size(v);
          off[u] = (uint32_t)(ptr_off(v) - estart[u]);
          act[u] = false;
        } else {
          const uint64_t link = ((uint64_t)khi << 32) | klo;
          if (link == 0 || ++hops > kMaxHops) {
            act[u] = false;
          } else {
            const uint64_t nb = key_vid(link);
            if (FRAMED) {
              const DevMeta &m = g.meta[pid[u]];
              bkt[u] = bucket_phys<true>(g, m, nb - istart[u] + psz[u]);
            } else {
              bkt[u] = nb;
            }
          }
        }
      }
    }
  }
  (void)vfb;
  if (hops_out) *hops_out = hops;
}

// Inclusive wave scan of u64 values.
__device__ __forceinline__ uint64_t wave_incl_scan(uint64_t x) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    uint64_t y = shfl64(x, (lane - d) & 63);
    if (lane >= (uint32_t)d) x += y;
  }
  return x;
}

// Exclusive block scan over kBlock threads; `total` gets the block sum.
// `tmp` is kBlock/kWave u64 of LDS.
__device__ __forceinline__ uint64_t block_excl_scan(uint64_t v, uint64_t *tmp, uint64_t &total) {
  const uint32_t lane = lane_id(), wave = threadIdx.x / kWave;
  const uint64_t inc = wave_incl_scan(v);
  if (lane == kWave - 1) tmp[wave] = inc;
  __syncthreads();
  uint64_t pre = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kBlock / kWave; w++) {
    const uint64_t t = tmp[w];
    if ((uint32_t)w < wave) pre += t;
    tot += t;
  }
  __syncthreads();
  total = tot;
  return pre + inc - v;
}

// Decoupled look-back (single pass) over tiles handed out in launch order
// by an atomic ticket, so a tile only ever waits on tiles that are already
// running. status[t] = value << 2 | flag, flag 1 = aggregate, 2 = inclusive.
// Called by one thread; returns the tile's exclusive prefix.
__device__ __forceinline__ uint64_t lookback(uint64_t *status, uint64_t tile, uint64_t agg) {
  if (tile == 0) {
    __hip_atomic_store(&status[0], (agg << 2) | 2u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    return 0;
  }
  __hip_atomic_store(&status[tile], (agg << 2) | 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  uint64_t excl = 0;
  uint64_t j = tile - 1;
  while (true) {
    const uint64_t s = __hip_atomic_load(&status[j], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t f = (uint32_t)(s & 3u);
    if (f == 0) {
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    excl += s >> 2;
    if (f == 2) break;
    j--;
  }
  __hip_atomic_store(&status[tile], ((excl + agg) << 2) | 2u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  return excl;
}

}  // namespace wkg
