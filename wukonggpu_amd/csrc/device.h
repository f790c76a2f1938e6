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
constexpr int kBlock = 256;            // threads per workgroup for every step kernel
constexpr int kTileRows = 8 * kBlock;  // probe tile: 8 rows per thread
constexpr int kExpandVT = 2;           // merged items per thread in the expand
constexpr int kExpandNV = kBlock * kExpandVT;
constexpr uint32_t kMaxHops = 1u << 20;  // chain-walk guard (images are validated)

// Read-back to the host without a stream synchronisation: a kernel writes
// counters into a pinned host block and then a sequence number into the
// block's flag word (system-scope release); the host polls the flag.
constexpr int kRbWords = 8;   // counters per read-back, at words [0, kRbWords)
constexpr int kRbFlag = 8;    // the flag word
constexpr uint64_t kStageOff = 4096;         // byte offset of the staging area for small tables
constexpr uint64_t kStageBytes = 64 << 10;   // tables up to this size reach the host through it
struct Signal {
  uint64_t *h;   // the pinned block (device address)
  uint64_t seq;
};
// One thread: v[0, n) then the flag.
__device__ __forceinline__ void publish(const Signal &sg, const uint64_t *v, int n) {
  for (int i = 0; i < n; i++) sg.h[i] = v[i];
  __hip_atomic_store(&sg.h[kRbFlag], sg.seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

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

// Lane i gets lane i-1's value, lane 0 gets 0 (DPP wave_shr:1, no LDS).
__device__ __forceinline__ uint32_t wave_shr1(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x138, 0xf, 0xf, false);
}
__device__ __forceinline__ uint64_t wave_shr1_64(uint64_t v) {
  return ((uint64_t)wave_shr1((uint32_t)(v >> 32)) << 32) | wave_shr1((uint32_t)v);
}
// Lane 63's value, in a scalar register.
__device__ __forceinline__ uint32_t lane63(uint32_t v) { return (uint32_t)__builtin_amdgcn_readlane((int)v, 63); }
__device__ __forceinline__ uint64_t lane63_64(uint64_t v) {
  return ((uint64_t)lane63((uint32_t)(v >> 32)) << 32) | lane63((uint32_t)v);
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

// Per-wave LDS scratch of wave_probe.
struct WaveProbeLds {
  uint4 kb[kWave];   // pending bucket walks: (key lo, key hi, bucket, leader index)
  uint2 res[kWave];  // run leaders' (degree, offset)
};

// Cluster-chaining probe of 64 consecutive rows, one per lane, all under one
// predicate (get_vertex_local, core/gstore.hpp:389-409; device form
// core/gpu_hash.cu:181-234).
//  0. The caller's carry holds the key of the row just before lane 0 (the
//     previous call's lane 63) and its result: a run of equal keys that
//     continues from the previous call is not probed again.
//  1. A run of equal keys in consecutive lanes is probed once: its first lane
//     (the run leader) hashes the key and parks (key, bucket, its rank among
//     the leaders) in LDS as a pending walk.
//  2. Pending walks are taken 8 lanes per walk: group g takes walks g, g+8,
//     g+16, ...; each lane of the group loads one 16-byte slot of the bucket
//     (one coalesced 128-byte read; a group's walks are loaded together, so
//     up to 8 bucket reads are in flight per lane). The lane whose slot
//     holds the key writes the leader's (degree, offset) to LDS; in a group
//     with no match, lane 7 (the link slot) either records a miss or appends
//     the walk, moved to the linked bucket, to the next hop's pending list.
//     The next hop only walks what is still pending, compacted.
//  3. Every lane reads its run leader's result.
// Buckets are addressed with 32 bits (images hold fewer than 2^32 buckets).
// Every lane of the wave must call this with the wave's own `w`. key == 0
// means "no key". Gives the lane's degree (0 on a miss) and list offset
// relative to the predicate's edge_start, moves the carry to lane 63 (it
// starts as key 0: valid keys are never 0) and returns the number of keys
// the wave probed (its run leaders).
struct ProbeCarry {
  uint64_t key = 0;
  uint32_t deg = 0, off = 0;
};

// A pointer read from memory (a parameter record in HBM) is generic to the
// compiler, which then accesses it with flat instructions. Those count
// against the LDS wait counter as well, so every later LDS wait or workgroup
// barrier also waits for them. The same pointer in the global address space
// keeps the accesses vector-memory instructions. (Kernel-argument pointers
// are global already; the cast costs nothing there.)
#define WKG_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ WKG_GLOBAL T *gmem(T *p) {
  return (WKG_GLOBAL T *)p;
}
// HIP's uint2/uint4 cannot be copied in another address space (their
// operators are members); these move them as plain clang vectors.
typedef uint32_t gu32x2 __attribute__((ext_vector_type(2)));
typedef uint32_t gu32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ uint2 gload(const uint2 *p) {
  const gu32x2 v = *(const WKG_GLOBAL gu32x2 *)p;
  return make_uint2(v.x, v.y);
}
__device__ __forceinline__ uint4 gload(const uint4 *p) {
  const gu32x4 v = *(const WKG_GLOBAL gu32x4 *)p;
  return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ void gstore(uint2 *p, uint2 v) { *(WKG_GLOBAL gu32x2 *)p = gu32x2{v.x, v.y}; }
__device__ __forceinline__ void gstore(uint4 *p, uint4 v) { *(WKG_GLOBAL gu32x4 *)p = gu32x4{v.x, v.y, v.z, v.w}; }

// Bit mask of the lanes below this one.
__device__ __forceinline__ uint64_t lanes_below() {
  const uint32_t lane = lane_id();
  return lane ? (~0ull >> (64 - lane)) : 0ull;
}

template <bool FRAMED>
__device__ __forceinline__ uint32_t wave_probe(const GraphView &g, const DevMeta &m, uint64_t key, WaveProbeLds &w,
                                               uint32_t &deg_out, uint32_t &off_out, ProbeCarry &carry) {
  const uint32_t lane = lane_id(), l8 = lane & 7, grp = lane >> 3, gbase = lane & ~7u;
  const bool valid = key != 0 && m.psz != 0;
  const uint64_t before = wave_shr1_64(key);
  const uint64_t prev = lane == 0 ? carry.key : before;
  const bool lead = valid && key != prev;
  const uint64_t leaders = wave_ballot(lead);
  const uint32_t L = (uint32_t)__builtin_popcountll(leaders);
  if (L == 0) {  // every valid lane continues the carried run (wave-uniform)
    deg_out = valid ? carry.deg : 0u;
    off_out = valid ? carry.off : 0u;
    if (lane63_64(key) != carry.key) carry = ProbeCarry{};  // lane 63 has no key
    return 0;
  }
  const uint64_t below = lanes_below();
  if (lead) {
    const uint32_t j = (uint32_t)__builtin_popcountll(leaders & below);
    const uint64_t local = fastmod(mix64(key), m.psz, m.magic);
    const uint64_t b = FRAMED ? bucket_phys<true>(g, m, local) : m.main_start + local;
    w.kb[j] = make_uint4((uint32_t)key, (uint32_t)(key >> 32), (uint32_t)b, j);
  }
  __builtin_amdgcn_wave_barrier();
  uint32_t n = L;  // pending walks (wave-uniform)
  for (uint32_t hop = 0; n; hop++) {
    const uint32_t umax = (n + 7) >> 3;
    uint4 e[8], sl[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      e[u] = make_uint4(0, 0, 0, 0);
      sl[u] = make_uint4(0, 0, 0, 0);
      if ((uint32_t)u < umax && grp + 8u * u < n) {
        e[u] = w.kb[grp + 8u * u];
        sl[u] = g.slots[(uint64_t)e[u].z * kAssoc + l8];
      }
    }
    // (every pending entry is read above before any is rewritten below: a
    // wave's LDS operations complete in order)
    uint32_t nn = 0;
#pragma unroll
    for (int u = 0; u < 8; u++) {
      if ((uint32_t)u >= umax) break;  // wave-uniform
      const bool in = grp + 8u * u < n;
      const bool hit = in && l8 < 7 && sl[u].x == e[u].x && sl[u].y == e[u].y;
      if (hit) {
        const uint64_t v = ((uint64_t)sl[u].w << 32) | sl[u].z;
        w.res[e[u].w] = make_uint2((uint32_t)ptr_size(v), (uint32_t)(ptr_off(v) - m.edge_start));
      }
      const uint32_t gb = (uint32_t)(wave_ballot(hit) >> gbase) & 0x7fu;
      const uint64_t link = ((uint64_t)sl[u].y << 32) | sl[u].x;
      const bool tail = in && l8 == 7 && gb == 0;  // lane 7 of a group without a match
      const bool follow = tail && link != 0 && hop + 1 < kMaxHops;
      if (tail && !follow) w.res[e[u].w] = make_uint2(0u, 0u);
      const uint64_t fb = wave_ballot(follow);
      if (follow) {
        const uint64_t nb = key_vid(link);
        const uint64_t bkt = FRAMED ? bucket_phys<true>(g, m, nb - m.indrct_start + m.psz) : nb;
        w.kb[nn + (uint32_t)__builtin_popcountll(fb & below)] = make_uint4(e[u].x, e[u].y, (uint32_t)bkt, e[u].w);
      }
      nn += (uint32_t)__builtin_popcountll(fb);
    }
    __builtin_amdgcn_wave_barrier();
    n = nn;
  }
  deg_out = 0;
  off_out = 0;
  if (valid) {
    // leaders at or below this lane; none: the lane continues the carried run
    const uint32_t nl = (uint32_t)__builtin_popcountll(leaders & (below | (1ull << lane)));
    if (nl) {
      const uint2 r = w.res[nl - 1];
      deg_out = r.x;
      off_out = r.y;
    } else {
      deg_out = carry.deg;
      off_out = carry.off;
    }
  }
  __builtin_amdgcn_wave_barrier();  // w is reused by the wave's next call
  carry.key = lane63_64(key);
  carry.deg = lane63(deg_out);
  carry.off = lane63(off_out);
  return L;
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

// LDS written by some lanes of a wave and read by others: order the
// accesses at wave scope.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// A workgroup barrier that orders LDS only: global memory operations still
// in flight are not waited for.
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Exclusive block scans of N values at once (one pair of barriers): v[k]
// becomes its exclusive prefix over the block (BS threads), total[k] the
// block sum. `tmp` is N * BS/kWave u64 of LDS. LDS_ONLY: the barriers order
// LDS only.
template <int N, int BS = kBlock, bool LDS_ONLY = false>
__device__ __forceinline__ void block_excl_scan_n(uint64_t (&v)[N], uint64_t *tmp, uint64_t (&total)[N]) {
  const uint32_t lane = lane_id(), wave = threadIdx.x / kWave;
  constexpr int W = BS / kWave;
  uint64_t inc[N];
#pragma unroll
  for (int k = 0; k < N; k++) inc[k] = wave_incl_scan(v[k]);
  if (lane == kWave - 1) {
#pragma unroll
    for (int k = 0; k < N; k++) tmp[k * W + wave] = inc[k];
  }
  if constexpr (LDS_ONLY) lds_barrier(); else __syncthreads();
#pragma unroll
  for (int k = 0; k < N; k++) {
    uint64_t pre = 0, tot = 0;
    if constexpr (W <= 4) {  // every thread sums the few wave totals
#pragma unroll
      for (int w = 0; w < W; w++) {
        const uint64_t t = tmp[k * W + w];
        if ((uint32_t)w < wave) pre += t;
        tot += t;
      }
    } else {  // lane j reads wave j's total; a wave scan of them
      const uint64_t ti = wave_incl_scan(lane < (uint32_t)W ? tmp[k * W + lane] : 0ull);
      pre = wave ? shfl64(ti, wave - 1) : 0ull;
      tot = shfl64(ti, W - 1);
    }
    v[k] = pre + inc[k] - v[k];
    total[k] = tot;
  }
  if constexpr (LDS_ONLY) lds_barrier(); else __syncthreads();
}

}  // namespace wkg
