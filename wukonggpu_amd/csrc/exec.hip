// Device-resident batched execution: whole queries advance through all of
// their pattern steps on the device, with no host round trip between steps.
//
// A batch is nq queries, each a compiled plan of pattern steps (the step
// GPU_Engine::execute_one_step dispatches, core/gpu_engine.hpp:395-467). Step
// level l (the l-th pattern of every query that has one) runs as
//
//   k_x_probe        persistent: probe tiles of all queries from the work queue
//   k_x_scan_chunks  persistent: tile sums -> chunk-local prefixes; the block
//                    that finishes a query's last chunk scans its chunk sums
//   k_x_finalize     persistent: compaction tiles, one wave per tile
//   k_x_plan         1 block: per query, read the totals, allocate the
//                    known_to_unknown output, commit the query's new table,
//                    give finishing queries their place in the host buffer;
//                    then size level l+1 (rows, tiles, scan chunks, scratch,
//                    filter output), carve it out of the HBM arena, write its
//                    step descriptors and tile prefix
//   k_x_partition    persistent: merge-path split points of all expands
//   k_x_expand       persistent: expand tiles from a second work queue
//   k_x_retire       persistent: finished queries' tables -> host buffer
//
// after a first k_x_plan that starts the query states and sizes level 0.
// Level l+1's plan only computes places in the arena; nothing writes them
// before level l's expand and retire are done (stream order), so planning it
// early is safe although its window overlaps level l's scratch and inputs.
// and k_x_records writes the per-query records that are copied back. The
// per-tile bodies are the plain path's (step_kernels.h), so each query's
// table is the one the per-step calls produce, cell for cell.
//
// Memory: one arena per batch object. Level l's output tables go to stack
// (l & 1) (stack 0 grows up from the bottom, stack 1 down from the top),
// whose previous contents -- level l-2's tables -- are dead by then; the
// level's scratch is carved from the other end of the free window between
// the stacks. So only one level's inputs, outputs and scratch are live at a
// time. A query whose needs do not fit ends with WKG_ENOMEM; which queries
// fit is decided in query order, so the outcome is deterministic.
//
// The launch sequence for a given level count is captured once as a hipGraph
// (no memset nodes: the plan kernels reset the counters) and replayed.
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <map>
#include <vector>

#include "../../include/wukong_gpu.h"
#include "step_kernels.h"
#include "wkg_internal.h"

using namespace wkg;

namespace {

constexpr uint64_t kAlign = 256;
__host__ __device__ __forceinline__ uint64_t al(uint64_t b) { return (b + kAlign - 1) & ~(kAlign - 1); }

constexpr int32_t kInactive = 1;         // QState::status of an unused slot
constexpr uint64_t kCopyUnit = 16384;    // bytes per retire work unit
constexpr uint64_t kRowLimit = 1ull << 32;
constexpr uint32_t kSubQueues = 8;       // one per XCD (blocks are dealt to the XCDs round-robin)

struct QState {
  const int32_t *T;   // current table (null before the first step)
  uint64_t rows;
  uint64_t expanded;
  uint64_t out_off;   // first cell in the host result buffer, once retired
  uint32_t cols;
  int32_t status;     // WKG_OK, a WKG_E* code, or kInactive
  uint32_t nsteps;
  uint32_t nchunks;   // tile-scan chunks of the current step
  uint32_t chunks_done;  // chunks scanned so far (the last one scans the chunk sums)
};

struct ExecState {
  uint64_t nq;
  uint64_t lo_top;     // table stack 0 (even levels' outputs): [0, lo_top)
  uint64_t hi_low;     // table stack 1 (odd levels' outputs): [hi_low, cap)
  uint64_t w0, w1;     // this level's free window
  uint64_t out_used;   // this level's outputs, bytes from the window's output end
  uint64_t scr_used;   // this level's scratch, bytes from the other end
  uint64_t nt, nc, np, nx, nu;  // probe tiles, scan chunks, partition entries, expand tiles, retire units
  uint64_t res_bytes;  // host result buffer bytes used
  uint64_t res_over;   // result bytes refused for lack of host buffer
  uint64_t peak;       // arena high-water mark
  uint64_t pad[2];
  // work queues of the probe (0) and expand (1) passes: 8 sub-queue heads
  // each, one 64-byte line apart so their atomics do not serialise on one line
  uint64_t queue[2][kSubQueues][8];
};

struct Item {
  StepParams p;
  ExpandParams x;
};

struct ExecArgs {
  GraphView g;
  ExecState *st;
  wkg_batch_result *res;  // [Q], right after st (one read-back copy)
  // query states as columns of 64-bit words (word k of query q at
  // [k * Q + q]), so that one thread per query reads or writes them with
  // consecutive addresses across a wave
  uint64_t *qsw;
  Item *items;            // [Q] the current passes' parameters (read per tile)
  uint64_t *qtot;         // [Q][4] each query's step totals (StepParams::totals points here)
  uint64_t *tile_base;    // [Q+1] probe / finalize tiles
  uint64_t *chunk_base;   // [Q+1] tile-scan chunks
  uint64_t *part_base;    // [Q+1]
  uint64_t *xtile_base;   // [Q+1]
  uint64_t *unit_base;    // [Q+1] retire copy units
  const uint64_t *up_nq;  // uploaded: query count
  const uint32_t *up_nsteps;       // uploaded: [Q]
  const wkg_plan_step *up_steps;   // uploaded: [Q * P], query q's level l at q * P + l
  char *arena;
  uint64_t cap;
  char *hres;             // host result buffer (device-visible pinned memory)
  uint64_t hcap;
  uint32_t Q, P;
};

template <typename T>
__device__ __forceinline__ T soa_get(const uint64_t *w, uint32_t Q, uint32_t q) {
  static_assert(sizeof(T) % 8 == 0, "whole words");
  uint64_t t[sizeof(T) / 8];
#pragma unroll
  for (uint32_t k = 0; k < sizeof(T) / 8; k++) t[k] = w[(uint64_t)k * Q + q];
  T v;
  __builtin_memcpy(&v, t, sizeof(T));
  return v;
}
template <typename T>
__device__ __forceinline__ void soa_put(uint64_t *w, uint32_t Q, uint32_t q, const T &v) {
  static_assert(sizeof(T) % 8 == 0, "whole words");
  uint64_t t[sizeof(T) / 8];
  __builtin_memcpy(t, &v, sizeof(T));
#pragma unroll
  for (uint32_t k = 0; k < sizeof(T) / 8; k++) w[(uint64_t)k * Q + q] = t[k];
}
// A 32-bit field of query q's QState (its byte offset in the struct).
__device__ __forceinline__ uint32_t *qs_field(const ExecArgs &a, uint32_t q, uint32_t off) {
  return (uint32_t *)((char *)(a.qsw + (uint64_t)(off / 8) * a.Q + q) + off % 8);
}
__device__ __forceinline__ QState qs_get(const ExecArgs &a, uint32_t q) { return soa_get<QState>(a.qsw, a.Q, q); }
__device__ __forceinline__ void qs_put(const ExecArgs &a, uint32_t q, const QState &s) { soa_put(a.qsw, a.Q, q, s); }

// Arena placement. Level l writes its output tables into stack (l & 1) --
// stack 0 grows up from 0, stack 1 down from cap -- after dropping that
// stack's old contents (level l-2's tables: dead, since their queries either
// moved on or were retired to the host). Scratch is carved from the other
// end of the window between the stacks, towards the outputs.
__device__ __forceinline__ char *out_at(const ExecArgs &a, uint32_t level, uint64_t w0, uint64_t w1, uint64_t off,
                                        uint64_t bytes) {
  return (level & 1) ? a.arena + w1 - off - bytes : a.arena + w0 + off;
}
__device__ __forceinline__ char *scr_at(const ExecArgs &a, uint32_t level, uint64_t w0, uint64_t w1, uint64_t off,
                                        uint64_t bytes) {
  return (level & 1) ? a.arena + w0 + off : a.arena + w1 - off - bytes;
}

// Last q in [0, n] with base[q] <= t, for base ascending with base[n] > t.
// Wave 0 runs a 64-ary search (each lane tests one pivot; the lanes that pass
// form a prefix) and the block gets the answer through LDS.
__device__ __forceinline__ uint32_t block_find(const uint64_t *base, uint32_t n, uint64_t t, uint32_t *s_q) {
  if (threadIdx.x < kWave) {
    const uint32_t lane = threadIdx.x;
    uint32_t lo = 0, hi = n;  // base[lo] <= t < base[hi]
    while (hi - lo > 1) {
      const uint32_t span = hi - lo;
      const uint32_t pos = lo + (uint32_t)(((uint64_t)span * (lane + 1)) / (kWave + 1));  // in [lo, hi)
      const bool le = base[pos] <= t;
      const uint32_t k = (uint32_t)__builtin_popcountll(wave_ballot(le));
      const uint32_t plo = shfl32(pos, k ? k - 1 : 0), phi = shfl32(pos, k < kWave ? k : 0);
      if (k) lo = plo;
      if (k < kWave) hi = phi;
    }
    if (lane == 0) *s_q = lo;
  }
  __syncthreads();
  return *s_q;
}

// The same search done by one wave for itself (the result is wave-uniform).
__device__ __forceinline__ uint32_t wave_find(const uint64_t *base, uint32_t n, uint64_t t) {
  const uint32_t lane = lane_id();
  uint32_t lo = 0, hi = n;
  while (hi - lo > 1) {
    const uint32_t span = hi - lo;
    const uint32_t pos = lo + (uint32_t)(((uint64_t)span * (lane + 1)) / (kWave + 1));
    const bool le = base[pos] <= t;
    const uint32_t k = (uint32_t)__builtin_popcountll(wave_ballot(le));
    const uint32_t plo = shfl32(pos, k ? k - 1 : 0), phi = shfl32(pos, k < kWave ? k : 0);
    if (k) lo = plo;
    if (k < kWave) hi = phi;
  }
  return lo;
}

// Thread-level binary search with the same contract.
__device__ __forceinline__ uint32_t thread_find(const uint64_t *base, uint32_t n, uint64_t t) {
  uint32_t lo = 0, hi = n;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (base[mid] <= t) lo = mid; else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ bool level_active(const QState &s, uint32_t level) {
  return s.status == WKG_OK && level < s.nsteps;
}

// The shared work queue over a pass's flattened tile space (all queries'
// tiles back to back). It is split into kSubQueues interleaved sub-queues --
// sub-queue k hands out chunks k, k + 8, k + 16, ... of K consecutive tiles
// -- and block b draws from sub-queue b % 8, so the heads' atomics spread
// over eight addresses and each draw covers K tiles. K gives each block
// about four draws. Thread 0 draws the next chunk while the current one is
// worked on. The block's queue position and the query of its current tile
// (consecutive tiles mostly belong to one query) live in LDS, so the tile
// bodies keep their registers.
struct QueueLds {
  uint64_t cur, end, nxt, t;
  uint64_t qlo, qhi;  // tile range of query q
  uint32_t q, sub;
  uint32_t need;      // tile t is outside [qlo, qhi)
};

// A block-uniform LDS value as a scalar (so what it indexes is loaded into
// scalar registers, not once per lane).
__device__ __forceinline__ uint64_t uniform64(uint64_t v) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

struct BlockQueue {
  QueueLds &s;
  unsigned long long *head;
  uint64_t n, K;
  __device__ BlockQueue(QueueLds &lds, uint64_t (*heads)[8], uint64_t total) : s(lds), n(total) {
    const uint32_t sub = blockIdx.x % kSubQueues;
    head = (unsigned long long *)heads[sub];
    const uint64_t per = (uint64_t)gridDim.x * 4;
    K = total > per ? total / per : 1;
    if (threadIdx.x == 0) {
      s.cur = s.end = 0;
      s.sub = sub;
      s.qlo = 1;
      s.qhi = 0;
      s.nxt = atomicAdd(head, 1ull);
    }
  }
  // Next tile and its query; false once the sub-queue is drained.
  __device__ bool next(const uint64_t *base, uint32_t nq, uint64_t &t, uint32_t &q, uint64_t &lt) {
    __syncthreads();  // the previous tile's LDS use is over
    if (threadIdx.x == 0) {
      if (s.cur >= s.end) {
        const uint64_t c = s.sub + (uint64_t)kSubQueues * s.nxt;
        s.cur = c * K;
        s.end = s.cur + K < n ? s.cur + K : n;
        if (s.cur < n) s.nxt = atomicAdd(head, 1ull);
      }
      s.t = s.cur < s.end ? s.cur++ : n;
      s.need = s.t < n && (s.t < s.qlo || s.t >= s.qhi);
    }
    __syncthreads();
    t = uniform64(s.t);
    if (t >= n) return false;
    if (s.need) {  // block-uniform: wave 0 finds the tile's query
      if (threadIdx.x < kWave) {
        const uint32_t qq = wave_find(base, nq, t);
        if (threadIdx.x == 0) {
          s.q = qq;
          s.qlo = base[qq];
          s.qhi = base[qq + 1];
        }
      }
      __syncthreads();
    }
    q = (uint32_t)__builtin_amdgcn_readfirstlane((int)s.q);
    lt = t - uniform64(s.qlo);
    return true;
  }
};

// The static deal: block b takes tiles b, b + G, b + 2G, ... of a pass's
// flat tile space (G blocks). No atomics, and blocks differ by at most one
// tile, which suits passes whose tiles cost about the same (large steps).
// (Equal contiguous runs instead of this interleave measured slower on the
// 512-query mix's first hop: tiles of one query cost alike, so a run of one
// expensive query's tiles lands on one block.)
struct StaticDeal {
  uint32_t &sq;  // LDS slot of the query search
  uint64_t n, t, qlo, qhi;
  uint32_t q;
  __device__ StaticDeal(uint32_t &lds, uint64_t total) : sq(lds), n(total), t(blockIdx.x), qlo(1), qhi(0), q(0) {}
  __device__ bool next(const uint64_t *base, uint32_t nq, uint64_t &tt, uint32_t &qq, uint64_t &lt) {
    __syncthreads();  // the previous tile's LDS use is over
    if (t >= n) return false;
    if (t < qlo || t >= qhi) {  // block-uniform
      q = (uint32_t)__builtin_amdgcn_readfirstlane((int)block_find(base, nq, t, &sq));
      qlo = base[q];
      qhi = base[q + 1];
    }
    tt = t;
    qq = q;
    lt = t - qlo;
    t += gridDim.x;
    return true;
  }
};

// ---- level plan ------------------------------------------------------------

struct LevelNeed {
  bool active;
  uint64_t R, nt, nc, scratch, table;
};

template <int TR>
__device__ __forceinline__ LevelNeed level_need(const QState &s, const wkg_plan_step &d, uint32_t level) {
  LevelNeed n{false, 0, 0, 0, 0, 0};
  if (!level_active(s, level)) return n;
  const bool key = d.kind == WKG_STEP_I2U || d.kind == WKG_STEP_C2U;
  n.active = true;
  n.R = key ? 1 : s.rows;
  n.nt = (n.R + TR - 1) / TR;
  n.nc = n.nt > kScanChunk ? (n.nt + kScanChunk - 1) / kScanChunk : 1;  // an empty step still writes totals
  // tile_sum | tile_nz | tile_ex | chunk sums x 3 | hits (TR per tile, 16 B) | K2U: cinfo, cexcl
  n.scratch = al(24 * n.nt + 24 * n.nc) + n.nt * TR * 16 + (d.kind == WKG_STEP_K2U || key ? 2 * al(n.R * 8) : 0);
  n.table = (d.kind == WKG_STEP_K2K || d.kind == WKG_STEP_K2C) ? al(n.R * s.cols * 4) : 0;
  return n;
}

// The one-block plan kernels run kPlanBlock threads, one query each per
// chunk of kPlanBlock queries; a chunk's needs are scanned once, and chunks
// follow one another with the running prefixes carried over.
constexpr int kPlanBlock = 512;

// A record read or written by its own thread in an array of structs puts
// up to 64 cache lines in one wave instruction, and the plan's one CU takes
// those lines one after another. So the query states are stored as columns
// (ExecArgs::qsw), and the parameter records, which the tile kernels read
// whole, pass through the wave's own LDS area: the wave's 64 records are
// copied as consecutive words (a few lines per instruction).
constexpr uint32_t kStageWords = sizeof(StepParams) / 8;  // the largest staged record
struct PlanLds {
  uint64_t w[kPlanBlock / kWave][kWave * kStageWords + kWave / 2];  // + the put flags
};

// Lanes' records of queries q0 + lane (those below n) into the field at `off`
// of an array of Items; only lanes with `put` are stored.
template <typename T>
__device__ __forceinline__ void wave_put(Item *items, uint32_t off, uint32_t q0, uint32_t n, const T &v, bool put,
                                         uint64_t *lw) {
  static_assert(sizeof(T) % 8 == 0 && sizeof(T) / 8 <= kStageWords, "record");
  constexpr uint32_t W = sizeof(T) / 8;
  uint32_t *flag = (uint32_t *)(lw + kWave * kStageWords);
  const uint32_t lane = lane_id();
  uint64_t t[W];
  __builtin_memcpy(t, &v, sizeof(T));
#pragma unroll
  for (uint32_t k = 0; k < W; k++) lw[lane * W + k] = t[k];
  flag[lane] = put && q0 + lane < n ? 1u : 0u;
  wave_lds_sync();
#pragma unroll
  for (uint32_t u = 0; u < W; u++) {
    const uint32_t m = u * kWave + lane, i = m / W, k = m - i * W;
    if (flag[i]) *(uint64_t *)((char *)&items[q0 + i] + off + 8 * k) = lw[m];
  }
  wave_lds_sync();  // the area is free again
}

// The field at `off` of items[q0 + lane] for lanes with q0 + lane < n (the
// others get T{}).
template <typename T>
__device__ __forceinline__ T wave_get(const Item *items, uint32_t off, uint32_t q0, uint32_t n, uint64_t *lw) {
  static_assert(sizeof(T) % 8 == 0 && sizeof(T) / 8 <= kStageWords, "record");
  constexpr uint32_t W = sizeof(T) / 8;
  const uint32_t lane = lane_id();
  uint64_t v[W];
#pragma unroll
  for (uint32_t u = 0; u < W; u++) {  // every load is issued before the first is used
    const uint32_t m = u * kWave + lane, i = m / W, k = m - i * W;
    v[u] = q0 + i < n ? *(const uint64_t *)((const char *)&items[q0 + i] + off + 8 * k) : 0ull;
  }
#pragma unroll
  for (uint32_t u = 0; u < W; u++) lw[u * kWave + lane] = v[u];
  wave_lds_sync();
  uint64_t t[W];
#pragma unroll
  for (uint32_t k = 0; k < W; k++) t[k] = lw[lane * W + k];
  wave_lds_sync();
  T r;
  __builtin_memcpy(&r, t, sizeof(T));
  return r;
}

// Starts a batch: query states from the upload, the arena stacks empty.
__device__ __forceinline__ uint32_t begin_phase(const ExecArgs &a) {
  const uint64_t nq = *a.up_nq;
  for (uint32_t c0 = 0; c0 < a.Q; c0 += kPlanBlock) {
    const uint32_t q = c0 + threadIdx.x;
    QState s{};
    s.T = nullptr;
    s.status = q < nq ? WKG_OK : kInactive;
    s.nsteps = q < nq ? a.up_nsteps[q] : 0;
    if (q < a.Q) qs_put(a, q, s);
  }
  if (threadIdx.x == 0) {  // (the level plan writes w0, w1, out_used, scr_used, nt and nc)
    ExecState &st = *a.st;
    st.nq = nq;
    st.lo_top = 0;
    st.hi_low = a.cap;
    st.np = st.nx = st.nu = 0;
    st.res_bytes = st.res_over = st.peak = 0;
  }
  return (uint32_t)nq;
}

// Plans level `level` in the window between the stacks: [lo_top, cap) for
// odd levels (stack 1 is refilled), [0, hi_low) for even ones.
template <int TR>
__device__ __forceinline__ void level_phase(const ExecArgs &a, uint32_t level, uint32_t nq, uint64_t lo_top,
                                            uint64_t hi_low, PlanLds &pl) {
  __shared__ uint64_t s_tmp[2 * (kPlanBlock / kWave)];
  uint64_t *lw = pl.w[threadIdx.x / kWave];
  __shared__ unsigned long long s_ut, s_us;
  ExecState *st = a.st;
  const uint64_t w0 = (level & 1) ? lo_top : 0, w1 = (level & 1) ? a.cap : hi_low;
  const uint64_t room = w1 - w0;
  if (threadIdx.x == 0) {  // ordered before the atomics by the scans' barriers
    s_ut = 0;
    s_us = 0;
  }
  const uint32_t lv = level < a.P ? level : a.P - 1;
  uint64_t carry_s = 0, carry_t = 0, carry_nt = 0, carry_nc = 0;
  for (uint32_t c0 = 0; c0 < nq; c0 += kPlanBlock) {
    const uint32_t q = c0 + threadIdx.x;
    const uint32_t qc = q < nq ? q : nq - 1;  // (nq >= 1 here)
    QState sv = qs_get(a, qc);
    const wkg_plan_step d = a.up_steps[(uint64_t)qc * a.P + lv];
    const LevelNeed n = q < nq ? level_need<TR>(sv, d, level) : LevelNeed{false, 0, 0, 0, 0, 0};
    uint64_t sz[2] = {n.scratch, n.table}, szt[2];
    block_excl_scan_n<2, kPlanBlock, true>(sz, s_tmp, szt);
    const uint64_t os = carry_s + sz[0], ot = carry_t + sz[1];
    // fitting queries (in query order) and their tiles and scan chunks
    const bool fit = n.active && ot + n.table + os + n.scratch <= room;
    uint64_t cnt[2] = {fit ? n.nt : 0, fit ? n.nc : 0}, cntt[2];
    block_excl_scan_n<2, kPlanBlock, true>(cnt, s_tmp, cntt);
    const uint64_t tbase = carry_nt + cnt[0], cbase = carry_nc + cnt[1];
    if (q < nq) {
      a.tile_base[q] = tbase;
      a.chunk_base[q] = cbase;
    }
    StepParams p{};
    if (n.active) {
      if (!fit) {
        sv.status = WKG_ENOMEM;
      } else {
        p.pid = (uint32_t)d.pid;
        p.dir = (uint32_t)d.dir;
        p.R = n.R;
        p.ntiles = n.nt;
        if (d.kind == WKG_STEP_I2U || d.kind == WKG_STEP_C2U) {
          p.key = d.kind == WKG_STEP_I2U ? make_key(0, (uint64_t)d.pid, (uint64_t)d.dir)
                                         : make_key((uint64_t)d.vid, (uint64_t)d.pid, (uint64_t)d.dir);
          p.mode = kModeK2U;
        } else {
          p.T = sv.T;
          p.c = sv.cols;
          p.sc = (uint32_t)d.start_col;
          p.mode = d.kind == WKG_STEP_K2U ? kModeK2U : d.kind == WKG_STEP_K2K ? kModeK2K : kModeK2C;
          p.ec = d.kind == WKG_STEP_K2K ? (uint32_t)d.end_col : 0u;
          p.target = (uint32_t)d.vid;
          p.target_ok = (d.kind == WKG_STEP_K2C && d.vid >= 0 && d.vid <= 0xffffffffLL) ? 1u : 0u;
        }
        char *sb = scr_at(a, level, w0, w1, os, n.scratch);
        p.totals = a.qtot + 4 * (uint64_t)q;
        p.tile_sum = (uint64_t *)sb;
        p.tile_nz = p.tile_sum + n.nt;
        p.tile_ex = p.tile_nz + n.nt;
        p.chunk_sum = p.tile_ex + n.nt;
        p.chunk_nz = p.chunk_sum + n.nc;
        p.chunk_ex = p.chunk_nz + n.nc;
        sb += al(24 * n.nt + 24 * n.nc);
        p.hits = (uint4 *)sb;
        sb += n.nt * TR * 16;
        if (p.mode == kModeK2U) {
          p.cinfo = (uint2 *)sb;
          p.cexcl = (uint64_t *)(sb + al(n.R * 8));
        } else {
          p.O = (int32_t *)out_at(a, level, w0, w1, ot, n.table);
        }
        sv.nchunks = (uint32_t)n.nc;
        sv.chunks_done = 0;
        // the fitting queries form a prefix (their cumulative needs are the
        // smallest), so the highest ends in use are the last fitting query's
        atomicMax(&s_ut, (unsigned long long)(ot + n.table));
        atomicMax(&s_us, (unsigned long long)(os + n.scratch));
      }
    }
    wave_put(a.items, offsetof(Item, p), q - lane_id(), nq, p, n.active && fit, lw);
    if (n.active) qs_put(a, q, sv);
    carry_s += szt[0];
    carry_t += szt[1];
    carry_nt += cntt[0];
    carry_nc += cntt[1];
  }
  lds_barrier();
  if (threadIdx.x == kPlanBlock - 1) {
    a.tile_base[nq] = carry_nt;
    a.chunk_base[nq] = carry_nc;
    st->w0 = w0;
    st->w1 = w1;
    st->out_used = s_ut;
    st->scr_used = s_us;
    st->nt = carry_nt;
    st->nc = carry_nc;
  }
  if (threadIdx.x < kSubQueues) st->queue[0][threadIdx.x][0] = 0;
}

// ---- probe, tile scan, finalize ----------------------------------------------

// 1024-row tiles: four waves per SIMD, like the per-step probe (the tile
// loop's state costs a few spills, all outside the probe's round loop; the
// 512-query mix's first hop went from 186 to 175 us). 2048-row tiles hold
// twice the rows' cells in registers: at four waves they spill inside the
// round loop (the heavy query's 145M-row probe took 2.8 ms instead of 1.5),
// so they keep three waves and no spills.
//
// The tile's step parameters come from the query's record in HBM. Read
// there by reference, every field is read again after each wave fence of
// the probe (a fence orders all memory), a memory round trip per round. So
// 2048-row tiles copy the record into registers once per tile (three waves
// leave room); 1024-row tiles, with no registers to spare at four waves,
// copy it into LDS and re-read it there.
template <int TR, bool SORTED, bool FRAMED>
__device__ __forceinline__ void x_probe_tile(const ExecArgs &a, uint32_t q, uint64_t lt, uint64_t *s_pw) {
  if constexpr (TR == 1024) {
    constexpr uint32_t kWords = sizeof(StepParams) / 8;
    static_assert(sizeof(StepParams) % 8 == 0, "the record is copied in 8-byte words");
    if (threadIdx.x < kWords) s_pw[threadIdx.x] = gmem((const uint64_t *)&a.items[q].p)[threadIdx.x];
    __syncthreads();  // (the next tile's copy comes after the deal's barrier)
    probe_tile<TR, SORTED, FRAMED>(a.g, *(const StepParams *)s_pw, lt);
  } else {
    const StepParams p = a.items[q].p;
    probe_tile<TR, SORTED, FRAMED>(a.g, p, lt);
  }
}

template <int TR, bool SORTED, bool FRAMED, bool QUEUE>
__global__ __launch_bounds__(kBlock, TR == 1024 ? 4 : 3) void k_x_probe(ExecArgs a) {
  __shared__ uint64_t s_pw[sizeof(StepParams) / 8];
  const uint32_t nq = (uint32_t)a.st->nq;
  uint64_t t, lt;
  uint32_t q;
  if constexpr (QUEUE) {
    __shared__ QueueLds s_wq;
    BlockQueue wq(s_wq, a.st->queue[0], a.st->nt);
    while (wq.next(a.tile_base, nq, t, q, lt)) x_probe_tile<TR, SORTED, FRAMED>(a, q, lt, s_pw);
  } else {
    __shared__ uint32_t s_q;
    StaticDeal wd(s_q, a.st->nt);
    while (wd.next(a.tile_base, nq, t, q, lt)) x_probe_tile<TR, SORTED, FRAMED>(a, q, lt, s_pw);
  }
}

__global__ __launch_bounds__(kBlock) void k_x_scan_chunks(ExecArgs a) {
  __shared__ uint32_t s_q, s_last;
  const uint64_t NC = a.st->nc;
  const uint32_t nq = (uint32_t)a.st->nq;
  for (uint64_t u = blockIdx.x; u < NC; u += gridDim.x) {
    __syncthreads();
    const uint32_t q = block_find(a.chunk_base, nq, u, &s_q);
    const uint32_t nch = *qs_field(a, q, offsetof(QState, nchunks));
    const StepParams p = a.items[q].p;
    scan_chunk(p, u - a.chunk_base[q], nch);
    if (nch > 1) {
      // thread 0 wrote this chunk's sums; the block that counts the query's
      // last chunk scans all of its chunk sums (release before the count,
      // acquire after it)
      __syncthreads();
      if (threadIdx.x == 0) {
        __threadfence();
        s_last = atomicAdd(qs_field(a, q, offsetof(QState, chunks_done)), 1u) + 1 == nch;
      }
      __syncthreads();
      if (s_last) {
        __threadfence();
        scan_top(p, nch);
      }
    }
  }
}

// One wave per tile, waves interleaved over the tile space (the tiles are
// short and alike, so a static deal balances).
template <int TR>
__global__ __launch_bounds__(kBlock) void k_x_finalize(ExecArgs a) {
  const uint32_t nq = (uint32_t)a.st->nq;
  const uint64_t NT = a.st->nt, W = (uint64_t)gridDim.x * (kBlock / kWave);
  uint32_t q = 0;
  uint64_t lo = 1, hi = 0;
  for (uint64_t t = blockIdx.x * (uint64_t)(kBlock / kWave) + threadIdx.x / kWave; t < NT; t += W) {
    if (t < lo || t >= hi) {  // wave-uniform
      q = wave_find(a.tile_base, nq, t);
      lo = a.tile_base[q];
      hi = a.tile_base[q + 1];
    }
    const StepParams p = a.items[q].p;
    finalize_tile<TR>(p, t - lo);
  }
}

// ---- expand plan: outputs, merge-path splits, retirement of finished queries ----

struct ExpandNeed {
  bool active, k2u, fin;  // fin: the query's last step
  bool bad;               // output beyond the row limit
  uint64_t table, scratch, nx, rows, cols;
};

template <int NV>
__device__ __forceinline__ ExpandNeed expand_need(const QState &s, const StepParams &p, const uint64_t (&tot)[4],
                                                  uint32_t level) {
  ExpandNeed n{false, false, false, false, 0, 0, 0, 0, 0};
  if (!level_active(s, level)) return n;
  n.active = true;
  n.fin = level + 1 == s.nsteps;
  n.k2u = p.mode == kModeK2U;
  n.rows = tot[0];
  if (n.k2u) {
    n.cols = p.c + 1;
    if (n.rows >= kRowLimit) {
      n.bad = true;
      return n;
    }
    const uint64_t nz = tot[2];
    n.table = al(n.rows * n.cols * 4);
    n.nx = n.rows ? (nz + n.rows + NV - 1) / NV : 0;
    n.scratch = n.nx ? al((n.nx + 1) * 8) : 0;
  } else {
    n.cols = p.c;
  }
  return n;
}

// Commits level `level`'s results; gives the stack tops after it.
template <int NV>
__device__ __forceinline__ void expand_phase(const ExecArgs &a, uint32_t level, uint32_t nq, uint64_t &lo_top,
                                             uint64_t &hi_low, PlanLds &pl) {
  __shared__ uint64_t s_tmp[3 * (kPlanBlock / kWave)];
  uint64_t *lw = pl.w[threadIdx.x / kWave];
  __shared__ unsigned long long s_ut, s_us, s_rb, s_over, s_lo, s_hi;
  ExecState *st = a.st;
  const uint64_t w0 = st->w0, w1 = st->w1, ou = st->out_used, su = st->scr_used;
  const uint64_t room = (w1 - w0) - ou - su;
  const uint64_t rb0 = st->res_bytes;
  const uint64_t lo_prev = st->lo_top, hi_prev = st->hi_low, peak_prev = st->peak, over_prev = st->res_over;
  if (threadIdx.x == 0) {
    s_ut = s_us = s_rb = s_over = 0;
  }
  uint64_t carry_t = 0, carry_s = 0, carry_p = 0, carry_x = 0, carry_r = 0, carry_u = 0;
  for (uint32_t c0 = 0; c0 < nq; c0 += kPlanBlock) {
    const uint32_t q = c0 + threadIdx.x;
    const uint32_t qc = q < nq ? q : nq - 1;
    QState sv = qs_get(a, qc);
    const StepParams p = wave_get<StepParams>(a.items, offsetof(Item, p), q - lane_id(), nq, lw);
    const uint64_t tot[4] = {a.qtot[4 * (uint64_t)qc], a.qtot[4 * (uint64_t)qc + 1], a.qtot[4 * (uint64_t)qc + 2], 0};
    const ExpandNeed n = q < nq ? expand_need<NV>(sv, p, tot, level) : ExpandNeed{false, false, false, false, 0, 0, 0, 0, 0};
    uint64_t sz[2] = {n.table, n.scratch}, szt[2];
    block_excl_scan_n<2, kPlanBlock, true>(sz, s_tmp, szt);
    const uint64_t ot = carry_t + sz[0], os = carry_s + sz[1];
    // arena fit; partition entries, expand tiles, result bytes of finishing queries
    const bool fit = n.active && !n.bad && ot + n.table + os + n.scratch <= room;
    const uint64_t rb = fit && n.fin ? al(n.rows * n.cols * 4) : 0;
    uint64_t cnt[3] = {fit && n.nx ? n.nx + 1 : 0, fit ? n.nx : 0, rb}, cntt[3];
    block_excl_scan_n<3, kPlanBlock, true>(cnt, s_tmp, cntt);
    const uint64_t orb = carry_r + cnt[2];
    // host-buffer fit (again a prefix) and copy units
    const bool host_fit = rb0 + orb + rb <= a.hcap;
    uint64_t nu[1] = {fit && n.fin && host_fit ? (rb + kCopyUnit - 1) / kCopyUnit : 0}, nut[1];
    block_excl_scan_n<1, kPlanBlock, true>(nu, s_tmp, nut);
    const uint64_t pbase = carry_p + cnt[0], xbase = carry_x + cnt[1], ubase = carry_u + nu[0];
    if (q < nq) {
      a.part_base[q] = pbase;
      a.xtile_base[q] = xbase;
      a.unit_base[q] = ubase;
    }
    ExpandParams x{};
    if (n.active) {
      if (n.bad) {
        sv.status = WKG_ERANGE;
      } else if (!fit) {
        sv.status = WKG_ENOMEM;
      } else {
        if (n.k2u) {
          x.pid = p.pid;
          x.T = p.T;
          x.c = p.c;
          x.R = tot[2];
          x.cinfo = p.cinfo;
          x.cexcl = p.cexcl;
          x.part = (const uint64_t *)scr_at(a, level, w0, w1, su + os, n.scratch);
          x.E = n.rows;
          x.O = (int32_t *)out_at(a, level, w0, w1, ou + ot, n.table);
          sv.T = x.O;
          sv.expanded += x.E;
          atomicMax(&s_ut, (unsigned long long)(ot + n.table));
          atomicMax(&s_us, (unsigned long long)(os + n.scratch));
        } else {  // filters: the kept rows were compacted into p.O
          sv.T = p.O;
          sv.expanded += tot[1];
        }
        sv.rows = n.rows;
        sv.cols = (uint32_t)n.cols;
        if (n.fin) {
          if (host_fit) {
            sv.out_off = (rb0 + orb) / 4;
            atomicMax(&s_rb, (unsigned long long)(orb + rb));
          } else {
            sv.status = WKG_ENOMEM;
            atomicAdd(&s_over, (unsigned long long)rb);
          }
        }
      }
    }
    wave_put(a.items, offsetof(Item, x), q - lane_id(), nq, x, n.active && fit && n.k2u, lw);
    if (n.active) qs_put(a, q, sv);
    carry_t += szt[0];
    carry_s += szt[1];
    carry_p += cntt[0];
    carry_x += cntt[1];
    carry_r += cntt[2];
    carry_u += nut[0];
  }
  lds_barrier();
  if (threadIdx.x == kPlanBlock - 1) {
    a.part_base[nq] = carry_p;
    a.xtile_base[nq] = carry_x;
    a.unit_base[nq] = carry_u;
    st->np = carry_p;
    st->nx = carry_x;
    st->nu = carry_u;
    const uint64_t out = ou + s_ut, scr = su + s_us;
    uint64_t lo = lo_prev, hi = hi_prev;
    if (level & 1) hi = w1 - out;
    else lo = w0 + out;
    st->lo_top = s_lo = lo;
    st->hi_low = s_hi = hi;
    const uint64_t in_use = a.cap - (w1 - w0) + out + scr;
    st->peak = in_use > peak_prev ? in_use : peak_prev;
    st->res_bytes = rb0 + s_rb;
    st->res_over = over_prev + s_over;
  }
  if (threadIdx.x < kSubQueues) st->queue[1][threadIdx.x][0] = 0;
  lds_barrier();
  lo_top = s_lo;
  hi_low = s_hi;
}

// The one-block plan between two levels: level `level`'s expand plan (none
// for level -1, which starts the batch instead), then the plan of level+1
// when it exists.
template <int TR, int NV>
__global__ __launch_bounds__(kPlanBlock) void k_x_plan(ExecArgs a, int level, int levels) {
  __shared__ PlanLds s_pl;
  uint32_t nq;
  uint64_t lo_top = 0, hi_low = a.cap;
  if (level < 0) {
    nq = begin_phase(a);
  } else {
    nq = (uint32_t)a.st->nq;
    expand_phase<NV>(a, (uint32_t)level, nq, lo_top, hi_low, s_pl);
  }
  // (each thread reads back only the states it stored itself)
  if (level + 1 < levels) level_phase<TR>(a, (uint32_t)(level + 1), nq, lo_top, hi_low, s_pl);
}

template <int NV>
__global__ __launch_bounds__(kBlock) void k_x_partition(ExecArgs a) {
  const uint64_t NP = a.st->np;
  const uint32_t nq = (uint32_t)a.st->nq;
  for (uint64_t k = blockIdx.x * (uint64_t)kBlock + threadIdx.x; k < NP; k += (uint64_t)gridDim.x * kBlock) {
    const uint32_t q = thread_find(a.part_base, nq, k);
    const ExpandParams &x = a.items[q].x;
    const uint64_t j = k - a.part_base[q];
    ((uint64_t *)x.part)[j] = merge_split<NV>(x.cexcl, x.R, x.E, j);
  }
}

template <int VT, bool FRAMED, bool QUEUE>
__global__ __launch_bounds__(kBlock) void k_x_expand(ExecArgs a) {
  const uint32_t nq = (uint32_t)a.st->nq;
  uint64_t t, lt;
  uint32_t q;
  if constexpr (QUEUE) {
    __shared__ QueueLds s_wq;
    BlockQueue wq(s_wq, a.st->queue[1], a.st->nx);
    while (wq.next(a.xtile_base, nq, t, q, lt)) {
      const ExpandParams x = a.items[q].x;
      expand_tile<VT, FRAMED>(a.g, x, lt);
    }
  } else {
    __shared__ uint32_t s_q;
    StaticDeal wd(s_q, a.st->nx);
    while (wd.next(a.xtile_base, nq, t, q, lt)) {
      const ExpandParams x = a.items[q].x;
      expand_tile<VT, FRAMED>(a.g, x, lt);
    }
  }
}

// Copies the tables of queries that finished at this level into the host
// result buffer (device-visible pinned memory). Tables and result slots are
// 256-byte aligned and padded, so 16-byte chunks stay inside both.
__global__ __launch_bounds__(kBlock) void k_x_retire(ExecArgs a) {
  __shared__ uint32_t s_q;
  const uint32_t nq = (uint32_t)a.st->nq;
  const uint64_t NU = a.st->nu;
  for (uint64_t u = blockIdx.x; u < NU; u += gridDim.x) {
    __syncthreads();
    const uint32_t q = block_find(a.unit_base, nq, u, &s_q);
    const QState s = qs_get(a, q);
    const uint64_t bytes = s.rows * s.cols * 4;
    const uint64_t b0 = (u - a.unit_base[q]) * kCopyUnit;
    if (b0 >= bytes) continue;  // (units are sized from the same bytes)
    const uint64_t b1 = b0 + kCopyUnit < bytes ? b0 + kCopyUnit : bytes;
    const uint4 *src = (const uint4 *)((const char *)s.T + b0);
    uint4 *dst = (uint4 *)(a.hres + s.out_off * 4 + b0);
    const uint64_t n16 = (b1 - b0 + 15) / 16;
    for (uint64_t i = threadIdx.x; i < n16; i += kBlock) dst[i] = src[i];
  }
}

__global__ __launch_bounds__(kBlock) void k_x_records(ExecArgs a) {
  const uint32_t nq = (uint32_t)a.st->nq;
  for (uint32_t q = threadIdx.x; q < nq; q += kBlock) {
    const QState s = qs_get(a, q);
    wkg_batch_result r{};
    r.status = s.status;
    r.expanded = s.expanded;
    r.cols = (int32_t)s.cols;
    if (s.status == WKG_OK) {
      r.offset = s.out_off;
      r.rows = s.rows;
    }
    a.res[q] = r;
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// host side

struct wkg_batch {
  wkg_ctx *ctx = nullptr;
  int Q = 0, P = 0;
  uint64_t cap = 0;
  hipStream_t s = nullptr;
  bool own_stream = false;
  int ncu = 0;
  char *arena = nullptr;
  char *dctl = nullptr;          // device control block
  size_t ctl_bytes = 0, down_bytes = 0, up_bytes = 0;
  char *d_up = nullptr;          // device copy of the upload block
  char *h_up = nullptr;          // pinned: nq | nsteps[Q] | steps[Q*P]
  char *h_down = nullptr;        // pinned: ExecState | results[Q]
  char *h_res = nullptr;         // pinned, device-visible: result tables
  uint64_t h_res_cap = 0;
  uint64_t grow_to = 0;          // result buffer size wanted at the next launch
  uint64_t last_peak = 0, last_res_bytes = 0;
  ExecArgs args{};
  wkg_batch_tuning tune{};
  std::map<int, hipGraphExec_t> graphs;  // by level count
  bool launched = false;
  int launched_nq = 0;
  std::vector<int> pids;

  void clear_graphs() {
    for (auto &kv : graphs) (void)hipGraphExecDestroy(kv.second);
    graphs.clear();
  }
  int set_result_buffer(uint64_t bytes);
  int enqueue(int levels);
  int graph_for(int levels, hipGraphExec_t *out);
};

namespace {

constexpr uint64_t kDefaultResultBuffer = 64ull << 20;

wkg_batch_tuning normalized(const wkg_batch_tuning &t) {
  wkg_batch_tuning r = t;
  if (r.probe_wg_per_cu <= 0) r.probe_wg_per_cu = 2;
  if (r.expand_wg_per_cu <= 0) r.expand_wg_per_cu = 2;
  if (r.aux_wg_per_cu <= 0) r.aux_wg_per_cu = 1;
  if (r.probe_tile_rows != 1024) r.probe_tile_rows = 2048;
  if (r.expand_tile_items != 1024 && r.expand_tile_items != 2048) r.expand_tile_items = 512;
  r.schedule = r.schedule == 1 ? 1 : 0;
  return r;
}

template <int TR, int VT, bool SORTED, bool FRAMED, bool QUEUE>
void enqueue_levels(const ExecArgs &a, hipStream_t s, int levels, unsigned gp, unsigned gx, unsigned ga,
                    unsigned gf) {
  constexpr int NV = kBlock * VT;
  k_x_plan<TR, NV><<<1, kPlanBlock, 0, s>>>(a, -1, levels);
  for (int l = 0; l < levels; l++) {
    k_x_probe<TR, SORTED, FRAMED, QUEUE><<<gp, kBlock, 0, s>>>(a);
    k_x_scan_chunks<<<ga, kBlock, 0, s>>>(a);
    k_x_finalize<TR><<<gf, kBlock, 0, s>>>(a);
    k_x_plan<TR, NV><<<1, kPlanBlock, 0, s>>>(a, l, levels);
    k_x_partition<NV><<<ga, kBlock, 0, s>>>(a);
    k_x_expand<VT, FRAMED, QUEUE><<<gx, kBlock, 0, s>>>(a);
    k_x_retire<<<ga, kBlock, 0, s>>>(a);
  }
}

template <int TR, int VT, bool QUEUE>
void enqueue_sched(const ExecArgs &a, hipStream_t s, int levels, bool sorted, bool framed, unsigned gp, unsigned gx,
                   unsigned ga, unsigned gf) {
  if (framed) {
    if (sorted) enqueue_levels<TR, VT, true, true, QUEUE>(a, s, levels, gp, gx, ga, gf);
    else enqueue_levels<TR, VT, false, true, QUEUE>(a, s, levels, gp, gx, ga, gf);
  } else {
    if (sorted) enqueue_levels<TR, VT, true, false, QUEUE>(a, s, levels, gp, gx, ga, gf);
    else enqueue_levels<TR, VT, false, false, QUEUE>(a, s, levels, gp, gx, ga, gf);
  }
}

template <int TR, int VT>
void enqueue_shape(const ExecArgs &a, hipStream_t s, int levels, bool sorted, bool framed, bool queue, unsigned gp,
                   unsigned gx, unsigned ga, unsigned gf) {
  if (queue) enqueue_sched<TR, VT, true>(a, s, levels, sorted, framed, gp, gx, ga, gf);
  else enqueue_sched<TR, VT, false>(a, s, levels, sorted, framed, gp, gx, ga, gf);
}

}  // namespace

// (Re)allocates the host result buffer; the captured sequences hold its
// address, so they are dropped.
int wkg_batch::set_result_buffer(uint64_t bytes) {
  bytes = al(std::max<uint64_t>(bytes, 4096));
  if (h_res) (void)hipHostFree(h_res);
  h_res = nullptr;
  h_res_cap = 0;
  args.hres = nullptr;
  args.hcap = 0;
  clear_graphs();
  hipError_t e = hipHostMalloc((void **)&h_res, bytes, hipHostMallocMapped);
  void *d = nullptr;
  if (e == hipSuccess) e = hipHostGetDevicePointer(&d, h_res, 0);
  if (e != hipSuccess) {
    if (h_res) (void)hipHostFree(h_res);
    h_res = nullptr;
    return ctx->fail(WKG_ENOMEM, "pinned result buffer of %llu bytes: %s", (unsigned long long)bytes,
                     hipGetErrorString(e));
  }
  h_res_cap = bytes;
  args.hres = (char *)d;
  args.hcap = bytes;
  return WKG_OK;
}

int wkg_batch::enqueue(int levels) {
  const wkg_batch_tuning t = normalized(tune);
  const unsigned gp = (unsigned)(ncu * t.probe_wg_per_cu), gx = (unsigned)(ncu * t.expand_wg_per_cu),
                 ga = (unsigned)(ncu * t.aux_wg_per_cu), gf = (unsigned)(ncu * 8);
  hipError_t e = hipMemcpyAsync(d_up, h_up, up_bytes, hipMemcpyHostToDevice, s);
  if (e != hipSuccess) return ctx->fail(WKG_EDEVICE, "plan upload failed: %s", hipGetErrorString(e));
  const bool sorted = ctx->sorted, framed = ctx->framed, queue = t.schedule == 0;
  if (t.probe_tile_rows == 1024) {
    if (t.expand_tile_items == 512) enqueue_shape<1024, 2>(args, s, levels, sorted, framed, queue, gp, gx, ga, gf);
    else if (t.expand_tile_items == 1024) enqueue_shape<1024, 4>(args, s, levels, sorted, framed, queue, gp, gx, ga, gf);
    else enqueue_shape<1024, 8>(args, s, levels, sorted, framed, queue, gp, gx, ga, gf);
  } else {
    if (t.expand_tile_items == 512) enqueue_shape<2048, 2>(args, s, levels, sorted, framed, queue, gp, gx, ga, gf);
    else if (t.expand_tile_items == 1024) enqueue_shape<2048, 4>(args, s, levels, sorted, framed, queue, gp, gx, ga, gf);
    else enqueue_shape<2048, 8>(args, s, levels, sorted, framed, queue, gp, gx, ga, gf);
  }
  k_x_records<<<1, kBlock, 0, s>>>(args);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(h_down, dctl, down_bytes, hipMemcpyDeviceToHost, s);
  if (e != hipSuccess) return ctx->fail(WKG_EDEVICE, "batch launch failed: %s", hipGetErrorString(e));
  return WKG_OK;
}

int wkg_batch::graph_for(int levels, hipGraphExec_t *out) {
  auto it = graphs.find(levels);
  if (it != graphs.end()) {
    *out = it->second;
    return WKG_OK;
  }
  hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return ctx->fail(WKG_EDEVICE, "stream capture failed: %s", hipGetErrorString(e));
  int rc = enqueue(levels);
  hipGraph_t g = nullptr;
  e = hipStreamEndCapture(s, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return ctx->fail(WKG_EDEVICE, "stream capture failed: %s", hipGetErrorString(e));
  hipGraphExec_t x = nullptr;
  e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return ctx->fail(WKG_EDEVICE, "graph instantiation failed: %s", hipGetErrorString(e));
  graphs[levels] = x;
  *out = x;
  return WKG_OK;
}

extern "C" {

int wkg_batch_create(wkg_ctx *ctx, int max_queries, int max_patterns, uint64_t arena_bytes, void *stream,
                     wkg_batch **out) {
  if (!ctx || !out || max_queries < 1 || max_patterns < 1 || max_patterns > 64 || arena_bytes < (1u << 20))
    return WKG_EINVAL;
  if (max_queries > (1 << 20)) return ctx->fail(WKG_ERANGE, "a batch holds at most 2^20 queries");
  if (!ctx->loaded) return ctx->fail(WKG_ESTATE, "no graph loaded");
  *out = nullptr;
  wkg_batch *b = new wkg_batch;
  b->ctx = ctx;
  b->Q = max_queries;
  b->P = max_patterns;
  b->cap = al(arena_bytes) - kAlign;
  hipError_t e = hipSuccess;
  if (stream) {
    b->s = (hipStream_t)stream;
  } else {
    e = hipStreamCreateWithFlags(&b->s, hipStreamNonBlocking);
    b->own_stream = e == hipSuccess;
  }
  int dev_cu = 0;
  if (e == hipSuccess) e = hipDeviceGetAttribute(&dev_cu, hipDeviceAttributeMultiprocessorCount, ctx->device);
  b->ncu = dev_cu > 0 ? dev_cu : 256;
  const uint64_t Q = (uint64_t)max_queries, P = (uint64_t)max_patterns;
  // control block: ExecState | results[Q] | query states | items[Q] | totals[Q] | 5 x base[Q+1] | upload
  const uint64_t o_res = al(sizeof(ExecState));
  const uint64_t o_qs = al(o_res + Q * sizeof(wkg_batch_result));
  const uint64_t o_items = al(o_qs + Q * sizeof(QState));
  const uint64_t o_tot = al(o_items + Q * sizeof(Item));
  const uint64_t o_base = al(o_tot + Q * 32);
  const uint64_t base_bytes = al((Q + 1) * 8);
  const uint64_t o_up = o_base + 5 * base_bytes;
  b->up_bytes = al(8 + Q * 4) + Q * P * sizeof(wkg_plan_step);
  b->ctl_bytes = o_up + b->up_bytes;
  b->down_bytes = o_res + Q * sizeof(wkg_batch_result);
  auto bail = [&](int rc) {
    wkg_batch_destroy(b);
    return rc;
  };
  if (e != hipSuccess) return bail(ctx->fail(WKG_EDEVICE, "batch stream setup failed: %s", hipGetErrorString(e)));
  b->dctl = (char *)ctx->pool->alloc(b->ctl_bytes, b->s, wkg_ctx::err_slot());
  if (!b->dctl) return bail(WKG_ENOMEM);
  b->arena = (char *)ctx->pool->alloc(b->cap, b->s, wkg_ctx::err_slot());
  if (!b->arena) return bail(WKG_ENOMEM);
  e = hipHostMalloc((void **)&b->h_up, b->up_bytes, hipHostMallocDefault);
  if (e == hipSuccess) e = hipHostMalloc((void **)&b->h_down, b->down_bytes, hipHostMallocDefault);
  if (e != hipSuccess) {
    b->h_up = b->h_down = nullptr;
    return bail(ctx->fail(WKG_ENOMEM, "pinned host buffers: %s", hipGetErrorString(e)));
  }
  memset(b->h_up, 0, b->up_bytes);
  b->d_up = b->dctl + o_up;
  ExecArgs &a = b->args;
  a.g = ctx->view();
  a.st = (ExecState *)b->dctl;
  a.res = (wkg_batch_result *)(b->dctl + o_res);
  a.qsw = (uint64_t *)(b->dctl + o_qs);
  a.items = (Item *)(b->dctl + o_items);
  a.qtot = (uint64_t *)(b->dctl + o_tot);
  a.tile_base = (uint64_t *)(b->dctl + o_base);
  a.chunk_base = (uint64_t *)(b->dctl + o_base + base_bytes);
  a.part_base = (uint64_t *)(b->dctl + o_base + 2 * base_bytes);
  a.xtile_base = (uint64_t *)(b->dctl + o_base + 3 * base_bytes);
  a.unit_base = (uint64_t *)(b->dctl + o_base + 4 * base_bytes);
  a.up_nq = (const uint64_t *)b->d_up;
  a.up_nsteps = (const uint32_t *)(b->d_up + 8);
  a.up_steps = (const wkg_plan_step *)(b->d_up + al(8 + Q * 4));
  a.arena = b->arena;
  a.cap = b->cap;
  a.Q = (uint32_t)Q;
  a.P = (uint32_t)P;
  if (int rc = b->set_result_buffer(kDefaultResultBuffer)) return bail(rc);
  *out = b;
  return WKG_OK;
}

int wkg_batch_destroy(wkg_batch *b) {
  if (!b) return WKG_OK;
  if (b->s) (void)hipStreamSynchronize(b->s);
  b->clear_graphs();
  if (b->arena) b->ctx->pool->release(b->arena, b->s);
  if (b->dctl) b->ctx->pool->release(b->dctl, b->s);
  if (b->h_up) (void)hipHostFree(b->h_up);
  if (b->h_down) (void)hipHostFree(b->h_down);
  if (b->h_res) (void)hipHostFree(b->h_res);
  if (b->own_stream) {
    // pooled blocks tagged with this stream go back to HIP before it is destroyed
    b->ctx->pool->trim();
    (void)hipStreamDestroy(b->s);
  }
  delete b;
  return WKG_OK;
}

int wkg_batch_set_tuning(wkg_batch *b, const wkg_batch_tuning *t) {
  if (!b || !t) return WKG_EINVAL;
  if (b->launched) return b->ctx->fail(WKG_ESTATE, "a batch is in flight");
  const wkg_batch_tuning n = normalized(*t);
  if (memcmp(&n, &b->tune, sizeof n) != 0) {
    b->clear_graphs();
    b->tune = n;
  }
  return WKG_OK;
}

int wkg_batch_get_tuning(const wkg_batch *b, wkg_batch_tuning *t) {
  if (!b || !t) return WKG_EINVAL;
  *t = normalized(b->tune);
  return WKG_OK;
}

int wkg_batch_reserve_results(wkg_batch *b, uint64_t bytes) {
  if (!b) return WKG_EINVAL;
  if (b->launched) return b->ctx->fail(WKG_ESTATE, "a batch is in flight");
  b->grow_to = 0;
  return b->set_result_buffer(bytes);
}

int wkg_batch_stats(const wkg_batch *b, uint64_t *arena_peak, uint64_t *result_bytes) {
  if (!b) return WKG_EINVAL;
  if (arena_peak) *arena_peak = b->last_peak;
  if (result_bytes) *result_bytes = b->last_res_bytes;
  return WKG_OK;
}

int wkg_batch_launch(wkg_batch *b, int nq, const wkg_plan_step *steps, const int32_t *step_off) {
  if (!b) return WKG_EINVAL;
  wkg_ctx *ctx = b->ctx;
  if (b->launched) return ctx->fail(WKG_ESTATE, "a batch is already in flight");
  if (nq < 0 || nq > b->Q || (nq && (!steps || !step_off)))
    return ctx->fail(WKG_EINVAL, "batch of %d queries (at most %d)", nq, b->Q);
  if (step_off && step_off[0] != 0) return ctx->fail(WKG_EINVAL, "step_off[0] must be 0");
  if (!ctx->loaded) return ctx->fail(WKG_ESTATE, "no graph loaded");
  {  // the captured sequences hold the graph view; a reloaded graph invalidates them
    const GraphView v = ctx->view(), &o = b->args.g;
    if (v.slots != o.slots || v.edges != o.edges || v.meta != o.meta || v.vframes != o.vframes ||
        v.eframes != o.eframes || v.vshift != o.vshift || v.eshift != o.eshift || v.num_metas != o.num_metas) {
      b->clear_graphs();
      b->args.g = v;
    }
  }
  // validate every plan (column bookkeeping as the steps will run) and stage it
  uint32_t *h_nsteps = (uint32_t *)(b->h_up + 8);
  wkg_plan_step *h_steps = (wkg_plan_step *)(b->h_up + al(8 + (uint64_t)b->Q * 4));
  int levels = 0;
  b->pids.clear();
  for (int q = 0; q < nq; q++) {
    const int n = step_off[q + 1] - step_off[q];
    if (n < 1 || n > b->P) return ctx->fail(WKG_EINVAL, "query %d has %d steps (1..%d)", q, n, b->P);
    levels = std::max(levels, n);
    int cols = 0;
    for (int l = 0; l < n; l++) {
      const wkg_plan_step &d = steps[step_off[q] + l];
      if (int rc = ctx->check_pattern(d.pid, d.dir)) return ctx->fail(rc, "query %d step %d: %s", q, l, wkg_ctx::err_slot().c_str());
      switch (d.kind) {
        case WKG_STEP_I2U:
        case WKG_STEP_C2U:
          if (l != 0) return ctx->fail(WKG_EINVAL, "query %d step %d: index/constant start after step 0", q, l);
          if (d.kind == WKG_STEP_I2U && !is_tpid(d.pid))
            return ctx->fail(WKG_EINVAL, "query %d: index id %d is not a predicate/type id", q, d.pid);
          if (d.kind == WKG_STEP_C2U && (d.vid < 0 || (uint64_t)d.vid > kMaxVid))
            return ctx->fail(WKG_EINVAL, "query %d: constant out of range", q);
          cols = 1;
          break;
        case WKG_STEP_K2U:
        case WKG_STEP_K2K:
        case WKG_STEP_K2C:
          if (l == 0) return ctx->fail(WKG_EINVAL, "query %d: step 0 needs an index or constant start", q);
          if (d.start_col < 0 || d.start_col >= cols)
            return ctx->fail(WKG_EINVAL, "query %d step %d: start column %d outside %d columns", q, l, d.start_col,
                             cols);
          if (d.kind == WKG_STEP_K2K && (d.end_col < 0 || d.end_col >= cols))
            return ctx->fail(WKG_EINVAL, "query %d step %d: end column %d outside %d columns", q, l, d.end_col, cols);
          if (d.kind == WKG_STEP_K2U) cols++;
          if (cols > kBlock) return ctx->fail(WKG_ERANGE, "query %d: tables are limited to %d columns", q, kBlock);
          break;
        default:
          return ctx->fail(WKG_EINVAL, "query %d step %d: unknown kind %d", q, l, d.kind);
      }
      if (std::find(b->pids.begin(), b->pids.end(), d.pid) == b->pids.end()) b->pids.push_back(d.pid);
    }
  }
  if (ctx->framed)  // every predicate of the batch must be resident for the whole run
    for (int p : b->pids)
      if (int rc = ctx->ensure_pred(p, b->s, &b->pids)) return rc;
  *(uint64_t *)b->h_up = (uint64_t)nq;
  for (int q = 0; q < b->Q; q++) h_nsteps[q] = q < nq ? (uint32_t)(step_off[q + 1] - step_off[q]) : 0u;
  for (int q = 0; q < nq; q++)
    memcpy(&h_steps[(uint64_t)q * b->P], &steps[step_off[q]], sizeof(wkg_plan_step) * (step_off[q + 1] - step_off[q]));
  if (levels == 0) levels = 1;
  if (b->grow_to > b->h_res_cap)
    if (int rc = b->set_result_buffer(b->grow_to)) return rc;
  b->grow_to = 0;
  hipGraphExec_t x;
  if (int rc = b->graph_for(levels, &x)) return rc;
  ProfScope ps(ctx, b->s, "batch_graph", 0.0);
  hipError_t e = hipGraphLaunch(x, b->s);
  if (e != hipSuccess) return ctx->fail(WKG_EDEVICE, "graph launch failed: %s", hipGetErrorString(e));
  b->launched = true;
  b->launched_nq = nq;
  return WKG_OK;
}

int wkg_batch_busy(wkg_batch *b) {
  if (!b || !b->launched) return 0;
  return hipStreamQuery(b->s) == hipErrorNotReady ? 1 : 0;
}

int wkg_batch_collect(wkg_batch *b, wkg_batch_result *res, const int32_t **cells, uint64_t *ncells) {
  if (!b || (!res && b->launched_nq)) return WKG_EINVAL;
  wkg_ctx *ctx = b->ctx;
  if (!b->launched) return ctx->fail(WKG_ESTATE, "no batch in flight");
  if (cells) *cells = nullptr;
  if (ncells) *ncells = 0;
  b->launched = false;
  hipError_t e = hipStreamSynchronize(b->s);
  if (e != hipSuccess) return ctx->fail(WKG_EDEVICE, "batch failed: %s", hipGetErrorString(e));
  const ExecState *st = (const ExecState *)b->h_down;
  const wkg_batch_result *r = (const wkg_batch_result *)(b->h_down + al(sizeof(ExecState)));
  if (b->launched_nq) memcpy(res, r, sizeof(wkg_batch_result) * b->launched_nq);
  b->last_peak = st->peak;
  b->last_res_bytes = st->res_bytes;
  if (cells) *cells = (const int32_t *)b->h_res;
  if (ncells) *ncells = st->res_bytes / 4;
  // some results did not fit: grow the buffer at the next launch (this
  // batch's results stay readable in the current one until then)
  if (st->res_over) b->grow_to = std::max(b->grow_to, (st->res_bytes + st->res_over) * 3 / 2);
  return WKG_OK;
}

int wkg_batch_run(wkg_batch *b, int nq, const wkg_plan_step *steps, const int32_t *step_off, wkg_batch_result *res,
                  const int32_t **cells, uint64_t *ncells) {
  if (int rc = wkg_batch_launch(b, nq, steps, step_off)) return rc;
  return wkg_batch_collect(b, res, cells, ncells);
}

int wkg_batch_autotune(wkg_batch *b, int nq, const wkg_plan_step *steps, const int32_t *step_off, int reps,
                       wkg_batch_tuning *best, double *best_ms) {
  if (!b || reps < 1) return WKG_EINVAL;
  std::vector<wkg_batch_result> res((size_t)std::max(nq, 1));
  std::vector<wkg_batch_tuning> cand;
  for (int tr : {1024, 2048})
    for (int xi : {512, 1024, 2048})
      for (int aw : {1, 2}) {
        // work queue: more workgroups than fit at once are fine (late ones
        // find the queue drained); static deal: at most what fits (four
        // probe workgroups per CU), or the late ones would run their share
        // after the others are done
        for (int pw : {1, 2, 4})
          for (int xw : {2, 4, 8}) cand.push_back(wkg_batch_tuning{pw, xw, aw, tr, xi, 0});
        for (int pw : {2, 3, 4})
          for (int xw : {4, 8}) cand.push_back(wkg_batch_tuning{pw, xw, aw, tr, xi, 1});
      }
  double best_t = 1e30;
  wkg_batch_tuning best_c = normalized(b->tune);
  for (const wkg_batch_tuning &c : cand) {
    if (int rc = wkg_batch_set_tuning(b, &c)) return rc;
    std::vector<double> ts;
    for (int i = 0; i <= reps; i++) {  // the first run captures the graph
      const auto t0 = std::chrono::steady_clock::now();
      int rc = wkg_batch_run(b, nq, steps, step_off, res.data(), nullptr, nullptr);
      if (rc) return rc;
      if (i) ts.push_back(std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
    }
    std::sort(ts.begin(), ts.end());
    const double med = ts[ts.size() / 2];
    if (med < best_t) {
      best_t = med;
      best_c = normalized(c);
    }
  }
  if (int rc = wkg_batch_set_tuning(b, &best_c)) return rc;
  if (best) *best = best_c;
  if (best_ms) *best_ms = best_t;
  return WKG_OK;
}

}  // extern "C"
