// Packed gstore formats shared by the host builder and the HIP kernels.
//
// The byte layout is fixed by the reference so that an image built by either
// side can be read by the other:
//   key  (ikey_t, core/gstore.hpp:117-149)  u64 = vid<<18 | pid<<1 | dir
//   ptr  (iptr_t, core/gstore.hpp:156-176)  u64 = size (bits 0-27) | off<<28
//   slot (vertex_t, core/gstore.hpp:179-182) = {key, ptr}, 16 bytes
//   bucket = 8 slots (128 bytes); slots 0..6 hold keys, slot 7 is the chain
//            link whose key.vid is the next bucket id (core/gstore.hpp:296-343)
//   edge (edge_t, core/gstore.hpp:185-187) = u32 vertex id
//   pred_meta_t (core/rdf_meta.hpp:31-54) = 8 x u64 per predicate/type id.
//
// Everything here is integer arithmetic; the header compiles under g++ (host
// builder) and hipcc (kernels).
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define WKG_HD __host__ __device__ __forceinline__
#else
#define WKG_HD static inline
#endif

namespace wkg {

enum : int { kDirBits = 1, kIdxBits = 17, kVidBits = 46 };
enum : int { kSizeBits = 28, kOffBits = 36 };
enum : int { kAssoc = 8 };               // slots per bucket (7 data + 1 link)
enum : int { kPredicateId = 0, kTypeId = 1 };
enum : int { kDirIn = 0, kDirOut = 1 };

constexpr uint64_t kSizeMask = (uint64_t(1) << kSizeBits) - 1;
constexpr uint64_t kMaxVid = (uint64_t(1) << kVidBits) - 1;
constexpr uint32_t kMaxIdx = (uint32_t(1) << kIdxBits) - 1;

// Predicate/type ids live in [2, 2^17); normal vertices start at 2^17
// (core/gstore.hpp:109, datagen/generate_data.cpp:88-98).
WKG_HD bool is_tpid(int64_t id) { return id > 1 && id < (int64_t(1) << kIdxBits); }

WKG_HD uint64_t make_key(uint64_t vid, uint64_t pid, uint64_t dir) {
  return (vid << (kIdxBits + kDirBits)) | (pid << kDirBits) | dir;
}
WKG_HD uint64_t key_vid(uint64_t k) { return k >> (kIdxBits + kDirBits); }
WKG_HD uint64_t key_pid(uint64_t k) { return (k >> kDirBits) & kMaxIdx; }
WKG_HD uint64_t key_dir(uint64_t k) { return k & 1; }

WKG_HD uint64_t make_ptr(uint64_t size, uint64_t off) { return size | (off << kSizeBits); }
WKG_HD uint64_t ptr_size(uint64_t p) { return p & kSizeMask; }
WKG_HD uint64_t ptr_off(uint64_t p) { return p >> kSizeBits; }

// Thomas Wang's 64-bit integer mix (hash64shift), applied to the packed key
// exactly as ikey_t::hash does (core/gstore.hpp:140-148 ->
// utils/mymath.hpp:62-71). The packed key value equals the "r" that ikey_t
// assembles from its bit fields.
WKG_HD uint64_t mix64(uint64_t x) {
  x = ~x + (x << 21);
  x ^= x >> 24;
  x = x + (x << 3) + (x << 8);
  x ^= x >> 14;
  x = x + (x << 2) + (x << 4);
  x ^= x >> 28;
  x += x << 31;
  return x;
}

// x mod d for d >= 1 using a precomputed m = floor((2^64-1)/d). The quotient
// estimate hi(x*m) is at most two below floor(x/d), so at most two
// corrections are needed.
WKG_HD uint64_t fastmod_magic(uint64_t d) { return ~uint64_t(0) / d; }
WKG_HD uint64_t fastmod(uint64_t x, uint64_t d, uint64_t m) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint64_t q = __umul64hi(x, m);
#else
  uint64_t q = (uint64_t)(((unsigned __int128)x * m) >> 64);
#endif
  uint64_t r = x - q * d;
  if (r >= d) r -= d;
  if (r >= d) r -= d;
  return r;
}

}  // namespace wkg

// Per predicate/type metadata, field for field the reference's pred_meta_t.
typedef struct wkg_pred_meta {
  uint64_t partition_sz;      // main-header buckets owned by this id
  uint64_t main_hdr_start;    // first main bucket (global bucket id)
  uint64_t main_hdr_end;
  uint64_t indrct_hdr_start;  // first indirect bucket (global bucket id)
  uint64_t indrct_hdr_end;
  uint64_t num_buckets;       // unused by the hot path (kept for layout)
  uint64_t edge_start;        // first edge entry of this id's segment
  uint64_t edge_end;
} wkg_pred_meta;
