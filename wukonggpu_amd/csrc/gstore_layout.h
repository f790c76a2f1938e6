// Layout of a gstore image from per-id counts: every decision the builder
// makes before it places a key (core/gstore.hpp:600-659, 957-978). Shared by
// the host builder (gstore.cpp) and the device builder (builder.hip), so the
// two produce the same image, slot for slot.
//
//   * main partitions proportional to each id's triple count
//     (core/gstore.hpp:618-651), at least one bucket for an id with triples;
//   * epochs in insert order: rdf:type (pid 1), then the type indexes, then
//     the predicates ascending; each epoch's edges are one contiguous segment
//     and its indirect buckets follow the previous epoch's.
// Host code, but chain_buckets, which the device builder's kernels use too.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "gstore_format.h"

namespace wkg {

// Per id, after duplicate triples are removed: (s,p,o) pairs and their
// distinct subjects under p (so), (o,p,s) pairs and distinct objects (os,
// predicates only), and type members (rdf:type triples with object t).
struct IdCounts {
  std::vector<uint64_t> so_cnt, os_cnt, nsubj, nobj, type_cnt;
};

// One insertion epoch: a contiguous run of keys sharing a main partition.
struct Epoch {
  uint32_t id;         // predicate or type id whose partition the keys use
  int kind;            // 0: rdf:type (s,1,OUT) keys, 1: type index, 2: predicate
  uint64_t edge_base;  // first edge of this epoch
  uint64_t nedges;
  uint64_t nkeys;
  uint64_t ext_count = 0;  // indirect buckets the epoch needs
  uint64_t ext_base = 0;   // its first indirect bucket (relative to num_buckets)
};

struct Layout {
  uint64_t num_buckets = 0;  // main-header buckets
  uint64_t num_keys = 0;
  uint64_t total_edges = 0;
  std::vector<wkg_pred_meta> metas;
  std::vector<Epoch> ep;
};

// Indirect buckets a bucket of `cnt` keys chains to (7 keys per bucket).
WKG_HD uint64_t chain_buckets(uint64_t cnt) { return cnt > 7 ? (cnt - 7 + 6) / 7 : 0; }

// Partition sizes, epochs and their edge segments. main_load is the target
// keys per data slot of the main region.
inline void plan_layout(const IdCounts &c, uint32_t num_metas, double main_load, Layout &L) {
  uint64_t num_edges_count = 0, num_keys = 0;
  std::vector<uint64_t> part_cnt(num_metas, 0);
  for (uint32_t p = 1; p < num_metas; p++) {
    part_cnt[p] = c.so_cnt[p] + c.os_cnt[p] + c.type_cnt[p];
    num_edges_count += part_cnt[p];
    num_keys += c.nsubj[p] + c.nobj[p];
    if (p >= 2 && c.so_cnt[p]) num_keys++;
    if (p >= 2 && c.os_cnt[p]) num_keys++;
    if (c.type_cnt[p]) num_keys++;
  }
  // rdf:type triples count once for pid 1 (so_cnt) and once for their type
  // (type_cnt), as core/gstore.hpp:619-626 does
  L.num_keys = num_keys;
  L.num_buckets = std::max<uint64_t>(1, (uint64_t)std::ceil((double)num_keys / (7.0 * main_load)));
  L.metas.assign(num_metas, wkg_pred_meta{});
  uint64_t main_cnt = 0;
  for (uint32_t p = 1; p < num_metas; p++) {
    uint64_t b = num_edges_count ? (uint64_t)((double)part_cnt[p] / (double)num_edges_count * (double)L.num_buckets)
                                 : 0;
    if (b == 0 && part_cnt[p] > 0) b = 1;
    L.metas[p].partition_sz = b;
    L.metas[p].main_hdr_start = main_cnt;
    L.metas[p].main_hdr_end = main_cnt + b;
    main_cnt += b;
  }
  // the "at least one bucket" rule can push the sum past num_buckets
  if (main_cnt > L.num_buckets) L.num_buckets = main_cnt;

  L.ep.clear();
  if (c.so_cnt[kTypeId]) L.ep.push_back(Epoch{(uint32_t)kTypeId, 0, 0, c.so_cnt[kTypeId], c.nsubj[kTypeId]});
  for (uint32_t p = 2; p < num_metas; p++)
    if (c.type_cnt[p]) L.ep.push_back(Epoch{p, 1, 0, c.type_cnt[p], 1});
  for (uint32_t p = 2; p < num_metas; p++)
    if (c.so_cnt[p] || c.os_cnt[p]) {
      const uint64_t nk = c.nsubj[p] + c.nobj[p] + (c.so_cnt[p] ? 1 : 0) + (c.os_cnt[p] ? 1 : 0);
      L.ep.push_back(Epoch{p, 2, 0, c.so_cnt[p] + c.os_cnt[p] + c.nsubj[p] + c.nobj[p], nk});
    }
  uint64_t total = 0;
  for (Epoch &e : L.ep) {
    e.edge_base = total;
    total += e.nedges;
  }
  L.total_edges = total;
}

// Once every epoch's ext_count is known: indirect bases in epoch order, and
// each id's indirect and edge ranges. Returns the indirect bucket total.
inline uint64_t place_epochs(Layout &L) {
  uint64_t ext_total = 0;
  for (Epoch &e : L.ep) {
    e.ext_base = ext_total;
    ext_total += e.ext_count;
  }
  for (const Epoch &e : L.ep) {
    wkg_pred_meta &m = L.metas[e.id];
    m.indrct_hdr_start = L.num_buckets + e.ext_base;
    m.indrct_hdr_end = L.num_buckets + e.ext_base + e.ext_count;
    m.edge_start = e.edge_base;
    m.edge_end = e.edge_base + e.nedges;
  }
  return ext_total;
}

}  // namespace wkg
