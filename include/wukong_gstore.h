/*
 * Host gstore: triple sets, the synthetic RDF generator and the gstore image
 * builder (C ABI of libwkg_gstore.so).
 *
 * The image is the reference's host key/value store (core/gstore.hpp:190-1195)
 * laid out byte for byte: a slot array of 16-byte {key, ptr} pairs grouped in
 * 128-byte buckets with cluster chaining, one u32 edge array, and one
 * pred_meta_t per predicate/type id. The GPU library (wukong_gpu.h) loads it
 * with wkg_load_graph().
 *
 * Every call returns 0 on success and a negative WKS_E* code on failure; the
 * message of the last failure on the calling thread is wks_last_error().
 */
#ifndef WUKONG_GSTORE_H
#define WUKONG_GSTORE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WKS_OK 0
#define WKS_EINVAL (-1)
#define WKS_ENOMEM (-2)
#define WKS_EIO (-3)
#define WKS_ERANGE (-4)

typedef struct wks_triples wks_triples;
typedef struct wks_gstore wks_gstore;

/* Seeded synthetic RDF graph. Vertex ids are 2^17 + [0, num_entities),
 * predicates are 2 .. num_preds+1, types are num_preds+2 .. num_preds+1+num_types
 * and rdf:type is id 1 (datagen/generate_data.cpp:88-98). For each predicate a
 * seeded permutation ranks the subjects and rank r gets out-degree
 * min(degree_cap, floor(C * r^-zipf_s)) with C solved so that the degrees sum to
 * about num_triples/num_preds; objects are uniform and distinct per subject.
 * Predicates 0..num_hubs-1 (in order) raise the degree of their rank-1 subject to
 * hub_degree. Type k gets type_sizes[k] distinct uniformly chosen members.
 *
 * Two planted correlations give multi-pattern cycles non-empty answers (on
 * uniform objects a closing edge almost never exists):
 *  - linked pairs: for k < num_linked, predicate 2+2k+1 is linked to
 *    predicate 2+2k. A subject s of the linked predicate with out-degree d
 *    that is also a subject of the source predicate takes
 *    min(ceil(link_frac * d), |source list of s|) of its d objects from its
 *    own source list (a seeded subset), the rest uniform and distinct as
 *    usual; so (s linked o) implies (s source o) for about link_frac of its
 *    edges. Degrees and subject ranks are unchanged.
 *  - sibling pairs: for each k < num_sibling, predicate sibling_src[k] (an
 *    id in 2..num_preds+1) plants edges of predicate sibling_dst[k]: for
 *    every subject of sibling_src[k] with ascending objects o_0 < o_1 < ...,
 *    the triples (o_2i, sibling_dst[k], o_2i+1). Every such pair of objects
 *    then closes a triangle. These triples come after the type triples; a
 *    duplicate of an existing triple is removed by the builders. */
typedef struct wks_gen_params {
  uint64_t seed;
  uint64_t num_entities;
  uint32_t num_preds;
  uint32_t num_types;
  uint64_t num_triples;
  double zipf_s;
  uint64_t degree_cap;        /* 0: uncapped */
  const uint64_t *type_sizes; /* num_types entries */
  uint32_t num_hubs;
  uint64_t hub_degree;
  uint32_t num_linked;        /* linked predicate pairs (0: none) */
  double link_frac;           /* share of a linked list taken from its source list */
  uint32_t num_sibling;       /* sibling plants (0: none) */
  const uint32_t *sibling_src; /* num_sibling predicate ids */
  const uint32_t *sibling_dst; /* num_sibling predicate ids */
} wks_gen_params;

int wks_generate(const wks_gen_params *params, wks_triples **out);
/* Copies n triples given as three u32 arrays. */
int wks_triples_from_arrays(const uint32_t *s, const uint32_t *p, const uint32_t *o,
                            uint64_t n, wks_triples **out);
/* Reads an ID triple file, one "s\tp\to" per line (datagen/generate_data.cpp:146-150). */
int wks_triples_read_file(const char *path, wks_triples **out);
int wks_triples_write_file(const wks_triples *t, const char *path);

/* A dataset directory in the reference's ID format (datagen/generate_data.cpp
 * :146-167, read by core/dgraph.hpp:362-417): every id_* file holds
 * "s\tp\to" lines; str_index maps the index ids (__PREDICATE__ 0, rdf:type
 * 1, predicates and types) and str_normal the normal vertex ids, one
 * "<string>\t<id>" per line. wks_dataset_read reads the id_* files in name
 * order, and, when str_index is there, checks that every predicate (and
 * every type a triple names) is one of its ids. */
typedef struct wks_dataset_info {
  uint64_t files, triples;
  uint64_t index_ids, max_index_id;   /* str_index, without __PREDICATE__ */
  uint64_t normal_ids, max_normal_id; /* str_normal */
  int32_t has_str_index, has_str_normal;
} wks_dataset_info;
int wks_dataset_read(const char *dir, wks_triples **out, wks_dataset_info *info);
/* Writes t as `files` id_* files plus str_index and str_normal (with
 * made-up strings) into dir (created if missing). */
int wks_dataset_write(const wks_triples *t, const char *dir, int files);
uint64_t wks_triples_count(const wks_triples *t);
const uint32_t *wks_triples_s(const wks_triples *t);
const uint32_t *wks_triples_p(const wks_triples *t);
const uint32_t *wks_triples_o(const wks_triples *t);
void wks_triples_free(wks_triples *t);

/* main_load: target keys per data slot of the main-header region (0 < x <= 1).
 * The indirect region is sized to exactly what the chains need. */
typedef struct wks_build_params {
  double main_load;
  int num_threads; /* 0: OpenMP default */
} wks_build_params;

int wks_build(const wks_triples *t, const wks_build_params *params, wks_gstore **out);
void wks_gstore_free(wks_gstore *g);

uint64_t wks_num_slots(const wks_gstore *g);        /* 8 x (main + indirect buckets) */
uint64_t wks_num_buckets(const wks_gstore *g);      /* main-header buckets */
uint64_t wks_num_buckets_ext(const wks_gstore *g);  /* indirect buckets */
uint64_t wks_num_keys(const wks_gstore *g);
const uint64_t *wks_slots(const wks_gstore *g);     /* 2 u64 (key, ptr) per slot */
uint64_t wks_num_edges(const wks_gstore *g);
const uint32_t *wks_edges(const wks_gstore *g);
int wks_num_metas(const wks_gstore *g);             /* max predicate/type id + 1 */
const void *wks_metas(const wks_gstore *g);         /* wkg_pred_meta[num_metas] */

/* Local lookup of key (vid, pid, dir): get_edges_local (core/gstore.hpp:432-450).
 * A missing key gives size 0 and a null list. */
int wks_get_edges(const wks_gstore *g, uint64_t vid, uint32_t pid, int dir,
                  const uint32_t **edges, uint64_t *size);

/* Persisted image: one binary file holding slots, edges and metas. */
int wks_gstore_save(const wks_gstore *g, const char *path);
int wks_gstore_load(const char *path, wks_gstore **out);

const char *wks_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
