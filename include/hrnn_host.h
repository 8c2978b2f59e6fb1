/*
 * hrnn_host.h -- C entry points over hnswlib::HRNN<float> (libhrnn_host.so).
 *
 * For callers that cannot use the C++ class directly (Python via ctypes, other
 * languages via FFI).  Each function wraps one method of the class in
 * include/hnswlib/hrnn.h; C++ exceptions become a negative status and the
 * message is kept in the handle (hrnn_index_last_error).
 *
 * Reference interfaces wrapped (file:line in the reference repository):
 *   hrnn_index_create        HRNN(space, max_elements, M, efc, seed, false, K)  include/hnswlib/hrnn.h:128
 *   hrnn_index_load          HRNN(space, location, ...) / loadIndex            include/hnswlib/hrnn.h:117, :930
 *   hrnn_index_add_points    addPoint (OpenMP loop as src/hrnn_build.cpp:75)   include/hnswlib/hrnn.h:1179
 *   hrnn_index_add_points_gpu addPoint for a whole batch, on the GPU           include/hnswlib/hrnn.h:1378-1517
 *   hrnn_index_refine        refineKNNG                                         include/hnswlib/hrnn.h:1725
 *   hrnn_index_build_rknng   buildRKNNG                                         include/hnswlib/hrnn.h:1869
 *   hrnn_index_build_rknng_mutable  buildRKNNG_mutable                          include/hnswlib/hrnn.h:1942
 *   hrnn_index_search        searchRknn (batched)                               include/hnswlib/hrnn.h:1966
 *   hrnn_index_insert_dynamic insertPointDynamic                                include/hnswlib/hrnn.h:2360
 *   hrnn_index_insert_dynamic_gpu insertPointDynamic for a whole batch, on the GPU include/hnswlib/hrnn.h:2360-2534
 *   hrnn_index_compact_for_serving compactForServing                            include/hnswlib/hrnn.h:299
 *   hrnn_index_save          saveIndex                                          include/hnswlib/hrnn.h:861
 */
#ifndef HRNN_HOST_H
#define HRNN_HOST_H

#include <stdint.h>

#include "hrnn_dev.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hrnn_index *hrnn_index_t;

int hrnn_index_create(uint64_t max_elements, uint32_t d, uint32_t M, uint32_t ef_construction,
                      uint64_t random_seed, uint32_t K_knng, hrnn_index_t *out);
int hrnn_index_load(const char *path, uint32_t d, uint64_t max_elements, hrnn_index_t *out);
int hrnn_index_free(hrnn_index_t ix);
const char *hrnn_index_last_error(hrnn_index_t ix);
uint64_t hrnn_index_count(hrnn_index_t ix);
/* Configuration of a built or loaded index: M, ef_construction, K_knng, max level. */
int hrnn_index_info(hrnn_index_t ix, uint32_t *M, uint32_t *efc, uint32_t *K, int *maxlevel);

/* addPoint for n rows; labels may be NULL (label = running count); nthreads
 * <= 1 inserts sequentially (deterministic). */
int hrnn_index_add_points(hrnn_index_t ix, const float *x, uint64_t n, const uint64_t *labels,
                          int nthreads);
/* refineKNNG on the GPU; stats may be NULL or hold num_iters records. */
/* HNSW construction with KNN seeding on the GPU for an empty index
 * (HRNN::addPointsDevice): rows x[0 .. n) become internal ids 0 .. n-1 with
 * labels[i] (i when labels is NULL); batch_permille (0: the default of 50)
 * sizes the insertion batches relative to the graph, seed_expansions (< 0:
 * the default of 32) the seed completion of each batch; stats (may be NULL)
 * receives the build's counters. */
int hrnn_index_add_points_gpu(hrnn_index_t ix, const float *x, uint64_t n, const uint64_t *labels,
                              int batch_permille, int seed_expansions, hrnn_hnsw_build_stats *stats);
int hrnn_index_refine(hrnn_index_t ix, int num_iters, int sample_size, hrnn_nnd_iter_stats *stats);
int hrnn_index_build_rknng(hrnn_index_t ix);
int hrnn_index_build_rknng_mutable(hrnn_index_t ix);
/* Copy the ranked KNN graph out / in (ids[n*K], dists[n*K], cnt[n]). */
int hrnn_index_get_knng(hrnn_index_t ix, uint32_t *ids, float *dists, uint16_t *cnt);
int hrnn_index_set_knng(hrnn_index_t ix, const uint32_t *ids, const float *dists, const uint16_t *cnt);
/* Reverse CSR as built by buildRKNNG: offsets[n+1], entries (4 or 8 bytes each). */
int hrnn_index_get_rknng(hrnn_index_t ix, uint32_t *rank_bits, int *wide, uint64_t *n_entries,
                         uint64_t *offsets, void *entries);
/* Phase A only: m proxies per query (internal ids, 0xFFFFFFFF-padded), on
 * the host (nthreads OpenMP threads) or, with on_gpu != 0, on the GPU. */
int hrnn_index_proxies(hrnn_index_t ix, const float *q, uint64_t nq, uint32_t m, uint32_t ef,
                       int nthreads, uint32_t *out);
int hrnn_index_proxies_gpu(hrnn_index_t ix, const float *q, uint64_t nq, uint32_t m, uint32_t ef,
                           uint32_t *out);
/* Host phase A with beam statistics, then the 32 learned-m features per query
 * (feats[nq*32]); proxies[nq*m_max] (m_max >= 16) receives the phase-A
 * proxies, closest first; raw[nq*20] (may be NULL) the beam statistics. */
int hrnn_index_query_features(hrnn_index_t ix, const float *q, uint64_t nq, uint32_t k, uint32_t ef,
                              uint32_t m_max, int nthreads, float *feats, uint32_t *proxies, float *raw);
/* searchRknn for nq queries: labels concatenated with offsets[nq+1].
 * per_query != 0 calls the single-query method once per query (the driver
 * loop of src/hrnn_search.cpp:168-179); 0 runs one batch. */
int hrnn_index_search(hrnn_index_t ix, const float *q, uint64_t nq, uint32_t m, float threshold,
                      uint32_t k, uint32_t K_prime, uint32_t ef, int nthreads, int per_query,
                      uint64_t *offsets, uint64_t *labels, uint64_t cap, uint64_t *needed);
/* Phases B+C for caller-given proxies (internal ids in, internal ids out). */
int hrnn_index_search_proxies(hrnn_index_t ix, const float *q, uint64_t nq, const uint32_t *prox,
                              uint32_t m, const uint32_t *k_per_q, uint32_t k, uint32_t K_prime,
                              float threshold, uint64_t *offsets, uint32_t *ids, uint64_t cap,
                              uint64_t *needed);
int hrnn_index_insert_dynamic(hrnn_index_t ix, const float *x, uint64_t n, const uint64_t *labels,
                              uint32_t m_update, int bidirectional, int nthreads);
/* insertPointDynamic for n points at once on the GPU
 * (HRNN::insertPointsDynamicDevice over hrnn_dev_insert_dynamic); batch_permille
 * (0: the default of 50) sizes the insertion batches relative to the graph;
 * stats may be NULL. */
int hrnn_index_insert_dynamic_gpu(hrnn_index_t ix, const float *x, uint64_t n, const uint64_t *labels,
                                  uint32_t m_update, int bidirectional, int batch_permille,
                                  hrnn_insert_stats *stats);
int hrnn_index_compact_for_serving(hrnn_index_t ix, uint32_t max_k_serve);
/* The mutable reverse lists (after buildRKNNG_mutable, kept up to date by
 * insertPointDynamic; reference include/hnswlib/hrnn.h:1942-1959, :2360-2534)
 * flattened: off[n+1] always, ids[off[n]] (each list in its own order) when
 * ids is not NULL. */
int hrnn_index_get_rknng_mutable(hrnn_index_t ix, uint64_t *off, uint32_t *ids);
int hrnn_index_save(hrnn_index_t ix, const char *path);
/* The index's device handle (opened on first use) with base, graph and CSR resident. */
int hrnn_index_device(hrnn_index_t ix, hrnn_dev_t *out);
/* The HNSW arrays in the layout hrnn_dev_set_hnsw uploads: l0[n*(maxM0+1)],
 * levels[n], up_off[n], up_pool[*pool_words]; with up_pool == NULL only
 * *pool_words is set.  entry / maxlevel may be NULL. */
int hrnn_index_export_hnsw(hrnn_index_t ix, uint32_t *l0, int *levels, uint32_t *up_off,
                           uint32_t *up_pool, uint64_t *pool_words, uint32_t *entry, int *maxlevel);
/* External label of every internal id (n entries).  Internal ids follow
 * insertion order, so they differ from labels after a multi-threaded build. */
int hrnn_index_get_labels(hrnn_index_t ix, uint64_t *labels);
/* Level-0 links (maxM0 + 1 words per node: count, then ids) and node levels. */
int hrnn_index_get_level0(hrnn_index_t ix, uint32_t *links, uint32_t words_per_node, int *levels);

#ifdef __cplusplus
}
#endif
#endif /* HRNN_HOST_H */
