/*
 * hrnn_dev.h -- C-ABI of libhrnn_hip.so, the MI355X (gfx950) device layer
 * beneath hnswlib::HRNN<float>.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, returns an
 * int status (HRNN_OK == 0, negative on error) and never throws.  The text of
 * the last error of a handle is available from hrnn_dev_last_error().
 *
 * A handle owns one HIP stream and all device memory of one index on one GPU:
 * base vectors (FP32 plus a BF16 mirror and FP32 row norms), the ranked KNN
 * graph (SoA: ids u32[n*K], dists f32[n*K], counts u32[n], new-flags u8[n*K])
 * and the reverse-KNN CSR.  Host buffers passed in are caller-owned and are
 * only read or written during the call.  A handle is not re-entrant: calls on
 * one handle must be serialised by the caller.
 *
 * Entry points whose name ends in _dptr take device pointers (memory already
 * resident in HBM) instead of host pointers.
 *
 * Reference interfaces replaced (file:line in the reference repository):
 *   hrnn_dev_nndescent        HRNN::refineKNNG              include/hnswlib/hrnn.h:1725
 *   hrnn_dev_build_rknn_csr   HRNN::buildRKNNG / _fill      include/hnswlib/hrnn.h:1869, :1912
 *   hrnn_dev_rknn_search      HRNN::searchRknn_impl phases B+C include/hnswlib/hrnn.h:2021-2069
 *   hrnn_dev_bf_kth           compute_kth_nn_dists          scripts/gen_rknn_gt.py:42-79
 *   hrnn_dev_bf_knn           BruteforceSearch::searchKnn   include/hnswlib/bruteforce.h:106-135
 *   hrnn_dev_bf_rknn          compute_rknn                  scripts/gen_rknn_gt.py:82-151
 *   hrnn_dev_l2sqr_pairs      L2Sqr / L2SqrSIMD16ExtAVX512  include/hnswlib/space_l2.h:6-54
 *   hrnn_dev_hnsw_build       HRNN::addPoint (KNN seeding)  include/hnswlib/hrnn.h:1378-1517
 *   hrnn_dev_insert_dynamic   HRNN::insertPointDynamic      include/hnswlib/hrnn.h:2360-2534
 */
#ifndef HRNN_DEV_H
#define HRNN_DEV_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hrnn_dev *hrnn_dev_t;

enum hrnn_status {
    HRNN_OK = 0,
    HRNN_E_ARG = -1,      /* invalid argument (null pointer, size out of range) */
    HRNN_E_HIP = -2,      /* a HIP runtime call failed */
    HRNN_E_STATE = -3,    /* call out of order (e.g. search before the CSR exists) */
    HRNN_E_OOM = -4,      /* device allocation failed */
    HRNN_E_CAP = -5,      /* an output capacity was too small; *needed is set */
    HRNN_E_RANGE = -6     /* packing overflow (node_bits + rank_bits > 64) */
};

enum hrnn_precision {
    HRNN_PREC_F32 = 0,    /* FP32 direct difference (reference summation order) */
    HRNN_PREC_BF16 = 1,   /* BF16 inputs on MFMA, FP32 accumulate, FP32 re-rank */
    HRNN_PREC_F16 = 2,    /* FP16 inputs on MFMA, FP32 accumulate, FP32 re-rank */
    HRNN_PREC_U8 = 3      /* vector storage only: the u8 mirror (8-bit data) */
};

/* ---- handle ---------------------------------------------------------- */
int hrnn_dev_device_count(int *count);
int hrnn_dev_open(int device, hrnn_dev_t *out);
int hrnn_dev_close(hrnn_dev_t h);
const char *hrnn_dev_last_error(hrnn_dev_t h);
/* Wait for all work queued on the handle's stream. */
int hrnn_dev_sync(hrnn_dev_t h);
/* The handle's HIP stream (a hipStream_t), for callers that time or order work. */
void *hrnn_dev_stream(hrnn_dev_t h);

/* ---- base vectors ------------------------------------------------------ */
/* Copy n x d row-major FP32 vectors into HBM; builds the BF16/FP16 mirrors and
 * FP32 norms.  Replaces any previous base and invalidates the KNN graph and CSR. */
int hrnn_dev_set_base(hrnn_dev_t h, const float *x, uint64_t n, uint32_t d);
int hrnn_dev_set_base_dptr(hrnn_dev_t h, const float *x_dev, uint64_t n, uint32_t d);
/* Append rows [n_old, n_old + n_add) (dynamic insert); keeps the KNN graph. */
int hrnn_dev_append_base(hrnn_dev_t h, const float *x, uint64_t n_add);
/* 1 if every base value is an integer in [-256, 256] (BF16/FP16-exact data). */
int hrnn_dev_base_is_small_int(hrnn_dev_t h, int *is_int);
/* 1 if every base value is an integer in [0, 255] (the u8 mirror and the int8
 * NNDescent join apply).  No reference counterpart: the autotuner's choice of
 * candidates. */
int hrnn_dev_base_is_u8(hrnn_dev_t h, int *is_u8);
/* Device pointer of the FP32 base (n*d floats). */
const float *hrnn_dev_base_ptr(hrnn_dev_t h);
/* Storage the search and phase-A kernels read vectors from: HRNN_PREC_BF16
 * (default) uses the BF16 mirror when every base value is exactly
 * representable in it (then distances are bit-identical to FP32 and the
 * vector bytes halve) and FP32 otherwise; HRNN_PREC_U8 uses the u8 mirror
 * when every base value is an integer in [0, 255] (a quarter of the FP32
 * bytes; distances of u8 queries by integer dot products, exact, so again
 * bit-identical) and falls back as BF16 otherwise; HRNN_PREC_F32 always
 * reads FP32. */
int hrnn_dev_set_vector_precision(hrnn_dev_t h, int prec);
/* Launch shapes set by the autotuner (hrnn_amd/autotune.py): "search_wg_per_cu"
 * (persistent RkNN search workgroups per CU, default 8); "search_lds_kib" (KiB
 * of LDS for the search's per-query deduplication table, default 16; a query
 * whose scanned entries exceed 3/4 of the table's slots uses a visited bitmap
 * in HBM instead, with identical results); "beam_wg_per_cu" is
 * accepted and unused (phase A launches one workgroup per query); values in
 * [1, 64].  "stage1_cap" (0 .. 1024, default 0 = none): the learned early
 * stop's stage-1 beam stops after this many expansions, and a query stopped
 * early whose class would keep it at stage 1 takes the first class with a
 * wider beam instead (the slow tail of stage 1 moves to stage 2, which
 * overlaps the other queries' search; results change only for those
 * queries).  "hnsw_batch_permille" (1 .. 1000, default 100) and
 * "hnsw_seed_exp" (0 .. 1000, default 32): the HNSW construction's batch size
 * relative to the graph and its seed completion.  "insert_offer_cap"
 * (0 .. 2^30, default 0 = max(2^22, 24n)): the incremental insertion's first
 * offer buffer (entries per batch; it grows when a batch needs more);
 * "insert_lds_bitmap" (0 or 1, default 1): the insertion's visited bitmap in
 * LDS when n bits fit there (160 KiB per workgroup on gfx950), else in HBM
 * (the same results).  "beam_vis_shift" (-1 .. 2, default 0): phase A's
 * LDS visited set per query is 2^(rule(ef) + shift) slots, within
 * 2^10 .. 2^13 (more slots cost occupancy, fewer overflow on long beams; the
 * same results).  "bf_row_frags" (1 or 2, default 1): the MFMA brute force's
 * 16-row fragments per wave (2 applies to 32-wide candidate lists at
 * d <= 256; the same results).  "nnd_join" (0 .. 3, default 0 = chosen from
 * d and the data): NNDescent's join staging, 1 chunked BF16, 2 whole-sample
 * BF16, 3 whole-sample int8 (8-bit data only; a mode that does not fit falls
 * back to chunked; the same lists).  "query_streams" (1 or 2, default 1):
 * hrnn_dev_rknn_query_dptr with 2 runs a batch of at least 2,048 queries as
 * two parts (each phase A then phases B+C) on two streams, so that their
 * kernels overlap (the same answers); "query_split" (100 .. 900, default
 * 500) is the first part's share in permille (a smaller first part's
 * phases B+C run beside the second part's phase A).  "csr_wide" (0 or 1, default 0): the
 * reverse CSR with u64 entries even where (node << rank_bits | rank) fits 32
 * bits (the same values; the layout large graphs need, testable on small
 * ones).  hrnn_amd/autotune.py times these
 * choices per workload and caches the fastest. */
int hrnn_dev_set_tuning(hrnn_dev_t h, const char *key, int value);
int hrnn_dev_get_tuning(hrnn_dev_t h, const char *key, int *value);

/* ---- ranked KNN graph --------------------------------------------------- */
/* Host layout (SoA): ids[n*K], dists[n*K] (squared L2), cnt[n] (u16, <= K).
 * Entries j >= cnt[i] are ignored.  All entries start with the "new" flag set,
 * as at the start of refineKNNG (hrnn.h:1744-1748). */
int hrnn_dev_set_knng(hrnn_dev_t h, uint32_t K, const uint32_t *ids,
                      const float *dists, const uint16_t *cnt);
int hrnn_dev_get_knng(hrnn_dev_t h, uint32_t *ids, float *dists, uint16_t *cnt);
/* The entries' NNDescent "new" flags, u8 [n*K] (the reference's per-entry
 * is_new vector, hrnn.h:1744-1748, :1775-1778): read after refinement, or
 * set after hrnn_dev_set_knng to resume refinement from a saved state
 * (entries j >= cnt[i] are ignored); flags NULL marks every entry new, the
 * state refineKNNG starts from. */
int hrnn_dev_get_knng_flags(hrnn_dev_t h, uint8_t *flags);
int hrnn_dev_set_knng_flags(hrnn_dev_t h, const uint8_t *flags);
/* Device pointers of the resident graph (ids n*K u32, dists n*K f32, cnt n u32). */
int hrnn_dev_knng_ptrs(hrnn_dev_t h, uint32_t **ids, float **dists, uint32_t **cnt);

/* ---- HNSW construction with KNN seeding --------------------------------- */
typedef struct hrnn_hnsw_build_stats {
    uint64_t batches;         /* insertion batches */
    uint64_t expansions;      /* beam expansions over all layers */
    uint64_t evaluations;     /* distances evaluated by the beams */
    uint64_t dropped_links;   /* links back beyond 64 per node, batch and layer (not made) */
    double ms;                /* wall time of the call */
    uint64_t seed_expansions; /* expansions of the seed completion */
    uint64_t seed_evaluations;
} hrnn_hnsw_build_stats;

/* Build the HNSW graph over all n resident base rows, inserted in id order
 * with levels[n] (drawn by the caller: floor(-ln U / ln M)), and, when
 * K > 0, the ranked KNN seed lists (the layer-0 beam sorted by (dist, id),
 * self skipped, the first K; every entry new).  M links per new node per
 * layer, lists of 2M at layer 0 and M above, beams of width ef_construction
 * (<= 512), 2 <= M <= 32.  Points go in batches of "hnsw_batch_permille"
 * (hrnn_dev_set_tuning, default 100) of the graph's size; each batch searches
 * the graph as it stood when the batch began, then every node's new links
 * back are merged into its list at once (appended, or the list pruned by the
 * diversity rule over entries and new links); each new point's seed list
 * then continues its layer-0 beam over the merged graph for
 * "hnsw_seed_exp" expansions (default 32; 0: none), admitting only lower
 * ids, so it sees the earlier points of its own batch as sequential
 * insertion would.  The graph stays resident for
 * phase A (hrnn_dev_proxies) and the KNN lists for hrnn_dev_nndescent;
 * *entry and *maxlevel receive the entry point and top layer.  Replaces
 * HRNN::addPoint (include/hnswlib/hrnn.h:1378-1517) with searchBaseLayer
 * (:401-481), getNeighborsByHeuristic2 (:619-666) and
 * mutuallyConnectNewElement (:682-806). */
int hrnn_dev_hnsw_build(hrnn_dev_t h, const int *levels, uint32_t M, uint32_t ef_construction, uint32_t K,
                        uint32_t *entry, int *maxlevel, hrnn_hnsw_build_stats *stats);
/* Download the resident graph: l0[n*(2M+1)] and the upper-layer pool
 * (pool_words = the sum over nodes of level * (M+1) words, blocks in id
 * order: node i's at the running sum over nodes before it). */
int hrnn_dev_get_hnsw(hrnn_dev_t h, uint32_t *l0, uint32_t *up_pool, uint64_t pool_words);

/* ---- incremental insertion ------------------------------------------- */
typedef struct hrnn_insert_stats {
    uint64_t batches;         /* insertion batches */
    uint64_t scanned;         /* reverse-list entries read while collecting candidates */
    uint64_t candidates;      /* distinct candidates (distances evaluated) */
    uint64_t offers;          /* offers that passed the batch-start filter */
    uint64_t updates;         /* entries the offers put into lists */
    uint64_t retries;         /* batch scans run again after the offer buffer grew */
    double ms;                /* wall time of the call */
    double hnsw_ms, csr_ms, affect_ms, merge_ms;   /* device time by stage (HIP events) */
    hrnn_hnsw_build_stats hnsw;                    /* the HNSW insertion's own counters */
} hrnn_insert_stats;

/* Incremental insertion of the resident rows [n_old, n) (appended with
 * hrnn_dev_append_base) into the resident index of rows [0, n_old): the HNSW
 * graph (from hrnn_dev_hnsw_build, hrnn_dev_set_hnsw or an earlier insertion,
 * linking exactly those rows, with the same M) and the ranked KNN lists
 * (K <= 1024).  levels[n]: every row's level (the new rows' drawn by the
 * caller).  Per new point x, as insertPointDynamic: HNSW insertion with KNN
 * seeding; candidates = the reverse lists of the first m_update nodes of x's
 * list (all when 0), x and its list excluded; each candidate c takes x when
 * its list is not full or d(x, c) < its last distance (the displaced entry
 * leaves); with bidirectional, x's list becomes the K closest of its list and
 * its candidates.  Points go in the HNSW construction's batches
 * (hnsw_batch_permille); each batch's reverse lists are rebuilt from the
 * graph after its seeding, its offers merged per list as the top K of (list
 * U offers).  The reverse CSR is stale afterwards (rebuild it for search);
 * the new points' entries carry the NNDescent new flag.  Replaces
 * HRNN::insertPointDynamic (include/hnswlib/hrnn.h:2360-2534) with
 * try_update_knng_evict (:1592-1655) and the mutable reverse lists of
 * buildRKNNG_mutable (:1942-1959). */
int hrnn_dev_insert_dynamic(hrnn_dev_t h, uint64_t n_old, const int *levels, uint32_t M, uint32_t ef_construction,
                            uint32_t m_update, int bidirectional, uint32_t *entry, int *maxlevel,
                            hrnn_insert_stats *stats);

/* The affected-area step of hrnn_dev_insert_dynamic alone, for the rows
 * [b0, b1) whose ranked lists already hold their seeds (no HNSW insertion):
 * the reverse lists rebuilt from the resident graph, each row's candidates,
 * offers filtered by the lists as they stand and merged (top K of list U
 * offers), with bidirectional each row's own list replaced first.  With
 * b1 = b0 + 1 this is phases 2-5 of the reference's insertPointDynamic for
 * one point (include/hnswlib/hrnn.h:2366-2532). */
int hrnn_dev_insert_affect(hrnn_dev_t h, uint64_t b0, uint64_t b1, uint32_t m_update, int bidirectional,
                           hrnn_insert_stats *stats);

/* ---- (a) NNDescent refinement ------------------------------------------ */
typedef struct hrnn_nnd_iter_stats {
    uint64_t dists;          /* distance evaluations in the local join */
    uint64_t updates;        /* entries inserted into K-lists (both directions) */
    uint64_t proposals;      /* proposals that passed the threshold filter */
    uint64_t dropped;        /* proposals beyond the proposal buffer (0 unless it overflows) */
    double ms;               /* device time of the iteration (HIP events around it) */
    uint64_t rows;           /* sample rows staged by the joins (sum over nodes of new + old samples) */
} hrnn_nnd_iter_stats;

/* num_iters rounds of UPDATE (sample S new / S old, reverse samples capped at
 * R = S, the smallest source ids of up to 256 offers kept) + JOIN (new x new, new x old) + MERGE
 * on the resident graph.  Every proposal of an iteration is kept in a proposal
 * CSR of bucket * n entries (bucket = 0: max(16384, 32 K)), capped at 60% of
 * the free device memory; proposals beyond it are counted as dropped.  stats
 * may be NULL or point to num_iters records.  The iterations run as one HIP
 * graph (environment HRNN_NND_GRAPH=0: launched one by one).  The join
 * stages each node's sample as int8 for 8-bit data at d <= 258, else as BF16
 * (whole, or 64 dims at a time when 4S rows do not fit 56 KiB of LDS);
 * HRNN_NND_JOIN=0/1/2 forces chunked BF16 / whole BF16 / int8 where the data
 * and LDS allow it (all three give the same lists). */
int hrnn_dev_nndescent(hrnn_dev_t h, int num_iters, int sample_size, uint32_t bucket,
                       hrnn_nnd_iter_stats *stats);

/* ---- (b) reverse-KNN CSR ----------------------------------------------- */
/* Transpose the resident KNN graph.  rank_bits = ceil(log2 K), node_bits =
 * ceil(log2 n); entries are u32 if node_bits + rank_bits <= 32, else u64
 * (*wide = 1); each entry is (src << rank_bits) | rank, and each list is in
 * (rank, src) order.  Offsets are u64[n+1].  The CSR stays resident. */
int hrnn_dev_build_rknn_csr(hrnn_dev_t h, uint32_t *rank_bits, int *wide,
                            uint64_t *n_entries);
/* Download offsets (n+1 u64) and entries (n_entries x 4 or 8 bytes). */
int hrnn_dev_get_rknn_csr(hrnn_dev_t h, uint64_t *offsets, void *entries);
/* Upload a CSR built elsewhere (e.g. loaded from an index file). */
int hrnn_dev_set_rknn_csr(hrnn_dev_t h, uint32_t rank_bits, int wide, uint64_t n,
                          const uint64_t *offsets, const void *entries, uint64_t n_entries);

/* ---- (c) RkNN search, phases B and C ----------------------------------- */
/* For each query q (nq rows of d floats) and its m proxies (internal ids,
 * proxies[q*m .. q*m+m)): collect the proxies' reverse-list prefixes with
 * rank < K_prime, keep candidates whose proxy count >= threshold, and accept
 * candidate c if L2Sqr(q, c) <= dist(c, k-1) (FLT_MAX when c has < k
 * neighbours).  k_per_q may be NULL (then every query uses k).  Accepted ids
 * are returned per query in ascending id order: res_off[nq+1] (u64) and
 * res_ids[res_off[nq]].  If cap < total, returns HRNN_E_CAP and sets *needed.
 * proxy ids equal to 0xFFFFFFFF are skipped (a query with fewer proxies). */
int hrnn_dev_rknn_search(hrnn_dev_t h, const float *q, uint32_t nq,
                         const uint32_t *proxies, uint32_t m,
                         const uint32_t *k_per_q, uint32_t k, uint32_t K_prime,
                         float threshold, uint64_t *res_off, uint32_t *res_ids,
                         uint64_t cap, uint64_t *needed);
/* Resident form: q_dev, proxies_dev, k_per_q_dev in HBM; per-query results
 * written to res_cnt_dev[nq] (u32) and res_ids_dev[q * cap_per_q ...].  A query
 * whose result count exceeds cap_per_q keeps the first cap_per_q ids and its
 * count is the true count.  work_dev (NULL or 2*nq u32) receives per query the
 * reverse-list entries scanned and the distinct candidates verified.  Runs as
 * one persistent launch that pulls queries from a device work queue.
 * Asynchronous on the handle's stream; proxy ids must be < n or 0xFFFFFFFF. */
int hrnn_dev_rknn_search_dptr(hrnn_dev_t h, const float *q_dev, uint32_t nq,
                              const uint32_t *proxies_dev, uint32_t m,
                              const uint32_t *k_per_q_dev, uint32_t k, uint32_t K_prime,
                              float threshold, uint32_t *res_cnt_dev,
                              uint32_t *res_ids_dev, uint32_t cap_per_q, uint32_t *work_dev);

/* The resident search without allocation, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  No counterpart in the reference.  Same arguments,
 * results and error behaviour as hrnn_dev_rknn_search_dptr; the search's
 * queue word and its per-workgroup fallback state (visited bitmaps / u16
 * counts, candidate buffers) live in a static device array of 128 MiB that
 * every GPU the library runs on holds for the life of the process, so the
 * call allocates nothing and clears its words with kernels, not memsets.
 * The persistent workgroups are capped so that their fallback state fits the
 * array (n/8 + 8 * min(m * longest reverse list, n) bytes each, n/2 in place
 * of n/8 when threshold > 1); a problem where one workgroup's state does not
 * fit is rejected with HRNN_E_ARG.  The r_k column is used only when an
 * earlier search built it for this k (else the list matrix is read; same
 * results).  Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (any handles, any streams) race and must be serialised
 * by the caller. */
int hrnn_dev_rknn_search_noalloc_dptr(hrnn_dev_t h, const float *q_dev, uint32_t nq,
                                      const uint32_t *proxies_dev, uint32_t m,
                                      const uint32_t *k_per_q_dev, uint32_t k, uint32_t K_prime,
                                      float threshold, uint32_t *res_cnt_dev, uint32_t *res_ids_dev,
                                      uint32_t cap_per_q, uint32_t *work_dev);

/* Download of a resident search's results, densely packed: on the handle's
 * stream the counts are scanned (each capped at cap_per_q) into offsets and
 * the ids gathered into one array on the device, then res_off[nq+1] (u64)
 * and res_ids[res_off[nq]] are copied to the host (pinned host memory makes
 * both copies run at the link's rate; a page-locked res_ids is written by the
 * gather itself, at most cap ids, and the call waits once).  Synchronous.
 * On HRNN_E_CAP res_ids may hold a partial result.  HRNN_E_CAP with
 * *needed set when a query's true count exceeds cap_per_q (its ids were cut
 * short on the device; *needed = the largest count) or when the total
 * exceeds cap (*needed = the total).  Replaces the per-query result vector
 * that searchRknn returns (reference include/hnswlib/hrnn.h:2055-2071). */
int hrnn_dev_fetch_results(hrnn_dev_t h, uint32_t nq, const uint32_t *res_cnt_dev,
                           const uint32_t *res_ids_dev, uint32_t cap_per_q, uint64_t *res_off,
                           uint32_t *res_ids, uint64_t cap, uint64_t *needed);

/* ---- RkNN phase A on the GPU -------------------------------------------- */
/* One query per workgroup (k_phase_a).  Graph ids and counts read by the
 * kernel are range-checked; a bad one sets the handle's device error word,
 * which hrnn_dev_sync (and hrnn_dev_proxies) report as HRNN_E_STATE. */
/* Upload the HNSW graph: l0[n*(maxM0+1)] layer-0 links (word 0: count in the
 * low 16 bits, then ids); levels[n]; up_off[n] offset (in words) of node i's
 * upper-layer block in up_pool (0xFFFFFFFF for level-0 nodes), the block
 * holding (maxM+1) words per layer 1..levels[i]; entry point and top layer.
 * The whole graph is validated on the host before upload.
 * Replaces: the layer-0 and upper link lists of HRNN (hrnn.h:160-174, :262-281). */
int hrnn_dev_set_hnsw(hrnn_dev_t h, uint32_t maxM0, uint32_t maxM, const uint32_t *l0,
                      const int *levels, const uint32_t *up_off, const uint32_t *up_pool,
                      uint64_t pool_words, uint32_t entry, int maxlevel);
/* Phase A of searchRknn_impl (hrnn.h:1992-2019): greedy descent over the
 * upper layers, layer-0 beam with width ef = max(ef_search, m)
 * (searchBaseLayerST<true>, hrnn.h:485-616), the m closest as proxies
 * (proxies[nq*m], 0xFFFFFFFF-padded).  Requires m <= ef <= 512. */
int hrnn_dev_proxies(hrnn_dev_t h, const float *q, uint32_t nq, uint32_t m, uint32_t ef,
                     uint32_t *proxies);
/* Resident form; expansions_dev (NULL or 2*nq u32) receives per query the
 * layer-0 nodes expanded and the distances evaluated at layer 0. */
int hrnn_dev_proxies_dptr(hrnn_dev_t h, const float *q_dev, uint32_t nq, uint32_t m, uint32_t ef,
                          uint32_t *proxies_dev, uint32_t *expansions_dev);
/* Whole searchRknn for a resident batch: phase A into proxies_dev (nq*m), then
 * phases B and C as hrnn_dev_rknn_search_dptr; one stream, no host round trip.
 * q_dev (here, in hrnn_dev_rknn_query_adaptive_dptr and in the other _dptr
 * calls) may also be the device address of page-locked host memory
 * (hipHostGetDevicePointer): the kernels then read the queries over the link
 * in place, with no copy before the call.  Here and in the adaptive call the
 * first phase-A launch reads each query once and writes it to a device
 * buffer of the handle (nq*d floats, kept for later calls), which the rest
 * of the call reads. */
int hrnn_dev_rknn_query_dptr(hrnn_dev_t h, const float *q_dev, uint32_t nq, uint32_t m,
                             uint32_t ef, const uint32_t *k_per_q_dev, uint32_t k,
                             uint32_t K_prime, float threshold, uint32_t *proxies_dev,
                             uint32_t *res_cnt_dev, uint32_t *res_ids_dev, uint32_t cap_per_q,
                             uint32_t *work_dev);

/* ---- (d) exact brute force --------------------------------------------- */
/* kth[i] = (k+1)-th smallest squared distance from base row i to all base rows
 * (self included), i.e. the k-th neighbour distance r_k(i).  prec selects the
 * candidate pass; the final value is the FP32 direct-difference distance after
 * re-ranking the top 2(k+1) candidates by (dist, id).  ids (may be NULL) gets
 * the (k+1) nearest ids per row in (dist, id) order, self included.
 * Every row's candidate list must carry a certificate (an error bound of the
 * candidate pass shows that no other row can rank among the first k+1); rows
 * without one are redone with a 64-wide list, then with the FP32 pass, so the
 * results equal an exhaustive FP32 re-rank. */
int hrnn_dev_bf_kth(hrnn_dev_t h, uint32_t k, int prec, float *kth, uint32_t *ids);

/* Certificate statistics of a brute-force top-k. */
typedef struct {
    uint64_t rows;         /* rows ranked */
    uint64_t certified;    /* rows certified after the first candidate pass */
    uint64_t redo_wide;    /* rows redone with a 64-wide list in the same format */
    uint64_t redo_f32;     /* rows redone with the FP32 candidate pass */
    uint64_t uncertified;  /* rows no pass certified (results of the last pass) */
    double ms_first;       /* device time of the first pass and its re-rank, ms */
    double ms_total;       /* device time of all passes, ms */
} hrnn_bf_stats;

/* hrnn_dev_bf_kth with the precision study's outputs (each may be NULL):
 * kth64[n] the (k+1)-th smallest FP64 distance (sequential FP64 sum, no fused
 * multiply-add), certified like kth; raw_kth[n] and raw_ids[n*(k+1)] the first
 * candidate pass's own (k+1)-th distance and first k+1 ids, before re-rank
 * (in prec's format: the error a BF16/FP16 oracle would make); stats the
 * certificate counts and device times.  No reference counterpart (the
 * reference's brute force is FP32 only, scripts/gen_rknn_gt.py:42-79). */
int hrnn_dev_bf_kth_ex(hrnn_dev_t h, uint32_t k, int prec, float *kth, double *kth64, uint32_t *ids,
                       float *raw_kth, uint32_t *raw_ids, hrnn_bf_stats *stats);
/* Exact kNN of nq host queries against the base: ids/d2 [nq*k] in (dist, id)
 * order (ties to the smaller id). */
int hrnn_dev_bf_knn(hrnn_dev_t h, const float *q, uint32_t nq, uint32_t k, int prec,
                    uint32_t *ids, float *d2);
/* RkNN groundtruth: for each query, every base row x with
 * d2(q, x) <= kth[x] (FP32 direct difference), sorted by (d2, id). */
int hrnn_dev_bf_rknn(hrnn_dev_t h, const float *q, uint32_t nq, const float *kth,
                     int prec, uint64_t *res_off, uint32_t *res_ids, float *res_d2,
                     uint64_t cap, uint64_t *needed);
/* The same with the FP64 re-check: a pair whose FP32 distance lies within
 * FP32 rounding of kth[x] (a rigorous bound for any summation order) is
 * decided by its FP64 distance against kth64[x] (kth and kth64 as
 * hrnn_dev_bf_kth_ex gives them); every other pair's FP32 decision provably
 * equals the FP64 one.  rechecked / flipped (may be NULL) count the pairs
 * re-checked and those whose FP64 decision differs from the FP32 one.
 * res_d2 holds FP32 distances.  Replaces the FP32 decision of compute_rknn
 * (scripts/gen_rknn_gt.py:116-124) with the FP64 judge's. */
int hrnn_dev_bf_rknn_f64(hrnn_dev_t h, const float *q, uint32_t nq, const float *kth,
                         const double *kth64, uint64_t *res_off, uint32_t *res_ids, float *res_d2,
                         uint64_t cap, uint64_t *needed, uint64_t *rechecked, uint64_t *flipped);

/* ---- learned proxy count (no reference counterpart) --------------------- */
/* A 32-64-64-16 MLP (ReLU, softmax) predicts per query the smallest m of a
 * 16-value grid that reaches full recall; plain fixed-m search stays the
 * fallback.  Parameters: HRNN_MLP_PARAMS floats laid out
 * W1[64][32] b1[64] W2[64][64] b2[64] W3[16][64] b3[16]; a model adds the
 * feature standardisation mu[32] sd[32] after them (HRNN_MLP_MODEL floats). */
#define HRNN_MLP_PARAMS 7312
#define HRNN_MLP_MODEL (HRNN_MLP_PARAMS + 64)
/* probs[n*16] = softmax(MLP(feats)). */
int hrnn_dev_mlp_forward(hrnn_dev_t h, const float *params, const float *feats, uint32_t n,
                         float *probs);
/* Mean cross-entropy gradient over the n samples (fixed reduction order). */
int hrnn_dev_mlp_grad(hrnn_dev_t h, const float *params, const float *feats,
                      const uint32_t *labels, uint32_t n, float *grad, float *loss);
/* Adam training in place on params (mini-batches, shuffled with seed);
 * epoch_loss may be NULL or hold one mean loss per epoch. */
int hrnn_dev_mlp_train(hrnn_dev_t h, float *params, const float *feats, const uint32_t *labels,
                       uint32_t n, int epochs, uint32_t batch, float lr, float weight_decay,
                       uint64_t seed, float *epoch_loss);
/* Phases B and C with learned m for a resident batch whose phase A ran on the
 * host: feats_dev (nq*32 unstandardised features, standardised in place with
 * the model's mu/sd), MLP forward, per-query m = the smallest value of the
 * 16-value grid (grid_dev) whose cumulative probability reaches tau (argmax
 * when tau <= 0; proxies_dev rows of m_max masked beyond m; m_dev, NULL or
 * nq u32, receives the chosen m), then phases B and C.  All work is queued on the
 * library stream; nothing is copied to the host. */
int hrnn_dev_rknn_search_learned_dptr(hrnn_dev_t h, const float *q_dev, uint32_t nq, float *feats_dev,
                                      const float *model_dev, const uint32_t *grid_dev, float tau,
                                      uint32_t *proxies_dev, uint32_t m_max, uint32_t *m_dev, uint32_t k,
                                      uint32_t K_prime, float threshold, uint32_t *res_cnt_dev,
                                      uint32_t *res_ids_dev, uint32_t cap_per_q, uint32_t *work_dev);

/* Learned early stop (no reference counterpart; it replaces the fixed
 * ef_search / m of searchRknn, hrnn.h:1979-2019, per query).
 *
 * Stage-1 phase A at beam width ef1 with the 32 learned-m features computed
 * inside the beam kernel (the definitions of HRNN::queryFeatures, over the
 * first 16 list entries; r_k with this k): proxies_dev[nq*m],
 * feats_dev[nq*32] unstandardised, expansions_dev (NULL or 2*nq u32:
 * expansions, evaluations).  Used to build training sets. */
int hrnn_dev_stage1_features_dptr(hrnn_dev_t h, const float *q_dev, uint32_t nq, uint32_t ef1,
                                  uint32_t m, uint32_t k, uint32_t *proxies_dev, float *feats_dev,
                                  uint32_t *expansions_dev);
/* Whole two-stage query for a resident batch: stage-1 phase A at ef1 for
 * every query (features standardised with the model's mu/sd); MLP forward;
 * per query the cheapest of 16 (ef, m) classes whose cumulative probability
 * reaches tau (argmax when tau <= 0), written to cls_dev[nq]; stage-2 phase A
 * at the class's ef for the queries whose class has ef > ef1; proxies beyond
 * the class's m masked (rows of m_max in proxies_dev); phases B and C.
 * classes_dev: 32 u32, the 16 classes' ef then their m, ordered by cost with
 * ef and m non-decreasing; ef_max >= every class ef (<= 512).  With
 * model_dev NULL the classes are read from cls_dev instead (forced classes).
 * work_dev as in hrnn_dev_rknn_search_dptr; expansions_dev (NULL or 4*nq u32)
 * receives per query the stage-1 (expansions, evaluations) in [0, 2nq) and
 * the stage-2 ones in [2nq, 4nq) (zero without a stage 2).  With res_cnt_dev
 * NULL only phase A runs (the classes and masked proxies are the outputs;
 * ef_max = ef1 leaves stage 1 alone, as the query runs it).
 * Asynchronous on the library stream; nothing is copied to the host. */
int hrnn_dev_rknn_query_adaptive_dptr(hrnn_dev_t h, const float *q_dev, uint32_t nq,
                                      const float *model_dev, const uint32_t *classes_dev,
                                      uint32_t ef_max, float tau, uint32_t ef1, uint32_t m_max,
                                      uint32_t k, uint32_t K_prime, float threshold,
                                      uint32_t *proxies_dev, uint32_t *cls_dev,
                                      uint32_t *res_cnt_dev, uint32_t *res_ids_dev,
                                      uint32_t cap_per_q, uint32_t *work_dev,
                                      uint32_t *expansions_dev);

/* ---- distances ---------------------------------------------------------- */
/* out[i] = L2Sqr(a[ia[i]], base[ib[i]]) in FP32 with the reference's 16-lane
 * partial-sum order (d % 16 == 0) or sequential order otherwise. */
int hrnn_dev_l2sqr_pairs(hrnn_dev_t h, const float *a, uint64_t na, const uint32_t *ia,
                         const uint32_t *ib, uint64_t npairs, float *out);

/* ---- synthetic data (host) ---------------------------------------------- */
/* Gaussian mixture: n_centres centres uniform in [lo, hi]^d; each row picks a
 * centre (uniform, or Zipf(zipf_s) when zipf_s > 0), adds sigma * N(0,1) per
 * coordinate, rounds to an integer if round_int, and clips to [lo, hi].
 * Counter-based (Philox4x32-10 + Box-Muller): the same seed gives the same
 * bytes on any host and thread count.  centre_of (may be NULL) gets the centre
 * index of each row. */
int hrnn_synth_mixture(uint64_t seed, uint64_t n, uint32_t d, uint32_t n_centres,
                       float sigma, float lo, float hi, int round_int, double zipf_s,
                       uint64_t centre_seed, float *out, uint32_t *centre_of);
/* |N(0,1)| * s_j with s_j ~ U(0.05, 1) per dimension (scale seeded by scale_seed). */
int hrnn_synth_halfnormal_scaled(uint64_t seed, uint64_t scale_seed, uint64_t n,
                                 uint32_t d, float *out);
/* L2-normalised vMF-like mixture: x = normalise(mu_c + z / sqrt(kappa)). */
int hrnn_synth_sphere_mixture(uint64_t seed, uint64_t centre_seed, uint64_t n, uint32_t d,
                              uint32_t n_centres, float kappa, float *out,
                              uint32_t *centre_of);
/* 0 .. n-1 in the order std::shuffle with std::mt19937(seed) leaves them: the
 * dynamic test's split order (reference src/hrnn_dynamic_test.cpp:200-203). */
int hrnn_synth_split_perm(uint64_t n, uint32_t seed, uint32_t *out);

#ifdef __cplusplus
}
#endif
#endif /* HRNN_DEV_H */
