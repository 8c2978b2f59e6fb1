/*
 * hrnn_oracle.c -- CPU oracle for the HRNN hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP kernels of
 * libhrnn_hip.so: only tests/, __graft_entry__.smoke() and the cpu_baseline leg
 * of bench.py may load it.  The product never calls it and never falls back
 * to it.
 *
 * Each function restates one piece of the reference algorithm in plain C and
 * cites the reference file:line it follows.  How each one is pinned (compiled
 * reference, closed forms, brute force) is written beside it; a function with
 * no such check says "parity unpinned".
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_OK 0
#define ORC_E_ARG -1
#define ORC_E_CAP -5
#define ORC_E_RANGE -6

/* ------------------------------------------------------------------------
 * Squared L2 in the reference's summation order.
 * space_l2.h:25-54 (AVX-512, W=16): element e is accumulated into partial
 * e % 16 in increasing e, then the 16 partials are summed left to right.
 * space_l2.h:60-91 (AVX, W=8): the same with 8 partials.  space_l2.h:6-20
 * (scalar, W=1).  use_fma selects whether diff*diff+acc is one rounding (GCC
 * contracts the intrinsic mul+add under -ffp-contract=fast) or two.
 * Pinned: bit-exact against the compiled reference's L2Space (tests).
 * ---------------------------------------------------------------------- */
float orc_l2sqr(const float *a, const float *b, uint32_t d, int W, int use_fma) {
    float part[16];
    if (W <= 1 || (d % (uint32_t)(W == 8 ? 16 : W)) != 0) {
        float r = 0.f;
        for (uint32_t i = 0; i < d; i++) {
            float t = a[i] - b[i];
            r = use_fma ? fmaf(t, t, r) : r + t * t;
        }
        return r;
    }
    for (int l = 0; l < W; l++) part[l] = 0.f;
    for (uint32_t e = 0; e < d; e++) {
        float t = a[e] - b[e];
        int l = (int)(e % (uint32_t)W);
        part[l] = use_fma ? fmaf(t, t, part[l]) : part[l] + t * t;
    }
    float r = part[0];
    for (int l = 1; l < W; l++) r += part[l];
    return r;
}

double orc_l2sqr_f64(const float *a, const float *b, uint32_t d) {
    double r = 0.0;
    for (uint32_t i = 0; i < d; i++) {
        double t = (double)a[i] - (double)b[i];
        r += t * t;
    }
    return r;
}

/* ---- (dist, id) helpers -------------------------------------------------- */
typedef struct { float d; uint32_t id; } orc_pair;

static int pair_less(orc_pair x, orc_pair y) {
    return x.d < y.d || (x.d == y.d && x.id < y.id);
}

/* Insert p into a sorted array of capacity cap (current size *sz), keeping the
 * cap smallest pairs by (d, id).  This is the order BruteforceSearch keeps:
 * its max-heap over (dist, label) pops the larger label on equal distance
 * (bruteforce.h:106-135). */
static void topk_push(orc_pair *arr, uint32_t *sz, uint32_t cap, orc_pair p) {
    if (*sz == cap) {
        if (!pair_less(p, arr[cap - 1])) return;
        (*sz)--;
    }
    uint32_t i = *sz;
    while (i > 0 && pair_less(p, arr[i - 1])) { arr[i] = arr[i - 1]; i--; }
    arr[i] = p;
    (*sz)++;
}

/* ------------------------------------------------------------------------
 * Exact kNN by linear scan: bruteforce.h:106-135 semantics (k smallest by
 * (dist, id)), distances by orc_l2sqr.  For the base-vs-base case
 * (q == x, nq == n) self is included, as compute_kth_nn_dists does
 * (gen_rknn_gt.py:62-75: topk(k+1) over all points, self at distance 0).
 * Pinned: brute force is its own definition; checked on tiny inputs against
 * an independent numpy sort in the tests.
 * ---------------------------------------------------------------------- */
int orc_knn(const float *x, uint64_t n, const float *q, uint64_t nq, uint32_t d,
            uint32_t k, int W, int use_fma, uint32_t *ids, float *d2) {
    if (!x || !q || !ids || k == 0 || k > n) return ORC_E_ARG;
#pragma omp parallel
    {
        orc_pair *buf = (orc_pair *)malloc(sizeof(orc_pair) * k);
#pragma omp for schedule(dynamic, 16)
        for (int64_t qi = 0; qi < (int64_t)nq; qi++) {
            uint32_t sz = 0;
            const float *qv = q + (size_t)qi * d;
            for (uint64_t i = 0; i < n; i++) {
                orc_pair p = { orc_l2sqr(qv, x + i * d, d, W, use_fma), (uint32_t)i };
                topk_push(buf, &sz, k, p);
            }
            for (uint32_t j = 0; j < k; j++) {
                ids[(size_t)qi * k + j] = buf[j].id;
                if (d2) d2[(size_t)qi * k + j] = buf[j].d;
            }
        }
        free(buf);
    }
    return ORC_OK;
}

/* kth[i] = r_k(i): the (k+1)-th smallest distance from row i over all rows,
 * self included (gen_rknn_gt.py:62-75). */
int orc_kth(const float *x, uint64_t n, uint32_t d, uint32_t k, int W, int use_fma,
            float *kth) {
    uint32_t kk = k + 1 < n ? k + 1 : (uint32_t)n;
    uint32_t *ids = (uint32_t *)malloc(sizeof(uint32_t) * n * kk);
    float *dd = (float *)malloc(sizeof(float) * n * kk);
    int rc = orc_knn(x, n, x, n, d, kk, W, use_fma, ids, dd);
    if (rc == ORC_OK)
        for (uint64_t i = 0; i < n; i++) kth[i] = dd[i * kk + (kk - 1)];
    free(ids);
    free(dd);
    return rc;
}

/* FP64 r_k for the precision study (no reference counterpart: the FP64 judge).
 * The (k+1)-th smallest FP64 distance (orc_l2sqr_f64: sequential, no fused
 * multiply-add) from row i over all rows, self included, counting equal
 * values separately.  Pinned: equal to kth on integer data; against a float64
 * numpy brute force on the precision form's data; an FP32-accumulation slip
 * fails that comparison (tests/test_oracle.py). */
static double kth_f64_row(const float *x, uint64_t n, uint32_t d, uint32_t kk, uint64_t i,
                          double *best) {
    uint32_t sz = 0;
    for (uint64_t j = 0; j < n; j++) {
        double v = orc_l2sqr_f64(x + (size_t)i * d, x + j * d, d);
        if (sz == kk) { if (v >= best[kk - 1]) continue; sz--; }
        uint32_t p = sz;
        while (p > 0 && v < best[p - 1]) { best[p] = best[p - 1]; p--; }
        best[p] = v; sz++;
    }
    return best[kk - 1];
}

int orc_kth_f64(const float *x, uint64_t n, uint32_t d, uint32_t k, double *kth) {
    uint32_t kk = k + 1 < n ? k + 1 : (uint32_t)n;
#pragma omp parallel
    {
        double *best = (double *)malloc(sizeof(double) * kk);
#pragma omp for schedule(dynamic, 16)
        for (int64_t i = 0; i < (int64_t)n; i++) kth[i] = kth_f64_row(x, n, d, kk, (uint64_t)i, best);
        free(best);
    }
    return ORC_OK;
}

/* The same for the listed rows only: kth[t] = r_k(rows[t]) in FP64. */
int orc_kth_f64_rows(const float *x, uint64_t n, uint32_t d, uint32_t k, const uint32_t *rows,
                     uint64_t nrows, double *kth) {
    uint32_t kk = k + 1 < n ? k + 1 : (uint32_t)n;
    for (uint64_t t = 0; t < nrows; t++)
        if (rows[t] >= n) return ORC_E_RANGE;
#pragma omp parallel
    {
        double *best = (double *)malloc(sizeof(double) * kk);
#pragma omp for schedule(dynamic, 1)
        for (int64_t t = 0; t < (int64_t)nrows; t++) kth[t] = kth_f64_row(x, n, d, kk, rows[t], best);
        free(best);
    }
    return ORC_OK;
}

/* ------------------------------------------------------------------------
 * RkNN groundtruth: compute_rknn (gen_rknn_gt.py:82-151): for each query,
 * every base x with d2(q, x) <= kth[x]; sorted by distance.  The script uses
 * np.argsort (unstable) on equal distances; this oracle sorts by (d2, id) and
 * comparisons of result sets ignore order.  Two passes: counts, then fill.
 * Pinned: closed form on integer data (every distance exact), brute force.
 * ---------------------------------------------------------------------- */
static int cmp_pair(const void *a, const void *b) {
    const orc_pair *x = (const orc_pair *)a, *y = (const orc_pair *)b;
    if (pair_less(*x, *y)) return -1;
    if (pair_less(*y, *x)) return 1;
    return 0;
}

int orc_rknn_gt(const float *x, uint64_t n, const float *q, uint64_t nq, uint32_t d,
                const float *kth, int W, int use_fma, uint64_t *off, uint32_t *ids,
                float *d2, uint64_t cap, uint64_t *needed) {
    uint64_t *cnt = (uint64_t *)calloc(nq + 1, sizeof(uint64_t));
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t qi = 0; qi < (int64_t)nq; qi++) {
        uint64_t c = 0;
        for (uint64_t i = 0; i < n; i++)
            if (orc_l2sqr(q + (size_t)qi * d, x + i * d, d, W, use_fma) <= kth[i]) c++;
        cnt[qi] = c;
    }
    off[0] = 0;
    for (uint64_t qi = 0; qi < nq; qi++) off[qi + 1] = off[qi] + cnt[qi];
    free(cnt);
    if (needed) *needed = off[nq];
    if (off[nq] > cap) return ORC_E_CAP;
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t qi = 0; qi < (int64_t)nq; qi++) {
        uint64_t c = off[qi + 1] - off[qi];
        orc_pair *buf = (orc_pair *)malloc(sizeof(orc_pair) * (c ? c : 1));
        uint64_t w = 0;
        for (uint64_t i = 0; i < n; i++) {
            float v = orc_l2sqr(q + (size_t)qi * d, x + i * d, d, W, use_fma);
            if (v <= kth[i]) { buf[w].d = v; buf[w].id = (uint32_t)i; w++; }
        }
        qsort(buf, w, sizeof(orc_pair), cmp_pair);
        for (uint64_t j = 0; j < w; j++) {
            ids[off[qi] + j] = buf[j].id;
            if (d2) d2[off[qi] + j] = buf[j].d;
        }
        free(buf);
    }
    return ORC_OK;
}

/* The FP64 judge of compute_rknn (no reference counterpart; the reference
 * decides in FP32, gen_rknn_gt.py:116-124): every base row x with
 * orc_l2sqr_f64(q, x) <= kth64[x], sorted by (FP64 distance, id).
 * Pinned: equal to rknn_gt on integer data (both exact there); against a
 * float64 numpy brute force on float data, including queries placed within
 * FP32 rounding of a radius where the FP32 decision differs. */
typedef struct { double d; uint32_t id; } orc_pair64;

static int cmp_pair64(const void *a, const void *b) {
    const orc_pair64 *x = (const orc_pair64 *)a, *y = (const orc_pair64 *)b;
    if (x->d != y->d) return x->d < y->d ? -1 : 1;
    return x->id < y->id ? -1 : (x->id > y->id ? 1 : 0);
}

int orc_rknn_gt_f64(const float *x, uint64_t n, const float *q, uint64_t nq, uint32_t d,
                    const double *kth64, uint64_t *off, uint32_t *ids, double *d2, uint64_t cap,
                    uint64_t *needed) {
    uint64_t *cnt = (uint64_t *)calloc(nq + 1, sizeof(uint64_t));
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t qi = 0; qi < (int64_t)nq; qi++) {
        uint64_t c = 0;
        for (uint64_t i = 0; i < n; i++)
            if (orc_l2sqr_f64(q + (size_t)qi * d, x + i * d, d) <= kth64[i]) c++;
        cnt[qi] = c;
    }
    off[0] = 0;
    for (uint64_t qi = 0; qi < nq; qi++) off[qi + 1] = off[qi] + cnt[qi];
    free(cnt);
    if (needed) *needed = off[nq];
    if (off[nq] > cap) return ORC_E_CAP;
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t qi = 0; qi < (int64_t)nq; qi++) {
        uint64_t c = off[qi + 1] - off[qi];
        orc_pair64 *buf = (orc_pair64 *)malloc(sizeof(orc_pair64) * (c ? c : 1));
        uint64_t w = 0;
        for (uint64_t i = 0; i < n; i++) {
            double v = orc_l2sqr_f64(q + (size_t)qi * d, x + i * d, d);
            if (v <= kth64[i]) { buf[w].d = v; buf[w].id = (uint32_t)i; w++; }
        }
        qsort(buf, w, sizeof(orc_pair64), cmp_pair64);
        for (uint64_t j = 0; j < w; j++) {
            ids[off[qi] + j] = buf[j].id;
            if (d2) d2[off[qi] + j] = buf[j].d;
        }
        free(buf);
    }
    return ORC_OK;
}

/* ------------------------------------------------------------------------
 * Reverse-KNN CSR: buildRKNNG (hrnn.h:1869-1909) + buildRKNNG_fill
 * (hrnn.h:1912-1935).  rank_bits = smallest b with 2^b >= K, node_bits =
 * smallest b with 2^b >= n; wide when their sum > 32; error when > 64.
 * Offsets: u64 exclusive scan of in-degree.  Entries (src << rank_bits) |
 * rank.  The reference sorts each list by rank with std::sort (unstable), so
 * the order among equal ranks is libstdc++'s; this oracle emits the
 * canonical (rank, src) order that a stable sort gives, by filling rank by
 * rank (outer) and source by source (inner).
 * Pinned: against the compiled reference's CSR after sorting each equal-rank
 * run by src; invariants (offsets = scan of in-degree, transpose multiset).
 * ---------------------------------------------------------------------- */
int orc_csr_params(uint64_t n, uint32_t K, uint32_t *rank_bits, uint32_t *node_bits,
                   int *wide) {
    uint32_t rb = 0, nb = 0;
    while ((1u << rb) < K) rb++;
    while ((1ull << nb) < n) nb++;
    *rank_bits = rb;
    if (node_bits) *node_bits = nb;
    *wide = (nb + rb > 32);
    if (nb + rb > 64) return ORC_E_RANGE;
    return ORC_OK;
}

int orc_csr(const uint32_t *knn_ids, const uint16_t *cnt, uint64_t n, uint32_t K,
            uint32_t *rank_bits_out, int *wide_out, uint64_t *offsets, void *entries) {
    uint32_t rb;
    int wide;
    int rc = orc_csr_params(n, K, &rb, NULL, &wide);
    if (rc) return rc;
    *rank_bits_out = rb;
    *wide_out = wide;
    memset(offsets, 0, sizeof(uint64_t) * (n + 1));
    for (uint64_t i = 0; i < n; i++)
        for (uint32_t j = 0; j < cnt[i]; j++) offsets[knn_ids[i * K + j] + 1]++;
    for (uint64_t i = 1; i <= n; i++) offsets[i] += offsets[i - 1];
    if (!entries) return ORC_OK;
    uint64_t *pos = (uint64_t *)malloc(sizeof(uint64_t) * (n + 1));
    memcpy(pos, offsets, sizeof(uint64_t) * (n + 1));
    for (uint32_t j = 0; j < K; j++)
        for (uint64_t i = 0; i < n; i++) {
            if (j >= cnt[i]) continue;
            uint32_t dst = knn_ids[i * K + j];
            if (wide)
                ((uint64_t *)entries)[pos[dst]++] = ((uint64_t)i << rb) | j;
            else
                ((uint32_t *)entries)[pos[dst]++] = (uint32_t)((i << rb) | j);
        }
    free(pos);
    return ORC_OK;
}

/* ------------------------------------------------------------------------
 * RkNN search phases B and C for given proxies: searchRknn_impl
 * (hrnn.h:2021-2069).  Phase B: for each proxy, walk its reverse list while
 * rank < K_prime, count each candidate (first sight appends it).  Phase C:
 * keep candidates with count >= threshold, sort ids ascending, accept c if
 * L2Sqr(q, c) <= dist(c, k-1) (FLT_MAX when c has <= k-1 neighbours,
 * getVerifyKdistSq hrnn.h:286-293).  Returns internal ids ascending.
 * Proxies equal to 0xFFFFFFFF are skipped.
 * Pinned: against the compiled reference's searchRknn with the same proxies
 * (tests), and brute force on tiny indexes.
 * ---------------------------------------------------------------------- */
static int cmp_u32(const void *a, const void *b) {
    uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return x < y ? -1 : (x > y);
}

int orc_search_bc(const uint64_t *offsets, const void *entries, int wide, uint32_t rank_bits,
                  const float *knn_d, const uint16_t *cnt, uint32_t K, const float *x,
                  uint64_t n, uint32_t d, const float *q, uint64_t nq,
                  const uint32_t *proxies, uint32_t m, const uint32_t *k_per_q, uint32_t k,
                  uint32_t K_prime, float threshold, int W, int use_fma, uint64_t *res_off,
                  uint32_t *res_ids, uint64_t cap, uint64_t *needed) {
    float *score = (float *)calloc(n, sizeof(float));
    uint32_t *cand = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint64_t mask = wide ? ((1ull << rank_bits) - 1) : (uint64_t)((1u << rank_bits) - 1);
    uint64_t total = 0;
    int overflow = 0;
    res_off[0] = 0;
    for (uint64_t qi = 0; qi < nq; qi++) {
        uint64_t nc = 0;
        for (uint32_t p = 0; p < m; p++) {
            uint32_t px = proxies[qi * m + p];
            if (px == 0xFFFFFFFFu) continue;
            for (uint64_t e = offsets[px]; e < offsets[px + 1]; e++) {
                uint64_t packed = wide ? ((const uint64_t *)entries)[e]
                                       : ((const uint32_t *)entries)[e];
                uint32_t rank = (uint32_t)(packed & mask);
                if (rank >= K_prime) break;
                uint32_t c = (uint32_t)(packed >> rank_bits);
                if (score[c] == 0.0f) cand[nc++] = c;
                score[c] += 1.0f;
            }
        }
        uint64_t w = 0;
        for (uint64_t i = 0; i < nc; i++) {
            uint32_t c = cand[i];
            int keep = score[c] >= threshold;
            score[c] = 0.0f;
            if (keep) cand[w++] = c;
        }
        qsort(cand, w, sizeof(uint32_t), cmp_u32);
        uint32_t kq = k_per_q ? k_per_q[qi] : k;
        uint32_t kidx = kq - 1;
        for (uint64_t i = 0; i < w; i++) {
            uint32_t c = cand[i];
            float dist = orc_l2sqr(q + qi * d, x + (size_t)c * d, d, W, use_fma);
            float kd = (kidx >= cnt[c]) ? FLT_MAX : knn_d[(size_t)c * K + kidx];
            if (dist <= kd) {
                if (total < cap) res_ids[total] = c; else overflow = 1;
                total++;
            }
        }
        res_off[qi + 1] = total;
    }
    free(score);
    free(cand);
    if (needed) *needed = total;
    return overflow ? ORC_E_CAP : ORC_OK;
}

/* ------------------------------------------------------------------------
 * NNDescent refinement, single-threaded: refineKNNG (hrnn.h:1725-1862) with
 * try_update_knng_nnd (hrnn.h:1660-1722).  Sequential order everywhere (node
 * ascending in UPDATE and JOIN), which is what the reference does with one
 * OpenMP thread, so the result is deterministic.
 *   UPDATE (hrnn.h:1765-1810): per node i, per entry j in rank order: a new
 *     entry goes to nn_new[i] while |nn_new[i]| < S and its flag is cleared
 *     (sampled or not); an old entry goes to nn_old[i] while < S; either
 *     pushes i into the matching reverse sample of nbr (capped at R = S) when
 *     dist(i, nbr) > nbr's K-th distance (FLT_MAX unless nbr's list is full).
 *   MERGE (hrnn.h:1813-1819): nn_new += rnn_new; nn_old += rnn_old, cut to 2R.
 *   JOIN (hrnn.h:1823-1853): new x new (i < j) and new x old pairs with
 *     a != b; each distance is offered to both lists.
 * stats_* (may be NULL) receive per-iteration update and distance counts.
 * Pinned: bit-exact against the compiled reference run with one thread.
 * ---------------------------------------------------------------------- */
static int nnd_try_update(uint32_t u, uint32_t w, float dd, uint32_t *ids, float *dist,
                          uint16_t *cnt, uint8_t *isnew, uint32_t K) {
    uint32_t c = cnt[u];
    uint32_t *li = ids + (size_t)u * K;
    float *ld = dist + (size_t)u * K;
    uint8_t *lf = isnew + (size_t)u * K;
    if (c >= K && dd >= ld[c - 1]) return 0;
    for (uint32_t i = 0; i < c; i++)
        if (li[i] == w) return 0;
    uint32_t pos;
    if (c < K) {
        pos = c;
        cnt[u] = (uint16_t)(c + 1);
    } else {
        pos = c - 1;
    }
    li[pos] = w;
    ld[pos] = dd;
    lf[pos] = 1;
    while (pos > 0 && ld[pos] < ld[pos - 1]) {
        uint32_t ti = li[pos]; li[pos] = li[pos - 1]; li[pos - 1] = ti;
        float tf = ld[pos]; ld[pos] = ld[pos - 1]; ld[pos - 1] = tf;
        uint8_t tb = lf[pos]; lf[pos] = lf[pos - 1]; lf[pos - 1] = tb;
        pos--;
    }
    return 1;
}

/* The same with the entries' "new" flags (u8 [n*K]) given and returned:
 * flags NULL starts every listed entry new, as refineKNNG does
 * (hrnn.h:1744-1748); otherwise the iterations start from flags and leave
 * the final flags there, so a run can resume from any state. */
int orc_nndescent_flags(const float *x, uint64_t n, uint32_t d, uint32_t K, uint32_t *ids,
                        float *dist, uint16_t *cnt, int iters, int sample, int W, int use_fma,
                        uint64_t *stats_updates, uint64_t *stats_dists, uint8_t *flags) {
    uint32_t S = (uint32_t)sample, R = S;
    uint8_t *isnew = (uint8_t *)malloc(n * K);
    uint32_t *nn_new = (uint32_t *)malloc(sizeof(uint32_t) * n * (S + R));
    uint32_t *nn_old = (uint32_t *)malloc(sizeof(uint32_t) * n * (S + R));
    uint32_t *rnn_new = (uint32_t *)malloc(sizeof(uint32_t) * n * R);
    uint32_t *rnn_old = (uint32_t *)malloc(sizeof(uint32_t) * n * R);
    uint32_t *c_new = (uint32_t *)malloc(sizeof(uint32_t) * n);
    uint32_t *c_old = (uint32_t *)malloc(sizeof(uint32_t) * n);
    uint32_t *c_rnew = (uint32_t *)malloc(sizeof(uint32_t) * n);
    uint32_t *c_rold = (uint32_t *)malloc(sizeof(uint32_t) * n);
    for (uint64_t i = 0; i < n; i++)
        for (uint32_t j = 0; j < K; j++) isnew[i * K + j] = j < cnt[i] ? (flags ? (flags[i * K + j] != 0) : 1) : 0;
    for (int it = 0; it < iters; it++) {
        uint64_t upd = 0, nd = 0;
        memset(c_new, 0, sizeof(uint32_t) * n);
        memset(c_old, 0, sizeof(uint32_t) * n);
        memset(c_rnew, 0, sizeof(uint32_t) * n);
        memset(c_rold, 0, sizeof(uint32_t) * n);
        for (uint64_t i = 0; i < n; i++) {
            uint32_t c = cnt[i];
            for (uint32_t j = 0; j < c; j++) {
                uint32_t nbr = ids[i * K + j];
                float dij = dist[i * K + j];
                int is_n = isnew[i * K + j];
                uint32_t cn = cnt[nbr];
                float worst = (cn >= K) ? dist[(size_t)nbr * K + cn - 1] : FLT_MAX;
                if (is_n) {
                    if (c_new[i] < S) nn_new[i * (S + R) + c_new[i]++] = nbr;
                    isnew[i * K + j] = 0;
                    if (cn > 0 && dij > worst && c_rnew[nbr] < R)
                        rnn_new[(size_t)nbr * R + c_rnew[nbr]++] = (uint32_t)i;
                } else {
                    if (c_old[i] < S) nn_old[i * (S + R) + c_old[i]++] = nbr;
                    if (cn > 0 && dij > worst && c_rold[nbr] < R)
                        rnn_old[(size_t)nbr * R + c_rold[nbr]++] = (uint32_t)i;
                }
            }
        }
        for (uint64_t i = 0; i < n; i++) {
            for (uint32_t r = 0; r < c_rnew[i]; r++)
                nn_new[i * (S + R) + c_new[i]++] = rnn_new[i * R + r];
            for (uint32_t r = 0; r < c_rold[i]; r++)
                nn_old[i * (S + R) + c_old[i]++] = rnn_old[i * R + r];
            if (c_old[i] > 2 * R) c_old[i] = 2 * R;
        }
        for (uint64_t node = 0; node < n; node++) {
            const uint32_t *nw = nn_new + node * (S + R);
            const uint32_t *ol = nn_old + node * (S + R);
            uint32_t a_n = c_new[node], o_n = c_old[node];
            for (uint32_t i = 0; i < a_n; i++)
                for (uint32_t j = i + 1; j < a_n; j++) {
                    uint32_t a = nw[i], b = nw[j];
                    if (a == b) continue;
                    float dd = orc_l2sqr(x + (size_t)a * d, x + (size_t)b * d, d, W, use_fma);
                    nd++;
                    upd += nnd_try_update(a, b, dd, ids, dist, cnt, isnew, K);
                    upd += nnd_try_update(b, a, dd, ids, dist, cnt, isnew, K);
                }
            for (uint32_t i = 0; i < a_n; i++)
                for (uint32_t j = 0; j < o_n; j++) {
                    uint32_t a = nw[i], b = ol[j];
                    if (a == b) continue;
                    float dd = orc_l2sqr(x + (size_t)a * d, x + (size_t)b * d, d, W, use_fma);
                    nd++;
                    upd += nnd_try_update(a, b, dd, ids, dist, cnt, isnew, K);
                    upd += nnd_try_update(b, a, dd, ids, dist, cnt, isnew, K);
                }
        }
        if (stats_updates) stats_updates[it] = upd;
        if (stats_dists) stats_dists[it] = nd;
    }
    if (flags) memcpy(flags, isnew, n * K);
    free(isnew); free(nn_new); free(nn_old); free(rnn_new); free(rnn_old);
    free(c_new); free(c_old); free(c_rnew); free(c_rold);
    return ORC_OK;
}

int orc_nndescent(const float *x, uint64_t n, uint32_t d, uint32_t K, uint32_t *ids,
                  float *dist, uint16_t *cnt, int iters, int sample, int W, int use_fma,
                  uint64_t *stats_updates, uint64_t *stats_dists) {
    return orc_nndescent_flags(x, n, d, K, ids, dist, cnt, iters, sample, W, use_fma, stats_updates,
                               stats_dists, NULL);
}

/* ------------------------------------------------------------------------
 * Phases 2-5 of insertPointDynamic (reference include/hnswlib/hrnn.h:2366-2532
 * with try_update_knng_evict :1592-1655) for the rows [b0, b1) of a ranked
 * KNN graph whose rows b0 .. b1-1 already hold their seed lists (phase 1,
 * addPoint).  The reverse lists are the transpose of the graph as given
 * (what buildRKNNG_mutable builds and phases 2 and 5 keep up to date when
 * bidir is 0).  Per new row x, in order: candidates = the ids on the
 * reverse lists of the first m_update entries of x's list (all when 0), each
 * once, x and every entry of x's list excluded; x is offered to each
 * candidate c at d = L2(x, c) when c's list (as given) is not full or
 * d < its last distance.  With bidir, x's list becomes the K smallest
 * (d, id) of its list and its candidates (x itself skipped).  Then every
 * list takes its offers: the K first of (list U offers) in distance order,
 * listed entries ahead of equal offers, offers by (d, id), offers of an id
 * already listed dropped.  For b1 = b0 + 1 this is the reference's step
 * for one point: each candidate gets at most one offer, and a list that
 * takes it replaces its last entry (or appends) and moves it up past every
 * strictly larger distance, as the evict insertion sort does.
 * Pinned: for one point at a time, the graph after each of the compiled
 * reference's insertPointDynamic calls (single-threaded, bidir 0, m_update
 * 0 and > 0) and, with bidir 1, after its first call (later calls read
 * reverse entries the reference leaves stale) (tests).
 * ---------------------------------------------------------------------- */
typedef struct {
    uint32_t tgt;
    float d;
    uint32_t id;
} orc_offer;

static int offer_cmp(const void *a, const void *b) {
    const orc_offer *x = (const orc_offer *)a, *y = (const orc_offer *)b;
    if (x->tgt != y->tgt) return x->tgt < y->tgt ? -1 : 1;
    if (x->d != y->d) return x->d < y->d ? -1 : 1;
    if (x->id != y->id) return x->id < y->id ? -1 : 1;
    return 0;
}

typedef struct {
    float d;
    uint32_t id;
} orc_key;

static int key_cmp(const void *a, const void *b) {
    const orc_key *x = (const orc_key *)a, *y = (const orc_key *)b;
    if (x->d != y->d) return x->d < y->d ? -1 : 1;
    if (x->id != y->id) return x->id < y->id ? -1 : 1;
    return 0;
}

int orc_insert_affect(const float *x, uint64_t n, uint32_t d, uint32_t *ids, float *dist, uint16_t *cnt,
                      uint32_t K, uint64_t b0, uint64_t b1, uint32_t m_update, int bidir, int W, int use_fma,
                      uint64_t *stats) {
    if (b0 >= b1 || b1 > n || K == 0) return ORC_E_ARG;
    /* reverse lists: transpose of the graph as given */
    uint64_t *roff = (uint64_t *)calloc(n + 1, sizeof(uint64_t));
    for (uint64_t i = 0; i < n; i++)
        for (uint32_t j = 0; j < cnt[i]; j++) roff[ids[i * K + j] + 1]++;
    for (uint64_t i = 0; i < n; i++) roff[i + 1] += roff[i];
    uint32_t *rent = (uint32_t *)malloc(sizeof(uint32_t) * (roff[n] + 1));
    uint64_t *fill = (uint64_t *)malloc(sizeof(uint64_t) * n);
    memcpy(fill, roff, sizeof(uint64_t) * n);
    for (uint64_t i = 0; i < n; i++)
        for (uint32_t j = 0; j < cnt[i]; j++) rent[fill[ids[i * K + j]]++] = (uint32_t)i;
    free(fill);
    uint8_t *seen = (uint8_t *)calloc(n, 1);
    uint32_t *cand = (uint32_t *)malloc(sizeof(uint32_t) * (roff[n] + 1));
    uint64_t ocap = 1024, no = 0, scanned = 0, ncand = 0;
    orc_offer *off = (orc_offer *)malloc(sizeof(orc_offer) * ocap);
    uint32_t *bid = bidir ? (uint32_t *)malloc(sizeof(uint32_t) * (b1 - b0) * K) : NULL;
    float *bd = bidir ? (float *)malloc(sizeof(float) * (b1 - b0) * K) : NULL;
    uint16_t *bc = bidir ? (uint16_t *)malloc(sizeof(uint16_t) * (b1 - b0)) : NULL;
    for (uint64_t xn = b0; xn < b1; xn++) {
        const uint32_t c0 = cnt[xn];
        const uint32_t np = (m_update > 0 && m_update < c0) ? m_update : c0;
        uint64_t nc = 0;
        seen[xn] = 1;
        for (uint32_t j = 0; j < c0; j++) seen[ids[xn * K + j]] = 1;
        for (uint32_t p = 0; p < np; p++) {
            const uint32_t y = ids[xn * K + p];
            for (uint64_t e = roff[y]; e < roff[y + 1]; e++) {
                const uint32_t c = rent[e];
                scanned++;
                if (!seen[c]) {
                    seen[c] = 1;
                    cand[nc++] = c;
                }
            }
        }
        seen[xn] = 0;
        for (uint32_t j = 0; j < c0; j++) seen[ids[xn * K + j]] = 0;
        for (uint64_t t = 0; t < nc; t++) seen[cand[t]] = 0;
        ncand += nc;
        orc_key *heap = bidir ? (orc_key *)malloc(sizeof(orc_key) * (c0 + nc + 1)) : NULL;
        uint64_t nh = 0;
        if (bidir)
            for (uint32_t j = 0; j < c0; j++) heap[nh++] = (orc_key){dist[xn * K + j], ids[xn * K + j]};
        for (uint64_t t = 0; t < nc; t++) {
            const uint32_t c = cand[t];
            const float dd = orc_l2sqr(x + xn * d, x + (uint64_t)c * d, d, W, use_fma);
            if (bidir) heap[nh++] = (orc_key){dd, c};
            const uint32_t cc = cnt[c];
            if (cc >= K && !(dd < dist[(uint64_t)c * K + cc - 1])) continue;
            if (no == ocap) {
                ocap *= 2;
                off = (orc_offer *)realloc(off, sizeof(orc_offer) * ocap);
            }
            off[no++] = (orc_offer){c, dd, (uint32_t)xn};
        }
        if (bidir) {
            qsort(heap, nh, sizeof(orc_key), key_cmp);
            uint32_t w = 0;
            for (uint64_t t = 0; t < nh && w < K; t++) {
                if (heap[t].id == (uint32_t)xn) continue;
                if (t > 0 && heap[t].id == heap[t - 1].id && heap[t].d == heap[t - 1].d) continue;
                bid[(xn - b0) * K + w] = heap[t].id;
                bd[(xn - b0) * K + w] = heap[t].d;
                w++;
            }
            bc[xn - b0] = (uint16_t)w;
            free(heap);
        }
    }
    if (bidir)
        for (uint64_t xn = b0; xn < b1; xn++) {
            memcpy(ids + xn * K, bid + (xn - b0) * K, sizeof(uint32_t) * bc[xn - b0]);
            memcpy(dist + xn * K, bd + (xn - b0) * K, sizeof(float) * bc[xn - b0]);
            cnt[xn] = bc[xn - b0];
        }
    /* merge the offers per list */
    qsort(off, no, sizeof(orc_offer), offer_cmp);
    uint64_t upd = 0;
    uint32_t *mi = (uint32_t *)malloc(sizeof(uint32_t) * (K + no + 1));
    float *md = (float *)malloc(sizeof(float) * (K + no + 1));
    for (uint64_t a = 0; a < no;) {
        uint64_t e = a;
        const uint32_t c = off[a].tgt;
        while (e < no && off[e].tgt == c) e++;
        const uint32_t lc = cnt[c];
        uint32_t i = 0, m = 0;
        uint64_t j = a;
        while ((i < lc || j < e) && m < K) {
            if (j < e) {   /* an offer whose id is listed (anywhere) is dropped */
                int listed = 0;
                for (uint32_t t = 0; t < lc; t++)
                    if (ids[(uint64_t)c * K + t] == off[j].id) listed = 1;
                if (listed) {
                    j++;
                    continue;
                }
            }
            if (j >= e || (i < lc && dist[(uint64_t)c * K + i] <= off[j].d)) {
                mi[m] = ids[(uint64_t)c * K + i];
                md[m] = dist[(uint64_t)c * K + i];
                i++;
            } else {
                mi[m] = off[j].id;
                md[m] = off[j].d;
                j++;
                upd++;
            }
            m++;
        }
        memcpy(ids + (uint64_t)c * K, mi, sizeof(uint32_t) * m);
        memcpy(dist + (uint64_t)c * K, md, sizeof(float) * m);
        cnt[c] = (uint16_t)m;
        a = e;
    }
    if (stats) {
        stats[0] = scanned;
        stats[1] = ncand;
        stats[2] = no;
        stats[3] = upd;
    }
    free(mi); free(md); free(off); free(cand); free(seen); free(roff); free(rent);
    free(bid); free(bd); free(bc);
    return ORC_OK;
}

/* ------------------------------------------------------------------------
 * KNN-graph recall against exact neighbour lists: for each node, the
 * fraction of its first kk exact neighbours (self excluded) present in its
 * first kk list entries.  No reference counterpart (quality measure).
 * ---------------------------------------------------------------------- */
double orc_knng_recall(const uint32_t *ids, const uint16_t *cnt, uint32_t K,
                       const uint32_t *exact, uint32_t kexact, uint64_t n, uint32_t kk) {
    uint64_t hit = 0, tot = 0;
    for (uint64_t i = 0; i < n; i++) {
        uint32_t lim = cnt[i] < kk ? cnt[i] : kk;
        for (uint32_t a = 0; a < kk && a < kexact; a++) {
            uint32_t e = exact[i * kexact + a];
            for (uint32_t b = 0; b < lim; b++)
                if (ids[i * K + b] == e) { hit++; break; }
            tot++;
        }
    }
    return tot ? (double)hit / (double)tot : 0.0;
}

/* Macro RkNN recall, compute_rknn_recall (utils.h:135-157): mean over queries
 * with non-empty groundtruth of |R intersect GT| / |GT|. */
double orc_rknn_recall(const uint64_t *gt_off, const uint32_t *gt_ids, const uint64_t *r_off,
                       const uint32_t *r_ids, uint64_t nq, uint64_t n, double *micro) {
    uint8_t *mark = (uint8_t *)calloc(n, 1);
    double sum = 0.0;
    uint64_t counted = 0, found_all = 0, total_gt = 0;
    for (uint64_t q = 0; q < nq; q++) {
        uint64_t g = gt_off[q + 1] - gt_off[q];
        if (g == 0) continue;
        for (uint64_t i = gt_off[q]; i < gt_off[q + 1]; i++) mark[gt_ids[i]] = 1;
        uint64_t f = 0;
        for (uint64_t i = r_off[q]; i < r_off[q + 1]; i++) f += mark[r_ids[i]];
        for (uint64_t i = gt_off[q]; i < gt_off[q + 1]; i++) mark[gt_ids[i]] = 0;
        sum += (double)f / (double)g;
        counted++;
        found_all += f;
        total_gt += g;
    }
    free(mark);
    if (micro) *micro = total_gt ? (double)found_all / (double)total_gt : 0.0;
    return counted ? sum / (double)counted : 0.0;
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_threads(int t) {
#ifdef _OPENMP
    if (t > 0) omp_set_num_threads(t);
#else
    (void)t;
#endif
}
