// beam.hip -- RkNN phase A on the GPU: HNSW greedy descent over the upper
// layers, then the layer-0 beam of width ef = max(ef_search, m), keeping the m
// closest nodes as proxies.
//
// Reference: searchRknn_impl phase A (include/hnswlib/hrnn.h:1992-2019) with
// searchBaseLayerST<true> (hrnn.h:485-616).  The result list is kept sorted
// by (dist, id); a node enters it when the list holds fewer than ef nodes or
// its distance is strictly below the worst; the closest unexpanded entry is
// expanded next; the search ends when every entry is expanded (= the closest
// open candidate is farther than the worst result).  Distances use the
// reference's 16-partial FMA order, so GPU and host beams agree up to the
// order of equal distances.
//
// k_phase_a: one query per workgroup of one wave (64 lanes), no state kept
// between workgroups.  The sorted list lives in LDS (double buffered); visited
// nodes go into an LDS hash set of 4096 ids (linear probing).  When a probe
// run is exhausted the node is treated as unseen and admission checks the
// list for it, so the result does not depend on the table size.  Per
// expansion the fresh neighbours' distances are computed by 16-lane groups,
// the admissible ones sorted across the wave and merged into the list in
// parallel (the first ef of the merged (dist, id) order: the same set a
// sequential insertion keeps).  When every base and query value is an integer
// in [0, 255] and d <= 258, every distance sum is an exact integer, so rows
// are read with 16-byte loads and summed in any order; otherwise the
// reference's 16-partial order is used.  Every id or count read from the graph
// is range-checked; a bad one is skipped and raises the handle's device error
// word (reported by hrnn_dev_sync).  The descent and the beam run bounded
// loops.
#include <algorithm>
#include <cfloat>

#include "handle.h"

namespace hrnn {

constexpr int kBeamMaxEf = 512;
constexpr int kBeamMaxDeg = 64;
constexpr uint32_t kVisSlots = 4096;
constexpr uint32_t kVisProbes = 32;

// Insert id into the LDS set; true when the id was not there (or could not be
// recorded: *lost is set then).
__device__ __forceinline__ bool vis_insert(uint32_t *tab, uint32_t id, uint32_t *lost) {
    uint32_t h = (id * 2654435761u) >> 20;   // 12 bits -> [0, 4096)
    for (uint32_t p = 0; p < kVisProbes; p++) {
        uint32_t old = atomicCAS(&tab[h], kInvalidId, id);
        if (old == kInvalidId) return true;
        if (old == id) return false;
        h = (h + 1) & (kVisSlots - 1);
    }
    *lost = 1u;
    return true;
}

template <class VT>
__device__ __forceinline__ float beam_dist(bool exact, const float *qv, const VT *x, uint64_t id, uint32_t d,
                                           uint32_t g) {
    return exact ? dist_row_exact<VT>(qv, x, id, d, g) : dist_row<VT>(qv, x, id, d, g);
}

__device__ __forceinline__ unsigned long long beam_key(float d, uint32_t id) {
    return ((unsigned long long)float_key(d) << 32) | id;
}

// Ascending bitonic sort of one 64-bit key per lane across the wave.
__device__ __forceinline__ unsigned long long wave_sort64(unsigned long long key, uint32_t lane) {
    for (uint32_t k = 2; k <= 64; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            const uint32_t lo = (uint32_t)key, hi = (uint32_t)(key >> 32);
            const unsigned long long other =
                ((unsigned long long)(uint32_t)__shfl_xor((int)hi, (int)j, 64) << 32) |
                (uint32_t)__shfl_xor((int)lo, (int)j, 64);
            const bool up = (lane & k) == 0;
            const bool keep_min = ((lane & j) == 0) == up;
            key = keep_min ? (key < other ? key : other) : (key > other ? key : other);
        }
    }
    return key;
}

template <class VT>
__global__ __launch_bounds__(64) void k_phase_a(
    const VT *__restrict__ x, uint32_t d, uint32_t n, const uint32_t *__restrict__ l0,
    uint32_t l0_words, const uint32_t *__restrict__ up_off, const uint32_t *__restrict__ up_pool,
    uint64_t pool_words, uint32_t up_words, uint32_t entry, int maxlevel, const float *__restrict__ q,
    uint32_t ef, uint32_t m, int base_exact, uint32_t *__restrict__ proxies,
    uint32_t *__restrict__ expansions, float *__restrict__ raw, uint32_t *__restrict__ err) {
    __shared__ float s_d[2][kBeamMaxEf];
    __shared__ uint32_t s_i[2][kBeamMaxEf];
    __shared__ uint8_t s_x[2][kBeamMaxEf];   // expanded flag
    __shared__ float s_nd[kBeamMaxDeg];
    __shared__ uint32_t s_ni[kBeamMaxDeg];
    __shared__ unsigned long long s_ck[kBeamMaxDeg];   // sorted candidate keys of one expansion
    __shared__ uint32_t s_vis[kVisSlots];
    __shared__ uint32_t s_lost, s_bad;
    extern __shared__ float s_qv[];

    const uint32_t lane = threadIdx.x;
    const uint32_t g = lane & 15u, grp = lane >> 4;
    const uint64_t qi = blockIdx.x;
    bool q_u8 = true;
    for (uint32_t e = lane; e < d; e += 64) {
        const float v = q[qi * d + e];
        s_qv[e] = v;
        q_u8 &= (v == rintf(v)) && v >= 0.f && v <= 255.f;
    }
    // exact sums for this query: every base and query value an integer in [0, 255]
    const bool exact = base_exact && __ballot(!q_u8) == 0;
    for (uint32_t t = lane; t < kVisSlots; t += 64) s_vis[t] = kInvalidId;
    if (lane == 0) {
        s_lost = 0;
        s_bad = 0;
    }
    __syncthreads();

    // ---- greedy descent over the upper layers ----
    uint32_t cur = entry < n ? entry : 0;
    float cd = __shfl(beam_dist<VT>(exact, s_qv, x, cur, d, g), 0, 64);
    const float d_entry = cd;
    for (int level = maxlevel; level > 0; level--) {
        for (uint32_t step = 0; step < n; step++) {   // each move strictly lowers cd
            const uint32_t uo = up_off[cur];
            const uint64_t base = (uint64_t)uo + (uint64_t)(level - 1) * up_words;
            if (uo == kInvalidId || base + up_words > pool_words) {
                if (lane == 0) s_bad = 1u;
                break;
            }
            const uint32_t *ll = up_pool + base;
            const uint32_t cnt = min(ll[0] & 0xFFFFu, up_words - 1);
            float best = cd;
            uint32_t best_id = cur;
            for (uint32_t b = 0; b < cnt; b += 4) {
                const uint32_t t = b + grp;
                float dv = FLT_MAX;
                uint32_t id = cur;
                if (t < cnt) {   // uniform per 16-lane group
                    id = ll[1 + t];
                    if (id < n) {
                        dv = beam_dist<VT>(exact, s_qv, x, id, d, g);
                    } else {
                        if (g == 0) s_bad = 1u;
                        id = cur;
                    }
                }
                // the host's sequential scan: the first strictly smaller wins
                for (uint32_t s = 0; s < 4; s++) {
                    float ds = __shfl(dv, (int)(s * 16), 64);
                    uint32_t is = __shfl(id, (int)(s * 16), 64);
                    if (b + s < cnt && ds < best) {
                        best = ds;
                        best_id = is;
                    }
                }
            }
            if (best_id == cur) break;
            cur = best_id;
            cd = best;
        }
    }

    // ---- layer-0 beam ----
    const float d_start = cd;
    if (lane == 0) {
        s_d[0][0] = cd;
        s_i[0][0] = cur;
        s_x[0][0] = 0;
        vis_insert(s_vis, cur, &s_lost);
    }
    __syncthreads();
    uint32_t L = 1, buf = 0, n_exp = 0, n_eval = 1;
    for (uint32_t iter = 0; iter <= n; iter++) {   // each pass expands one list entry
        uint32_t pick = kInvalidId;
        for (uint32_t b0 = 0; b0 < L; b0 += 64) {
            const uint32_t t = b0 + lane;
            const uint64_t mask = __ballot(t < L && s_x[buf][t] == 0);
            if (mask) {
                pick = b0 + (uint32_t)__ffsll((long long)mask) - 1;
                break;
            }
        }
        if (pick == kInvalidId) break;
        const uint32_t c = s_i[buf][pick];
        __syncthreads();
        if (lane == 0) s_x[buf][pick] = 1;
        n_exp++;
        const uint32_t *ll = l0 + (uint64_t)c * l0_words;
        const uint32_t cnt = min(ll[0] & 0xFFFFu, l0_words - 1);
        uint32_t nb = 0;
        bool fresh = false;
        if (lane < cnt) {
            nb = ll[1 + lane];
            if (nb < n) fresh = vis_insert(s_vis, nb, &s_lost);
            else s_bad = 1u;
        }
        const uint64_t fm = __ballot(fresh);
        const uint32_t nf = (uint32_t)__popcll(fm);
        n_eval += nf;
        if (fresh) s_ni[__popcll(fm & lanemask_lt())] = nb;
        __syncthreads();
        // distances of the fresh neighbours, 4 at a time (one per 16-lane group)
        for (uint32_t b = 0; b < nf; b += 4) {
            const uint32_t t = b + grp;
            if (t < nf) {
                const float dv = beam_dist<VT>(exact, s_qv, x, s_ni[t], d, g);
                if (g == 0) s_nd[t] = dv;
            }
        }
        __syncthreads();
        if (nf == 0) continue;
        // candidates that can enter: the list is not full, or (dist, id) below
        // its worst entry; not already listed (checked only when the visited
        // set lost an insertion)
        unsigned long long key = ~0ull;
        if (lane < nf) {
            const float dv = s_nd[lane];
            const uint32_t id = s_ni[lane];
            bool ok = L < ef || pair_less(dv, id, s_d[buf][L - 1], s_i[buf][L - 1]);
            if (ok && s_lost) {
                for (uint32_t j = 0; j < L && ok; j++) ok = s_i[buf][j] != id;
            }
            if (ok) key = beam_key(dv, id);
        }
        const uint32_t nv = (uint32_t)__popcll(__ballot(key != ~0ull));
        if (nv == 0) continue;   // uniform
        key = wave_sort64(key, lane);
        s_ck[lane] = key;
        __syncthreads();
        // merge the sorted candidates into the sorted list, keeping the first ef
        const uint32_t nbuf = buf ^ 1u;
        for (uint32_t e = lane; e < L; e += 64) {
            const unsigned long long ke = beam_key(s_d[buf][e], s_i[buf][e]);
            uint32_t lo = 0, hi = nv;   // candidates below entry e
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (s_ck[mid] < ke) lo = mid + 1; else hi = mid;
            }
            const uint32_t pos = e + lo;
            if (pos < ef) {
                s_d[nbuf][pos] = s_d[buf][e];
                s_i[nbuf][pos] = s_i[buf][e];
                s_x[nbuf][pos] = s_x[buf][e];
            }
        }
        if (lane < nv) {
            const unsigned long long kc = s_ck[lane];
            uint32_t lo = 0, hi = L;    // list entries below candidate lane
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (beam_key(s_d[buf][mid], s_i[buf][mid]) < kc) lo = mid + 1; else hi = mid;
            }
            const uint32_t pos = lane + lo;
            if (pos < ef) {
                s_d[nbuf][pos] = key_float((uint32_t)(kc >> 32));
                s_i[nbuf][pos] = (uint32_t)kc;
                s_x[nbuf][pos] = 0;
            }
        }
        __syncthreads();
        L = min(L + nv, ef);
        buf = nbuf;
    }
    // ---- output: the m closest ----
    for (uint32_t t = lane; t < m; t += 64) proxies[qi * m + t] = (t < L) ? s_i[buf][t] : kInvalidId;
    if (expansions && lane == 0) {
        expansions[2 * qi] = n_exp;
        expansions[2 * qi + 1] = n_eval;
    }
    if (raw) {   // per-query statistics for the learned m (kRawWords floats)
        float *r = raw + qi * kRawWords;
        if (lane < 16) r[2 + lane] = (lane < L) ? s_d[buf][lane] : s_d[buf][L - 1];
        if (lane == 0) {
            r[0] = d_entry;
            r[1] = d_start;
            r[18] = (float)n_exp;
            r[19] = (float)L;
        }
    }
    if (lane == 0 && s_bad) atomicOr(err, 1u);
}

void beam_launch(hrnn_dev_t h, const float *q_dev, uint32_t nq, uint32_t m, uint32_t ef,
                 uint32_t *proxies_dev, uint32_t *expansions_dev, float *raw_dev) {
    REQUIRE(h->hnsw_valid, HRNN_E_STATE, "proxies: upload the HNSW graph first");
    REQUIRE(ef >= 1 && ef <= (uint32_t)kBeamMaxEf, HRNN_E_ARG, "proxies: ef must be in [1, 512]");
    REQUIRE(m >= 1 && m <= ef, HRNN_E_ARG, "proxies: m must be in [1, ef]");
    REQUIRE(h->l0_words >= 1 && h->l0_words - 1 <= (uint32_t)kBeamMaxDeg, HRNN_E_ARG,
            "proxies: layer-0 degree must be in [0, 64]");
    REQUIRE(h->d >= 1 && h->d <= 4096, HRNN_E_ARG, "proxies: d must be in [1, 4096]");
    REQUIRE(q_dev && proxies_dev, HRNN_E_ARG, "proxies: null argument");
    if (nq == 0) return;
    uint32_t *err = h->counters.as<uint32_t>() + kErrWord;
    // every sum exact in any order: values in [0, 255] and d <= 258, with the
    // row width the wide loads need (8 BF16 or 4 FP32 elements)
    const int base_exact = (h->base_u8 && h->d <= 258 && h->d % 8 == 0) ? 1 : 0;
#define LAUNCH(VT, XP)                                                                            \
    hipLaunchKernelGGL(k_phase_a<VT>, dim3(nq), dim3(64), h->d * sizeof(float), h->stream, XP, h->d, \
                       (uint32_t)h->n, h->l0.as<uint32_t>(), h->l0_words, h->up_off.as<uint32_t>(),   \
                       h->up_pool.as<uint32_t>(), h->up_pool_words, h->up_words, h->entry,            \
                       h->maxlevel, q_dev, ef, m, base_exact, proxies_dev, expansions_dev, raw_dev, err)
    // BF16 mirror where it is exact (same distances, half the bytes)
    if (h->small_int && h->vec_bf16)
        LAUNCH(__hip_bfloat16, h->xb.as<__hip_bfloat16>());
    else
        LAUNCH(float, h->x.as<float>());
#undef LAUNCH
    HIPC(hipGetLastError());
}

}  // namespace hrnn

using namespace hrnn;

extern "C" {

int hrnn_dev_set_hnsw(hrnn_dev_t h, uint32_t maxM0, uint32_t maxM, const uint32_t *l0,
                      const int *levels, const uint32_t *up_off, const uint32_t *up_pool,
                      uint64_t pool_words, uint32_t entry, int maxlevel) {
    if (!h) return HRNN_E_ARG;
    return guarded(h, [&] {
        REQUIRE(h->n > 0, HRNN_E_STATE, "set_hnsw: set_base first");
        REQUIRE(l0 && levels && up_off && maxM0 >= 1 && maxM0 <= (uint32_t)kBeamMaxDeg && maxM >= 1,
                HRNN_E_ARG, "set_hnsw: bad arguments");
        REQUIRE(entry < h->n && maxlevel >= 0 && levels[entry] == maxlevel, HRNN_E_ARG,
                "set_hnsw: bad entry point");
        const uint64_t n = h->n;
        const uint32_t w0 = maxM0 + 1, wu = maxM + 1;
        // Validate every access the kernel can make: layer-0 ids < n; a node of
        // level L >= 1 owns L blocks inside the pool; a link at layer l points
        // to a node whose level is >= l.
        for (uint64_t i = 0; i < n; i++) {
            const uint32_t *ll = l0 + i * w0;
            uint32_t c = ll[0] & 0xFFFFu;
            REQUIRE(c <= maxM0, HRNN_E_ARG, "set_hnsw: layer-0 count > maxM0");
            for (uint32_t t = 0; t < c; t++) REQUIRE(ll[1 + t] < n, HRNN_E_ARG, "set_hnsw: link id >= n");
            int lv = levels[i];
            REQUIRE(lv >= 0 && lv <= maxlevel, HRNN_E_ARG, "set_hnsw: level out of range");
            if (lv == 0) continue;
            REQUIRE(up_pool && up_off[i] != kInvalidId && (uint64_t)up_off[i] + (uint64_t)lv * wu <= pool_words,
                    HRNN_E_ARG, "set_hnsw: upper block out of pool");
            for (int l = 1; l <= lv; l++) {
                const uint32_t *ul = up_pool + up_off[i] + (uint64_t)(l - 1) * wu;
                uint32_t uc = ul[0] & 0xFFFFu;
                REQUIRE(uc <= maxM, HRNN_E_ARG, "set_hnsw: upper count > maxM");
                for (uint32_t t = 0; t < uc; t++)
                    REQUIRE(ul[1 + t] < n && levels[ul[1 + t]] >= l, HRNN_E_ARG,
                            "set_hnsw: upper link to a node below that layer");
            }
        }
        h->l0.reserve(n * w0 * 4);
        h->up_off.reserve(n * 4);
        h->up_pool.reserve(std::max<uint64_t>(pool_words, 1) * 4);
        h2d(h, h->l0.p, l0, n * w0 * 4);
        h2d(h, h->up_off.p, up_off, n * 4);
        if (pool_words) h2d(h, h->up_pool.p, up_pool, pool_words * 4);
        h->l0_words = w0;
        h->up_words = wu;
        h->up_pool_words = pool_words;
        h->entry = entry;
        h->maxlevel = maxlevel;
        h->hnsw_valid = true;
    });
}

int hrnn_dev_proxies(hrnn_dev_t h, const float *q, uint32_t nq, uint32_t m, uint32_t ef,
                     uint32_t *proxies) {
    if (!h) return HRNN_E_ARG;
    return guarded(h, [&] {
        REQUIRE(q && proxies, HRNN_E_ARG, "proxies: null argument");
        if (!nq) return;
        DevBuf dq, dp;
        dq.reserve((uint64_t)nq * h->d * 4);
        dp.reserve((uint64_t)nq * m * 4);
        h2d(h, dq.p, q, (uint64_t)nq * h->d * 4);
        beam_launch(h, dq.as<float>(), nq, m, ef, dp.as<uint32_t>(), nullptr);
        d2h(h, proxies, dp.p, (uint64_t)nq * m * 4);
        dq.release();
        dp.release();
        check_device_error(h);
    });
}

int hrnn_dev_proxies_dptr(hrnn_dev_t h, const float *q_dev, uint32_t nq, uint32_t m, uint32_t ef,
                          uint32_t *proxies_dev, uint32_t *expansions_dev) {
    if (!h) return HRNN_E_ARG;
    return guarded(h, [&] { beam_launch(h, q_dev, nq, m, ef, proxies_dev, expansions_dev); });
}

int hrnn_dev_rknn_query_dptr(hrnn_dev_t h, const float *q_dev, uint32_t nq, uint32_t m,
                             uint32_t ef, const uint32_t *k_per_q_dev, uint32_t k,
                             uint32_t K_prime, float threshold, uint32_t *proxies_dev,
                             uint32_t *res_cnt_dev, uint32_t *res_ids_dev, uint32_t cap_per_q,
                             uint32_t *work_dev) {
    if (!h) return HRNN_E_ARG;
    return guarded(h, [&] {
        REQUIRE(h->csr_valid && h->K > 0, HRNN_E_STATE, "rknn_query: build the CSR first");
        REQUIRE(k_per_q_dev || (k >= 1 && k <= h->K), HRNN_E_ARG, "rknn_query: k out of [1, K]");
        if (nq == 0) return;
        beam_launch(h, q_dev, nq, m, std::max(ef, m), proxies_dev, nullptr);
        rknn_search_resident(h, q_dev, nq, proxies_dev, m, k_per_q_dev, k, K_prime, threshold,
                             res_cnt_dev, res_ids_dev, cap_per_q, work_dev);
    });
}

}  // extern "C"
