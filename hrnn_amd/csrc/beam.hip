// beam.hip -- RkNN phase A on the GPU: HNSW greedy descent over the upper
// layers, then the layer-0 beam of width ef = max(ef_search, m), keeping the m
// closest nodes as proxies.
//
// Reference: searchRknn_impl phase A (include/hnswlib/hrnn.h:1992-2019) with
// searchBaseLayerST<true> (hrnn.h:485-616).  Ordering rules are those of the
// host beam of this build (include/hnswlib/hrnn.h, beam_query): candidates
// and results ordered by (dist, id); a node enters the result list when the
// list holds fewer than ef nodes or its distance is strictly below the worst;
// the closest unexpanded result is expanded next; the search ends when every
// result is expanded (= the closest open candidate is farther than the worst
// result).  Distances use the reference's 16-partial FMA order, so GPU and
// host beams visit the same nodes.
//
// One wave per query (persistent blocks of 64 lanes pull queries from a
// device queue); the sorted result list lives in LDS; visited nodes are
// marked in a per-block bitmap over the n base ids (cleared after each query).
#include <algorithm>
#include <cfloat>

#include "handle.h"

namespace hrnn {

constexpr int kBeamMaxEf = 512;
constexpr int kBeamMaxDeg = 64;

template <class VT>
__global__ __launch_bounds__(64) void k_beam(
    const VT *__restrict__ x, uint32_t d, uint64_t n, const uint32_t *__restrict__ l0,
    uint32_t l0_words, const uint32_t *__restrict__ up_off, const uint32_t *__restrict__ up_pool,
    uint32_t up_words, uint32_t entry, int maxlevel, const float *__restrict__ q, uint32_t nq,
    uint32_t ef, uint32_t m, uint32_t *__restrict__ visited_all, uint64_t vis_words,
    uint32_t *__restrict__ queue, uint32_t *__restrict__ proxies, uint32_t *__restrict__ expansions,
    float *__restrict__ raw) {
    __shared__ float s_d[2][kBeamMaxEf];
    __shared__ uint32_t s_i[2][kBeamMaxEf];
    __shared__ uint8_t s_x[2][kBeamMaxEf];   // expanded flag
    __shared__ float s_nd[kBeamMaxDeg];
    __shared__ uint32_t s_ni[kBeamMaxDeg];
    __shared__ uint32_t s_q;
    extern __shared__ float s_qv[];

    const uint32_t lane = threadIdx.x;
    const uint32_t g = lane & 15u, grp = lane >> 4;
    uint32_t *vis = visited_all + (uint64_t)blockIdx.x * vis_words;

    for (;;) {
        if (lane == 0) s_q = atomicAdd(queue, 1u);
        __syncthreads();
        const uint32_t qi = s_q;
        if (qi >= nq) break;
        for (uint32_t e = lane; e < d; e += 64) s_qv[e] = q[(uint64_t)qi * d + e];
        __syncthreads();

        // ---- greedy descent over the upper layers ----
        uint32_t cur = entry;
        float cd;
        {
            float t = dist_row<VT>(s_qv, x, cur, d, g);
            cd = __shfl(t, 0, 64);
        }
        const float d_entry = cd;
        for (int level = maxlevel; level > 0; level--) {
            for (;;) {
                const uint32_t *ll = up_pool + (uint64_t)up_off[cur] + (uint64_t)(level - 1) * up_words;
                uint32_t cnt = ll[0] & 0xFFFFu;
                float best = cd;
                uint32_t best_id = cur;
                for (uint32_t b = 0; b < cnt; b += 4) {
                    uint32_t t = b + grp;
                    float dv = FLT_MAX;
                    uint32_t id = 0;
                    if (t < cnt) {   // uniform per 16-lane group
                        id = ll[1 + t];
                        dv = dist_row<VT>(s_qv, x, id, d, g);
                    }
                    // sequential scan order of the host: first strictly smaller wins
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
        int buf = 0;
        uint32_t L = 1;
        if (lane == 0) {
            s_d[0][0] = cd;
            s_i[0][0] = cur;
            s_x[0][0] = 0;
        }
        if (lane == 0) atomicOr(&vis[cur >> 5], 1u << (cur & 31u));
        __syncthreads();
        uint32_t n_exp = 0, n_eval = 1;
        for (;;) {
            // closest unexpanded entry of the sorted list
            uint32_t pick = 0xFFFFFFFFu;
            for (uint32_t base = 0; base < L; base += 64) {
                uint32_t t = base + lane;
                bool open = t < L && s_x[buf][t] == 0;
                uint64_t mask = __ballot(open);
                if (mask) {
                    pick = base + (uint32_t)__ffsll((long long)mask) - 1;
                    break;
                }
            }
            if (pick == 0xFFFFFFFFu) break;
            const uint32_t c = s_i[buf][pick];
            __syncthreads();
            if (lane == 0) s_x[buf][pick] = 1;
            n_exp++;
            const uint32_t *ll = l0 + (uint64_t)c * l0_words;
            const uint32_t cnt = ll[0] & 0xFFFFu;
            // neighbours not visited yet
            uint32_t nb = (lane < cnt) ? ll[1 + lane] : 0u;
            bool fresh = false;
            if (lane < cnt) {
                uint32_t bit = 1u << (nb & 31u);
                fresh = (atomicOr(&vis[nb >> 5], bit) & bit) == 0;
            }
            uint64_t fm = __ballot(fresh);
            uint32_t nf = (uint32_t)__popcll(fm);
            n_eval += nf;
            uint32_t slot = (uint32_t)__popcll(fm & lanemask_lt());
            if (fresh) s_ni[slot] = nb;
            __syncthreads();
            // distances of the fresh neighbours, 4 at a time (one per 16-lane group)
            for (uint32_t b = 0; b < nf; b += 4) {
                uint32_t t = b + grp;
                if (t < nf) {
                    float dv = dist_row<VT>(s_qv, x, s_ni[t], d, g);
                    if (g == 0) s_nd[t] = dv;
                }
            }
            __syncthreads();
            // admission: list not full, or strictly closer than the worst entry
            // (candidates are admitted in the host's scan order, so the worst
            // entry moves as earlier candidates are inserted; emulate with a
            // sequential pass by lane 0 over at most 64 fresh neighbours)
            if (lane == 0) {
                uint32_t LL = L;
                float *dd = s_d[buf];
                uint32_t *ii = s_i[buf];
                uint8_t *xx = s_x[buf];
                for (uint32_t t = 0; t < nf; t++) {
                    float dv = s_nd[t];
                    uint32_t id = s_ni[t];
                    if (LL >= ef && !(dv < dd[LL - 1])) continue;
                    uint32_t p = (LL < ef) ? LL : LL - 1;
                    if (LL < ef) LL++;
                    while (p > 0 && pair_less(dv, id, dd[p - 1], ii[p - 1])) {
                        dd[p] = dd[p - 1];
                        ii[p] = ii[p - 1];
                        xx[p] = xx[p - 1];
                        p--;
                    }
                    dd[p] = dv;
                    ii[p] = id;
                    xx[p] = 0;
                }
                s_q = LL;   // reuse as broadcast slot
            }
            __syncthreads();
            L = s_q;
            __syncthreads();
        }
        // ---- output: the m closest ----
        for (uint32_t t = lane; t < m; t += 64)
            proxies[(uint64_t)qi * m + t] = (t < L) ? s_i[buf][t] : kInvalidId;
        if (expansions && lane == 0) {
            expansions[2 * (uint64_t)qi] = n_exp;
            expansions[2 * (uint64_t)qi + 1] = n_eval;
        }
        if (raw) {   // per-query statistics for the learned m: see kRawWords
            float *r = raw + (uint64_t)qi * kRawWords;
            if (lane < 16) r[2 + lane] = (lane < L) ? s_d[buf][lane] : s_d[buf][L - 1];
            if (lane == 0) {
                r[0] = d_entry;
                r[1] = d_start;
                r[18] = (float)n_exp;
                r[19] = (float)L;
            }
        }
        // clear the visited bitmap for the next query
        __syncthreads();
        for (uint64_t w = lane; w < vis_words; w += 64) vis[w] = 0;
        __syncthreads();
    }
}

// The GPU beam is switched off: its first GPU run (2000 queries, ef 200, m 64,
// with the learned-m feature kernel after it) ended in a memory fault whose
// cause has not been found; a CPU emulation of this kernel on the same index
// keeps every access in bounds.  Phase A runs on the host until the cause is
// shown; every entry point that would launch k_beam returns HRNN_E_STATE.
constexpr bool kBeamEnabled = false;

void beam_launch(hrnn_dev_t h, const float *q_dev, uint32_t nq, uint32_t m, uint32_t ef,
                 uint32_t *proxies_dev, uint32_t *expansions_dev, float *raw_dev) {
    REQUIRE(kBeamEnabled, HRNN_E_STATE,
            "GPU phase A (HNSW beam) is switched off pending an unresolved memory fault; "
            "run phase A on the host (hrnn_index_proxies) and pass the proxies to rknn_search");
    REQUIRE(h->hnsw_valid, HRNN_E_STATE, "proxies: upload the HNSW graph first");
    REQUIRE(ef >= 1 && ef <= (uint32_t)kBeamMaxEf, HRNN_E_ARG, "proxies: ef must be in [1, 512]");
    REQUIRE(m >= 1 && m <= ef, HRNN_E_ARG, "proxies: m must be in [1, ef]");
    REQUIRE(h->l0_words - 1 <= (uint32_t)kBeamMaxDeg, HRNN_E_ARG, "proxies: layer-0 degree > 64");
    if (nq == 0) return;
    uint64_t vis_words = (h->n + 31) / 32;
    uint32_t blocks = (uint32_t)std::min<uint64_t>(nq, (uint64_t)h->n_cu * h->tune_beam_wg_per_cu);
    uint64_t need = (uint64_t)blocks * vis_words * 4;
    if (h->beam_vis.bytes < need) {
        h->beam_vis.reserve(need);
        HIPC(hipMemsetAsync(h->beam_vis.p, 0, h->beam_vis.bytes, h->stream));
    }
    uint32_t *queue = h->counters.as<uint32_t>() + 80;
    HIPC(hipMemsetAsync(queue, 0, 4, h->stream));
#define LAUNCH(VT, XP)                                                                          \
    hipLaunchKernelGGL(k_beam<VT>, dim3(blocks), dim3(64), h->d * sizeof(float), h->stream, XP, h->d, \
                       h->n, h->l0.as<uint32_t>(), h->l0_words, h->up_off.as<uint32_t>(),            \
                       h->up_pool.as<uint32_t>(), h->up_words, h->entry, h->maxlevel, q_dev, nq, ef, m, \
                       h->beam_vis.as<uint32_t>(), vis_words, queue, proxies_dev, expansions_dev, raw_dev)
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
