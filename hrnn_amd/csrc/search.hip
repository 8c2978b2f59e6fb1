// search.hip -- (c) RkNN search phases B and C for batches of queries.
//
// Reference: HRNN::searchRknn_impl (include/hnswlib/hrnn.h:1979-2071):
//   B (:2021-2037)  for each proxy, walk its reverse list while rank < K';
//                   count every candidate, remembering first sightings;
//   C (:2043-2069)  keep candidates with count >= threshold, verify
//                   L2Sqr(q, c) <= dist(c, k-1) (getVerifyKdistSq :286-293),
//                   return accepted ids in ascending order.
//
// One persistent launch: each workgroup (256 lanes) owns a private visited
// bitmap (threshold <= 1) or u16 count array (threshold > 1) over the n base
// ids and a candidate buffer, and pulls queries from a device queue, so
// queries with 100x different candidate counts balance across the chip.
// Per query: the m reverse-list prefixes are located by binary search on the
// rank field, gathered with coalesced reads, deduplicated in the private
// array, and verified by 16-lane groups reading each candidate's vector as
// contiguous 64-byte segments; the distance uses the reference's 16-partial
// FMA order so accept/reject decisions match the CPU path bit for bit.
#include <algorithm>
#include <cfloat>

#include "handle.h"

namespace hrnn {

constexpr int kSearchBlock = 256;
constexpr int kMaxM = 1024;
constexpr uint32_t kResLds = 4096;   // results sorted in LDS up to this count

template <class EntryT, bool kCount, class VT>
__global__ __launch_bounds__(kSearchBlock) void k_rknn_search(
    const VT *__restrict__ x, uint32_t d, const uint64_t *__restrict__ off,
    const EntryT *__restrict__ ent, uint32_t rank_bits, const float *__restrict__ knn_d,
    const uint32_t *__restrict__ knn_cnt, uint32_t K, const float *__restrict__ q, uint32_t nq,
    const uint32_t *__restrict__ proxies, uint32_t m, const uint32_t *__restrict__ k_per_q,
    uint32_t k, uint32_t K_prime, float threshold, uint32_t *__restrict__ mark_all,
    uint64_t mark_words, uint32_t *__restrict__ cand_all, uint64_t cand_cap,
    uint32_t *__restrict__ queue, uint32_t *__restrict__ res_cnt, uint32_t *__restrict__ res_ids,
    uint32_t cap_per_q, uint32_t *__restrict__ work) {
    __shared__ uint32_t s_q;
    __shared__ uint32_t s_ncand, s_nkeep, s_nres;
    __shared__ uint64_t s_start[kMaxM];
    __shared__ uint32_t s_pos[kMaxM + 1];
    __shared__ uint32_t s_res[kResLds];
    extern __shared__ float s_qv[];

    uint32_t *mark = mark_all + (uint64_t)blockIdx.x * mark_words;
    uint32_t *cand = cand_all + (uint64_t)blockIdx.x * cand_cap * 2;
    uint32_t *keep = cand + cand_cap;
    const EntryT rmask = ((EntryT)1 << rank_bits) - 1;

    for (;;) {
        if (threadIdx.x == 0) {
            s_q = atomicAdd(queue, 1u);
            s_ncand = 0;
            s_nkeep = 0;
            s_nres = 0;
        }
        __syncthreads();
        uint32_t qi = s_q;
        if (qi >= nq) break;
        for (uint32_t e = threadIdx.x; e < d; e += blockDim.x) s_qv[e] = q[(uint64_t)qi * d + e];
        // locate each proxy's prefix with rank < K'
        for (uint32_t p = threadIdx.x; p < m; p += blockDim.x) {
            uint32_t px = proxies[(uint64_t)qi * m + p];
            uint64_t b = 0, len = 0;
            if (px != kInvalidId) {
                b = off[px];
                uint64_t lo = b, hi = off[px + 1];
                while (lo < hi) {   // first entry with rank >= K'
                    uint64_t mid = (lo + hi) >> 1;
                    if ((uint32_t)(ent[mid] & rmask) < K_prime) lo = mid + 1; else hi = mid;
                }
                len = lo - b;
            }
            s_start[p] = b;
            s_pos[p + 1] = (uint32_t)len;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            s_pos[0] = 0;
            for (uint32_t p = 0; p < m; p++) s_pos[p + 1] += s_pos[p];
        }
        __syncthreads();
        uint32_t total = s_pos[m];
        // phase B: gather, dedupe and count
        for (uint32_t t = threadIdx.x; t < total; t += blockDim.x) {
            uint32_t lo = 0, hi = m;   // proxy p with s_pos[p] <= t < s_pos[p+1]
            while (hi - lo > 1) {
                uint32_t mid = (lo + hi) >> 1;
                if (s_pos[mid] <= t) lo = mid; else hi = mid;
            }
            EntryT e = ent[s_start[lo] + (t - s_pos[lo])];
            uint32_t c = (uint32_t)(e >> rank_bits);
            bool first;
            if (kCount) {
                uint32_t sh = (c & 1u) * 16u;
                uint32_t old = atomicAdd(&mark[c >> 1], 1u << sh);
                first = ((old >> sh) & 0xFFFFu) == 0;
            } else {
                uint32_t bit = 1u << (c & 31u);
                uint32_t old = atomicOr(&mark[c >> 5], bit);
                first = (old & bit) == 0;
            }
            if (first) {
                uint32_t s = atomicAdd(&s_ncand, 1u);
                cand[s] = c;
            }
        }
        __syncthreads();
        uint32_t ncand = s_ncand;
        // threshold filter and reset of the private array
        for (uint32_t i = threadIdx.x; i < ncand; i += blockDim.x) {
            uint32_t c = cand[i];
            bool ok = true;
            if (kCount) {
                uint32_t sc = (mark[c >> 1] >> ((c & 1u) * 16u)) & 0xFFFFu;
                ok = (float)sc >= threshold;
            }
            if (ok) {
                uint32_t s = atomicAdd(&s_nkeep, 1u);
                keep[s] = c;
            }
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < ncand; i += blockDim.x) {
            uint32_t c = cand[i];
            mark[kCount ? (c >> 1) : (c >> 5)] = 0;
        }
        uint32_t nkeep = s_nkeep;
        uint32_t kq = k_per_q ? k_per_q[qi] : k;
        uint32_t kidx = kq - 1;
        // phase C: verification, one 16-lane group per candidate
        uint32_t grp = threadIdx.x / 16, g = threadIdx.x & 15u;
        for (uint32_t i = grp; i < nkeep; i += blockDim.x / 16) {
            uint32_t c = keep[i];
            float dist = dist_row<VT>(s_qv, x, c, d, g);
            if (g == 0) {
                float kd = (kidx < knn_cnt[c]) ? knn_d[(uint64_t)c * K + kidx] : FLT_MAX;
                if (dist <= kd) {
                    uint32_t r = atomicAdd(&s_nres, 1u);
                    if (r < kResLds) s_res[r] = c;
                    if (r < cap_per_q) res_ids[(uint64_t)qi * cap_per_q + r] = c;
                }
            }
        }
        __syncthreads();
        uint32_t nres = s_nres;
        uint32_t nout = nres < cap_per_q ? nres : cap_per_q;
        if (nres <= kResLds) {
            uint32_t P = 1;
            while (P < nres) P <<= 1;
            for (uint32_t t = nres + threadIdx.x; t < P; t += blockDim.x) s_res[t] = 0xFFFFFFFFu;
            if (P > 1) bitonic_sort_lds(s_res, P);
            for (uint32_t t = threadIdx.x; t < nout; t += blockDim.x)
                res_ids[(uint64_t)qi * cap_per_q + t] = s_res[t];
        } else {
            // rare: more results than the LDS buffer; odd-even transposition sort in place
            uint32_t *rr = res_ids + (uint64_t)qi * cap_per_q;
            for (uint32_t pass = 0; pass < nout; pass++) {
                for (uint32_t t = 2 * threadIdx.x + (pass & 1u); t + 1 < nout; t += 2 * blockDim.x) {
                    uint32_t a = rr[t], b = rr[t + 1];
                    if (a > b) { rr[t] = b; rr[t + 1] = a; }
                }
                __syncthreads();
            }
        }
        if (threadIdx.x == 0) {
            res_cnt[qi] = nres;
            if (work) {
                work[2 * (uint64_t)qi] = total;
                work[2 * (uint64_t)qi + 1] = nkeep;
            }
        }
        __syncthreads();
    }
}

void rknn_search_resident(hrnn_dev_t h, const float *q, uint32_t nq, const uint32_t *proxies,
                          uint32_t m, const uint32_t *k_per_q, uint32_t k, uint32_t K_prime,
                          float threshold, uint32_t *res_cnt, uint32_t *res_ids,
                          uint32_t cap_per_q, uint32_t *work) {
    REQUIRE(h->csr_valid, HRNN_E_STATE, "search: no CSR");
    REQUIRE(m <= (uint32_t)kMaxM, HRNN_E_ARG, "search: m > 1024");
    REQUIRE(cap_per_q > 0, HRNN_E_ARG, "search: cap_per_q == 0");
    bool count_mode = threshold > 1.0f;
    uint64_t n = h->n;
    uint64_t words = count_mode ? (n + 1) / 2 : (n + 31) / 32;
    // distinct candidates <= sum of the m prefix lengths <= m * longest list
    uint64_t cand_cap = std::max<uint64_t>(1, (uint64_t)m * std::max<uint64_t>(h->max_deg, 1));
    cand_cap = std::min<uint64_t>(cand_cap, n);
    uint32_t slots = (uint32_t)std::min<uint64_t>(nq, (uint64_t)h->n_cu * h->tune_search_wg_per_cu);
    // bound the private state to ~4 GB
    uint64_t per_slot = words * 4 + cand_cap * 8;
    uint64_t max_slots = std::max<uint64_t>(1, (4ull << 30) / per_slot);
    slots = (uint32_t)std::min<uint64_t>(slots, max_slots);
    // private arrays (kept all-zero between launches by the kernel itself) and
    // candidate buffers are owned by the search alone
    uint64_t mark_bytes = (uint64_t)slots * words * 4;
    if (h->search_mark.bytes < mark_bytes) {
        h->search_mark.reserve(mark_bytes);
        HIPC(hipMemsetAsync(h->search_mark.p, 0, h->search_mark.bytes, h->stream));
    }
    h->search_cand.reserve((uint64_t)slots * cand_cap * 8);
    uint32_t *queue = h->counters.as<uint32_t>() + 64;
    HIPC(hipMemsetAsync(queue, 0, 4, h->stream));
    size_t dyn = (size_t)h->d * sizeof(float);
    const uint64_t *off = h->csr_off.as<uint64_t>();
    const float *kd = h->knn_d.as<float>();
    const uint32_t *kc = h->knn_cnt.as<uint32_t>();
    uint32_t *mark = h->search_mark.as<uint32_t>();
    uint32_t *cand = h->search_cand.as<uint32_t>();
    // Candidate vectors are read from the BF16 mirror when every base value is
    // exactly representable in it (same distances, half the bytes).
    const bool bf16 = h->small_int && h->vec_bf16;
#define LAUNCH(ET, CM, VT, XP)                                                                      \
    hipLaunchKernelGGL((k_rknn_search<ET, CM, VT>), dim3(slots), dim3(kSearchBlock), dyn, h->stream, \
                       XP, h->d, off, h->csr_ent.as<ET>(), h->rank_bits, kd, kc, h->K, q, nq, proxies, \
                       m, k_per_q, k, K_prime, threshold, mark, words, cand, cand_cap, queue, res_cnt, \
                       res_ids, cap_per_q, work)
#define LAUNCH_VT(ET, CM)                                                  \
    if (bf16) LAUNCH(ET, CM, __hip_bfloat16, h->xb.as<__hip_bfloat16>());  \
    else LAUNCH(ET, CM, float, h->x.as<float>())
    if (h->wide) {
        if (count_mode) { LAUNCH_VT(uint64_t, true); } else { LAUNCH_VT(uint64_t, false); }
    } else {
        if (count_mode) { LAUNCH_VT(uint32_t, true); } else { LAUNCH_VT(uint32_t, false); }
    }
#undef LAUNCH_VT
#undef LAUNCH
    HIPC(hipGetLastError());
}

}  // namespace hrnn
