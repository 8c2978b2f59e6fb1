// search.hip -- (c) RkNN search phases B and C for batches of queries.
//
// Reference: HRNN::searchRknn_impl (include/hnswlib/hrnn.h:1979-2071):
//   B (:2021-2037)  for each proxy, walk its reverse list while rank < K';
//                   count every candidate, remembering first sightings;
//   C (:2043-2069)  keep candidates with count >= threshold, verify
//                   L2Sqr(q, c) <= dist(c, k-1) (getVerifyKdistSq :286-293),
//                   return accepted ids in ascending order.
//
// One persistent launch of 256-lane workgroups that pull queries from a
// device queue, so queries with 100x different candidate counts balance
// across the chip.  Per query the m reverse-list prefixes are located by
// binary search on the rank field and gathered with coalesced reads.
//
// Deduplication and counting happen in an LDS hash table (open addressing,
// linear probing, sized to the query's scanned-entry count at load <= 3/4):
// u32 ids when only presence matters (threshold <= 1), u64 (count << 32 | id)
// when counts are compared with the threshold.  The surviving ids are
// compacted in place, verified by 16-lane groups (a rejected id is
// overwritten with 0xFFFFFFFF by the one group that owns it), compacted again
// and sorted, so HBM sees each scanned entry once and each candidate's vector
// and k-distance once; with one k for the batch the k-distances come from an
// n-float r_k column (4 MB at n = 1M, cache resident) instead of the n x K
// list matrix.  A query whose scanned entries do not fit the table
// falls back to a per-workgroup visited bitmap / u16 count array in HBM.
//
// Distances use the reference's 16-partial FMA order, so accept/reject
// decisions match the CPU path bit for bit; when every base value and query
// value is an integer in [0, 255] and d <= 258, all partial sums are exact
// integers, the order is immaterial, and each lane reads 16 contiguous bytes.
#include <algorithm>
#include <cfloat>
#include <mutex>
#include <type_traits>

#include "handle.h"

namespace hrnn {

constexpr int kSearchBlock = 256;
constexpr int kSearchWaves = kSearchBlock / 64;
constexpr int kMaxM = 1024;
constexpr unsigned long long kEmpty64 = 0xFFFFFFFFull;   // count 0, id 0xFFFFFFFF

// Insert candidate c into the LDS table of 2^logT slots.  Presence mode
// stores the id; count mode adds one to the count in the high word.  The
// table always holds fewer distinct ids than slots, so an empty or matching
// slot is always found.
template <bool kCount>
__device__ __forceinline__ void tab_insert(void *tab, uint32_t logT, uint32_t c) {
    const uint32_t mask = (1u << logT) - 1u;
    uint32_t s = (c * 0x9E3779B1u) >> (32u - logT);
    for (uint32_t probe = 0; probe <= mask; probe++, s = (s + 1u) & mask) {
        if (kCount) {
            unsigned long long *t = static_cast<unsigned long long *>(tab);
            const unsigned long long old = atomicCAS(&t[s], kEmpty64, (1ull << 32) | c);
            if (old == kEmpty64) return;
            if ((uint32_t)old == c) {
                atomicAdd(&t[s], 1ull << 32);
                return;
            }
        } else {
            uint32_t *t = static_cast<uint32_t *>(tab);
            const uint32_t old = atomicCAS(&t[s], kInvalidId, c);
            if (old == kInvalidId || old == c) return;
        }
    }
}

// Block-wide in-place compaction: for i in [0, n), pred(i, &v) decides whether
// entry i survives and gives its u32 value; survivors are written, in order, to
// dst[0..).  Entry i is read before any write to an index >= the current
// 256-entry chunk, so dst may overlay the storage that pred reads as long as
// index i of dst does not lie beyond the bytes pred(i) reads.  Returns the
// survivor count (the same in every lane).  n must be uniform in the block.
template <class Pred>
__device__ __forceinline__ uint32_t block_compact(uint32_t n, uint32_t *dst, uint32_t *s_wc,
                                                  Pred pred) {
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t out = 0;
    for (uint32_t b = 0; b < n; b += kSearchBlock) {
        const uint32_t i = b + threadIdx.x;
        uint32_t v = 0;
        const bool keep = (i < n) && pred(i, &v);
        const uint64_t bal = __ballot(keep);
        if (lane == 0) s_wc[wave] = (uint32_t)__popcll(bal);
        __syncthreads();
        uint32_t before = 0, chunk = 0;
#pragma unroll
        for (int w = 0; w < kSearchWaves; w++) {
            const uint32_t c = s_wc[w];
            before += ((uint32_t)w < wave) ? c : 0u;
            chunk += c;
        }
        if (keep) dst[out + before + (uint32_t)__popcll(bal & lanemask_lt())] = v;
        out += chunk;
        __syncthreads();
    }
    return out;
}

// r_k column: rk[i] = the k-th distance of i's ranked list (FLT_MAX when the
// list is shorter), the value getVerifyKdistSq returns (hrnn.h:286-293).
__global__ void k_rk_column(const float *__restrict__ knn_d, const uint32_t *__restrict__ knn_cnt,
                            uint64_t n, uint32_t K, uint32_t k, float *__restrict__ rk) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    rk[i] = (k - 1u < knn_cnt[i]) ? knn_d[i * K + k - 1u] : FLT_MAX;
}

// dst[0 .. count) = value (clears words without a memset node, so a caller
// may capture the launch in a graph)
__global__ void k_fill_u32(uint32_t *__restrict__ dst, uint32_t count, uint32_t value) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x)
        dst[i] = value;
}

template <class EntryT, bool kCount, class VT>
__global__ __launch_bounds__(kSearchBlock) void k_rknn_search(
    const VT *__restrict__ x, uint32_t d, int base_exact, uint64_t x_bytes, const uint64_t *__restrict__ off,
    const EntryT *__restrict__ ent, uint32_t rank_bits, const float *__restrict__ knn_d,
    const uint32_t *__restrict__ knn_cnt, uint32_t K, const float *__restrict__ rk_col,
    const float *__restrict__ q, uint32_t nq, const uint32_t *__restrict__ qlist,
    const uint32_t *__restrict__ qcount,
    const uint32_t *__restrict__ proxies, uint32_t m, const uint32_t *__restrict__ k_per_q,
    uint32_t k, uint32_t K_prime, float threshold, uint32_t log_tmax, uint32_t *__restrict__ mark_all,
    uint64_t mark_words, uint32_t *__restrict__ cand_all, uint64_t cand_cap,
    uint32_t *__restrict__ queue, uint32_t *__restrict__ res_cnt, uint32_t *__restrict__ res_ids,
    uint32_t cap_per_q, uint32_t *__restrict__ work) {
    __shared__ uint32_t s_q, s_ncand, s_nkeep, s_nres, s_qq;
    __shared__ uint32_t s_wc[kSearchWaves];
    // dynamic LDS: table (2^log_tmax slots) | query vector (d floats) | the
    // query as packed bytes (d) | prefix starts (m u64) | prefix positions (m + 1 u32)
    extern __shared__ __align__(16) unsigned char s_dyn[];
    typedef typename std::conditional<kCount, unsigned long long, uint32_t>::type SlotT;
    SlotT *s_tab = reinterpret_cast<SlotT *>(s_dyn);
    uint32_t *s_t32 = reinterpret_cast<uint32_t *>(s_dyn);
    const uint32_t tmax = 1u << log_tmax;
    const uint32_t res_lds = tmax * (uint32_t)(sizeof(SlotT) / 4);   // u32 capacity of the table
    float *s_qv = reinterpret_cast<float *>(s_dyn + (size_t)tmax * sizeof(SlotT));
    uint32_t *s_q8 = reinterpret_cast<uint32_t *>(s_dyn + (size_t)tmax * sizeof(SlotT) +
                                                  (((size_t)d * 4 + 7) & ~(size_t)7));
    uint64_t *s_start = reinterpret_cast<uint64_t *>(
        s_dyn + (size_t)tmax * sizeof(SlotT) + (((size_t)d * 4 + 7) & ~(size_t)7) + (((size_t)d + 7) & ~(size_t)7));
    uint32_t *s_pos = reinterpret_cast<uint32_t *>(s_start + m);

    uint32_t *mark = mark_all + (uint64_t)blockIdx.x * mark_words;
    uint32_t *cand = cand_all + (uint64_t)blockIdx.x * cand_cap * 2;
    uint32_t *keep = cand + cand_cap;
    const EntryT rmask = ((EntryT)1 << rank_bits) - 1;
    const uint32_t grp = threadIdx.x / 16, g = threadIdx.x & 15u;
    // the base as a buffer for the exact path (base_exact only below 4 GiB)
    const __amdgpu_buffer_rsrc_t xrs = make_rsrc(x, base_exact ? (uint32_t)(x_bytes) : 0u);
    // Phase C over list[0..cnt): 16-lane group grp takes candidates
    // i0 + grp + 16 r (r < 4) of each 64-candidate pass, so four row reads
    // (and their k-distances) are in flight per group; on_result(i, c, ok) runs
    // on the group's lane 0.  Padding rows read row 0 (list entries may be
    // rewritten by their owners during a pass).
    auto verify = [&](const uint32_t *list, uint32_t cnt, bool exact, uint32_t kidx, auto &&on_result) {
        for (uint32_t i0 = 0; i0 < cnt; i0 += 4u * (kSearchBlock / 16)) {
            uint32_t id4[4];
            bool v4[4];
            float kd4[4];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const uint32_t i = i0 + grp + (kSearchBlock / 16) * (uint32_t)r;
                v4[r] = i < cnt;
                id4[r] = v4[r] ? list[i] : 0u;
                kd4[r] = 0.f;
                if (v4[r] && g == 0)
                    kd4[r] = rk_col ? rk_col[id4[r]]
                                    : (kidx < knn_cnt[id4[r]]) ? knn_d[(uint64_t)id4[r] * K + kidx] : FLT_MAX;
            }
            float o[4];
            if (exact) {
                if constexpr (std::is_same<VT, uint8_t>::value)
                    dist4_u8_buf(s_q8, s_qq, xrs, id4, d, g, o);
                else
                    dist4_exact_buf<VT>(s_qv, xrs, id4, d * (uint32_t)sizeof(VT), d, g, o);
            } else if (std::is_same<VT, float>::value && (d & 15u) == 0) {
                // quad r of the group takes row r; its first lane holds the distance
                const float s4 = dist4_quad_f32(s_qv, reinterpret_cast<const float *>(x), id4, v4, d, g);
#pragma unroll
                for (int r = 0; r < 4; r++) o[r] = __shfl(s4, (int)((threadIdx.x & 63u & ~15u) + 4u * r), 64);
            } else {
#pragma unroll
                for (int r = 0; r < 4; r++) o[r] = v4[r] ? dist_row<VT>(s_qv, x, id4[r], d, g) : FLT_MAX;
            }
            if (g == 0) {
#pragma unroll
                for (int r = 0; r < 4; r++)
                    if (v4[r]) on_result(i0 + grp + (kSearchBlock / 16) * (uint32_t)r, id4[r], o[r] <= kd4[r]);
            }
        }
    };

    for (;;) {
        if (threadIdx.x == 0) {
            s_q = atomicAdd(queue, 1u);
            s_qq = 0;
            s_ncand = 0;
            s_nkeep = 0;
            s_nres = 0;
        }
        __syncthreads();
        const uint32_t idx = s_q;
        if (idx >= (qcount ? *qcount : nq)) break;
        const uint32_t qi = qlist ? qlist[idx] : idx;
        // The query read once (16-byte reads where its row allows: it may sit
        // in page-locked host memory): to LDS as FP32 and, for u8 storage,
        // packed 4 bytes per word with its sum of squares (used only when
        // every value is an integer in [0, 255]).
        int not_u8 = 0;
        const float *qrow = q + (uint64_t)qi * d;
        uint32_t part = 0;
        if ((d & 3u) == 0 && (reinterpret_cast<uintptr_t>(qrow) & 15u) == 0) {
            for (uint32_t w = threadIdx.x; 4 * w < d; w += blockDim.x) {
                const float4 v = reinterpret_cast<const float4 *>(qrow)[w];
                reinterpret_cast<float4 *>(s_qv)[w] = v;
                not_u8 |= !(is_u8_value(v.x) && is_u8_value(v.y) && is_u8_value(v.z) && is_u8_value(v.w));
                if constexpr (std::is_same<VT, uint8_t>::value) {
                    const uint32_t b0 = clamp_u8(v.x), b1 = clamp_u8(v.y), b2 = clamp_u8(v.z), b3 = clamp_u8(v.w);
                    s_q8[w] = b0 | (b1 << 8) | (b2 << 16) | (b3 << 24);
                    part += b0 * b0 + b1 * b1 + b2 * b2 + b3 * b3;
                }
            }
        } else {
            for (uint32_t e = threadIdx.x; e < d; e += blockDim.x) {
                const float v = qrow[e];
                s_qv[e] = v;
                not_u8 |= !is_u8_value(v);
            }
            if constexpr (std::is_same<VT, uint8_t>::value) {
                for (uint32_t w = threadIdx.x; 4 * w + 3 < d; w += blockDim.x) {
                    const float *qw = qrow + 4 * w;
                    const uint32_t b0 = clamp_u8(qw[0]), b1 = clamp_u8(qw[1]), b2 = clamp_u8(qw[2]), b3 = clamp_u8(qw[3]);
                    s_q8[w] = b0 | (b1 << 8) | (b2 << 16) | (b3 << 24);
                    part += b0 * b0 + b1 * b1 + b2 * b2 + b3 * b3;
                }
            }
        }
        if constexpr (std::is_same<VT, uint8_t>::value) atomicAdd(&s_qq, part);
        // locate each proxy's prefix with rank < K'
        for (uint32_t p = threadIdx.x; p < m; p += blockDim.x) {
            const uint32_t px = proxies[(uint64_t)qi * m + p];
            uint64_t b = 0, len = 0;
            if (px != kInvalidId) {
                b = off[px];
                uint64_t lo = b, hi = off[px + 1];
                while (lo < hi) {   // first entry with rank >= K'
                    const uint64_t mid = (lo + hi) >> 1;
                    if ((uint32_t)(ent[mid] & rmask) < K_prime) lo = mid + 1; else hi = mid;
                }
                len = lo - b;
            }
            s_start[p] = b;
            s_pos[p + 1] = (uint32_t)len;
        }
        const bool exact = base_exact && !__syncthreads_or(not_u8);
        if (threadIdx.x == 0) {
            s_pos[0] = 0;
            for (uint32_t p = 0; p < m; p++) s_pos[p + 1] += s_pos[p];
        }
        __syncthreads();
        const uint32_t total = s_pos[m];
        const uint32_t kq = k_per_q ? k_per_q[qi] : k;
        const uint32_t kidx = kq - 1;
        uint32_t nkeep, nres;

        if ((uint64_t)total * 4 <= (uint64_t)tmax * 3) {
            // ---- LDS path: table of T >= 4/3 total slots ----
            uint32_t logT = 8;
            while ((1u << logT) * 3u < total * 4u) logT++;
            const uint32_t T = 1u << logT;
            for (uint32_t s = threadIdx.x; s < T; s += blockDim.x)
                s_tab[s] = kCount ? (SlotT)kEmpty64 : (SlotT)kInvalidId;
            __syncthreads();
            for (uint32_t t = threadIdx.x; t < total; t += blockDim.x) {
                uint32_t lo = 0, hi = m;   // proxy p with s_pos[p] <= t < s_pos[p+1]
                while (hi - lo > 1) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (s_pos[mid] <= t) lo = mid; else hi = mid;
                }
                const EntryT e = ent[s_start[lo] + (t - s_pos[lo])];
                tab_insert<kCount>(s_tab, logT, (uint32_t)(e >> rank_bits));
            }
            __syncthreads();
            // threshold filter: surviving ids to s_t32[0..nkeep)
            nkeep = block_compact(T, s_t32, s_wc, [&](uint32_t i, uint32_t *v) {
                const SlotT sl = s_tab[i];
                *v = (uint32_t)sl;
                if (kCount) return (uint32_t)sl != kInvalidId && (float)(uint32_t)(sl >> 32) >= threshold;
                return (uint32_t)sl != kInvalidId;
            });
            // phase C: rejected ids become 0xFFFFFFFF in place (slot i is read
            // and written by its owner only)
            verify(s_t32, nkeep, exact, kidx, [&](uint32_t i, uint32_t, bool ok) {
                if (!ok) s_t32[i] = kInvalidId;
            });
            __syncthreads();
            nres = block_compact(nkeep, s_t32, s_wc, [&](uint32_t i, uint32_t *v) {
                *v = s_t32[i];
                return *v != kInvalidId;
            });
            const uint32_t P = next_pow2_dev(nres);
            for (uint32_t t = nres + threadIdx.x; t < P; t += blockDim.x) s_t32[t] = kInvalidId;
            if (P > 1) bitonic_sort_lds(s_t32, P);
            const uint32_t nout = nres < cap_per_q ? nres : cap_per_q;
            for (uint32_t t = threadIdx.x; t < nout; t += blockDim.x)
                res_ids[(uint64_t)qi * cap_per_q + t] = s_t32[t];
        } else {
            // ---- HBM fallback: private visited bitmap / u16 counts ----
            for (uint32_t t = threadIdx.x; t < total; t += blockDim.x) {
                uint32_t lo = 0, hi = m;
                while (hi - lo > 1) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (s_pos[mid] <= t) lo = mid; else hi = mid;
                }
                const EntryT e = ent[s_start[lo] + (t - s_pos[lo])];
                const uint32_t c = (uint32_t)(e >> rank_bits);
                bool first;
                if (kCount) {
                    const uint32_t sh = (c & 1u) * 16u;
                    const uint32_t old = atomicAdd(&mark[c >> 1], 1u << sh);
                    first = ((old >> sh) & 0xFFFFu) == 0;
                } else {
                    const uint32_t bit = 1u << (c & 31u);
                    const uint32_t old = atomicOr(&mark[c >> 5], bit);
                    first = (old & bit) == 0;
                }
                if (first) cand[atomicAdd(&s_ncand, 1u)] = c;
            }
            __syncthreads();
            const uint32_t ncand = s_ncand;
            for (uint32_t i = threadIdx.x; i < ncand; i += blockDim.x) {
                const uint32_t c = cand[i];
                bool ok = true;
                if (kCount) {
                    const uint32_t sc = (mark[c >> 1] >> ((c & 1u) * 16u)) & 0xFFFFu;
                    ok = (float)sc >= threshold;
                }
                if (ok) keep[atomicAdd(&s_nkeep, 1u)] = c;
            }
            __syncthreads();
            for (uint32_t i = threadIdx.x; i < ncand; i += blockDim.x) {
                const uint32_t c = cand[i];
                mark[kCount ? (c >> 1) : (c >> 5)] = 0;
            }
            nkeep = s_nkeep;
            verify(keep, nkeep, exact, kidx, [&](uint32_t, uint32_t c, bool ok) {
                if (ok) {
                    const uint32_t r = atomicAdd(&s_nres, 1u);
                    if (r < res_lds) s_t32[r] = c;
                    if (r < cap_per_q) res_ids[(uint64_t)qi * cap_per_q + r] = c;
                }
            });
            __syncthreads();
            nres = s_nres;
            const uint32_t nout = nres < cap_per_q ? nres : cap_per_q;
            if (nres <= res_lds) {
                const uint32_t P = next_pow2_dev(nres);
                for (uint32_t t = nres + threadIdx.x; t < P; t += blockDim.x) s_t32[t] = kInvalidId;
                if (P > 1) bitonic_sort_lds(s_t32, P);
                for (uint32_t t = threadIdx.x; t < nout; t += blockDim.x)
                    res_ids[(uint64_t)qi * cap_per_q + t] = s_t32[t];
            } else {
                // rare: more results than the LDS table holds; odd-even
                // transposition sort of the kept ids in place
                uint32_t *rr = res_ids + (uint64_t)qi * cap_per_q;
                for (uint32_t pass = 0; pass < nout; pass++) {
                    for (uint32_t t = 2 * threadIdx.x + (pass & 1u); t + 1 < nout; t += 2 * blockDim.x) {
                        const uint32_t a = rr[t], b = rr[t + 1];
                        if (a > b) { rr[t] = b; rr[t + 1] = a; }
                    }
                    __syncthreads();
                }
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

void rk_column_ready(hrnn_dev_t h, uint32_t k, hipStream_t st) {
    const uint64_t n = h->n;
    if (k < 1 || k > h->K) return;
    if (h->rk_col_k == k && h->rk_col_version == h->knn_version && h->rk_col_n == n) return;
    h->rk_col.reserve(n * 4);
    hipLaunchKernelGGL(k_rk_column, dim3(ceil_div(n, 256)), dim3(256), 0, st, h->knn_d.as<float>(),
                       h->knn_cnt.as<uint32_t>(), n, h->K, k, h->rk_col.as<float>());
    HIPC(hipGetLastError());
    h->rk_col_k = k;
    h->rk_col_version = h->knn_version;
    h->rk_col_n = n;
}

void ensure_stream2(hrnn_dev_t h) {
    if (h->stream2) return;
    HIPC(hipStreamCreateWithFlags(&h->stream2, hipStreamNonBlocking));
    HIPC(hipEventCreateWithFlags(&h->ev_a, hipEventDisableTiming));
    HIPC(hipEventCreateWithFlags(&h->ev_b, hipEventDisableTiming));
}

void rknn_search_resident(hrnn_dev_t h, const float *q, uint32_t nq, const uint32_t *proxies,
                          uint32_t m, const uint32_t *k_per_q, uint32_t k, uint32_t K_prime,
                          float threshold, uint32_t *res_cnt, uint32_t *res_ids,
                          uint32_t cap_per_q, uint32_t *work, const SearchSel *sel) {
    REQUIRE(h->csr_valid, HRNN_E_STATE, "search: no CSR");
    REQUIRE(m >= 1 && m <= (uint32_t)kMaxM, HRNN_E_ARG, "search: m must be in [1, 1024]");
    REQUIRE(cap_per_q > 0, HRNN_E_ARG, "search: cap_per_q == 0");
    REQUIRE(h->n < kInvalidId, HRNN_E_ARG, "search: n must be below 2^32 - 1");
    const SearchSel so = sel ? *sel : SearchSel{};
    hipStream_t st = so.stream ? so.stream : h->stream;
    const bool count_mode = threshold > 1.0f;
    const uint64_t n = h->n;
    const uint64_t words = count_mode ? (n + 1) / 2 : (n + 31) / 32;
    // distinct candidates <= sum of the m prefix lengths <= m * longest list
    uint64_t cand_cap = std::max<uint64_t>(1, (uint64_t)m * std::max<uint64_t>(h->max_deg, 1));
    cand_cap = std::min<uint64_t>(cand_cap, n);
    uint32_t slots = (uint32_t)std::min<uint64_t>(nq, (uint64_t)h->n_cu * h->tune_search_wg_per_cu);
    if (so.max_slots) slots = std::min(slots, so.max_slots);
    slots = std::max<uint32_t>(slots, 1);
    DevBuf &bmark = so.buffer_set ? h->search_mark2 : h->search_mark;
    DevBuf &bcand = so.buffer_set ? h->search_cand2 : h->search_cand;
    const uint64_t per_slot = words * 4 + cand_cap * 8;
    uint32_t *queue = nullptr, *mark = nullptr, *cand = nullptr;
    if (so.scratch) {
        // caller-free form: [queue word | pad to 256 B | slots x bitmap | slots x
        // candidate pairs]; the bitmaps are all-zero between launches (the
        // kernel clears what it sets), the queue is cleared by a kernel
        REQUIRE(so.scratch_bytes > 256 + per_slot, HRNN_E_ARG,
                "search (no allocation): one workgroup's fallback state (" + std::to_string(per_slot) +
                    " B at n = " + std::to_string(n) + ", m = " + std::to_string(m) +
                    ") exceeds the scratch array of " + std::to_string(so.scratch_bytes) + " B");
        REQUIRE(so.scratch_clean, HRNN_E_ARG, "search (no allocation): no clean-extent record");
        slots = (uint32_t)std::min<uint64_t>(slots, (so.scratch_bytes - 256) / per_slot);
        queue = reinterpret_cast<uint32_t *>(so.scratch);
        mark = reinterpret_cast<uint32_t *>(so.scratch + 256);
        cand = mark + (uint64_t)slots * words;
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, queue, 1u, 0u);
        const uint64_t mark_bytes = (uint64_t)slots * words * 4;
        if (mark_bytes > *so.scratch_clean) {
            const uint64_t w0 = *so.scratch_clean / 4, nw = (mark_bytes - *so.scratch_clean) / 4;
            hipLaunchKernelGGL(k_fill_u32, dim3(std::min<uint32_t>(ceil_div(nw, 256), 2048)), dim3(256), 0, st,
                               mark + w0, (uint32_t)nw, 0u);
        }
        HIPC(hipGetLastError());
        *so.scratch_clean = mark_bytes;
    } else {
        // bound the fallback's private state to ~4 GB
        const uint64_t max_slots = std::max<uint64_t>(1, (4ull << 30) / per_slot);
        slots = (uint32_t)std::min<uint64_t>(slots, max_slots);
        // private arrays (kept all-zero between launches by the kernel itself) and
        // candidate buffers are owned by the search alone
        const uint64_t mark_bytes = (uint64_t)slots * words * 4;
        if (bmark.bytes < mark_bytes) {
            bmark.reserve(mark_bytes);
            HIPC(hipMemsetAsync(bmark.p, 0, bmark.bytes, st));
        }
        bcand.reserve((uint64_t)slots * cand_cap * 8);
        queue = h->counters.as<uint32_t>() + so.queue_word;
        if (!so.queue_zeroed) HIPC(hipMemsetAsync(queue, 0, 4, st));
        mark = bmark.as<uint32_t>();
        cand = bcand.as<uint32_t>();
    }
    // LDS table: the largest power-of-two slot count within the configured
    // bytes (u32 slots in presence mode, u64 in count mode), at least 256
    const uint32_t slot_bytes = count_mode ? 8u : 4u;
    uint32_t log_tmax = 8;
    while ((2ull << log_tmax) * slot_bytes <= (uint64_t)h->tune_search_lds_kib * 1024) log_tmax++;
    const size_t dyn = ((size_t)1 << log_tmax) * slot_bytes + (((size_t)h->d * 4 + 7) & ~(size_t)7) +
                       (((size_t)h->d + 7) & ~(size_t)7) + (size_t)m * 8 + ((size_t)m + 1) * 4;
    REQUIRE(dyn <= std::min<size_t>(h->lds_per_block, 65536), HRNN_E_ARG,
            "search: LDS need " + std::to_string(dyn) + " B exceeds the workgroup limit");
    const uint64_t *off = h->csr_off.as<uint64_t>();
    const float *kd = h->knn_d.as<float>();
    const uint32_t *kc = h->knn_cnt.as<uint32_t>();
    // one k for the whole batch: verify against the r_k column (built once per
    // graph version and k)
    const float *rk = nullptr;
    const bool rk_ready = h->rk_col_k == k && h->rk_col_version == h->knn_version && h->rk_col_n == n;
    if (so.scratch) {
        if (!k_per_q && rk_ready) rk = h->rk_col.as<float>();   // else read the list matrix
    } else if (!k_per_q && k >= 1 && k <= h->K) {
        rk_column_ready(h, k, st);
        rk = h->rk_col.as<float>();
    }
    // Candidate vectors are read from the u8 mirror (HRNN_PREC_U8) or the BF16
    // mirror when every base value is exactly representable in it (same
    // distances, a quarter or half of the bytes).
    const bool use_u8 = h->vec_u8 && h->base_u8;
    const bool bf16 = !use_u8 && h->small_int && h->vec_bf16;
    const uint64_t x_bytes = n * h->d * (use_u8 ? 1 : bf16 ? 2 : 4);
    const int base_exact = (h->base_u8 && h->d <= 258 && h->d % 8 == 0 && x_bytes < 0xFFFFFFF0ull) ? 1 : 0;
#define LAUNCH(ET, CM, VT, XP)                                                                      \
    hipLaunchKernelGGL((k_rknn_search<ET, CM, VT>), dim3(slots), dim3(kSearchBlock), dyn, st,       \
                       XP, h->d, base_exact, x_bytes, off, h->csr_ent.as<ET>(), h->rank_bits, kd, kc, h->K, rk, q, \
                       nq, so.qlist, so.qcount, proxies, m, k_per_q, k, K_prime, threshold, log_tmax,  \
                       mark, words, cand, cand_cap, queue, res_cnt, res_ids, cap_per_q, work)
#define LAUNCH_VT(ET, CM)                                                      \
    if (use_u8) LAUNCH(ET, CM, uint8_t, h->xu8.as<uint8_t>());                 \
    else if (bf16) LAUNCH(ET, CM, __hip_bfloat16, h->xb.as<__hip_bfloat16>()); \
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

// ---- search without allocation ----
//
// hrnn_dev_rknn_search_noalloc_dptr keeps the search's queue word and its
// per-workgroup fallback state (visited bitmaps or u16 counts, candidate
// buffers) in this array, of which every GPU the library runs on holds one
// copy for the life of the process.  Host code never uses the array's name
// as an address: hrnn_dev_open looks up this device's copy once
// (hipGetSymbolAddress) and every use goes through that pointer.
constexpr size_t kNoAllocScratchBytes = 128ull << 20;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];
static std::mutex g_scratch_mu;
static uint8_t *g_scratch_ptr[kMaxDevices] = {};
static hipError_t g_scratch_err[kMaxDevices] = {};
static size_t g_scratch_clean[kMaxDevices] = {};   // zero bytes at the start of the bitmap area

void noalloc_scratch_init(hrnn_dev_t h) {
    std::lock_guard<std::mutex> g(g_scratch_mu);
    const int dv = h->device;
    if (dv < 0 || dv >= kMaxDevices || g_scratch_ptr[dv]) return;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        g_scratch_err[dv] = e != hipSuccess ? e : hipErrorInvalidSymbol;
        (void)hipGetLastError();
        return;
    }
    // the bitmaps must start all-zero; the kernel keeps them so afterwards
    hipLaunchKernelGGL(k_fill_u32, dim3(2048), dim3(256), 0, h->stream, static_cast<uint32_t *>(p),
                       (uint32_t)(kNoAllocScratchBytes / 4), 0u);
    HIPC(hipGetLastError());
    HIPC(hipStreamSynchronize(h->stream));
    g_scratch_ptr[dv] = static_cast<uint8_t *>(p);
    g_scratch_clean[dv] = kNoAllocScratchBytes - 256;
}

// ---- result compaction (the download side of searchRknn's return value) ----
//
// The resident search leaves query q's ids at res_ids[q * cap .. q * cap +
// min(cnt[q], cap)).  k_results_scan forms off[q] = sum over j < q of
// min(cnt[j], cap) (u64), off[nq] = the total and off[nq + 1] = the largest
// true count, in one workgroup (8 consecutive counts per lane, a wave scan,
// per-wave totals in LDS, a running carry between chunks of 8,192 queries),
// and, given host_off / host_max (device addresses of page-locked host
// memory), writes off[0..nq] and the largest count there too;
// k_results_gather then packs the ids densely.
constexpr int kResScanThreads = 1024;
constexpr uint32_t kResScanPer = 8;

__global__ __launch_bounds__(kResScanThreads) void k_results_scan(const uint32_t *__restrict__ cnt,
                                                                   uint32_t nq, uint32_t cap,
                                                                   uint64_t *__restrict__ off,
                                                                   uint64_t *__restrict__ host_off,
                                                                   uint64_t *__restrict__ host_max) {
    __shared__ unsigned long long s_w[kResScanThreads / 64];
    __shared__ unsigned long long s_carry;
    __shared__ uint32_t s_max;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    if (threadIdx.x == 0) {
        s_carry = 0;
        s_max = 0;
    }
    __syncthreads();
    uint32_t mx = 0;
    for (uint32_t base = 0; base < nq; base += kResScanThreads * kResScanPer) {
        const uint32_t i0 = base + threadIdx.x * kResScanPer;
        uint32_t v[kResScanPer];
        unsigned long long sum = 0;
#pragma unroll
        for (uint32_t j = 0; j < kResScanPer; j++) {
            const uint32_t i = i0 + j;
            const uint32_t c = i < nq ? cnt[i] : 0u;
            mx = c > mx ? c : mx;
            v[j] = c < cap ? c : cap;
            sum += v[j];
        }
        unsigned long long inc = sum;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const unsigned long long t = __shfl_up(inc, (unsigned)o, 64);
            if ((int)lane >= o) inc += t;
        }
        if (lane == 63u) s_w[wave] = inc;
        __syncthreads();
        unsigned long long run = s_carry;
        for (uint32_t w = 0; w < wave; w++) run += s_w[w];
        run += inc - sum;
#pragma unroll
        for (uint32_t j = 0; j < kResScanPer; j++) {
            const uint32_t i = i0 + j;
            if (i < nq) {
                off[i] = run;
                if (host_off) host_off[i] = run;
            }
            run += v[j];
        }
        __syncthreads();   // every lane has read s_carry and s_w
        if (threadIdx.x == kResScanThreads - 1) s_carry = run;
        __syncthreads();
    }
    for (int o = 32; o > 0; o >>= 1) {
        const uint32_t y = (uint32_t)__shfl_xor((int)mx, o, 64);
        mx = y > mx ? y : mx;
    }
    if (lane == 0) atomicMax(&s_max, mx);
    __syncthreads();
    if (threadIdx.x == 0) {
        off[nq] = s_carry;
        off[nq + 1] = s_max;
        if (host_off) host_off[nq] = s_carry;
        if (host_max) *host_max = s_max;
    }
}

// One thread per packed output position (grid-stride): the query that owns
// position o is the last q with off[q] <= o (a binary search over the
// offsets), so consecutive lanes store consecutive ids -- full-line writes,
// also over the link when out is page-locked host memory.
constexpr uint32_t kGatherBlocks = 1024;
__global__ __launch_bounds__(256) void k_results_gather(const uint32_t *__restrict__ ids, uint32_t nq,
                                                        uint32_t cap, const uint64_t *__restrict__ off,
                                                        uint32_t *__restrict__ out, uint64_t out_cap) {
    const uint64_t total = off[nq] < out_cap ? off[nq] : out_cap;
    for (uint64_t o = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; o < total;
         o += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t lo = 0, hi = nq;   // off[lo] <= o < off[hi]
        while (hi - lo > 1) {
            const uint32_t mid = lo + (hi - lo) / 2;
            if (off[mid] <= o) lo = mid;
            else hi = mid;
        }
        out[o] = ids[(uint64_t)lo * cap + (o - off[lo])];
    }
}

// The device address of page-locked host memory (hipHostMalloc), or null for
// memory the device cannot reach (pageable).
static void *pinned_device_ptr(const void *p) {
    if (!p) return nullptr;
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return a.type == hipMemoryTypeHost ? a.devicePointer : nullptr;
}

}  // namespace hrnn

using namespace hrnn;

extern "C" {

int hrnn_dev_rknn_search_noalloc_dptr(hrnn_dev_t h, const float *q_dev, uint32_t nq,
                                      const uint32_t *proxies_dev, uint32_t m,
                                      const uint32_t *k_per_q_dev, uint32_t k, uint32_t K_prime,
                                      float threshold, uint32_t *res_cnt_dev, uint32_t *res_ids_dev,
                                      uint32_t cap_per_q, uint32_t *work_dev) {
    if (!h) return HRNN_E_ARG;
    return guarded(h, [&] {
        REQUIRE(h->csr_valid, HRNN_E_STATE, "rknn_search_noalloc: build the CSR first");
        REQUIRE(h->K > 0, HRNN_E_STATE, "rknn_search_noalloc: no KNN graph");
        REQUIRE(k_per_q_dev || (k >= 1 && k <= h->K), HRNN_E_ARG, "rknn_search_noalloc: k out of [1, K]");
        REQUIRE(m >= 1 && m <= 1024, HRNN_E_ARG, "rknn_search_noalloc: m out of [1, 1024]");
        const int dv = h->device;
        REQUIRE(dv >= 0 && dv < kMaxDevices, HRNN_E_ARG, "rknn_search_noalloc: device ordinal out of range");
        std::lock_guard<std::mutex> g(g_scratch_mu);   // held while the launches are queued
        uint8_t *p = g_scratch_ptr[dv];
        REQUIRE(p, HRNN_E_HIP, std::string("rknn_search_noalloc: scratch array not available on this device (") +
                                   hipGetErrorString(g_scratch_err[dv]) + ")");
        if (nq == 0) return;
        SearchSel sel;
        sel.scratch = p;
        sel.scratch_bytes = kNoAllocScratchBytes;
        sel.scratch_clean = &g_scratch_clean[dv];
        rknn_search_resident(h, q_dev, nq, proxies_dev, m, k_per_q_dev, k, K_prime, threshold, res_cnt_dev,
                             res_ids_dev, cap_per_q, work_dev, &sel);
    });
}

int hrnn_dev_fetch_results(hrnn_dev_t h, uint32_t nq, const uint32_t *res_cnt_dev,
                           const uint32_t *res_ids_dev, uint32_t cap_per_q, uint64_t *res_off,
                           uint32_t *res_ids, uint64_t cap, uint64_t *needed) {
    if (!h) return HRNN_E_ARG;
    return guarded(h, [&] {
        REQUIRE(res_off && (nq == 0 || (res_cnt_dev && res_ids_dev)), HRNN_E_ARG,
                "fetch_results: null argument");
        REQUIRE(cap_per_q > 0, HRNN_E_ARG, "fetch_results: cap_per_q == 0");
        if (needed) *needed = 0;
        res_off[0] = 0;
        if (nq == 0) return;
        h->fetch_off.reserve(((uint64_t)nq + 2) * 8);
        if (!h->pin_small) HIPC(hipHostMalloc(&h->pin_small, 64, hipHostMallocDefault));
        uint64_t *off = h->fetch_off.as<uint64_t>();
        uint64_t *small = static_cast<uint64_t *>(h->pin_small);
        // Page-locked offsets: the scan writes them (and the largest count)
        // straight into host memory as well, with no copies after it.
        uint64_t *off_direct = static_cast<uint64_t *>(pinned_device_ptr(res_off));
        uint64_t *small_direct = off_direct ? static_cast<uint64_t *>(pinned_device_ptr(small)) : nullptr;
        hipLaunchKernelGGL(k_results_scan, dim3(1), dim3(kResScanThreads), 0, h->stream, res_cnt_dev, nq,
                           cap_per_q, off, small_direct ? off_direct : nullptr, small_direct);
        HIPC(hipGetLastError());
        if (!small_direct) {
            HIPC(hipMemcpyAsync(res_off, off, ((uint64_t)nq + 1) * 8, hipMemcpyDeviceToHost, h->stream));
            HIPC(hipMemcpyAsync(small, off + nq + 1, 8, hipMemcpyDeviceToHost, h->stream));
        }
        // Page-locked id buffer: the gather writes the packed ids straight
        // into it (at most cap of them), so one wait covers the whole fetch.
        uint32_t *ids_direct = res_ids && cap ? static_cast<uint32_t *>(pinned_device_ptr(res_ids)) : nullptr;
        if (ids_direct) {
            hipLaunchKernelGGL(k_results_gather, dim3(kGatherBlocks), dim3(256), 0, h->stream, res_ids_dev, nq,
                               cap_per_q, off, ids_direct, cap);
            HIPC(hipGetLastError());
        }
        HIPC(hipStreamSynchronize(h->stream));
        const uint64_t total = res_off[nq], mx = small[0];
        if (mx > cap_per_q) {
            if (needed) *needed = mx;
            throw Error{HRNN_E_CAP, "fetch_results: a query has " + std::to_string(mx) +
                                        " results, more than cap_per_q = " + std::to_string(cap_per_q) +
                                        " (its list was cut short on the device)"};
        }
        if (needed) *needed = total;
        if (total > cap || (total && !res_ids))
            throw Error{HRNN_E_CAP, "fetch_results: result capacity too small"};
        if (!total || ids_direct) return;
        h->fetch_ids.reserve(total * 4);
        hipLaunchKernelGGL(k_results_gather, dim3(kGatherBlocks), dim3(256), 0, h->stream, res_ids_dev, nq,
                           cap_per_q, off, h->fetch_ids.as<uint32_t>(), total);
        HIPC(hipGetLastError());
        HIPC(hipMemcpyAsync(res_ids, h->fetch_ids.p, total * 4, hipMemcpyDeviceToHost, h->stream));
        HIPC(hipStreamSynchronize(h->stream));
    });
}

}  // extern "C"
