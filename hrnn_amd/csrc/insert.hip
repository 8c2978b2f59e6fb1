// insert.hip -- incremental insertion on the GPU.
//
// Reference: HRNN::insertPointDynamic (include/hnswlib/hrnn.h:2360-2534) with
// try_update_knng_evict (:1592-1655) over the mutable reverse lists of
// buildRKNNG_mutable (:1942-1959).  Per inserted point x the reference
//   1. inserts x into the HNSW graph and seeds its ranked KNN list from the
//      layer-0 beam (addPoint);
//   2. appends x to the reverse list of every node on its list;
//   3. collects the candidates: every id on the reverse lists of the first
//      m_update nodes of its list (all of them when m_update is 0), each
//      once, x and the nodes on its list excluded;
//   4. offers x to each candidate c at d(x, c): a full list takes it only
//      when d < its last distance, replacing that entry (whose reverse list
//      loses c); x's reverse list gains c.
// Here the points go in batches.  Step 1 is the batched HNSW insertion of
// hnsw_build.hip.  The reverse lists of steps 2 and 4 are always the exact
// transpose of the lists, so after each batch's seeding they are rebuilt as
// the reverse CSR of the whole graph (csr.hip): every new point's seeds are
// then on their nodes' reverse lists, as step 2 leaves them.  Step 3 and the
// distance of step 4 run in k_ins_affect (one workgroup per new point, a
// per-workgroup visited bitmap in HBM for the exclusions and the dedup), its
// offers that pass the batch-start filter (d < the last distance of a full
// list) are gathered per target (count, scan, scatter) and merged into the
// lists by the NNDescent merge, which keeps the top K of (list U offers) and
// leaves equal distances already listed ahead of new ones -- what the
// reference's strict-< insertion sort leaves.  An offer the filter rejects
// would be rejected by the reference as well (the last distance of a list
// only falls), so for the offers of a batch the lists end as the reference
// would leave them after applying the same offers one by one, up to the
// order of equal distances.
//
// With bidirectional (Opt-2, hrnn.h:2410-2418 and 2471-2530) x's own list
// becomes the K closest of its list and its candidates: the workgroup keeps
// x's best keys in LDS (a buffer of 2K keys compacted to the best K by a
// bitonic sort whenever it fills, admitting only keys below the K-th best
// kept so far) and writes them when the scan ends.  The reference does not
// remove x from the reverse lists of nodes its list drops (stale entries it
// calls benign); the CSR rebuilt here has no stale entries.
//
// Distances are the reference's: exact integer sums for 8-bit data (d <= 258,
// any order), else FP32 in the reference's 16-partial order.
#include <algorithm>
#include <cfloat>
#include <chrono>
#include <vector>

#include "handle.h"

namespace hrnn {

constexpr uint32_t kInsThreads = 512;
constexpr uint32_t kInsMaxK = 1024;     // longest list (the merge's limit)
constexpr uint32_t kInsPerThread = 2;   // reverse-list entries per thread per scan round
constexpr uint32_t kInsRound = kInsThreads * kInsPerThread;
constexpr uint32_t kInsQueue = 2 * kInsRound;   // LDS queue of fresh candidates
constexpr uint32_t kInsBest = 2 * kInsMaxK;     // bidirectional: LDS key buffer

enum { IS_SCANNED = 0, IS_CAND = 1, IS_PROP = 2, IS_N = 3 };
// static LDS of k_ins_affect (bytes), for the host's LDS budget
constexpr size_t kInsStaticLds = (kInsMaxK + 1) * 4 + kInsMaxK * 8 + kInsQueue * 4 + 16 * 4 + 64;

__device__ __forceinline__ unsigned long long bcast_u64(unsigned long long v) {
    const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)v, 0, 64);
    const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(v >> 32), 0, 64);
    return ((unsigned long long)hi << 32) | lo;
}

// Sum of squared byte differences of two 16-byte vectors (exact integer).
__device__ __forceinline__ uint32_t sqdiff16(uint4 a, uint4 b) {
    uint32_t acc = 0;
    const uint32_t aw[4] = {a.x, a.y, a.z, a.w}, bw[4] = {b.x, b.y, b.z, b.w};
#pragma unroll
    for (int w = 0; w < 4; w++)
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int t = (int)((aw[w] >> (8 * k)) & 255u) - (int)((bw[w] >> (8 * k)) & 255u);
            acc += (uint32_t)(t * t);
        }
    return acc;
}

// Bitonic sort of the u64 keys s[0 .. P) ascending (P a power of two, the
// whole workgroup).
__device__ void ins_sort_keys(unsigned long long *s, uint32_t P) {
    for (uint32_t k = 2; k <= P; k <<= 1)
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < P; i += blockDim.x) {
                const uint32_t l = i ^ j;
                if (l > i) {
                    const unsigned long long a = s[i], b = s[l];
                    const bool up = (i & k) == 0;
                    if ((a > b) == up) {
                        s[i] = b;
                        s[l] = a;
                    }
                }
            }
            __syncthreads();
        }
}

// One workgroup (512 lanes) per new point (points b0 + w, b0 + w +
// gridDim.x, ...).  The visited bitmap of n bits lives in LDS (kLdsBm) or,
// for graphs too large for that, in HBM (bm_words u32 per workgroup, a
// multiple of 4; all zero between points).  Offers go to (out_tgt, out_key)
// at a slot taken from *out_n; slots at or past out_cap are counted but not
// written.  With bidir the point's new list goes to row p - b0 of (bd_ids,
// bd_d, bd_cnt) [K per row]; the point's own list is not changed here.
//
// Distances: 8-bit rows with d % 16 == 0, d <= 256 go L = d/16 rounded up to
// a power of two lanes per row, 16 bytes per lane, four rows per lane group
// in flight (exact integers); other 8-bit rows by 16-lane groups 4 bytes at a
// time (exact); FP32 rows in the reference's 16-partial order.
template <bool kU8, class EntT, bool kLdsBm>
__global__ __launch_bounds__(kInsThreads) void k_ins_affect(
    const uint8_t *__restrict__ xu8, const float *__restrict__ xf, uint32_t d, uint32_t n,
    uint32_t *__restrict__ knn_ids, float *__restrict__ knn_d, uint32_t *__restrict__ knn_cnt, uint32_t K,
    const float *__restrict__ worst, const uint64_t *__restrict__ csr_off, const EntT *__restrict__ csr_ent,
    uint32_t rank_bits, uint32_t b0, uint32_t b1, uint32_t m_update, int bidir, uint32_t *__restrict__ bitmaps,
    uint32_t bm_words, uint32_t *__restrict__ out_tgt, unsigned long long *__restrict__ out_key,
    unsigned long long out_cap, unsigned long long *__restrict__ out_n, uint32_t *__restrict__ bd_ids,
    float *__restrict__ bd_d, uint32_t *__restrict__ bd_cnt, unsigned long long *__restrict__ stats,
    uint32_t *__restrict__ err) {
    __shared__ uint32_t s_pref[kInsMaxK + 1];   // prefix of the proxies' reverse-list lengths
    __shared__ uint64_t s_lo[kInsMaxK];         // start of each proxy's reverse list
    __shared__ uint32_t s_q[kInsQueue];         // fresh candidates
    __shared__ uint32_t s_qn, s_bad, s_bn;
    __shared__ uint32_t s_wsum[kInsThreads / 64];
    extern __shared__ __align__(16) unsigned char s_dyn[];
    // dynamic LDS: the point (FP32 [d] or packed bytes), with bidir the key
    // buffer, with kLdsBm the bitmap
    float *s_qf = reinterpret_cast<float *>(s_dyn);
    uint32_t *s_q8 = reinterpret_cast<uint32_t *>(s_dyn);
    const uint32_t q_bytes = kU8 ? ((d + 15) / 16) * 16 : (d * 4 + 15) / 16 * 16;
    unsigned long long *s_best = reinterpret_cast<unsigned long long *>(s_dyn + q_bytes);
    uint32_t *bm = kLdsBm ? reinterpret_cast<uint32_t *>(s_dyn + q_bytes + (bidir ? kInsBest * 8 : 0))
                          : bitmaps + (uint64_t)blockIdx.x * bm_words;

    const uint32_t tid = threadIdx.x, lane = tid & 63u, wv = tid >> 6;
    const uint32_t g = lane & 15u, grp = lane >> 4;
    constexpr uint32_t kWaves = kInsThreads / 64;
    // 8-bit rows 16 bytes per lane: L lanes per row, G groups per wave
    const bool wide_u8 = kU8 && (d % 16 == 0) && d <= 256;
    uint32_t L = 1;
    while (L * 16 < d) L <<= 1;
    const uint32_t G = 64 / L, gl = lane & (L - 1), gi = lane / L;
    unsigned long long st_scan = 0, st_cand = 0, st_prop = 0;
    if (tid == 0) s_bad = 0;
    if constexpr (kLdsBm) {   // the LDS bitmap starts clear
        for (uint32_t w = tid; w < bm_words / 4; w += kInsThreads)
            reinterpret_cast<uint4 *>(bm)[w] = make_uint4(0u, 0u, 0u, 0u);
        __syncthreads();
    }

    for (uint32_t p = b0 + blockIdx.x; p < b1; p += gridDim.x) {
        // ---- stage the point, mark it and its list (excluded) ----
        if constexpr (kU8) {
            for (uint32_t w = tid; 4 * w < d; w += kInsThreads)
                s_q8[w] = reinterpret_cast<const uint32_t *>(xu8 + (uint64_t)p * d)[w];
        } else {
            for (uint32_t e = tid; e < d; e += kInsThreads) s_qf[e] = xf[(uint64_t)p * d + e];
        }
        const uint32_t cnt = min(knn_cnt[p], K);
        const uint32_t np = (m_update > 0 && m_update < cnt) ? m_update : cnt;
        if (tid == 0) atomicOr(&bm[p >> 5], 1u << (p & 31));
        for (uint32_t t = tid; t < cnt; t += kInsThreads) {
            const uint32_t y = knn_ids[(uint64_t)p * K + t];
            if (y >= n) {
                s_bad = 1u;
                continue;
            }
            atomicOr(&bm[y >> 5], 1u << (y & 31));
        }
        // proxies' reverse lists: lengths, then an exclusive scan (2 per thread)
        constexpr uint32_t kPP = (kInsMaxK + kInsThreads - 1) / kInsThreads;
        uint32_t lenk[kPP], part = 0;
#pragma unroll
        for (uint32_t k = 0; k < kPP; k++) {
            const uint32_t j = tid * kPP + k;
            lenk[k] = 0;
            if (j < np) {
                const uint32_t y = knn_ids[(uint64_t)p * K + j];
                if (y < n) {
                    const uint64_t lo = csr_off[y], hi = csr_off[y + 1];
                    s_lo[j] = lo;
                    lenk[k] = (uint32_t)(hi - lo);
                }
            }
            part += lenk[k];
        }
        uint32_t incl = part;
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t y = (uint32_t)__shfl_up((int)incl, off, 64);
            if (lane >= (uint32_t)off) incl += y;
        }
        if (lane == 63) s_wsum[wv] = incl;
        if (tid == 0) {
            s_qn = 0;
            s_bn = 0;
        }
        if constexpr (!kLdsBm) __threadfence();   // the exclusion marks are in place before any scan reads them
        __syncthreads();
        uint32_t run = incl - part, tot = 0;
        for (uint32_t w = 0; w < kWaves; w++) {
            if (w < wv) run += s_wsum[w];
            tot += s_wsum[w];
        }
#pragma unroll
        for (uint32_t k = 0; k < kPP; k++) {
            const uint32_t j = tid * kPP + k;
            if (j < np) s_pref[j] = run;
            run += lenk[k];
        }
        if (tid == 0) s_pref[np] = tot;
        if (bidir) {
            // x's list as the first keys of its best set
            for (uint32_t t = tid; t < cnt; t += kInsThreads)
                s_best[t] = ((unsigned long long)float_key(knn_d[(uint64_t)p * K + t]) << 32) | knn_ids[(uint64_t)p * K + t];
            if (tid == 0) s_bn = cnt;
        }
        __syncthreads();
        const uint32_t T = s_pref[np];
        st_scan += T;
        float thr = FLT_MAX;   // bidirectional: the K-th best key's distance once K are kept
        if (bidir && cnt >= K) thr = knn_d[(uint64_t)p * K + K - 1];

        // ---- scan in rounds of kInsRound entries; fresh ids queue up in LDS ----
        for (uint32_t base = 0; base < T; base += kInsRound) {
            uint32_t cc[kInsPerThread];
            bool ok[kInsPerThread];
            EntT ent[kInsPerThread];
#pragma unroll
            for (uint32_t k = 0; k < kInsPerThread; k++) {   // every load issued before any is used
                const uint32_t pos = base + k * kInsThreads + tid;
                ok[k] = pos < T;
                const uint32_t ps = ok[k] ? pos : 0;
                uint32_t lo = 0, hi = np;   // the proxy j with s_pref[j] <= pos < s_pref[j + 1]
                while (hi - lo > 1) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (s_pref[mid] <= ps) lo = mid;
                    else hi = mid;
                }
                ent[k] = ok[k] ? csr_ent[s_lo[lo] + (ps - s_pref[lo])] : (EntT)0;
            }
#pragma unroll
            for (uint32_t k = 0; k < kInsPerThread; k++) {
                bool fresh = false;
                cc[k] = (uint32_t)(ent[k] >> rank_bits);
                if (ok[k]) {
                    if (cc[k] >= n) {
                        s_bad = 1u;
                    } else {
                        const uint32_t bit = 1u << (cc[k] & 31);
                        fresh = (atomicOr(&bm[cc[k] >> 5], bit) & bit) == 0;
                    }
                }
                const uint64_t bal = __ballot(fresh);
                if (bal) {
                    uint32_t at = 0;
                    if (lane == 0) at = atomicAdd(&s_qn, (uint32_t)__popcll(bal));
                    at = (uint32_t)__shfl((int)at, 0, 64);
                    if (fresh) s_q[at + (uint32_t)__popcll(bal & lanemask_lt())] = cc[k];
                }
            }
            __syncthreads();
            const uint32_t qn = s_qn;
            __syncthreads();   // read by every thread before the next round adds to it
            if (qn + kInsRound <= kInsQueue && base + kInsRound < T) continue;   // uniform
            // ---- distances of the queued candidates ----
            st_cand += qn;
            const uint32_t per_pass = wide_u8 ? kWaves * G * 4 : kWaves * 4;
            for (uint32_t q0 = 0; q0 < qn; q0 += per_pass) {
                uint32_t cidx[4];
                float dd[4];
                bool lead;   // this lane reports the group's candidates
                if (wide_u8) {
                    // group gi of wave wv: candidates q0 + (wv * G + gi) * 4 + r, r < 4
                    uint4 rows[4];
                    const uint4 qv = gl * 16 < d ? reinterpret_cast<const uint4 *>(s_q8)[gl] : make_uint4(0, 0, 0, 0);
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const uint32_t qi = q0 + (wv * G + gi) * 4 + r;
                        cidx[r] = qi < qn ? qi : kInvalidId;
                        const uint32_t c = s_q[qi < qn ? qi : q0];
                        rows[r] = gl * 16 < d ? reinterpret_cast<const uint4 *>(xu8 + (uint64_t)c * d)[gl]
                                              : make_uint4(0, 0, 0, 0);
                    }
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        uint32_t acc = sqdiff16(qv, rows[r]);
                        for (uint32_t s2 = 1; s2 < L; s2 <<= 1) acc += (uint32_t)__shfl_xor((int)acc, (int)s2, 64);
                        dd[r] = (float)acc;
                    }
                    lead = gl == 0;
                } else {
                    // 16-lane group grp of wave wv: candidate q0 + wv * 4 + grp
                    const uint32_t qi = q0 + wv * 4 + grp;
                    const uint32_t c = s_q[qi < qn ? qi : q0];
                    float dv;
                    if constexpr (kU8) {
                        const uint32_t *row = reinterpret_cast<const uint32_t *>(xu8 + (uint64_t)c * d);
                        uint32_t acc = 0;
                        for (uint32_t w = g; 4 * w < d; w += 16) {
                            const uint32_t a = s_q8[w], b = row[w];
#pragma unroll
                            for (int k = 0; k < 4; k++) {
                                const int t = (int)((a >> (8 * k)) & 255u) - (int)((b >> (8 * k)) & 255u);
                                acc += (uint32_t)(t * t);
                            }
                        }
#pragma unroll
                        for (int s2 = 1; s2 < 16; s2 <<= 1) acc += (uint32_t)__shfl_xor((int)acc, s2, 64);
                        dv = (float)acc;
                    } else {
                        const float *row = xf + (uint64_t)c * d;
                        dv = ((d & 15u) == 0) ? l2_ref16_group(s_qf, row, d, (int)g) : l2_seq(s_qf, row, d);
                    }
                    cidx[0] = qi < qn ? qi : kInvalidId;
                    cidx[1] = cidx[2] = cidx[3] = kInvalidId;
                    dd[0] = dv;
                    dd[1] = dd[2] = dd[3] = 0.f;
                    lead = g == 0;
                }
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const bool valid = lead && cidx[r] != kInvalidId;
                    const uint32_t c = valid ? s_q[cidx[r]] : 0u;
                    const bool offer = valid && dd[r] < worst[c];
                    const uint64_t ob = __ballot(offer);
                    if (ob) {
                        unsigned long long at = 0;
                        if (lane == 0) at = atomicAdd(out_n, (unsigned long long)__popcll(ob));
                        at = bcast_u64(at);
                        if (offer) {
                            const unsigned long long slot = at + (unsigned long long)__popcll(ob & lanemask_lt());
                            if (slot < out_cap) {
                                out_tgt[slot] = c;
                                out_key[slot] = ((unsigned long long)float_key(dd[r]) << 32) | p;
                            }
                        }
                        if (lane == 0) st_prop += (unsigned long long)__popcll(ob);
                    }
                    if (bidir) {
                        // keys below the running K-th best join x's buffer
                        const bool keep = valid && dd[r] < thr;
                        const uint64_t kb = __ballot(keep);
                        if (kb) {
                            uint32_t at = 0;
                            if (lane == 0) at = atomicAdd(&s_bn, (uint32_t)__popcll(kb));
                            at = (uint32_t)__shfl((int)at, 0, 64);
                            const uint32_t slot = at + (uint32_t)__popcll(kb & lanemask_lt());
                            if (keep && slot < kInsBest)
                                s_best[slot] = ((unsigned long long)float_key(dd[r]) << 32) | c;
                        }
                    }
                }
                if (bidir) {
                    __syncthreads();
                    const uint32_t bn_now = s_bn;
                    __syncthreads();   // read by every thread before the next pass adds to it
                    if (bn_now + per_pass > kInsBest) {   // uniform: compact to the best K
                        const uint32_t bn = min(bn_now, kInsBest);
                        for (uint32_t t = bn + tid; t < kInsBest; t += kInsThreads) s_best[t] = ~0ull;
                        __syncthreads();
                        ins_sort_keys(s_best, kInsBest);
                        if (bn >= K) thr = key_float((uint32_t)(s_best[K - 1] >> 32));
                        __syncthreads();
                        if (tid == 0) s_bn = min(bn, K);
                        __syncthreads();
                    }
                }
            }
            __syncthreads();
            if (tid == 0) s_qn = 0;
            __syncthreads();
        }

        // ---- bidirectional: x's list = the K best of its buffer ----
        if (bidir) {
            const uint32_t bn = min(s_bn, kInsBest);
            for (uint32_t t = bn + tid; t < kInsBest; t += kInsThreads) s_best[t] = ~0ull;
            __syncthreads();
            ins_sort_keys(s_best, kInsBest);
            // drop repeats (an id listed and met again has the same distance,
            // so the copies are adjacent) and x itself, keep the first K
            constexpr uint32_t kPB = kInsBest / kInsThreads;
            uint32_t keepf[kPB];
            uint32_t part2 = 0;
#pragma unroll
            for (uint32_t k = 0; k < kPB; k++) {
                const uint32_t t = tid * kPB + k;
                const unsigned long long v = s_best[t];
                keepf[k] = (v != ~0ull && (uint32_t)v != p && (t == 0 || s_best[t - 1] != v)) ? 1u : 0u;
                part2 += keepf[k];
            }
            uint32_t inc2 = part2;
            for (int off = 1; off < 64; off <<= 1) {
                const uint32_t y = (uint32_t)__shfl_up((int)inc2, off, 64);
                if (lane >= (uint32_t)off) inc2 += y;
            }
            __syncthreads();
            if (lane == 63) s_wsum[wv] = inc2;
            __syncthreads();
            uint32_t r2 = inc2 - part2;
            for (uint32_t w = 0; w < wv; w++) r2 += s_wsum[w];
            uint32_t total = 0;
            for (uint32_t w = 0; w < kWaves; w++) total += s_wsum[w];
            const uint64_t row = (uint64_t)(p - b0) * K;
#pragma unroll
            for (uint32_t k = 0; k < kPB; k++) {
                const uint32_t t = tid * kPB + k;
                if (keepf[k] && r2 < K) {
                    const unsigned long long v = s_best[t];
                    bd_ids[row + r2] = (uint32_t)v;
                    bd_d[row + r2] = key_float((uint32_t)(v >> 32));
                }
                r2 += keepf[k];
            }
            if (tid == 0) bd_cnt[p - b0] = min(total, K);
        }

        // ---- clear the bitmap: whole when that is cheaper than the walk ----
        __syncthreads();
        if (kLdsBm || (uint64_t)bm_words <= (uint64_t)T + cnt) {
            uint4 *bm4 = reinterpret_cast<uint4 *>(bm);
            for (uint32_t w = tid; w < bm_words / 4; w += kInsThreads) bm4[w] = make_uint4(0u, 0u, 0u, 0u);
        } else {
            if (tid == 0) bm[p >> 5] = 0;
            for (uint32_t t = tid; t < cnt; t += kInsThreads) {
                const uint32_t y = knn_ids[(uint64_t)p * K + t];
                if (y < n) bm[y >> 5] = 0;
            }
            for (uint32_t pos = tid; pos < T; pos += kInsThreads) {
                uint32_t lo = 0, hi = np;
                while (hi - lo > 1) {
                    const uint32_t mid = (lo + hi) >> 1;
                    if (s_pref[mid] <= pos) lo = mid;
                    else hi = mid;
                }
                const uint32_t c = (uint32_t)(csr_ent[s_lo[lo] + (pos - s_pref[lo])] >> rank_bits);
                if (c < n) bm[c >> 5] = 0;
            }
        }
        if constexpr (!kLdsBm) __threadfence();   // cleared before the next point's marks
        __syncthreads();
    }
    if (tid == 0) {   // st_scan and st_cand are the same in every lane: counted once
        atomicAdd(&stats[IS_SCANNED], st_scan);
        atomicAdd(&stats[IS_CAND], st_cand);
    }
    if (lane == 0 && st_prop) atomicAdd(&stats[IS_PROP], st_prop);
    if (tid == 0 && s_bad) atomicOr(err, 1u);
}

// Bidirectional: the batch's new lists replace the points' lists.
__global__ void k_ins_commit(const uint32_t *__restrict__ bd_ids, const float *__restrict__ bd_d,
                             const uint32_t *__restrict__ bd_cnt, uint32_t b0, uint32_t nb, uint32_t K,
                             uint32_t *__restrict__ ids, float *__restrict__ kd, uint32_t *__restrict__ cnt,
                             uint8_t *__restrict__ isnew) {
    const uint32_t r = blockIdx.x;
    if (r >= nb) return;
    const uint32_t c = min(bd_cnt[r], K);
    const uint64_t dst = (uint64_t)(b0 + r) * K, src = (uint64_t)r * K;
    for (uint32_t t = threadIdx.x; t < c; t += blockDim.x) {
        ids[dst + t] = bd_ids[src + t];
        kd[dst + t] = bd_d[src + t];
        isnew[dst + t] = 1;
    }
    if (threadIdx.x == 0) cnt[b0 + r] = c;
}

// Offers per target.
__global__ void k_ins_count(const uint32_t *__restrict__ tgt, uint64_t m, uint32_t *__restrict__ pcnt) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (uint64_t)gridDim.x * blockDim.x)
        atomicAdd(&pcnt[tgt[i]], 1u);
}

// Offers into their target's range of the proposal CSR.
__global__ void k_ins_scatter(const uint32_t *__restrict__ tgt, const unsigned long long *__restrict__ key,
                              uint64_t m, const uint64_t *__restrict__ poff, uint32_t *__restrict__ pfill,
                              unsigned long long *__restrict__ pkey) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t t = tgt[i];
        pkey[poff[t] + atomicAdd(&pfill[t], 1u)] = key[i];
    }
}

// Last distance of each full list (FLT_MAX for the others).
__global__ void k_ins_worst(const float *__restrict__ kd, const uint32_t *__restrict__ cnt, uint64_t n, uint32_t K,
                            float *__restrict__ worst) {
    const uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (u < n) worst[u] = cnt[u] >= K ? kd[u * K + (K - 1)] : FLT_MAX;
}

// Buffers and counters of one insertion call; the affected-area step of
// one batch of rows [b, b1) whose lists hold their seeds runs in run().
struct InsertCall {
    hrnn_dev_t h;
    uint64_t n;
    uint32_t K, d, m_update, slots_max, bm_words;
    int bidir;
    bool u8, lds_bm;   // lds_bm: the visited bitmap in LDS (n bits fit beside the rest)
    size_t lds;        // dynamic LDS of a launch
    uint64_t cap;
    DevBuf bitmaps, worst, pcnt, poff, pkey, otgt, okey, ctr, bdi, bdd, bdc;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    hrnn_insert_stats s{};

    InsertCall(hrnn_dev_t hh, uint32_t mu, int bd) : h(hh), m_update(mu), bidir(bd ? 1 : 0) {
        REQUIRE(h->K >= 1 && h->K <= kInsMaxK, HRNN_E_STATE, "insert: the resident KNN graph needs 1 <= K <= 1024");
        n = h->n;
        K = h->K;
        d = h->d;
        REQUIRE(n * K < 0xFFFFFFFFull, HRNN_E_ARG, "insert: n * K must stay below 2^32");
        u8 = h->base_u8 && d <= 258 && d % 4 == 0;
        const size_t base_lds = (u8 ? ((d + 15) / 16) * 16 : ((size_t)d * 4 + 15) / 16 * 16) + (bidir ? kInsBest * 8 : 0);
        REQUIRE(base_lds + kInsStaticLds <= 65536, HRNN_E_ARG,
                "insert: the point's row does not fit in LDS (d too large)");
        hipStream_t st = h->stream;
        bm_words = (uint32_t)((n + 127) / 128) * 4;
        const size_t lds_cap = std::max<size_t>(h->lds_per_block, 65536);
        lds_bm = h->tune_insert_lds_bitmap != 0 && base_lds + (size_t)bm_words * 4 + kInsStaticLds <= lds_cap;
        const uint32_t n_cu = (uint32_t)std::max(1, h->n_cu);
        if (lds_bm) {
            lds = base_lds + (size_t)bm_words * 4;
            // workgroups that fit a CU's LDS at once (two waves per SIMD at most)
            const uint32_t per_cu = (uint32_t)std::min<size_t>(2, std::max<size_t>(1, lds_cap / (lds + kInsStaticLds)));
            slots_max = n_cu * per_cu;
            if (lds > 65536) {
                const hipError_t e1 = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_ins_affect<true, uint32_t, true>),
                                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
                const hipError_t e2 = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_ins_affect<true, uint64_t, true>),
                                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
                const hipError_t e3 = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_ins_affect<false, uint32_t, true>),
                                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
                const hipError_t e4 = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_ins_affect<false, uint64_t, true>),
                                                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
                if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess || e4 != hipSuccess) {
                    (void)hipGetLastError();
                    lds_bm = false;   // the runtime refuses that much: bitmaps in HBM
                }
            }
        }
        if (!lds_bm) {
            lds = base_lds;
            slots_max = n_cu * 2;
            bitmaps.reserve((uint64_t)slots_max * bm_words * 4);
            HIPC(hipMemsetAsync(bitmaps.p, 0, (uint64_t)slots_max * bm_words * 4, st));
        }
        worst.reserve(n * 4);
        pcnt.reserve(n * 8 + 64);   // counts, then fill positions
        poff.reserve((n + 1 + scan_block_count(n) + 1) * 8 + 64);
        cap = h->tune_insert_offer_cap > 0 ? (uint64_t)h->tune_insert_offer_cap
                                           : std::max<uint64_t>(1 << 22, n * 24);   // offers per batch (grows on demand)
        otgt.reserve(cap * 4);
        okey.reserve(cap * 8);
        pkey.reserve(cap * 8);
        ctr.reserve((16 + nnd_stat_words()) * 8);
        HIPC(hipMemsetAsync(ctr.p, 0, (16 + nnd_stat_words()) * 8, st));
        for (auto &e : ev) HIPC(hipEventCreate(&e));
    }
    ~InsertCall() {
        for (DevBuf *b : {&bitmaps, &worst, &pcnt, &poff, &pkey, &otgt, &okey, &ctr, &bdi, &bdd, &bdc}) b->release();
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
    }
    double elapsed(int a, int b) {
        float ms = 0.f;
        HIPC(hipEventElapsedTime(&ms, ev[a], ev[b]));
        return (double)ms;
    }

    void run(uint32_t b, uint32_t b1) {
        hipStream_t st = h->stream;
        unsigned long long *c_out = ctr.as<unsigned long long>();   // [0] the batch's offers
        unsigned long long *c_st = c_out + 8;                       // [8..10] its scanned / candidates / offers
        unsigned long long *c_merge = c_out + 16;                   // merge counters (whole call)
        uint32_t *err = h->counters.as<uint32_t>() + kErrWord;
        uint32_t *pc = pcnt.as<uint32_t>(), *pf = pc + n;
        const uint32_t nblk = scan_block_count(n);
        uint64_t *po = poff.as<uint64_t>(), *bsum = po + n + 1, *ptot = bsum + nblk;

        // the reverse lists of the graph as it stands (the batch's seeds included)
        HIPC(hipEventRecord(ev[0], st));
        csr_build(h);
        HIPC(hipEventRecord(ev[1], st));
        hipLaunchKernelGGL(k_ins_worst, dim3(ceil_div(n, 256)), dim3(256), 0, st, h->knn_d.as<float>(),
                           h->knn_cnt.as<uint32_t>(), n, K, worst.as<float>());
        const uint32_t nb = b1 - b;
        const uint32_t grid = std::min<uint32_t>(slots_max, nb);
        if (bidir) {
            bdi.reserve((uint64_t)nb * K * 4);
            bdd.reserve((uint64_t)nb * K * 4);
            bdc.reserve((uint64_t)nb * 4);
        }
        unsigned long long total = 0, bc[16];
        for (;;) {
            HIPC(hipMemsetAsync(c_out, 0, 16 * 8, st));
#define LAUNCH(U8, ET, LB)                                                                                           \
    hipLaunchKernelGGL((k_ins_affect<U8, ET, LB>), dim3(grid), dim3(kInsThreads), lds, st, h->xu8.as<uint8_t>(),     \
                       h->x.as<float>(), d, (uint32_t)n, h->knn_ids.as<uint32_t>(), h->knn_d.as<float>(),            \
                       h->knn_cnt.as<uint32_t>(), K, worst.as<float>(), h->csr_off.as<uint64_t>(),                    \
                       h->csr_ent.as<ET>(), h->rank_bits, b, b1, m_update, bidir, bitmaps.as<uint32_t>(), bm_words,   \
                       otgt.as<uint32_t>(), okey.as<unsigned long long>(), (unsigned long long)cap, c_out,           \
                       bdi.as<uint32_t>(), bdd.as<float>(), bdc.as<uint32_t>(), c_st, err)
            if (lds_bm) {
                if (h->wide) {
                    if (u8) LAUNCH(true, uint64_t, true);
                    else LAUNCH(false, uint64_t, true);
                } else {
                    if (u8) LAUNCH(true, uint32_t, true);
                    else LAUNCH(false, uint32_t, true);
                }
            } else {
                if (h->wide) {
                    if (u8) LAUNCH(true, uint64_t, false);
                    else LAUNCH(false, uint64_t, false);
                } else {
                    if (u8) LAUNCH(true, uint32_t, false);
                    else LAUNCH(false, uint32_t, false);
                }
            }
#undef LAUNCH
            HIPC(hipGetLastError());
            d2h(h, bc, c_out, sizeof(bc));
            total = bc[0];
            if (total <= cap) break;
            // the offers did not fit: a larger buffer and the batch's scan
            // again (it changed nothing but the offer buffer, the staged
            // lists and the batch's counters)
            cap = total + total / 4;
            otgt.release();
            okey.release();
            pkey.release();
            otgt.reserve(cap * 4);
            okey.reserve(cap * 8);
            pkey.reserve(cap * 8);
            s.retries++;
        }
        if (bidir)
            hipLaunchKernelGGL(k_ins_commit, dim3(nb), dim3(256), 0, st, bdi.as<uint32_t>(), bdd.as<float>(),
                               bdc.as<uint32_t>(), b, nb, K, h->knn_ids.as<uint32_t>(), h->knn_d.as<float>(),
                               h->knn_cnt.as<uint32_t>(), h->knn_new.as<uint8_t>());
        HIPC(hipEventRecord(ev[2], st));
        // offers per target -> proposal CSR -> merge
        HIPC(hipMemsetAsync(pc, 0, n * 8, st));
        if (total) {
            const unsigned cg = std::min<unsigned>(ceil_div(total, 256), 8192);
            hipLaunchKernelGGL(k_ins_count, dim3(cg), dim3(256), 0, st, otgt.as<uint32_t>(), (uint64_t)total, pc);
            scan_u32(h, pc, n, po, bsum, ptot);
            hipLaunchKernelGGL(k_ins_scatter, dim3(cg), dim3(256), 0, st, otgt.as<uint32_t>(),
                               okey.as<unsigned long long>(), (uint64_t)total, po, pf, pkey.as<unsigned long long>());
            HIPC(hipGetLastError());
            nnd_merge_launch(h, n, pc, po, pkey.as<unsigned long long>(), total, c_merge);
        }
        HIPC(hipEventRecord(ev[3], st));
        HIPC(hipEventSynchronize(ev[3]));
        s.csr_ms += elapsed(0, 1);
        s.affect_ms += elapsed(1, 2);
        s.merge_ms += elapsed(2, 3);
        s.offers += total;
        s.scanned += bc[8 + IS_SCANNED];
        s.candidates += bc[8 + IS_CAND];
        s.batches++;
        h->knn_version++;
    }

    // counters of the whole call; the reverse CSR is stale (the lists
    // changed after its last rebuild)
    void finish(hrnn_insert_stats *out, std::chrono::steady_clock::time_point t0) {
        std::vector<unsigned long long> hc(nnd_stat_words());
        d2h(h, hc.data(), ctr.as<unsigned long long>() + 16, hc.size() * 8);
        check_device_error(h);
        h->csr_valid = false;
        h->knn_version++;
        if (out) {
            s.updates = nnd_stat_updates(hc.data());
            s.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            *out = s;
        }
    }
};

}  // namespace hrnn

using namespace hrnn;

extern "C" {

int hrnn_dev_insert_dynamic(hrnn_dev_t h, uint64_t n_old, const int *levels, uint32_t M, uint32_t ef_construction,
                            uint32_t m_update, int bidirectional, uint32_t *entry_out, int *maxlevel_out,
                            hrnn_insert_stats *stats) {
    if (!h) return HRNN_E_ARG;
    return guarded(h, [&] {
        const auto t_start = std::chrono::steady_clock::now();
        REQUIRE(levels && entry_out && maxlevel_out, HRNN_E_ARG, "insert: null argument");
        REQUIRE(n_old > 0 && n_old < h->n, HRNN_E_ARG, "insert: append the new rows first (n_old < n)");
        InsertCall call(h, m_update, bidirectional);
        double t_hnsw = 0.0;
        auto t_batch = std::chrono::steady_clock::now();
        auto after = [&](uint32_t b, uint32_t b1) {
            // the batch's HNSW links and seeds are enqueued; the step below
            // synchronises, so wall time up to its start is the batch's HNSW part
            HIPC(hipStreamSynchronize(h->stream));
            t_hnsw += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_batch).count();
            call.run(b, b1);
            t_batch = std::chrono::steady_clock::now();
        };
        hnsw_insert_rows(h, levels, n_old, M, ef_construction, h->K, &call.s.hnsw, after);
        call.s.hnsw_ms = t_hnsw;
        call.finish(stats, t_start);
        *entry_out = h->entry;
        *maxlevel_out = h->maxlevel;
    });
}

int hrnn_dev_insert_affect(hrnn_dev_t h, uint64_t b0, uint64_t b1, uint32_t m_update, int bidirectional,
                           hrnn_insert_stats *stats) {
    if (!h) return HRNN_E_ARG;
    return guarded(h, [&] {
        const auto t_start = std::chrono::steady_clock::now();
        REQUIRE(h->K >= 1, HRNN_E_STATE, "insert_affect: no KNN graph on the device");
        REQUIRE(b0 < b1 && b1 <= h->n && h->n < kInvalidId, HRNN_E_ARG, "insert_affect: rows out of range");
        InsertCall call(h, m_update, bidirectional);
        call.run((uint32_t)b0, (uint32_t)b1);
        call.finish(stats, t_start);
    });
}

}  // extern "C"
