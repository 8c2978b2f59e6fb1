// nndescent.hip -- (a) NNDescent refinement of the ranked KNN graph.
//
// Reference: HRNN::refineKNNG (include/hnswlib/hrnn.h:1725-1862) with
// try_update_knng_nnd (hrnn.h:1660-1722).  Per iteration:
//   UPDATE (hrnn.h:1765-1810)  k_nnd_sample: one wave per node walks its list
//       in rank order; the first S new entries become nn_new, the first S old
//       ones nn_old; every new flag is cleared; entry (i -> nbr) is offered to
//       nbr's reverse sample (new or old) when dist(i, nbr) > nbr's K-th
//       distance (FLT_MAX unless nbr's list is full).  Two passes (new
//       entries, then old ones): the old pass skips nodes whose join has no
//       new member, which changes no join.  The reference keeps the
//       first R = S offers, which a single thread makes in increasing i; here
//       up to 256 offers are collected per node and the join keeps the R
//       smallest source ids, the same set whenever a node gets <= 256 offers.
//   MERGE (hrnn.h:1813-1819)   the join reads forward + reverse samples
//       directly (old side cut at 2R).
//   JOIN (hrnn.h:1823-1853)    k_nnd_join: one workgroup per node stages the
//       <= 4S sample vectors in LDS (BF16, 64-dim chunks) and forms their Gram
//       matrix on MFMA (v_mfma_f32_16x16x32_bf16): rows = new samples, columns
//       = new (upper triangle) and old samples.  d2 = |a|^2 + |b|^2 - 2 a.b.
//       A pair (a, b), a != b, is proposed to a's list when d2 < worst(a) and
//       to b's list when d2 < worst(b), worst taken at the iteration start.
//       The join runs twice: a counting pass (proposals per target, summed
//       per sample slot in LDS, one global atomic per slot), an exclusive
//       scan into a proposal CSR, then a filling pass in which each slot
//       reserves its range with one atomic and writes its proposals there.
//       Nothing is dropped unless the whole iteration exceeds the proposal
//       buffer (reported as dropped).
//   k_nnd_merge: per node, proposals in chunks of 1024: keep those below the
//       current worst entry, sort them by (d2, id), drop duplicates and ids
//       already listed (found in the run of equal list distances at the
//       proposal's rank), merge with the sorted list
//       (existing entries first on equal distance, as the reference's strict-<
//       insertion leaves them); the first K survive, inserted entries flagged
//       new.  Chunk after chunk this is the top-K of (list U all proposals).
// The reference applies each proposal at once under a per-node lock: an
// offer rejected at some point would be rejected later too (the K-th
// distance only falls), so its final list is the same top-K of (list U
// proposals), up to the order of equal distances.
//
// Distances: for integer data in [0, 255] at d <= 128 (the integer configs)
// the BF16 Gram form is exact FP32 arithmetic (every norm, product and sum an
// integer below 2^24), identical to the reference's direct-difference L2Sqr.  Otherwise each proposal that passes the BF16
// filter is re-computed in the reference's FP32 order before it is counted.
#include <algorithm>
#include <cfloat>
#include <chrono>
#include <vector>

#include "handle.h"

namespace hrnn {

typedef __bf16 bf16x8_n __attribute__((ext_vector_type(8)));
typedef float f32x4_n __attribute__((ext_vector_type(4)));

enum { ST_DISTS = 0, ST_PROPOSALS = 1, ST_DROPPED = 2, ST_UPDATES = 3, ST_ROWS = 4, ST_N = 5 };
// Each counter is spread over kStatSlots words (workgroup b adds to slot
// b % kStatSlots; the host sums them): a million workgroups adding to one
// word would serialise on it.
constexpr uint32_t kStatSlots = 256;
__device__ __forceinline__ unsigned long long *stat_at(unsigned long long *stats, int s) {
    return stats + (uint32_t)s * kStatSlots + (blockIdx.x & (kStatSlots - 1));
}

// worst(u) and per-iteration counter reset
__global__ void k_nnd_prepare(const float *__restrict__ kd, const uint32_t *__restrict__ cnt,
                              uint64_t n, uint32_t K, float *__restrict__ worst,
                              uint32_t *__restrict__ rn_cnt, uint32_t *__restrict__ ro_cnt,
                              uint32_t *__restrict__ prop_cnt, uint32_t *__restrict__ prop_fill,
                              unsigned long long *__restrict__ stats) {
    uint64_t u = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (uint64_t i = u; i < (uint64_t)ST_N * kStatSlots; i += (uint64_t)gridDim.x * blockDim.x) stats[i] = 0;
    if (u >= n) return;
    uint32_t c = cnt[u];
    worst[u] = (c >= K) ? kd[u * K + (K - 1)] : FLT_MAX;
    rn_cnt[u] = 0;
    ro_cnt[u] = 0;
    prop_cnt[u] = 0;
    prop_fill[u] = 0;
}

// *active = 1 when the previous iteration inserted anything (or there is no
// previous one).  After an iteration without insertions no entry carries the
// "new" flag (sampling clears every flag, only insertions set it), so no
// node has a new sample, the join evaluates no pair and the lists stay as
// they are: the iteration can skip its sampling pass.
__global__ void k_nnd_active(const unsigned long long *__restrict__ prev_stats, uint32_t *__restrict__ active) {
    unsigned long long v = prev_stats ? prev_stats[ST_UPDATES * kStatSlots + threadIdx.x] : 1ull;
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    __shared__ unsigned long long w[4];
    if ((threadIdx.x & 63u) == 0) w[threadIdx.x / 64] = v;
    __syncthreads();
    if (threadIdx.x == 0) *active = (w[0] + w[1] + w[2] + w[3]) != 0ull ? 1u : 0u;
}

// Sampling runs in two passes over each node's list (one wave per node).
// Pass 1 (kOld = false) takes the new entries: nn_new and the reverse new
// offers.  k_nnd_need then marks every node whose join has work, i.e. one
// with a new sample or a reverse new offer.  Pass 2 (kOld = true) takes the
// old entries and clears the new flags; it writes nn_old only for a marked
// node and makes a reverse old offer only to a marked node.  Every pair a
// node's join evaluates has a member from that node's new samples or
// reverse new offers, so an unmarked node's join evaluates nothing and what
// pass 2 leaves out never reaches a join: the joins are those of a single
// pass over all entries.  After convergence starts most nodes are unmarked
// and pass 2 skips their random worst[] reads and reverse-offer atomics.
template <bool kOld>
__global__ __launch_bounds__(256) void k_nnd_sample(
    const uint32_t *__restrict__ ids, const float *__restrict__ kd, uint8_t *__restrict__ isnew,
    const uint32_t *__restrict__ cnt, uint64_t n, uint32_t K, uint32_t S, uint32_t RB,
    const float *__restrict__ worst, uint32_t *__restrict__ fs, uint32_t *__restrict__ fs_cnt,
    uint32_t *__restrict__ rs, uint32_t *__restrict__ rs_cnt, const uint32_t *__restrict__ need,
    const uint32_t *__restrict__ active) {
    uint64_t i = (uint64_t)blockIdx.x * 4 + threadIdx.x / 64;
    if (i >= n) return;
    const uint32_t lane = threadIdx.x & 63u;
    if (!*active) {   // nothing new anywhere: empty samples (the join then skips the node)
        if (lane == 0) fs_cnt[i] = 0;
        return;
    }
    const bool need_i = !kOld || ((need[i >> 5] >> (i & 31u)) & 1u);
    const uint64_t lt = lanemask_lt();
    const uint32_t c = cnt[i];
    uint32_t before = 0;
    for (uint32_t base = 0; base < c; base += 64) {
        const uint32_t j = base + lane;
        const bool valid = j < c;
        const bool f = valid && isnew[i * K + j] != 0;
        const bool mine = kOld ? (valid && !f) : f;
        const uint64_t mm = __ballot(mine);
        const uint32_t rnk = before + __popcll(mm & lt);
        if (kOld && f) isnew[i * K + j] = 0;
        if (mine) {
            const uint32_t nbr = ids[i * K + j];
            if (need_i && rnk < S) fs[i * S + rnk] = nbr;
            const bool offer = !kOld || ((need[nbr >> 5] >> (nbr & 31u)) & 1u);
            if (offer && kd[i * K + j] > worst[nbr]) {
                uint32_t s = atomicAdd(&rs_cnt[nbr], 1u);
                if (s < RB) rs[(uint64_t)nbr * RB + s] = (uint32_t)i;
            }
        }
        before += __popcll(mm);
    }
    if (lane == 0) fs_cnt[i] = need_i ? (before < S ? before : S) : 0u;
}

// need bit i = node i has a new sample or a reverse new offer (after pass 1)
__global__ __launch_bounds__(256) void k_nnd_need(const uint32_t *__restrict__ fn_cnt,
                                                  const uint32_t *__restrict__ rn_cnt, uint64_t n,
                                                  uint32_t *__restrict__ need) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool b = i < n && (fn_cnt[i] | rn_cnt[i]) != 0u;
    const uint64_t m = __ballot(b);
    const uint32_t lane = threadIdx.x & 63u;
    if ((lane & 31u) == 0 && i < n) need[i >> 5] = (uint32_t)(lane ? (m >> 32) : m);
}

constexpr int kJMax = 256;      // sample members per node (4S, S <= 64)
constexpr int kJDC = 64;        // dims per LDS chunk
constexpr int kJDP = kJDC + 8;  // padded row stride
constexpr int kJMaxB = 32;      // Gram blocks per wave
constexpr int kJRev = 256;      // reverse offers kept per node and kind

// Sample staging modes of the join.  kJoinChunked: BF16 rows staged 64 dims
// at a time, each wave keeping the accumulators of all its Gram blocks across
// the chunks.  kJoinFullBf16: the whole sample (T rows x d BF16 dims) staged
// at once, each wave evaluating its 16x16 blocks one after another.
// kJoinFullI8: 8-bit data (every value an integer in [0, 255], d <= 258):
// rows staged as int8 after subtracting 128 (x ^ 0x80), half the LDS of BF16,
// Gram blocks on the int8 MFMA with int32 accumulation and squared norms of
// the shifted rows from int8 dot products.  |a - b|^2 = |a'|^2 + |b'|^2 -
// 2 a'.b' is then an exact integer below 2^24: the reference's FP32 value.
enum { kJoinChunked = 0, kJoinFullBf16 = 1, kJoinFullI8 = 2 };
typedef int i32x4_n __attribute__((ext_vector_type(4)));

template <bool kFill, int kMode>
__global__ __launch_bounds__(256) void k_nnd_join(
    const __hip_bfloat16 *__restrict__ xb, const uint8_t *__restrict__ xu8, const float *__restrict__ xn,
    const float *__restrict__ x32,
    uint32_t d, uint64_t n, uint32_t S, uint32_t R, uint32_t RB, const uint32_t *__restrict__ fn,
    const uint32_t *__restrict__ fn_cnt, const uint32_t *__restrict__ fo,
    const uint32_t *__restrict__ fo_cnt, const uint32_t *__restrict__ rn,
    const uint32_t *__restrict__ rn_cnt, const uint32_t *__restrict__ ro,
    const uint32_t *__restrict__ ro_cnt, const float *__restrict__ worst,
    uint32_t *__restrict__ prop_cnt, const uint64_t *__restrict__ prop_off,
    unsigned long long *__restrict__ prop_key, uint64_t cap, int exact, unsigned long long *__restrict__ stats) {
    __shared__ uint32_t sid[kJMax];
    __shared__ float sthr[kJMax], snorm[kJMax];
    __shared__ int s_n8[kJMax];   // kJoinFullI8: squared norms of the shifted rows
    extern __shared__ __align__(16) unsigned char s_dyn[];
    // kJoinFullBf16: Tpad x (d + 8) BF16; kJoinChunked: kJMax x kJDP BF16;
    // kJoinFullI8: Tpad x (roundup(d, 64) + 16) bytes
    __hip_bfloat16 *vec = reinterpret_cast<__hip_bfloat16 *>(s_dyn);
    int8_t *vec8 = reinterpret_cast<int8_t *>(s_dyn);
    constexpr bool kFull = kMode != kJoinChunked;
    __shared__ uint32_t rsel[2][kJRev];   // reverse offers (new, old), sorted by source id
    __shared__ uint8_t blk_r[kJMaxB * 4], blk_c[kJMaxB * 4];
    __shared__ uint32_t s_pcnt[kJMax], s_pcur[kJMax], s_pbase[kJMax];   // per-slot proposals
    __shared__ uint32_t s_a, s_b, s_c, s_nn, s_T, s_nb;

    const uint64_t node = blockIdx.x;
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x / 64;
    if (threadIdx.x == 0) {
        uint32_t a = fn_cnt[node], b = min(rn_cnt[node], R);
        uint32_t c = fo_cnt[node], e = min(ro_cnt[node], R);
        uint32_t nn = a + b;
        uint32_t no = c + e;
        if (no > 2 * R) no = 2 * R;
        s_a = a;
        s_b = b;
        s_c = c;
        s_nn = nn;
        s_T = nn + no;
    }
    __syncthreads();
    const uint32_t nn = s_nn, T = s_T;
    if (nn == 0) return;   // no new sample: no pair to evaluate
    if (!kFill && threadIdx.x == 0) atomicAdd(stat_at(stats, ST_ROWS), (unsigned long long)T);   // rows staged
    // The R smallest source ids among the (<= RB) reverse offers of each kind.
    {
        const uint32_t cn = min(rn_cnt[node], RB), co = min(ro_cnt[node], RB);
        for (uint32_t t = threadIdx.x; t < kJRev; t += blockDim.x) {
            rsel[0][t] = t < cn ? rn[node * RB + t] : kInvalidId;
            rsel[1][t] = t < co ? ro[node * RB + t] : kInvalidId;
        }
        for (uint32_t size = 2; size <= (uint32_t)kJRev; size <<= 1) {
            for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
                __syncthreads();
                for (uint32_t t = threadIdx.x; t < (uint32_t)kJRev; t += blockDim.x) {   // 2 x kJRev/2 pairs
                    const uint32_t arr = t / (kJRev / 2), i = t % (kJRev / 2);
                    const uint32_t lo = 2 * i - (i & (stride - 1)), hi = lo + stride;
                    const bool up = (lo & size) == 0;
                    uint32_t x0 = rsel[arr][lo], x1 = rsel[arr][hi];
                    if ((x0 > x1) == up) {
                        rsel[arr][lo] = x1;
                        rsel[arr][hi] = x0;
                    }
                }
            }
        }
        __syncthreads();
    }
    const uint32_t a = s_a, nnew = s_nn, c = s_c;
    for (uint32_t t = threadIdx.x; t < T; t += blockDim.x) {
        uint32_t id;
        if (t < a) id = fn[node * S + t];
        else if (t < nnew) id = rsel[0][t - a];
        else if (t < nnew + c) id = fo[node * S + (t - nnew)];
        else id = rsel[1][t - nnew - c];
        sid[t] = id;
        sthr[t] = worst[id];
        if (kMode == kJoinFullI8) s_n8[t] = 0;
        else snorm[t] = xn[id];
    }
    const uint32_t RBk = (nn + 15) / 16, CB = (T + 15) / 16;
    if (threadIdx.x == 0) {
        uint32_t nb = 0;
        for (uint32_t rb = 0; rb < RBk; rb++)
            for (uint32_t cb = rb; cb < CB; cb++) {
                blk_r[nb] = (uint8_t)rb;
                blk_c[nb] = (uint8_t)cb;
                nb++;
            }
        s_nb = nb;
    }
    __syncthreads();
    const uint32_t NB = s_nb;
    const uint32_t Tpad = CB * 16;

    const uint32_t dP = d + 8;                       // kJoinFullBf16 row stride (elements)
    const uint32_t d64 = (d + 63) & ~63u, dP8 = d64 + 16;   // kJoinFullI8 row (bytes): padded dims, stride
    f32x4_n acc[kFull ? 1 : kJMaxB];
    if constexpr (kMode == kJoinFullI8) {
        // 16 bytes per step (d is a multiple of 32); dims d .. d64 are zero
        // (no term in any product)
        const uint32_t parts = d64 / 16;
        for (uint32_t u = threadIdx.x; u < Tpad * parts; u += blockDim.x) {
            const uint32_t t = u / parts, part = u % parts;
            i32x4_n v = {0, 0, 0, 0};
            if (t < T && part * 16 < d) {
                v = *reinterpret_cast<const i32x4_n *>(xu8 + (uint64_t)sid[t] * d + part * 16);
                v ^= 0x80808080;
                int sq = 0;
#pragma unroll
                for (int e = 0; e < 4; e++) sq = __builtin_amdgcn_sdot4(v[e], v[e], sq, false);
                atomicAdd(&s_n8[t], sq);
            }
            *reinterpret_cast<i32x4_n *>(&vec8[t * dP8 + part * 16]) = v;
        }
        __syncthreads();
    } else if constexpr (kFull) {
        const uint32_t parts = d / 8;
        for (uint32_t u = threadIdx.x; u < Tpad * parts; u += blockDim.x) {
            const uint32_t t = u / parts, part = u % parts;
            bf16x8_n v;
            if (t < T) {
                v = *reinterpret_cast<const bf16x8_n *>(xb + (uint64_t)sid[t] * d + part * 8);
            } else {
#pragma unroll
                for (int e = 0; e < 8; e++) v[e] = 0;
            }
            *reinterpret_cast<bf16x8_n *>(&vec[t * dP + part * 8]) = v;
        }
        __syncthreads();
    } else {
#pragma unroll
        for (int j = 0; j < kJMaxB; j++) acc[j] = f32x4_n{0.f, 0.f, 0.f, 0.f};

        for (uint32_t ch = 0; ch < d; ch += kJDC) {
            const uint32_t dims = (d - ch) < (uint32_t)kJDC ? (d - ch) : (uint32_t)kJDC;   // 32 or 64
            __syncthreads();
            const uint32_t parts = dims / 8;
            for (uint32_t u = threadIdx.x; u < Tpad * parts; u += blockDim.x) {
                uint32_t t = u / parts, part = u % parts;
                bf16x8_n v;
                if (t < T) {
                    v = *reinterpret_cast<const bf16x8_n *>(xb + (uint64_t)sid[t] * d + ch + part * 8);
                } else {
#pragma unroll
                    for (int e = 0; e < 8; e++) v[e] = 0;
                }
                *reinterpret_cast<bf16x8_n *>(&vec[t * kJDP + part * 8]) = v;
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < kJMaxB; j++) {
                uint32_t b = w + 4 * j;
                if (b < NB) {
                    uint32_t rb = blk_r[b], cb = blk_c[b];
                    for (uint32_t kl = 0; kl < dims / 32; kl++) {
                        bf16x8_n av = *reinterpret_cast<const bf16x8_n *>(
                            &vec[(rb * 16 + (lane & 15u)) * kJDP + kl * 32 + 8 * (lane >> 4)]);
                        bf16x8_n bv = *reinterpret_cast<const bf16x8_n *>(
                            &vec[(cb * 16 + (lane & 15u)) * kJDP + kl * 32 + 8 * (lane >> 4)]);
                        acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[j], 0, 0, 0);
                    }
                }
            }
        }
    }

    // Float data (not exact in BF16): the Gram value v_g of a pair decides
    // which pairs need their exact distance, with a margin that bounds its
    // error, |v_g - v| <= g_rel (|a|^2 + |b|^2): BF16 rounding of both rows
    // (2u + u^2 with u = 2^-8, times 2 |a.b| <= |a|^2 + |b|^2), FP32
    // accumulation over d products and of the norms (about d 2^-24 each), and
    // the expression's own roundings, with a quarter to spare.  The counting
    // pass reserves a slot for every pair whose v_g - margin is below a
    // threshold; the filling pass computes those pairs' distances exactly
    // (FP32 rows, the reference's order, one quad of lanes per pair) and
    // writes the key of each that passes, an empty key (~0, which the merge
    // skips) in the slot of each that does not.  No pair below its threshold
    // is missed, as in the exact modes.
    const bool approx = kMode != kJoinFullI8 && !exact;
    const float g_rel = 0.01f + 1.6e-7f * (float)d;

    // Proposals go to sample slots (target = sid[slot]).  In the MFMA output
    // layout the 16 lanes of a lane group share the row slot and the 4 lane
    // groups x 4 rows of a lane share the column slot, so proposals are summed
    // over those lanes first and each slot's LDS counter takes one add per
    // group; after the pass each slot takes one global atomic for its whole
    // count (a node's ~10^4 proposals reach at most 4S targets).  The filling
    // pass walks the pairs again and writes each proposal into the range its
    // slot reserved.  Both passes derive the same pa/pb from the same Gram
    // values and thresholds, so the ranges are filled exactly.
    uint32_t ndist = 0, nprop = 0, dropped = 0;
    const uint32_t grp = lane >> 4, gl = lane & 15u;
    // The 4 pairs a lane holds in Gram block (rb, cb) (squared distances dg
    // from the Gram values): pa/pb, a slot reserved at the row / column
    // target; ea/eb, a proposal made there (pa/pb themselves unless the data
    // are approximate); dv, the exact distance where ea or eb; then on_block.
    auto block_props = [&](uint32_t rb, uint32_t cb, const float *dg, auto &&on_block) {
        const uint32_t col = cb * 16 + gl;
        bool pa[4], pb[4], ea[4], eb[4];
        float dv[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const uint32_t r = rb * 16 + 4 * grp + i;
            pa[i] = pb[i] = false;
            dv[i] = 0.f;
            if (r < nn && col > r && col < T) {
                const uint32_t ia = sid[r], ib = sid[col];
                if (ia != ib) {
                    if (!kFill) ndist++;
                    const float v = dg[i];
                    const float lo = approx ? v - g_rel * (snorm[r] + snorm[col]) : v;
                    pa[i] = lo < sthr[r];
                    pb[i] = lo < sthr[col];
                    dv[i] = v;
                }
            }
            ea[i] = pa[i];
            eb[i] = pb[i];
        }
        if (kFill && approx) {
            // the exact distances of the reserved pairs: per pair index i, the
            // lanes' pairs in a per-wave LDS queue (the int8 norms' array,
            // unused here: 64 entries per wave), 16 pairs per round, one per
            // quad; each quad's first lane writes its distance over its entry
            uint32_t *qw = reinterpret_cast<uint32_t *>(s_n8) + 64u * w;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const bool nd = pa[i] || pb[i];
                const uint64_t m = __ballot(nd);
                const uint32_t cnt = (uint32_t)__popcll(m);   // uniform
                if (cnt == 0) continue;
                const uint32_t pos = (uint32_t)__popcll(m & lanemask_lt());
                const uint32_t r = rb * 16 + 4 * grp + i;
                if (nd) qw[pos] = (r << 8) | col;
                wave_lds_sync();
                for (uint32_t b0 = 0; b0 < cnt; b0 += 16) {
                    const uint32_t it = b0 + (lane >> 2);
                    const bool ok = it < cnt;
                    const uint32_t pr = ok ? qw[it] : 0u;
                    const float v = pair_quad_f32(x32, sid[pr >> 8], sid[pr & 255u], d, lane & 3u, ok);
                    if (ok && (lane & 3u) == 0) qw[it] = __float_as_uint(v);
                }
                wave_lds_sync();
                if (nd) {
                    const float v = __uint_as_float(qw[pos]);
                    dv[i] = v;
                    ea[i] = pa[i] && v < sthr[r];
                    eb[i] = pb[i] && v < sthr[col];
                }
                wave_lds_sync();
            }
        }
        on_block(rb * 16 + 4 * grp, col, pa, pb, ea, eb, dv);
    };
    auto each_block = [&](auto &&on_block) {
        // squared distances of a lane's 4 pairs from the Gram values g
        auto from_gram = [&](uint32_t rb, uint32_t cb, const f32x4_n &g, float *dg) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const float v = snorm[rb * 16 + 4 * grp + i] + snorm[cb * 16 + gl] - 2.f * g[i];
                dg[i] = v > 0.f ? v : 0.f;
            }
        };
        if constexpr (kMode == kJoinFullI8) {
            for (uint32_t b = w; b < NB; b += 4) {   // uniform over the wave
                const uint32_t rb = blk_r[b], cb = blk_c[b];
                i32x4_n g = {0, 0, 0, 0};
                const int8_t *ra = &vec8[(rb * 16 + gl) * dP8 + 16 * grp];
                const int8_t *rc = &vec8[(cb * 16 + gl) * dP8 + 16 * grp];
                for (uint32_t kl = 0; kl < d64; kl += 64) {
                    const i32x4_n av = *reinterpret_cast<const i32x4_n *>(ra + kl);
                    const i32x4_n bv = *reinterpret_cast<const i32x4_n *>(rc + kl);
                    g = __builtin_amdgcn_mfma_i32_16x16x64_i8(av, bv, g, 0, 0, 0);
                }
                float dg[4];
                const int nc = s_n8[cb * 16 + gl];
#pragma unroll
                for (int i = 0; i < 4; i++) dg[i] = (float)(s_n8[rb * 16 + 4 * grp + i] + nc - 2 * g[i]);
                block_props(rb, cb, dg, on_block);
            }
        } else if constexpr (kFull) {
            for (uint32_t b = w; b < NB; b += 4) {   // uniform over the wave
                const uint32_t rb = blk_r[b], cb = blk_c[b];
                f32x4_n g = {0.f, 0.f, 0.f, 0.f};
                const __hip_bfloat16 *ra = &vec[(rb * 16 + gl) * dP + 8 * grp];
                const __hip_bfloat16 *rc = &vec[(cb * 16 + gl) * dP + 8 * grp];
                for (uint32_t kl = 0; kl < d; kl += 32) {
                    const bf16x8_n av = *reinterpret_cast<const bf16x8_n *>(ra + kl);
                    const bf16x8_n bv = *reinterpret_cast<const bf16x8_n *>(rc + kl);
                    g = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, g, 0, 0, 0);
                }
                float dg[4];
                from_gram(rb, cb, g, dg);
                block_props(rb, cb, dg, on_block);
            }
        } else {
#pragma unroll
            for (int j = 0; j < kJMaxB; j++) {
                const uint32_t b = w + 4 * j;
                if (b < NB) {   // uniform over the wave
                    float dg[4];
                    from_gram(blk_r[b], blk_c[b], acc[j], dg);
                    block_props(blk_r[b], blk_c[b], dg, on_block);
                }
            }
        }
    };
    for (uint32_t t = threadIdx.x; t < T; t += blockDim.x) {
        s_pcnt[t] = 0;
        s_pcur[t] = 0;
    }
    __syncthreads();
    each_block([&](uint32_t r0, uint32_t col, const bool *pa, const bool *pb, const bool *ea, const bool *eb,
                   const float *) {
        uint32_t cb_cnt = 0;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const uint32_t gm = (uint32_t)(__ballot(pa[i]) >> (16 * grp)) & 0xFFFFu;
            if (gl == 0 && gm) atomicAdd(&s_pcnt[r0 + i], (uint32_t)__popc(gm));
            cb_cnt += pb[i] ? 1u : 0u;
            if (kFill) nprop += (ea[i] ? 1u : 0u) + (eb[i] ? 1u : 0u);
        }
        cb_cnt += __shfl_xor(cb_cnt, 16, 64);
        cb_cnt += __shfl_xor(cb_cnt, 32, 64);
        if (grp == 0 && cb_cnt) atomicAdd(&s_pcnt[col], cb_cnt);
    });
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < T; t += blockDim.x) {
        const uint32_t c = s_pcnt[t];
        if (c) {
            const uint32_t base = atomicAdd(&prop_cnt[sid[t]], c);   // count, or fill cursor
            if (kFill) s_pbase[t] = base;
        }
    }
    if (kFill) {
        __syncthreads();
        // a reserved slot: the (d2, id) sort key of a proposal, or the empty
        // key where the exact distance did not pass
        auto put = [&](uint32_t target, uint64_t at, uint32_t cand, float v, bool made) {
            const uint64_t pos = prop_off[target] + at;
            if (pos < cap) {
                prop_key[pos] = made ? ((unsigned long long)float_key(v) << 32) | cand : ~0ull;
            } else if (made) {
                dropped++;
            }
        };
        each_block([&](uint32_t r0, uint32_t col, const bool *pa, const bool *pb, const bool *ea, const bool *eb,
                       const float *dv) {
            const uint32_t ib = col < T ? sid[col] : 0u;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const uint32_t gm = (uint32_t)(__ballot(pa[i]) >> (16 * grp)) & 0xFFFFu;
                uint32_t base = 0;
                if (gl == 0 && gm) base = atomicAdd(&s_pcur[r0 + i], (uint32_t)__popc(gm));
                base = __shfl(base, (int)(16 * grp), 64);
                if (pa[i]) {
                    const uint32_t r = r0 + i;
                    put(sid[r], (uint64_t)s_pbase[r] + base + __popc(gm & ((1u << gl) - 1u)), ib, dv[i], ea[i]);
                }
            }
            uint32_t mine = 0;
#pragma unroll
            for (int i = 0; i < 4; i++) mine += pb[i] ? 1u : 0u;
            const uint32_t c0 = __shfl(mine, (int)gl, 64), c1 = __shfl(mine, (int)gl + 16, 64);
            const uint32_t c2 = __shfl(mine, (int)gl + 32, 64), c3 = __shfl(mine, (int)gl + 48, 64);
            const uint32_t total = c0 + c1 + c2 + c3;
            uint32_t base = 0;
            if (grp == 0 && total) base = atomicAdd(&s_pcur[col], total);
            base = __shfl(base, (int)gl, 64);
            uint32_t k = base + (grp > 0 ? c0 : 0u) + (grp > 1 ? c1 : 0u) + (grp > 2 ? c2 : 0u);
#pragma unroll
            for (int i = 0; i < 4; i++) {
                if (pb[i]) {
                    put(ib, (uint64_t)s_pbase[col] + k, sid[r0 + i], dv[i], eb[i]);
                    k++;
                }
            }
        });
    }
    // wave-level reduction of the counters (distances and proposals from the
    // counting pass, drops from the filling pass)
    for (int off = 32; off > 0; off >>= 1) {
        ndist += __shfl_down(ndist, off, 64);
        nprop += __shfl_down(nprop, off, 64);
        dropped += __shfl_down(dropped, off, 64);
    }
    if (lane == 0) {
        if (!kFill && ndist) atomicAdd(stat_at(stats, ST_DISTS), (unsigned long long)ndist);
        if (kFill && nprop) atomicAdd(stat_at(stats, ST_PROPOSALS), (unsigned long long)nprop);
        if (kFill && dropped) atomicAdd(stat_at(stats, ST_DROPPED), (unsigned long long)dropped);
    }
}

constexpr uint32_t kMergeChunk = 1024;   // proposals per merge round (1,024 and 512 measured alike, 2,048 17 % slower)
constexpr uint32_t kMergeThreads = 256;  // k_nnd_merge's workgroup size
constexpr uint32_t kSeenSlots = 2048;    // ids listed during one node's merge (LDS set)

__device__ __forceinline__ void seen_insert(uint32_t *set, uint32_t *n, uint32_t id) {
    if (atomicAdd(n, 1u) >= kSeenSlots * 3 / 4) return;   // full enough: stop growing (exactness kept)
    uint32_t h = (id * 2654435761u) & (kSeenSlots - 1);
    for (uint32_t probe = 0; probe < kSeenSlots; probe++) {
        const uint32_t old = atomicCAS(&set[h], kInvalidId, id);
        if (old == kInvalidId || old == id) return;
        h = (h + 1) & (kSeenSlots - 1);
    }
}
__device__ __forceinline__ bool seen_has(const uint32_t *set, uint32_t id) {
    uint32_t h = (id * 2654435761u) & (kSeenSlots - 1);
    for (uint32_t probe = 0; probe < kSeenSlots; probe++) {
        const uint32_t e = set[h];
        if (e == id) return true;
        if (e == kInvalidId) return false;
        h = (h + 1) & (kSeenSlots - 1);
    }
    return false;
}
constexpr uint32_t kMergeMaxK = 1024;

// Exclusive scan of flags[0..P) in place (P <= kMergeChunk, 256 threads);
// flags[P] receives the total.
__device__ void block_excl_scan(uint32_t *flags, uint32_t P, uint32_t *wsum) {
    const uint32_t per = (P + blockDim.x - 1) / blockDim.x;
    const uint32_t beg = threadIdx.x * per, end = min(beg + per, P);
    uint32_t local = 0;
    for (uint32_t t = beg; t < end; t++) local += flags[t];
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x / 64;
    uint32_t incl = local;
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t y = __shfl_up(incl, off, 64);
        if (lane >= (uint32_t)off) incl += y;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    uint32_t woff = 0;
    for (uint32_t k = 0; k < w; k++) woff += wsum[k];
    uint32_t run = woff + incl - local;
    __syncthreads();
    for (uint32_t t = beg; t < end; t++) {
        uint32_t v = flags[t];
        flags[t] = run;
        run += v;
    }
    if (threadIdx.x == blockDim.x - 1) flags[P] = run;
    __syncthreads();
}

// One workgroup (256 lanes) per node with proposals: the list lives in LDS
// while the node's proposals are merged into it 2048 at a time.
__global__ __launch_bounds__(kMergeThreads) void k_nnd_merge(
    uint32_t *__restrict__ ids, float *__restrict__ kd, uint8_t *__restrict__ isnew,
    uint32_t *__restrict__ cnt, uint32_t K, const uint32_t *__restrict__ prop_cnt,
    const uint64_t *__restrict__ prop_off, const unsigned long long *__restrict__ prop_key, uint64_t cap,
    unsigned long long *__restrict__ stats) {
    __shared__ unsigned long long keys[kMergeChunk];
    __shared__ uint32_t scan[kMergeChunk + 1];
    __shared__ uint32_t lid[kMergeMaxK], oid[kMergeMaxK];
    __shared__ float ld[kMergeMaxK], od[kMergeMaxK];
    __shared__ uint8_t lnw[kMergeMaxK], onw[kMergeMaxK];
    __shared__ uint32_t wsum[4];
    __shared__ uint32_t s_nk;
    // Ids listed at any point of this merge (insert-only, linear probing):
    // a proposal whose id is here is skipped before the sort.  That is exact:
    // a still-listed id is a duplicate, and a listed-then-evicted id has the
    // same distance as when it was evicted, never below the worst entry since.
    // Past 3/4 load the set stops growing and the check after the sort (the
    // run of equal list distances) alone keeps the merge exact.
    __shared__ uint32_t s_seen[kSeenSlots];
    __shared__ uint32_t s_seen_n;

    const uint64_t u = blockIdx.x;
    const uint64_t beg = prop_off[u];
    uint64_t np64 = prop_cnt[u];
    if (beg >= cap) np64 = 0;
    else if (beg + np64 > cap) np64 = cap - beg;   // proposals past the buffer were dropped
    const uint32_t np = (uint32_t)np64;
    if (np == 0) return;
    uint32_t c = cnt[u];
    for (uint32_t t = threadIdx.x; t < kSeenSlots; t += kMergeThreads) s_seen[t] = kInvalidId;
    if (threadIdx.x == 0) s_seen_n = 0;
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < c; t += kMergeThreads) {
        lid[t] = ids[u * K + t];
        ld[t] = kd[u * K + t];
        lnw[t] = isnew[u * K + t];
        seen_insert(s_seen, &s_seen_n, lid[t]);
    }
    uint32_t upd = 0;
    for (uint32_t base = 0; base < np; base += kMergeChunk) {
        const uint32_t nb = min(kMergeChunk, np - base);
        // Proposals that can still enter (below the current worst entry; the
        // worst falls as chunks merge) are compacted, then only they are
        // sorted: the first chunk keeps nearly all, later ones few.
        __syncthreads();   // the list (loaded, or rewritten by the last chunk) is complete
        const float worst_full = (c >= K) ? ld[K - 1] : FLT_MAX;
        if (threadIdx.x == 0) s_nk = 0;
        __syncthreads();
        // all of the thread's proposals are loaded at once (one round trip)
        constexpr uint32_t kPer = kMergeChunk / kMergeThreads;
        unsigned long long pk[kPer];
#pragma unroll
        for (uint32_t i = 0; i < kPer; i++) {   // unconditional (clamped) loads: no branch between them
            const uint32_t t = min(i * kMergeThreads + threadIdx.x, nb - 1);
            pk[i] = prop_key[beg + base + t];
        }
        const uint32_t worst_key = float_key(worst_full);   // order-preserving: d < worst <=> key < worst_key
#pragma unroll
        for (uint32_t i = 0; i < kPer; i++) {
            const bool in = i * kMergeThreads + threadIdx.x < nb && (uint32_t)(pk[i] >> 32) < worst_key &&
                            !seen_has(s_seen, (uint32_t)pk[i]);
            const unsigned long long key = in ? pk[i] : ~0ull;
            const uint64_t bm = __ballot(key != ~0ull);
            uint32_t at = 0;
            if ((threadIdx.x & 63u) == 0 && bm) at = atomicAdd(&s_nk, (uint32_t)__popcll(bm));
            at = __shfl(at, 0, 64);
            if (key != ~0ull) keys[at + (uint32_t)__popcll(bm & lanemask_lt())] = key;
        }
        __syncthreads();
        const uint32_t nk = s_nk;
        if (nk == 0) continue;   // uniform
        uint32_t PP = 256;   // the register sort's smallest size
        while (PP < nk) PP <<= 1;
        for (uint32_t t = nk + threadIdx.x; t < PP; t += kMergeThreads) keys[t] = ~0ull;
        __syncthreads();
        static_assert(kMergeChunk == 1024, "the sort dispatch below covers 256, 512 and 1024 keys");
        if (PP == 256) block_sort_u64_regs<1>(keys, PP);
        else if (PP == 512) block_sort_u64_regs<2>(keys, PP);
        else block_sort_u64_regs<4>(keys, PP);
        __syncthreads();
        // Each of the thread's keys (positions t = j * 256 + tid): lo = the
        // number of list entries with distance <= its own (its merge rank in
        // the list), by branchless binary searches run in lockstep.  A
        // proposal counts once: not a repeat of the previous key (the same
        // pair always has the same distance) and not already listed -- a
        // listed id has exactly this distance, so it sits in the run of
        // entries equal to it just below lo.
        constexpr uint32_t kPerT = kMergeChunk / kMergeThreads;
        const uint32_t np_t = PP / kMergeThreads;   // keys per thread, uniform
        uint32_t top = 1;                           // largest power of two <= c (1 when c <= 1)
        while (2 * top <= c) top <<= 1;
        unsigned long long kj[kPerT];
        uint32_t loj[kPerT];
#pragma unroll
        for (uint32_t j = 0; j < kPerT; j++) {
            kj[j] = j < np_t ? keys[j * kMergeThreads + threadIdx.x] : ~0ull;
            loj[j] = 0;
        }
        for (uint32_t step = top; step > 0; step >>= 1) {
#pragma unroll
            for (uint32_t j = 0; j < kPerT; j++) {
                const float dv = key_float((uint32_t)(kj[j] >> 32));
                if (kj[j] != ~0ull && loj[j] + step <= c && ld[loj[j] + step - 1] <= dv) loj[j] += step;
            }
        }
#pragma unroll
        for (uint32_t j = 0; j < kPerT; j++) {
            if (j >= np_t) break;
            const uint32_t t = j * kMergeThreads + threadIdx.x;
            const unsigned long long key = kj[j];
            uint32_t ok = 0;
            if (key != ~0ull && (t == 0 || keys[t - 1] != key)) {
                const float dv = key_float((uint32_t)(key >> 32));
                const uint32_t id = (uint32_t)key;
                ok = 1;
                for (int e = (int)loj[j] - 1; e >= 0 && ld[e] == dv; e--)
                    if (lid[e] == id) {
                        ok = 0;
                        break;
                    }
            }
            scan[t] = ok;
        }
        __syncthreads();
        block_excl_scan(scan, PP, wsum);
        const uint32_t nvalid = scan[PP];
        if (nvalid == 0) continue;   // uniform
        // proposals: pos = #valid before + #listed with dist <= d
#pragma unroll
        for (uint32_t j = 0; j < kPerT; j++) {
            if (j >= np_t) break;
            const uint32_t t = j * kMergeThreads + threadIdx.x;
            if (kj[j] == ~0ull || scan[t + 1] == scan[t]) continue;
            const uint32_t pos = scan[t] + loj[j];
            if (pos < K) {
                oid[pos] = (uint32_t)kj[j];
                od[pos] = key_float((uint32_t)(kj[j] >> 32));
                onw[pos] = 1;
                upd++;
                seen_insert(s_seen, &s_seen_n, (uint32_t)kj[j]);
            }
        }
        // listed entries: pos = e + #valid proposals with dist < ld[e]
        // (branchless searches over the sorted keys, in lockstep; padding
        // keys are ~0 and never counted)
        constexpr uint32_t kLPerT = kMergeMaxK / kMergeThreads;
        uint32_t kfe[kLPerT], loe[kLPerT];
#pragma unroll
        for (uint32_t j = 0; j < kLPerT; j++) {
            const uint32_t e = j * kMergeThreads + threadIdx.x;
            kfe[j] = e < c ? float_key(ld[e]) : 0u;
            loe[j] = 0;
        }
        for (uint32_t step = PP; step > 0; step >>= 1) {
#pragma unroll
            for (uint32_t j = 0; j < kLPerT; j++) {
                if (loe[j] + step <= PP && (uint32_t)(keys[loe[j] + step - 1] >> 32) < kfe[j]) loe[j] += step;
            }
        }
#pragma unroll
        for (uint32_t j = 0; j < kLPerT; j++) {
            const uint32_t e = j * kMergeThreads + threadIdx.x;
            if (e >= c) continue;
            const uint32_t pos = e + scan[loe[j]];
            if (pos < K) {
                oid[pos] = lid[e];
                od[pos] = ld[e];
                onw[pos] = lnw[e];
            }
        }
        __syncthreads();
        c = min(c + nvalid, K);
        for (uint32_t t = threadIdx.x; t < c; t += kMergeThreads) {
            lid[t] = oid[t];
            ld[t] = od[t];
            lnw[t] = onw[t];
        }
    }
    __syncthreads();
    for (uint32_t t = threadIdx.x; t < c; t += kMergeThreads) {
        ids[u * K + t] = lid[t];
        kd[u * K + t] = ld[t];
        isnew[u * K + t] = lnw[t];
    }
    for (int off = 32; off > 0; off >>= 1) upd += __shfl_down(upd, off, 64);
    if ((threadIdx.x & 63u) == 0 && upd) atomicAdd(stat_at(stats, ST_UPDATES), (unsigned long long)upd);
    if (threadIdx.x == 0) cnt[u] = c;
}

void nndescent_run(hrnn_dev_t h, int iters, int S, uint32_t per_node_cap, hrnn_nnd_iter_stats *st) {
    const uint64_t n = h->n;
    const uint32_t K = h->K;
    REQUIRE(S >= 1 && S <= 64, HRNN_E_ARG, "nndescent: sample_size must be in [1, 64]");
    REQUIRE(K <= kMergeMaxK, HRNN_E_ARG, "nndescent: K_knng must be <= 1024");
    REQUIRE(h->d % 32 == 0, HRNN_E_ARG, "nndescent: d must be a multiple of 32");
    const uint32_t R = (uint32_t)S;
    const uint32_t RB = kJRev;   // reverse offers kept per node and kind
    const int exact = (h->base_u8 && h->d <= 128) ? 1 : 0;
    // Whole-sample staging when 4S rows fit in 56 KiB of LDS: as int8 for
    // 8-bit data at d <= 258 (S = 50: d <= 256), else as BF16 (S = 50:
    // d <= 128); wider vectors go through 64-dim chunks.
    const uint64_t rows = (uint64_t)((4 * S + 15) / 16 * 16);
    const uint64_t i8_lds = rows * (((h->d + 63) & ~63u) + 16);
    const uint64_t bf16_lds = rows * (h->d + 8) * 2;
    int mode = kJoinChunked;
    if (h->base_u8 && h->d <= 258 && i8_lds <= 56 * 1024) mode = kJoinFullI8;
    else if (bf16_lds <= 56 * 1024) mode = kJoinFullBf16;
    if (h->tune_nnd_join) mode = h->tune_nnd_join - 1;   // the tuning key (same lists in every mode)
    if (const char *jm = std::getenv("HRNN_NND_JOIN")) mode = std::min(std::max(std::atoi(jm), 0), 2);
    if (mode == kJoinFullI8 && !(h->base_u8 && h->d <= 258 && i8_lds <= 64 * 1024)) mode = kJoinChunked;
    if (mode == kJoinFullBf16 && bf16_lds > 64 * 1024) mode = kJoinChunked;
    const uint32_t join_lds = mode == kJoinFullI8    ? (uint32_t)i8_lds
                              : mode == kJoinFullBf16 ? (uint32_t)bf16_lds
                                                      : (uint32_t)(kJMax * kJDP * 2);
    auto join_count = mode == kJoinFullI8    ? k_nnd_join<false, kJoinFullI8>
                      : mode == kJoinFullBf16 ? k_nnd_join<false, kJoinFullBf16>
                                              : k_nnd_join<false, kJoinChunked>;
    auto join_fill = mode == kJoinFullI8    ? k_nnd_join<true, kJoinFullI8>
                     : mode == kJoinFullBf16 ? k_nnd_join<true, kJoinFullBf16>
                                             : k_nnd_join<true, kJoinChunked>;

    // Proposal buffer: per_node_cap entries per node (default max(16384, 32K)),
    // at most 60% of the free device memory (8 B per proposal).
    uint64_t want = n * (uint64_t)(per_node_cap ? per_node_cap : std::max<uint32_t>(16384, 32 * K));
    size_t free_b = 0, total_b = 0;
    HIPC(hipMemGetInfo(&free_b, &total_b));
    uint64_t budget = std::max<uint64_t>(h->s3.bytes, free_b / 10 * 6) / 8;
    const uint64_t cap = std::max<uint64_t>(1, std::min(want, budget));

    // scratch layout (u32 words unless noted)
    DevBuf &w = h->s0;    // worst f32[n]
    DevBuf &fwd = h->s1;  // fn[n*S], fo[n*S]
    DevBuf &rev = h->s2;  // rn[n*RB], ro[n*RB]
    DevBuf &cnts = h->s4; // fn_cnt, fo_cnt, rn_cnt, ro_cnt, prop_cnt, prop_fill [n each], need bits
    DevBuf &prop = h->s3; // prop_key[cap]: (float_key(d2) << 32) | id
    const uint32_t nblk = scan_block_count(n);
    w.reserve(n * 4 + 64);
    fwd.reserve(n * S * 8 + 64);
    rev.reserve(n * RB * 8 + 64);
    cnts.reserve(n * 24 + (n + 31) / 32 * 4 + 64);
    prop.reserve(cap * 8 + 64);
    h->nnd_aux.reserve((n + 1 + nblk + 1) * 8 + 64);
    h->csr_valid = false;   // the graph changes: any CSR built before is stale
    float *worst = w.as<float>();
    uint32_t *fn = fwd.as<uint32_t>(), *fo = fn + n * S;
    uint32_t *rn = rev.as<uint32_t>(), *ro = rn + n * RB;
    uint32_t *fn_cnt = cnts.as<uint32_t>(), *fo_cnt = fn_cnt + n, *rn_cnt = fo_cnt + n,
             *ro_cnt = rn_cnt + n, *prop_cnt = ro_cnt + n, *prop_fill = prop_cnt + n,
             *need = prop_fill + n;   // need bits, one word per 32 nodes
    unsigned long long *prop_key = prop.as<unsigned long long>();
    uint64_t *prop_off = h->nnd_aux.as<uint64_t>();
    uint64_t *bsum = prop_off + n + 1;
    uint64_t *ptotal = bsum + nblk;

    uint32_t *ids = h->knn_ids.as<uint32_t>();
    float *kd = h->knn_d.as<float>();
    uint32_t *cnt = h->knn_cnt.as<uint32_t>();
    uint8_t *isnew = h->knn_new.as<uint8_t>();
    // per-iteration counters: ST_N x kStatSlots u64 per iteration (zeroed by k_nnd_prepare)
    constexpr uint64_t kStatWords = (uint64_t)ST_N * kStatSlots;
    h->s5.reserve((uint64_t)std::max(iters, 1) * kStatWords * 8 + 64);
    unsigned long long *all_stats = h->s5.as<unsigned long long>();
    uint32_t *active = h->counters.as<uint32_t>() + kNndActiveWord;

    auto enqueue_iteration = [&](int it) {
        unsigned long long *stats = all_stats + (uint64_t)it * kStatWords;
        hipLaunchKernelGGL(k_nnd_prepare, dim3(ceil_div(n, 256)), dim3(256), 0, h->stream, kd, cnt, n,
                           K, worst, rn_cnt, ro_cnt, prop_cnt, prop_fill, stats);
        hipLaunchKernelGGL(k_nnd_active, dim3(1), dim3(kStatSlots), 0, h->stream,
                           it > 0 ? all_stats + (uint64_t)(it - 1) * kStatWords : nullptr, active);
        hipLaunchKernelGGL(k_nnd_sample<false>, dim3(ceil_div(n, 4)), dim3(256), 0, h->stream, ids, kd,
                           isnew, cnt, n, K, (uint32_t)S, RB, worst, fn, fn_cnt, rn, rn_cnt, need, active);
        hipLaunchKernelGGL(k_nnd_need, dim3(ceil_div(n, 256)), dim3(256), 0, h->stream, fn_cnt, rn_cnt, n,
                           need);
        hipLaunchKernelGGL(k_nnd_sample<true>, dim3(ceil_div(n, 4)), dim3(256), 0, h->stream, ids, kd,
                           isnew, cnt, n, K, (uint32_t)S, RB, worst, fo, fo_cnt, ro, ro_cnt, need, active);
        hipLaunchKernelGGL(join_count, dim3((unsigned)n), dim3(256), join_lds, h->stream,
                           h->xb.as<__hip_bfloat16>(), h->xu8.as<uint8_t>(), h->norms.as<float>(),
                           h->x.as<float>(), h->d,
                           n, (uint32_t)S, R, RB, fn, fn_cnt, fo, fo_cnt, rn, rn_cnt, ro, ro_cnt, worst,
                           prop_cnt, prop_off, prop_key, cap, exact, stats);
        scan_u32(h, prop_cnt, n, prop_off, bsum, ptotal);
        hipLaunchKernelGGL(join_fill, dim3((unsigned)n), dim3(256), join_lds, h->stream,
                           h->xb.as<__hip_bfloat16>(), h->xu8.as<uint8_t>(), h->norms.as<float>(),
                           h->x.as<float>(), h->d,
                           n, (uint32_t)S, R, RB, fn, fn_cnt, fo, fo_cnt, rn, rn_cnt, ro, ro_cnt, worst,
                           prop_fill, prop_off, prop_key, cap, exact, stats);
        hipLaunchKernelGGL(k_nnd_merge, dim3((unsigned)n), dim3(kMergeThreads), 0, h->stream, ids, kd, isnew,
                           cnt, K, prop_cnt, prop_off, prop_key, cap, stats);
    };

    const char *env = std::getenv("HRNN_NND_GRAPH");
    const bool use_graph = !(env && env[0] == '0') && iters > 0;
    std::vector<unsigned long long> sv((uint64_t)std::max(iters, 1) * kStatWords, 0);
    if (use_graph) {
        // Each iteration captured as a HIP graph (10 kernels, no host
        // synchronisation; counters are reset by kernels only), all of them
        // instantiated before the first is launched, then launched back to
        // back with an event between consecutive iterations: each
        // iteration's device time is measured, not apportioned.
        std::vector<hipGraph_t> graphs(iters, nullptr);
        std::vector<hipGraphExec_t> execs(iters, nullptr);
        std::vector<hipEvent_t> evs(iters + 1, nullptr);
        auto cleanup = [&] {
            for (auto &e : execs)
                if (e) (void)hipGraphExecDestroy(e);
            for (auto &g : graphs)
                if (g) (void)hipGraphDestroy(g);
            for (auto &e : evs)
                if (e) (void)hipEventDestroy(e);
        };
        try {
            for (auto &e : evs) HIPC(hipEventCreate(&e));
            for (int it = 0; it < iters; it++) {
                HIPC(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
                enqueue_iteration(it);
                hipError_t ce = hipStreamEndCapture(h->stream, &graphs[it]);
                if (ce != hipSuccess)
                    throw Error{HRNN_E_HIP, std::string("nndescent capture: ") + hipGetErrorString(ce)};
                hipError_t ie = hipGraphInstantiate(&execs[it], graphs[it], nullptr, nullptr, 0);
                if (ie != hipSuccess)
                    throw Error{HRNN_E_HIP, std::string("nndescent graph instantiate: ") + hipGetErrorString(ie)};
            }
            HIPC(hipEventRecord(evs[0], h->stream));
            for (int it = 0; it < iters; it++) {
                HIPC(hipGraphLaunch(execs[it], h->stream));
                HIPC(hipEventRecord(evs[it + 1], h->stream));
            }
            d2h(h, sv.data(), all_stats, sv.size() * 8);
            if (st)
                for (int it = 0; it < iters; it++) {
                    float ms = 0.f;
                    HIPC(hipEventElapsedTime(&ms, evs[it], evs[it + 1]));
                    st[it].ms = ms;
                }
        } catch (...) {
            cleanup();
            throw;
        }
        cleanup();
    } else {
        for (int it = 0; it < iters; it++) {
            auto t0 = std::chrono::steady_clock::now();
            enqueue_iteration(it);
            HIPC(hipGetLastError());
            d2h(h, sv.data() + (uint64_t)it * kStatWords, all_stats + (uint64_t)it * kStatWords, kStatWords * 8);
            if (st) st[it].ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        }
    }
    if (st)
        for (int it = 0; it < iters; it++) {
            unsigned long long s[ST_N] = {0, 0, 0, 0, 0};
            for (int c = 0; c < ST_N; c++)
                for (uint32_t j = 0; j < kStatSlots; j++) s[c] += sv[(uint64_t)it * kStatWords + c * kStatSlots + j];
            st[it].dists = s[ST_DISTS];
            st[it].proposals = s[ST_PROPOSALS];
            st[it].dropped = s[ST_DROPPED];
            st[it].updates = s[ST_UPDATES];
            st[it].rows = s[ST_ROWS];
        }
}

void nnd_merge_launch(hrnn_dev_t h, uint64_t n, const uint32_t *prop_cnt, const uint64_t *prop_off,
                      const unsigned long long *prop_key, uint64_t cap, unsigned long long *stats) {
    REQUIRE(h->K >= 1 && h->K <= kMergeMaxK, HRNN_E_ARG, "merge: K_knng must be in [1, 1024]");
    if (!n) return;
    hipLaunchKernelGGL(k_nnd_merge, dim3((unsigned)n), dim3(kMergeThreads), 0, h->stream, h->knn_ids.as<uint32_t>(),
                       h->knn_d.as<float>(), h->knn_new.as<uint8_t>(), h->knn_cnt.as<uint32_t>(), h->K, prop_cnt,
                       prop_off, prop_key, cap, stats);
    HIPC(hipGetLastError());
}

uint64_t nnd_stat_words() { return (uint64_t)ST_N * kStatSlots; }

uint64_t nnd_stat_updates(const unsigned long long *blk) {
    uint64_t s = 0;
    for (uint32_t j = 0; j < kStatSlots; j++) s += blk[(uint64_t)ST_UPDATES * kStatSlots + j];
    return s;
}

}  // namespace hrnn
