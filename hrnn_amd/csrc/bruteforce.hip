// bruteforce.hip -- (d) exact brute-force kNN / r_k / RkNN on the whole base.
//
// Reference semantics: compute_kth_nn_dists and compute_rknn
// (scripts/gen_rknn_gt.py:42-151): d2 = |x|^2 + |y|^2 - 2 x.y clamped at 0,
// topk(k+1) over all points (self included), RkNN(q) = {x : d2(q,x) <= r_k(x)};
// kNN ties go to the smaller id (BruteforceSearch::searchKnn,
// include/hnswlib/bruteforce.h:106-135).
//
// Candidate pass (the only dense contraction of the path):
//   BF16/FP16: 64 query rows x 64 base columns per tile on MFMA
//     (v_mfma_f32_16x16x32_{bf16,f16}, FP32 accumulate); each wave owns 16
//     rows and keeps its A fragments in registers for the whole sweep; base
//     tiles are staged through LDS in 64-dim chunks.
//   FP32: the same tiling on VALU with direct differences.
// Every distance is compared with its row's running C-th best; survivors are
// staged in LDS and merged into per-row sorted (dist, id) lists of length C
// (C >= 2 * kk).  A re-rank pass then recomputes the C candidates with the
// reference's FP32 direct-difference order and sorts by (dist, id).
// For integer data in [0, 255] at d <= 128 (the integer configs) BF16/FP16
// inputs with FP32 accumulation are exact (every norm, product and partial sum
// an integer below 2^24), so candidate distances equal the reference's bit for
// bit and the RkNN pass (mode 1) can decide on them directly.
#include <algorithm>
#include <cfloat>
#include <vector>

#include "handle.h"

namespace hrnn {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

constexpr int kRB = 64;        // rows per block
constexpr int kCT = 64;        // columns per tile
constexpr int kCH = 64;        // dims per LDS chunk (2 MFMA k-steps)
constexpr int kCHP = kCH + 8;  // padded LDS row stride (bf16 elements)
constexpr int kHalf = 32;      // columns staged per half-tile

// Per-thread squared L2 in the reference's 16-partial FMA order
// (space_l2.h:25-54); sequential order when d % 16 != 0.
__device__ float l2_ref16_thread(const float *__restrict__ a, const float *__restrict__ b, uint32_t d) {
    if (d & 15u) return l2_seq(a, b, d);
    float p[16];
#pragma unroll
    for (int l = 0; l < 16; l++) p[l] = 0.f;
    for (uint32_t e = 0; e < d; e += 16) {
#pragma unroll
        for (int l = 0; l < 16; l++) {
            float t = a[e + l] - b[e + l];
            p[l] = __builtin_fmaf(t, t, p[l]);
        }
    }
    float r = p[0];
#pragma unroll
    for (int l = 1; l < 16; l++) r += p[l];
    return r;
}

template <int C, int R = kRB>
struct TopState {
    float top_d[R][C];
    uint32_t top_i[R][C];
    uint32_t top_n[R];
    float thr[R];
    float st_d[R][kHalf];
    uint32_t st_i[R][kHalf];
    uint32_t st_n[R];
};

template <int C, int R>
__device__ void top_init(TopState<C, R> &s) {
    for (int r = threadIdx.x; r < R; r += blockDim.x) {
        s.top_n[r] = 0;
        s.st_n[r] = 0;
        s.thr[r] = FLT_MAX;
    }
}

// rows with staged candidates merge them into their sorted top list
template <int C, int R>
__device__ void top_merge(TopState<C, R> &s) {
    __syncthreads();
    for (int r = threadIdx.x; r < R; r += blockDim.x) {
        uint32_t ns = s.st_n[r];
        if (!ns) continue;
        uint32_t n = s.top_n[r];
        for (uint32_t t = 0; t < ns; t++) {
            float dv = s.st_d[r][t];
            uint32_t iv = s.st_i[r][t];
            if (n == C && !pair_less(dv, iv, s.top_d[r][C - 1], s.top_i[r][C - 1])) continue;
            uint32_t p = (n == C) ? C - 1 : n;
            if (n < C) n++;
            while (p > 0 && pair_less(dv, iv, s.top_d[r][p - 1], s.top_i[r][p - 1])) {
                s.top_d[r][p] = s.top_d[r][p - 1];
                s.top_i[r][p] = s.top_i[r][p - 1];
                p--;
            }
            s.top_d[r][p] = dv;
            s.top_i[r][p] = iv;
        }
        s.top_n[r] = n;
        s.st_n[r] = 0;
        s.thr[r] = (n == C) ? s.top_d[r][C - 1] : FLT_MAX;
    }
    __syncthreads();
}

// The same merge for the 16 rows of one wave (rows 16 w .. 16 w + 15, one
// lane each), without block barriers: only this wave stages candidates for
// these rows, and one wave's LDS accesses are performed in program order.
template <int C, int R>
__device__ void top_merge_wave(TopState<C, R> &s, uint32_t w, uint32_t lane) {
    if (lane >= 16) return;
    const int r = (int)(16 * w + lane);
    const uint32_t ns = s.st_n[r];
    if (!ns) return;
    uint32_t n = s.top_n[r];
    for (uint32_t t = 0; t < ns; t++) {
        float dv = s.st_d[r][t];
        uint32_t iv = s.st_i[r][t];
        if (n == C && !pair_less(dv, iv, s.top_d[r][C - 1], s.top_i[r][C - 1])) continue;
        uint32_t p = (n == C) ? C - 1 : n;
        if (n < C) n++;
        while (p > 0 && pair_less(dv, iv, s.top_d[r][p - 1], s.top_i[r][p - 1])) {
            s.top_d[r][p] = s.top_d[r][p - 1];
            s.top_i[r][p] = s.top_i[r][p - 1];
            p--;
        }
        s.top_d[r][p] = dv;
        s.top_i[r][p] = iv;
    }
    s.top_n[r] = n;
    s.st_n[r] = 0;
    s.thr[r] = (n == C) ? s.top_d[r][C - 1] : FLT_MAX;
}

template <int C, int R>
__device__ __forceinline__ void top_offer(TopState<C, R> &s, int r, float dv, uint32_t col) {
    // columns arrive in increasing id, so an equal distance never displaces a kept id
    if (dv < s.thr[r]) {
        uint32_t t = atomicAdd(&s.st_n[r], 1u);
        s.st_d[r][t] = dv;
        s.st_i[r][t] = col;
    }
}

template <int C, int R>
__device__ void top_store(TopState<C, R> &s, uint64_t row0, uint32_t nq, uint32_t *cand_i, uint32_t *cand_n) {
    for (int t = threadIdx.x; t < R * C; t += blockDim.x) {
        int r = t / C, j = t % C;
        uint64_t row = row0 + r;
        if (row >= nq) continue;
        cand_i[row * C + j] = (uint32_t)j < s.top_n[r] ? s.top_i[r][j] : kInvalidId;
        if (j == 0) cand_n[row] = s.top_n[r];
    }
}

template <class T>
struct Mfma;
template <>
struct Mfma<__hip_bfloat16> {
    typedef bf16x8_t vec;
    __device__ static f32x4_t run(vec a, vec b, f32x4_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    }
};
template <>
struct Mfma<__half> {
    typedef f16x8_t vec;
    __device__ static f32x4_t run(vec a, vec b, f32x4_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
    }
};

// MFMA candidate pass.  KS = d / 32 k-steps (A fragments live in registers).
// mode 0: per-row top-C;  mode 1: RkNN hits d2 <= kth[col] (exact data only).
// FR 16-row fragments per wave: a block covers 64 FR rows, so each staged
// base tile feeds FR times as many MFMAs.
template <class T, int KS, int C, int MODE, int FR>
__global__ __launch_bounds__(256) void k_bf_mfma(
    const T *__restrict__ qa, const float *__restrict__ qn, uint32_t nq,
    const T *__restrict__ xb, const float *__restrict__ xn, uint64_t n, uint32_t d,
    uint32_t *__restrict__ cand_i, uint32_t *__restrict__ cand_n,
    const float *__restrict__ kth, uint32_t *__restrict__ hit_cnt, uint32_t *__restrict__ hit_i,
    float *__restrict__ hit_d, uint32_t cap) {
    typedef typename Mfma<T>::vec vec;
    constexpr int RB = 64 * FR;
    __shared__ TopState<C, RB> st;
    // d <= 128 (KS <= 4): the whole base tile (64 columns x d) in LDS, one
    // barrier pair per tile; wider rows: one 64-dim chunk at a time
    constexpr bool kOneStage = KS <= 4;
    constexpr int kBSP = kOneStage ? KS * 32 + 8 : kCHP;   // bs row stride (elements)
    __shared__ __attribute__((aligned(16))) T bs[kCT * kBSP];
    __shared__ float s_xn[kCT];
    __shared__ float s_qn[RB];

    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x / 64;
    const uint64_t row0 = (uint64_t)blockIdx.x * RB;
    if (MODE == 0) top_init(st);
    for (int r = threadIdx.x; r < RB; r += blockDim.x) s_qn[r] = (row0 + r < nq) ? qn[row0 + r] : 0.f;

    // A fragments: fragment f of wave w holds row row0 + 16 (FR w + f) +
    // lane%16, dims ks*32 + 8*(lane/16) .. +8
    vec afrag[FR][KS];
#pragma unroll
    for (int f = 0; f < FR; f++) {
        uint64_t row = row0 + 16 * (FR * w + f) + (lane & 15u);
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
            if (row < nq) {
                afrag[f][ks] = *reinterpret_cast<const vec *>(qa + row * d + ks * 32 + 8 * (lane >> 4));
            } else {
#pragma unroll
                for (int e = 0; e < 8; e++) afrag[f][ks][e] = 0;
            }
        }
    }
    __syncthreads();

    float qn_r[FR][4], thr_r[FR][4];   // this lane's rows: squared norms, running C-th best
#pragma unroll
    for (int f = 0; f < FR; f++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            qn_r[f][i] = s_qn[16 * (FR * w + f) + 4 * (lane >> 4) + i];
            thr_r[f][i] = FLT_MAX;
        }
    // Base tiles are prefetched into registers one tile ahead (d <= 256: the
    // whole next tile, 64 columns x d dims, KS / 2 chunks of two 16-byte
    // loads per thread) or one chunk ahead (wider rows), so the global-load
    // latency overlaps the MFMAs and the epilogue instead of stalling every
    // chunk.  (A barrier does not wait for outstanding global loads.)
    constexpr int NCH = (KS + 1) / 2;                 // 64-dim chunks per tile
    constexpr bool kWholeTile = KS <= 8;
    constexpr int NPV = kWholeTile ? 2 * NCH : 2;     // 16-byte vectors held per thread
    vec pv[NPV];
    float pxn = 0.f;
    // global -> registers for chunk ch of the tile at column c; slot base b
    auto fetch = [&](uint64_t c, int ch, int b) {
        const int dims = (KS - 2 * ch >= 2) ? 64 : 32;
#pragma unroll
        for (int v2 = 0; v2 < 2; v2++) {
            const int u = threadIdx.x + v2 * 256;
            const int col = u / (dims / 8), part = u % (dims / 8);
            const uint64_t gc = c + col;
            if (u < kCT * (dims / 8) && gc < n) {
                pv[b + v2] = *reinterpret_cast<const vec *>(xb + gc * d + ch * 64 + part * 8);
            } else {
#pragma unroll
                for (int e = 0; e < 8; e++) pv[b + v2][e] = 0;
            }
        }
    };
    auto put = [&](int ch, int b) {   // registers -> LDS chunk buffer
        const int dims = (KS - 2 * ch >= 2) ? 64 : 32;
#pragma unroll
        for (int v2 = 0; v2 < 2; v2++) {
            const int u = threadIdx.x + v2 * 256;
            if (u < kCT * (dims / 8)) {
                const int col = u / (dims / 8), part = u % (dims / 8);
                *reinterpret_cast<vec *>(&bs[col * kBSP + (kOneStage ? ch * 64 : 0) + part * 8]) = pv[b + v2];
            }
        }
    };
    if (kWholeTile) {
#pragma unroll
        for (int ch = 0; ch < NCH; ch++) fetch(0, ch, 2 * ch);
    } else {
        fetch(0, 0, 0);
    }
    if (threadIdx.x < kCT) pxn = threadIdx.x < n ? xn[threadIdx.x] : 0.f;

    for (uint64_t c0 = 0; c0 < n; c0 += kCT) {
        f32x4_t acc[FR][4];
#pragma unroll
        for (int f = 0; f < FR; f++)
#pragma unroll
            for (int cb = 0; cb < 4; cb++) acc[f][cb] = f32x4_t{0.f, 0.f, 0.f, 0.f};
        const uint64_t c1 = c0 + kCT;   // next tile
#pragma unroll
        for (int ch = 0; ch < NCH; ch++) {
            if (!kOneStage || ch == 0)
                __syncthreads();   // the previous chunk's MFMAs (and epilogue) are done with bs and s_xn
            if (ch == 0 && threadIdx.x < kCT) s_xn[threadIdx.x] = pxn;
            if (kWholeTile) {
                put(ch, 2 * ch);
                if (c1 < n) fetch(c1, ch, 2 * ch);   // this chunk of the next tile
                if (ch == 0 && threadIdx.x < kCT && c1 < n)
                    pxn = (c1 + threadIdx.x < n) ? xn[c1 + threadIdx.x] : 0.f;
            } else {
                put(ch, 0);
                if (ch + 1 < NCH) fetch(c0, ch + 1, 0);
                else if (c1 < n) fetch(c1, 0, 0);
                if (ch == 0 && threadIdx.x < kCT && c1 < n)
                    pxn = (c1 + threadIdx.x < n) ? xn[c1 + threadIdx.x] : 0.f;
            }
            if (kOneStage && ch + 1 < NCH) continue;   // stage every chunk, then one barrier
            __syncthreads();
#pragma unroll
            for (int kc = 0; kc < (kOneStage ? NCH : 1); kc++) {
            const int cc = kOneStage ? kc : ch;   // chunk the MFMAs read
#pragma unroll
            for (int kl = 0; kl < 2; kl++) {
                if (2 * cc + kl < KS) {
#pragma unroll
                    for (int cb = 0; cb < 4; cb++) {
                        vec b = *reinterpret_cast<const vec *>(
                            &bs[(16 * cb + (lane & 15u)) * kBSP + (kOneStage ? cc * 64 : 0) + kl * 32 + 8 * (lane >> 4)]);
#pragma unroll
                        for (int f = 0; f < FR; f++) acc[f][cb] = Mfma<T>::run(afrag[f][2 * cc + kl], b, acc[f][cb]);
                    }
                }
            }
            }
        }
        // epilogue: lane holds rows 16 (FR w + f) + 4*(lane/16) + i, column
        // 16cb + lane%16; row norms and thresholds come from registers, the
        // four column norms (and r_k) are read once per tile
        float xn_c[4], kth_c[4];
#pragma unroll
        for (int cb = 0; cb < 4; cb++) {
            const uint64_t col = c0 + 16 * cb + (lane & 15u);
            xn_c[cb] = s_xn[16 * cb + (lane & 15u)];
            kth_c[cb] = (MODE == 1 && col < n) ? kth[col] : 0.f;
        }
        for (int half = 0; half < 2; half++) {
#pragma unroll
            for (int cbh = 0; cbh < 2; cbh++) {
                const int cb = half * 2 + cbh;
                uint32_t cl = 16 * cb + (lane & 15u);
                uint64_t col = c0 + cl;
#pragma unroll
                for (int f = 0; f < FR; f++)
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    int rl = 16 * (FR * w + f) + 4 * (lane >> 4) + i;
                    if (col < n && row0 + rl < nq) {
                        float dv = qn_r[f][i] + xn_c[cb] - 2.f * acc[f][cb][i];
                        dv = dv > 0.f ? dv : 0.f;
                        if (MODE == 0) {
                            // columns arrive in increasing id, so an equal
                            // distance never displaces a kept id
                            if (dv < thr_r[f][i]) {
                                const uint32_t t = atomicAdd(&st.st_n[rl], 1u);
                                st.st_d[rl][t] = dv;
                                st.st_i[rl][t] = (uint32_t)col;
                            }
                        } else if (dv <= kth_c[cb]) {
                            uint64_t q = row0 + rl;
                            uint32_t s = atomicAdd(&hit_cnt[q], 1u);
                            if (s < cap) {
                                hit_i[q * cap + s] = (uint32_t)col;
                                hit_d[q * cap + s] = dv;
                            }
                        }
                    }
                }
            }
            if (MODE == 0) {
                // each wave merges its own rows (FR fragments of 16): no block barrier
#pragma unroll
                for (int f = 0; f < FR; f++) top_merge_wave(st, FR * w + f, lane);
#pragma unroll
                for (int f = 0; f < FR; f++)
#pragma unroll
                    for (int i = 0; i < 4; i++) thr_r[f][i] = st.thr[16 * (FR * w + f) + 4 * (lane >> 4) + i];
            }
        }
    }
    if (MODE == 0) {
        __syncthreads();   // every wave's rows are merged
        top_store(st, row0, nq, cand_i, cand_n);
    }
}

// FP32 VALU candidate pass: 4x4 micro-tile per lane, direct differences.
// mode 0: top-C; mode 1: RkNN hits, boundary cases re-decided with the
// reference order.
constexpr int kFCH = 16;   // dims per chunk
template <int C, int MODE>
__global__ __launch_bounds__(256) void k_bf_f32(
    const float *__restrict__ qa, uint32_t nq, const float *__restrict__ x, uint64_t n, uint32_t d,
    uint32_t *__restrict__ cand_i, uint32_t *__restrict__ cand_n, const float *__restrict__ kth,
    uint32_t *__restrict__ hit_cnt, uint32_t *__restrict__ hit_i, float *__restrict__ hit_d,
    uint32_t cap) {
    __shared__ TopState<C> st;
    __shared__ float as_[kFCH][kRB + 4];
    __shared__ float bs_[kFCH][kCT + 4];
    const uint64_t row0 = (uint64_t)blockIdx.x * kRB;
    if (MODE == 0) top_init(st);
    const int tr = threadIdx.x / 16, tc = threadIdx.x % 16;   // rows tr*4.., cols tc*4..
    for (uint64_t c0 = 0; c0 < n; c0 += kCT) {
        float acc[4][4];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) acc[i][j] = 0.f;
        for (uint32_t k0 = 0; k0 < d; k0 += kFCH) {
            __syncthreads();
            for (int u = threadIdx.x; u < kRB * kFCH; u += blockDim.x) {
                int r = u / kFCH, e = u % kFCH;
                uint64_t row = row0 + r;
                as_[e][r] = (row < nq && k0 + e < d) ? qa[row * d + k0 + e] : 0.f;
                uint64_t col = c0 + r;
                bs_[e][r] = (col < n && k0 + e < d) ? x[col * d + k0 + e] : 0.f;
            }
            __syncthreads();
            uint32_t kmax = (d - k0) < (uint32_t)kFCH ? (d - k0) : (uint32_t)kFCH;
            for (uint32_t e = 0; e < kmax; e++) {
                float av[4], bv[4];
#pragma unroll
                for (int i = 0; i < 4; i++) av[i] = as_[e][tr * 4 + i];
#pragma unroll
                for (int j = 0; j < 4; j++) bv[j] = bs_[e][tc * 4 + j];
#pragma unroll
                for (int i = 0; i < 4; i++)
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        float t = av[i] - bv[j];
                        acc[i][j] = __builtin_fmaf(t, t, acc[i][j]);
                    }
            }
        }
        for (int half = 0; half < 2; half++) {
            if ((tc * 4) / kHalf == half) {
#pragma unroll
                for (int i = 0; i < 4; i++)
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        int rl = tr * 4 + i;
                        uint64_t col = c0 + tc * 4 + j;
                        uint64_t row = row0 + rl;
                        if (col >= n || row >= nq) continue;
                        float dv = acc[i][j];
                        if (MODE == 0) {
                            top_offer(st, rl, dv, (uint32_t)col);
                        } else {
                            float kv = kth[col];
                            bool near = fabsf(dv - kv) <= 1e-5f * fmaxf(kv, 1.f);
                            if (near) dv = l2_ref16_thread(qa + row * d, x + col * d, d);
                            if (dv <= kv) {
                                if (!near) dv = l2_ref16_thread(qa + row * d, x + col * d, d);
                                uint32_t s = atomicAdd(&hit_cnt[row], 1u);
                                if (s < cap) {
                                    hit_i[row * cap + s] = (uint32_t)col;
                                    hit_d[row * cap + s] = dv;
                                }
                            }
                        }
                    }
            }
            if (MODE == 0) top_merge(st);
        }
    }
    if (MODE == 0) top_store(st, row0, nq, cand_i, cand_n);
}

// Re-rank: one wave per row recomputes its C candidates in the reference's
// FP32 order, sorts (dist, id) and writes the first kk.
template <int C>
__global__ __launch_bounds__(256) void k_rerank(const float *__restrict__ qa, uint32_t nq,
                                                const float *__restrict__ x, uint32_t d,
                                                const uint32_t *__restrict__ cand_i,
                                                uint32_t kk, uint32_t *__restrict__ ids,
                                                float *__restrict__ dist) {
    __shared__ float sd[4][C];
    __shared__ uint32_t si[4][C];
    const uint32_t w = threadIdx.x / 64, lane = threadIdx.x & 63u;
    uint64_t row = (uint64_t)blockIdx.x * 4 + w;
    const bool valid = row < nq;
    const float *qv = qa + (valid ? row : 0) * d;
    for (int c = lane / 16; valid && c < C; c += 4) {
        uint32_t id = cand_i[row * C + c];
        float dv = FLT_MAX;
        if (id != kInvalidId) {   // uniform across the 16-lane group
            if ((d & 15u) == 0) dv = l2_ref16_group(qv, x + (uint64_t)id * d, d, lane & 15u);
            else dv = l2_seq(qv, x + (uint64_t)id * d, d);
        }
        if ((lane & 15u) == 0) {
            sd[w][c] = dv;
            si[w][c] = id;
        }
    }
    __syncthreads();
    if (valid && lane == 0) {
        // insertion sort of C pairs (C <= 64)
        for (int a = 1; a < C; a++) {
            float dv = sd[w][a];
            uint32_t iv = si[w][a];
            int p = a;
            while (p > 0 && pair_less(dv, iv, sd[w][p - 1], si[w][p - 1])) {
                sd[w][p] = sd[w][p - 1];
                si[w][p] = si[w][p - 1];
                p--;
            }
            sd[w][p] = dv;
            si[w][p] = iv;
        }
        for (uint32_t j = 0; j < kk; j++) {
            ids[row * kk + j] = si[w][j];
            dist[row * kk + j] = sd[w][j];
        }
    }
}

// FP32 -> BF16/FP16 copies and norms of external queries.
__global__ void k_query_prep(const float *__restrict__ q, uint64_t nq, uint32_t d,
                             __hip_bfloat16 *__restrict__ qb, __half *__restrict__ qh,
                             float *__restrict__ qn) {
    uint64_t r = (uint64_t)blockIdx.x * 4 + threadIdx.x / 64;
    if (r >= nq) return;
    uint32_t lane = threadIdx.x & 63u;
    float acc = 0.f;
    for (uint32_t e = lane; e < d; e += 64) {
        float v = q[r * d + e];
        qb[r * d + e] = __float2bfloat16(v);
        qh[r * d + e] = __float2half(v);
        acc = __builtin_fmaf(v, v, acc);
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
    if (lane == 0) qn[r] = acc;
}

// ---- launchers -------------------------------------------------------------

template <class T, int C, int MODE>
static void launch_mfma(hrnn_dev_t h, const T *qa, const float *qn, uint32_t nq, const T *xb,
                        uint32_t *ci, uint32_t *cn, const float *kth, uint32_t *hc, uint32_t *hi,
                        float *hd, uint32_t cap) {
    // One 16-row fragment per wave (64 rows per block): the top state's LDS
    // then allows three blocks per CU, which measured faster on the 1M x 1M
    // r_k pass (1.05 s) than two fragments (128 rows per block, half the
    // base-tile traffic per MFMA, two blocks per CU: 1.11 s).  The kernel
    // keeps FR as a template parameter.
    constexpr int FR = 1;
    unsigned grid = ceil_div(nq, 64 * FR);
    uint32_t KS = h->d / 32;
#define L(ks)                                                                                     \
    case ks:                                                                                      \
        hipLaunchKernelGGL((k_bf_mfma<T, ks, C, MODE, FR>), dim3(grid), dim3(256), 0, h->stream, qa, qn, \
                           nq, xb, h->norms.as<float>(), h->n, h->d, ci, cn, kth, hc, hi, hd, cap);  \
        break;
    switch (KS) {
        L(1) L(2) L(3) L(4) L(5) L(6) L(8) L(12) L(16) L(24) L(30) L(32)
        default:
            throw Error{HRNN_E_ARG, "bruteforce: MFMA path supports d/32 in {1..6,8,12,16,24,30,32}"};
    }
#undef L
    HIPC(hipGetLastError());
}

static bool mfma_ok(hrnn_dev_t h) {
    uint32_t KS = h->d / 32;
    if (h->d % 32) return false;
    switch (KS) {
        case 1: case 2: case 3: case 4: case 5: case 6: case 8: case 12: case 16: case 24: case 30: case 32:
            return true;
        default:
            return false;
    }
}

template <int C>
static void topk_c(hrnn_dev_t h, const float *q32, const void *q16, const float *qn, uint32_t nq,
                   uint32_t kk, int prec, uint32_t *ids_dev, float *d_dev) {
    uint64_t nqC = (uint64_t)nq * C;
    h->s6.reserve(nqC * 4 + (uint64_t)nq * 4 + 64);
    uint32_t *ci = h->s6.as<uint32_t>();
    uint32_t *cn = ci + nqC;
    unsigned grid = ceil_div(nq, kRB);
    if (prec == HRNN_PREC_BF16 && mfma_ok(h)) {
        launch_mfma<__hip_bfloat16, C, 0>(h, (const __hip_bfloat16 *)q16, qn, nq,
                                           h->xb.as<__hip_bfloat16>(), ci, cn, nullptr, nullptr,
                                           nullptr, nullptr, 0);
    } else if (prec == HRNN_PREC_F16 && mfma_ok(h)) {
        launch_mfma<__half, C, 0>(h, (const __half *)q16, qn, nq, h->xh.as<__half>(), ci, cn,
                                   nullptr, nullptr, nullptr, nullptr, 0);
    } else {
        hipLaunchKernelGGL((k_bf_f32<C, 0>), dim3(grid), dim3(256), 0, h->stream, q32, nq,
                           h->x.as<float>(), h->n, h->d, ci, cn, nullptr, nullptr, nullptr, nullptr, 0u);
        HIPC(hipGetLastError());
    }
    hipLaunchKernelGGL((k_rerank<C>), dim3(ceil_div(nq, 4)), dim3(256), 0, h->stream, q32, nq,
                       h->x.as<float>(), h->d, ci, kk, ids_dev, d_dev);
    HIPC(hipGetLastError());
}

// Top-kk of nq query rows (or of the base itself when self_base) over the base.
void bf_topk_base(hrnn_dev_t h, const float *q_dev, uint32_t nq, bool self_base, uint32_t kk,
                  int prec, uint32_t *ids_dev, float *d_dev) {
    REQUIRE(kk >= 1 && kk <= 64 && kk <= h->n, HRNN_E_ARG, "bruteforce: k+1 must be in [1, 64] and <= n");
    const float *q32;
    const void *q16;
    const float *qn;
    if (self_base) {
        q32 = h->x.as<float>();
        q16 = (prec == HRNN_PREC_F16) ? (const void *)h->xh.p : (const void *)h->xb.p;
        qn = h->norms.as<float>();
    } else {
        h->s7.reserve((uint64_t)nq * h->d * 4 + (uint64_t)nq * 4 + 256);
        __hip_bfloat16 *qb = h->s7.as<__hip_bfloat16>();
        __half *qh = reinterpret_cast<__half *>(qb + (uint64_t)nq * h->d);
        float *qnn = reinterpret_cast<float *>(qh + (uint64_t)nq * h->d);
        hipLaunchKernelGGL(k_query_prep, dim3(ceil_div(nq, 4)), dim3(256), 0, h->stream, q_dev,
                           (uint64_t)nq, h->d, qb, qh, qnn);
        HIPC(hipGetLastError());
        q32 = q_dev;
        q16 = (prec == HRNN_PREC_F16) ? (const void *)qh : (const void *)qb;
        qn = qnn;
    }
    uint32_t want = std::max<uint32_t>(2 * kk, 1);
    if (want <= 32)
        topk_c<32>(h, q32, q16, qn, nq, kk, prec, ids_dev, d_dev);
    else
        topk_c<64>(h, q32, q16, qn, nq, kk, prec, ids_dev, d_dev);
}

void bf_rknn_count_fill(hrnn_dev_t h, const float *q_dev, uint32_t nq, const float *kth_dev,
                        int prec, uint32_t *cnt_dev, uint32_t *ids_dev, float *d_dev,
                        uint32_t cap_per_q) {
    HIPC(hipMemsetAsync(cnt_dev, 0, (uint64_t)nq * 4, h->stream));
    unsigned grid = ceil_div(nq, kRB);
    // the Gram form decides exactly when every value is an integer in [0, 255]
    // and d <= 128 (all norms and products below 2^24)
    bool exact16 = h->base_u8 && h->d <= 128 && mfma_ok(h) && prec != HRNN_PREC_F32;
    if (exact16) {
        h->s7.reserve((uint64_t)nq * h->d * 4 + (uint64_t)nq * 4 + 256);
        __hip_bfloat16 *qb = h->s7.as<__hip_bfloat16>();
        __half *qh = reinterpret_cast<__half *>(qb + (uint64_t)nq * h->d);
        float *qnn = reinterpret_cast<float *>(qh + (uint64_t)nq * h->d);
        hipLaunchKernelGGL(k_query_prep, dim3(ceil_div(nq, 4)), dim3(256), 0, h->stream, q_dev,
                           (uint64_t)nq, h->d, qb, qh, qnn);
        HIPC(hipGetLastError());
        // queries must be exact in BF16 too; checked by the caller
        launch_mfma<__hip_bfloat16, 32, 1>(h, qb, qnn, nq, h->xb.as<__hip_bfloat16>(), nullptr, nullptr,
                                           kth_dev, cnt_dev, ids_dev, d_dev, cap_per_q);
    } else {
        hipLaunchKernelGGL((k_bf_f32<32, 1>), dim3(grid), dim3(256), 0, h->stream, q_dev, nq,
                           h->x.as<float>(), h->n, h->d, nullptr, nullptr, kth_dev, cnt_dev,
                           ids_dev, d_dev, cap_per_q);
        HIPC(hipGetLastError());
    }
}

}  // namespace hrnn

using namespace hrnn;

extern "C" {

int hrnn_dev_bf_kth(hrnn_dev_t h, uint32_t k, int prec, float *kth, uint32_t *ids) {
    if (!h) return HRNN_E_ARG;
    return guarded(h, [&] {
        REQUIRE(h->n > 0, HRNN_E_STATE, "bf_kth: set_base first");
        REQUIRE(kth, HRNN_E_ARG, "bf_kth: null output");
        uint32_t kk = std::min<uint64_t>(k + 1, h->n);
        uint64_t n = h->n;
        DevBuf di, dd;
        di.reserve(n * kk * 4);
        dd.reserve(n * kk * 4);
        bf_topk_base(h, nullptr, (uint32_t)n, true, kk, prec, di.as<uint32_t>(), dd.as<float>());
        std::vector<float> all(n * kk);
        d2h(h, all.data(), dd.p, n * kk * 4);
        for (uint64_t i = 0; i < n; i++) kth[i] = all[i * kk + (kk - 1)];
        if (ids) d2h(h, ids, di.p, n * kk * 4);
        di.release();
        dd.release();
    });
}

int hrnn_dev_bf_knn(hrnn_dev_t h, const float *q, uint32_t nq, uint32_t k, int prec,
                    uint32_t *ids, float *d2) {
    if (!h) return HRNN_E_ARG;
    return guarded(h, [&] {
        REQUIRE(h->n > 0, HRNN_E_STATE, "bf_knn: set_base first");
        REQUIRE(q && ids, HRNN_E_ARG, "bf_knn: null argument");
        if (!nq) return;
        DevBuf dq, di, dd;
        dq.reserve((uint64_t)nq * h->d * 4);
        di.reserve((uint64_t)nq * k * 4);
        dd.reserve((uint64_t)nq * k * 4);
        h2d(h, dq.p, q, (uint64_t)nq * h->d * 4);
        bf_topk_base(h, dq.as<float>(), nq, false, k, prec, di.as<uint32_t>(), dd.as<float>());
        d2h(h, ids, di.p, (uint64_t)nq * k * 4);
        if (d2) d2h(h, d2, dd.p, (uint64_t)nq * k * 4);
        dq.release(); di.release(); dd.release();
    });
}

int hrnn_dev_bf_rknn(hrnn_dev_t h, const float *q, uint32_t nq, const float *kth, int prec,
                     uint64_t *res_off, uint32_t *res_ids, float *res_d2, uint64_t cap,
                     uint64_t *needed) {
    if (!h) return HRNN_E_ARG;
    return guarded(h, [&] {
        REQUIRE(h->n > 0, HRNN_E_STATE, "bf_rknn: set_base first");
        REQUIRE(q && kth && res_off, HRNN_E_ARG, "bf_rknn: null argument");
        res_off[0] = 0;
        if (!nq) { if (needed) *needed = 0; return; }
        uint64_t n = h->n;
        // Gram-form path only when the queries are integers in [0, 255] as well
        bool q_int = true;
        for (uint64_t i = 0; i < (uint64_t)nq * h->d && q_int; i++) {
            float v = q[i];
            q_int = (v == std::nearbyint(v)) && v >= 0.f && v <= 255.f;
        }
        int p = q_int ? prec : HRNN_PREC_F32;
        const uint32_t chunk = 8192;
        uint32_t cap_q = 1024;
        DevBuf dq, dk, dc, di, dd;
        dk.reserve(n * 4);
        h2d(h, dk.p, kth, n * 4);
        std::vector<uint32_t> ids_all;
        std::vector<float> d_all;
        std::vector<uint64_t> off(nq + 1, 0);
        for (uint32_t q0 = 0; q0 < nq; q0 += chunk) {
            uint32_t nqc = std::min(chunk, nq - q0);
            dq.reserve((uint64_t)nqc * h->d * 4);
            dc.reserve((uint64_t)nqc * 4);
            h2d(h, dq.p, q + (uint64_t)q0 * h->d, (uint64_t)nqc * h->d * 4);
            std::vector<uint32_t> cnt(nqc);
            for (;;) {
                di.reserve((uint64_t)nqc * cap_q * 4);
                dd.reserve((uint64_t)nqc * cap_q * 4);
                bf_rknn_count_fill(h, dq.as<float>(), nqc, dk.as<float>(), p, dc.as<uint32_t>(),
                                   di.as<uint32_t>(), dd.as<float>(), cap_q);
                d2h(h, cnt.data(), dc.p, (uint64_t)nqc * 4);
                uint32_t mx = 0;
                for (uint32_t c : cnt) mx = std::max(mx, c);
                if (mx <= cap_q) break;
                cap_q = mx;
            }
            std::vector<uint32_t> hi((uint64_t)nqc * cap_q);
            std::vector<float> hd((uint64_t)nqc * cap_q);
            d2h(h, hi.data(), di.p, hi.size() * 4);
            d2h(h, hd.data(), dd.p, hd.size() * 4);
            for (uint32_t i = 0; i < nqc; i++) {
                std::vector<std::pair<float, uint32_t>> v(cnt[i]);
                for (uint32_t j = 0; j < cnt[i]; j++)
                    v[j] = {hd[(uint64_t)i * cap_q + j], hi[(uint64_t)i * cap_q + j]};
                std::sort(v.begin(), v.end());
                for (auto &pr : v) {
                    ids_all.push_back(pr.second);
                    d_all.push_back(pr.first);
                }
                off[q0 + i + 1] = ids_all.size();
            }
        }
        for (uint32_t i = 0; i <= nq; i++) res_off[i] = off[i];
        uint64_t total = ids_all.size();
        if (needed) *needed = total;
        if (total > cap) throw Error{HRNN_E_CAP, "bf_rknn: result capacity too small"};
        if (total && res_ids) std::copy(ids_all.begin(), ids_all.end(), res_ids);
        if (total && res_d2) std::copy(d_all.begin(), d_all.end(), res_d2);
        dq.release(); dk.release(); dc.release(); di.release(); dd.release();
    });
}

}  // extern "C"
