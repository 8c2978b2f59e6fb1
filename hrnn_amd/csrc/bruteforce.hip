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

// FP64 squared L2 in the oracle's order: float64 differences, squares and
// sums rounded one by one, left to right, never fused.
__device__ double l2_seq_f64(const float *__restrict__ a, const float *__restrict__ b, uint32_t d) {
#pragma clang fp contract(off)
    double r = 0.0;
    if ((d & 3u) == 0) {
        const float4 *a4 = reinterpret_cast<const float4 *>(a);
        const float4 *b4 = reinterpret_cast<const float4 *>(b);
        for (uint32_t e = 0; e < d / 4; e++) {
            const float4 u = a4[e], v = b4[e];
            double t = (double)u.x - (double)v.x;
            r += t * t;
            t = (double)u.y - (double)v.y;
            r += t * t;
            t = (double)u.z - (double)v.z;
            r += t * t;
            t = (double)u.w - (double)v.w;
            r += t * t;
        }
    } else {
        for (uint32_t e = 0; e < d; e++) {
            double t = (double)a[e] - (double)b[e];
            r += t * t;
        }
    }
    return r;
}

// Candidate-pass formats
enum : int { kPassF32 = 0, kPassBF16 = 1, kPassF16 = 2 };

__device__ __forceinline__ float round_to(float v, int pass) {
    if (pass == kPassBF16) return __bfloat162float(__float2bfloat16(v));
    if (pass == kPassF16) return __half2float(__float2half(v));
    return v;
}

// Per-format rounding statistics of the base for the certified top-C:
// squared norms of the BF16- and FP16-rounded rows, and over all rows the
// largest of each and the largest squared rounding error sum_e (r(x_e) -
// x_e)^2 (mx[0..3]: BF16 norm, FP16 norm, BF16 error, FP16 error, as float
// bits; non-negative floats order as unsigned integers).  r(x) - x is exact
// in FP32 (Sterbenz); an FP16 overflow gives +inf and so no certificate.
__global__ __launch_bounds__(256) void k_bf_prep(const float *__restrict__ x, uint64_t n, uint32_t d,
                                                 float *__restrict__ nb, float *__restrict__ nh,
                                                 unsigned *__restrict__ mx) {
    __shared__ unsigned s_mx[4][4];
    const uint32_t w = threadIdx.x / 64, lane = threadIdx.x & 63u;
    unsigned m[4] = {0u, 0u, 0u, 0u};
    for (uint64_t r = (uint64_t)blockIdx.x * 4 + w; r < n; r += (uint64_t)gridDim.x * 4) {
        float sb = 0.f, sh = 0.f, eb = 0.f, eh = 0.f;
        for (uint32_t e = lane; e < d; e += 64) {
            const float v = x[r * d + e];
            const float b = round_to(v, kPassBF16), hv = round_to(v, kPassF16);
            sb = __builtin_fmaf(b, b, sb);
            sh = __builtin_fmaf(hv, hv, sh);
            eb = __builtin_fmaf(b - v, b - v, eb);
            eh = __builtin_fmaf(hv - v, hv - v, eh);
        }
        for (int off = 32; off > 0; off >>= 1) {
            sb += __shfl_xor(sb, off, 64);
            sh += __shfl_xor(sh, off, 64);
            eb += __shfl_xor(eb, off, 64);
            eh += __shfl_xor(eh, off, 64);
        }
        if (lane == 0) {
            nb[r] = sb;
            nh[r] = sh;
        }
        m[0] = max(m[0], __float_as_uint(sb));
        m[1] = max(m[1], __float_as_uint(sh));
        m[2] = max(m[2], __float_as_uint(eb));
        m[3] = max(m[3], __float_as_uint(eh));
    }
    if (lane == 0)
        for (int i = 0; i < 4; i++) s_mx[w][i] = m[i];
    __syncthreads();
    if (threadIdx.x < 4) {
        unsigned v = 0u;
        for (int i = 0; i < 4; i++) v = max(v, s_mx[i][threadIdx.x]);
        atomicMax(&mx[threadIdx.x], v);
    }
}

// Rounded norms and maxima for the current base (recomputed when it changed).
static void ensure_bf_prep(hrnn_dev_t h) {
    if (h->bf_prep_version == h->base_version && h->bf_norm16.p) return;
    h->bf_norm16.reserve(h->n * 8 + 64);
    float *nb = h->bf_norm16.as<float>();
    unsigned *mx = reinterpret_cast<unsigned *>(nb + 2 * h->n);
    HIPC(hipMemsetAsync(mx, 0, 16, h->stream));
    unsigned grid = (unsigned)std::min<uint64_t>(ceil_div(h->n, 4), 4096);
    hipLaunchKernelGGL(k_bf_prep, dim3(grid), dim3(256), 0, h->stream, h->x.as<float>(), h->n, h->d,
                       nb, nb + h->n, mx);
    HIPC(hipGetLastError());
    h->bf_prep_version = h->base_version;
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

// Writes each row's candidates (sorted by the pass's own distance, then id),
// their count, the row's final threshold (its C-th distance; FLT_MAX while the
// list is not full) and, when kraw > 0, its kraw-th distance.  Every column
// not kept has a pass distance >= the final threshold: it was rejected
// against a threshold at least as large, or displaced from the list.
template <int C, int R>
__device__ void top_store(TopState<C, R> &s, uint64_t row0, uint32_t nq, uint32_t *cand_i, uint32_t *cand_n,
                          float *cand_thr, float *raw_kth, uint32_t kraw) {
    for (int t = threadIdx.x; t < R * C; t += blockDim.x) {
        int r = t / C, j = t % C;
        uint64_t row = row0 + r;
        if (row >= nq) continue;
        cand_i[row * C + j] = (uint32_t)j < s.top_n[r] ? s.top_i[r][j] : kInvalidId;
        if (j == 0) {
            cand_n[row] = s.top_n[r];
            cand_thr[row] = s.thr[r];
            if (kraw) raw_kth[row] = kraw <= s.top_n[r] ? s.top_d[r][kraw - 1] : FLT_MAX;
        }
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
    uint32_t *__restrict__ cand_i, uint32_t *__restrict__ cand_n, float *__restrict__ cand_thr,
    float *__restrict__ raw_kth, uint32_t kraw,
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
        top_store(st, row0, nq, cand_i, cand_n, cand_thr, raw_kth, kraw);
    }
}

// Column-split form of k_bf_mfma for d <= 128 (KS <= 4): wave w takes the
// columns 16 w .. 16 w + 15 of every 64-column tile for all 64 rows of the
// block (the A fragments of all four 16-row blocks in registers), so each
// staged base tile is read from LDS once per block instead of once per wave
// (the row-split form's LDS reads of the B fragments are four times the
// MFMAs' worth).  A row's candidates now come from all four waves: they are
// staged for the whole tile (up to 64 per row), merged after a barrier by the
// wave that owns the row (16 rows per wave, one lane each), and the
// thresholds reloaded after the next barrier.  A threshold a lane holds is
// never below the row's current one (it only falls), so a stale one stages
// more candidates, never fewer, and the merge decides each by (distance, id)
// against the list: the same lists as the row-split form.
template <int C>
struct TopStateCS {
    float top_d[kRB][C];
    uint32_t top_i[kRB][C];
    uint32_t top_n[kRB];
    float thr[kRB];
    float st_d[kRB][kCT];
    uint8_t st_o[kRB][kCT];   // staged column's offset in the tile
    uint32_t st_n[kRB];
};

template <class T, int KS, int C, int MODE>
__global__ __launch_bounds__(256) void k_bf_mfma_cs(
    const T *__restrict__ qa, const float *__restrict__ qn, uint32_t nq,
    const T *__restrict__ xb, const float *__restrict__ xn, uint64_t n, uint32_t d,
    uint32_t *__restrict__ cand_i, uint32_t *__restrict__ cand_n, float *__restrict__ cand_thr,
    float *__restrict__ raw_kth, uint32_t kraw,
    const float *__restrict__ kth, uint32_t *__restrict__ hit_cnt, uint32_t *__restrict__ hit_i,
    float *__restrict__ hit_d, uint32_t cap) {
    static_assert(KS <= 4, "the column-split form stages a whole tile of d <= 128");
    typedef typename Mfma<T>::vec vec;
    __shared__ __attribute__((aligned(16))) TopStateCS<C> st;
    constexpr int kBSP = KS * 32 + 8;   // bs row stride (elements)
    __shared__ __attribute__((aligned(16))) T bs[kCT * kBSP];
    __shared__ float s_xn[kCT];
    __shared__ float s_qn[kRB];

    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x / 64, gq = lane >> 4;
    const uint64_t row0 = (uint64_t)blockIdx.x * kRB;
    for (int r = threadIdx.x; r < kRB; r += blockDim.x) {
        s_qn[r] = (row0 + r < nq) ? qn[row0 + r] : 0.f;
        if (MODE == 0) {
            st.top_n[r] = 0;
            st.st_n[r] = 0;
            st.thr[r] = FLT_MAX;
        }
    }
    // A fragments of the four row blocks: rows 16 rb + lane % 16, dims
    // ks * 32 + 8 * (lane / 16) .. + 8
    vec afrag[4][KS];
#pragma unroll
    for (int rb = 0; rb < 4; rb++) {
        const uint64_t row = row0 + 16 * rb + (lane & 15u);
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
            if (row < nq) {
                afrag[rb][ks] = *reinterpret_cast<const vec *>(qa + row * d + ks * 32 + 8 * gq);
            } else {
#pragma unroll
                for (int e = 0; e < 8; e++) afrag[rb][ks][e] = 0;
            }
        }
    }
    __syncthreads();
    // this lane's output rows 16 rb + 4 (lane / 16) + i: norms and thresholds
    float qn_r[4][4], thr_r[4][4];
#pragma unroll
    for (int rb = 0; rb < 4; rb++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            qn_r[rb][i] = s_qn[16 * rb + 4 * gq + i];
            thr_r[rb][i] = FLT_MAX;
        }
    // the whole next tile prefetched into registers (64 columns x d dims:
    // NCH 64-dim chunks of two 16-byte loads per thread)
    constexpr int NCH = (KS + 1) / 2;
    vec pv[2 * NCH];
    float pxn = 0.f;
    auto fetch = [&](uint64_t c, int ch) {
        const int dims = (KS - 2 * ch >= 2) ? 64 : 32;
#pragma unroll
        for (int v2 = 0; v2 < 2; v2++) {
            const int u = threadIdx.x + v2 * 256;
            const int col = u / (dims / 8), part = u % (dims / 8);
            const uint64_t gc = c + col;
            if (u < kCT * (dims / 8) && gc < n) {
                pv[2 * ch + v2] = *reinterpret_cast<const vec *>(xb + gc * d + ch * 64 + part * 8);
            } else {
#pragma unroll
                for (int e = 0; e < 8; e++) pv[2 * ch + v2][e] = 0;
            }
        }
    };
    auto put = [&](int ch) {
        const int dims = (KS - 2 * ch >= 2) ? 64 : 32;
#pragma unroll
        for (int v2 = 0; v2 < 2; v2++) {
            const int u = threadIdx.x + v2 * 256;
            if (u < kCT * (dims / 8)) {
                const int col = u / (dims / 8), part = u % (dims / 8);
                *reinterpret_cast<vec *>(&bs[col * kBSP + ch * 64 + part * 8]) = pv[2 * ch + v2];
            }
        }
    };
#pragma unroll
    for (int ch = 0; ch < NCH; ch++) fetch(0, ch);
    if (threadIdx.x < kCT) pxn = threadIdx.x < n ? xn[threadIdx.x] : 0.f;

    const uint32_t cl = 16 * w + (lane & 15u);   // this lane's column in every tile
    for (uint64_t c0 = 0; c0 < n; c0 += kCT) {
        const uint64_t c1 = c0 + kCT;
        __syncthreads();   // the last tile's MFMAs read bs; its merges set thr
        if (threadIdx.x < kCT) s_xn[threadIdx.x] = pxn;
#pragma unroll
        for (int ch = 0; ch < NCH; ch++) {
            put(ch);
            if (c1 < n) fetch(c1, ch);
        }
        if (threadIdx.x < kCT && c1 < n) pxn = (c1 + threadIdx.x < n) ? xn[c1 + threadIdx.x] : 0.f;
        if (MODE == 0 && c0 > 0) {
#pragma unroll
            for (int rb = 0; rb < 4; rb++) {
                const float4 t4 = *reinterpret_cast<const float4 *>(&st.thr[16 * rb + 4 * gq]);
                thr_r[rb][0] = t4.x;
                thr_r[rb][1] = t4.y;
                thr_r[rb][2] = t4.z;
                thr_r[rb][3] = t4.w;
            }
        }
        __syncthreads();
        f32x4_t acc[4];
#pragma unroll
        for (int rb = 0; rb < 4; rb++) acc[rb] = f32x4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
            const vec b = *reinterpret_cast<const vec *>(&bs[cl * kBSP + ks * 32 + 8 * gq]);
#pragma unroll
            for (int rb = 0; rb < 4; rb++) acc[rb] = Mfma<T>::run(afrag[rb][ks], b, acc[rb]);
        }
        const uint64_t col = c0 + cl;
        const float xn_c = s_xn[cl];
        const float kth_c = (MODE == 1 && col < n) ? kth[col] : 0.f;
        if (col < n) {
#pragma unroll
            for (int rb = 0; rb < 4; rb++)
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int rl = 16 * rb + 4 * (int)gq + i;
                    if (row0 + rl < nq) {
                        float dv = qn_r[rb][i] + xn_c - 2.f * acc[rb][i];
                        dv = dv > 0.f ? dv : 0.f;
                        if (MODE == 0) {
                            if (dv < thr_r[rb][i]) {
                                const uint32_t t = atomicAdd(&st.st_n[rl], 1u);
                                st.st_d[rl][t] = dv;
                                st.st_o[rl][t] = (uint8_t)cl;
                            }
                        } else if (dv <= kth_c) {
                            const uint64_t q = row0 + rl;
                            const uint32_t s = atomicAdd(&hit_cnt[q], 1u);
                            if (s < cap) {
                                hit_i[q * cap + s] = (uint32_t)col;
                                hit_d[q * cap + s] = dv;
                            }
                        }
                    }
                }
        }
        if (MODE == 0) {
            __syncthreads();   // every wave's candidates of this tile are staged
            if (lane < 16) {   // wave w merges rows 16 w .. 16 w + 15
                const int r = (int)(16 * w + lane);
                const uint32_t ns = st.st_n[r];
                if (ns) {
                    uint32_t nn = st.top_n[r];
                    for (uint32_t t = 0; t < ns; t++) {
                        const float dv = st.st_d[r][t];
                        const uint32_t iv = (uint32_t)(c0 + st.st_o[r][t]);
                        if (nn == C && !pair_less(dv, iv, st.top_d[r][C - 1], st.top_i[r][C - 1])) continue;
                        uint32_t p = (nn == C) ? C - 1 : nn;
                        if (nn < C) nn++;
                        while (p > 0 && pair_less(dv, iv, st.top_d[r][p - 1], st.top_i[r][p - 1])) {
                            st.top_d[r][p] = st.top_d[r][p - 1];
                            st.top_i[r][p] = st.top_i[r][p - 1];
                            p--;
                        }
                        st.top_d[r][p] = dv;
                        st.top_i[r][p] = iv;
                    }
                    st.top_n[r] = nn;
                    st.st_n[r] = 0;
                    st.thr[r] = (nn == C) ? st.top_d[r][C - 1] : FLT_MAX;
                }
            }
        }
    }
    if (MODE == 0) {
        __syncthreads();   // every row's merges are done
        for (int t = threadIdx.x; t < kRB * C; t += blockDim.x) {
            const int r = t / C, j = t % C;
            const uint64_t row = row0 + r;
            if (row >= nq) continue;
            cand_i[row * C + j] = (uint32_t)j < st.top_n[r] ? st.top_i[r][j] : kInvalidId;
            if (j == 0) {
                cand_n[row] = st.top_n[r];
                cand_thr[row] = st.thr[r];
                if (kraw) raw_kth[row] = kraw <= st.top_n[r] ? st.top_d[r][kraw - 1] : FLT_MAX;
            }
        }
    }
}

// FP32 VALU candidate pass: 4x4 micro-tile per lane, direct differences.
// mode 0: top-C; mode 1: RkNN hits.  In mode 1 a pair whose distance lies
// within band_rel * max(d2, r_k) of r_k (a bound on the FP32 rounding of
// both, any summation order) is re-decided: with kth64, by its FP64
// distance against the FP64 radius (chk[0] counts these pairs, chk[1] those
// whose FP64 decision differs from the reference-order FP32 one); without,
// by its reference-order FP32 distance.
constexpr int kFCH = 16;   // dims per chunk
template <int C, int MODE>
__global__ __launch_bounds__(256) void k_bf_f32(
    const float *__restrict__ qa, uint32_t nq, const float *__restrict__ x, uint64_t n, uint32_t d,
    uint32_t *__restrict__ cand_i, uint32_t *__restrict__ cand_n, float *__restrict__ cand_thr,
    float *__restrict__ raw_kth, uint32_t kraw, const float *__restrict__ kth,
    uint32_t *__restrict__ hit_cnt, uint32_t *__restrict__ hit_i, float *__restrict__ hit_d,
    uint32_t cap, float band_rel, const double *__restrict__ kth64,
    unsigned long long *__restrict__ chk) {
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
                            const float kv = kth[col];
                            const bool near = fabsf(dv - kv) <= band_rel * fmaxf(dv, kv) + 1e-30f;
                            bool in = dv <= kv;
                            if (near) {
                                dv = l2_ref16_thread(qa + row * d, x + col * d, d);
                                in = dv <= kv;
                                if (kth64) {
                                    const bool in64 = l2_seq_f64(qa + row * d, x + col * d, d) <= kth64[col];
                                    atomicAdd(&chk[0], 1ull);
                                    if (in64 != in) atomicAdd(&chk[1], 1ull);
                                    in = in64;
                                }
                            }
                            if (in) {
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
    if (MODE == 0) top_store(st, row0, nq, cand_i, cand_n, cand_thr, raw_kth, kraw);
}

// Certificate of a candidate pass (see k_rerank_cert).  Unit roundoff 2^-23
// throughout, so that FP32 sums in any order and with round-to-nearest or
// truncation are covered.
struct CertParam {
    int pass = kPassF32;          // format of the candidate pass
    int exact_base = 0;           // every base value an integer in [0, 255] and d <= 128
    const unsigned *mx = nullptr; // base maxima of k_bf_prep (BF16 / FP16 passes)
    double g = 0.0;               // gamma_{d+2} = (d + 2) u / (1 - (d + 2) u)
};

// Re-rank and certificate: one wave per row.
// 1. The C candidates' FP32 distances in the reference's order, sorted by
//    (dist, id): ids/dist get the first kk, r32 = the kk-th.
// 2. Their FP64 distances (l2_seq_f64): r64 = the kk-th smallest.
// 3. Certificate: every column left out of the candidates has, in the
//    pass's format, a distance >= thr (the row's C-th).  A lower bound L on
//    the exact distance of every such column follows:
//    * FP32 pass: d~ <= D^2 (1 + g), so D^2 >= thr / (1 + g);
//    * BF16/FP16 Gram pass with rounded rows q~, x~ and their norms:
//      |d~ - ||q~ - x~||^2| <= A = (g + 2u(1 + g)) (|q~| + X)^2, X the
//      largest rounded base norm; ||q - x|| >= ||q~ - x~|| - e_q - E, e_q =
//      ||q~ - q|| and E the largest ||x~ - x||; so
//      D >= sqrt(max(thr - A, 0)) - e_q - E.
//    The row is certified when L^2 (1 - g) > r32 and L^2 (1 - 1e-12) > r64:
//    no column outside the candidates can then be among the first kk in
//    FP32 (reference order, any tie order) or in FP64, so ids, dist, r32 and
//    r64 equal those of an exhaustive re-rank.  Also certified: rows with
//    fewer than C candidates (every column is one), and rows of integer
//    queries in [0, 255] over an exact base (every distance exact).
template <int C>
__global__ __launch_bounds__(256) void k_rerank_cert(
    const float *__restrict__ qa, uint32_t nq, const float *__restrict__ x, uint32_t d,
    const uint32_t *__restrict__ cand_i, const uint32_t *__restrict__ cand_n,
    const float *__restrict__ cand_thr, uint32_t kk, CertParam cp, uint32_t *__restrict__ ids,
    float *__restrict__ dist, double *__restrict__ r64, uint8_t *__restrict__ cert,
    uint32_t *__restrict__ raw_ids) {
    __shared__ float sd[4][C];
    __shared__ uint32_t si[4][C];
    __shared__ double s64[4][C];
    __shared__ double s_r64[4];
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
    // the query row's rounding statistics (BF16/FP16 passes) and whether it
    // is made of integers in [0, 255]
    double qq = 0.0, ee = 0.0;
    bool q_u8 = true;
    for (uint32_t e = lane; valid && e < d; e += 64) {
        const float v = qv[e];
        const float r = round_to(v, cp.pass);
        qq += (double)r * (double)r;
        ee += ((double)r - (double)v) * ((double)r - (double)v);
        q_u8 &= v == rintf(v) && v >= 0.f && v <= 255.f;
    }
    for (int off = 32; off > 0; off >>= 1) {
        qq += __shfl_xor(qq, off, 64);
        ee += __shfl_xor(ee, off, 64);
    }
    q_u8 = __all(q_u8);
    __syncthreads();
    if (valid && lane < C) {
        const uint32_t id = si[w][lane];
        s64[w][lane] = id != kInvalidId ? l2_seq_f64(qv, x + (uint64_t)id * d, d) : DBL_MAX;
    }
    __syncthreads();
    if (valid && lane < C) {   // the kk-th smallest FP64 distance (equal values counted apart)
        const double v = s64[w][lane];
        uint32_t rank = 0;
        for (int j = 0; j < C; j++) {
            const double u = s64[w][j];
            rank += (u < v || (u == v && j < (int)lane)) ? 1u : 0u;
        }
        if (rank == kk - 1) {
            s_r64[w] = v;
            if (r64) r64[row] = v;
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
        if (raw_ids)
            for (uint32_t j = 0; j < kk; j++) raw_ids[row * kk + j] = cand_i[row * C + j];
    }
    if (valid && lane == 0 && cert) {
        const uint32_t cn = cand_n[row];
        bool ok = cn < (uint32_t)C || (cp.exact_base && q_u8);
        if (!ok) {
            const double r64v = s_r64[w];
            const double r32 = sd[w][kk - 1];
            const double thr = cand_thr[row];
            const double g = cp.g, u = 0x1p-23;
            double lb2;
            if (cp.pass == kPassF32) {
                lb2 = thr / (1.0 + g);
            } else {
                const int o = cp.pass == kPassBF16 ? 0 : 1;
                const double X = sqrt((double)__uint_as_float(cp.mx[o]) * (1.0 + 2.0 * g));
                const double E = sqrt((double)__uint_as_float(cp.mx[2 + o]) * (1.0 + 2.0 * g)) * (1.0 + 1e-9);
                const double qn = sqrt(qq) * (1.0 + 1e-9), eq = sqrt(ee) * (1.0 + 1e-9);
                const double A = (g + 2.0 * u * (1.0 + g)) * (qn + X) * (qn + X) * (1.0 + 1e-9) + 1e-30;
                const double t = sqrt(fmax(thr - A, 0.0)) - eq - E;
                lb2 = t > 0.0 ? t * t : 0.0;
            }
            ok = lb2 * (1.0 - g) > r32 && lb2 * (1.0 - 1e-12) > r64v;
        }
        cert[row] = ok ? 1 : 0;
    }
}

// FP32 -> BF16/FP16 copies of external queries and the squared norms of the
// rounded rows (the Gram form then measures ||q~ - x~||^2).
__global__ void k_query_prep(const float *__restrict__ q, uint64_t nq, uint32_t d,
                             __hip_bfloat16 *__restrict__ qb, __half *__restrict__ qh,
                             float *__restrict__ qnb, float *__restrict__ qnh) {
    uint64_t r = (uint64_t)blockIdx.x * 4 + threadIdx.x / 64;
    if (r >= nq) return;
    uint32_t lane = threadIdx.x & 63u;
    float ab = 0.f, ah = 0.f;
    for (uint32_t e = lane; e < d; e += 64) {
        float v = q[r * d + e];
        const __hip_bfloat16 b = __float2bfloat16(v);
        const __half hv = __float2half(v);
        qb[r * d + e] = b;
        qh[r * d + e] = hv;
        const float bf = __bfloat162float(b), hf = __half2float(hv);
        ab = __builtin_fmaf(bf, bf, ab);
        ah = __builtin_fmaf(hf, hf, ah);
    }
    for (int off = 32; off > 0; off >>= 1) {
        ab += __shfl_down(ab, off, 64);
        ah += __shfl_down(ah, off, 64);
    }
    if (lane == 0) {
        qnb[r] = ab;
        qnh[r] = ah;
    }
}

// dst row t = src row idx[t] (rows of `words` 32-bit words)
__global__ void k_gather_rows(const uint32_t *__restrict__ src, const uint32_t *__restrict__ idx,
                              uint32_t m, uint32_t words, uint32_t *__restrict__ dst) {
    const uint64_t t = blockIdx.x;
    if (t >= m) return;
    const uint64_t s = idx[t];
    for (uint32_t e = threadIdx.x; e < words; e += blockDim.x) dst[t * words + e] = src[s * words + e];
}

// The redone rows' results back into the full arrays.
__global__ void k_scatter_topk(const uint32_t *__restrict__ idx, uint32_t m, uint32_t kk,
                               const uint32_t *__restrict__ sids, const float *__restrict__ sdist,
                               const double *__restrict__ sr64, const uint8_t *__restrict__ scert,
                               uint32_t *__restrict__ ids, float *__restrict__ dist,
                               double *__restrict__ r64, uint8_t *__restrict__ cert) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (uint64_t)m * kk) return;
    const uint64_t i = t / kk, j = t % kk, row = idx[i];
    ids[row * kk + j] = sids[t];
    dist[row * kk + j] = sdist[t];
    if (j == 0) {
        r64[row] = sr64[i];
        cert[row] = scert[i];
    }
}

// ---- launchers -------------------------------------------------------------

template <class T, int C, int MODE>
static void launch_mfma(hrnn_dev_t h, const T *qa, const float *qn, uint32_t nq, const T *xb,
                        const float *xn, uint32_t *ci, uint32_t *cn, float *cthr, float *raw_kth,
                        uint32_t kraw, const float *kth, uint32_t *hc, uint32_t *hi, float *hd,
                        uint32_t cap) {
    // 16-row fragments per wave, "bf_row_frags": one by default (64 rows per
    // block: the top state's LDS allows three blocks per CU, which measured
    // faster on the 1M x 1M r_k pass, 1.05 s, than two fragments, 128 rows
    // per block with half the base-tile traffic per MFMA and two blocks per
    // CU: 1.11 s); two where the autotuner finds them faster, for 32-wide
    // candidate lists and d <= 256 (wider rows or lists would not fit).
    const bool fr2 = h->tune_bf_frags == 2;
    const bool cs = h->tune_bf_layout == 1;
    uint32_t KS = h->d / 32;
#define L(ks)                                                                                     \
    case ks:                                                                                      \
        if constexpr (ks <= 4) {                                                                  \
            if (cs) {                                                                             \
                hipLaunchKernelGGL((k_bf_mfma_cs<T, ks, C, MODE>), dim3(ceil_div(nq, 64)), dim3(256), 0, \
                                   h->stream, qa, qn, nq, xb, xn, h->n, h->d, ci, cn, cthr, raw_kth, kraw, \
                                   kth, hc, hi, hd, cap);                                         \
                break;                                                                            \
            }                                                                                     \
        }                                                                                         \
        if constexpr (C == 32 && ks <= 8) {                                                       \
            if (fr2) {                                                                            \
                hipLaunchKernelGGL((k_bf_mfma<T, ks, C, MODE, 2>), dim3(ceil_div(nq, 128)), dim3(256), 0, \
                                   h->stream, qa, qn, nq, xb, xn, h->n, h->d, ci, cn, cthr, raw_kth, kraw, \
                                   kth, hc, hi, hd, cap);                                         \
                break;                                                                            \
            }                                                                                     \
        }                                                                                         \
        hipLaunchKernelGGL((k_bf_mfma<T, ks, C, MODE, 1>), dim3(ceil_div(nq, 64)), dim3(256), 0, h->stream, \
                           qa, qn, nq, xb, xn, h->n, h->d, ci, cn, cthr, raw_kth, kraw, kth, hc, hi, hd, cap); \
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

// One candidate pass (format `pass`, top-C) over nq rows, then the re-rank
// and certificate.  q16/qn16: the rows rounded to the pass's format and their
// squared norms (BF16/FP16 passes).
template <int C>
static void bf_pass(hrnn_dev_t h, int pass, const float *q32, const void *q16, const float *qn16,
                    uint32_t nq, uint32_t kk, const CertParam &cp, uint32_t *ids, float *dist,
                    double *r64, uint8_t *cert, float *raw_kth, uint32_t *raw_ids) {
    const uint64_t nqC = (uint64_t)nq * C;
    h->s6.reserve(nqC * 4 + (uint64_t)nq * 8 + 64);
    uint32_t *ci = h->s6.as<uint32_t>();
    uint32_t *cn = ci + nqC;
    float *cthr = reinterpret_cast<float *>(cn + nq);
    const uint32_t kraw = raw_kth ? kk : 0;
    const float *nb = h->bf_norm16.as<float>();
    if (pass == kPassBF16) {
        launch_mfma<__hip_bfloat16, C, 0>(h, (const __hip_bfloat16 *)q16, qn16, nq, h->xb.as<__hip_bfloat16>(),
                                           nb, ci, cn, cthr, raw_kth, kraw, nullptr, nullptr, nullptr,
                                           nullptr, 0);
    } else if (pass == kPassF16) {
        launch_mfma<__half, C, 0>(h, (const __half *)q16, qn16, nq, h->xh.as<__half>(), nb + h->n, ci, cn,
                                   cthr, raw_kth, kraw, nullptr, nullptr, nullptr, nullptr, 0);
    } else {
        hipLaunchKernelGGL((k_bf_f32<C, 0>), dim3(ceil_div(nq, kRB)), dim3(256), 0, h->stream, q32, nq,
                           h->x.as<float>(), h->n, h->d, ci, cn, cthr, raw_kth, kraw, nullptr, nullptr,
                           nullptr, nullptr, 0u, 0.f, nullptr, nullptr);
        HIPC(hipGetLastError());
    }
    hipLaunchKernelGGL((k_rerank_cert<C>), dim3(ceil_div(nq, 4)), dim3(256), 0, h->stream, q32, nq,
                       h->x.as<float>(), h->d, ci, cn, cthr, kk, cp, ids, dist, r64, cert, raw_ids);
    HIPC(hipGetLastError());
}

static double gamma_u23(uint32_t m) {
    const double mu = (double)m * 0x1p-23;
    return mu / (1.0 - mu);
}

// Top-kk of nq query rows (or of the base itself when self_base) over the
// base, exact: after the first candidate pass (prec's format, C = 32 when
// 2 kk <= 32, else 64) every row must carry a certificate; the rows without
// one are redone with a 64-wide candidate list in the same format, then, if
// still uncertified, with the FP32 pass (64 wide).  Rows no pass certifies
// keep the last pass's results and are counted.
void bf_topk_base(hrnn_dev_t h, const float *q_dev, uint32_t nq, bool self_base, uint32_t kk,
                  int prec, uint32_t *ids_dev, float *d_dev, const BfExtra *ex) {
    REQUIRE(kk >= 1 && kk <= 64 && kk <= h->n, HRNN_E_ARG, "bruteforce: k+1 must be in [1, 64] and <= n");
    if (!nq) return;
    const bool mf = mfma_ok(h);
    const int pass0 = (prec == HRNN_PREC_BF16 && mf) ? kPassBF16 : (prec == HRNN_PREC_F16 && mf) ? kPassF16 : kPassF32;
    if (pass0 != kPassF32) ensure_bf_prep(h);
    const float *q32;
    const void *q16 = nullptr;
    const float *qn16 = nullptr;
    if (self_base) {
        q32 = h->x.as<float>();
        if (pass0 != kPassF32) {
            q16 = pass0 == kPassF16 ? (const void *)h->xh.p : (const void *)h->xb.p;
            qn16 = h->bf_norm16.as<float>() + (pass0 == kPassF16 ? h->n : 0);
        }
    } else {
        h->s7.reserve((uint64_t)nq * h->d * 4 + (uint64_t)nq * 8 + 256);
        __hip_bfloat16 *qb = h->s7.as<__hip_bfloat16>();
        __half *qh = reinterpret_cast<__half *>(qb + (uint64_t)nq * h->d);
        float *qnb = reinterpret_cast<float *>(qh + (uint64_t)nq * h->d);
        hipLaunchKernelGGL(k_query_prep, dim3(ceil_div(nq, 4)), dim3(256), 0, h->stream, q_dev,
                           (uint64_t)nq, h->d, qb, qh, qnb, qnb + nq);
        HIPC(hipGetLastError());
        q32 = q_dev;
        q16 = pass0 == kPassF16 ? (const void *)qh : (const void *)qb;
        qn16 = pass0 == kPassF16 ? qnb + nq : qnb;
    }
    CertParam cp;
    cp.exact_base = h->base_u8 && h->d <= 128;
    if (pass0 != kPassF32) cp.mx = reinterpret_cast<const unsigned *>(h->bf_norm16.as<float>() + 2 * h->n);
    cp.g = gamma_u23(h->d + 2);
    DevBuf r64b, certb;
    double *r64 = ex && ex->r64 ? ex->r64 : nullptr;
    if (!r64) {
        r64b.reserve((uint64_t)nq * 8);
        r64 = r64b.as<double>();
    }
    certb.reserve(nq);
    uint8_t *cert = certb.as<uint8_t>();
    struct Events {
        hipEvent_t e[3] = {nullptr, nullptr, nullptr};
        ~Events() {
            for (auto &x : e)
                if (x) (void)hipEventDestroy(x);
        }
    } evs;
    hipEvent_t *ev = evs.e;
    for (int i = 0; i < 3; i++) HIPC(hipEventCreate(&ev[i]));
    HIPC(hipEventRecord(ev[0], h->stream));
    const uint32_t C0 = 2 * kk <= 32 ? 32 : 64;
    cp.pass = pass0;
    float *raw_kth = ex ? ex->raw_kth : nullptr;
    uint32_t *raw_ids = ex ? ex->raw_ids : nullptr;
    if (C0 == 32)
        bf_pass<32>(h, pass0, q32, q16, qn16, nq, kk, cp, ids_dev, d_dev, r64, cert, raw_kth, raw_ids);
    else
        bf_pass<64>(h, pass0, q32, q16, qn16, nq, kk, cp, ids_dev, d_dev, r64, cert, raw_kth, raw_ids);
    HIPC(hipEventRecord(ev[1], h->stream));
    std::vector<uint8_t> ok(nq);
    d2h(h, ok.data(), cert, nq);
    std::vector<uint32_t> fail;
    for (uint32_t i = 0; i < nq; i++)
        if (!ok[i]) fail.push_back(i);
    hrnn_bf_stats st{};
    st.rows = nq;
    st.certified = nq - fail.size();
    // the rungs after the first pass: 64 wide in the same format, then FP32
    std::vector<int> rungs;
    if (pass0 != kPassF32 && C0 == 32) rungs.push_back(pass0);
    if (pass0 != kPassF32 || C0 == 32) rungs.push_back(kPassF32);
    for (int pass : rungs) {
        if (fail.empty()) break;
        const uint32_t m = (uint32_t)fail.size();
        if (pass == kPassF32) st.redo_f32 += m;
        else st.redo_wide += m;
        DevBuf gi, g32, g16, gn, sids, sdist, sr64, scert;
        gi.reserve((uint64_t)m * 4);
        h2d(h, gi.p, fail.data(), (uint64_t)m * 4);
        g32.reserve((uint64_t)m * h->d * 4);
        hipLaunchKernelGGL(k_gather_rows, dim3(m), dim3(256), 0, h->stream, (const uint32_t *)q32,
                           gi.as<uint32_t>(), m, h->d, g32.as<uint32_t>());
        HIPC(hipGetLastError());
        if (pass != kPassF32) {
            g16.reserve((uint64_t)m * h->d * 2);
            gn.reserve((uint64_t)m * 4);
            hipLaunchKernelGGL(k_gather_rows, dim3(m), dim3(256), 0, h->stream, (const uint32_t *)q16,
                               gi.as<uint32_t>(), m, h->d / 2, g16.as<uint32_t>());
            hipLaunchKernelGGL(k_gather_rows, dim3(m), dim3(64), 0, h->stream, (const uint32_t *)qn16,
                               gi.as<uint32_t>(), m, 1u, gn.as<uint32_t>());
            HIPC(hipGetLastError());
        }
        sids.reserve((uint64_t)m * kk * 4);
        sdist.reserve((uint64_t)m * kk * 4);
        sr64.reserve((uint64_t)m * 8);
        scert.reserve(m);
        CertParam cq = cp;
        cq.pass = pass;
        bf_pass<64>(h, pass, g32.as<float>(), g16.p, gn.as<float>(), m, kk, cq, sids.as<uint32_t>(),
                    sdist.as<float>(), sr64.as<double>(), scert.as<uint8_t>(), nullptr, nullptr);
        hipLaunchKernelGGL(k_scatter_topk, dim3(ceil_div((uint64_t)m * kk, 256)), dim3(256), 0, h->stream,
                           gi.as<uint32_t>(), m, kk, sids.as<uint32_t>(), sdist.as<float>(), sr64.as<double>(),
                           scert.as<uint8_t>(), ids_dev, d_dev, r64, cert);
        HIPC(hipGetLastError());
        std::vector<uint8_t> sok(m);
        d2h(h, sok.data(), scert.p, m);
        std::vector<uint32_t> next;
        for (uint32_t t = 0; t < m; t++)
            if (!sok[t]) next.push_back(fail[t]);
        fail.swap(next);
    }
    st.uncertified = fail.size();
    HIPC(hipEventRecord(ev[2], h->stream));
    HIPC(hipEventSynchronize(ev[2]));
    float ms1 = 0.f, ms2 = 0.f;
    HIPC(hipEventElapsedTime(&ms1, ev[0], ev[1]));
    HIPC(hipEventElapsedTime(&ms2, ev[0], ev[2]));
    st.ms_first = ms1;
    st.ms_total = ms2;
    if (ex && ex->st) *ex->st = st;
}

void bf_rknn_count_fill(hrnn_dev_t h, const float *q_dev, uint32_t nq, const float *kth_dev,
                        int prec, uint32_t *cnt_dev, uint32_t *ids_dev, float *d_dev,
                        uint32_t cap_per_q, const double *kth64_dev, unsigned long long *chk_dev) {
    HIPC(hipMemsetAsync(cnt_dev, 0, (uint64_t)nq * 4, h->stream));
    if (chk_dev) HIPC(hipMemsetAsync(chk_dev, 0, 16, h->stream));
    unsigned grid = ceil_div(nq, kRB);
    // the Gram form decides exactly when every value is an integer in [0, 255]
    // and d <= 128 (all norms and products below 2^24); the FP64 decision is
    // then the same as well
    bool exact16 = h->base_u8 && h->d <= 128 && mfma_ok(h) && prec != HRNN_PREC_F32;
    if (exact16) {
        h->s7.reserve((uint64_t)nq * h->d * 4 + (uint64_t)nq * 8 + 256);
        __hip_bfloat16 *qb = h->s7.as<__hip_bfloat16>();
        __half *qh = reinterpret_cast<__half *>(qb + (uint64_t)nq * h->d);
        float *qnb = reinterpret_cast<float *>(qh + (uint64_t)nq * h->d);
        hipLaunchKernelGGL(k_query_prep, dim3(ceil_div(nq, 4)), dim3(256), 0, h->stream, q_dev,
                           (uint64_t)nq, h->d, qb, qh, qnb, qnb + nq);
        HIPC(hipGetLastError());
        // queries must be exact in BF16 too (checked by the caller): the
        // rounded norms are then the FP32 norms
        launch_mfma<__hip_bfloat16, 32, 1>(h, qb, qnb, nq, h->xb.as<__hip_bfloat16>(), h->norms.as<float>(),
                                           nullptr, nullptr, nullptr, nullptr, 0, kth_dev, cnt_dev,
                                           ids_dev, d_dev, cap_per_q);
    } else {
        // the FP32 rounding of a distance and of r_k, any order: gamma_{d+2} each
        const double g = gamma_u23(h->d + 2);
        const float band = (float)(2.0 * g * (1.0 + 2.0 * g) + 1e-9);
        hipLaunchKernelGGL((k_bf_f32<32, 1>), dim3(grid), dim3(256), 0, h->stream, q_dev, nq,
                           h->x.as<float>(), h->n, h->d, nullptr, nullptr, nullptr, nullptr, 0u, kth_dev,
                           cnt_dev, ids_dev, d_dev, cap_per_q, band, kth64_dev, chk_dev);
        HIPC(hipGetLastError());
    }
}

}  // namespace hrnn

using namespace hrnn;

extern "C" {

int hrnn_dev_bf_kth_ex(hrnn_dev_t h, uint32_t k, int prec, float *kth, double *kth64, uint32_t *ids,
                       float *raw_kth, uint32_t *raw_ids, hrnn_bf_stats *stats) {
    if (!h) return HRNN_E_ARG;
    return guarded(h, [&] {
        REQUIRE(h->n > 0, HRNN_E_STATE, "bf_kth: set_base first");
        REQUIRE(kth, HRNN_E_ARG, "bf_kth: null output");
        uint32_t kk = std::min<uint64_t>(k + 1, h->n);
        uint64_t n = h->n;
        DevBuf di, dd, d64, drk, dri;
        di.reserve(n * kk * 4);
        dd.reserve(n * kk * 4);
        d64.reserve(n * 8);
        BfExtra ex;
        ex.r64 = d64.as<double>();
        ex.st = stats;
        if (raw_kth) {
            drk.reserve(n * 4);
            ex.raw_kth = drk.as<float>();
        }
        if (raw_ids) {
            dri.reserve(n * kk * 4);
            ex.raw_ids = dri.as<uint32_t>();
        }
        bf_topk_base(h, nullptr, (uint32_t)n, true, kk, prec, di.as<uint32_t>(), dd.as<float>(), &ex);
        std::vector<float> all(n * kk);
        d2h(h, all.data(), dd.p, n * kk * 4);
        for (uint64_t i = 0; i < n; i++) kth[i] = all[i * kk + (kk - 1)];
        if (ids) d2h(h, ids, di.p, n * kk * 4);
        if (kth64) d2h(h, kth64, d64.p, n * 8);
        if (raw_kth) d2h(h, raw_kth, drk.p, n * 4);
        if (raw_ids) d2h(h, raw_ids, dri.p, n * kk * 4);
    });
}

int hrnn_dev_bf_kth(hrnn_dev_t h, uint32_t k, int prec, float *kth, uint32_t *ids) {
    return hrnn_dev_bf_kth_ex(h, k, prec, kth, nullptr, ids, nullptr, nullptr, nullptr);
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

static int bf_rknn_impl(hrnn_dev_t h, const float *q, uint32_t nq, const float *kth, const double *kth64,
                        int prec, uint64_t *res_off, uint32_t *res_ids, float *res_d2, uint64_t cap,
                        uint64_t *needed, uint64_t *chk_out) {
    if (!h) return HRNN_E_ARG;
    return guarded(h, [&] {
        REQUIRE(h->n > 0, HRNN_E_STATE, "bf_rknn: set_base first");
        REQUIRE(q && kth && res_off, HRNN_E_ARG, "bf_rknn: null argument");
        res_off[0] = 0;
        if (chk_out) chk_out[0] = chk_out[1] = 0;
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
        DevBuf dq, dk, dc, di, dd, dk64, dchk;
        dk.reserve(n * 4);
        h2d(h, dk.p, kth, n * 4);
        if (kth64) {
            dk64.reserve(n * 8);
            h2d(h, dk64.p, kth64, n * 8);
            dchk.reserve(16);
        }
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
                                   di.as<uint32_t>(), dd.as<float>(), cap_q, dk64.as<double>(),
                                   dchk.as<unsigned long long>());
                d2h(h, cnt.data(), dc.p, (uint64_t)nqc * 4);
                uint32_t mx = 0;
                for (uint32_t c : cnt) mx = std::max(mx, c);
                if (mx <= cap_q) break;
                cap_q = mx;
            }
            if (kth64 && chk_out) {
                unsigned long long c2[2];
                d2h(h, c2, dchk.p, 16);
                chk_out[0] += c2[0];
                chk_out[1] += c2[1];
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
    });
}

int hrnn_dev_bf_rknn(hrnn_dev_t h, const float *q, uint32_t nq, const float *kth, int prec,
                     uint64_t *res_off, uint32_t *res_ids, float *res_d2, uint64_t cap,
                     uint64_t *needed) {
    return bf_rknn_impl(h, q, nq, kth, nullptr, prec, res_off, res_ids, res_d2, cap, needed, nullptr);
}

int hrnn_dev_bf_rknn_f64(hrnn_dev_t h, const float *q, uint32_t nq, const float *kth, const double *kth64,
                         uint64_t *res_off, uint32_t *res_ids, float *res_d2, uint64_t cap,
                         uint64_t *needed, uint64_t *rechecked, uint64_t *flipped) {
    if (!h) return HRNN_E_ARG;
    if (!kth64) {
        h->err = "bf_rknn_f64: null kth64";
        return HRNN_E_ARG;
    }
    uint64_t chk[2] = {0, 0};
    int rc = bf_rknn_impl(h, q, nq, kth, kth64, HRNN_PREC_BF16, res_off, res_ids, res_d2, cap, needed, chk);
    if (rechecked) *rechecked = chk[0];
    if (flipped) *flipped = chk[1];
    return rc;
}

}  // extern "C"
