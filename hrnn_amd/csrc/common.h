// common.h -- shared definitions of the gfx950 kernels and the C-ABI layer.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "../../include/hrnn_dev.h"

namespace hrnn {

struct Error {
    int code;
    std::string msg;
};

#define HIPC(expr)                                                                        \
    do {                                                                                  \
        hipError_t e_ = (expr);                                                           \
        if (e_ != hipSuccess)                                                             \
            throw ::hrnn::Error{HRNN_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_) + \
                                    " (" + __FILE__ + ":" + std::to_string(__LINE__) + ")"}; \
    } while (0)

#define REQUIRE(cond, code, text)                                                        \
    do {                                                                                  \
        if (!(cond)) throw ::hrnn::Error{(code), std::string(text)};                      \
    } while (0)

// Growable device buffer owned by a handle.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    void reserve(size_t b) {
        if (b <= bytes) return;
        if (p) HIPC(hipFree(p));
        p = nullptr;
        bytes = 0;
        hipError_t e = hipMalloc(&p, b ? b : 1);
        if (e != hipSuccess) {
            p = nullptr;
            throw Error{HRNN_E_OOM, "hipMalloc(" + std::to_string(b) + " B): " + hipGetErrorString(e)};
        }
        bytes = b;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T>
    T *as() const { return static_cast<T *>(p); }
};

inline unsigned ceil_div(uint64_t a, uint64_t b) { return (unsigned)((a + b - 1) / b); }

constexpr uint32_t kInvalidId = 0xFFFFFFFFu;

// Squared L2 with the reference's summation order (space_l2.h:25-54): 16
// partial sums, element e into partial e % 16 with a fused multiply-add, then
// the partials added left to right.  Computed by a group of 16 consecutive
// lanes: lane g (0..15) of the group owns partial g.  Every lane of the group
// returns the full sum.  d must be a multiple of 16.
__device__ __forceinline__ float l2_ref16_group(const float *__restrict__ a,
                                                const float *__restrict__ b, uint32_t d,
                                                int g) {
    float p = 0.f;
    for (uint32_t e = (uint32_t)g; e < d; e += 16) {
        float t = a[e] - b[e];
        p = __builtin_fmaf(t, t, p);
    }
    // left-to-right sum of the 16 partials of this group
    int base = (int)(threadIdx.x & 63u) & ~15;
    float r = __shfl(p, base + 0, 64);
#pragma unroll
    for (int l = 1; l < 16; l++) r += __shfl(p, base + l, 64);
    return r;
}

// Same order with the second operand stored as BF16 (exactly representable
// data only: integers in [-256, 256], where the result equals the FP32 one
// bit for bit).
__device__ __forceinline__ float l2_ref16_group_bf16(const float *__restrict__ a,
                                                     const __hip_bfloat16 *__restrict__ b,
                                                     uint32_t d, int g) {
    float p = 0.f;
    for (uint32_t e = (uint32_t)g; e < d; e += 16) {
        float t = a[e] - __bfloat162float(b[e]);
        p = __builtin_fmaf(t, t, p);
    }
    int base = (int)(threadIdx.x & 63u) & ~15;
    float r = __shfl(p, base + 0, 64);
#pragma unroll
    for (int l = 1; l < 16; l++) r += __shfl(p, base + l, 64);
    return r;
}

// Sequential-order squared L2 (space_l2.h:6-20 with contraction), used when
// d is not a multiple of 16.
__device__ __forceinline__ float l2_seq(const float *__restrict__ a, const float *__restrict__ b,
                                        uint32_t d) {
    float r = 0.f;
    for (uint32_t e = 0; e < d; e++) {
        float t = a[e] - b[e];
        r = __builtin_fmaf(t, t, r);
    }
    return r;
}

// Distance from an FP32 vector (LDS) to row `id` of the base stored as VT
// (float, or __hip_bfloat16 for exactly representable data), computed by a
// 16-lane group in the reference's order (sequential when d % 16 != 0).
template <class VT>
__device__ __forceinline__ float dist_row(const float *__restrict__ qv, const VT *__restrict__ base,
                                          uint64_t id, uint32_t d, uint32_t g);
template <>
__device__ __forceinline__ float dist_row<float>(const float *__restrict__ qv,
                                                 const float *__restrict__ base, uint64_t id,
                                                 uint32_t d, uint32_t g) {
    const float *xv = base + id * d;
    return ((d & 15u) == 0) ? l2_ref16_group(qv, xv, d, (int)g) : l2_seq(qv, xv, d);
}
template <>
__device__ __forceinline__ float dist_row<__hip_bfloat16>(const float *__restrict__ qv,
                                                          const __hip_bfloat16 *__restrict__ base,
                                                          uint64_t id, uint32_t d, uint32_t g) {
    const __hip_bfloat16 *xv = base + id * d;
    if ((d & 15u) == 0) return l2_ref16_group_bf16(qv, xv, d, (int)g);
    float r = 0.f;
    for (uint32_t e = 0; e < d; e++) {
        float t = qv[e] - __bfloat162float(xv[e]);
        r = __builtin_fmaf(t, t, r);
    }
    return r;
}

template <>
__device__ __forceinline__ float dist_row<uint8_t>(const float *__restrict__ qv,
                                                   const uint8_t *__restrict__ base, uint64_t id,
                                                   uint32_t d, uint32_t g) {
    const uint8_t *xv = base + id * d;
    if ((d & 15u) == 0) {   // the reference's 16-partial order over exact u8 values
        float p = 0.f;
        for (uint32_t e = g; e < d; e += 16) {
            const float t = qv[e] - (float)xv[e];
            p = __builtin_fmaf(t, t, p);
        }
        const int base_lane = (int)(threadIdx.x & 63u) & ~15;
        float r = __shfl(p, base_lane + 0, 64);
#pragma unroll
        for (int l = 1; l < 16; l++) r += __shfl(p, base_lane + l, 64);
        return r;
    }
    float r = 0.f;
    for (uint32_t e = 0; e < d; e++) {
        const float t = qv[e] - (float)xv[e];
        r = __builtin_fmaf(t, t, r);
    }
    return r;
}

// An integer in [0, 255], exactly (the u8 mirror represents it).
__device__ __forceinline__ bool is_u8_value(float v) { return v == rintf(v) && v >= 0.f && v <= 255.f; }
// v clamped to [0, 255] and truncated (the byte of an is_u8_value v).
__device__ __forceinline__ uint32_t clamp_u8(float v) { return (uint32_t)fminf(fmaxf(v, 0.f), 255.f); }

// Quad lane k's value in every lane of its quad (DPP quad_perm [k,k,k,k]).
__device__ __forceinline__ float quad_bcast(float v, int k) {
    switch (k) {
        case 1: return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x55, 0xF, 0xF, false));
        case 2: return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xAA, 0xF, 0xF, false));
        default: return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xFF, 0xF, 0xF, false));
    }
}

// Four rows of an FP32 base against an FP32 query (16-byte aligned) by one
// 16-lane group, d % 16 == 0: quad r (lanes 4r..4r+3 of the group) takes row
// id[r] (valid iff val[r]).  Lane 4r+c keeps the reference's partials
// 4c..4c+3 -- partial j sums (q[e] - x[e])^2 over e = j, j+16, j+32, ... in
// increasing e with FMAs, the order of l2_ref16_group -- from 16-byte loads
// (eight of them in flight per lane), then the quad's first lane adds the 16
// partials left to right, its neighbours' through quad DPP moves.  Returns
// row r's distance in lane 4r of the group (other lanes: unspecified).
__device__ __forceinline__ float dist4_quad_f32(const float *__restrict__ qv, const float *__restrict__ x,
                                                const uint32_t *id, const bool *val, uint32_t d, uint32_t g) {
    const uint32_t r = g >> 2, c = g & 3u;
    const uint32_t my_id = r == 0 ? id[0] : r == 1 ? id[1] : r == 2 ? id[2] : id[3];
    const bool my_v = r == 0 ? val[0] : r == 1 ? val[1] : r == 2 ? val[2] : val[3];
    float p0 = 0.f, p1 = 0.f, p2 = 0.f, p3 = 0.f;
    if (my_v) {
        const float4 *xr = reinterpret_cast<const float4 *>(x + (uint64_t)my_id * d) + c;
        const float4 *qr = reinterpret_cast<const float4 *>(qv) + c;
        const uint32_t nb = d >> 4;
#pragma unroll 8
        for (uint32_t i = 0; i < nb; i++) {
            const float4 xv = xr[4 * i];
            const float4 qq = qr[4 * i];
            float t = qq.x - xv.x;
            p0 = __builtin_fmaf(t, t, p0);
            t = qq.y - xv.y;
            p1 = __builtin_fmaf(t, t, p1);
            t = qq.z - xv.z;
            p2 = __builtin_fmaf(t, t, p2);
            t = qq.w - xv.w;
            p3 = __builtin_fmaf(t, t, p3);
        }
    }
    float s = p0;
    s += p1;
    s += p2;
    s += p3;
#pragma unroll
    for (int k = 1; k < 4; k++) {
        s += quad_bcast(p0, k);
        s += quad_bcast(p1, k);
        s += quad_bcast(p2, k);
        s += quad_bcast(p3, k);
    }
    return s;
}

// Squared L2 between two FP32 rows a and b of a base (d % 16 == 0) by the four
// lanes of a quad (c = lane & 3): the reference's 16-partial order, as
// dist4_quad_f32 (t = a - b).  ok uniform within the quad; the quad's first
// lane returns the distance.
__device__ __forceinline__ float pair_quad_f32(const float *__restrict__ x, uint32_t ia, uint32_t ib, uint32_t d,
                                               uint32_t c, bool ok) {
    float p0 = 0.f, p1 = 0.f, p2 = 0.f, p3 = 0.f;
    if (ok) {
        const float4 *ar = reinterpret_cast<const float4 *>(x + (uint64_t)ia * d) + c;
        const float4 *br = reinterpret_cast<const float4 *>(x + (uint64_t)ib * d) + c;
        const uint32_t nb = d >> 4;
#pragma unroll 6
        for (uint32_t i = 0; i < nb; i++) {
            const float4 av = ar[4 * i];
            const float4 bv = br[4 * i];
            float t = av.x - bv.x;
            p0 = __builtin_fmaf(t, t, p0);
            t = av.y - bv.y;
            p1 = __builtin_fmaf(t, t, p1);
            t = av.z - bv.z;
            p2 = __builtin_fmaf(t, t, p2);
            t = av.w - bv.w;
            p3 = __builtin_fmaf(t, t, p3);
        }
    }
    float s = p0;
    s += p1;
    s += p2;
    s += p3;
#pragma unroll
    for (int k = 1; k < 4; k++) {
        s += quad_bcast(p0, k);
        s += quad_bcast(p1, k);
        s += quad_bcast(p2, k);
        s += quad_bcast(p3, k);
    }
    return s;
}

// LDS written by some lanes of a wave, then read by others: keep the
// compiler from moving memory operations across this point (a wave's LDS
// operations execute in order).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Sum over each row of 16 lanes, every lane of the row receiving it, by DPP
// moves (no LDS round trip): swap within quads, then with the other quad of
// the 8, then with the other 8.  The order of the additions is fixed but not
// the reference's, so it serves sums that are exact in any order.
__device__ __forceinline__ int row16_sum_i(int v) {
    v += __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false);    // quad_perm [1,0,3,2]
    v += __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false);    // quad_perm [2,3,0,1]
    v += __builtin_amdgcn_update_dpp(0, v, 0x141, 0xF, 0xF, false);   // row_half_mirror
    v += __builtin_amdgcn_update_dpp(0, v, 0x140, 0xF, 0xF, false);   // row_mirror
    return v;
}
__device__ __forceinline__ float row16_sum_f(float v) {
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, false));
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x4E, 0xF, 0xF, false));
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x141, 0xF, 0xF, false));
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x140, 0xF, 0xF, false));
    return v;
}

typedef __bf16 ex_bf16x8 __attribute__((ext_vector_type(8)));
typedef float ex_f32x4 __attribute__((ext_vector_type(4)));

// Squared L2 from the LDS query to base row id by a 16-lane group, for data
// whose every sum is an exact integer (values in [0, 255], d <= 258): the
// order of the additions does not matter, so each lane reads 16 contiguous
// bytes per step and the group reduces with DPP moves.  Every lane of the
// group returns the sum.  d must be a multiple of 8 (BF16) or 4 (FP32).
template <class VT>
__device__ __forceinline__ float dist_row_exact(const float *__restrict__ qv, const VT *__restrict__ base,
                                                uint64_t id, uint32_t d, uint32_t g);
template <>
__device__ __forceinline__ float dist_row_exact<__hip_bfloat16>(const float *__restrict__ qv,
                                                                const __hip_bfloat16 *__restrict__ base,
                                                                uint64_t id, uint32_t d, uint32_t g) {
    const __hip_bfloat16 *xv = base + id * d;
    float acc = 0.f;
    for (uint32_t e = 8 * g; e < d; e += 128) {
        const ex_bf16x8 v = *reinterpret_cast<const ex_bf16x8 *>(xv + e);
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const float t = qv[e + j] - (float)v[j];
            acc = __builtin_fmaf(t, t, acc);
        }
    }
    return row16_sum_f(acc);
}
template <>
__device__ __forceinline__ float dist_row_exact<float>(const float *__restrict__ qv,
                                                       const float *__restrict__ base, uint64_t id,
                                                       uint32_t d, uint32_t g) {
    const float *xv = base + id * d;
    float acc = 0.f;
    for (uint32_t e = 4 * g; e < d; e += 64) {
        const ex_f32x4 v = *reinterpret_cast<const ex_f32x4 *>(xv + e);
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const float t = qv[e + j] - v[j];
            acc = __builtin_fmaf(t, t, acc);
        }
    }
    return row16_sum_f(acc);
}

// Row vector type of the exact-sum distance: 16 bytes per lane and step.
template <class VT>
struct ExactRow;
template <>
struct ExactRow<__hip_bfloat16> {
    typedef ex_bf16x8 V;
    static constexpr uint32_t W = 8;
};
template <>
struct ExactRow<float> {
    typedef ex_f32x4 V;
    static constexpr uint32_t W = 4;
};

// dist_row_exact for four rows at once: each 16-lane group issues the loads
// of all four rows before it accumulates, so four row reads are in flight per
// group.  id[0..3] must all be valid rows (repeat one to pad).  Every lane of
// the group returns the four sums.
template <class VT>
__device__ __forceinline__ void dist4_exact(const float *__restrict__ qv, const VT *__restrict__ base,
                                            const uint32_t *id, uint32_t d, uint32_t g, float *out) {
    typedef typename ExactRow<VT>::V V;
    constexpr uint32_t W = ExactRow<VT>::W;
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    for (uint32_t e = W * g; e < d; e += 16 * W) {
        V v[4];
#pragma unroll
        for (int r = 0; r < 4; r++) v[r] = *reinterpret_cast<const V *>(base + (uint64_t)id[r] * d + e);
#pragma unroll
        for (int r = 0; r < 4; r++) {
#pragma unroll
            for (uint32_t j = 0; j < W; j++) {
                const float t = qv[e + j] - (float)v[r][j];
                acc[r] = __builtin_fmaf(t, t, acc[r]);
            }
            // one row's arithmetic at a time: the loads above are already in
            // flight, and interleaving the rows only raises register pressure
            __builtin_amdgcn_sched_barrier(0);
        }
    }
#pragma unroll
    for (int r = 0; r < 4; r++) out[r] = row16_sum_f(acc[r]);
}

// Raw buffer resource over `bytes` bytes from base (hardware range check:
// a load past the end returns zeros).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void *base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes, 0x00020000);
}

typedef uint32_t ex_u32x4 __attribute__((ext_vector_type(4)));

// dist4_exact through a buffer resource: 32-bit byte offsets instead of
// 64-bit row addresses (fewer registers and address instructions).  Rows are
// row_bytes apart; the whole base must lie below 4 GiB of offset.
template <class VT>
__device__ __forceinline__ void dist4_exact_buf(const float *__restrict__ qv, __amdgpu_buffer_rsrc_t rs,
                                                const uint32_t *id, uint32_t row_bytes, uint32_t d,
                                                uint32_t g, float *out) {
    typedef typename ExactRow<VT>::V V;
    constexpr uint32_t W = ExactRow<VT>::W;
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    uint32_t off[4];
#pragma unroll
    for (int r = 0; r < 4; r++) off[r] = id[r] * row_bytes + W * g * (uint32_t)sizeof(VT);
    for (uint32_t e = W * g; e < d; e += 16 * W) {
        V v[4];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const ex_u32x4 raw = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)off[r], 0, 0);
            v[r] = __builtin_bit_cast(V, raw);
            off[r] += 16u * W * (uint32_t)sizeof(VT);
        }
#pragma unroll
        for (int r = 0; r < 4; r++) {
#pragma unroll
            for (uint32_t j = 0; j < W; j++) {
                const float t = qv[e + j] - (float)v[r][j];
                acc[r] = __builtin_fmaf(t, t, acc[r]);
            }
            __builtin_amdgcn_sched_barrier(0);
        }
    }
#pragma unroll
    for (int r = 0; r < 4; r++) out[r] = row16_sum_f(acc[r]);
}

typedef uint32_t ex_u32x2 __attribute__((ext_vector_type(2)));

// Squared L2 from a u8 query to R u8 rows (8-bit data, exact): each 16-lane
// group reads 8 bytes per lane and step through a buffer resource (the R
// loads of a step are issued together) and forms sum (q - x)^2 = sum q^2 +
// sum x^2 - 2 sum q x with integer dot products (v_dot4_u32_u8); every term
// is an exact integer and, for d <= 258, so is the FP32 result.  q8 holds the
// query packed 4 bytes per word (LDS), qq = sum q^2.  d must be a multiple of
// 8; rows are d bytes apart.
template <int R>
__device__ __forceinline__ void distR_u8_buf(const uint32_t *__restrict__ q8, uint32_t qq,
                                             __amdgpu_buffer_rsrc_t rs, const uint32_t *id, uint32_t d,
                                             uint32_t g, float *out) {
    int acc[R];
    uint32_t off[R];
#pragma unroll
    for (int r = 0; r < R; r++) {
        acc[r] = 0;
        off[r] = id[r] * d + 8u * g;
    }
    for (uint32_t e = 8u * g; e < d; e += 128u) {
        ex_u32x2 v[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            v[r] = __builtin_amdgcn_raw_buffer_load_b64(rs, (int)off[r], 0, 0);
            off[r] += 128u;
        }
        const uint32_t qa = q8[e >> 2], qb = q8[(e >> 2) + 1];
#pragma unroll
        for (int r = 0; r < R; r++) {
            const uint32_t xx = __builtin_amdgcn_udot4(v[r].x, v[r].x, __builtin_amdgcn_udot4(v[r].y, v[r].y, 0u, false), false);
            const uint32_t qx = __builtin_amdgcn_udot4(qa, v[r].x, __builtin_amdgcn_udot4(qb, v[r].y, 0u, false), false);
            acc[r] += (int)xx - 2 * (int)qx;
        }
    }
#pragma unroll
    for (int r = 0; r < R; r++) out[r] = (float)((int)qq + row16_sum_i(acc[r]));
}

__device__ __forceinline__ void dist4_u8_buf(const uint32_t *__restrict__ q8, uint32_t qq,
                                             __amdgpu_buffer_rsrc_t rs, const uint32_t *id, uint32_t d,
                                             uint32_t g, float *out) {
    distR_u8_buf<4>(q8, qq, rs, id, d, g, out);
}

__device__ __forceinline__ uint64_t lanemask_lt() {
    uint32_t lane = threadIdx.x & 63u;
    return lane ? (~0ull >> (64 - lane)) : 0ull;
}

// (dist, id) ordering used everywhere a ranked list is built: ascending
// distance, ties to the smaller id.
__device__ __forceinline__ bool pair_less(float da, uint32_t ia, float db, uint32_t ib) {
    return da < db || (da == db && ia < ib);
}

// Order-preserving float -> u32 map for non-negative and negative floats.
__device__ __forceinline__ uint32_t float_key(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key_float(uint32_t k) {
    uint32_t u = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
    return __uint_as_float(u);
}

// In-LDS bitonic sort of n (power of two) keys, ascending, by the whole block.
template <class T>
__device__ void bitonic_sort_lds(T *keys, uint32_t n) {
    for (uint32_t size = 2; size <= n; size <<= 1) {
        for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            __syncthreads();
            for (uint32_t t = threadIdx.x; t < (n >> 1); t += blockDim.x) {
                uint32_t lo = 2 * t - (t & (stride - 1));
                uint32_t hi = lo + stride;
                bool up = ((lo & size) == 0);
                T a = keys[lo], b = keys[hi];
                if ((a > b) == up) {
                    keys[lo] = b;
                    keys[hi] = a;
                }
            }
        }
    }
    __syncthreads();
}

// x from lane (lane ^ lj) of the wave.  Distances 1, 2 (quad permutes) and 8
// (a rotation by half a row) are DPP moves on the VALU; distance 4 takes the
// row shift whose source lane is lane ^ 4 (checked on the lane ids, so the
// shift direction need not be assumed); 16 and 32 cross rows through the
// LDS permute path (ds_bpermute).  lj must be a power of two below 64.
__device__ __forceinline__ uint32_t lane_xor32(uint32_t x, uint32_t lj) {
    const int xi = (int)x;
    if (lj == 1) return (uint32_t)__builtin_amdgcn_update_dpp(0, xi, 0xB1, 0xF, 0xF, true);
    if (lj == 2) return (uint32_t)__builtin_amdgcn_update_dpp(0, xi, 0x4E, 0xF, 0xF, true);
    if (lj == 8) return (uint32_t)__builtin_amdgcn_update_dpp(0, xi, 0x128, 0xF, 0xF, true);
    if (lj == 4) {
        const int lane = (int)(threadIdx.x & 63u);
        const int src_a = __builtin_amdgcn_update_dpp(-1, lane, 0x104, 0xF, 0xF, false);   // row_shl:4
        const int a = __builtin_amdgcn_update_dpp(0, xi, 0x104, 0xF, 0xF, true);
        const int b = __builtin_amdgcn_update_dpp(0, xi, 0x114, 0xF, 0xF, true);          // row_shr:4
        return (uint32_t)(src_a == (lane ^ 4) ? a : b);
    }
    return __shfl_xor(x, (int)lj, 64);
}
__device__ __forceinline__ unsigned long long lane_xor64(unsigned long long x, uint32_t lj) {
    const uint32_t lo = lane_xor32((uint32_t)x, lj), hi = lane_xor32((uint32_t)(x >> 32), lj);
    return ((unsigned long long)hi << 32) | lo;
}

// Ascending bitonic sort of the first n (<= 256 E) keys of `keys` (LDS) by a
// 256-thread block, E keys per thread in registers (thread t holds positions
// t E .. t E + E - 1; positions >= n sort as ~0).  Levels whose partner
// distance is below E stay in registers, below 64 E they cross lanes of one
// wave (shuffles), and only the few longer ones go through LDS, so the LDS
// traffic of the in-LDS network (one read and one write per key and level)
// falls to a handful of levels.  keys holds the sorted n keys afterwards.
template <int E>
__device__ void block_sort_u64_regs(unsigned long long *keys, uint32_t n) {
    const uint32_t t = threadIdx.x;
    unsigned long long v[E];
#pragma unroll
    for (int r = 0; r < E; r++) v[r] = (t * E + r < n) ? keys[t * E + r] : ~0ull;
#pragma unroll
    for (uint32_t k = 2; k <= 256u * E; k <<= 1) {
#pragma unroll
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            if (j < (uint32_t)E) {
#pragma unroll
                for (int r = 0; r < E; r++) {
                    const int p = r ^ (int)j;
                    if (p > r) {
                        const bool asc = ((t * E + r) & k) == 0;
                        const unsigned long long a = v[r], b = v[p];
                        const bool sw = asc ? (a > b) : (a < b);
                        v[r] = sw ? b : a;
                        v[p] = sw ? a : b;
                    }
                }
            } else {
                const bool cross_wave = j >= 64u * E;
                if (cross_wave) {
                    __syncthreads();
#pragma unroll
                    for (int r = 0; r < E; r++) keys[t * E + r] = v[r];
                    __syncthreads();
                }
#pragma unroll
                for (int r = 0; r < E; r++) {
                    const uint32_t i = t * E + r;
                    unsigned long long o;
                    if (cross_wave) {
                        o = keys[i ^ j];
                    } else {
                        const uint32_t lj = j / E;   // lane distance inside the wave
                        o = lane_xor64(v[r], lj);
                    }
                    const bool keep_min = ((i & j) == 0) == ((i & k) == 0);
                    v[r] = keep_min ? (v[r] < o ? v[r] : o) : (v[r] < o ? o : v[r]);
                }
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < E; r++)
        if (t * E + r < n) keys[t * E + r] = v[r];
    __syncthreads();
}

__device__ __forceinline__ uint32_t next_pow2_dev(uint32_t v) {
    return v <= 1u ? 1u : (1u << (32 - __clz(v - 1u)));
}

inline uint32_t next_pow2(uint32_t v) {
    uint32_t p = 1;
    while (p < v) p <<= 1;
    return p;
}

}  // namespace hrnn
