// beam_core.h -- building blocks shared by the GPU beams over the HNSW graph:
// the query beam of phase A (beam.hip) and the construction beams
// (hnsw_build.hip).  One wave (64 lanes) per beam.
#pragma once
#include <cfloat>
#include <type_traits>

#include "handle.h"

namespace hrnn {

constexpr int kBeamMaxEf = 512;
constexpr int kBeamMaxDeg = 64;
constexpr uint32_t kVisProbes = 32;

// Insert id into the LDS set of 2^log_slots ids; true when the id was not
// there (or could not be recorded: *lost is set then).  Triangular probing
// (offsets 0, 1, 3, 6, ...) visits every slot of a power-of-two table and
// avoids the long runs of linear probing.
__device__ __forceinline__ bool vis_insert(uint32_t *tab, uint32_t log_slots, uint32_t id,
                                           uint32_t *lost) {
    const uint32_t mask = (1u << log_slots) - 1u;
    uint32_t h = (id * 2654435761u) >> (32u - log_slots);
    for (uint32_t p = 0; p < kVisProbes; p++) {
        const uint32_t old = atomicCAS(&tab[h], kInvalidId, id);
        if (old == kInvalidId) return true;
        if (old == id) return false;
        h = (h + p + 1u) & mask;
    }
    *lost = 1u;
    return true;
}

// Distances of R rows per 16-lane group (4 R rows per wave; row r valid iff
// val[r], val uniform within the group).  Exact-sum data issues the R row
// reads together; otherwise the reference's 16-partial order, one row at a
// time.  R > 4 only for u8 storage.
template <class VT, int R = 4>
__device__ __forceinline__ void beam_dist4(bool exact, const float *qv, const uint32_t *q8, uint32_t qq,
                                           const VT *x, __amdgpu_buffer_rsrc_t xrs, const uint32_t *id,
                                           const bool *val, uint32_t d, uint32_t g, float *o) {
    if (exact) {   // exact-sum data; the host allows it only for a base below 4 GiB
        if constexpr (std::is_same<VT, uint8_t>::value) {
            distR_u8_buf<R>(q8, qq, xrs, id, d, g, o);
        } else {
            static_assert(R == 4, "wide row groups need u8 storage");
            dist4_exact_buf<VT>(qv, xrs, id, d * (uint32_t)sizeof(VT), d, g, o);
        }
    } else {
#pragma unroll
        for (int r = 0; r < R; r++) o[r] = val[r] ? dist_row<VT>(qv, x, id[r], d, g) : FLT_MAX;
    }
}

__device__ __forceinline__ unsigned long long beam_key(float d, uint32_t id) {
    return ((unsigned long long)float_key(d) << 32) | id;
}

// Ascending bitonic sort of one 64-bit key per lane over lanes [0, W) (W a
// power of two <= 64, uniform); lanes >= W keep their key.
__device__ __forceinline__ unsigned long long wave_sort64(unsigned long long key, uint32_t lane,
                                                          uint32_t W) {
    for (uint32_t k = 2; k <= W; k <<= 1) {
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

// A value the whole wave holds alike, moved to a scalar register (so the
// compiler keeps it, and what is computed from it, out of vector registers).
__device__ __forceinline__ uint32_t uni(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }
__device__ __forceinline__ float uni_f(float v) { return __uint_as_float(uni(__float_as_uint(v))); }
__device__ __forceinline__ unsigned long long uni64(unsigned long long v) {
    return ((unsigned long long)uni((uint32_t)(v >> 32)) << 32) | uni((uint32_t)v);
}

__device__ __forceinline__ unsigned long long shfl_xor64(unsigned long long v, int m) {
    const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, m, 64);
    const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), m, 64);
    return ((unsigned long long)hi << 32) | lo;
}

// Lane j's 64-bit value, j uniform.
__device__ __forceinline__ unsigned long long read_lane64(unsigned long long v, uint32_t j) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, (int)j);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), (int)j);
    return ((unsigned long long)hi << 32) | lo;
}

// Layer-0 row of node c, issued as vector loads that nothing waits on here:
// lane j receives word j of the row in *w (lane 0: the count word) and
// neighbour j in *nb.  l0_count() reads the count where it is used, so a row
// loaded ahead of its use does not stall the work in between (a uniform read
// of the count word would be a scalar load, whose wait also holds every LDS
// access behind it).
__device__ __forceinline__ void l0_row_issue(const uint32_t *__restrict__ l0, uint32_t l0_words, uint32_t c,
                                             uint32_t lane, uint32_t *w, uint32_t *nb) {
    const uint32_t *ll = l0 + (uint64_t)c * l0_words;
    *w = (lane < l0_words) ? ll[lane] : 0u;
    *nb = (lane + 1u < l0_words) ? ll[1 + lane] : 0u;
}
__device__ __forceinline__ uint32_t l0_count(uint32_t w, uint32_t l0_words) {
    return min((uint32_t)__builtin_amdgcn_readlane((int)w, 0) & 0xFFFFu, l0_words - 1u);
}

// LDS layout of a beam workgroup (dynamic part), in bytes: query vector, the query
// as packed bytes, visited set, list keys [2][ef_cap], expanded flags
// [2][ef_cap].
struct BeamLds {
    uint32_t qv_bytes, q8_off, vis_off, k_off, x_off, total;
    __host__ __device__ BeamLds(uint32_t d, uint32_t ef_cap, uint32_t log_vis) {
        qv_bytes = (d * 4u + 15u) & ~15u;
        q8_off = qv_bytes;
        vis_off = q8_off + ((d + 15u) & ~15u);
        k_off = vis_off + (4u << log_vis);
        x_off = k_off + 2u * ef_cap * 8u;
        total = x_off + 2u * ef_cap;
    }
};

}  // namespace hrnn
