// space_l2.h -- squared L2 for the HRNN host path.
//
// Semantics follow the reference's L2Space (include/hnswlib/space_l2.h:6-54,
// 208-253): FP32 squared distance by direct difference.  When d % 16 == 0 the
// sum is formed in 16 partials (element e into partial e % 16, fused
// multiply-add) that are then added left to right -- the order the
// reference's AVX-512 build produces -- so host, GPU and reference agree bit
// for bit.  With AVX2 the 16 partials are two 8-wide accumulators; otherwise
// a scalar loop keeps the same order.  d % 16 != 0 uses one sequential sum.
#pragma once

#include <cmath>
#include <cstddef>
#if defined(__AVX2__) && defined(__FMA__)
#include <immintrin.h>
#endif

namespace hnswlib {

static inline float l2sqr_seq(const float *a, const float *b, size_t d) {
    float r = 0.f;
    for (size_t i = 0; i < d; i++) {
        float t = a[i] - b[i];
        r = std::fma(t, t, r);
    }
    return r;
}

static inline float l2sqr_16lane(const float *a, const float *b, size_t d) {
#if defined(__AVX512F__)
    __m512 acc = _mm512_setzero_ps();
    for (size_t e = 0; e < d; e += 16) {
        __m512 t = _mm512_sub_ps(_mm512_loadu_ps(a + e), _mm512_loadu_ps(b + e));
        acc = _mm512_fmadd_ps(t, t, acc);
    }
    alignas(64) float p[16];
    _mm512_store_ps(p, acc);
#elif defined(__AVX2__) && defined(__FMA__)
    __m256 lo = _mm256_setzero_ps(), hi = _mm256_setzero_ps();
    for (size_t e = 0; e < d; e += 16) {
        __m256 t0 = _mm256_sub_ps(_mm256_loadu_ps(a + e), _mm256_loadu_ps(b + e));
        __m256 t1 = _mm256_sub_ps(_mm256_loadu_ps(a + e + 8), _mm256_loadu_ps(b + e + 8));
        lo = _mm256_fmadd_ps(t0, t0, lo);
        hi = _mm256_fmadd_ps(t1, t1, hi);
    }
    alignas(32) float p[16];
    _mm256_store_ps(p, lo);
    _mm256_store_ps(p + 8, hi);
#else
    float p[16] = {0};
    for (size_t e = 0; e < d; e += 16)
        for (int l = 0; l < 16; l++) {
            float t = a[e + l] - b[e + l];
            p[l] = std::fma(t, t, p[l]);
        }
#endif
    float r = p[0];
    for (int l = 1; l < 16; l++) r += p[l];
    return r;
}

static float L2Sqr(const void *pa, const void *pb, const void *param) {
    size_t d = *static_cast<const size_t *>(param);
    const float *a = static_cast<const float *>(pa), *b = static_cast<const float *>(pb);
    return (d % 16 == 0) ? l2sqr_16lane(a, b, d) : l2sqr_seq(a, b, d);
}

class L2Space : public SpaceInterface<float> {
    size_t dim_;
    size_t data_size_;

 public:
    explicit L2Space(size_t dim) : dim_(dim), data_size_(dim * sizeof(float)) {}
    size_t get_data_size() override { return data_size_; }
    DISTFUNC<float> get_dist_func() override { return L2Sqr; }
    void *get_dist_func_param() override { return &dim_; }
    ~L2Space() override {}
};

}  // namespace hnswlib
