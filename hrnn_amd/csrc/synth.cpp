// synth.cpp -- deterministic synthetic datasets (host side of libhrnn_hip.so).
//
// The reference ships no generator (datasets are downloaded); these generators
// produce the synthetic workloads with a counter-based RNG, Philox4x32-10
// (Salmon, Moraes, Dror, Shaw: "Parallel random numbers: as easy as 1, 2, 3",
// SC'11), so a value depends only on (seed, stream, index) and the bytes are
// the same for any thread count.  Normals come from the Box-Muller transform.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <random>
#include <vector>

#include "../../include/hrnn_dev.h"

namespace {

struct Philox {
    uint32_t k0, k1;
    explicit Philox(uint64_t seed) : k0((uint32_t)seed), k1((uint32_t)(seed >> 32)) {}
    // 4 x u32 for counter (c0, c1, c2, c3)
    void gen(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t out[4]) const {
        const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
        const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
        uint32_t a = k0, b = k1;
        for (int r = 0; r < 10; r++) {
            uint64_t p0 = (uint64_t)M0 * c0, p1 = (uint64_t)M1 * c2;
            uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
            uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
            uint32_t n0 = hi1 ^ c1 ^ a, n1 = lo1, n2 = hi0 ^ c3 ^ b, n3 = lo0;
            c0 = n0; c1 = n1; c2 = n2; c3 = n3;
            a += W0;
            b += W1;
        }
        out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
    }
    // two doubles in (0, 1) for index idx of stream s
    void uniform2(uint64_t idx, uint32_t s, double &u1, double &u2) const {
        uint32_t o[4];
        gen((uint32_t)idx, (uint32_t)(idx >> 32), s, 0x48524E4Eu, o);
        uint64_t a = ((uint64_t)o[0] << 32) | o[1], b = ((uint64_t)o[2] << 32) | o[3];
        u1 = ((double)(a >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        u2 = ((double)(b >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    }
    void normal2(uint64_t idx, uint32_t s, double &z0, double &z1) const {
        double u1, u2;
        uniform2(idx, s, u1, u2);
        double r = std::sqrt(-2.0 * std::log(u1));
        double t = 6.283185307179586476925 * u2;
        z0 = r * std::cos(t);
        z1 = r * std::sin(t);
    }
};

enum : uint32_t { S_CENTRE = 1, S_PICK = 2, S_NOISE = 3, S_SCALE = 4 };

std::vector<double> zipf_cdf(uint32_t n, double s) {
    std::vector<double> cdf(n);
    double acc = 0;
    for (uint32_t r = 0; r < n; r++) {
        acc += std::pow((double)(r + 1), -s);
        cdf[r] = acc;
    }
    for (auto &v : cdf) v /= acc;
    return cdf;
}

uint32_t pick_centre(const Philox &pr, uint64_t i, uint32_t n_centres, const std::vector<double> *cdf) {
    double u1, u2;
    pr.uniform2(i, S_PICK, u1, u2);
    if (!cdf) {
        uint32_t c = (uint32_t)(u1 * n_centres);
        return c < n_centres ? c : n_centres - 1;
    }
    uint32_t lo = 0, hi = n_centres - 1;
    while (lo < hi) {
        uint32_t mid = (lo + hi) / 2;
        if ((*cdf)[mid] < u1) lo = mid + 1; else hi = mid;
    }
    return lo;
}

}  // namespace

extern "C" {

int hrnn_synth_mixture(uint64_t seed, uint64_t n, uint32_t d, uint32_t n_centres, float sigma,
                       float lo, float hi, int round_int, double zipf_s, uint64_t centre_seed,
                       float *out, uint32_t *centre_of) {
    if (!out || !n || !d || !n_centres || !(hi >= lo)) return HRNN_E_ARG;
    Philox pc(centre_seed), pr(seed);
    std::vector<float> mu((size_t)n_centres * d);
    for (uint32_t c = 0; c < n_centres; c++)
        for (uint32_t j = 0; j < d; j += 2) {
            double u1, u2;
            pc.uniform2((uint64_t)c * ((d + 1) / 2) + j / 2, S_CENTRE, u1, u2);
            mu[(size_t)c * d + j] = (float)(lo + (hi - lo) * u1);
            if (j + 1 < d) mu[(size_t)c * d + j + 1] = (float)(lo + (hi - lo) * u2);
        }
    std::vector<double> cdf;
    if (zipf_s > 0) cdf = zipf_cdf(n_centres, zipf_s);
    const uint64_t half = (d + 1) / 2;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; i++) {
        uint32_t c = pick_centre(pr, (uint64_t)i, n_centres, zipf_s > 0 ? &cdf : nullptr);
        if (centre_of) centre_of[i] = c;
        for (uint32_t j = 0; j < d; j += 2) {
            double z0, z1;
            pr.normal2((uint64_t)i * half + j / 2, S_NOISE, z0, z1);
            for (int t = 0; t < 2 && j + t < d; t++) {
                double v = mu[(size_t)c * d + j + t] + sigma * (t ? z1 : z0);
                if (round_int) v = std::nearbyint(v);
                if (v < lo) v = lo;
                if (v > hi) v = hi;
                out[(size_t)i * d + j + t] = (float)v;
            }
        }
    }
    return HRNN_OK;
}

int hrnn_synth_halfnormal_scaled(uint64_t seed, uint64_t scale_seed, uint64_t n, uint32_t d,
                                 float *out) {
    if (!out || !n || !d) return HRNN_E_ARG;
    Philox ps(scale_seed), pr(seed);
    std::vector<double> s(d);
    for (uint32_t j = 0; j < d; j++) {
        double u1, u2;
        ps.uniform2(j, S_SCALE, u1, u2);
        s[j] = 0.05 + 0.95 * u1;
    }
    const uint64_t half = (d + 1) / 2;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; i++)
        for (uint32_t j = 0; j < d; j += 2) {
            double z0, z1;
            pr.normal2((uint64_t)i * half + j / 2, S_NOISE, z0, z1);
            out[(size_t)i * d + j] = (float)(std::fabs(z0) * s[j]);
            if (j + 1 < d) out[(size_t)i * d + j + 1] = (float)(std::fabs(z1) * s[j + 1]);
        }
    return HRNN_OK;
}

int hrnn_synth_sphere_mixture(uint64_t seed, uint64_t centre_seed, uint64_t n, uint32_t d,
                              uint32_t n_centres, float kappa, float *out, uint32_t *centre_of) {
    if (!out || !n || !d || !n_centres || !(kappa > 0)) return HRNN_E_ARG;
    Philox pc(centre_seed), pr(seed);
    const uint64_t half = (d + 1) / 2;
    std::vector<double> mu((size_t)n_centres * d);
    for (uint32_t c = 0; c < n_centres; c++) {
        double nrm = 0;
        for (uint32_t j = 0; j < d; j += 2) {
            double z0, z1;
            pc.normal2((uint64_t)c * half + j / 2, S_CENTRE, z0, z1);
            mu[(size_t)c * d + j] = z0;
            if (j + 1 < d) mu[(size_t)c * d + j + 1] = z1;
        }
        for (uint32_t j = 0; j < d; j++) nrm += mu[(size_t)c * d + j] * mu[(size_t)c * d + j];
        nrm = std::sqrt(nrm);
        for (uint32_t j = 0; j < d; j++) mu[(size_t)c * d + j] /= nrm;
    }
    const double sc = 1.0 / std::sqrt((double)kappa);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; i++) {
        uint32_t c = pick_centre(pr, (uint64_t)i, n_centres, nullptr);
        if (centre_of) centre_of[i] = c;
        std::vector<double> v(d);
        double nrm = 0;
        for (uint32_t j = 0; j < d; j += 2) {
            double z0, z1;
            pr.normal2((uint64_t)i * half + j / 2, S_NOISE, z0, z1);
            v[j] = mu[(size_t)c * d + j] + sc * z0;
            if (j + 1 < d) v[j + 1] = mu[(size_t)c * d + j + 1] + sc * z1;
        }
        for (uint32_t j = 0; j < d; j++) nrm += v[j] * v[j];
        nrm = std::sqrt(nrm);
        for (uint32_t j = 0; j < d; j++) out[(size_t)i * d + j] = (float)(v[j] / nrm);
    }
    return HRNN_OK;
}

// The insertion order of the dynamic test: 0 .. n-1 shuffled by std::shuffle
// with std::mt19937(seed) (reference src/hrnn_dynamic_test.cpp:200-203), the
// same permutation wherever both are built against libstdc++.
int hrnn_synth_split_perm(uint64_t n, uint32_t seed, uint32_t *out) {
    if (!out || n >= 0xFFFFFFFFull) return HRNN_E_ARG;
    std::vector<uint32_t> perm(n);
    std::iota(perm.begin(), perm.end(), 0u);
    std::mt19937 rng(seed);
    std::shuffle(perm.begin(), perm.end(), rng);
    std::copy(perm.begin(), perm.end(), out);
    return HRNN_OK;
}

}  // extern "C"
