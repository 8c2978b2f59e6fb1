// utils.h -- file formats, recall metrics and timers of the HRNN drivers.
//
// Same names and formats as the reference's include/utils.h:
//   fbin            [n u32][d u32][n*d f32]                      (utils.h:31-50)
//   RkNN GT         [nq u32][(nq+1) u32 offsets][ids u32][dists f32] (utils.h:76-108,
//                   written by scripts/gen_rknn_gt.py:154-168)
//   macro recall    mean over queries with non-empty GT of |R n GT| / |GT| (utils.h:135-157)
//   kNN GT          [nq u32][k u32][nq*k ids u32][nq*k dists f32]   (utils.h:52-73)
// and the reference's `po` namespace (utils.h:20-21): Boost.Program_options
// where it is installed, else the subset in hrnn_po.h, so the reference's
// drivers compile unchanged against this directory.
#ifndef HRNN_UTILS_H
#define HRNN_UTILS_H

#include <algorithm>
#include <chrono>
#include <cstdint>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_set>
#include <vector>

#include "hnswlib/hnswlib.h"

#if __has_include(<boost/program_options.hpp>)
#include <boost/program_options.hpp>
namespace po = boost::program_options;
#else
#include "hrnn_po.h"
#endif

#ifdef _OPENMP
#include <omp.h>
#endif

struct FbinData {
    uint32_t n = 0, d = 0;
    std::vector<float> data;
};

inline FbinData load_fbin(const std::string &path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("Cannot open: " + path);
    FbinData fb;
    in.read(reinterpret_cast<char *>(&fb.n), 4);
    in.read(reinterpret_cast<char *>(&fb.d), 4);
    if (!in) throw std::runtime_error("Read error: " + path);
    fb.data.resize((size_t)fb.n * fb.d);
    in.read(reinterpret_cast<char *>(fb.data.data()), (std::streamsize)(fb.data.size() * 4));
    if (!in) throw std::runtime_error("Read error: " + path);
    return fb;
}

inline void save_fbin(const std::string &path, const float *x, uint32_t n, uint32_t d) {
    std::ofstream out(path, std::ios::binary);
    if (!out) throw std::runtime_error("Cannot open: " + path);
    out.write(reinterpret_cast<const char *>(&n), 4);
    out.write(reinterpret_cast<const char *>(&d), 4);
    out.write(reinterpret_cast<const char *>(x), (std::streamsize)((size_t)n * d * 4));
    if (!out) throw std::runtime_error("Write error: " + path);
}

struct KnnGroundtruth {
    uint32_t nq = 0, k = 0;
    std::vector<uint32_t> ids;   // [nq][k]
    std::vector<float> dists;    // [nq][k]
};

inline KnnGroundtruth load_knn_groundtruth(const std::string &path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("Cannot open: " + path);
    KnnGroundtruth gt;
    in.read(reinterpret_cast<char *>(&gt.nq), 4);
    in.read(reinterpret_cast<char *>(&gt.k), 4);
    if (!in) throw std::runtime_error("Read error: " + path);
    const size_t total = (size_t)gt.nq * gt.k;
    gt.ids.resize(total);
    gt.dists.resize(total);
    in.read(reinterpret_cast<char *>(gt.ids.data()), (std::streamsize)(total * 4));
    in.read(reinterpret_cast<char *>(gt.dists.data()), (std::streamsize)(total * 4));
    if (!in) throw std::runtime_error("Read error: " + path);
    return gt;
}

// kNN recall: hits of results[q] among the first min(k, gt.k) GT ids, over
// nq * min(k, gt.k) (utils.h:115-132).
inline double compute_knn_recall(const KnnGroundtruth &gt,
                                 const std::vector<std::vector<hnswlib::labeltype>> &results, int k) {
    const size_t kk = std::min<size_t>((size_t)std::max(k, 0), gt.k);
    size_t found = 0;
    for (size_t q = 0; q < gt.nq; q++) {
        std::unordered_set<uint32_t> s(gt.ids.begin() + q * gt.k, gt.ids.begin() + q * gt.k + kk);
        for (auto l : results[q]) found += s.count((uint32_t)l);
    }
    const size_t total = (size_t)gt.nq * kk;
    return total ? (double)found / (double)total : 0.0;
}

struct RknnGroundtruth {
    uint32_t nq = 0;
    std::vector<uint32_t> offsets;
    std::vector<uint32_t> ids;
    std::vector<float> dists;
    size_t count(size_t q) const { return offsets[q + 1] - offsets[q]; }
    const uint32_t *ids_for(size_t q) const { return ids.data() + offsets[q]; }
    const float *dists_for(size_t q) const { return dists.data() + offsets[q]; }
    size_t total() const { return offsets[nq]; }
};

inline RknnGroundtruth load_rknn_groundtruth(const std::string &path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("Cannot open: " + path);
    RknnGroundtruth gt;
    in.read(reinterpret_cast<char *>(&gt.nq), 4);
    gt.offsets.resize((size_t)gt.nq + 1);
    in.read(reinterpret_cast<char *>(gt.offsets.data()), (std::streamsize)((gt.nq + 1) * 4));
    if (!in) throw std::runtime_error("Read error: " + path);
    for (uint32_t q = 0; q < gt.nq; q++)
        if (gt.offsets[q + 1] < gt.offsets[q]) throw std::runtime_error("Bad GT offsets: " + path);
    size_t total = gt.offsets[gt.nq];
    gt.ids.resize(total);
    gt.dists.resize(total);
    in.read(reinterpret_cast<char *>(gt.ids.data()), (std::streamsize)(total * 4));
    in.read(reinterpret_cast<char *>(gt.dists.data()), (std::streamsize)(total * 4));
    if (!in) throw std::runtime_error("Read error: " + path);
    return gt;
}

inline void save_rknn_groundtruth(const std::string &path, const RknnGroundtruth &gt) {
    std::ofstream out(path, std::ios::binary);
    if (!out) throw std::runtime_error("Cannot open: " + path);
    out.write(reinterpret_cast<const char *>(&gt.nq), 4);
    out.write(reinterpret_cast<const char *>(gt.offsets.data()), (std::streamsize)(gt.offsets.size() * 4));
    out.write(reinterpret_cast<const char *>(gt.ids.data()), (std::streamsize)(gt.ids.size() * 4));
    out.write(reinterpret_cast<const char *>(gt.dists.data()), (std::streamsize)(gt.dists.size() * 4));
    if (!out) throw std::runtime_error("Write error: " + path);
}

inline double compute_rknn_recall(const RknnGroundtruth &gt,
                                  const std::vector<std::vector<hnswlib::labeltype>> &results) {
    double sum = 0;
    size_t counted = 0;
    for (size_t q = 0; q < gt.nq; q++) {
        size_t g = gt.count(q);
        if (g == 0) continue;
        std::unordered_set<uint32_t> s(gt.ids_for(q), gt.ids_for(q) + g);
        size_t f = 0;
        for (auto l : results[q]) f += s.count((uint32_t)l);
        sum += (double)f / (double)g;
        counted++;
    }
    return counted ? sum / (double)counted : 0.0;
}

// Micro recall: total hits over total GT size (src/hrnn_search.cpp:189-206).
inline double compute_rknn_micro_recall(const RknnGroundtruth &gt,
                                        const std::vector<std::vector<hnswlib::labeltype>> &results) {
    size_t found = 0, total = 0;
    for (size_t q = 0; q < gt.nq; q++) {
        size_t g = gt.count(q);
        if (g == 0) continue;
        total += g;
        std::unordered_set<uint32_t> s(gt.ids_for(q), gt.ids_for(q) + g);
        for (auto l : results[q]) found += s.count((uint32_t)l);
    }
    return total ? (double)found / (double)total : 0.0;
}

using Clock = std::chrono::high_resolution_clock;
using Ms = std::chrono::milliseconds;
using Us = std::chrono::microseconds;

static inline long elapsed_ms(Clock::time_point t0) {
    return (long)std::chrono::duration_cast<Ms>(Clock::now() - t0).count();
}
static inline long elapsed_us(Clock::time_point t0) {
    return (long)std::chrono::duration_cast<Us>(Clock::now() - t0).count();
}

#endif  // HRNN_UTILS_H
