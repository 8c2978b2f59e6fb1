// hnswlib.h -- interfaces of the HRNN host API (drop-in for the reference's
// include/hnswlib/hnswlib.h as used by the HRNN drivers).
//
// Kept alike because callers bind to them: labeltype = size_t
// (reference hnswlib.h:125), SpaceInterface (hnswlib.h:173-184),
// AlgorithmInterface (hnswlib.h:186-201), DISTFUNC (hnswlib.h:170-171),
// BaseFilterFunctor (hnswlib.h:128-132) and the POD stream helpers used by the
// index file format (hnswlib.h:160-168).
#pragma once

#include <cstddef>
#include <cstdint>
#include <iostream>
#include <queue>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace hnswlib {

typedef size_t labeltype;

class BaseFilterFunctor {
 public:
    virtual bool operator()(labeltype) { return true; }
    virtual ~BaseFilterFunctor() {}
};

template <typename MTYPE>
using DISTFUNC = MTYPE (*)(const void *, const void *, const void *);

template <typename MTYPE>
class SpaceInterface {
 public:
    virtual size_t get_data_size() = 0;
    virtual DISTFUNC<MTYPE> get_dist_func() = 0;
    virtual void *get_dist_func_param() = 0;
    virtual ~SpaceInterface() {}
};

template <typename dist_t>
class AlgorithmInterface {
 public:
    virtual void addPoint(const void *datapoint, labeltype label, bool replace_deleted = false) = 0;
    virtual std::priority_queue<std::pair<dist_t, labeltype>> searchKnn(
        const void *query, size_t k, BaseFilterFunctor *filter = nullptr) const = 0;
    virtual void saveIndex(const std::string &location) = 0;
    virtual ~AlgorithmInterface() {}

    // k nearest, closest first
    std::vector<std::pair<dist_t, labeltype>> searchKnnCloserFirst(
        const void *query, size_t k, BaseFilterFunctor *filter = nullptr) const {
        auto pq = searchKnn(query, k, filter);
        std::vector<std::pair<dist_t, labeltype>> out(pq.size());
        for (size_t i = out.size(); i-- > 0;) {
            out[i] = pq.top();
            pq.pop();
        }
        return out;
    }
};

template <typename T>
static void writeBinaryPOD(std::ostream &out, const T &v) {
    out.write(reinterpret_cast<const char *>(&v), sizeof(T));
}

template <typename T>
static void readBinaryPOD(std::istream &in, T &v) {
    in.read(reinterpret_cast<char *>(&v), sizeof(T));
}

}  // namespace hnswlib

#include "space_l2.h"
