// ref_harness.cpp -- C entry points around the reference's own HRNN<float>
// (include/hnswlib/hrnn.h of the reference repository), compiled from the
// reference sources where they lie by oracle/Makefile into oracle/_ref/.
//
// TEST INFRASTRUCTURE ONLY: the compiled reference pins the oracle
// restatement (hrnn_oracle.c) and serves as the CPU baseline of bench.py.
// Only the reference's public members are used; nothing here changes what
// the reference computes.
#include "hnswlib/hnswlib.h"
#include "hnswlib/hrnn.h"

#include <omp.h>
#include <cstdint>
#include <cstring>
#include <exception>
#include <memory>
#include <string>
#include <vector>

namespace {
struct RefIndex {
    std::unique_ptr<hnswlib::L2Space> space;
    std::unique_ptr<hnswlib::HRNN<float>> idx;
    uint32_t d = 0;
};
thread_local std::string g_err;

template <class F>
int guarded(F &&f) {
    try {
        f();
        return 0;
    } catch (std::exception &e) {
        g_err = e.what();
        return -1;
    }
}
}  // namespace

extern "C" {

const char *ref_last_error() { return g_err.c_str(); }

void *ref_new(uint64_t max_elements, uint32_t d, uint32_t M, uint32_t efc, uint32_t K,
              uint64_t seed) {
    auto *r = new RefIndex();
    r->d = d;
    r->space.reset(new hnswlib::L2Space(d));
    r->idx.reset(new hnswlib::HRNN<float>(r->space.get(), max_elements, M, efc, seed, false, K));
    return r;
}

void ref_free(void *h) { delete static_cast<RefIndex *>(h); }

float ref_l2(void *h, const float *a, const float *b) {
    auto *r = static_cast<RefIndex *>(h);
    return r->idx->fstdistfunc_(a, b, r->idx->dist_func_param_);
}

// addPoint for rows [0, n) with labels (or i); nthreads > 1 uses the
// driver's OpenMP schedule (hrnn_build.cpp:75).
int ref_add_points(void *h, const float *x, uint64_t n, int nthreads, const uint64_t *labels) {
    auto *r = static_cast<RefIndex *>(h);
    return guarded([&] {
        if (nthreads <= 1) {
            for (uint64_t i = 0; i < n; i++)
                r->idx->addPoint(x + i * r->d, labels ? labels[i] : i);
        } else {
            omp_set_num_threads(nthreads);
            std::exception_ptr ep = nullptr;
#pragma omp parallel for schedule(dynamic, 512)
            for (int64_t i = 0; i < (int64_t)n; i++) {
                try {
                    r->idx->addPoint(x + i * r->d, labels ? labels[i] : (uint64_t)i);
                } catch (...) {
#pragma omp critical
                    ep = std::current_exception();
                }
            }
            if (ep) std::rethrow_exception(ep);
        }
    });
}

// Place vectors without building HNSW (for refine / CSR runs on a given graph).
int ref_set_raw(void *h, const float *x, uint64_t n) {
    auto *r = static_cast<RefIndex *>(h);
    return guarded([&] {
        if (n > r->idx->max_elements_) throw std::runtime_error("n > max_elements");
        for (uint64_t i = 0; i < n; i++) {
            std::memset(r->idx->data_level0_memory_ + i * r->idx->size_data_per_element_, 0,
                        r->idx->size_data_per_element_);
            std::memcpy(r->idx->getDataByInternalId((hnswlib::tableint)i), x + i * r->d,
                        r->d * sizeof(float));
            r->idx->setExternalLabel((hnswlib::tableint)i, i);
        }
        r->idx->cur_element_count = n;
    });
}

uint64_t ref_count(void *h) { return static_cast<RefIndex *>(h)->idx->cur_element_count; }

uint32_t ref_knn_depth(void *h) { return (uint32_t)static_cast<RefIndex *>(h)->idx->K_knng_; }

int ref_set_knng(void *h, const uint32_t *ids, const float *dists, const uint16_t *cnt) {
    auto *r = static_cast<RefIndex *>(h);
    return guarded([&] {
        size_t n = r->idx->cur_element_count, K = r->idx->K_knng_;
        for (size_t i = 0; i < n; i++) {
            auto *ll = r->idx->get_knng_linklist((hnswlib::tableint)i);
            char *base = (char *)(ll + 1);
            for (size_t j = 0; j < cnt[i]; j++) {
                std::memcpy(base + j * 8, ids + i * K + j, 4);
                std::memcpy(base + j * 8 + 4, dists + i * K + j, 4);
            }
            r->idx->setListCount(ll, cnt[i]);
        }
    });
}

int ref_get_knng(void *h, uint32_t *ids, float *dists, uint16_t *cnt) {
    auto *r = static_cast<RefIndex *>(h);
    return guarded([&] {
        size_t n = r->idx->cur_element_count, K = r->idx->K_knng_;
        for (size_t i = 0; i < n; i++) {
            unsigned c = r->idx->getListCount(r->idx->get_knng_linklist((hnswlib::tableint)i));
            cnt[i] = (uint16_t)c;
            for (size_t j = 0; j < K; j++) {
                ids[i * K + j] = j < c ? r->idx->get_knng_neighbor((hnswlib::tableint)i, (unsigned)j) : 0xFFFFFFFFu;
                dists[i * K + j] = j < c ? r->idx->get_knng_dist((hnswlib::tableint)i, (unsigned)j) : 0.f;
            }
        }
    });
}

int ref_refine(void *h, int iters, int sample, int nthreads) {
    auto *r = static_cast<RefIndex *>(h);
    return guarded([&] {
        omp_set_num_threads(nthreads > 0 ? nthreads : 1);
        r->idx->refineKNNG(iters, sample);
    });
}

int ref_mark_knng_built(void *h) {
    static_cast<RefIndex *>(h)->idx->knng_built_ = true;
    return 0;
}

int ref_build_rknng(void *h) {
    auto *r = static_cast<RefIndex *>(h);
    return guarded([&] { r->idx->buildRKNNG(); });
}

int ref_csr_info(void *h, uint32_t *rank_bits, int *wide, uint64_t *n_entries) {
    auto *r = static_cast<RefIndex *>(h);
    *rank_bits = r->idx->rknng_rank_bits_;
    *wide = r->idx->rknng_wide_ ? 1 : 0;
    *n_entries = r->idx->rknng_offsets_.empty() ? 0 : r->idx->rknng_offsets_.back();
    return r->idx->rknng_built_ ? 0 : -1;
}

int ref_get_csr(void *h, uint64_t *offsets, void *entries) {
    auto *r = static_cast<RefIndex *>(h);
    auto &off = r->idx->rknng_offsets_;
    std::memcpy(offsets, off.data(), off.size() * sizeof(uint64_t));
    if (r->idx->rknng_wide_)
        std::memcpy(entries, r->idx->rknng_entries_64_.data(), r->idx->rknng_entries_64_.size() * 8);
    else
        std::memcpy(entries, r->idx->rknng_entries_32_.data(), r->idx->rknng_entries_32_.size() * 4);
    return 0;
}

// Full searchRknn per query; results are labels, concatenated with offsets.
int ref_search(void *h, const float *q, uint64_t nq, uint32_t m, float threshold, uint32_t k,
               uint32_t K_prime, uint32_t ef, int nthreads, uint64_t *off, uint64_t *labels,
               uint64_t cap, uint64_t *needed) {
    auto *r = static_cast<RefIndex *>(h);
    return guarded([&] {
        std::vector<std::vector<hnswlib::labeltype>> res(nq);
        omp_set_num_threads(nthreads > 0 ? nthreads : 1);
#pragma omp parallel for schedule(dynamic, 4) if (nthreads > 1)
        for (int64_t i = 0; i < (int64_t)nq; i++)
            res[i] = r->idx->searchRknn(q + i * r->d, m, threshold, k, K_prime, ef);
        uint64_t t = 0;
        off[0] = 0;
        for (uint64_t i = 0; i < nq; i++) {
            for (auto l : res[i]) {
                if (t < cap) labels[t] = l;
                t++;
            }
            off[i + 1] = t;
        }
        *needed = t;
        if (t > cap) throw std::runtime_error("cap");
    });
}

// Phase A alone: greedy descent over the upper layers, then the layer-0
// beam of width max(ef, m) via the reference's searchBaseLayerST<true>,
// keeping the m closest (the first steps of searchRknn_impl).
int ref_proxies(void *h, const float *q, uint64_t nq, uint32_t m, uint32_t ef, uint32_t *out) {
    auto *r = static_cast<RefIndex *>(h);
    auto &ix = *r->idx;
    return guarded([&] {
        for (uint64_t qi = 0; qi < nq; qi++) {
            const float *qv = q + qi * r->d;
            hnswlib::tableint cur = ix.enterpoint_node_;
            float cd = ix.fstdistfunc_(qv, ix.getDataByInternalId(cur), ix.dist_func_param_);
            for (int level = ix.maxlevel_; level > 0; level--) {
                bool moved = true;
                while (moved) {
                    moved = false;
                    unsigned *ll = (unsigned *)ix.get_linklist(cur, level);
                    int sz = ix.getListCount(ll);
                    hnswlib::tableint *nb = (hnswlib::tableint *)(ll + 1);
                    for (int t = 0; t < sz; t++) {
                        float dd = ix.fstdistfunc_(qv, ix.getDataByInternalId(nb[t]), ix.dist_func_param_);
                        if (dd < cd) { cd = dd; cur = nb[t]; moved = true; }
                    }
                }
            }
            auto top = ix.searchBaseLayerST<true>(cur, qv, std::max<size_t>(ef, m));
            while (top.size() > m) top.pop();
            uint32_t c = 0;
            std::vector<uint32_t> tmp;
            while (!top.empty()) { tmp.push_back(top.top().second); top.pop(); }
            for (auto t : tmp) out[qi * m + c++] = t;
            for (; c < m; c++) out[qi * m + c] = 0xFFFFFFFFu;
        }
    });
}

int ref_save(void *h, const char *path) {
    auto *r = static_cast<RefIndex *>(h);
    return guarded([&] { r->idx->saveIndex(path); });
}

void *ref_load(const char *path, uint32_t d, uint64_t max_elements) {
    auto *r = new RefIndex();
    r->d = d;
    r->space.reset(new hnswlib::L2Space(d));
    try {
        r->idx.reset(new hnswlib::HRNN<float>(r->space.get(), std::string(path), false, max_elements));
    } catch (std::exception &e) {
        g_err = e.what();
        delete r;
        return nullptr;
    }
    return r;
}

int ref_build_rknng_mutable(void *h) {
    auto *r = static_cast<RefIndex *>(h);
    return guarded([&] { r->idx->buildRKNNG_mutable(); });
}

int ref_insert_dynamic(void *h, const float *x, uint64_t n, const uint64_t *labels,
                       uint32_t m_update, int bidir, int nthreads) {
    auto *r = static_cast<RefIndex *>(h);
    return guarded([&] {
        omp_set_num_threads(nthreads > 0 ? nthreads : 1);
        std::exception_ptr ep = nullptr;
#pragma omp parallel for schedule(dynamic, 64) if (nthreads > 1)
        for (int64_t i = 0; i < (int64_t)n; i++) {
            try {
                r->idx->insertPointDynamic(x + i * r->d, labels[i], m_update, bidir != 0);
            } catch (...) {
#pragma omp critical
                ep = std::current_exception();
            }
        }
        if (ep) std::rethrow_exception(ep);
    });
}

// The mutable reverse lists (rknng_lists_) flattened: off[n+1] always, the
// ids (in each list's own order) when ids is not null.
int ref_get_rknng_mutable(void *h, uint64_t *off, uint32_t *ids) {
    auto *r = static_cast<RefIndex *>(h);
    return guarded([&] {
        auto &ix = *r->idx;
        const size_t n = ix.cur_element_count;
        off[0] = 0;
        for (size_t i = 0; i < n; i++)
            off[i + 1] = off[i] + (i < ix.rknng_lists_.size() ? ix.rknng_lists_[i].size() : 0);
        if (!ids) return;
        for (size_t i = 0; i < n && i < ix.rknng_lists_.size(); i++)
            std::memcpy(ids + off[i], ix.rknng_lists_[i].data(), ix.rknng_lists_[i].size() * 4);
    });
}

// HNSW layer-0 adjacency for structure checks: out[i*maxM0 ..] links, cnt[i].
int ref_get_level0(void *h, uint32_t *links, uint32_t *cnt, uint32_t maxM0) {
    auto *r = static_cast<RefIndex *>(h);
    auto &ix = *r->idx;
    size_t n = ix.cur_element_count;
    for (size_t i = 0; i < n; i++) {
        auto *ll = ix.get_linklist0((hnswlib::tableint)i);
        unsigned c = ix.getListCount(ll);
        cnt[i] = c;
        auto *nb = (hnswlib::tableint *)(ll + 1);
        for (unsigned j = 0; j < maxM0; j++) links[i * maxM0 + j] = j < c ? nb[j] : 0xFFFFFFFFu;
    }
    return 0;
}

int ref_get_labels(void *h, uint64_t *labels) {
    auto *r = static_cast<RefIndex *>(h);
    for (size_t i = 0; i < r->idx->cur_element_count; i++)
        labels[i] = r->idx->getExternalLabel((hnswlib::tableint)i);
    return 0;
}

int ref_levels(void *h, int *levels, int *maxlevel, uint32_t *entry) {
    auto *r = static_cast<RefIndex *>(h);
    auto &ix = *r->idx;
    for (size_t i = 0; i < ix.cur_element_count; i++) levels[i] = ix.element_levels_[i];
    *maxlevel = ix.maxlevel_;
    *entry = ix.enterpoint_node_;
    return 0;
}

}  // extern "C"
