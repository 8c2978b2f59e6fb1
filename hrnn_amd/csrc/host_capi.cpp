// host_capi.cpp -- C entry points of libhrnn_host.so over hnswlib::HRNN<float>.
#include <omp.h>

#include <cstring>
#include <exception>
#include <memory>
#include <string>

#include "hnswlib/hnswlib.h"
#include "hnswlib/hrnn.h"
#include "hrnn_host.h"

struct hrnn_index {
    std::unique_ptr<hnswlib::L2Space> space;
    std::unique_ptr<hnswlib::HRNN<float>> idx;
    std::string err;
    uint32_t d = 0;
};

namespace {
template <class F>
int guard(hrnn_index_t ix, F &&f) {
    try {
        f();
        return HRNN_OK;
    } catch (std::exception &e) {
        if (ix) ix->err = e.what();
        return HRNN_E_STATE;
    }
}
thread_local std::string g_create_err;
}  // namespace

extern "C" {

int hrnn_index_create(uint64_t max_elements, uint32_t d, uint32_t M, uint32_t efc, uint64_t seed,
                      uint32_t K, hrnn_index_t *out) {
    if (!out || !d || !max_elements) return HRNN_E_ARG;
    auto *ix = new hrnn_index();
    ix->d = d;
    int rc = guard(ix, [&] {
        ix->space.reset(new hnswlib::L2Space(d));
        ix->idx.reset(new hnswlib::HRNN<float>(ix->space.get(), max_elements, M, efc, seed, false, K));
    });
    if (rc) {
        g_create_err = ix->err;
        delete ix;
        return rc;
    }
    *out = ix;
    return HRNN_OK;
}

int hrnn_index_load(const char *path, uint32_t d, uint64_t max_elements, hrnn_index_t *out) {
    if (!out || !path || !d) return HRNN_E_ARG;
    auto *ix = new hrnn_index();
    ix->d = d;
    int rc = guard(ix, [&] {
        ix->space.reset(new hnswlib::L2Space(d));
        ix->idx.reset(new hnswlib::HRNN<float>(ix->space.get(), std::string(path), false, max_elements));
    });
    if (rc) {
        g_create_err = ix->err;
        delete ix;
        return rc;
    }
    *out = ix;
    return HRNN_OK;
}

int hrnn_index_free(hrnn_index_t ix) {
    delete ix;
    return HRNN_OK;
}

const char *hrnn_index_last_error(hrnn_index_t ix) { return ix ? ix->err.c_str() : g_create_err.c_str(); }

uint64_t hrnn_index_count(hrnn_index_t ix) { return ix ? ix->idx->getCurrentElementCount() : 0; }

int hrnn_index_info(hrnn_index_t ix, uint32_t *M, uint32_t *efc, uint32_t *K, int *maxlevel) {
    if (!ix) return HRNN_E_ARG;
    if (M) *M = (uint32_t)ix->idx->M_;
    if (efc) *efc = (uint32_t)ix->idx->ef_construction_;
    if (K) *K = (uint32_t)ix->idx->K_knng_;
    if (maxlevel) *maxlevel = ix->idx->maxlevel_;
    return HRNN_OK;
}

int hrnn_index_add_points(hrnn_index_t ix, const float *x, uint64_t n, const uint64_t *labels,
                          int nthreads) {
    if (!ix || (!x && n)) return HRNN_E_ARG;
    return guard(ix, [&] {
        uint64_t base = ix->idx->getCurrentElementCount();
        if (nthreads <= 1) {
            for (uint64_t i = 0; i < n; i++)
                ix->idx->addPoint(x + i * ix->d, labels ? labels[i] : base + i);
            return;
        }
        std::exception_ptr ep = nullptr;
#pragma omp parallel for schedule(dynamic, 512) num_threads(nthreads)
        for (int64_t i = 0; i < (int64_t)n; i++) {
            try {
                ix->idx->addPoint(x + i * ix->d, labels ? labels[i] : base + (uint64_t)i);
            } catch (...) {
#pragma omp critical
                ep = std::current_exception();
            }
        }
        if (ep) std::rethrow_exception(ep);
    });
}

int hrnn_index_add_points_gpu(hrnn_index_t ix, const float *x, uint64_t n, const uint64_t *labels,
                              int batch_permille, int seed_expansions, hrnn_hnsw_build_stats *stats) {
    if (!ix || (!x && n) || batch_permille < 0 || batch_permille > 1000 || seed_expansions > 1000) return HRNN_E_ARG;
    return guard(ix, [&] {
        ix->idx->setBuildBatchPermille(batch_permille);
        ix->idx->setBuildSeedExpansions(seed_expansions);
        std::vector<hnswlib::labeltype> lab;
        if (labels) lab.assign(labels, labels + n);
        ix->idx->addPointsDevice(x, n, labels ? lab.data() : nullptr);
        if (stats) *stats = ix->idx->lastBuildStats();
    });
}

int hrnn_index_refine(hrnn_index_t ix, int iters, int sample, hrnn_nnd_iter_stats *stats) {
    if (!ix) return HRNN_E_ARG;
    return guard(ix, [&] {
        ix->idx->refineKNNG(iters, sample);
        if (stats) {
            auto &s = ix->idx->lastRefineStats();
            for (size_t i = 0; i < s.size(); i++) stats[i] = s[i];
        }
    });
}

int hrnn_index_build_rknng(hrnn_index_t ix) {
    if (!ix) return HRNN_E_ARG;
    return guard(ix, [&] { ix->idx->buildRKNNG(); });
}

int hrnn_index_build_rknng_mutable(hrnn_index_t ix) {
    if (!ix) return HRNN_E_ARG;
    return guard(ix, [&] { ix->idx->buildRKNNG_mutable(); });
}

int hrnn_index_get_knng(hrnn_index_t ix, uint32_t *ids, float *dists, uint16_t *cnt) {
    if (!ix) return HRNN_E_ARG;
    return guard(ix, [&] {
        size_t n = ix->idx->getCurrentElementCount(), K = ix->idx->K_knng_;
        if (ix->idx->serve_mode_) throw std::runtime_error("get_knng: serve mode");
        if (ids) std::memcpy(ids, ix->idx->knnIds().data(), n * K * 4);
        if (dists) std::memcpy(dists, ix->idx->knnDists().data(), n * K * 4);
        if (cnt) std::memcpy(cnt, ix->idx->knnCounts().data(), n * 2);
    });
}

int hrnn_index_set_knng(hrnn_index_t ix, const uint32_t *ids, const float *dists, const uint16_t *cnt) {
    if (!ix || !ids || !dists || !cnt) return HRNN_E_ARG;
    return guard(ix, [&] { ix->idx->setKnng(ids, dists, cnt); });
}

int hrnn_index_get_rknng(hrnn_index_t ix, uint32_t *rank_bits, int *wide, uint64_t *n_entries,
                         uint64_t *offsets, void *entries) {
    if (!ix) return HRNN_E_ARG;
    return guard(ix, [&] {
        auto &I = *ix->idx;
        if (!I.rknng_built_) throw std::runtime_error("get_rknng: RKNNG not built");
        if (rank_bits) *rank_bits = I.rknng_rank_bits_;
        if (wide) *wide = I.rknng_wide_ ? 1 : 0;
        uint64_t ne = I.rknng_offsets_.back();
        if (n_entries) *n_entries = ne;
        if (offsets) std::memcpy(offsets, I.rknng_offsets_.data(), I.rknng_offsets_.size() * 8);
        if (entries) {
            if (I.rknng_wide_) std::memcpy(entries, I.rknng_entries_64_.data(), ne * 8);
            else std::memcpy(entries, I.rknng_entries_32_.data(), ne * 4);
        }
    });
}

int hrnn_index_proxies(hrnn_index_t ix, const float *q, uint64_t nq, uint32_t m, uint32_t ef,
                       int nthreads, uint32_t *out) {
    if (!ix || !q || !out) return HRNN_E_ARG;
    return guard(ix, [&] {
        auto p = ix->idx->proxies(q, nq, m, ef, nthreads);
        std::memcpy(out, p.data(), p.size() * 4);
    });
}

int hrnn_index_query_features(hrnn_index_t ix, const float *q, uint64_t nq, uint32_t k, uint32_t ef,
                              uint32_t m_max, int nthreads, float *feats, uint32_t *proxies, float *raw) {
    if (!ix || !q || !feats || !proxies) return HRNN_E_ARG;
    return guard(ix, [&] {
        if (m_max < 16) throw std::runtime_error("query_features: m_max must be >= 16");
        std::vector<float> r;
        float *rp = raw;
        if (!rp) {
            r.resize(nq * 20);
            rp = r.data();
        }
        ix->idx->proxiesWithStats(q, nq, m_max, ef, nthreads, proxies, rp);
        ix->idx->queryFeatures(rp, proxies, nq, m_max, k, feats);
    });
}

int hrnn_index_proxies_gpu(hrnn_index_t ix, const float *q, uint64_t nq, uint32_t m, uint32_t ef,
                           uint32_t *out) {
    if (!ix || !q || !out) return HRNN_E_ARG;
    return guard(ix, [&] {
        auto p = ix->idx->proxiesDevice(q, nq, m, ef);
        std::memcpy(out, p.data(), p.size() * 4);
    });
}

int hrnn_index_search(hrnn_index_t ix, const float *q, uint64_t nq, uint32_t m, float thr,
                      uint32_t k, uint32_t Kp, uint32_t ef, int nthreads, int per_query,
                      uint64_t *offsets, uint64_t *labels, uint64_t cap, uint64_t *needed) {
    if (!ix || !q || !offsets) return HRNN_E_ARG;
    int rc = guard(ix, [&] {
        std::vector<std::vector<hnswlib::labeltype>> res;
        if (per_query) {
            res.resize(nq);
            for (uint64_t i = 0; i < nq; i++)
                res[i] = ix->idx->searchRknn(q + i * ix->d, m, thr, k, Kp, ef);
        } else {
            res = ix->idx->searchRknnBatch(q, nq, m, thr, k, Kp, ef, nthreads);
        }
        uint64_t t = 0;
        offsets[0] = 0;
        for (uint64_t i = 0; i < nq; i++) {
            for (auto l : res[i]) {
                if (t < cap && labels) labels[t] = l;
                t++;
            }
            offsets[i + 1] = t;
        }
        if (needed) *needed = t;
        if (t > cap) throw std::runtime_error("search: result capacity too small");
    });
    if (rc && needed && *needed > cap) return HRNN_E_CAP;
    return rc;
}

int hrnn_index_search_proxies(hrnn_index_t ix, const float *q, uint64_t nq, const uint32_t *prox,
                              uint32_t m, const uint32_t *k_per_q, uint32_t k, uint32_t Kp,
                              float thr, uint64_t *offsets, uint32_t *ids, uint64_t cap,
                              uint64_t *needed) {
    if (!ix || !q || !prox || !offsets) return HRNN_E_ARG;
    uint64_t total = 0;
    int rc = guard(ix, [&] {
        std::vector<uint64_t> off;
        std::vector<uint32_t> v;
        ix->idx->searchWithProxies(q, nq, prox, m, k_per_q, k, Kp, thr, off, v);
        std::memcpy(offsets, off.data(), off.size() * 8);
        total = v.size();
        if (needed) *needed = total;
        if (total > cap) throw std::runtime_error("search_proxies: result capacity too small");
        if (total && ids) std::memcpy(ids, v.data(), total * 4);
    });
    if (rc && total > cap) return HRNN_E_CAP;
    return rc;
}

int hrnn_index_insert_dynamic(hrnn_index_t ix, const float *x, uint64_t n, const uint64_t *labels,
                              uint32_t m_update, int bidir, int nthreads) {
    if (!ix || !x || !labels) return HRNN_E_ARG;
    return guard(ix, [&] {
        std::exception_ptr ep = nullptr;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads > 0 ? nthreads : 1)
        for (int64_t i = 0; i < (int64_t)n; i++) {
            try {
                ix->idx->insertPointDynamic(x + i * ix->d, labels[i], m_update, bidir != 0);
            } catch (...) {
#pragma omp critical
                ep = std::current_exception();
            }
        }
        if (ep) std::rethrow_exception(ep);
    });
}

int hrnn_index_insert_dynamic_gpu(hrnn_index_t ix, const float *x, uint64_t n, const uint64_t *labels,
                                  uint32_t m_update, int bidir, int batch_permille, hrnn_insert_stats *stats) {
    if (!ix || !x || !labels || batch_permille < 0 || batch_permille > 1000) return HRNN_E_ARG;
    return guard(ix, [&] {
        ix->idx->setBuildBatchPermille(batch_permille);
        std::vector<hnswlib::labeltype> lab(labels, labels + n);
        ix->idx->insertPointsDynamicDevice(x, n, lab.data(), m_update, bidir != 0);
        if (stats) *stats = ix->idx->lastInsertStats();
    });
}

int hrnn_index_get_rknng_mutable(hrnn_index_t ix, uint64_t *off, uint32_t *ids) {
    if (!ix || !off) return HRNN_E_ARG;
    return guard(ix, [&] {
        auto &idx = *ix->idx;
        if (!idx.rknng_mutable_) throw std::runtime_error("get_rknng_mutable: call buildRKNNG_mutable first");
        const size_t n = idx.getCurrentElementCount();
        off[0] = 0;
        for (size_t i = 0; i < n; i++) off[i + 1] = off[i] + idx.rknng_lists_[i].size();
        if (!ids) return;
        for (size_t i = 0; i < n; i++)
            std::memcpy(ids + off[i], idx.rknng_lists_[i].data(), idx.rknng_lists_[i].size() * 4);
    });
}

int hrnn_index_compact_for_serving(hrnn_index_t ix, uint32_t max_k) {
    if (!ix) return HRNN_E_ARG;
    return guard(ix, [&] { ix->idx->compactForServing(max_k); });
}

int hrnn_index_save(hrnn_index_t ix, const char *path) {
    if (!ix || !path) return HRNN_E_ARG;
    return guard(ix, [&] { ix->idx->saveIndex(path); });
}

int hrnn_index_device(hrnn_index_t ix, hrnn_dev_t *out) {
    if (!ix || !out) return HRNN_E_ARG;
    return guard(ix, [&] {
        *out = ix->idx->device();
        ix->idx->syncDevice();
    });
}

int hrnn_index_get_labels(hrnn_index_t ix, uint64_t *labels) {
    if (!ix || !labels) return HRNN_E_ARG;
    return guard(ix, [&] {
        size_t n = ix->idx->getCurrentElementCount();
        for (size_t i = 0; i < n; i++) labels[i] = ix->idx->getExternalLabel((hnswlib::tableint)i);
    });
}

int hrnn_index_export_hnsw(hrnn_index_t ix, uint32_t *l0, int *levels, uint32_t *up_off,
                           uint32_t *up_pool, uint64_t *pool_words, uint32_t *entry, int *maxlevel) {
    if (!ix || !pool_words) return HRNN_E_ARG;
    return guard(ix, [&] {
        std::vector<uint32_t> off, pool;
        std::vector<int> lv;
        ix->idx->exportUpper(off, pool, lv);
        *pool_words = pool.size();
        if (entry) *entry = ix->idx->entryPoint();
        if (maxlevel) *maxlevel = ix->idx->maxLevel();
        if (!up_pool) return;
        size_t n = ix->idx->getCurrentElementCount();
        const size_t w0 = ix->idx->maxM0_ + 1;
        if (l0)
            for (size_t i = 0; i < n; i++) std::memcpy(l0 + i * w0, ix->idx->level0Links((hnswlib::tableint)i), w0 * 4);
        if (levels) std::copy(lv.begin(), lv.end(), levels);
        if (up_off) std::copy(off.begin(), off.end(), up_off);
        std::copy(pool.begin(), pool.end(), up_pool);
    });
}

int hrnn_index_get_level0(hrnn_index_t ix, uint32_t *links, uint32_t words, int *levels) {
    if (!ix) return HRNN_E_ARG;
    return guard(ix, [&] {
        size_t n = ix->idx->getCurrentElementCount();
        uint32_t need = (uint32_t)ix->idx->maxM0_ + 1;
        if (words < need) throw std::runtime_error("get_level0: words_per_node too small");
        for (size_t i = 0; i < n; i++) {
            const uint32_t *ll = ix->idx->level0Links((hnswlib::tableint)i);
            uint32_t c = ll[0] & 0xFFFFu;
            links[i * words] = c;
            for (uint32_t t = 0; t < words - 1; t++) links[i * words + 1 + t] = t < c ? ll[1 + t] : 0xFFFFFFFFu;
        }
        if (levels) {
            auto lv = ix->idx->levels();
            for (size_t i = 0; i < n; i++) levels[i] = lv[i];
        }
    });
}

}  // extern "C"
