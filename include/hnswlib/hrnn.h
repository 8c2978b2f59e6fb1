// hrnn.h -- hnswlib::HRNN<float>, the host API of the MI355X HRNN index.
//
// Drop-in for the reference's include/hnswlib/hrnn.h as used by the
// hrnn_build / hrnn_search / hrnn_dynamic_test drivers: same class name,
// constructors, methods, public fields and index file format.  The
// data-parallel stages run on the GPU through the C-ABI of libhrnn_hip.so
// (include/hrnn_dev.h):
//   refineKNNG  -> hrnn_dev_nndescent      (reference hrnn.h:1725-1862)
//   buildRKNNG  -> hrnn_dev_build_rknn_csr (reference hrnn.h:1869-1935)
//   searchRknn  -> host phase A (HNSW descent + layer-0 beam) then
//                  hrnn_dev_rknn_search for phases B and C
//                  (reference hrnn.h:1966-2071)
// HNSW construction with KNN seeding (addPoint), dynamic insertion, the
// mutable reverse lists, serve-mode compaction and persistence are host C++.
//
// Host memory layout (differs from the reference's AoS, which is kept only in
// the file format): vectors contiguous [max_elements x d] f32; level-0 links
// [max_elements x (4 + 4*maxM0)] with a u16 count in the first word; upper
// links per element; the ranked KNN graph as SoA ids u32[n*K], dists f32[n*K],
// counts u16[n]; the reverse CSR as u64 offsets + u32/u64 packed entries.
#pragma once

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <iostream>
#include <limits>
#include <memory>
#include <mutex>
#include <queue>
#include <random>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../hrnn_dev.h"
#include "hnswlib.h"

namespace hnswlib {

typedef unsigned int tableint;
typedef unsigned int linklistsizeint;

template <typename dist_t>
class HRNN : public AlgorithmInterface<dist_t> {
 public:
    static constexpr tableint INVALID_KNNG_ID = std::numeric_limits<tableint>::max();
    static constexpr uint32_t kFileMagic = 0x4E4E5248u;   // "HRNN"

    // ---- parameters and state (public names kept for callers) ----------
    size_t max_elements_{0};
    std::atomic<size_t> cur_element_count{0};
    size_t M_{0}, maxM_{0}, maxM0_{0};
    size_t ef_construction_{0};
    size_t ef_{10};
    double mult_{0.0};
    int maxlevel_{-1};
    tableint enterpoint_node_{0};

    size_t K_knng_{0};
    bool knng_built_{false};
    bool rknng_built_{false};

    bool rknng_wide_{false};
    unsigned rknng_rank_bits_{0};
    std::vector<uint64_t> rknng_offsets_;
    std::vector<uint32_t> rknng_entries_32_;
    std::vector<uint64_t> rknng_entries_64_;

    std::vector<std::vector<uint32_t>> rknng_lists_;
    bool rknng_mutable_{false};

    bool serve_mode_{false};
    size_t max_k_serve_{0};

    // ---- construction ------------------------------------------------------
    HRNN(SpaceInterface<dist_t> *s, const std::string &location, bool nmslib = false,
         size_t max_elements = 0, bool allow_replace_deleted = false) {
        (void)nmslib;
        if (allow_replace_deleted)
            throw std::runtime_error("HRNN: replacement of deleted elements is not supported");
        bind_space(s);
        loadIndex(location, s, max_elements);
    }

    HRNN(SpaceInterface<dist_t> *s, size_t max_elements, size_t M = 16,
         size_t ef_construction = 200, size_t random_seed = 100,
         bool allow_replace_deleted = false, size_t K_knng = 0) {
        if (allow_replace_deleted)
            throw std::runtime_error("HRNN: replacement of deleted elements is not supported");
        bind_space(s);
        M_ = std::min<size_t>(M, 10000);
        maxM_ = M_;
        maxM0_ = 2 * M_;
        ef_construction_ = std::max(ef_construction, M_);
        mult_ = 1.0 / std::log((double)M_);
        K_knng_ = K_knng;
        if (K_knng_ > 65535) throw std::runtime_error("HRNN: K_knng must fit in 16 bits");
        level_rng_.seed(random_seed);
        allocate(max_elements);
    }

    ~HRNN() {
        release_upper();
        close_device();
    }

    HRNN(const HRNN &) = delete;
    HRNN &operator=(const HRNN &) = delete;

    void setEf(size_t ef) { ef_ = ef; }
    size_t getMaxElements() const { return max_elements_; }
    size_t getCurrentElementCount() const { return cur_element_count; }
    size_t getDeletedCount() const { return 0; }
    size_t dim() const { return dim_; }

    const float *getDataByInternalId(tableint id) const { return vecs_.data() + (size_t)id * dim_; }
    labeltype getExternalLabel(tableint id) const { return labels_[id]; }

    // ---- HNSW insertion with KNN seeding ----------------------------------
    // reference: addPoint (hrnn.h:1179, 1378-1517).  Thread-safe.
    void addPoint(const void *data_point, labeltype label, bool replace_deleted = false) override {
        if (replace_deleted)
            throw std::runtime_error("HRNN: replacement of deleted elements is not supported");
        insert_internal(static_cast<const float *>(data_point), label);
    }

    // ---- kNN search over the HNSW graph -----------------------------------
    std::priority_queue<std::pair<dist_t, labeltype>> searchKnn(
        const void *query, size_t k, BaseFilterFunctor *filter = nullptr) const override {
        std::priority_queue<std::pair<dist_t, labeltype>> out;
        if (cur_element_count == 0) return out;
        const float *q = static_cast<const float *>(query);
        tableint ep = descend(q);
        auto W = beam_query(ep, q, std::max(ef_, k));
        std::vector<std::pair<dist_t, tableint>> v;
        while (!W.empty()) {
            v.push_back(W.top());
            W.pop();
        }
        std::sort(v.begin(), v.end());
        for (auto &p : v) {
            if (filter && !(*filter)(labels_[p.second])) continue;
            out.emplace(p.first, labels_[p.second]);
            if (out.size() == k) break;
        }
        return out;
    }

    // ---- HNSW construction with KNN seeding on the GPU ------------------------
    // The batch form of addPoint (reference hrnn.h:1378-1517) into an empty
    // index: rows X[0 .. nx) get internal ids 0 .. nx-1 and labels labels[i]
    // (i when labels is null); levels are drawn in id order with this index's
    // generator, as nx sequential addPoint calls would draw them; the graph
    // and the KNN seed lists are built by hrnn_dev_hnsw_build and copied back,
    // and stay resident on the device.
    void addPointsDevice(const float *X, size_t nx, const labeltype *labels = nullptr) {
        if (cur_element_count != 0)
            throw std::runtime_error("addPointsDevice: the index must be empty (use addPoint to extend it)");
        if (nx > max_elements_) throw std::runtime_error("The number of elements exceeds the specified limit");
        if (nx == 0) return;
        if (maxM0_ > 64 || M_ < 2) throw std::runtime_error("addPointsDevice: M must be in [2, 32]");
        {
            std::lock_guard<std::mutex> g(label_mu_);
            for (size_t i = 0; i < nx; i++) {
                const labeltype lab = labels ? labels[i] : (labeltype)i;
                if (!label_map_.emplace(lab, (tableint)i).second) {
                    label_map_.clear();
                    throw std::runtime_error("addPoint: label already present (updates are not supported)");
                }
                labels_[i] = lab;
            }
        }
        std::memcpy(vecs_.data(), X, nx * data_size_);
        for (size_t i = 0; i < nx; i++) levels_[i] = random_level();
        cur_element_count = nx;
        std::lock_guard<std::mutex> g(dev_mu_);
        sync_device_base();
        if (build_batch_permille_ > 0)
            dev_check(hrnn_dev_set_tuning(dev_, "hnsw_batch_permille", build_batch_permille_));
        if (build_seed_exp_ >= 0) dev_check(hrnn_dev_set_tuning(dev_, "hnsw_seed_exp", build_seed_exp_));
        uint32_t entry = 0;
        int maxl = 0;
        hrnn_hnsw_build_stats st{};
        dev_check(hrnn_dev_hnsw_build(dev_, levels_.data(), (uint32_t)M_, (uint32_t)ef_construction_,
                                      (uint32_t)K_knng_, &entry, &maxl, &st));
        last_build_stats_ = st;
        // copy the graph back: layer 0 as is, upper blocks in id order
        size_t pool = 0;
        for (size_t i = 0; i < nx; i++) pool += (size_t)levels_[i] * (maxM_ + 1);
        std::vector<uint32_t> up(std::max<size_t>(pool, 1));
        dev_check(hrnn_dev_get_hnsw(dev_, level0_.data(), up.data(), pool));
        size_t o = 0;
        for (size_t i = 0; i < nx; i++) {
            if (levels_[i] <= 0) continue;
            upper_[i] = static_cast<char *>(std::malloc((size_t)levels_[i] * links_bytes_));
            if (!upper_[i]) throw std::runtime_error("addPoint: out of memory");
            std::memcpy(upper_[i], up.data() + o, (size_t)levels_[i] * links_bytes_);
            o += (size_t)levels_[i] * (maxM_ + 1);
        }
        enterpoint_node_ = entry;
        maxlevel_ = maxl;
        dev_hnsw_rows_ = nx;
        if (K_knng_ > 0) {
            dev_check(hrnn_dev_get_knng(dev_, knn_ids_.data(), knn_d_.data(), knn_cnt_.data()));
            dev_knng_ok_ = true;
        }
        dev_csr_ok_ = false;
    }
    const hrnn_hnsw_build_stats &lastBuildStats() const { return last_build_stats_; }
    // Batch size of addPointsDevice as a fraction of the graph, per mille
    // (0: the device default).
    void setBuildBatchPermille(int permille) { build_batch_permille_ = permille; }
    // Expansions of each new point's seed completion (-1: the device default).
    void setBuildSeedExpansions(int e) { build_seed_exp_ = e; }

    // ---- (a) NNDescent on the GPU -----------------------------------------
    // reference: refineKNNG (hrnn.h:1725-1862)
    void refineKNNG(int num_iters = 3, int sample_size = 20) {
        require_knn("refineKNNG");
        size_t n = cur_element_count;
        std::cout << "    refineKNNG (NNDescent, GPU): n=" << n << " K=" << K_knng_
                  << " iters=" << num_iters << " S=" << sample_size << std::endl;
        std::lock_guard<std::mutex> g(dev_mu_);
        sync_device_base();
        sync_device_knng();
        // every entry starts new, as the reference's refineKNNG starts each
        // call (hrnn.h:1744-1748), also when the device graph is already current
        dev_check(hrnn_dev_set_knng_flags(dev_, nullptr));
        std::vector<hrnn_nnd_iter_stats> st(std::max(num_iters, 1));
        dev_check(hrnn_dev_nndescent(dev_, num_iters, sample_size, 0, st.data()));
        dev_check(hrnn_dev_get_knng(dev_, knn_ids_.data(), knn_d_.data(), knn_cnt_.data()));
        dev_knng_ok_ = true;
        dev_csr_ok_ = false;
        for (int it = 0; it < num_iters; it++)
            std::cout << "      iter " << it << ": " << st[it].updates << " updates, "
                      << st[it].dists << " dists  (" << (long)st[it].ms << " ms)"
                      << (st[it].dropped ? "  dropped=" + std::to_string(st[it].dropped) : "")
                      << std::endl;
        last_nnd_stats_.assign(st.begin(), st.begin() + num_iters);
        knng_built_ = true;
    }

    // ---- (b) reverse CSR on the GPU ------------------------------------------
    // reference: buildRKNNG / buildRKNNG_fill (hrnn.h:1869-1935)
    void buildRKNNG() {
        if (!knng_built_) throw std::runtime_error("buildRKNNG: KNNG not built yet");
        require_knn("buildRKNNG");
        std::lock_guard<std::mutex> g(dev_mu_);
        sync_device_base();
        sync_device_knng();
        uint32_t rb = 0;
        int wide = 0;
        uint64_t ne = 0;
        dev_check(hrnn_dev_build_rknn_csr(dev_, &rb, &wide, &ne));
        size_t n = cur_element_count;
        unsigned node_bits = 0;
        while ((1ull << node_bits) < n) node_bits++;
        std::cout << "    buildRKNNG (GPU): node_bits=" << node_bits << " rank_bits=" << rb
                  << " wide=" << wide << std::endl;
        rknng_rank_bits_ = rb;
        rknng_wide_ = wide != 0;
        rknng_offsets_.assign(n + 1, 0);
        if (rknng_wide_) {
            rknng_entries_64_.resize(ne);
            rknng_entries_32_.clear();
            dev_check(hrnn_dev_get_rknn_csr(dev_, rknng_offsets_.data(), rknng_entries_64_.data()));
        } else {
            rknng_entries_32_.resize(ne);
            rknng_entries_64_.clear();
            dev_check(hrnn_dev_get_rknn_csr(dev_, rknng_offsets_.data(), rknng_entries_32_.data()));
        }
        dev_csr_ok_ = true;
        rknng_built_ = true;
        std::cout << "    buildRKNNG: total_entries=" << rknng_offsets_[n] << std::endl;
    }

    // reference: buildRKNNG_mutable (hrnn.h:1942-1959): per-node reverse lists
    // in (source, rank) order, sized to max_elements.
    void buildRKNNG_mutable() {
        if (!knng_built_) throw std::runtime_error("buildRKNNG_mutable: KNNG not built yet");
        require_knn("buildRKNNG_mutable");
        size_t n = cur_element_count, total = 0;
        rknng_lists_.assign(max_elements_, {});
        for (size_t i = 0; i < n; i++) {
            const uint32_t *row = knn_ids_.data() + i * K_knng_;
            for (unsigned j = 0; j < knn_cnt_[i]; j++) rknng_lists_[row[j]].push_back((uint32_t)i);
            total += knn_cnt_[i];
        }
        rknng_mutable_ = true;
        std::cout << "    buildRKNNG_mutable: n=" << n << " total_entries=" << total << std::endl;
    }

    // ---- (c) RkNN search --------------------------------------------------------
    // reference: searchRknn (hrnn.h:1966-2071).  Labels in ascending internal-id order.
    std::vector<labeltype> searchRknn(const void *query_data, size_t m, float threshold,
                                      size_t k_rknn, size_t K_prime, size_t ef_search) const {
        auto r = searchRknnBatch(static_cast<const float *>(query_data), 1, m, threshold, k_rknn,
                                 K_prime, ef_search, 1);
        return std::move(r[0]);
    }

    // Batched form: phase A for all queries on the host (OpenMP), one GPU
    // launch for phases B and C.
    std::vector<std::vector<labeltype>> searchRknnBatch(const float *Q, size_t nq, size_t m,
                                                        float threshold, size_t k_rknn,
                                                        size_t K_prime, size_t ef_search,
                                                        int threads = 0) const {
        if (!rknng_built_) throw std::runtime_error("searchRknn: RKNNG not built");
        std::vector<std::vector<labeltype>> out(nq);
        if (cur_element_count == 0 || nq == 0) return out;
        if (k_rknn == 0 || k_rknn > knn_depth())
            throw std::runtime_error("searchRknn: k must be in [1, K]");
        std::vector<uint32_t> prox = phase_a_on_gpu(nq) ? proxiesDevice(Q, nq, m, ef_search)
                                                        : proxies(Q, nq, m, ef_search, threads);
        std::vector<uint64_t> off(nq + 1);
        std::vector<uint32_t> ids;
        device_search(Q, nq, prox.data(), m, nullptr, (uint32_t)k_rknn, (uint32_t)K_prime,
                      threshold, off, ids);
        for (size_t q = 0; q < nq; q++) {
            out[q].reserve(off[q + 1] - off[q]);
            for (uint64_t i = off[q]; i < off[q + 1]; i++) out[q].push_back(labels_[ids[i]]);
        }
        return out;
    }

    // Phase A only: m proxies per query (internal ids, 0xFFFFFFFF-padded).
    std::vector<uint32_t> proxies(const float *Q, size_t nq, size_t m, size_t ef_search,
                                  int threads = 0) const {
        std::vector<uint32_t> prox(nq * m, INVALID_KNNG_ID);
        if (cur_element_count == 0) return prox;
        size_t ef = std::max(ef_search, m);
        (void)threads;
#pragma omp parallel for schedule(dynamic, 8) num_threads(threads > 0 ? threads : omp_threads_default()) if (nq > 1)
        for (long long qi = 0; qi < (long long)nq; qi++) {
            const float *q = Q + (size_t)qi * dim_;
            tableint ep = descend(q);
            auto W = beam_query(ep, q, ef);
            while (W.size() > m) W.pop();
            std::vector<tableint> tmp;
            while (!W.empty()) {
                tmp.push_back(W.top().second);
                W.pop();
            }
            for (size_t j = 0; j < tmp.size(); j++) prox[(size_t)qi * m + j] = tmp[tmp.size() - 1 - j];
        }
        return prox;
    }

    // Phase A with the per-query beam statistics the learned m uses
    // (raw[nq x 20], same layout as the device library's beam statistics):
    //   0 d(q, entry)   1 d(q, layer-0 start)   2..17 the 16 closest beam
    //   distances (the last one repeated when the beam holds fewer)
    //   18 expansions   19 beam size
    void proxiesWithStats(const float *Q, size_t nq, size_t m, size_t ef_search, int threads,
                          uint32_t *prox, float *raw) const {
        std::fill(prox, prox + nq * m, INVALID_KNNG_ID);
        if (cur_element_count == 0) {
            std::fill(raw, raw + nq * 20, 0.f);
            return;
        }
        size_t ef = std::max(ef_search, m);
#pragma omp parallel for schedule(dynamic, 8) num_threads(threads > 0 ? threads : omp_threads_default()) if (nq > 1)
        for (long long qi = 0; qi < (long long)nq; qi++) {
            const float *q = Q + (size_t)qi * dim_;
            float *r = raw + (size_t)qi * 20;
            tableint ep = descend(q);
            unsigned expanded = 0;
            auto W = beam_query(ep, q, ef, &expanded);
            const size_t beam = W.size();
            while (W.size() > m) W.pop();       // the m proxies, as proxies() keeps them
            std::vector<std::pair<dist_t, tableint>> v;
            v.reserve(W.size());
            while (!W.empty()) {
                v.push_back(W.top());
                W.pop();
            }
            std::reverse(v.begin(), v.end());   // closest first
            for (size_t j = 0; j < v.size(); j++) prox[(size_t)qi * m + j] = v[j].second;
            r[0] = dist_to(q, enterpoint_node_);
            r[1] = dist_to(q, ep);
            for (size_t j = 0; j < 16; j++) r[2 + j] = v[std::min(j, v.size() - 1)].first;
            r[18] = (float)expanded;
            r[19] = (float)beam;
        }
    }

    // The 32 learned-m features of each query from its beam statistics (raw,
    // as proxiesWithStats) and the ranked KNN lists of its 16 closest proxies
    // (prox rows of mstride >= 16 ids, closest first):
    //   0 log1p d(q,entry)   1 log1p d(q,start)   2..17 log1p beam distances
    //   18..25 log1p r_k of proxies 1..8 (0 for a missing proxy)
    //   26..28 mean / min / max of log1p r_k over proxies 1..16
    //   29 fraction of proxies 1..16 with d(q,p) <= r_k(p)
    //   30 log1p d1 - log1p r_k(p1)   31 log1p(expansions)
    // r_k(p) is the k-th list distance, 3e38 when p's list is shorter.
    void queryFeatures(const float *raw, const uint32_t *prox, size_t nq, size_t mstride, size_t k,
                       float *feats) const {
        if (mstride < 16) throw std::runtime_error("queryFeatures: need at least 16 proxies per query");
        if (k == 0 || k > K_knng_) throw std::runtime_error("queryFeatures: k must be in [1, K]");
        const size_t n = cur_element_count;
        for (size_t qi = 0; qi < nq; qi++) {
            const float *r = raw + qi * 20;
            float *f = feats + qi * 32;
            f[0] = std::log1p(r[0]);
            f[1] = std::log1p(r[1]);
            for (int j = 0; j < 16; j++) f[2 + j] = std::log1p(std::min(r[2 + j], 3.0e38f));
            float sum = 0.f, mn = FLT_MAX, mx = -FLT_MAX, inside = 0.f, rk1 = 0.f;
            int cnt = 0;
            for (int j = 0; j < 16; j++) {
                uint32_t p = prox[qi * mstride + j];
                float rk = 0.f;
                if (p != INVALID_KNNG_ID) {
                    if (p >= n) throw std::runtime_error("queryFeatures: proxy id >= n");
                    float v = (k - 1 < knn_cnt_[p]) ? knn_d_[(size_t)p * K_knng_ + k - 1] : 3.0e38f;
                    rk = std::log1p(v);
                    inside += (r[2 + j] <= v) ? 1.f : 0.f;
                    cnt++;
                    sum += rk;
                    mn = std::min(mn, rk);
                    mx = std::max(mx, rk);
                }
                if (j == 0) rk1 = rk;
                if (j < 8) f[18 + j] = rk;
            }
            f[26] = cnt ? sum / cnt : 0.f;
            f[27] = cnt ? mn : 0.f;
            f[28] = cnt ? mx : 0.f;
            f[29] = cnt ? inside / cnt : 0.f;
            f[30] = f[2] - rk1;
            f[31] = std::log1p(r[18]);
        }
    }

    // Phase A on the GPU (hrnn_dev_proxies): same ordering rules as proxies().
    std::vector<uint32_t> proxiesDevice(const float *Q, size_t nq, size_t m, size_t ef_search) const {
        std::vector<uint32_t> prox(nq * m, INVALID_KNNG_ID);
        if (cur_element_count == 0 || nq == 0) return prox;
        std::lock_guard<std::mutex> g(dev_mu_);
        sync_device_base();
        sync_device_hnsw();
        dev_check(hrnn_dev_proxies(dev_, Q, (uint32_t)nq, (uint32_t)m,
                                   (uint32_t)std::max(ef_search, m), prox.data()));
        return prox;
    }

    // Phases B + C for caller-given proxies; results are internal ids.
    void searchWithProxies(const float *Q, size_t nq, const uint32_t *prox, size_t m,
                           const uint32_t *k_per_q, size_t k_rknn, size_t K_prime,
                           float threshold, std::vector<uint64_t> &off,
                           std::vector<uint32_t> &ids) const {
        if (!rknng_built_) throw std::runtime_error("searchRknn: RKNNG not built");
        device_search(Q, nq, prox, m, k_per_q, (uint32_t)k_rknn, (uint32_t)K_prime, threshold,
                      off, ids);
    }

    // r_k(c) used by verification (reference getVerifyKdistSq, hrnn.h:286-293).
    float getVerifyKdistSq(tableint id, unsigned k_idx) const {
        if (serve_mode_) return kdist_[(size_t)id * max_k_serve_ + k_idx];
        if (k_idx >= knn_cnt_[id]) return std::numeric_limits<float>::max();
        return knn_d_[(size_t)id * K_knng_ + k_idx];
    }

    // reference: compactForServing (hrnn.h:299-364): dense r_k matrix for
    // k <= max_k_serve, KNN lists dropped from memory.
    void compactForServing(size_t max_k_serve) {
        if (serve_mode_) return;
        if (K_knng_ == 0 || !knng_built_)
            throw std::runtime_error("compactForServing: KNNG not built");
        if (max_k_serve == 0 || max_k_serve > K_knng_)
            throw std::runtime_error("compactForServing: max_k_serve must be in [1, K_knng]");
        size_t n = cur_element_count;
        kdist_.assign(n * max_k_serve, std::numeric_limits<float>::max());
        for (size_t i = 0; i < n; i++)
            for (size_t j = 0; j < std::min<size_t>(max_k_serve, knn_cnt_[i]); j++)
                kdist_[i * max_k_serve + j] = knn_d_[i * K_knng_ + j];
        max_k_serve_ = max_k_serve;
        serve_mode_ = true;
        std::vector<uint32_t>().swap(knn_ids_);
        std::vector<float>().swap(knn_d_);
        std::lock_guard<std::mutex> g(dev_mu_);
        dev_knng_ok_ = false;
        std::fprintf(stderr, "[compactForServing] max_k_serve=%zu  kdist_matrix: %.1f MB\n",
                     max_k_serve, (double)(n * max_k_serve * 4) / (1024.0 * 1024.0));
    }

    // ---- dynamic insertion -------------------------------------------------------
    // reference: insertPointDynamic (hrnn.h:2360-2534) with try_update_knng_evict
    // (hrnn.h:1592-1655).  Needs buildRKNNG_mutable().  Thread-safe.
    tableint insertPointDynamic(const void *data_point, labeltype label, unsigned m_update = 0,
                                bool bidirectional = false) {
        if (!rknng_mutable_) throw std::runtime_error("insertPointDynamic: call buildRKNNG_mutable first");
        const float *x = static_cast<const float *>(data_point);
        tableint xn = insert_internal(x, label);
        const size_t K = K_knng_;
        unsigned cnt = knn_cnt_[xn];
        std::vector<tableint> nbrs(knn_ids_.begin() + (size_t)xn * K,
                                   knn_ids_.begin() + (size_t)xn * K + cnt);
        for (tableint y : nbrs) {
            std::lock_guard<std::mutex> lk(node_locks_[y]);
            rknng_lists_[y].push_back(xn);
        }
        unsigned nprox = (m_update > 0 && m_update < cnt) ? m_update : cnt;
        thread_local std::vector<uint8_t> seen;
        if (seen.size() < max_elements_) seen.resize(max_elements_, 0);
        std::vector<uint32_t> cand;
        seen[xn] = 1;
        for (tableint y : nbrs) seen[y] = 1;
        for (unsigned p = 0; p < nprox; p++) {
            tableint y = nbrs[p];
            std::lock_guard<std::mutex> lk(node_locks_[y]);
            for (uint32_t c : rknng_lists_[y])
                if (!seen[c]) {
                    seen[c] = 1;
                    cand.push_back(c);
                }
        }
        seen[xn] = 0;
        for (tableint y : nbrs) seen[y] = 0;
        for (uint32_t c : cand) seen[c] = 0;

        std::priority_queue<std::pair<dist_t, tableint>> best;   // x's improved list (bidirectional)
        if (bidirectional)
            for (unsigned j = 0; j < cnt; j++) best.emplace(knn_d_[(size_t)xn * K + j], nbrs[j]);

        for (uint32_t c : cand) {
            dist_t dd = dist_(x, getDataByInternalId(c), dist_param_);
            if (bidirectional) {
                if (best.size() < K) best.emplace(dd, c);
                else if (dd < best.top().first) {
                    best.pop();
                    best.emplace(dd, c);
                }
            }
            unsigned cc = knn_cnt_[c];   // racy read: a stale value only costs a lock
            if (cc >= K && dd >= knn_d_[(size_t)c * K + cc - 1]) continue;
            tableint evicted = INVALID_KNNG_ID;
            if (!offer_evict(c, xn, dd, evicted)) continue;
            {
                std::lock_guard<std::mutex> lk(node_locks_[xn]);
                rknng_lists_[xn].push_back(c);
            }
            if (evicted != INVALID_KNNG_ID) {
                std::lock_guard<std::mutex> lk(node_locks_[evicted]);
                auto &rl = rknng_lists_[evicted];
                auto it = std::find(rl.begin(), rl.end(), c);
                if (it != rl.end()) {
                    *it = rl.back();
                    rl.pop_back();
                }
            }
        }

        if (bidirectional) {
            std::vector<std::pair<dist_t, tableint>> v;
            while (!best.empty()) {
                v.push_back(best.top());
                best.pop();
            }
            std::sort(v.begin(), v.end());
            std::vector<tableint> gained;
            {
                std::lock_guard<std::mutex> lk(node_locks_[xn]);
                // x's list as it is now (a concurrent insertion may have
                // changed it since nbrs was read): marked and cleared from
                // this copy, so every mark set here is cleared here
                const unsigned old_cnt = knn_cnt_[xn];
                const std::vector<tableint> old_ids(knn_ids_.begin() + (size_t)xn * K,
                                                    knn_ids_.begin() + (size_t)xn * K + old_cnt);
                for (tableint id : old_ids) seen[id] = 1;
                unsigned w = 0;
                for (auto &pr : v) {
                    if (pr.second == xn || w >= K || seen[pr.second] == 2) continue;
                    knn_ids_[(size_t)xn * K + w] = pr.second;
                    knn_d_[(size_t)xn * K + w] = pr.first;
                    if (seen[pr.second] == 0) gained.push_back(pr.second);
                    seen[pr.second] = 2;
                    w++;
                }
                knn_cnt_[xn] = (uint16_t)w;
                for (tableint id : old_ids) seen[id] = 0;
                for (auto &pr : v) seen[pr.second] = 0;
            }
            for (tableint y : gained) {
                std::lock_guard<std::mutex> lk(node_locks_[y]);
                rknng_lists_[y].push_back(xn);
            }
        }
        dev_knng_ok_ = false;
        dev_csr_ok_ = false;
        return xn;
    }

    // Batched incremental insertion on the GPU: insertPointDynamic
    // (hrnn.h:2360-2534) for nx points, in batches (hrnn_dev_insert_dynamic:
    // HNSW insertion with KNN seeding, candidates from the reverse lists of
    // each point's first m_update neighbours, offers merged into the affected
    // lists, with bidirectional the points' own lists improved).  Needs a
    // built KNN graph; not for concurrent use with insertPointDynamic.  The
    // mutable reverse lists, when built, are rebuilt from the new graph; the
    // flat reverse CSR must be rebuilt (buildRKNNG) before searching.
    void insertPointsDynamicDevice(const float *X, size_t nx, const labeltype *labels, unsigned m_update = 0,
                                   bool bidirectional = false) {
        if (!knng_built_) throw std::runtime_error("insertPointsDynamicDevice: KNNG not built");
        require_knn("insertPointsDynamicDevice");
        const size_t n_old = cur_element_count;
        if (n_old == 0) throw std::runtime_error("insertPointsDynamicDevice: the index is empty (use addPointsDevice)");
        if (nx == 0) return;
        if (!X || !labels) throw std::runtime_error("insertPointsDynamicDevice: null argument");
        if (n_old + nx > max_elements_) throw std::runtime_error("The number of elements exceeds the specified limit");
        if (maxM0_ > 64 || M_ < 2) throw std::runtime_error("insertPointsDynamicDevice: M must be in [2, 32]");
        std::lock_guard<std::mutex> g(dev_mu_);
        // the device holds the base rows, the graph and the lists of [0, n_old)
        sync_device_base();
        sync_device_knng();
        sync_device_hnsw();
        {
            std::lock_guard<std::mutex> gl(label_mu_);
            size_t i = 0;
            for (; i < nx; i++)
                if (!label_map_.emplace(labels[i], (tableint)(n_old + i)).second) break;
            if (i < nx) {
                for (size_t j = 0; j < i; j++) label_map_.erase(labels[j]);
                throw std::runtime_error("addPoint: label already present (updates are not supported)");
            }
            for (size_t t = 0; t < nx; t++) labels_[n_old + t] = labels[t];
        }
        std::memcpy(vecs_.data() + n_old * dim_, X, nx * data_size_);
        for (size_t t = 0; t < nx; t++) levels_[n_old + t] = random_level();
        cur_element_count = n_old + nx;
        const size_t n = cur_element_count;
        dev_check(hrnn_dev_append_base(dev_, vecs_.data() + n_old * dim_, nx));
        dev_base_rows_ = n;
        uint32_t entry = 0;
        int maxl = 0;
        hrnn_insert_stats st{};
        dev_check(hrnn_dev_insert_dynamic(dev_, n_old, levels_.data(), (uint32_t)M_, (uint32_t)ef_construction_,
                                          m_update, bidirectional ? 1 : 0, &entry, &maxl, &st));
        last_insert_stats_ = st;
        // copy the graph back: layer 0 and every upper block (links back
        // changed old nodes' lists too), then the KNN lists
        size_t pool = 0;
        for (size_t i = 0; i < n; i++) pool += (size_t)std::max(0, levels_[i]) * (maxM_ + 1);
        std::vector<uint32_t> up(std::max<size_t>(pool, 1));
        dev_check(hrnn_dev_get_hnsw(dev_, level0_.data(), up.data(), pool));
        size_t o = 0;
        for (size_t i = 0; i < n; i++) {
            if (levels_[i] <= 0) continue;
            if (i >= n_old) {
                upper_[i] = static_cast<char *>(std::malloc((size_t)levels_[i] * links_bytes_));
                if (!upper_[i]) throw std::runtime_error("addPoint: out of memory");
            }
            std::memcpy(upper_[i], up.data() + o, (size_t)levels_[i] * links_bytes_);
            o += (size_t)levels_[i] * (maxM_ + 1);
        }
        enterpoint_node_ = entry;
        maxlevel_ = maxl;
        dev_hnsw_rows_ = n;
        dev_check(hrnn_dev_get_knng(dev_, knn_ids_.data(), knn_d_.data(), knn_cnt_.data()));
        dev_knng_ok_ = true;
        dev_csr_ok_ = false;
        rknng_built_ = false;
        if (rknng_mutable_) {
            rknng_lists_.assign(max_elements_, {});
            for (size_t i = 0; i < n; i++) {
                const uint32_t *row = knn_ids_.data() + i * K_knng_;
                for (unsigned j = 0; j < knn_cnt_[i]; j++) rknng_lists_[row[j]].push_back((uint32_t)i);
            }
        }
    }
    const hrnn_insert_stats &lastInsertStats() const { return last_insert_stats_; }

    // ---- persistence (reference file format, hrnn.h:861-1047) ----------------
    void saveIndex(const std::string &location) override {
        if (serve_mode_) throw std::runtime_error("saveIndex: not available in serve mode");
        std::ofstream out(location, std::ios::binary);
        if (!out) throw std::runtime_error("saveIndex: cannot open " + location);
        const size_t n = cur_element_count;
        const FileLayout L = file_layout();
        writeBinaryPOD(out, kFileMagic);
        writeBinaryPOD(out, (size_t)0);   // offsetLevel0_
        writeBinaryPOD(out, max_elements_);
        writeBinaryPOD(out, n);
        writeBinaryPOD(out, L.elem);
        writeBinaryPOD(out, L.label_off);
        writeBinaryPOD(out, L.data_off);
        writeBinaryPOD(out, maxlevel_);
        writeBinaryPOD(out, enterpoint_node_);
        writeBinaryPOD(out, maxM_);
        writeBinaryPOD(out, maxM0_);
        writeBinaryPOD(out, M_);
        writeBinaryPOD(out, mult_);
        writeBinaryPOD(out, ef_construction_);
        writeBinaryPOD(out, K_knng_);
        writeBinaryPOD(out, L.knng_slot);
        writeBinaryPOD(out, L.knng_off);
        writeBinaryPOD(out, knng_built_);
        writeBinaryPOD(out, rknng_built_);
        // level-0 AoS: [links | KNN slot | vector | label] per element, in blocks
        const size_t blk = 4096;
        std::vector<char> buf(blk * L.elem);
        for (size_t b0 = 0; b0 < n; b0 += blk) {
            size_t nb = std::min(blk, n - b0);
            std::memset(buf.data(), 0, nb * L.elem);
            for (size_t t = 0; t < nb; t++) {
                size_t i = b0 + t;
                char *e = buf.data() + t * L.elem;
                std::memcpy(e, links0(i), links0_bytes_);
                if (K_knng_) {
                    uint32_t c = knn_cnt_[i];
                    std::memcpy(e + L.knng_off, &c, 4);
                    char *slot = e + L.knng_off + 4;
                    for (uint32_t j = 0; j < c; j++) {
                        std::memcpy(slot + j * 8, &knn_ids_[i * K_knng_ + j], 4);
                        std::memcpy(slot + j * 8 + 4, &knn_d_[i * K_knng_ + j], 4);
                    }
                }
                std::memcpy(e + L.data_off, getDataByInternalId((tableint)i), data_size_);
                std::memcpy(e + L.label_off, &labels_[i], sizeof(labeltype));
            }
            out.write(buf.data(), (std::streamsize)(nb * L.elem));
        }
        for (size_t i = 0; i < n; i++) {
            unsigned int sz = levels_[i] > 0 ? (unsigned)(links_bytes_ * levels_[i]) : 0u;
            writeBinaryPOD(out, sz);
            if (sz) out.write(upper_[i], sz);
        }
        if (rknng_built_) {
            uint8_t eb = rknng_wide_ ? 8 : 4;
            writeBinaryPOD(out, eb);
            writeBinaryPOD(out, rknng_rank_bits_);
            size_t no = rknng_offsets_.size();
            writeBinaryPOD(out, no);
            out.write(reinterpret_cast<const char *>(rknng_offsets_.data()), (std::streamsize)(no * 8));
            if (rknng_wide_) {
                size_t ne = rknng_entries_64_.size();
                writeBinaryPOD(out, ne);
                out.write(reinterpret_cast<const char *>(rknng_entries_64_.data()), (std::streamsize)(ne * 8));
                uint64_t mask = (uint64_t(1) << rknng_rank_bits_) - 1;
                writeBinaryPOD(out, mask);
            } else {
                size_t ne = rknng_entries_32_.size();
                writeBinaryPOD(out, ne);
                out.write(reinterpret_cast<const char *>(rknng_entries_32_.data()), (std::streamsize)(ne * 4));
                uint32_t mask = (uint32_t(1) << rknng_rank_bits_) - 1;
                writeBinaryPOD(out, mask);
            }
        }
        if (!out) throw std::runtime_error("saveIndex: write failed for " + location);
    }

    // Validating loader of the reference format: every read is checked and the
    // header must describe this space's dimension.
    void loadIndex(const std::string &location, SpaceInterface<dist_t> *s, size_t max_elements_i = 0) {
        std::ifstream in(location, std::ios::binary);
        if (!in.is_open()) throw std::runtime_error("Cannot open file");
        bind_space(s);
        auto rd = [&](auto &v, const char *what) {
            readBinaryPOD(in, v);
            if (!in) throw std::runtime_error(std::string("loadIndex: truncated file at ") + what);
        };
        uint32_t magic = 0;
        rd(magic, "magic");
        if (magic != kFileMagic) throw std::runtime_error("Not an HRNN index file (bad magic)");
        size_t off0, maxe, n, elem, label_off, data_off, knng_slot, knng_off;
        rd(off0, "offsetLevel0");
        rd(maxe, "max_elements");
        rd(n, "cur_element_count");
        rd(elem, "size_data_per_element");
        rd(label_off, "label_offset");
        rd(data_off, "offsetData");
        rd(maxlevel_, "maxlevel");
        rd(enterpoint_node_, "enterpoint");
        rd(maxM_, "maxM");
        rd(maxM0_, "maxM0");
        rd(M_, "M");
        rd(mult_, "mult");
        rd(ef_construction_, "ef_construction");
        rd(K_knng_, "K_knng");
        rd(knng_slot, "knng_slot");
        rd(knng_off, "knng_offset");
        rd(knng_built_, "knng_built");
        rd(rknng_built_, "rknng_built");
        if (off0 != 0) throw std::runtime_error("loadIndex: unsupported offsetLevel0");
        if (K_knng_ > 65535) throw std::runtime_error("loadIndex: K_knng out of range");
        links0_bytes_ = 4 + 4 * maxM0_;
        links_bytes_ = 4 + 4 * maxM_;
        const FileLayout L = file_layout();
        if (elem != L.elem || label_off != L.label_off || data_off != L.data_off ||
            knng_slot != L.knng_slot || knng_off != L.knng_off)
            throw std::runtime_error("loadIndex: element layout does not match this space (dimension?)");
        if (n > maxe || (n > 0 && enterpoint_node_ >= n))
            throw std::runtime_error("loadIndex: inconsistent header");
        size_t cap = std::max(max_elements_i, n);
        if (max_elements_i < n) cap = std::max(maxe, n);
        allocate(cap);
        cur_element_count = n;
        const size_t blk = 4096;
        std::vector<char> buf(blk * L.elem);
        for (size_t b0 = 0; b0 < n; b0 += blk) {
            size_t nb = std::min(blk, n - b0);
            in.read(buf.data(), (std::streamsize)(nb * L.elem));
            if (!in) throw std::runtime_error("loadIndex: truncated level-0 data");
            for (size_t t = 0; t < nb; t++) {
                size_t i = b0 + t;
                const char *e = buf.data() + t * L.elem;
                std::memcpy(links0(i), e, links0_bytes_);
                if (count_of(links0(i)) > maxM0_) throw std::runtime_error("loadIndex: bad link count");
                if (K_knng_) {
                    uint32_t c;
                    std::memcpy(&c, e + L.knng_off, 4);
                    c &= 0xFFFFu;
                    if (c > K_knng_) throw std::runtime_error("loadIndex: bad KNN count");
                    knn_cnt_[i] = (uint16_t)c;
                    const char *slot = e + L.knng_off + 4;
                    for (uint32_t j = 0; j < c; j++) {
                        std::memcpy(&knn_ids_[i * K_knng_ + j], slot + j * 8, 4);
                        std::memcpy(&knn_d_[i * K_knng_ + j], slot + j * 8 + 4, 4);
                        if (knn_ids_[i * K_knng_ + j] >= n) throw std::runtime_error("loadIndex: bad KNN id");
                    }
                }
                std::memcpy(vecs_.data() + i * dim_, e + L.data_off, data_size_);
                std::memcpy(&labels_[i], e + L.label_off, sizeof(labeltype));
                label_map_[labels_[i]] = (tableint)i;
            }
        }
        for (size_t i = 0; i < n; i++) {
            unsigned int sz = 0;
            rd(sz, "upper links");
            if (sz % links_bytes_) throw std::runtime_error("loadIndex: bad upper link size");
            levels_[i] = (int)(sz / links_bytes_);
            if (sz) {
                upper_[i] = static_cast<char *>(std::malloc(sz));
                if (!upper_[i]) throw std::runtime_error("loadIndex: out of memory");
                in.read(upper_[i], sz);
                if (!in) throw std::runtime_error("loadIndex: truncated upper links");
            }
        }
        if (rknng_built_) {
            uint8_t eb = 0;
            rd(eb, "csr entry width");
            if (eb != 4 && eb != 8) throw std::runtime_error("loadIndex: bad CSR entry width");
            rknng_wide_ = (eb == 8);
            rd(rknng_rank_bits_, "rank bits");
            size_t no = 0, ne = 0;
            rd(no, "offsets size");
            if (no != n + 1) throw std::runtime_error("loadIndex: CSR offsets size mismatch");
            rknng_offsets_.resize(no);
            in.read(reinterpret_cast<char *>(rknng_offsets_.data()), (std::streamsize)(no * 8));
            if (!in) throw std::runtime_error("loadIndex: truncated CSR offsets");
            rd(ne, "entries size");
            if (ne != rknng_offsets_[n]) throw std::runtime_error("loadIndex: CSR size mismatch");
            if (rknng_wide_) {
                rknng_entries_64_.resize(ne);
                in.read(reinterpret_cast<char *>(rknng_entries_64_.data()), (std::streamsize)(ne * 8));
                uint64_t mask;
                rd(mask, "mask");
            } else {
                rknng_entries_32_.resize(ne);
                in.read(reinterpret_cast<char *>(rknng_entries_32_.data()), (std::streamsize)(ne * 4));
                uint32_t mask;
                rd(mask, "mask");
            }
            if (!in) throw std::runtime_error("loadIndex: truncated CSR entries");
        }
        ef_ = 10;
        dev_base_rows_ = 0;
        dev_hnsw_rows_ = 0;
        dev_knng_ok_ = false;
        dev_csr_ok_ = false;
    }

    // ---- accessors used by the C entry points and tools ------------------------
    const std::vector<uint32_t> &knnIds() const { return knn_ids_; }
    const std::vector<float> &knnDists() const { return knn_d_; }
    const std::vector<uint16_t> &knnCounts() const { return knn_cnt_; }
    void setKnng(const uint32_t *ids, const float *d, const uint16_t *cnt) {
        require_knn("setKnng");
        size_t n = cur_element_count;
        std::memcpy(knn_ids_.data(), ids, n * K_knng_ * 4);
        std::memcpy(knn_d_.data(), d, n * K_knng_ * 4);
        std::memcpy(knn_cnt_.data(), cnt, n * 2);
        knng_built_ = true;
        dev_knng_ok_ = false;
        dev_csr_ok_ = false;
    }
    std::vector<int> levels() const { return std::vector<int>(levels_.begin(), levels_.begin() + cur_element_count); }
    const uint32_t *level0Links(tableint i) const { return links0(i); }
    const std::vector<hrnn_nnd_iter_stats> &lastRefineStats() const { return last_nnd_stats_; }
    // Device handle of this index (opened on first use).
    hrnn_dev_t device() const {
        std::lock_guard<std::mutex> g(dev_mu_);
        open_device();
        return dev_;
    }
    // Make base, graph and CSR resident on the device (for callers that time
    // the device alone).
    void syncDevice() const {
        std::lock_guard<std::mutex> g(dev_mu_);
        sync_device_base();
        sync_device_knng();
        if (rknng_built_) sync_device_csr();
        if (cur_element_count > 0 && M_ * 2 <= 64) sync_device_hnsw();
    }
    // Phase A of a batch runs on the GPU (k_phase_a) from 32 queries up,
    // unless HRNN_PHASE_A=host; single queries and small batches stay on the
    // host, where a launch would cost more than the beam.
    bool phase_a_on_gpu(size_t nq) const {
        const char *e = std::getenv("HRNN_PHASE_A");
        if (e && std::string(e) == "host") return false;
        if (e && std::string(e) == "gpu") return maxM0_ <= 64;
        return nq >= 32 && maxM0_ <= 64;
    }

 private:
    // ---- space ------------------------------------------------------------------
    size_t dim_{0}, data_size_{0};
    DISTFUNC<dist_t> dist_{nullptr};
    void *dist_param_{nullptr};

    // ---- graph storage ------------------------------------------------------------
    size_t links0_bytes_{0}, links_bytes_{0};
    std::vector<float> vecs_;
    std::vector<uint32_t> level0_;
    std::vector<char *> upper_;
    std::vector<int> levels_;
    std::vector<labeltype> labels_;
    std::unique_ptr<std::mutex[]> node_locks_;
    std::mutex label_mu_, entry_mu_, rng_mu_;
    std::unordered_map<labeltype, tableint> label_map_;
    std::default_random_engine level_rng_;

    std::vector<uint32_t> knn_ids_;
    std::vector<float> knn_d_;
    std::vector<uint16_t> knn_cnt_;
    std::vector<float> kdist_;
    std::vector<hrnn_nnd_iter_stats> last_nnd_stats_;
    hrnn_hnsw_build_stats last_build_stats_{};
    int build_batch_permille_{0};
    int build_seed_exp_{-1};

    // ---- device state ---------------------------------------------------------------
    mutable hrnn_dev_t dev_{nullptr};
    mutable std::mutex dev_mu_;
    mutable size_t dev_base_rows_{0};
    mutable bool dev_knng_ok_{false};
    mutable bool dev_csr_ok_{false};
    mutable size_t dev_hnsw_rows_{0};
    hrnn_insert_stats last_insert_stats_{};   // element count when the HNSW graph was uploaded

    struct FileLayout {
        size_t elem, label_off, data_off, knng_slot, knng_off;
    };
    FileLayout file_layout() const {
        FileLayout L;
        L.knng_slot = K_knng_ ? 4 + K_knng_ * 8 : 0;
        L.knng_off = K_knng_ ? links0_bytes_ : 0;
        L.data_off = links0_bytes_ + L.knng_slot;
        L.label_off = L.data_off + data_size_;
        L.elem = L.label_off + sizeof(labeltype);
        return L;
    }

    static int omp_threads_default() {
        const char *e = std::getenv("OMP_NUM_THREADS");
        int t = e ? std::atoi(e) : 0;
        return t > 0 ? t : 8;
    }

    void bind_space(SpaceInterface<dist_t> *s) {
        data_size_ = s->get_data_size();
        dim_ = data_size_ / sizeof(float);
        dist_ = s->get_dist_func();
        dist_param_ = s->get_dist_func_param();
    }

    void allocate(size_t max_elements) {
        release_upper();
        max_elements_ = max_elements;
        links0_bytes_ = 4 + 4 * maxM0_;
        links_bytes_ = 4 + 4 * maxM_;
        vecs_.assign(max_elements * dim_, 0.f);
        level0_.assign(max_elements * (links0_bytes_ / 4), 0u);
        upper_.assign(max_elements, nullptr);
        levels_.assign(max_elements, 0);
        labels_.assign(max_elements, 0);
        node_locks_.reset(new std::mutex[std::max<size_t>(max_elements, 1)]);
        knn_ids_.assign(max_elements * K_knng_, 0u);
        knn_d_.assign(max_elements * K_knng_, 0.f);
        knn_cnt_.assign(max_elements, 0);
        label_map_.clear();
        cur_element_count = 0;
    }

    void release_upper() {
        for (char *p : upper_) std::free(p);
        upper_.clear();
    }

    void require_knn(const char *who) const {
        if (K_knng_ == 0) throw std::runtime_error(std::string(who) + ": index built with K_knng = 0");
        if (serve_mode_) throw std::runtime_error(std::string(who) + ": not available in serve mode");
    }

    size_t knn_depth() const { return serve_mode_ ? max_k_serve_ : K_knng_; }

    uint32_t *links0(size_t i) { return level0_.data() + i * (links0_bytes_ / 4); }
    const uint32_t *links0(size_t i) const { return level0_.data() + i * (links0_bytes_ / 4); }
    uint32_t *links_at(tableint i, int level) {
        return level == 0 ? links0(i) : reinterpret_cast<uint32_t *>(upper_[i] + (level - 1) * links_bytes_);
    }
    const uint32_t *links_at(tableint i, int level) const {
        return level == 0 ? links0(i) : reinterpret_cast<const uint32_t *>(upper_[i] + (level - 1) * links_bytes_);
    }
    static unsigned count_of(const uint32_t *ll) { return ll[0] & 0xFFFFu; }
    static void set_count(uint32_t *ll, unsigned c) { ll[0] = (ll[0] & 0xFFFF0000u) | (c & 0xFFFFu); }

    dist_t dist_to(const float *q, tableint id) const { return dist_(q, getDataByInternalId(id), dist_param_); }
    dist_t dist_ids(tableint a, tableint b) const {
        return dist_(getDataByInternalId(a), getDataByInternalId(b), dist_param_);
    }

    // per-thread visited tags (epoch counter; cleared on wrap or growth)
    struct Visited {
        std::vector<uint32_t> tag;
        uint32_t epoch = 0;
    };
    uint32_t next_epoch(Visited &v) const {
        if (v.tag.size() < max_elements_) {
            v.tag.assign(max_elements_, 0);
            v.epoch = 0;
        }
        if (++v.epoch == 0) {
            std::fill(v.tag.begin(), v.tag.end(), 0);
            v.epoch = 1;
        }
        return v.epoch;
    }

    typedef std::pair<dist_t, tableint> DP;
    typedef std::priority_queue<DP> MaxHeap;
    typedef std::priority_queue<DP, std::vector<DP>, std::greater<DP>> MinHeap;
    // Heap ordered by the distance alone: equal distances leave in the order
    // the standard library's heap operations give them.  The query beam uses
    // it for both of its queues, with the same sequence of pushes and pops as
    // the reference's searchBaseLayerST (hrnn.h:485-616, CompareByFirst at
    // hrnn.h:218-223), so distance ties resolve the same way.
    struct ByDist {
        bool operator()(const DP &a, const DP &b) const noexcept { return a.first < b.first; }
    };
    typedef std::priority_queue<DP, std::vector<DP>, ByDist> DistHeap;

    int random_level() {
        std::lock_guard<std::mutex> g(rng_mu_);
        std::uniform_real_distribution<double> u(0.0, 1.0);
        return (int)(-std::log(u(level_rng_)) * mult_);
    }

    // Greedy descent over the upper layers (reference hrnn.h:1993-2013).
    tableint descend(const float *q, int stop_level = 0) const {
        tableint cur = enterpoint_node_;
        dist_t cd = dist_to(q, cur);
        for (int level = maxlevel_; level > stop_level; level--) {
            bool moved = true;
            while (moved) {
                moved = false;
                uint32_t buf[1024];
                unsigned c = copy_links(cur, level, buf);
                for (unsigned t = 0; t < c; t++) {
                    dist_t dd = dist_to(q, buf[t]);
                    if (dd < cd) {
                        cd = dd;
                        cur = buf[t];
                        moved = true;
                    }
                }
            }
        }
        return cur;
    }

    unsigned copy_links(tableint id, int level, uint32_t *buf) const {
        std::lock_guard<std::mutex> lk(node_locks_[id]);
        const uint32_t *ll = links_at(id, level);
        unsigned c = count_of(ll);
        std::memcpy(buf, ll + 1, c * 4);
        return c;
    }

    // Construction beam at one layer (reference searchBaseLayer, hrnn.h:401-481):
    // stop when the closest open candidate is farther than the ef-th result.
    MaxHeap beam_build(tableint ep, const float *q, int level, size_t ef) const {
        thread_local Visited vis;
        uint32_t tag = next_epoch(vis);
        MaxHeap W;
        MinHeap C;
        dist_t d0 = dist_to(q, ep);
        W.emplace(d0, ep);
        C.emplace(d0, ep);
        vis.tag[ep] = tag;
        dist_t lb = d0;
        uint32_t buf[1024];
        while (!C.empty()) {
            DP c = C.top();
            if (c.first > lb && W.size() >= ef) break;
            C.pop();
            unsigned cnt = copy_links(c.second, level, buf);
            for (unsigned t = 0; t < cnt; t++) {
                tableint e = buf[t];
                if (vis.tag[e] == tag) continue;
                vis.tag[e] = tag;
                dist_t de = dist_to(q, e);
                if (W.size() < ef || de < lb) {
                    C.emplace(de, e);
                    W.emplace(de, e);
                    if (W.size() > ef) W.pop();
                    lb = W.top().first;
                }
            }
        }
        return W;
    }

    // Query beam at layer 0 with the bare stop rule of the reference's
    // searchBaseLayerST<true> (hrnn.h:485-616): stop when the closest open
    // candidate is farther than the current worst result.
    // W holds the results (largest distance on top); C the open candidates
    // keyed by negated distance (closest on top).
    DistHeap beam_query(tableint ep, const float *q, size_t ef, unsigned *expanded = nullptr) const {
        thread_local Visited vis;
        uint32_t tag = next_epoch(vis);
        DistHeap W, C;
        dist_t d0 = dist_to(q, ep);
        W.emplace(d0, ep);
        C.emplace(-d0, ep);
        vis.tag[ep] = tag;
        dist_t lb = d0;
        while (!C.empty()) {
            DP c = C.top();
            if (-c.first > lb) break;
            C.pop();
            if (expanded) ++*expanded;
            const uint32_t *ll = links0(c.second);   // layer 0 is immutable at query time
            unsigned cnt = count_of(ll);
            for (unsigned t = 1; t <= cnt; t++) {
                tableint e = ll[t];
                if (vis.tag[e] == tag) continue;
                vis.tag[e] = tag;
                dist_t de = dist_to(q, e);
                if (W.size() < ef || lb > de) {
                    C.emplace(-de, e);
                    W.emplace(de, e);
                    if (W.size() > ef) W.pop();
                    lb = W.top().first;
                }
            }
        }
        return W;
    }

    // Diversity pruning (reference getNeighborsByHeuristic2, hrnn.h:619-658):
    // candidates ascending by distance to the base; keep c unless a kept r is
    // closer to c than the base is.  Fewer than M candidates are all kept.
    std::vector<DP> select_diverse(std::vector<DP> cand, size_t M) const {
        std::sort(cand.begin(), cand.end());
        if (cand.size() < M) return cand;
        std::vector<DP> kept;
        for (const DP &c : cand) {
            if (kept.size() >= M) break;
            bool ok = true;
            for (const DP &r : kept)
                if (dist_ids(r.second, c.second) < c.first) {
                    ok = false;
                    break;
                }
            if (ok) kept.push_back(c);
        }
        return kept;
    }

    // Link cur at `level` to a diverse subset of W and link back, pruning full
    // lists (reference mutuallyConnectNewElement, hrnn.h:682-805).  Returns the
    // closest selected node (entry point of the next layer).
    tableint connect(tableint cur, MaxHeap &W, int level) {
        std::vector<DP> cand;
        cand.reserve(W.size());
        while (!W.empty()) {
            cand.push_back(W.top());
            W.pop();
        }
        std::vector<DP> sel = select_diverse(std::move(cand), M_);
        const size_t cap = level ? maxM_ : maxM0_;
        {
            // cur's own lock is held by insert_internal for the whole insertion
            uint32_t *ll = links_at(cur, level);
            set_count(ll, (unsigned)sel.size());
            for (size_t t = 0; t < sel.size(); t++) ll[1 + t] = sel[t].second;
        }
        for (const DP &s : sel) {
            tableint nb = s.second;
            std::lock_guard<std::mutex> lk(node_locks_[nb]);
            uint32_t *ll = links_at(nb, level);
            unsigned c = count_of(ll);
            if (c < cap) {
                ll[1 + c] = cur;
                set_count(ll, c + 1);
                continue;
            }
            std::vector<DP> pool;
            pool.reserve(c + 1);
            pool.emplace_back(s.first, cur);
            for (unsigned t = 0; t < c; t++) pool.emplace_back(dist_ids(ll[1 + t], nb), ll[1 + t]);
            std::vector<DP> keep = select_diverse(std::move(pool), cap);
            for (size_t t = 0; t < keep.size(); t++) ll[1 + t] = keep[t].second;
            set_count(ll, (unsigned)keep.size());
        }
        return sel.empty() ? cur : sel.front().second;
    }

    tableint insert_internal(const float *x, labeltype label) {
        tableint cur;
        {
            std::lock_guard<std::mutex> g(label_mu_);
            if (label_map_.count(label))
                throw std::runtime_error("addPoint: label already present (updates are not supported)");
            if (cur_element_count >= max_elements_)
                throw std::runtime_error("The number of elements exceeds the specified limit");
            cur = (tableint)cur_element_count.load();
            label_map_[label] = cur;
            cur_element_count++;
        }
        // Held until cur is linked at every level (as the reference's lock_el,
        // hrnn.h:1411): a concurrent insertion that reaches cur through an
        // upper layer waits here instead of starting its layer-0 beam from a
        // node whose layer-0 list is still empty.
        std::unique_lock<std::mutex> lock_el(node_locks_[cur]);
        int lvl = random_level();
        std::memcpy(vecs_.data() + (size_t)cur * dim_, x, data_size_);
        labels_[cur] = label;
        if (lvl > 0) {
            upper_[cur] = static_cast<char *>(std::calloc((size_t)lvl, links_bytes_));
            if (!upper_[cur]) throw std::runtime_error("addPoint: out of memory");
        }
        levels_[cur] = lvl;
        std::unique_lock<std::mutex> entry(entry_mu_);
        int maxl = maxlevel_;
        tableint ep = enterpoint_node_;
        if (maxl < 0) {   // first element
            enterpoint_node_ = cur;
            maxlevel_ = lvl;
            return cur;
        }
        if (lvl <= maxl) entry.unlock();   // keep the lock only when cur becomes the entry point
        if (lvl < maxl) {
            dist_t cd = dist_to(x, ep);
            for (int level = maxl; level > lvl; level--) {
                bool moved = true;
                while (moved) {
                    moved = false;
                    uint32_t buf[1024];
                    unsigned c = copy_links(ep, level, buf);
                    for (unsigned t = 0; t < c; t++) {
                        dist_t dd = dist_to(x, buf[t]);
                        if (dd < cd) {
                            cd = dd;
                            ep = buf[t];
                            moved = true;
                        }
                    }
                }
            }
        }
        for (int level = std::min(lvl, maxl); level >= 0; level--) {
            MaxHeap W = beam_build(ep, x, level, ef_construction_);
            if (level == 0 && K_knng_ > 0) seed_knn(cur, W);
            ep = connect(cur, W, level);
        }
        if (lvl > maxl) {
            enterpoint_node_ = cur;
            maxlevel_ = lvl;
        }
        return cur;
    }

    // KNN seed: the layer-0 beam sorted by (dist, id), self skipped, first K
    // (reference hrnn.h:1479-1501).
    void seed_knn(tableint cur, const MaxHeap &W) {
        MaxHeap copy = W;
        std::vector<DP> v;
        v.reserve(copy.size());
        while (!copy.empty()) {
            v.push_back(copy.top());
            copy.pop();
        }
        std::sort(v.begin(), v.end());
        unsigned c = 0;
        uint32_t *ids = knn_ids_.data() + (size_t)cur * K_knng_;
        float *ds = knn_d_.data() + (size_t)cur * K_knng_;
        for (const DP &p : v) {
            if (p.second == cur) continue;
            if (c >= K_knng_) break;
            ids[c] = p.second;
            ds[c] = p.first;
            c++;
        }
        knn_cnt_[cur] = (uint16_t)c;
    }

    // try_update_knng_evict semantics (reference hrnn.h:1592-1655): reject a
    // duplicate; append when not full; otherwise replace the farthest if
    // strictly closer; keep ascending order with equal distances in arrival order.
    bool offer_evict(tableint u, tableint w, dist_t dd, tableint &evicted) {
        std::lock_guard<std::mutex> lk(node_locks_[u]);
        const size_t K = K_knng_;
        uint32_t *ids = knn_ids_.data() + (size_t)u * K;
        float *ds = knn_d_.data() + (size_t)u * K;
        unsigned c = knn_cnt_[u];
        for (unsigned t = 0; t < c; t++)
            if (ids[t] == w) return false;
        unsigned pos;
        if (c < K) {
            pos = c;
            knn_cnt_[u] = (uint16_t)(c + 1);
            evicted = INVALID_KNNG_ID;
        } else {
            if (dd >= ds[c - 1]) return false;
            pos = c - 1;
            evicted = ids[pos];
        }
        while (pos > 0 && dd < ds[pos - 1]) {
            ids[pos] = ids[pos - 1];
            ds[pos] = ds[pos - 1];
            pos--;
        }
        ids[pos] = w;
        ds[pos] = dd;
        return true;
    }

    // ---- device plumbing (callers hold dev_mu_) ---------------------------------------
    static int device_ordinal() {
        const char *e = std::getenv("HRNN_DEVICE");
        return e ? std::atoi(e) : 0;
    }
    void open_device() const {
        if (dev_) return;
        int rc = hrnn_dev_open(device_ordinal(), &dev_);
        if (rc != HRNN_OK) {
            dev_ = nullptr;
            throw std::runtime_error("HRNN: no usable GPU (hrnn_dev_open returned " + std::to_string(rc) + ")");
        }
    }
    void close_device() {
        if (dev_) hrnn_dev_close(dev_);
        dev_ = nullptr;
    }
    void dev_check(int rc) const {
        if (rc != HRNN_OK) throw std::runtime_error(std::string("HRNN device: ") + hrnn_dev_last_error(dev_));
    }
    void sync_device_base() const {
        open_device();
        size_t n = cur_element_count;
        if (dev_base_rows_ == n) return;
        if (dev_base_rows_ > 0 && dev_base_rows_ < n && dev_knng_ok_) {
            dev_check(hrnn_dev_append_base(dev_, vecs_.data() + dev_base_rows_ * dim_, n - dev_base_rows_));
        } else {
            dev_check(hrnn_dev_set_base(dev_, vecs_.data(), n, (uint32_t)dim_));
            dev_knng_ok_ = false;
        }
        dev_base_rows_ = n;
        dev_csr_ok_ = false;
        dev_hnsw_rows_ = 0;   // the device drops its HNSW copy when the base changes
    }
    void sync_device_knng() const {
        if (dev_knng_ok_) return;
        size_t n = cur_element_count;
        if (serve_mode_) {
            std::vector<uint32_t> zid(n * max_k_serve_, 0);
            std::vector<uint16_t> c(n, (uint16_t)max_k_serve_);
            dev_check(hrnn_dev_set_knng(dev_, (uint32_t)max_k_serve_, zid.data(), kdist_.data(), c.data()));
        } else {
            dev_check(hrnn_dev_set_knng(dev_, (uint32_t)K_knng_, knn_ids_.data(), knn_d_.data(), knn_cnt_.data()));
        }
        dev_knng_ok_ = true;
        dev_csr_ok_ = false;
    }
    // Upload the HNSW graph for phase A on the GPU (after any insertion).
    void sync_device_hnsw() const {
        size_t n = cur_element_count;
        if (dev_hnsw_rows_ == n) return;   // every insertion changes n
        std::vector<uint32_t> off, pool;
        std::vector<int> lv;
        exportUpper(off, pool, lv);
        dev_check(hrnn_dev_set_hnsw(dev_, (uint32_t)maxM0_, (uint32_t)maxM_, level0_.data(), lv.data(),
                                    off.data(), pool.data(), pool.size(), enterpoint_node_, maxlevel_));
        dev_hnsw_rows_ = n;
    }

 public:
    // Upper layers packed for the device: node i of level L >= 1 owns L
    // blocks of maxM + 1 words at pool[off[i]] (off = 0xFFFFFFFF at level 0).
    void exportUpper(std::vector<uint32_t> &off, std::vector<uint32_t> &pool, std::vector<int> &lv) const {
        size_t n = cur_element_count;
        const size_t wu = maxM_ + 1;
        off.assign(n, INVALID_KNNG_ID);
        pool.clear();
        lv.assign(levels_.begin(), levels_.begin() + n);
        size_t total = 0;
        for (size_t i = 0; i < n; i++) total += (size_t)std::max(0, levels_[i]) * wu;
        pool.reserve(total);
        for (size_t i = 0; i < n; i++) {
            if (levels_[i] <= 0) continue;
            off[i] = (uint32_t)pool.size();
            const uint32_t *b = reinterpret_cast<const uint32_t *>(upper_[i]);
            pool.insert(pool.end(), b, b + (size_t)levels_[i] * wu);
        }
    }
    tableint entryPoint() const { return enterpoint_node_; }
    int maxLevel() const { return maxlevel_; }

 private:
    void sync_device_csr() const {
        if (dev_csr_ok_) return;
        size_t n = cur_element_count;
        if (rknng_wide_)
            dev_check(hrnn_dev_set_rknn_csr(dev_, rknng_rank_bits_, 1, n, rknng_offsets_.data(),
                                            rknng_entries_64_.data(), rknng_entries_64_.size()));
        else
            dev_check(hrnn_dev_set_rknn_csr(dev_, rknng_rank_bits_, 0, n, rknng_offsets_.data(),
                                            rknng_entries_32_.data(), rknng_entries_32_.size()));
        dev_csr_ok_ = true;
    }
    void device_search(const float *Q, size_t nq, const uint32_t *prox, size_t m,
                       const uint32_t *k_per_q, uint32_t k, uint32_t Kp, float thr,
                       std::vector<uint64_t> &off, std::vector<uint32_t> &ids) const {
        std::lock_guard<std::mutex> g(dev_mu_);
        sync_device_base();
        sync_device_knng();
        sync_device_csr();
        off.assign(nq + 1, 0);
        uint64_t cap = std::max<uint64_t>(1024, nq * 64), need = 0;
        for (;;) {
            ids.resize(cap);
            int rc = hrnn_dev_rknn_search(dev_, Q, (uint32_t)nq, prox, (uint32_t)m, k_per_q, k, Kp,
                                          thr, off.data(), ids.data(), cap, &need);
            if (rc == HRNN_E_CAP && need > cap) {
                cap = need;
                continue;
            }
            dev_check(rc);
            break;
        }
        ids.resize(off[nq]);
    }
};

}  // namespace hnswlib
