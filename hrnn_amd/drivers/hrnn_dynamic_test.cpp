// hrnn_dynamic_test -- streaming-insert harness (CLI of the reference's
// src/hrnn_dynamic_test.cpp: same flags, defaults, split and report lines).
// Initial split = first n*split ids of a mt19937(42) permutation; build (or
// load a snapshot; HNSW + KNN seeds on the GPU unless --hnsw_device 0) ->
// insertion of the rest: on the GPU in batches (insertPointsDynamicDevice;
// --insert_device 0: mutable reverse lists and insertPointDynamic per point on
// the host threads) -> flat buildRKNNG (GPU) -> searchRknn evaluation.
#include <omp.h>

#include <algorithm>
#include <atomic>
#include <filesystem>
#include <iomanip>
#include <iostream>
#include <memory>
#include <numeric>
#include <random>

#include "cli.h"
#include "hnswlib/hrnn.h"
#include "utils.h"

using HRNN = hnswlib::HRNN<float>;

struct Eval {
    double macro = 0, micro = 0, avg = 0;
    long ms = 0;
};

static Eval evaluate(HRNN &index, const FbinData &query, const RknnGroundtruth &gt, int m, int k, int ef) {
    index.setEf((size_t)ef);
    auto t0 = Clock::now();
    auto res = index.searchRknnBatch(query.data.data(), query.n, (size_t)m, 1.0f, (size_t)k,
                                     index.K_knng_, (size_t)ef);
    Eval e;
    e.ms = elapsed_ms(t0);
    e.macro = compute_rknn_recall(gt, res);
    e.micro = compute_rknn_micro_recall(gt, res);
    for (auto &r : res) e.avg += (double)r.size();
    e.avg /= std::max<double>(1, query.n);
    return e;
}

int main(int argc, char **argv) {
    Cli cli("HRNN Dynamic Update Test");
    cli.add("base", "", "Path to base vectors (.fbin)", true);
    cli.add("query", "", "Path to query vectors (.fbin)", true);
    cli.add("gt", "", "Path to RkNN groundtruth (.bin)", true);
    cli.add("snapshot", "", "Snapshot index path (load if present, else build and save)");
    cli.add("oracle_snapshot", "", "Optional full-build snapshot for recall-gap comparison");
    cli.add("batch_size", "0", "Inserted batch size. 0 = all points after split");
    cli.add("M", "16", "HNSW M parameter");
    cli.add("ef_construction", "200", "HNSW ef_construction");
    cli.add("K_knng", "500", "KNNG depth");
    cli.add("nn_iters", "3", "NNDescent iterations for the initial build");
    cli.add("nn_sample", "20", "NNDescent sample size");
    cli.add("num_threads", "64", "Number of threads");
    cli.add("k", "10", "RkNN k parameter");
    cli.add("m", "50", "Proxy count for search");
    cli.add("ef_search", "200", "HNSW ef_search");
    cli.add("split", "0.5", "Fraction of data for the initial build");
    cli.add("m_update", "0", "Proxy count for the insert self-query (0 = all KNNG neighbours)");
    cli.add("bidirectional", "0", "1 = also improve the new point's KNNG from the self-query");
    cli.add("hnsw_device", "1", "1 = initial HNSW construction on the GPU, 0 = addPoint on the host threads");
    cli.add("insert_device", "1", "1 = insertion on the GPU in batches, 0 = insertPointDynamic per point on the host threads");
    try {
        if (!cli.parse(argc, argv)) return 0;
        const int M = cli.i("M"), efc = cli.i("ef_construction"), K = cli.i("K_knng");
        const int k = cli.i("k"), m = cli.i("m"), ef = cli.i("ef_search");
        const int m_update = cli.i("m_update"), bidir = cli.i("bidirectional");
        const double split = cli.f("split");
        std::cout << "=== HRNN Dynamic Update Test ===" << std::endl;
        std::cout << "  M=" << M << "  ef_construction=" << efc << "  K_knng=" << K << std::endl;
        std::cout << "  split=" << split << "  k=" << k << "  m=" << m << "  ef_search=" << ef
                  << "  m_update=" << m_update << "  bidirectional=" << bidir << std::endl;
        omp_set_num_threads(cli.i("num_threads"));
        std::cout << "  OpenMP threads: " << omp_get_max_threads() << std::endl;

        std::cout << "\n[1] Loading data" << std::endl;
        auto t0 = Clock::now();
        FbinData base = load_fbin(cli.str("base"));
        std::cout << "      base: n=" << base.n << "  d=" << base.d << "  (" << elapsed_ms(t0) << " ms)" << std::endl;
        FbinData query = load_fbin(cli.str("query"));
        std::cout << "      query: n=" << query.n << "  d=" << query.d << std::endl;
        RknnGroundtruth gt = load_rknn_groundtruth(cli.str("gt"));
        std::cout << "      GT: nq=" << gt.nq << "  total=" << gt.total() << std::endl;

        const size_t n = base.n;
        const size_t split_n = (size_t)((double)n * split);
        std::vector<uint32_t> perm(n);
        std::iota(perm.begin(), perm.end(), 0u);
        std::mt19937 rng(42);
        std::shuffle(perm.begin(), perm.end(), rng);
        const size_t rest = n - split_n;
        const size_t batch_n = cli.i("batch_size") > 0 ? std::min<size_t>((size_t)cli.i("batch_size"), rest) : rest;
        std::cout << "\n[2] Split: initial=" << split_n << "  incremental_total=" << rest
                  << "  active_batch=" << batch_n << std::endl;

        hnswlib::L2Space space(base.d);
        std::unique_ptr<HRNN> index;
        long build_ms = 0;
        const std::string snap = cli.str("snapshot");
        if (!snap.empty() && std::filesystem::exists(snap)) {
            std::cout << "\n[3] Loading snapshot: " << snap << std::endl;
            t0 = Clock::now();
            index = std::make_unique<HRNN>(&space, snap, false, n);
            long load_ms = elapsed_ms(t0);
            std::cout << "      Loaded (" << load_ms << " ms)  cur_elements=" << index->getCurrentElementCount()
                      << "  K_knng=" << index->K_knng_ << "  knng_built=" << index->knng_built_ << std::endl;
            t0 = Clock::now();
            if (!cli.i("insert_device")) index->buildRKNNG_mutable();
            build_ms = load_ms + elapsed_ms(t0);
        } else {
            std::cout << "\n[3] Building HRNN on initial " << split_n << " points" << std::endl;
            index = std::make_unique<HRNN>(&space, n, M, efc, 100, false, (size_t)K);
            t0 = Clock::now();
            if (cli.i("hnsw_device")) {
                std::vector<float> xs(split_n * base.d);
                std::vector<hnswlib::labeltype> ls(split_n);
                for (size_t i = 0; i < split_n; ++i) {
                    std::copy_n(base.data.data() + (size_t)perm[i] * base.d, base.d, xs.data() + i * base.d);
                    ls[i] = perm[i];
                }
                index->addPointsDevice(xs.data(), split_n, ls.data());
            } else {
#pragma omp parallel for schedule(dynamic, 512)
                for (long long i = 0; i < (long long)split_n; ++i)
                    index->addPoint(base.data.data() + (size_t)perm[i] * base.d, (hnswlib::labeltype)perm[i]);
            }
            long hnsw_ms = elapsed_ms(t0);
            std::cout << "      HNSW + initial KNNG: " << hnsw_ms << " ms" << std::endl;
            std::cout << "\n[3b] NNDescent refinement (" << cli.i("nn_iters") << " iters)" << std::endl;
            t0 = Clock::now();
            index->refineKNNG(cli.i("nn_iters"), cli.i("nn_sample"));
            long refine_ms = elapsed_ms(t0);
            std::cout << "      Refine: " << refine_ms << " ms" << std::endl;
            if (!snap.empty()) {
                std::cout << "\n[3s] Saving snapshot to " << snap << std::endl;
                index->saveIndex(snap);
            }
            long rk_ms = 0;
            if (!cli.i("insert_device")) {
                std::cout << "\n[3c] Building mutable RKNNG" << std::endl;
                t0 = Clock::now();
                index->buildRKNNG_mutable();
                rk_ms = elapsed_ms(t0);
                std::cout << "      RKNNG: " << rk_ms << " ms" << std::endl;
            }
            build_ms = hnsw_ms + refine_ms + rk_ms;
        }
        std::cout << "\n      Total initial build/load: " << build_ms << " ms" << std::endl;

        std::cout << "\n[4] Inserting " << batch_n << " points incrementally" << std::endl;
        std::vector<long> per_us;
        t0 = Clock::now();
        if (cli.i("insert_device")) {
            // one batched call: per-point latency is the amortised time
            std::vector<float> xs(batch_n * base.d);
            std::vector<hnswlib::labeltype> ls(batch_n);
            for (size_t i = 0; i < batch_n; ++i) {
                std::copy_n(base.data.data() + (size_t)perm[split_n + i] * base.d, base.d, xs.data() + i * base.d);
                ls[i] = perm[split_n + i];
            }
            index->insertPointsDynamicDevice(xs.data(), batch_n, ls.data(), (unsigned)m_update, bidir != 0);
            const auto &st = index->lastInsertStats();
            std::cout << "      GPU insertion: " << st.batches << " batches, " << st.candidates
                      << " candidates, " << st.offers << " offers, " << st.updates << " updates  (HNSW "
                      << (long)st.hnsw_ms << " ms, reverse lists " << (long)st.csr_ms << " ms, scan "
                      << (long)st.affect_ms << " ms, merge " << (long)st.merge_ms << " ms)" << std::endl;
        } else {
            per_us.assign(batch_n, 0);
#pragma omp parallel for schedule(dynamic, 64)
            for (long long i = 0; i < (long long)batch_n; ++i) {
                uint32_t id = perm[split_n + i];
                auto ts = Clock::now();
                index->insertPointDynamic(base.data.data() + (size_t)id * base.d, (hnswlib::labeltype)id,
                                          (unsigned)m_update, bidir != 0);
                per_us[i] = elapsed_us(ts);
            }
        }
        long maint_ms = elapsed_ms(t0);
        double per_ms = batch_n ? (double)maint_ms / (double)batch_n : 0.0;
        std::cout << "      Total insert: " << maint_ms << " ms" << std::endl;
        if (!per_us.empty()) {
            std::vector<long> s = per_us;
            std::sort(s.begin(), s.end());
            auto pct = [&](double p) { return s[std::min(s.size() - 1, (size_t)(p * (double)s.size()))]; };
            double mean = (double)std::accumulate(s.begin(), s.end(), 0LL) / (double)s.size();
            std::cout << "      Per-insert wall (measured, us):  mean=" << std::fixed << std::setprecision(1)
                      << mean << "  p50=" << pct(0.5) << "  p95=" << pct(0.95) << "  p99=" << pct(0.99)
                      << "  max=" << s.back() << std::endl;
        }
        std::cout << "      Per-insert: " << std::fixed << std::setprecision(3) << per_ms << " ms" << std::endl;
        std::cout << "      Index size after maintenance: " << index->getCurrentElementCount() << std::endl;

        std::cout << "\n[5] Searching (m=" << m << ", threshold=1, ef_search=" << ef << ")" << std::endl;
        t0 = Clock::now();
        index->buildRKNNG();
        std::cout << "      Flat RKNNG for search: " << elapsed_ms(t0) << " ms" << std::endl;
        Eval e = evaluate(*index, query, gt, m, k, ef);
        std::cout << std::fixed << std::setprecision(6);
        std::cout << "      Macro Recall = " << e.macro << std::endl;
        std::cout << "      Micro Recall = " << e.micro << std::endl;
        std::cout << std::fixed << std::setprecision(2);
        std::cout << "      Avg results/query = " << e.avg << std::endl;
        std::cout << "      Search time = " << e.ms << " ms ("
                  << (double)query.n * 1000.0 / (double)std::max<long>(1, e.ms) << " QPS)" << std::endl;

        bool has_oracle = false;
        Eval oe;
        const std::string osnap = cli.str("oracle_snapshot");
        if (!osnap.empty()) {
            std::cout << "\n[6] Loading oracle snapshot: " << osnap << std::endl;
            auto oracle = std::make_unique<HRNN>(&space, osnap, false, n);
            oracle->buildRKNNG();
            oe = evaluate(*oracle, query, gt, m, k, ef);
            has_oracle = true;
            std::cout << std::fixed << std::setprecision(6);
            std::cout << "      Oracle Macro Recall = " << oe.macro << std::endl;
            std::cout << "      Oracle Micro Recall = " << oe.micro << std::endl;
        }
        std::cout << "\n=== Summary ===" << std::endl;
        std::cout << "  Initial build/load (" << split_n << " pts): " << build_ms << " ms" << std::endl;
        std::cout << "  Maintained batch size: " << batch_n << std::endl;
        std::cout << "  Maintenance time: " << maint_ms << " ms" << std::endl;
        std::cout << "  Per-insert latency: " << std::fixed << std::setprecision(3) << per_ms << " ms" << std::endl;
        std::cout << "  Macro Recall = " << std::fixed << std::setprecision(6) << e.macro << std::endl;
        std::cout << "  Micro Recall = " << std::fixed << std::setprecision(6) << e.micro << std::endl;
        if (has_oracle) {
            std::cout << "  Oracle macro gap = " << (e.macro - oe.macro) << std::endl;
            std::cout << "  Oracle micro gap = " << (e.micro - oe.micro) << std::endl;
        }
    } catch (std::exception &ex) {
        std::cerr << "error: " << ex.what() << std::endl;
        return 1;
    }
    return 0;
}
