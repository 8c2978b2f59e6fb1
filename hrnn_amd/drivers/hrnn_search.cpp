// hrnn_search -- RkNN search sweep over (m, threshold) (CLI of the reference's
// src/hrnn_search.cpp: same flags, defaults and report lines).  Queries run one
// at a time through searchRknn, as the reference times them; --batch 1 runs
// each sweep point as one batched call instead (phase A on host threads, one
// GPU launch for phases B and C) and reports batch throughput.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <iomanip>
#include <iostream>
#include <memory>
#include <sstream>

#include "cli.h"
#include "hnswlib/hrnn.h"
#include "utils.h"

struct LatencyStats {
    double mean = 0, p50 = 0, p95 = 0, p99 = 0, max = 0;
};

static LatencyStats latency_stats(std::vector<double> t) {
    LatencyStats s;
    if (t.empty()) return s;
    std::sort(t.begin(), t.end());
    double sum = 0;
    for (double v : t) sum += v;
    s.mean = sum / (double)t.size();
    auto pct = [&](double p) {   // linear interpolation between closest ranks
        double r = p * (double)(t.size() - 1);
        size_t lo = (size_t)std::floor(r), hi = (size_t)std::ceil(r);
        return t[lo] + (t[hi] - t[lo]) * (r - (double)lo);
    };
    s.p50 = pct(0.5);
    s.p95 = pct(0.95);
    s.p99 = pct(0.99);
    s.max = t.back();
    return s;
}

template <class T>
static std::vector<T> parse_list(const std::string &s) {
    std::vector<T> out;
    std::stringstream ss(s);
    std::string item;
    while (std::getline(ss, item, ','))
        if (!item.empty()) out.push_back((T)std::stod(item));
    return out;
}

int main(int argc, char **argv) {
    Cli cli("HRNN Search Options");
    cli.add("index_path", "", "HRNN index path", true);
    cli.add("query", "", "Query vectors (.fbin)", true);
    cli.add("gt", "", "RkNN groundtruth file", true);
    cli.add("k", "10", "RkNN k parameter");
    cli.add("K_prime", "500", "KNNG K' truncation (use first K' reverse entries)");
    cli.add("m", "1,3,5,10,20,50", "Proxy counts to sweep (comma-separated)");
    cli.add("threshold", "1", "Binary score thresholds (comma-separated)");
    cli.add("ef_search", "200", "HNSW ef_search");
    cli.add("max_queries", "0", "0=all");
    cli.add("max_k_serve", "0", "Enable serve mode: dense kdist matrix for k<=max_k_serve (0=off)");
    cli.add("batch", "0", "1 = one batched call per sweep point");
    cli.add("num_threads", "0", "host threads for batched phase A (0 = OMP default)");
    try {
        if (!cli.parse(argc, argv)) return 0;
        const int k = cli.i("k"), Kp = cli.i("K_prime"), ef = cli.i("ef_search");
        auto ms = parse_list<int>(cli.str("m"));
        if (ms.empty()) ms.push_back(20);
        auto ths = parse_list<float>(cli.str("threshold"));
        if (ths.empty()) ths.push_back(1.0f);
        std::cout << "=== HRNN RkNN Search ===" << std::endl;
        std::cout << "  k=" << k << "  K'=" << Kp << "  ef_search=" << ef << std::endl;
        std::cout << "  m values:";
        for (int m : ms) std::cout << " " << m;
        std::cout << std::endl << "  thresholds:";
        for (float t : ths) std::cout << " " << t;
        std::cout << std::endl;

        std::cout << "\n[1] Loading queries " << cli.str("query") << std::endl;
        FbinData query = load_fbin(cli.str("query"));
        std::cout << "      nq=" << query.n << "  d=" << query.d << std::endl;
        std::cout << "[2] Loading RkNN GT " << cli.str("gt") << std::endl;
        RknnGroundtruth gt = load_rknn_groundtruth(cli.str("gt"));
        std::cout << "      nq=" << gt.nq << "  total=" << gt.total() << std::endl;
        std::cout << "[3] Loading HRNN index " << cli.str("index_path") << std::endl;
        auto t0 = Clock::now();
        hnswlib::L2Space space(query.d);
        auto index = std::make_unique<hnswlib::HRNN<float>>(&space, cli.str("index_path"));
        index->setEf((size_t)ef);
        std::cout << "      Done (" << elapsed_ms(t0) << " ms)" << std::endl;
        std::cout << "      K_knng=" << index->K_knng_ << "  knng_built=" << index->knng_built_
                  << "  rknng_built=" << index->rknng_built_ << std::endl;
        if (cli.i("max_k_serve") > 0) {
            std::cout << "[3b] Compacting for serve mode (max_k_serve=" << cli.i("max_k_serve") << ")" << std::endl;
            auto tc = Clock::now();
            index->compactForServing((size_t)cli.i("max_k_serve"));
            std::cout << "      Done (" << elapsed_ms(tc) << " ms)" << std::endl;
        }
        size_t nq = query.n;
        if (cli.i("max_queries") > 0) nq = std::min<size_t>((size_t)cli.i("max_queries"), query.n);
        const bool batch = cli.i("batch") != 0;
        index->syncDevice();   // upload base, graph and CSR before timing

        for (int m : ms)
            for (float thr : ths) {
                std::cout << "\n[4] Search: nq=" << nq << "  m=" << m << "  threshold=" << thr << std::endl;
                std::vector<std::vector<hnswlib::labeltype>> res(nq);
                std::vector<double> times(nq, 0.0);
                double wall_ms = 0;
                if (batch) {
                    auto ta = Clock::now();
                    res = index->searchRknnBatch(query.data.data(), nq, (size_t)m, thr, (size_t)k,
                                                 (size_t)Kp, (size_t)ef, cli.i("num_threads"));
                    wall_ms = std::chrono::duration<double, std::milli>(Clock::now() - ta).count();
                    for (auto &t : times) t = wall_ms / (double)nq;
                } else {
                    for (size_t q = 0; q < nq; q++) {
                        auto ta = Clock::now();
                        res[q] = index->searchRknn(query.data.data() + q * query.d, (size_t)m, thr,
                                                   (size_t)k, (size_t)Kp, (size_t)ef);
                        times[q] = std::chrono::duration<double, std::milli>(Clock::now() - ta).count();
                    }
                }
                if (nq == query.n) {
                    std::cout << "      Macro Recall = " << std::fixed << std::setprecision(6)
                              << compute_rknn_recall(gt, res) << std::endl;
                    std::cout << "      Micro Recall = " << std::fixed << std::setprecision(6)
                              << compute_rknn_micro_recall(gt, res) << std::endl;
                }
                double avg = 0;
                for (auto &r : res) avg += (double)r.size();
                std::cout << "      Avg results/query = " << std::fixed << std::setprecision(2)
                          << avg / (double)nq << std::endl;
                LatencyStats ls = latency_stats(times);
                std::cout << std::fixed << std::setprecision(3);
                std::cout << "      Latency (ms): mean=" << ls.mean << "  p50=" << ls.p50 << "  p95=" << ls.p95
                          << "  p99=" << ls.p99 << "  max=" << ls.max << std::endl;
                double total = 0;
                for (double t : times) total += t;
                std::cout << "      QPS = " << std::fixed << std::setprecision(2)
                          << (double)nq * 1000.0 / total << std::endl;
            }
    } catch (std::exception &e) {
        std::cerr << "error: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
