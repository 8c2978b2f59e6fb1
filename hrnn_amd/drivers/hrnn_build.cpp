// hrnn_build -- build an HRNN index (CLI of the reference's src/hrnn_build.cpp:
// same flags, defaults and report lines).  Stages: [1] load fbin, [2] HNSW +
// KNN seeding (GPU, addPointsDevice; --hnsw_device 0: host addPoint under
// OpenMP), [3] refineKNNG (GPU), [4] buildRKNNG (GPU), [5] saveIndex
// (reference file format).
#include <omp.h>

#include <atomic>
#include <iostream>
#include <memory>

#include "cli.h"
#include "hnswlib/hrnn.h"
#include "utils.h"

int main(int argc, char **argv) {
    Cli cli("HRNN Build Options");
    cli.add("base", "", "Path to base vectors (.fbin)", true);
    cli.add("index_path", "", "Index path to save", true);
    cli.add("M", "16", "HNSW M parameter");
    cli.add("ef_construction", "200", "HNSW ef_construction parameter");
    cli.add("K_knng", "500", "KNNG depth (K nearest neighbors to store per node)");
    cli.add("nn_iters", "3", "NNDescent refinement iterations");
    cli.add("nn_sample", "20", "NNDescent sample size per node per iteration");
    cli.add("num_threads", "64", "Number of threads");
    cli.add("hnsw_device", "1", "1 = HNSW construction on the GPU, 0 = addPoint on the host threads");
    try {
        if (!cli.parse(argc, argv)) return 0;
        const int M = cli.i("M"), efc = cli.i("ef_construction"), K = cli.i("K_knng");
        const int iters = cli.i("nn_iters"), sample = cli.i("nn_sample"), threads = cli.i("num_threads");
        std::cout << "=== HRNN Index Build ===" << std::endl;
        std::cout << "  M=" << M << "  ef_construction=" << efc << "  K_knng=" << K << std::endl;
        std::cout << "  nn_iters=" << iters << "  nn_sample=" << sample << std::endl;
        omp_set_num_threads(threads);
        std::cout << "  OpenMP threads: " << omp_get_max_threads() << std::endl;

        std::cout << "\n[1] Loading base  " << cli.str("base") << std::endl;
        auto t0 = Clock::now();
        FbinData base = load_fbin(cli.str("base"));
        std::cout << "      n=" << base.n << "  d=" << base.d << "  (" << elapsed_ms(t0) << " ms)" << std::endl;

        std::cout << "\n[2] Building HNSW + initial KNNG (addPoint saves candidates to KNNG slot) ..." << std::endl;
        hnswlib::L2Space space(base.d);
        auto index = std::make_unique<hnswlib::HRNN<float>>(&space, base.n, M, efc, 100, false, (size_t)K);
        t0 = Clock::now();
        const int total = (int)base.n;
        if (cli.i("hnsw_device") != 0) {
            index->addPointsDevice(base.data.data(), base.n);
            const hrnn_hnsw_build_stats &bs = index->lastBuildStats();
            std::cout << "      HNSW on GPU: " << total << " points in " << bs.batches << " batches, "
                      << bs.expansions << " expansions, " << bs.evaluations << " distances" << std::endl;
        } else {
            std::atomic<int> progress{0};
            const int every = std::max(1, total / 20);
#pragma omp parallel for schedule(dynamic, 512)
            for (int i = 0; i < total; ++i) {
                index->addPoint(base.data.data() + (size_t)i * base.d, (hnswlib::labeltype)i);
                int done = progress.fetch_add(1, std::memory_order_relaxed) + 1;
                if (done % every == 0 || done == total) {
#pragma omp critical
                    std::cout << "\r      HNSW addPoint: " << done << " / " << total << " ("
                              << (100LL * done / total) << "%)  [" << elapsed_ms(t0) << " ms]" << std::flush;
                }
            }
            std::cout << std::endl;
        }
        std::cout << "      HNSW + initial KNNG done  (" << elapsed_ms(t0) << " ms)" << std::endl;

        std::cout << "\n[3] Refining KNNG (NNDescent local join, " << iters << " iterations) ..." << std::endl;
        t0 = Clock::now();
        index->refineKNNG(iters, sample);
        std::cout << "      Refine done  (" << elapsed_ms(t0) << " ms)" << std::endl;

        std::cout << "\n[4] Building RKNNG (CSR transpose) ..." << std::endl;
        t0 = Clock::now();
        index->buildRKNNG();
        std::cout << "      RKNNG done  (" << elapsed_ms(t0) << " ms)" << std::endl;

        std::cout << "\n[5] Saving index to " << cli.str("index_path") << std::endl;
        t0 = Clock::now();
        index->saveIndex(cli.str("index_path"));
        std::cout << "      Saved  (" << elapsed_ms(t0) << " ms)" << std::endl;
    } catch (std::exception &e) {
        std::cerr << "error: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
