// hrnn_gt -- exact RkNN groundtruth on the GPU (replaces the reference's
// scripts/gen_rknn_gt.py: same --base/--query/--k/--output flags and output
// format, gen_rknn_gt.py:154-168).  r_k(x) is the (k+1)-th smallest distance
// over all base points (self included, gen_rknn_gt.py:62-75); RkNN(q) =
// {x : d2(q,x) <= r_k(x)} sorted by distance.  Distances are FP32 direct
// differences in the reference's summation order; --precision picks the
// candidate pass (bf16 MFMA is exact for integer data in 0..255).
#include <cstring>
#include <iostream>
#include <vector>

#include "cli.h"
#include "hrnn_dev.h"
#include "utils.h"

static void check(hrnn_dev_t h, int rc, const char *what) {
    if (rc != HRNN_OK) throw std::runtime_error(std::string(what) + ": " + hrnn_dev_last_error(h));
}

int main(int argc, char **argv) {
    Cli cli("Generate RkNN groundtruth (GPU)");
    cli.add("base", "", "Base vectors (.fbin)", true);
    cli.add("query", "", "Query vectors (.fbin)", true);
    cli.add("k", "", "k for RkNN", true);
    cli.add("output", "", "Output groundtruth file", true);
    cli.add("precision", "bf16", "candidate pass: bf16 | f16 | f32");
    cli.add("device", "0", "GPU ordinal");
    try {
        if (!cli.parse(argc, argv)) return 0;
        auto t0 = Clock::now();
        FbinData base = load_fbin(cli.str("base"));
        FbinData query = load_fbin(cli.str("query"));
        if (base.d != query.d) throw std::runtime_error("Dimension mismatch");
        std::cout << "  base n=" << base.n << " d=" << base.d << "  query n=" << query.n << std::endl;
        std::string p = cli.str("precision");
        int prec = p == "f32" ? HRNN_PREC_F32 : p == "f16" ? HRNN_PREC_F16 : HRNN_PREC_BF16;
        hrnn_dev_t h = nullptr;
        if (hrnn_dev_open(cli.i("device"), &h) != HRNN_OK) throw std::runtime_error("no usable GPU");
        check(h, hrnn_dev_set_base(h, base.data.data(), base.n, base.d), "set_base");
        std::vector<float> kth(base.n);
        check(h, hrnn_dev_bf_kth(h, (uint32_t)cli.i("k"), prec, kth.data(), nullptr), "bf_kth");
        std::cout << "  r_k done (" << elapsed_ms(t0) << " ms)" << std::endl;
        RknnGroundtruth gt;
        gt.nq = query.n;
        std::vector<uint64_t> off(query.n + 1);
        uint64_t need = 0;
        int rc = hrnn_dev_bf_rknn(h, query.data.data(), query.n, kth.data(), prec, off.data(), nullptr,
                                  nullptr, 0, &need);
        if (rc != HRNN_OK && rc != HRNN_E_CAP) check(h, rc, "bf_rknn");
        gt.ids.resize(need);
        gt.dists.resize(need);
        check(h, hrnn_dev_bf_rknn(h, query.data.data(), query.n, kth.data(), prec, off.data(),
                                  gt.ids.data(), gt.dists.data(), need, &need), "bf_rknn");
        gt.offsets.resize(query.n + 1);
        for (size_t i = 0; i <= query.n; i++) gt.offsets[i] = (uint32_t)off[i];
        save_rknn_groundtruth(cli.str("output"), gt);
        std::cout << "  Saved " << cli.str("output") << ": nq=" << gt.nq << ", total_results=" << gt.total()
                  << "  (" << elapsed_ms(t0) << " ms)" << std::endl;
        hrnn_dev_close(h);
    } catch (std::exception &e) {
        std::cerr << "error: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
