"""The drop-in boundary on the CPU: the reference's own drivers compile
unchanged against include/ (reference src/hrnn_build.cpp:1-2 includes
"hnswlib/hrnn.h" and "utils.h"; src/*.cpp use the `po` namespace of
include/utils.h:20-21), the `po` subset parses command lines as Boost's
program_options does, and the file formats of include/utils.h (fbin, kNN and
RkNN groundtruth, reference utils.h:31-108) and its recall definitions
(:115-157) read and write the bytes the reference's files hold."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
LIB = os.path.join(ROOT, "hrnn_amd", "lib")
REF_SRC = "/root/reference/src"


def _compile(tmp_path, src_text_or_path, name, extra=()):
    if os.path.exists(str(src_text_or_path)):
        src = str(src_text_or_path)
    else:
        src = str(tmp_path / (name + ".cpp"))
        with open(src, "w") as f:
            f.write(src_text_or_path)
    out = str(tmp_path / name)
    cmd = ["g++", "-O1", "-std=gnu++20", "-fopenmp", "-I" + INC, src, "-o", out] + list(extra)
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return out


PO_PROBE = r'''
#include "utils.h"
#include <iostream>
int main(int argc, char **argv) {
    std::string s; int i; double f;
    po::options_description desc("Probe Options");
    desc.add_options()("help,h", "Show help")
        ("name", po::value<std::string>(&s)->required(), "a string")
        ("count,c", po::value<int>(&i)->default_value(7), "an int")
        ("ratio", po::value<double>(&f)->default_value(0.5), "a double");
    try {
        po::variables_map vm;
        po::store(po::parse_command_line(argc, argv, desc), vm);
        if (vm.count("help")) { std::cout << desc; return 0; }
        po::notify(vm);
        std::cout << s << "|" << i << "|" << f << "|" << vm.count("count") << std::endl;
    } catch (po::error &e) {
        std::cout << "ERR " << e.what() << std::endl;
        return 3;
    }
    return 0;
}
'''


def test_po_subset_parses_like_program_options(tmp_path):
    exe = _compile(tmp_path, PO_PROBE, "po_probe")
    run = lambda *a: subprocess.run([exe, *a], capture_output=True, text=True)
    r = run("--name", "abc")
    assert r.returncode == 0 and r.stdout.strip() == "abc|7|0.5|1"      # defaults count as present
    r = run("--name=x y", "-c", "12", "--ratio", "2.25")
    assert r.stdout.strip() == "x y|12|2.25|1"
    r = run("--name", "q", "-c3")
    assert r.stdout.strip() == "q|3|0.5|1"
    r = run("--name", "a", "--name", "b")                                 # the last occurrence wins
    assert r.stdout.strip() == "b|7|0.5|1"
    r = run("-c", "4")
    assert r.returncode == 3 and "'--name' is required" in r.stdout
    r = run("--name", "a", "--count", "4x")
    assert r.returncode == 3 and "invalid" in r.stdout
    r = run("--name", "a", "--bogus", "1")
    assert r.returncode == 3 and "unrecognised option '--bogus'" in r.stdout
    r = run("--name")
    assert r.returncode == 3 and "missing" in r.stdout
    r = run("--help")
    assert r.returncode == 0
    assert "Probe Options:" in r.stdout and "-c [ --count ] arg (=7)" in r.stdout
    assert "--name arg" in r.stdout and "-h [ --help ]" in r.stdout


@pytest.mark.skipif(not os.path.isdir(REF_SRC), reason="reference sources absent")
@pytest.mark.parametrize("driver,required", [("hrnn_build", "base"), ("hrnn_search", "index_path"),
                                             ("hrnn_dynamic_test", "base")])
def test_reference_drivers_compile_unchanged(tmp_path, driver, required):
    """The reference's driver source, as it lies, against this include/ and
    libhrnn_hip.so; its option table prints, and a missing required option
    stops it before any index work (no GPU needed for either)."""
    if not os.path.exists(os.path.join(LIB, "libhrnn_hip.so")):
        pytest.skip("libhrnn_hip.so not built")
    exe = _compile(tmp_path, os.path.join(REF_SRC, driver + ".cpp"), driver,
                   ["-L" + LIB, "-lhrnn_hip", "-Wl,-rpath," + LIB])
    r = subprocess.run([exe, "--help"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    assert "--" + required + " arg" in r.stdout
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode != 0 and f"'--{required}' is required" in r.stderr


IO_PROBE = r'''
#include "utils.h"
#include <cstdio>
#include <iostream>
int main(int argc, char **argv) {
    std::string dir = argv[1];
    FbinData fb = load_fbin(dir + "/x.fbin");
    double s = 0;
    for (float v : fb.data) s += v;
    std::printf("fbin %u %u %.6f\n", fb.n, fb.d, s);
    save_fbin(dir + "/x2.fbin", fb.data.data(), fb.n, fb.d);
    RknnGroundtruth gt = load_rknn_groundtruth(dir + "/gt.bin");
    std::printf("gt %u %zu\n", gt.nq, gt.total());
    save_rknn_groundtruth(dir + "/gt2.bin", gt);
    KnnGroundtruth kg = load_knn_groundtruth(dir + "/knn.bin");
    // results: query q answers with the ids listed in res.txt line q
    std::vector<std::vector<hnswlib::labeltype>> res(gt.nq);
    FILE *f = std::fopen((dir + "/res.txt").c_str(), "r");
    for (uint32_t q = 0; q < gt.nq; q++) {
        unsigned c = 0;
        if (std::fscanf(f, "%u", &c) != 1) return 4;
        for (unsigned j = 0; j < c; j++) { unsigned long long v; if (std::fscanf(f, "%llu", &v) != 1) return 4; res[q].push_back(v); }
    }
    std::fclose(f);
    std::printf("macro %.9f\n", compute_rknn_recall(gt, res));
    std::printf("micro %.9f\n", compute_rknn_micro_recall(gt, res));
    std::printf("knn %u %u %.9f\n", kg.nq, kg.k, compute_knn_recall(kg, res, 3));
    return 0;
}
'''


def test_file_formats_and_recall(tmp_path):
    """fbin, RkNN GT and kNN GT files written here byte by byte in the
    reference's layouts are read by include/utils.h; the files it writes back
    are identical; macro recall skips queries with an empty GT, micro recall
    pools all hits, kNN recall counts the first min(k, gt.k) GT ids."""
    rng = np.random.default_rng(3)
    x = rng.integers(0, 256, (37, 16)).astype(np.float32)
    with open(tmp_path / "x.fbin", "wb") as f:
        f.write(np.array([37, 16], np.uint32).tobytes() + x.tobytes())
    nq = 5
    gt_ids = [[3, 1], [], [7, 8, 9, 10], [2], [5, 6]]
    off = np.cumsum([0] + [len(g) for g in gt_ids]).astype(np.uint32)
    ids = np.array(sum(gt_ids, []), np.uint32)
    dists = rng.random(ids.size).astype(np.float32)
    with open(tmp_path / "gt.bin", "wb") as f:
        f.write(np.array([nq], np.uint32).tobytes() + off.tobytes() + ids.tobytes() + dists.tobytes())
    kg = np.array([[3, 4, 5, 6], [0, 1, 2, 3], [7, 8, 9, 1], [2, 4, 6, 8], [5, 6, 7, 8]], np.uint32)
    with open(tmp_path / "knn.bin", "wb") as f:
        f.write(np.array([nq, 4], np.uint32).tobytes() + kg.tobytes() + np.zeros(kg.size, np.float32).tobytes())
    res = [[3, 4], [0], [7, 9, 11], [], [6, 5, 99]]
    with open(tmp_path / "res.txt", "w") as f:
        for r in res:
            f.write(" ".join(map(str, [len(r)] + r)) + "\n")
    exe = _compile(tmp_path, IO_PROBE, "io_probe")
    out = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True)
    assert out.returncode == 0, out.stderr
    lines = dict(l.split(" ", 1) for l in out.stdout.strip().splitlines())
    assert lines["fbin"].split()[:2] == ["37", "16"]
    assert float(lines["fbin"].split()[2]) == pytest.approx(float(x.sum(dtype=np.float64)), rel=1e-12)
    assert lines["gt"] == f"{nq} {ids.size}"
    assert open(tmp_path / "x2.fbin", "rb").read() == open(tmp_path / "x.fbin", "rb").read()
    assert open(tmp_path / "gt2.bin", "rb").read() == open(tmp_path / "gt.bin", "rb").read()
    # macro: q0 1/2, q1 skipped, q2 2/4, q3 0/1, q4 2/2 -> (0.5 + 0.5 + 0 + 1) / 4
    assert float(lines["macro"]) == pytest.approx(0.5, abs=1e-9)
    # micro: (1 + 2 + 0 + 2) / (2 + 4 + 1 + 2)
    assert float(lines["micro"]) == pytest.approx(5 / 9, abs=1e-9)
    # knn@3: q0 {3,4,5}: 2, q1 {0,1,2}: 1, q2 {7,8,9}: 2, q3: 0, q4 {5,6,7}: 2 -> 7 / 15
    assert lines["knn"].split()[:2] == [str(nq), "4"]
    assert float(lines["knn"].split()[2]) == pytest.approx(7 / 15, abs=1e-9)
