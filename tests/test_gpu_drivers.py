"""The CLI drivers end to end on the GPU at the reference's small
configuration (10k x 128 mixture, 100 queries; M=16, ef_c=200, K=100,
10 x S=50; k=10, K'=100, m in {1,5,10}, ef_search=200):

  hrnn_gt     (replaces scripts/gen_rknn_gt.py; output format :154-168)
  hrnn_build  (flags and stages of src/hrnn_build.cpp)
  hrnn_search (flags, report lines and recall of src/hrnn_search.cpp:160-228)

The groundtruth file must equal the oracle's exact RkNN (offsets, ids and
distances), it must load through load_rknn_groundtruth (hrnn_search reads it
so), and the recall hrnn_search prints must equal the recall of the same
searches made through the library and scored by the oracle."""
import os
import re
import subprocess

import numpy as np
import pytest

import hrnn_amd
import oracle
from hrnn_amd.index import Index

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "hrnn_amd", "bin")


def _fbin(path, x):
    with open(path, "wb") as f:
        f.write(np.array(x.shape, np.uint32).tobytes() + np.ascontiguousarray(x, np.float32).tobytes())


def _run(args, timeout=300):
    print("run:", " ".join(args), flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-2000:])
    return r.stdout


def test_gt_build_search_drivers_c1(tmp_path):
    x = hrnn_amd.synth_mixture(1, 10_000, 128, 50, 20.0, 0.0, 255.0, True, 0.0, 1001)
    q = hrnn_amd.synth_mixture(2, 100, 128, 50, 20.0, 0.0, 255.0, True, 0.0, 1001)
    xb, qb = str(tmp_path / "base.fbin"), str(tmp_path / "query.fbin")
    _fbin(xb, x)
    _fbin(qb, q)
    gt_path, idx = str(tmp_path / "gt.bin"), str(tmp_path / "c1.idx")

    _run([os.path.join(BIN, "hrnn_gt"), "--base", xb, "--query", qb, "--k", "10", "--output", gt_path])
    raw = open(gt_path, "rb").read()
    nq = int(np.frombuffer(raw[:4], np.uint32)[0])
    off = np.frombuffer(raw[4:4 + 4 * (nq + 1)], np.uint32)
    tot = int(off[-1])
    ids = np.frombuffer(raw[4 + 4 * (nq + 1):4 + 4 * (nq + 1) + 4 * tot], np.uint32)
    d2 = np.frombuffer(raw[4 + 4 * (nq + 1) + 4 * tot:], np.float32)
    assert nq == 100 and d2.size == tot
    kth = oracle.kth(x, 10)
    e_off, e_ids, e_d2 = oracle.rknn_gt(x, q, kth)
    assert np.array_equal(off.astype(np.int64), e_off.astype(np.int64))
    for i in range(nq):
        a = sorted(zip(d2[off[i]:off[i + 1]].tolist(), ids[off[i]:off[i + 1]].tolist()))
        b = sorted(zip(e_d2[e_off[i]:e_off[i + 1]].tolist(), e_ids[e_off[i]:e_off[i + 1]].tolist()))
        assert a == b, i
        assert np.all(np.diff(d2[off[i]:off[i + 1]]) >= 0), i        # sorted by distance

    out = _run([os.path.join(BIN, "hrnn_build"), "--base", xb, "--index_path", idx, "--M", "16",
                "--ef_construction", "200", "--K_knng", "100", "--nn_iters", "10", "--nn_sample", "50",
                "--num_threads", "8"])
    assert os.path.getsize(idx) > 0 and "refineKNNG" in out

    out = _run([os.path.join(BIN, "hrnn_search"), "--index_path", idx, "--query", qb, "--gt", gt_path,
                "--k", "10", "--K_prime", "100", "--m", "1,5,10", "--ef_search", "200"])
    printed = [float(v) for v in re.findall(r"Macro Recall = ([0-9.]+)", out)]
    assert len(printed) == 3 and "nq=100  total=%d" % tot in out
    ix = Index.load(idx, 128)
    for m, rec_p in zip((1, 5, 10), printed):
        r_off, r_lab = ix.search(q, m, 1.0, 10, 100, 200, per_query=True)
        rec = oracle.rknn_recall(e_off, e_ids, r_off, r_lab.astype(np.uint32), 10_000)[0]
        assert rec_p == pytest.approx(rec, abs=2e-6), (m, rec_p, rec)
    assert printed[-1] >= 0.95


def test_dynamic_test_driver_gpu_and_host_insertion(tmp_path):
    """hrnn_dynamic_test (flags, split and report lines of
    src/hrnn_dynamic_test.cpp): the initial half built and saved as a
    snapshot, the rest inserted on the GPU in batches; then from the same
    snapshot inserted per point on the host threads; the full-build index as
    the oracle snapshot.  Both insertions reach the recall the printed
    summary reports within half a point of each other, and the printed
    recall equals the library's search of the same index scored by the
    oracle (the gap line is the difference to the full build)."""
    n = 8000
    x = hrnn_amd.synth_mixture(5, n, 128, 40, 20.0, 0.0, 255.0, True, 0.0, 1005)
    q = hrnn_amd.synth_mixture(6, 100, 128, 40, 20.0, 0.0, 255.0, True, 0.0, 1005)
    xb, qb = str(tmp_path / "base.fbin"), str(tmp_path / "query.fbin")
    _fbin(xb, x)
    _fbin(qb, q)
    gt_path, full_idx, snap = str(tmp_path / "gt.bin"), str(tmp_path / "full.idx"), str(tmp_path / "snap.idx")
    _run([os.path.join(BIN, "hrnn_gt"), "--base", xb, "--query", qb, "--k", "10", "--output", gt_path])
    common = ["--M", "16", "--ef_construction", "100", "--K_knng", "64", "--nn_iters", "4", "--nn_sample", "20",
              "--num_threads", "8"]
    _run([os.path.join(BIN, "hrnn_build"), "--base", xb, "--index_path", full_idx] + common)
    dyn = [os.path.join(BIN, "hrnn_dynamic_test"), "--base", xb, "--query", qb, "--gt", gt_path,
           "--snapshot", snap, "--oracle_snapshot", full_idx, "--k", "10", "--m", "10", "--ef_search", "100"] + common
    out_gpu = _run(dyn + ["--insert_device", "1"])
    assert os.path.exists(snap) and "GPU insertion:" in out_gpu and "Oracle macro gap" in out_gpu
    out_host = _run(dyn + ["--insert_device", "0"])   # loads the snapshot the first run saved
    assert "Loading snapshot" in out_host and "Per-insert wall" in out_host

    def summary(out):
        macro = float(re.findall(r"  Macro Recall = ([0-9.]+)", out)[-1])
        gap = float(re.findall(r"Oracle macro gap = (-?[0-9.]+)", out)[-1])
        return macro, gap

    (mg, gg), (mh, gh) = summary(out_gpu), summary(out_host)
    print(f"dynamic driver: GPU insertion recall {mg:.4f} (gap {gg:+.4f}), host {mh:.4f} (gap {gh:+.4f})")
    assert mg >= mh - 0.005, (mg, mh)
    assert mg > 0.9


def test_search_driver_serve_mode(tmp_path):
    """hrnn_search --max_k_serve (src/hrnn_search.cpp): the recall printed in
    serve mode equals the recall printed without it, for the same index,
    queries and groundtruth."""
    x = hrnn_amd.synth_mixture(7, 6000, 128, 40, 20.0, 0.0, 255.0, True, 0.0, 1007)
    q = hrnn_amd.synth_mixture(8, 100, 128, 40, 20.0, 0.0, 255.0, True, 0.0, 1007)
    xb, qb = str(tmp_path / "base.fbin"), str(tmp_path / "query.fbin")
    _fbin(xb, x)
    _fbin(qb, q)
    gt_path, idx = str(tmp_path / "gt.bin"), str(tmp_path / "s.idx")
    _run([os.path.join(BIN, "hrnn_gt"), "--base", xb, "--query", qb, "--k", "10", "--output", gt_path])
    _run([os.path.join(BIN, "hrnn_build"), "--base", xb, "--index_path", idx, "--M", "16", "--ef_construction",
          "100", "--K_knng", "64", "--nn_iters", "4", "--nn_sample", "20", "--num_threads", "8"])
    args = [os.path.join(BIN, "hrnn_search"), "--index_path", idx, "--query", qb, "--gt", gt_path, "--k", "10",
            "--K_prime", "64", "--m", "1,5", "--ef_search", "100"]
    plain = [float(v) for v in re.findall(r"Macro Recall = ([0-9.]+)", _run(args))]
    out = _run(args + ["--max_k_serve", "10"])
    served = [float(v) for v in re.findall(r"Macro Recall = ([0-9.]+)", out)]
    assert "serve mode" in out and len(plain) == 2 and served == plain, (plain, served)
