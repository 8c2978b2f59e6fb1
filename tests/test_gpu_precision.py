"""The precision form on the GPU against the oracle (the generator of the
1M x 960 configuration at a size the oracle finishes in seconds):
|N(0,1)| scaled per dimension by U(0.05, 1), d = 960, k = 10.

1. Certified brute force (hrnn_dev_bf_kth_ex): for the FP32, BF16 and FP16
   candidate passes, r_k, the (k+1) ids and the FP64 r_k equal the oracle's
   exhaustive FP32 ranking (kth, knn: ids bit-exact, ties to the smaller id)
   and FP64 judge (kth_f64) bit for bit.  Two planted clusters of near-copies
   (48 and 70 rows, sigma 1e-3) defeat the first pass's certificate: the
   48-row cluster must be resolved by the 64-wide rung, the 70-row one by the
   FP32 rung, and no row may stay uncertified.
2. FP64 re-check of RkNN (hrnn_dev_bf_rknn_f64): queries placed within FP32
   rounding of a row's radius, where the FP32 and FP64 decisions differ
   (tests/boundary_cases.py), plus ordinary queries: the result sets equal
   the oracle's FP64 groundtruth (rknn_gt_f64) and the planted pairs are
   counted as flipped; the FP32 decision (hrnn_dev_bf_rknn) equals the
   oracle's FP32 groundtruth on the same queries."""
import numpy as np
import pytest

import hrnn_amd
import oracle
from boundary_cases import boundary_queries

pytestmark = pytest.mark.gpu

K = 10


def _data():
    x = hrnn_amd.synth_halfnormal_scaled(5, 55, 4000, 960)
    rng = np.random.default_rng(9)
    a = x[0] + rng.normal(0, 1e-3, (48, 960)).astype(np.float32)
    b = x[1] + rng.normal(0, 1e-3, (70, 960)).astype(np.float32)
    return np.ascontiguousarray(np.concatenate([x, a, b]).astype(np.float32))


@pytest.fixture(scope="module")
def prec_case():
    x = _data()
    e_ids, e_d = oracle.knn(x, x, K + 1)
    k64 = oracle.kth_f64(x, K)
    print(f"precision case: n={x.shape[0]} d=960, oracle ready", flush=True)
    return x, e_ids, e_d, k64


@pytest.mark.parametrize("prec", ["f32", "bf16", "f16"])
def test_certified_kth_equals_exhaustive(prec_case, prec):
    x, e_ids, e_d, k64 = prec_case
    n = x.shape[0]
    dev = hrnn_amd.Device(0)
    try:
        dev.set_base(x)
        r = dev.bf_kth_ex(K, prec, raw=True)
    finally:
        dev.close()
    st = r["stats"]
    print(prec, st, flush=True)
    assert st["rows"] == n and st["uncertified"] == 0
    assert np.array_equal(r["ids"], e_ids)
    assert np.array_equal(r["kth"], e_d[:, K])
    assert np.array_equal(r["kth64"], k64)
    clusters = np.arange(4000, n)
    if prec == "f32":
        assert st["redo_wide"] == 0
    else:
        # the cluster rows cannot be certified from 32 candidates in 16 bits
        assert st["certified"] <= n - len(clusters) and st["redo_wide"] >= len(clusters)
        assert st["redo_f32"] >= 70
    rel = np.abs(r["raw_kth"].astype(np.float64) - k64) / k64
    print(f"{prec}: first-pass r_k error vs FP64 max {rel[:4000].max():.3e} "
          f"median {np.median(rel[:4000]):.3e}", flush=True)


def test_f64_recheck_rknn_boundary_queries(prec_case):
    x, e_ids, e_d, k64 = prec_case
    k32 = e_d[:, K].copy()
    qb, pb, inb = boundary_queries(x, K, k32, k64, range(2, 4000, 61), want=8)
    assert len(qb) >= 4, len(qb)
    qr = hrnn_amd.synth_halfnormal_scaled(6, 55, 40, 960)
    q = np.ascontiguousarray(np.concatenate([qb, qr]))
    print(f"{len(qb)} boundary queries", flush=True)
    dev = hrnn_amd.Device(0)
    try:
        dev.set_base(x)
        r = dev.bf_kth_ex(K, "bf16")
        assert np.array_equal(r["kth"], k32) and np.array_equal(r["kth64"], k64)
        off, ids, d2, rechecked, flipped = dev.bf_rknn_f64(q, r["kth"], r["kth64"])
        f_off, f_ids, _ = dev.bf_rknn(q, r["kth"], "f32")
    finally:
        dev.close()
    print(f"rechecked {rechecked} flipped {flipped}", flush=True)
    g_off, g_ids, _ = oracle.rknn_gt_f64(x, q, k64)
    o_off, o_ids, _ = oracle.rknn_gt(x, q, k32)
    assert np.array_equal(off, g_off)
    for i in range(q.shape[0]):
        assert np.array_equal(np.sort(ids[off[i]:off[i + 1]]), np.sort(g_ids[g_off[i]:g_off[i + 1]])), i
    assert np.array_equal(f_off, o_off)
    for i in range(q.shape[0]):
        assert np.array_equal(np.sort(f_ids[f_off[i]:f_off[i + 1]]), np.sort(o_ids[o_off[i]:o_off[i + 1]])), i
    for i in range(len(qb)):
        got = set(ids[off[i]:off[i + 1]].tolist())
        f32 = set(f_ids[f_off[i]:f_off[i + 1]].tolist())
        assert (int(pb[i]) in got) == bool(inb[i]) and (int(pb[i]) in f32) != bool(inb[i]), i
    assert flipped >= len(qb) and rechecked >= flipped


def test_nndescent_d960_same_seed_graph_vs_reference():
    """NNDescent on the precision form's data (d = 960, float, the chunked
    BF16-staged join with FP32 distances) against the compiled reference's
    own refineKNNG (hrnn.h:1725-1862) from the same seed graph: 20,000
    points, M=16, ef_c=200, K=200, 10 x S=50.  KNN recall at 10 and at K
    (500-node sample, exact FP32 neighbours) within 0.5 points."""
    from hrnn_amd.index import Index
    import os
    import tempfile
    if not oracle.ref_available():
        pytest.skip("oracle/_ref not built")
    n, K = 20_000, 200
    oracle.set_threads(16)
    x = hrnn_amd.synth_halfnormal_scaled(11, 55, n, 960)
    ix = Index(n, 960, 16, 200, 100, K)
    ix.add_points(x, labels=np.arange(n, dtype=np.uint64), nthreads=16)
    with tempfile.TemporaryDirectory() as tmp:
        p = os.path.join(tmp, "seed.idx")
        ix.save(p)
        ref = oracle.RefIndex.load(p, 960)
    ref.K = K
    lab = ix.labels()
    st = ix.refine(10, 50)
    assert sum(s["dropped"] for s in st) == 0
    ref.refine(10, 50, nthreads=16)
    samp = np.random.default_rng(1).choice(n, 500, replace=False)
    ex, _ = oracle.knn(x, x[samp], K + 1)
    ex = ex[:, 1:].copy()

    def rec(ids, cnt, kk):
        g = np.zeros_like(ids)
        live = np.where(np.arange(ids.shape[1]) < cnt[:, None], ids, 0)
        g[lab] = lab[live]
        c = np.zeros_like(cnt)
        c[lab] = cnt
        return oracle.knng_recall(g[samp], c[samp], K, ex, kk)

    gid, _, gcnt = ix.get_knng()
    rid, _, rcnt = ref.get_knng()
    ref.close()
    for kk in (10, K):
        ours, theirs = rec(gid, gcnt, kk), rec(rid, rcnt, kk)
        print(f"d=960 KNN recall@{kk}: GPU {ours:.5f} reference {theirs:.5f}", flush=True)
        assert abs(ours - theirs) <= 0.005, (kk, ours, theirs)
