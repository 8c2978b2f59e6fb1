"""Incremental insertion on the GPU (hrnn_dev_insert_affect,
hrnn_dev_insert_dynamic, HRNN::insertPointsDynamicDevice) against the
oracle's restatement of insertPointDynamic (include/hnswlib/hrnn.h:2360-2534),
which tests/test_oracle_insert.py pins against the compiled reference.

* The affected-area step (phases 2-5) is exact: on the same graph the device
  leaves every list -- ids, distances, counts -- as the oracle does, one point
  at a time (states produced by the reference's own insertions) and in batches;
  8-bit data (integer distances, many ties) and float data (FP32 in the
  reference's order), bidirectional on float data, a sparse graph (the
  bitmap cleared by walking the lists) with a tiny offer buffer (every batch
  scanned again after it grows).
* The whole insertion (HNSW insertion with KNN seeding, then the step, batch
  by batch) is randomised like the reference's multi-threaded one; on the
  dynamic test's flow (half the points built, half inserted) its graph and
  RkNN recall at every m are within half a point of the reference's own
  insertions, or better."""
import numpy as np
import pytest

import hrnn_amd
import oracle
from hrnn_amd.index import Index

pytestmark = pytest.mark.gpu


def _same(a, b, what=""):
    ids_a, d_a, c_a = a
    ids_b, d_b, c_b = b
    assert np.array_equal(c_a, c_b), what
    K = ids_a.shape[1]
    live = np.arange(K)[None, :] < c_a[:, None]
    assert np.array_equal(np.where(live, ids_a, 0), np.where(live, ids_b, 0)), what
    assert np.array_equal(np.where(live, d_a, 0), np.where(live, d_b, 0)), what


def _pad(state, n, K):
    ids, d, c = state
    m = ids.shape[0]
    out = (np.zeros((n, K), np.uint32), np.zeros((n, K), np.float32), np.zeros(n, np.uint16))
    out[0][:m], out[1][:m], out[2][:m] = ids, d, c
    return out


def _gpu_step(dev, x, state, K, b0, b1, m_update=0, bidir=False):
    dev.set_base(x)
    dev.set_knng(*state)
    st = dev.insert_affect(b0, b1, m_update, bidir)
    return dev.get_knng(), st


@pytest.mark.parametrize("m_update", [0, 3])
def test_insert_steps_equal_oracle(gpu_dev, m_update):
    if not oracle.ref_available():
        pytest.skip("compiled reference (oracle/_ref) not available")
    n_old, nx, K = 1200, 12, 40
    x = hrnn_amd.synth_mixture(21, n_old + nx, 128, 30, 20.0, 0.0, 255.0, True, 0.0, 1021)
    r = oracle.RefIndex(n_old + nx, 128, M=16, efc=64, K=K)
    r.add_points(x[:n_old], nthreads=1)
    r.refine(2, 12, nthreads=1)
    r.build_rknng_mutable()
    prev = r.get_knng()
    upd = 0
    for t in range(nx):
        xn = n_old + t
        r.insert_dynamic(x[xn:xn + 1], np.array([xn], np.uint64), m_update, False)
        after = r.get_knng()
        s = _pad(prev, xn + 1, K)
        s[0][xn], s[1][xn], s[2][xn] = after[0][xn], after[1][xn], after[2][xn]
        want = oracle.insert_affect(x[:xn + 1], *s, K, xn, xn + 1, m_update)
        got, st = _gpu_step(gpu_dev, x[:xn + 1], s, K, xn, xn + 1, m_update)
        _same(got, want[:3], f"step {t}")
        _same(got, after, f"step {t} vs reference")
        assert st["offers"] == int(want[3][2]) and st["candidates"] == int(want[3][1])
        upd += st["updates"]
        prev = after
    assert upd > 0


def _batch_state(x, n_old, nb, K, seed_len):
    """A graph after the reference's insertion of nb points, the new rows cut
    back to their first seed_len entries (seed lists)."""
    r = oracle.RefIndex(n_old + nb, x.shape[1], M=16, efc=64, K=K)
    r.add_points(x[:n_old], nthreads=1)
    r.refine(2, 12, nthreads=1)
    r.build_rknng_mutable()
    r.insert_dynamic(x[n_old:], np.arange(n_old, n_old + nb, dtype=np.uint64), 0, False, nthreads=1)
    ids, d, c = r.get_knng()
    c = c.copy()
    c[n_old:] = np.minimum(c[n_old:], seed_len)
    return ids, d, c


@pytest.mark.parametrize("kind,bidir,m_update,lds_bm", [("u8", False, 0, 1), ("u8", False, 5, 1), ("f32", False, 0, 1),
                                                        ("f32", True, 0, 1), ("f32", True, 4, 1), ("u8", False, 0, 0),
                                                        ("f32", True, 0, 0)])
def test_insert_batch_equals_oracle(gpu_dev, kind, bidir, m_update, lds_bm):
    if not oracle.ref_available():
        pytest.skip("compiled reference (oracle/_ref) not available")
    n_old, nb, K = 1500, 60, 40
    if kind == "u8":
        x = hrnn_amd.synth_mixture(23, n_old + nb, 128, 30, 20.0, 0.0, 255.0, True, 0.0, 1023)
    else:
        x = np.random.default_rng(9).standard_normal((n_old + nb, 96)).astype(np.float32)
    s = _batch_state(x, n_old, nb, K, 24)
    want = oracle.insert_affect(x, *s, K, n_old, n_old + nb, m_update, bidir)
    gpu_dev.set_tuning("insert_lds_bitmap", lds_bm)   # visited bitmap in LDS or in HBM
    try:
        got, st = _gpu_step(gpu_dev, x, s, K, n_old, n_old + nb, m_update, bidir)
    finally:
        gpu_dev.set_tuning("insert_lds_bitmap", 1)
    _same(got, want[:3])
    assert st["offers"] == int(want[3][2]) > 0
    assert st["scanned"] == int(want[3][0]) and st["candidates"] == int(want[3][1])


@pytest.mark.parametrize("lds_bm", [0, 1])
def test_insert_sparse_graph_walk_clear_and_offer_retry(gpu_dev, lds_bm):
    """n = 100k rows with 8-entry lists: with the bitmap in HBM a point scans
    far fewer entries than its bitmap has words, so the kernel clears by
    walking the lists; an offer buffer of 64 entries makes every batch grow it
    and scan again."""
    rng = np.random.default_rng(3)
    n, d, K, nb = 100_000, 32, 8, 3000
    x = rng.standard_normal((n, d)).astype(np.float32)
    # K distinct other rows per row (offsets 1..64 times a stride, 64 * stride < n),
    # sorted by their distance (the lists' distances are data here)
    off = np.argsort(rng.random((n, 64)), axis=1)[:, :K] + 1
    ids = ((np.arange(n)[:, None] + off * 1543) % n).astype(np.uint32)
    dd = ((x[ids] - x[:, None, :]) ** 2).sum(-1).astype(np.float32)
    order = np.lexsort((ids, dd), axis=1)
    ids = np.ascontiguousarray(np.take_along_axis(ids, order, 1))
    dd = np.ascontiguousarray(np.take_along_axis(dd, order, 1))
    cnt = np.full(n, K, np.uint16)
    b0 = n - nb
    want = oracle.insert_affect(x, ids, dd, cnt, K, b0, n, 0)
    gpu_dev.set_tuning("insert_offer_cap", 64)
    gpu_dev.set_tuning("insert_lds_bitmap", lds_bm)
    try:
        got, st = _gpu_step(gpu_dev, x, (ids, dd, cnt), K, b0, n)
    finally:
        gpu_dev.set_tuning("insert_offer_cap", 0)
        gpu_dev.set_tuning("insert_lds_bitmap", 1)
    _same(got, want[:3])
    assert st["retries"] >= 1 and st["offers"] == int(want[3][2]) > 64


def _c4_like(n, nq, seed):
    x = hrnn_amd.synth_mixture(seed, n, 128, 100, 25.0, 0.0, 255.0, True, 0.0, 1000 + seed)
    q = hrnn_amd.synth_mixture(seed + 1, nq, 128, 100, 25.0, 0.0, 255.0, True, 0.0, 1000 + seed)
    return x, q


def _recalls(ix, q, gt_off, gt_ids, n, ms, Kp):
    out = {}
    for m in ms:
        off, lab = ix.search(q, m, 1.0, 10, Kp, 200)
        out[m] = oracle.rknn_recall(gt_off, gt_ids, off, lab.astype(np.uint32), n)[0]
    return out


def _knn_recall_labels(ids, cnt, labels, exact, K, kk):
    lab = np.zeros_like(ids)
    lab[labels] = labels[np.where(np.arange(K) < cnt[:, None], ids, 0)]
    c = np.zeros_like(cnt)
    c[labels] = cnt
    return oracle.knng_recall(lab, c, K, exact, kk)


@pytest.mark.parametrize("gpu_build", [True, False])
def test_dynamic_flow_recall_vs_reference(gpu_build):
    """hrnn_dynamic_test's flow (src/hrnn_dynamic_test.cpp:197-331): split 0.5
    with mt19937(42)'s shuffle order replaced by numpy's; M=16, ef_c=200,
    K=100, 10 x S=50; insert the rest (m_update 0, bidirectional off); flat
    CSR; search k=10, K'=100, m in {1, 5, 10}.  The device insertion against
    the reference's own insertions on 8 threads from its own build."""
    if not oracle.ref_available():
        pytest.skip("compiled reference (oracle/_ref) not available")
    n, nq, K, efc = 16_000, 200, 100, 200
    x, q = _c4_like(n, nq, 51)
    perm = np.random.default_rng(42).permutation(n)
    half = n // 2
    init, rest = perm[:half].astype(np.uint64), perm[half:].astype(np.uint64)
    ix = Index(n, 128, 16, efc, 100, K)
    if gpu_build:
        ix.add_points_gpu(x[init], labels=init)
    else:
        ix.add_points(x[init], labels=init, nthreads=8)
    ix.refine(10, 50)
    st = ix.insert_dynamic_gpu(x[rest], rest)
    assert ix.count == n and st["batches"] >= 1 and st["updates"] > 0
    print("insert stats:", {k: v for k, v in st.items() if k != "hnsw"}, flush=True)
    ix.build_rknng()
    r = oracle.RefIndex(n, 128, M=16, efc=efc, K=K)
    r.add_points(x[init], nthreads=8, labels=init)
    r.refine(10, 50, nthreads=8)
    r.build_rknng_mutable()
    r.insert_dynamic(x[rest], rest, 0, False, nthreads=8)
    r.build_rknng()
    kth = oracle.kth(x, 10)
    gt_off, gt_ids, _ = oracle.rknn_gt(x, q, kth)
    exact, _ = oracle.knn(x, x, 11)
    exact = exact[:, 1:].copy()
    g_ids, _, g_cnt = ix.get_knng()
    r_ids, _, r_cnt = r.get_knng()
    kg = _knn_recall_labels(g_ids, g_cnt, ix.labels(), exact, K, 10)
    kr = _knn_recall_labels(r_ids, r_cnt, r.labels(), exact, K, 10)
    print(f"KNN recall@10 after insertion: GPU {kg:.5f} reference {kr:.5f}", flush=True)
    assert kg >= kr - 0.005, (kg, kr)
    ms = (1, 5, 10)
    rg = _recalls(ix, q, gt_off, gt_ids, n, ms, 100)
    rr = {}
    for m in ms:
        off, lab = r.search(q, m, 1.0, 10, 100, 200)
        rr[m] = oracle.rknn_recall(gt_off, gt_ids, off, lab.astype(np.uint32), n)[0]
    for m in ms:
        print(f"RkNN recall m={m}: GPU {rg[m]:.5f} reference {rr[m]:.5f}", flush=True)
        assert rg[m] >= rr[m] - 0.005, (m, rg[m], rr[m])
