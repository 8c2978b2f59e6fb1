"""Learned early stop on the GPU: the stage-1 features fused into the beam
kernel against the host's HRNN::queryFeatures, the two-stage query against
per-class runs of the plain path, and what the trained network saves over the
fixed ef = 200 search at equal recall."""
import numpy as np
import pytest

import hrnn_amd
from hrnn_amd.index import Index
from hrnn_amd.learned import CLASSES, EF1, LearnedEarlyStop, label_classes, per_query_recall

pytestmark = pytest.mark.gpu

N, D, K = 10_000, 128, 100


@pytest.fixture(scope="module")
def c1_index():
    """10k x 128 integer mixture (50 centres, sigma 20), M 16, ef_c 200, K 100,
    NNDescent 10 x 50 on the GPU, reverse CSR."""
    x = hrnn_amd.synth_mixture(1, N, D, 50, 20.0, 0.0, 255.0, True, 0.0, 1001)
    ix = Index(N, D, 16, 200, 100, K)
    ix.add_points(x, labels=np.arange(N, dtype=np.uint64), nthreads=8)
    ix.refine(10, 50)
    ix.build_rknng()
    dev = hrnn_amd.Device.borrow(ix.device_handle(), N, D, K)
    return ix, dev


def _queries(seed, nq):
    return hrnn_amd.synth_mixture(seed, nq, D, 50, 20.0, 0.0, 255.0, True, 0.0, 1001)


def test_stage1_features_match_host(c1_index):
    """Same beam (ef 16) on both sides: where the 16 proxies are equal (the
    integer data has many equal distances, which the host's distance-only
    heaps and the GPU's (dist, id) order may keep differently at the list
    boundary), the fused device features equal the host's
    HRNN::queryFeatures within FP32 rounding of log1p."""
    ix, dev = c1_index
    q = _queries(31, 400)
    f_dev, p_dev, ex = dev.stage1_features(q, EF1, 16, 10)
    f_host, p_host = ix.query_features(q, 10, EF1, 16)
    same = np.all(p_dev == p_host, axis=1)
    assert same.mean() >= 0.95, same.mean()
    assert np.isfinite(f_dev).all()
    a, b = f_dev[same], f_host[same]
    assert np.allclose(a, b, rtol=1e-5, atol=1e-5), float(np.abs(a - b).max())
    assert np.all(ex[:, 0] >= 1) and np.all(ex[:, 1] >= ex[:, 0])
    # feature 31 is log1p(expansions) of the beam
    assert np.allclose(f_dev[:, 31], np.log1p(ex[:, 0].astype(np.float32)), rtol=1e-6)


def test_adaptive_forced_classes_equal_per_class_runs(c1_index):
    """Classes forced per query: stage 1 at ef 16, stage 2 at the class's ef
    for the others, masked to the class's m -- result by result equal to the
    plain phase A at that ef and phases B+C with that m."""
    _, dev = c1_index
    q = _queries(32, 300)
    rng = np.random.default_rng(5)
    cls = rng.integers(0, 16, q.shape[0]).astype(np.uint32)
    off, ids, cls_out, prox = dev.query_adaptive(q, CLASSES, EF1, 16, 10, 100, cls=cls)
    assert np.array_equal(cls_out, cls)
    # phase A alone (no result buffers) leaves the same classes and proxies
    _, _, cls_a, prox_a = dev.query_adaptive(q, CLASSES, EF1, 16, 10, 100, cls=cls, phase_a_only=True)
    assert np.array_equal(cls_a, cls) and np.array_equal(prox_a, prox)
    # queries in page-locked host memory (stage 1 reads them over the link and
    # leaves the device copy the later stages read): the same answers
    off_p, ids_p, cls_p, prox_p = dev.query_adaptive(q, CLASSES, EF1, 16, 10, 100, cls=cls, pinned=True)
    assert np.array_equal(cls_p, cls) and np.array_equal(prox_p, prox)
    assert np.array_equal(off_p, off) and np.array_equal(ids_p, ids)
    for c in range(16):
        sel = np.nonzero(cls == c)[0]
        if sel.size == 0:
            continue
        ef, m = (int(v) for v in CLASSES[c])
        p = dev.proxies(q[sel], 16, ef)
        p[:, m:] = 0xFFFFFFFF
        assert np.array_equal(prox[sel], p), c
        eo, ei = dev.rknn_search(q[sel], p[:, :max(m, 1)], 10, 100)
        for j, qi in enumerate(sel):
            got = ids[off[qi]:off[qi + 1]]
            exp = ei[eo[j]:eo[j + 1]]
            assert np.array_equal(got, exp), (c, qi)


def test_adaptive_learned_saves_phase_a_work(c1_index):
    """Train on 2000 queries (labels from GPU runs of every class against the
    GPU brute-force groundtruth), choose tau on 500 of them, evaluate on 300
    held-out queries: the two-stage path keeps the recall of the fixed
    ef = 200, m = 6 search within two points with well under half its phase-A
    expansions."""
    ix, dev = c1_index
    k, Kp = 10, 100
    kth = dev.bf_kth(k, "bf16")
    qt, qe = _queries(41, 2000), _queries(42, 300)
    ft, _, _ = dev.stage1_features(qt, EF1, 16, k)
    gt_off, gt_ids, _ = dev.bf_rknn(qt, kth, "bf16")
    labels, rec_t, _ = label_classes(dev, qt, gt_off, gt_ids, k, Kp)
    tr, va = slice(0, 1500), slice(1500, 2000)
    model, losses = LearnedEarlyStop.train(dev, ft[tr], labels[tr], epochs=60, batch=128, lr=2e-3)
    assert losses[-1] < losses[0]
    fixed_c = 12                                   # (200, 6)
    tau = model.pick_tau(dev, ft[va], rec_t[:, va], np.nanmean(rec_t[fixed_c, va]) - 0.002)
    ge_off, ge_ids, _ = dev.bf_rknn(qe, kth, "bf16")
    _, rec_e, exp_e = label_classes(dev, qe, ge_off, ge_ids, k, Kp)
    off, ids, cls, _ = model.query(dev, qe, k, Kp, tau)
    r_q = per_query_recall(ge_off, ge_ids, off, ids)
    rec_l = np.nanmean(r_q)
    # the device's choice is the one the host rule makes from the same
    # probabilities, and each query's answer is its class's plain-path answer
    fe, _, _ = dev.stage1_features(qe, EF1, 16, k)
    assert np.array_equal(cls, model.choose(model.probs(dev, fe), tau))
    cols = np.arange(qe.shape[0])
    assert np.array_equal(np.isnan(r_q), np.isnan(rec_e[cls, cols]))
    assert np.allclose(r_q, rec_e[cls, cols], rtol=0, atol=1e-12, equal_nan=True)
    cost_l = exp_e[0] + np.where(CLASSES[cls, 0] > EF1, exp_e[cls, cols], 0)
    cost_f = exp_e[fixed_c]
    rec_f = np.nanmean(rec_e[fixed_c])
    print(f"learned (tau {tau}): recall {rec_l:.4f}, phase-A expansions {cost_l.mean():.1f}; "
          f"fixed (200, 6): recall {rec_f:.4f}, expansions {cost_f.mean():.1f}; "
          f"classes {np.bincount(cls, minlength=16).tolist()}")
    # 300 held-out queries: one query's recall moves the mean by up to 0.0033
    assert rec_l >= rec_f - 0.02
    assert cost_l.mean() <= 0.5 * cost_f.mean()
    # Stage-1 cap ("stage1_cap"): a query whose ef-16 beam needs more
    # expansions than the cap leaves stage 1 early and escalates (class with
    # ef > 16), its answer again its class's plain-path answer; a query that
    # converges below the cap is untouched.
    e1 = exp_e[0]                                  # uncapped stage-1 expansions per query
    cap1 = int(np.percentile(e1, 60))              # an ef-16 beam expands at least ~16 entries
    dev.set_tuning("stage1_cap", cap1)
    try:
        off_c, ids_c, cls_c, _ = model.query(dev, qe, k, Kp, tau)
    finally:
        dev.set_tuning("stage1_cap", 0)
    under, over = e1 < cap1, e1 > cap1
    assert under.any() and over.any()
    assert np.array_equal(cls_c[under], cls[under])
    assert np.all(CLASSES[cls_c[over], 0] > EF1)
    r_c = per_query_recall(ge_off, ge_ids, off_c, ids_c)
    assert np.allclose(r_c, rec_e[cls_c, cols], rtol=0, atol=1e-12, equal_nan=True)
    for qi in np.nonzero(under)[0]:
        assert np.array_equal(ids_c[off_c[qi]:off_c[qi + 1]], ids[off[qi]:off[qi + 1]])
