"""Learned m on the GPU: forward/backward kernels against the FP64 oracle,
training on the GPU, and what the trained network gains over fixed m."""
import numpy as np
import pytest

import hrnn_amd
import oracle
from hrnn_amd.learned import GRID, LearnedM, label_queries, mask_proxies, per_query_recall
from hrnn_amd.index import Index
from oracle import mlp_oracle as mo

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    return hrnn_amd.Device(0)


def test_forward_matches_oracle(dev):
    p = mo.init_params(7)
    X = np.random.default_rng(0).standard_normal((300, 32)).astype(np.float32)
    got = dev.mlp_forward(p, X)
    exp, _ = mo.forward(p, X)
    assert np.allclose(got, exp, rtol=1e-4, atol=1e-6)


def test_gradient_matches_oracle(dev):
    p = mo.init_params(8)
    rng = np.random.default_rng(1)
    X = rng.standard_normal((257, 32)).astype(np.float32)
    y = rng.integers(0, 16, 257).astype(np.uint32)
    g, loss = dev.mlp_grad(p, X, y)
    eg = mo.grad(p, X, y.astype(np.int64))
    assert loss == pytest.approx(mo.loss(p, X, y.astype(np.int64)), rel=1e-5)
    err = np.abs(g - eg)
    assert np.all(err <= 1e-5 + 1e-3 * np.abs(eg)), float(err.max())
    # deterministic reduction: identical on a second call
    g2, _ = dev.mlp_grad(p, X, y)
    assert np.array_equal(g, g2)


def test_training_fits_a_learnable_rule(dev):
    rng = np.random.default_rng(2)
    X = rng.standard_normal((4000, 32)).astype(np.float32)
    y = (np.clip((X[:, 0] * 3 + 8), 0, 15)).astype(np.uint32)     # monotone in one feature
    model, losses = LearnedM.train(dev, X, y, epochs=30, batch=128, lr=3e-3)
    assert losses[-1] < 0.5 * losses[0]
    pred = np.argmax(model.probs(dev, X), 1)
    assert float((np.abs(pred.astype(int) - y.astype(int)) <= 1).mean()) > 0.9


@pytest.mark.skip(reason="its first GPU run ended in a memory fault (the persistent phase-A beam "
                         "k_beam with the feature kernel k_features); both kernels and the device "
                         "feature entry points are removed, the cause was not found; the learned m "
                         "is tested with host features (test_learned_m_gain_host_features) and with "
                         "features fused into k_phase_a (tests/test_gpu_adaptive.py)")
def test_learned_m_gain_over_fixed_m50(dev):
    """10k x 128 mixture index (K=100): train on 2000 queries, evaluate on 300
    held-out ones against fixed m = 50 (the proxy count the learned form is
    judged against)."""
    x = hrnn_amd.synth_mixture(1, 10_000, 128, 50, 20.0, 0.0, 255.0, True, 0.0, 1001)
    qt = hrnn_amd.synth_mixture(21, 2000, 128, 50, 20.0, 0.0, 255.0, True, 0.0, 1001)
    qe = hrnn_amd.synth_mixture(22, 300, 128, 50, 20.0, 0.0, 255.0, True, 0.0, 1001)
    ix = Index(10_000, 128, 16, 200, 100, 100)
    ix.add_points(x, labels=np.arange(10_000, dtype=np.uint64), nthreads=1)
    ix.refine(10, 50)
    ix.build_rknng()
    h = ix.device_handle()
    d = hrnn_amd.Device.borrow(h, 10_000, 128, 100)
    kth = d.bf_kth(10, "bf16")
    k, Kp, ef = 10, 100, 200
    ft, pt = d.query_features(qt, k, ef, 64)
    gt_off, gt_ids, _ = d.bf_rknn(qt, kth, "bf16")
    labels, _ = label_queries(d, qt, pt, gt_off, gt_ids, k, Kp)
    model, losses = LearnedM.train(d, ft, labels, epochs=60, batch=128, lr=2e-3)
    assert losses[-1] < losses[0]
    fe, pe = d.query_features(qe, k, ef, 64)
    ge_off, ge_ids, _ = d.bf_rknn(qe, kth, "bf16")
    probs = model.probs(d, fe)
    m_q = model.choose_m(probs, 0.9)
    off, ids = d.rknn_search(qe, mask_proxies(pe, m_q), k, Kp)
    rec_l = np.nanmean(per_query_recall(ge_off, ge_ids, off, ids))
    off50, ids50 = d.rknn_search(qe, np.ascontiguousarray(pe[:, :50]), k, Kp)
    rec_50 = np.nanmean(per_query_recall(ge_off, ge_ids, off50, ids50))
    print(f"learned: recall {rec_l:.4f} mean m {m_q.mean():.1f}; fixed m=50: recall {rec_50:.4f}")
    assert rec_l >= rec_50 - 0.02
    assert m_q.mean() <= 0.6 * 50


def test_learned_m_gain_host_features(dev):
    """10k x 128 mixture index (K=100), features from the host phase A: train
    the MLP on the GPU on 2000 queries labelled by GPU searches against the
    GPU brute-force groundtruth; on 300 held-out queries the learned m keeps
    the recall of fixed m = 50 with far fewer proxies."""
    x = hrnn_amd.synth_mixture(1, 10_000, 128, 50, 20.0, 0.0, 255.0, True, 0.0, 1001)
    qt = hrnn_amd.synth_mixture(21, 2000, 128, 50, 20.0, 0.0, 255.0, True, 0.0, 1001)
    qe = hrnn_amd.synth_mixture(22, 300, 128, 50, 20.0, 0.0, 255.0, True, 0.0, 1001)
    ix = Index(10_000, 128, 16, 200, 100, 100)
    ix.add_points(x, labels=np.arange(10_000, dtype=np.uint64), nthreads=8)
    ix.refine(10, 50)
    ix.build_rknng()
    d = hrnn_amd.Device.borrow(ix.device_handle(), 10_000, 128, 100)
    kth = d.bf_kth(10, "bf16")               # internal-id space throughout
    k, Kp, ef = 10, 100, 200
    ft, pt = ix.query_features(qt, k, ef, 64)
    assert np.isfinite(ft).all()
    gt_off, gt_ids, _ = d.bf_rknn(qt, kth, "bf16")
    labels, rec_grid = label_queries(d, qt, pt, gt_off, gt_ids, k, Kp)
    tr, va = slice(0, 1500), slice(1500, 2000)      # fit / choose tau on separate queries
    model, losses = LearnedM.train(d, ft[tr], labels[tr], epochs=60, batch=128, lr=2e-3)
    assert losses[-1] < losses[0]
    # tau: the smallest one that keeps fixed m = 50's recall on the validation queries
    pv = model.probs(d, ft[va])
    rv = rec_grid[:, va]
    rec50_v = rv[list(GRID).index(50)].mean()
    tau_star = 0.995
    for tau in (0.5, 0.7, 0.8, 0.9, 0.95, 0.98, 0.995):
        gi = np.searchsorted(GRID, model.choose_m(pv, tau))
        if rv[gi, np.arange(len(gi))].mean() >= rec50_v - 0.002:
            tau_star = tau
            break
    fe, pe = ix.query_features(qe, k, ef, 64)
    ge_off, ge_ids, _ = d.bf_rknn(qe, kth, "bf16")
    probs = model.probs(d, fe)
    m_q = model.choose_m(probs, tau_star)
    off, ids = d.rknn_search(qe, mask_proxies(pe, m_q), k, Kp)
    rec_l = np.nanmean(per_query_recall(ge_off, ge_ids, off, ids))
    off50, ids50 = d.rknn_search(qe, np.ascontiguousarray(pe[:, :50]), k, Kp)
    rec_50 = np.nanmean(per_query_recall(ge_off, ge_ids, off50, ids50))
    print(f"learned (tau {tau_star}): recall {rec_l:.4f} mean m {m_q.mean():.1f}; "
          f"fixed m=50: recall {rec_50:.4f}")
    assert rec_l >= rec_50 - 0.02
    assert m_q.mean() <= 0.6 * 50
