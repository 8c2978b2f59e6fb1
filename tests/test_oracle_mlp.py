"""Pin the FP64 MLP oracle by finite differences (the gradient must be the
derivative of the loss), before it judges the HIP training kernels."""
import numpy as np

from oracle import mlp_oracle as mo


def _data(n=40, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, 32))
    y = rng.integers(0, 16, n)
    return X, y


def test_gradient_matches_finite_differences():
    p = mo.init_params(3).astype(np.float64)
    X, y = _data()
    g = mo.grad(p, X, y)
    rng = np.random.default_rng(1)
    idx = np.concatenate([rng.choice(mo.N_PARAMS, 60, replace=False),
                          [0, 2048, 2048 + 64, 7311]])
    eps = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (mo.loss(pp, X, y) - mo.loss(pm, X, y)) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_a_dropped_relu_mask_is_caught():
    """A plausible slip (no ReLU derivative in layer 2) disagrees with finite differences."""
    p = mo.init_params(4).astype(np.float64)
    X, y = _data(seed=2)
    P = mo.unpack(p)
    prob, (Xd, z1, h1, z2, h2) = mo.forward(p, X)
    d = prob.copy()
    d[np.arange(len(y)), y] -= 1
    d /= len(y)
    dz2_wrong = d @ P["W3"]                     # missing * (z2 > 0)
    gW2_wrong = dz2_wrong.T @ h1
    g = mo.grad(p, X, y)
    o = 2048 + 64
    assert not np.allclose(gW2_wrong.ravel(), g[o:o + 4096])


def test_softmax_rows_sum_to_one():
    p = mo.init_params(5)
    X, _ = _data(seed=3)
    prob, _ = mo.forward(p, X)
    assert np.allclose(prob.sum(1), 1.0)
