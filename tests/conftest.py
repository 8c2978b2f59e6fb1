import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import hrnn_amd  # noqa: E402
import oracle  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libhrnn_hip.so)")
    config.addinivalue_line("markers", "slow: takes more than a few seconds on the CPU")


def _have_gpu() -> bool:
    try:
        return hrnn_amd.device_count() > 0
    except Exception:
        return False


def pytest_collection_modifyitems(config, items):
    if _have_gpu():
        return
    skip = pytest.mark.skip(reason="no GPU visible")
    for it in items:
        if "gpu" in it.keywords:
            it.add_marker(skip)


# ---- shared datasets ------------------------------------------------------
# C1-like mixture: values 0..255 integers, 50 centres, sigma 20.

def mixture(n, d=128, centres=50, sigma=20.0, seed=1, centre_seed=1001):
    return hrnn_amd.synth_mixture(seed, n, d, centres, sigma, 0.0, 255.0, True, 0.0, centre_seed)


@pytest.fixture(scope="session")
def small_data():
    """n=3000 base + 60 queries, d=128, integer mixture (C1 generator, smaller)."""
    x = mixture(3000, seed=11, centre_seed=1011)
    q = mixture(60, seed=12, centre_seed=1011)
    return x, q


@pytest.fixture(scope="session")
def small_ref(small_data):
    """Reference HRNN built single-threaded on small_data (M=16, efc=100, K=48), refined."""
    x, _ = small_data
    if not oracle.ref_available():
        pytest.skip("compiled reference (oracle/_ref) not available")
    r = oracle.RefIndex(x.shape[0], x.shape[1], M=16, efc=100, K=48)
    r.add_points(x, nthreads=1)
    seed = r.get_knng()
    r.refine(3, 12, nthreads=1)
    r.build_rknng()
    return r, seed


@pytest.fixture(scope="session")
def float_data():
    rng = np.random.default_rng(5)
    x = rng.standard_normal((1500, 96)).astype(np.float32)
    q = rng.standard_normal((40, 96)).astype(np.float32)
    return x, q


@pytest.fixture(scope="session")
def gpu_dev():
    if not _have_gpu():
        pytest.skip("no GPU visible")
    return hrnn_amd.Device(0)
