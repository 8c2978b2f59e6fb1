"""CPU checks of the product libraries: load, exports, host-only helpers."""
import ctypes
import hashlib
import os
import re

import numpy as np

import hrnn_amd
from hrnn_amd import dev

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(hrnn_[a-z0-9_]+)\s*\(", txt)))


def test_device_lib_exports_every_declared_symbol():
    L = ctypes.CDLL(dev.LIB_PATH)
    names = _declared("hrnn_dev.h")
    assert len(names) >= 25
    missing = [s for s in names if not hasattr(L, s)]
    assert not missing, missing
    assert sorted(dev.EXPORTS) == names


def test_host_lib_exports_every_declared_symbol():
    path = os.path.join(ROOT, "hrnn_amd", "lib", "libhrnn_host.so")
    hdr = os.path.join(ROOT, "include", "hrnn_host.h")
    if not os.path.exists(hdr):
        return
    L = ctypes.CDLL(path)
    names = _declared("hrnn_host.h")
    missing = [s for s in names if not hasattr(L, s)]
    assert not missing, missing


def test_open_without_gpu_fails_cleanly():
    if hrnn_amd.device_count() > 0:
        return
    try:
        hrnn_amd.Device(0)
    except hrnn_amd.HrnnError as e:
        assert e.code in (-1, -2)
    else:
        raise AssertionError("Device() must fail without a GPU")


def test_synth_deterministic_and_pinned():
    """The C1 generator (50 centres, sigma 20, 0..255 integers) is pinned by SHA-256."""
    a = hrnn_amd.synth_mixture(1, 10_000, 128, 50, 20.0, 0.0, 255.0, True, 0.0, 1001)
    b = hrnn_amd.synth_mixture(1, 10_000, 128, 50, 20.0, 0.0, 255.0, True, 0.0, 1001)
    assert np.array_equal(a, b)
    assert a.min() >= 0 and a.max() <= 255 and np.array_equal(a, np.round(a))
    h = hashlib.sha256(a.tobytes()).hexdigest()
    assert h == C1_BASE_SHA256, h


def test_synth_row_independent_of_n():
    """Counter-based: row i does not depend on how many rows are generated."""
    a = hrnn_amd.synth_mixture(3, 100, 64, 10, 25.0, 0.0, 255.0, True, 0.0, 1003)
    b = hrnn_amd.synth_mixture(3, 1000, 64, 10, 25.0, 0.0, 255.0, True, 0.0, 1003)
    assert np.array_equal(a, b[:100])


def test_synth_zipf_skews_centres():
    _, c = hrnn_amd.synth_mixture(7, 20000, 8, 1000, 25.0, 0.0, 255.0, True, 1.1, 1003,
                                  return_centres=True)
    counts = np.bincount(c, minlength=1000)
    assert counts[0] > 50 * max(1, counts[500])


def test_synth_other_generators():
    x = hrnn_amd.synth_halfnormal_scaled(5, 55, 100, 960)
    assert (x >= 0).all() and x.shape == (100, 960)
    s = hrnn_amd.synth_sphere_mixture(8, 88, 200, 96, 50, 50.0)
    assert np.allclose(np.linalg.norm(s, axis=1), 1.0, atol=1e-5)


C1_BASE_SHA256 = "2ab33ce066ec9675392e6321ed2bca0dd1f4b2224efd01386854acfb38700daa"
