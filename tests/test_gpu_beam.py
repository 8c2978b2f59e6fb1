"""GPU phase A (k_phase_a: HNSW descent + layer-0 beam, one query per wave,
LDS visited set) against the host beam, which follows the reference's
searchBaseLayerST<true> (hrnn.h:485-616).  Both use the reference's distance
order, so the proxy sets agree up to the order of equal distances at the
list boundary."""
import numpy as np
import pytest

import hrnn_amd
import oracle
from conftest import mixture
from hrnn_amd.index import Index

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_index():
    x = mixture(3000, seed=61, centre_seed=1061)
    q = mixture(80, seed=62, centre_seed=1061)
    ix = Index(3000, 128, 16, 100, 100, 48)
    ix.add_points(x, nthreads=1)
    return ix, x, q


def _agree(a, b):
    return float(np.mean([len(set(a[i].tolist()) & set(b[i].tolist())) / a.shape[1]
                          for i in range(a.shape[0])]))


@pytest.mark.parametrize("m,ef", [(1, 10), (10, 64), (20, 200), (50, 200), (64, 64)])
def test_phase_a_matches_host_beam_small(small_index, m, ef):
    ix, x, q = small_index
    hp = ix.proxies(q, m, ef, 1)
    gp = ix.proxies_gpu(q, m, ef)
    assert gp.shape == hp.shape
    assert np.all(gp < 3000)
    assert _agree(hp, gp) >= 0.995, _agree(hp, gp)
    exact = oracle.knn(x, q, m)[0]
    assert _agree(exact.astype(np.uint32), gp) > 0.9


def test_phase_a_and_query_c1_scale():
    """An index of the reference's small configuration (10k x 128, K=100):
    2000 queries with m=64,
    ef=200 on the GPU beam vs the host beam, then the whole resident query
    (GPU phase A + phases B and C in one call) vs host phase A + GPU B and C."""
    from hrnn_amd import hiprt
    x = hrnn_amd.synth_mixture(1, 10_000, 128, 50, 20.0, 0.0, 255.0, True, 0.0, 1001)
    q = hrnn_amd.synth_mixture(21, 2000, 128, 50, 20.0, 0.0, 255.0, True, 0.0, 1001)
    ix = Index(10_000, 128, 16, 200, 100, 100)
    ix.add_points(x, labels=np.arange(10_000, dtype=np.uint64), nthreads=8)
    ix.refine(10, 50)
    ix.build_rknng()
    hp = ix.proxies(q, 64, 200, 8)
    gp = ix.proxies_gpu(q, 64, 200)
    assert np.all(gp < 10_000)
    assert _agree(hp, gp) >= 0.995, _agree(hp, gp)
    # resident whole query at m = 5 against host proxies + the same B/C kernel
    d = hrnn_amd.Device.borrow(ix.device_handle(), 10_000, 128, 100)
    L = d.L
    nq, m, cap = q.shape[0], 5, 1024
    qd = hiprt.DeviceArray.from_host(q)
    pd = hiprt.DeviceArray((nq * m,), np.uint32)
    cd = hiprt.DeviceArray((nq,), np.uint32)
    idd = hiprt.DeviceArray((nq * cap,), np.uint32)
    rc = L.hrnn_dev_rknn_query_dptr(d.h, qd.ptr, nq, m, 200, None, 10, 100, 1.0, pd.ptr, cd.ptr, idd.ptr,
                                    cap, None)
    assert rc == 0, L.hrnn_dev_last_error(d.h)
    assert L.hrnn_dev_sync(d.h) == 0, L.hrnn_dev_last_error(d.h)
    cnt = cd.to_host().astype(np.int64)
    ids = idd.to_host().reshape(nq, cap)
    off_h, ids_h = d.rknn_search(q, np.ascontiguousarray(hp[:, :m]), 10, 100)
    same = np.mean([set(ids[i, :cnt[i]].tolist()) == set(ids_h[off_h[i]:off_h[i + 1]].tolist())
                    for i in range(nq)])
    assert same >= 0.99, same


def test_phase_a_u8_storage_equals_bf16(small_index):
    """The u8 mirror gives the same distances as BF16 on 8-bit data, so the
    GPU beam returns the same proxies, ties included; for shifted queries the
    reference-order path over bytes agrees as well."""
    ix, x, q = small_index
    d = hrnn_amd.Device.borrow(ix.device_handle(), 3000, 128, 48)
    for qq in (q, (q + np.float32(0.5)).astype(np.float32)):
        d.set_vector_precision("bf16")
        a = ix.proxies_gpu(qq, 20, 64)
        d.set_vector_precision("u8")
        b = ix.proxies_gpu(qq, 20, 64)
        d.set_vector_precision("bf16")
        assert np.array_equal(a, b)
