"""The search entry point that allocates nothing
(hrnn_dev_rknn_search_noalloc_dptr) against the allocating one
(hrnn_dev_rknn_search_dptr): through the C-ABI, on the same resident inputs,
identical counts, ids and work counters; two calls in a row, the second on a
smaller batch, with the LDS table at its default size and at 1 KiB (most
queries then take the fallback whose state lives in the static scratch
array), in presence mode (threshold 1) and count mode (threshold 2)."""
import ctypes as C

import numpy as np
import pytest

import hrnn_amd
from hrnn_amd import hiprt

pytestmark = pytest.mark.gpu


def _run(dev, fn, q, prox, k, Kp, thr, cap):
    nq, m = prox.shape
    qd = hiprt.DeviceArray.from_host(q)
    pd = hiprt.DeviceArray.from_host(prox)
    cd = hiprt.DeviceArray((nq,), np.uint32)
    idd = hiprt.DeviceArray((nq * cap,), np.uint32)
    wd = hiprt.DeviceArray((2 * nq,), np.uint32)
    rc = fn(dev.h, qd.ptr, nq, pd.ptr, m, None, k, Kp, C.c_float(thr), cd.ptr, idd.ptr, cap, wd.ptr)
    if rc != 0:
        raise hrnn_amd.HrnnError(rc, dev.L.hrnn_dev_last_error(dev.h).decode())
    dev.sync()
    cnt = cd.to_host()
    ids = idd.to_host().reshape(nq, cap)
    return cnt, [ids[i, :min(int(cnt[i]), cap)].copy() for i in range(nq)], wd.to_host()


@pytest.mark.parametrize("lds_kib,thr", [(16, 1.0), (1, 1.0), (1, 2.0)])
def test_noalloc_search_equals_allocating_search(small_data, small_ref, lds_kib, thr):
    x, q = small_data
    r, _ = small_ref
    ids, dists, cnt = r.get_knng()
    dev = hrnn_amd.Device(0)
    dev.set_base(x)
    dev.set_knng(ids, dists, cnt)
    dev.build_rknn_csr()
    dev.set_tuning("search_lds_kib", lds_kib)
    L = dev.L
    prox = r.proxies(q, 10, 64)
    for nq in (len(q), len(q) // 3):                       # the second call is smaller
        p = np.ascontiguousarray(prox[:nq])
        qq = np.ascontiguousarray(q[:nq])
        a = _run(dev, L.hrnn_dev_rknn_search_noalloc_dptr, qq, p, 10, 48, thr, 4096)
        b = _run(dev, L.hrnn_dev_rknn_search_dptr, qq, p, 10, 48, thr, 4096)
        assert np.array_equal(a[0], b[0]), nq
        assert all(np.array_equal(u, v) for u, v in zip(a[1], b[1])), nq
        assert np.array_equal(a[2], b[2]), nq
        assert int(a[0].sum()) > 0
        if lds_kib == 1:
            assert int((a[2][0::2] * 4 > 3 * 256).sum()) > nq // 4   # the fallback is taken
    dev.close()


def test_noalloc_search_layout_changes_between_calls(small_data, small_ref):
    """The scratch array's bitmap area moves with the batch and the mode (its
    size is workgroups x n/32 words in presence mode, x n/2 in count mode),
    so one call's bitmaps can land where an earlier call kept candidate ids:
    alternating modes and batch sizes, every call still equals the
    allocating search."""
    x, q = small_data
    r, _ = small_ref
    ids, dists, cnt = r.get_knng()
    dev = hrnn_amd.Device(0)
    dev.set_base(x)
    dev.set_knng(ids, dists, cnt)
    dev.build_rknn_csr()
    dev.set_tuning("search_lds_kib", 1)
    L = dev.L
    prox = r.proxies(q, 10, 64)
    for nq, thr in [(len(q), 1.0), (len(q), 2.0), (len(q) // 3, 1.0), (len(q), 3.0), (7, 2.0), (len(q), 1.0)]:
        p = np.ascontiguousarray(prox[:nq])
        qq = np.ascontiguousarray(q[:nq])
        a = _run(dev, L.hrnn_dev_rknn_search_noalloc_dptr, qq, p, 10, 48, thr, 4096)
        b = _run(dev, L.hrnn_dev_rknn_search_dptr, qq, p, 10, 48, thr, 4096)
        assert np.array_equal(a[0], b[0]), (nq, thr)
        assert all(np.array_equal(u, v) for u, v in zip(a[1], b[1])), (nq, thr)
        assert np.array_equal(a[2], b[2]), (nq, thr)
    dev.close()
