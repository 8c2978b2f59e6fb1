"""The autotuner's keys and tuners (hrnn_amd/autotune.py).  Its premise is
that every candidate of a key gives the same results, only faster or
slower; each key is checked for that here, and each tuner runs, returns
one of its candidates and leaves the device state as it found it."""
import numpy as np
import pytest

import hrnn_amd
from hrnn_amd import autotune, hiprt
from hrnn_amd.dev import Device
from hrnn_amd.index import Index

pytestmark = pytest.mark.gpu


def _sphere(n, d, seed):
    return hrnn_amd.synth_sphere_mixture(seed, 88, n, d, 200, 50.0)


def test_tuning_keys_ranges(gpu_dev):
    for key, good, bad in (("beam_vis_shift", (-1, 0, 2), (-2, 3)), ("bf_row_frags", (1, 2), (0, 3)),
                           ("bf_layout", (0, 1), (-1, 2)),
                           ("nnd_join", (0, 1, 3), (-1, 4)), ("query_streams", (1, 2), (0, 3)),
                           ("query_split", (100, 400, 900), (99, 901))):
        for v in good:
            gpu_dev.set_tuning(key, v)
            assert gpu_dev.get_tuning(key) == v
        for v in bad:
            with pytest.raises(Exception):
                gpu_dev.set_tuning(key, v)
    gpu_dev.set_tuning("beam_vis_shift", 0)
    gpu_dev.set_tuning("bf_row_frags", 1)
    gpu_dev.set_tuning("nnd_join", 0)
    gpu_dev.set_tuning("query_streams", 1)
    gpu_dev.set_tuning("query_split", 500)
    gpu_dev.set_tuning("bf_layout", 0)


@pytest.fixture(scope="module")
def sphere_index():
    n, d = 20_000, 96
    x = _sphere(n, d, 5)
    q = _sphere(500, d, 6)
    ix = Index(n, d, 16, 100, 100, 32)
    ix.add_points_gpu(x)
    ix.refine(2, 8)
    ix.build_rknng()
    dev = Device.borrow(ix.device_handle(), n, d, 32)
    yield ix, dev, x, q


def test_beam_vis_shift_same_proxies(sphere_index):
    """Phase A at ef 200 (float rows, about 2,000 ids evaluated per query on
    these data) with the visited set at 2^10 .. 2^13 slots: the same proxies."""
    ix, dev, x, q = sphere_index
    out = []
    try:
        for s in (-1, 0, 1, 2):
            dev.set_tuning("beam_vis_shift", s)
            out.append(dev.proxies(q, 50, 200))
    finally:
        dev.set_tuning("beam_vis_shift", 0)
    for p in out[1:]:
        assert np.array_equal(p, out[0])


@pytest.mark.parametrize("kind", ["int", "sphere"])
@pytest.mark.parametrize("prec", ["bf16", "f16"])
def test_bf_row_frags_same_results(gpu_dev, kind, prec):
    """The MFMA brute force with one and two 16-row fragments per wave (32-wide
    candidate lists: k + 1 <= 16): the same ids and distances, k-NN of
    queries and the base's own r_k."""
    if kind == "int":
        x = hrnn_amd.synth_mixture(7, 6000, 128, 40, 20.0, 0.0, 255.0, True, 0.0, 1007)
    else:
        x = _sphere(6000, 96, 7)
    q = x[:700] + np.float32(0.25)
    gpu_dev.set_base(x)
    res = []
    try:
        for fr in (1, 2):
            gpu_dev.set_tuning("bf_row_frags", fr)
            ids, d2 = gpu_dev.bf_knn(q, 11, prec)
            kth, kids = gpu_dev.bf_kth(10, prec, with_ids=True)
            res.append((ids, d2, kth, kids))
    finally:
        gpu_dev.set_tuning("bf_row_frags", 1)
    for a, b in zip(res[0], res[1]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("kind", ["int", "sphere"])
@pytest.mark.parametrize("prec", ["bf16", "f16"])
def test_bf_layout_same_results(gpu_dev, kind, prec):
    """The MFMA brute force with waves splitting each tile's rows (0) or its
    columns (1): the same k-NN ids and distances with 32- and 64-wide
    candidate lists (k + 1 = 11 and 20), the same certified r_k, and, on
    exact data, the same RkNN hits."""
    if kind == "int":
        x = hrnn_amd.synth_mixture(17, 5000, 128, 40, 20.0, 0.0, 255.0, True, 0.0, 1017)
    else:
        x = _sphere(5000, 96, 17)
    q = x[:900] + np.float32(0.25)
    gpu_dev.set_base(x)
    res = []
    try:
        for lay in (0, 1):
            gpu_dev.set_tuning("bf_layout", lay)
            r = [*gpu_dev.bf_knn(q, 11, prec), *gpu_dev.bf_knn(q, 20, prec)]
            kth, kids = gpu_dev.bf_kth(10, prec, with_ids=True)
            r += [kth, kids]
            if kind == "int":
                off, ids, d2 = gpu_dev.bf_rknn(q, kth, prec)
                r += [off, ids, d2]
            res.append(r)
    finally:
        gpu_dev.set_tuning("bf_layout", 0)
    for a, b in zip(res[0], res[1]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("kind", ["u8", "sphere"])
def test_nnd_join_key_same_lists(gpu_dev, kind):
    """NNDescent from the same seeds with each join staging chosen by the key:
    on float data (distinct distances) two iterations end with the same lists,
    ids included, and the same counters; on 8-bit data (many equal
    distances, where which of two equal ids enters may differ, and with it
    the next iteration's samples) one iteration ends with the same distance
    lists after the same number of evaluations."""
    n, K, S = 3000, 24, 12
    if kind == "u8":
        x = hrnn_amd.synth_mixture(9, n, 128, 30, 20.0, 0.0, 255.0, True, 0.0, 1009)
        modes = (1, 2, 3)
    else:
        x = _sphere(n, 96, 9)
        modes = (1, 2)
    ix = Index(n, x.shape[1], 8, 40, 40, K)
    ix.add_points(x, nthreads=1)
    seed = ix.get_knng()
    gpu_dev.set_base(x)
    assert gpu_dev.base_is_u8() == (kind == "u8")
    out = []
    try:
        for mode in modes:
            gpu_dev.set_tuning("nnd_join", mode)
            gpu_dev.set_knng(*seed)
            st = gpu_dev.nndescent(1 if kind == "u8" else 2, S)
            out.append((gpu_dev.get_knng(), [(s_["dists"], s_["updates"] if kind == "sphere" else 0)
                                             for s_ in st]))
    finally:
        gpu_dev.set_tuning("nnd_join", 0)
    (ids0, d0, c0), st0 = out[0]
    for (ids, d, c), st in out[1:]:
        assert st == st0
        assert np.array_equal(c, c0)
        live = np.arange(K)[None, :] < c0[:, None]
        assert np.array_equal(np.where(live, d, 0), np.where(live, d0, 0))
        if kind == "sphere":
            assert np.array_equal(np.where(live, ids, 0), np.where(live, ids0, 0))


@pytest.mark.parametrize("per_query_k", [False, True])
def test_query_streams_same_answers(sphere_index, per_query_k):
    """A resident batch of 3,001 queries through hrnn_dev_rknn_query_dptr on
    one stream and split in two parts on two streams (halves, and the first
    part 30 % or 90 % of the batch): the same proxies, counts, answers and
    work counters, with one k for the batch and with a k per query."""
    ix, dev, x, q0 = sphere_index
    L = hrnn_amd.lib()
    h = ix.device_handle()
    q = np.concatenate([q0] * 6 + [q0[:1]])
    nq, m, ef, cap = len(q), 8, 64, 2048
    kq = (np.arange(nq) % 3 * 4 + 2).astype(np.uint32)
    qd = hiprt.DeviceArray.from_host(q)
    kd = hiprt.DeviceArray.from_host(kq)
    outs = []
    try:
        for ns, sp in ((1, 500), (2, 500), (2, 300), (2, 900)):
            dev.set_tuning("query_streams", ns)
            dev.set_tuning("query_split", sp)
            pd = hiprt.DeviceArray((nq * m,), np.uint32)
            cd = hiprt.DeviceArray((nq,), np.uint32)
            idd = hiprt.DeviceArray((nq * cap,), np.uint32)
            wd = hiprt.DeviceArray((2 * nq,), np.uint32)
            rc = L.hrnn_dev_rknn_query_dptr(h, qd.ptr, nq, m, ef, kd.ptr if per_query_k else None,
                                            0 if per_query_k else 10, 32, 1.0, pd.ptr, cd.ptr, idd.ptr, cap, wd.ptr)
            assert rc == 0, L.hrnn_dev_last_error(h).decode()
            assert L.hrnn_dev_sync(h) == 0
            cnt = cd.to_host()
            ids = idd.to_host().reshape(nq, cap)
            outs.append((pd.to_host(), cnt, [ids[i, :cnt[i]].copy() for i in range(nq)], wd.to_host()))
    finally:
        dev.set_tuning("query_streams", 1)
        dev.set_tuning("query_split", 500)
    a = outs[0]
    for b in outs[1:]:
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and np.array_equal(a[3], b[3])
        assert all(np.array_equal(u, v) for u, v in zip(a[2], b[2]))
    assert a[1].sum() > 0


def test_tuners_run_and_restore(sphere_index):
    ix, dev, x, q = sphere_index
    L = hrnn_amd.lib()
    h = ix.device_handle()
    nq, m, ef = len(q), 20, 100
    qd = hiprt.DeviceArray.from_host(q)
    pd = hiprt.DeviceArray((nq * m,), np.uint32)
    r = autotune.tune_beam(h, qd, nq, m, ef, pd, {"test": 1}, reps=2, use_cache=False)
    assert r["beam_vis_shift"] in autotune.CANDIDATES["beam_vis_shift"]
    assert dev.get_tuning("beam_vis_shift") == r["beam_vis_shift"]
    r = autotune.tune_bruteforce(dev, x[:2048], 10, "bf16", {"test": 1}, reps=1, use_cache=False)
    assert r["applies"] and r["bf_row_frags"] in (1, 2) and r["bf_layout"] in (0, 1) and len(r["timings_ms"]) == 3
    seed = dev.get_knng()
    flags = dev.get_knng_flags()
    r = autotune.tune_nndescent(dev, 8, {"test": 1}, use_cache=False)
    assert r["nnd_join"] in (1, 2) and len(r["timings_ms"]) == 2
    after = dev.get_knng()
    for a, b in zip(seed, after):
        assert np.array_equal(a, b)
    assert np.array_equal(flags, dev.get_knng_flags())
    ix.build_rknng()
    cd = hiprt.DeviceArray((nq,), np.uint32)
    idd = hiprt.DeviceArray((nq * 1024,), np.uint32)
    pd2 = hiprt.DeviceArray((len(x[:4096]) * m,), np.uint32)
    qd2 = hiprt.DeviceArray.from_host(np.ascontiguousarray(x[:4096]))
    cd2 = hiprt.DeviceArray((4096,), np.uint32)
    idd2 = hiprt.DeviceArray((4096 * 1024,), np.uint32)
    r = autotune.tune_query(h, qd2, 4096, m, ef, 10, 32, pd2, cd2, idd2, 1024, {"test": 1}, reps=2,
                            use_cache=False)
    assert r["query_streams"] in (1, 2) and r["query_split"] in autotune.CANDIDATES["query_split"]
    assert len(r["timings_ms"]) == 1 + len(autotune.CANDIDATES["query_split"])
    assert dev.get_tuning("query_split") == r["query_split"]
    for key in ("beam_vis_shift", "bf_row_frags", "nnd_join", "query_streams", "bf_layout"):
        dev.set_tuning(key, 0 if key not in ("bf_row_frags", "query_streams") else 1)
    dev.set_tuning("query_split", 500)
