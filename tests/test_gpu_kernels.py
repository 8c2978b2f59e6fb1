"""GPU parity: each HIP kernel through the C-ABI against the CPU oracle.

Bars: bit-exact for ids, CSR entries and integer-data distances; FP32
distances in the reference's summation order bit-exact; NNDescent (a
randomised, race-ordered algorithm in the reference too) within 0.5 recall
points of the oracle's single-threaded restatement.
"""
import ctypes

import numpy as np
import pytest

import hrnn_amd
import oracle
from conftest import mixture

pytestmark = pytest.mark.gpu


def _canon(off, ent, rb):
    out = ent.copy()
    mask = (1 << rb) - 1
    for dst in range(len(off) - 1):
        seg = ent[off[dst]:off[dst + 1]].astype(np.int64)
        keys = np.lexsort((seg >> rb, seg & mask))
        out[off[dst]:off[dst + 1]] = seg[keys].astype(ent.dtype)
    return out


# ---- distances -------------------------------------------------------------------

@pytest.mark.parametrize("d", [96, 128, 960, 40])
def test_l2_pairs_reference_order(d):
    rng = np.random.default_rng(d)
    dev = hrnn_amd.Device(0)
    x = rng.standard_normal((300, d)).astype(np.float32)
    a = rng.standard_normal((50, d)).astype(np.float32) * 2
    dev.set_base(x)
    ia = rng.integers(0, 50, 2000).astype(np.uint32)
    ib = rng.integers(0, 300, 2000).astype(np.uint32)
    got = dev.l2sqr_pairs(a, ia, ib)
    W = 16 if d % 16 == 0 else 1
    exp = np.array([oracle.l2sqr(a[i], x[j], W=W) for i, j in zip(ia, ib)], np.float32)
    assert np.array_equal(got, exp)


# ---- reverse CSR -------------------------------------------------------------------

def _graph(n, K, seed, full=False):
    rng = np.random.default_rng(seed)
    cnt = np.full(n, K, np.uint16) if full else rng.integers(0, K + 1, n).astype(np.uint16)
    ids = np.argsort(rng.random((n, n)) if n <= 3000 else rng.random((n, K * 2)), axis=1)[:, :K]
    return ids.astype(np.uint32), cnt


def _csr_check(dev, x, ids, cnt, K):
    dev.set_base(x)
    dev.set_knng(ids, np.zeros(ids.shape, np.float32), cnt)
    info = dev.build_rknn_csr()
    off, ent = dev.get_rknn_csr()
    eoff, eent, erb, ewide = oracle.csr(ids, cnt, K)
    assert (info.rank_bits, info.wide) == (erb, ewide)
    assert np.array_equal(off, eoff)
    assert np.array_equal(ent, eent)
    return off, ent


def test_csr_bit_exact_random():
    dev = hrnn_amd.Device(0)
    x = np.zeros((2500, 16), np.float32)
    ids, cnt = _graph(2500, 100, 1)
    _csr_check(dev, x, ids, cnt, 100)


def test_csr_bit_exact_reference_graph(small_data, small_ref):
    x, _ = small_data
    r, _ = small_ref
    ids, dists, cnt = r.get_knng()
    dev = hrnn_amd.Device(0)
    off, ent = _csr_check(dev, x, ids, cnt, r.K)
    roff, rent, rb, _ = r.get_csr()
    assert np.array_equal(ent, _canon(roff, rent, rb))


def test_csr_long_lists_and_empty():
    """Hub lists longer than the LDS sort capacity, and nodes with no in-edges."""
    n, K = 40000, 8
    rng = np.random.default_rng(9)
    ids = rng.integers(5, n, (n, K)).astype(np.uint32)
    ids[:, 0] = 0                       # node 0: in-degree n at rank 0
    ids[: n // 2, 1] = 1                # node 1: n/2 at rank 1
    ids[:, 2] = np.where(np.arange(n) % 2 == 0, 2, ids[:, 2])
    for i in range(n):                  # keep each row duplicate-free
        row = ids[i]
        _, first = np.unique(row, return_index=True)
        dup = np.setdiff1d(np.arange(K), first)
        row[dup] = (row[dup] + 7 + i) % (n - 5) + 5
    cnt = np.full(n, K, np.uint16)
    cnt[::3] = 3
    dev = hrnn_amd.Device(0)
    _csr_check(dev, np.zeros((n, 16), np.float32), ids, cnt, K)


def test_csr_k500_shape():
    """K = 500 (rank_bits 9) on a graph with ties of in-degree across ranks."""
    n, K = 3000, 500
    ids, cnt = _graph(n, K, 4, full=True)
    dev = hrnn_amd.Device(0)
    _csr_check(dev, np.zeros((n, 16), np.float32), ids, cnt, K)


@pytest.mark.parametrize("n,K", [(4300, 4096), (4600, 4200)])
def test_csr_rows_at_and_past_a_radix_tile(n, K):
    """K = 4096: one KNN row fills a 4,096-element radix tile exactly; K =
    4200 (rank_bits 13, u32 entries): a row is longer than a tile, so the
    build takes the count / fill / per-list sort path; bit-exact both."""
    rng = np.random.default_rng(12)
    ids = np.argsort(rng.random((n, n), dtype=np.float32), axis=1)[:, :K].astype(np.uint32)
    cnt = rng.integers(K - 300, K + 1, n).astype(np.uint16)
    dev = hrnn_amd.Device(0)
    _csr_check(dev, np.zeros((n, 16), np.float32), ids, cnt, K)


def test_csr_deterministic_across_runs():
    dev = hrnn_amd.Device(0)
    x = np.zeros((2000, 16), np.float32)
    ids, cnt = _graph(2000, 64, 2)
    dev.set_base(x)
    dev.set_knng(ids, np.zeros(ids.shape, np.float32), cnt)
    dev.build_rknn_csr()
    a = dev.get_rknn_csr()
    dev.build_rknn_csr()
    b = dev.get_rknn_csr()
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


# ---- search phases B + C ------------------------------------------------------------

@pytest.mark.parametrize("m,Kp,k,thr", [(1, 48, 10, 1.0), (5, 20, 10, 1.0), (10, 48, 5, 1.0),
                                         (10, 48, 10, 2.0), (20, 48, 48, 1.0), (3, 1, 1, 1.0)])
def test_search_vs_oracle(small_data, small_ref, m, Kp, k, thr):
    x, q = small_data
    r, _ = small_ref
    ids, dists, cnt = r.get_knng()
    dev = hrnn_amd.Device(0)
    dev.set_base(x)
    dev.set_knng(ids, dists, cnt)
    dev.build_rknn_csr()
    off, ent = dev.get_rknn_csr()
    prox = r.proxies(q, m, 64)
    goff, gids = dev.rknn_search(q, prox, k, Kp, thr)
    eoff, eids = oracle.search_bc(off, ent, dev.csr.wide, dev.csr.rank_bits, dists, cnt, r.K, x,
                                  q, prox, k, Kp, thr)
    assert np.array_equal(goff, eoff)
    assert np.array_equal(gids, eids)
    if thr == 1.0 and Kp == r.K:
        # end to end against the reference's own searchRknn
        roff, rlab = r.search(q, m, thr, k, Kp, 64)
        assert np.array_equal(gids.astype(np.uint64), rlab)


@pytest.mark.parametrize("lds_kib,thr,qshift", [(1, 1.0, 0.0), (1, 2.0, 0.0), (32, 1.0, 0.5),
                                                 (1, 3.0, 0.25), (8, 2.0, 0.0)])
def test_search_lds_table_and_hbm_fallback(small_data, small_ref, lds_kib, thr, qshift):
    """A 256-slot LDS table sends every query with more than 192 scanned entries
    to the HBM visited-bitmap path, in the same launch as the queries (every
    second one, with 2 proxies) that stay in LDS; shifted queries are not
    integers, so their distances take the reference's 16-partial order instead
    of the exact 16-byte-row form."""
    x, q = small_data
    q = (q + np.float32(qshift)).astype(np.float32)
    r, _ = small_ref
    ids, dists, cnt = r.get_knng()
    dev = hrnn_amd.Device(0)
    dev.set_base(x)
    dev.set_knng(ids, dists, cnt)
    dev.build_rknn_csr()
    dev.set_tuning("search_lds_kib", lds_kib)
    assert dev.get_tuning("search_lds_kib") == lds_kib
    off, ent = dev.get_rknn_csr()
    prox = r.proxies(q, 8, 64)
    prox[::2, 2:] = 0xFFFFFFFF
    goff, gids = dev.rknn_search(q, prox, 10, 48, thr)
    eoff, eids = oracle.search_bc(off, ent, dev.csr.wide, dev.csr.rank_bits, dists, cnt, r.K, x,
                                  q, prox, 10, 48, thr)
    assert np.array_equal(goff, eoff)
    assert np.array_equal(gids, eids)


@pytest.mark.parametrize("m,thr", [(5, 1.0), (10, 2.0)])
def test_search_resident_and_fetch_results(small_data, small_ref, m, thr):
    """The resident search with its answers packed on the device and
    downloaded by hrnn_dev_fetch_results equals the oracle; a per-query
    capacity below the largest answer is reported (HRNN_E_CAP, needed = the
    largest count) instead of returning cut lists; an empty batch returns
    off = [0]."""
    x, q = small_data
    r, _ = small_ref
    ids, dists, cnt = r.get_knng()
    dev = hrnn_amd.Device(0)
    dev.set_base(x)
    dev.set_knng(ids, dists, cnt)
    dev.build_rknn_csr()
    off, ent = dev.get_rknn_csr()
    prox = r.proxies(q, m, 64)
    goff, gids = dev.rknn_search_resident(q, prox, 10, 48, thr)
    eoff, eids = oracle.search_bc(off, ent, dev.csr.wide, dev.csr.rank_bits, dists, cnt, r.K, x,
                                  q, prox, 10, 48, thr)
    assert np.array_equal(goff, eoff) and np.array_equal(gids, eids)
    big = int(np.diff(eoff).max())
    assert big >= 2
    with pytest.raises(hrnn_amd.HrnnError, match="cap_per_q"):
        dev.rknn_search_resident(q, prox, 10, 48, thr, cap_per_q=big - 1)
    goff2, gids2 = dev.rknn_search_resident(q, prox, 10, 48, thr, cap_per_q=big)
    assert np.array_equal(goff2, eoff) and np.array_equal(gids2, eids)
    z_off, z_ids = dev.rknn_search_resident(q[:0], prox[:0], 10, 48, thr)
    assert z_off.tolist() == [0] and z_ids.size == 0


def test_search_hub_results_above_m_times_kprime():
    """A proxy whose reverse list is far longer than K' (node 0 is everyone's
    nearest neighbour): one proxy and K' = 1 return ~2000 ids, which the
    host-buffer entry point must size by the longest reverse list."""
    n, K, d = 2000, 4, 16
    rng = np.random.default_rng(9)
    x = rng.integers(0, 256, (n, d)).astype(np.float32)
    ids = np.zeros((n, K), np.uint32)
    for i in range(n):
        ids[i] = [0 if i else 1] + list(rng.choice(np.arange(2, n), K - 1, replace=False))
    dists = np.full((n, K), 1e30, np.float32)            # every candidate is accepted
    cnt = np.full(n, K, np.uint16)
    dev = hrnn_amd.Device(0)
    dev.set_base(x)
    dev.set_knng(ids, dists, cnt)
    dev.build_rknn_csr()
    off, ent = dev.get_rknn_csr()
    q = rng.integers(0, 256, (3, d)).astype(np.float32)
    prox = np.zeros((3, 1), np.uint32)
    goff, gids = dev.rknn_search(q, prox, 1, 1, 1.0)
    eoff, eids = oracle.search_bc(off, ent, dev.csr.wide, dev.csr.rank_bits, dists, cnt, K, x, q, prox,
                                  1, 1, 1.0)
    assert int(goff[1] - goff[0]) == n - 1
    assert np.array_equal(goff, eoff) and np.array_equal(gids, eids)


def test_search_float_data_and_k_per_query(float_data):
    """Float vectors: accept/reject uses the reference's FP32 order, so results are identical."""
    x, q = float_data
    n, K = x.shape[0], 24
    ids, d2 = oracle.knn(x, x, K + 1)
    ids, d2 = ids[:, 1:].copy(), d2[:, 1:].copy()       # drop self
    cnt = np.full(n, K, np.uint16)
    cnt[::7] = 5                                         # some short lists
    dev = hrnn_amd.Device(0)
    dev.set_base(x)
    dev.set_knng(ids, d2, cnt)
    dev.build_rknn_csr()
    off, ent = dev.get_rknn_csr()
    rng = np.random.default_rng(3)
    prox = rng.integers(0, n, (q.shape[0], 12)).astype(np.uint32)
    prox[0, 5:] = 0xFFFFFFFF                             # a query with fewer proxies
    kq = rng.choice([1, 5, 10], q.shape[0]).astype(np.uint32)
    goff, gids = dev.rknn_search(q, prox, 0, 16, 1.0, k_per_q=kq)
    eoff, eids = oracle.search_bc(off, ent, 0, dev.csr.rank_bits, d2, cnt, K, x, q, prox, 0, 16,
                                  1.0, k_per_q=kq)
    assert np.array_equal(goff, eoff) and np.array_equal(gids, eids)


def test_search_rejects_bad_k(small_data, small_ref):
    x, q = small_data
    r, _ = small_ref
    ids, dists, cnt = r.get_knng()
    dev = hrnn_amd.Device(0)
    dev.set_base(x)
    dev.set_knng(ids, dists, cnt)
    with pytest.raises(hrnn_amd.HrnnError):
        dev.rknn_search(q, np.zeros((q.shape[0], 2), np.uint32), 10, 10)   # no CSR yet
    dev.build_rknn_csr()
    with pytest.raises(hrnn_amd.HrnnError):
        dev.rknn_search(q, np.zeros((q.shape[0], 2), np.uint32), r.K + 1, 10)


def test_set_knng_rejects_out_of_range_ids(small_data, small_ref):
    """Listed ids >= n are refused (checked on the device after the upload);
    ids past a row's count are not looked at; a valid graph loads again."""
    x, _ = small_data
    r, _ = small_ref
    ids, dists, cnt = r.get_knng()
    dev = hrnn_amd.Device(0)
    dev.set_base(x)
    bad = ids.copy()
    bad[1234, int(cnt[1234]) - 1] = x.shape[0]
    with pytest.raises(hrnn_amd.HrnnError):
        dev.set_knng(bad, dists, cnt)
    with pytest.raises(hrnn_amd.HrnnError):
        dev.nndescent(1, 8)   # no usable graph after the refusal
    tail = ids.copy()
    short = cnt.copy()
    short[7] = cnt[7] - 1
    tail[7, int(short[7])] = 0xFFFFFFFF   # past the row's count: ignored
    dev.set_knng(tail, dists, short)
    dev.set_knng(ids, dists, cnt)
    dev.build_rknn_csr()


# ---- brute force ---------------------------------------------------------------------

@pytest.mark.parametrize("prec", ["bf16", "f16", "f32"])
def test_bf_kth_and_ids_integer_data(small_data, prec):
    x, _ = small_data
    dev = hrnn_amd.Device(0)
    dev.set_base(x)
    assert dev.base_is_small_int()
    k = 10
    kth, ids = dev.bf_kth(k, prec, with_ids=True)
    eids, ed = oracle.knn(x, x, k + 1)
    assert np.array_equal(ids, eids)
    assert np.array_equal(kth, ed[:, k])


@pytest.mark.parametrize("prec", ["bf16", "f32"])
def test_bf_knn_queries(small_data, prec):
    x, q = small_data
    dev = hrnn_amd.Device(0)
    dev.set_base(x)
    ids, d2 = dev.bf_knn(q, 17, prec)
    eids, ed = oracle.knn(x, q, 17)
    assert np.array_equal(ids, eids) and np.array_equal(d2, ed)


def test_bf_kth_float_data_f32(float_data):
    x, _ = float_data
    dev = hrnn_amd.Device(0)
    dev.set_base(x)
    kth, ids = dev.bf_kth(10, "f32", with_ids=True)
    eids, ed = oracle.knn(x, x, 11)
    assert np.array_equal(ids, eids)
    assert np.array_equal(kth, ed[:, 10])


@pytest.mark.parametrize("prec", ["bf16", "f16"])
def test_bf_kth_float_data_certified(float_data, prec):
    """BF16/FP16 candidates with the error-bound certificate (rows it does not
    certify are redone wider, then in FP32) and the FP32 re-rank: on float
    data every row's ids and r_k equal the exhaustive FP32 brute force."""
    x, _ = float_data
    dev = hrnn_amd.Device(0)
    dev.set_base(x)
    kth, ids = dev.bf_kth(10, prec, with_ids=True)
    eids, ed = oracle.knn(x, x, 11)
    assert np.array_equal(ids, eids)
    assert np.array_equal(kth, ed[:, 10])


@pytest.mark.parametrize("prec", ["bf16", "f32"])
def test_bf_rknn_gt(small_data, prec):
    x, q = small_data
    dev = hrnn_amd.Device(0)
    dev.set_base(x)
    kth = oracle.kth(x, 10)
    off, ids, d2 = dev.bf_rknn(q, kth, prec)
    eoff, eids, ed = oracle.rknn_gt(x, q, kth)
    assert np.array_equal(off, eoff)
    assert np.array_equal(ids, eids) and np.array_equal(d2, ed)


def test_bf_rknn_float_data(float_data):
    x, q = float_data
    dev = hrnn_amd.Device(0)
    dev.set_base(x)
    kth = oracle.kth(x, 10)
    off, ids, d2 = dev.bf_rknn(q, kth, "f32")
    eoff, eids, ed = oracle.rknn_gt(x, q, kth)
    assert np.array_equal(off, eoff) and np.array_equal(ids, eids)


# ---- NNDescent --------------------------------------------------------------------------

def _knn_recall(ids, cnt, K, exact, kk):
    return oracle.knng_recall(ids, cnt, K, exact, kk)


def test_nndescent_recall_parity(small_data, small_ref):
    """GPU refineKNNG vs the oracle's single-threaded restatement from the same HNSW seeds."""
    x, _ = small_data
    r, seed = small_ref
    sid, sd, sc = seed
    K, iters, S = r.K, 3, 12
    oid, od, oc, _, _ = oracle.nndescent(x, sid, sd, sc, K, iters, S)
    dev = hrnn_amd.Device(0)
    dev.set_base(x)
    dev.set_knng(sid, sd, sc)
    st = dev.nndescent(iters, S)
    assert all(t["dropped"] == 0 for t in st)
    gid, gd, gc = dev.get_knng()
    exact, _ = oracle.knn(x, x, K + 1)
    exact = exact[:, 1:].copy()
    for kk in (10, K):
        ro = _knn_recall(oid, oc, K, exact, kk)
        rg = _knn_recall(gid, gc, K, exact, kk)
        assert abs(rg - ro) <= 0.005, (kk, rg, ro)
    # every list is sorted, duplicate-free, self-free, with exact distances
    for i in range(0, x.shape[0], 17):
        c = int(gc[i])
        assert np.all(np.diff(gd[i, :c]) >= 0)
        assert len(set(gid[i, :c].tolist())) == c and i not in gid[i, :c]
        dd = ((x[gid[i, :c]].astype(np.int64) - x[i].astype(np.int64)) ** 2).sum(1)
        assert np.array_equal(dd.astype(np.float32), gd[i, :c])
    assert st[0]["dists"] > 0 and st[0]["updates"] > 0


def _first_iteration_matches_oracle(x, sid, sd, sc, K, S):
    n = x.shape[0]
    # reverse offers per node in this iteration (all seed entries are new)
    worst = np.where(sc >= K, sd[np.arange(n), np.maximum(sc.astype(np.int64) - 1, 0)], np.inf)
    offers = np.zeros(n, np.int64)
    for i in range(n):
        c = int(sc[i])
        nb = sid[i, :c]
        np.add.at(offers, nb[sd[i, :c] > worst[nb]], 1)
    assert offers.max() <= 256        # every offer is kept: the reverse samples match exactly
    oid, od, oc, ou, odist = oracle.nndescent(x, sid, sd, sc, K, 1, S)
    dev = hrnn_amd.Device(0)
    dev.set_base(x)
    dev.set_knng(sid, sd, sc)
    st = dev.nndescent(1, S)
    gid, gd, gc = dev.get_knng()
    assert st[0]["dropped"] == 0
    assert st[0]["dists"] == int(odist[0])
    assert np.array_equal(gc, oc)
    m = np.arange(K) < gc[:, None]
    assert np.array_equal(np.where(m, gd, 0), np.where(m, od, 0))
    same_ids = float((np.where(m, gid, 0) == np.where(m, oid, 0)).all(1).mean())
    print("nodes with identical id lists:", same_ids)


def _offers_fit(ids, d, cnt, flags, K):
    """Reverse offers per node in the next iteration (new and old separately)
    all fit the 256 the sampling keeps, so the reverse samples are exact."""
    n = ids.shape[0]
    worst = np.where(cnt >= K, d[np.arange(n), np.maximum(cnt.astype(np.int64) - 1, 0)], np.inf)
    for want in (1, 0):
        offers = np.zeros(n, np.int64)
        for i in range(n):
            c = int(cnt[i])
            nb, dd, fl = ids[i, :c], d[i, :c], flags[i, :c]
            np.add.at(offers, nb[(dd > worst[nb]) & (fl == want)], 1)
        assert offers.max() <= 256, (want, offers.max())


def _resume_matches_oracle(x, ids, d, cnt, flags, K, S, iters):
    """From a state with flags: the GPU and the oracle evaluate the same
    number of pairs per iteration and end with the same distance lists."""
    _offers_fit(ids, d, cnt, flags, K)
    oid, od, oc, ofl, ou, odist = oracle.nndescent_flags(x, ids, d, cnt, flags, K, iters, S)
    dev = hrnn_amd.Device(0)
    dev.set_base(x)
    dev.set_knng(ids, d, cnt)
    dev.set_knng_flags(flags)
    st = dev.nndescent(iters, S)
    gid, gd, gc = dev.get_knng()
    assert all(t["dropped"] == 0 for t in st)
    assert [int(t["dists"]) for t in st] == [int(v) for v in odist]
    assert np.array_equal(gc, oc)
    m = np.arange(K) < gc[:, None]
    assert np.array_equal(np.where(m, gd, 0), np.where(m, od, 0))
    return st, (gid, gd, gc, dev.get_knng_flags()), (oid, od, oc, ofl)


def test_nndescent_second_iteration_exact(small_data, small_ref):
    """Iteration 2 from the GPU's own state after iteration 1 (lists and "new"
    flags read back): the old-entry sampling pass, the need bits and the
    merge's already-listed filter all act, and the GPU still evaluates
    exactly the oracle's pairs and ends with its distance lists.  Where the
    two first iterations left identical id lists, the flags agree as well
    (entries inserted in the iteration are new, all others old)."""
    x, _ = small_data
    r, seed = small_ref
    sid, sd, sc = seed
    K, S = r.K, 12
    dev = hrnn_amd.Device(0)
    dev.set_base(x)
    dev.set_knng(sid, sd, sc)
    dev.nndescent(1, S)
    g1 = dev.get_knng()
    f1 = dev.get_knng_flags()
    oid, od, oc, ofl, _, _ = oracle.nndescent_flags(x, sid, sd, sc, np.ones_like(f1), K, 1, S)
    m = np.arange(K) < oc[:, None]
    same = (np.where(m, g1[0], 0) == np.where(m, oid, 0)).all(1) & (g1[2] == oc)
    assert same.mean() > 0.5
    assert np.array_equal(np.where(m, f1, 0)[same], np.where(m, ofl, 0)[same])
    assert f1[m].any() and not f1[m].all()
    st, _, _ = _resume_matches_oracle(x, g1[0], g1[1], g1[2], f1, K, S, 1)
    assert st[0]["dists"] > 0 and st[0]["updates"] > 0


def test_nndescent_iteration_without_updates(small_data, small_ref):
    """From the oracle's state after 10 iterations (lists and flags), the
    next iteration evaluates a few pairs and inserts nothing, so the one
    after has nothing new: the GPU evaluates the oracle's pairs, skips the
    second iteration's sampling pass (nothing is new after an iteration
    without insertions), both sides evaluate zero pairs there, and the lists
    stay as they were."""
    x, _ = small_data
    r, seed = small_ref
    sid, sd, sc = seed
    K, S = r.K, 12
    ones = np.ones((len(x), K), np.uint8)
    cid, cd, cc, cf, _, _ = oracle.nndescent_flags(x, sid, sd, sc, ones, K, 10, S)
    _, _, _, _, ou, odist = oracle.nndescent_flags(x, cid, cd, cc, cf, K, 2, S)
    assert int(ou[0]) == 0 and int(odist[0]) > 0 and int(odist[1]) == 0, (ou, odist)
    st, (gid, gd, gc, gf), _ = _resume_matches_oracle(x, cid, cd, cc, cf, K, S, 2)
    assert st[0]["updates"] == 0 and st[1]["dists"] == 0 and st[1]["updates"] == 0
    assert np.array_equal(gc, cc)
    m = np.arange(K) < cc[:, None]
    assert np.array_equal(np.where(m, gid, 0), np.where(m, cid, 0))
    assert not gf[m].any()


@pytest.mark.parametrize("join_mode", [None, "1", "0"])
def test_nndescent_first_iteration_exact(small_data, small_ref, join_mode, monkeypatch):
    """One iteration from the same seeds: the GPU evaluates exactly the pairs
    of the oracle's single-threaded join (same forward and reverse samples)
    and every node ends with exactly the oracle's sorted distance list (the
    top-K of list U proposals is unique as distances; ids may differ only
    among equal distances).  Nothing is dropped.  At d=128, S=12 the join's
    three staging modes are run: int8 rows (the default for 8-bit data), the
    whole sample as BF16 (HRNN_NND_JOIN=1) and 64-dim BF16 chunks (=0)."""
    if join_mode is not None:
        monkeypatch.setenv("HRNN_NND_JOIN", join_mode)
    x, _ = small_data
    r, seed = small_ref
    sid, sd, sc = seed
    _first_iteration_matches_oracle(x, sid, sd, sc, r.K, 12)


@pytest.mark.parametrize("join_mode", [None, "0"])
def test_nndescent_first_iteration_exact_wide_dims(join_mode, monkeypatch):
    """As above at d=256, S=32: int8 rows (default; every distance an exact
    integer below 2^24) and 64-dim BF16 chunks with one accumulator per Gram
    block (HRNN_NND_JOIN=0), where d > 128 also takes the FP32 re-evaluation
    of every candidate pair."""
    if not oracle.ref_available():
        pytest.skip("compiled reference (oracle/_ref) not available")
    if join_mode is not None:
        monkeypatch.setenv("HRNN_NND_JOIN", join_mode)
    x = mixture(1000, d=256, seed=41)
    r = oracle.RefIndex(1000, 256, M=8, efc=40, K=40)
    r.add_points(x, nthreads=1)
    sid, sd, sc = r.get_knng()
    _first_iteration_matches_oracle(x, sid, sd, sc, 40, 32)


def test_nndescent_graph_capture_matches_eager(small_data, small_ref, monkeypatch):
    """The iterations captured as one HIP graph (default) and launched eagerly
    (HRNN_NND_GRAPH=0) reach the same KNN recall; the graph is launched twice
    in a row with host work in between."""
    x, _ = small_data
    r, seed = small_ref
    sid, sd, sc = seed
    exact, _ = oracle.knn(x, x, r.K + 1)
    exact = exact[:, 1:].copy()
    rec = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("HRNN_NND_GRAPH", mode)
        dev = hrnn_amd.Device(0)
        dev.set_base(x)
        dev.set_knng(sid, sd, sc)
        st1 = dev.nndescent(2, 12)
        mid = dev.get_knng()
        st2 = dev.nndescent(2, 12)
        gid, gd, gc = dev.get_knng()
        rec[mode] = _knn_recall(gid, gc, r.K, exact, r.K)
        assert st1[0]["dists"] > 0 and st2[0]["dists"] > 0
        assert _knn_recall(mid[0], mid[2], r.K, exact, r.K) <= rec[mode] + 1e-9
    assert abs(rec["1"] - rec["0"]) <= 0.005, rec


def test_nndescent_non_full_seed():
    """Lists that start below K (the K=500 > ef_construction=400 regime) fill to K."""
    x = mixture(1200, d=64, seed=31)
    r = oracle.RefIndex(1200, 64, M=8, efc=24, K=40)
    r.add_points(x, nthreads=1)
    sid, sd, sc = r.get_knng()
    oid, od, oc, _, _ = oracle.nndescent(x, sid, sd, sc, 40, 3, 8)
    dev = hrnn_amd.Device(0)
    dev.set_base(x)
    dev.set_knng(sid, sd, sc)
    dev.nndescent(3, 8)
    gid, gd, gc = dev.get_knng()
    exact, _ = oracle.knn(x, x, 41)
    exact = exact[:, 1:].copy()
    ro = _knn_recall(oid, oc, 40, exact, 40)
    rg = _knn_recall(gid, gc, 40, exact, 40)
    assert abs(rg - ro) <= 0.005, (rg, ro)
    assert int(gc.min()) == int(oc.min())


@pytest.mark.parametrize("qshift", [0.0, 0.5])
def test_search_u8_storage_matches_oracle(small_data, small_ref, qshift):
    """u8 vector storage (HRNN_PREC_U8) on 8-bit data: integer queries use the
    v_dot4 form, shifted queries the reference's 16-partial order over the
    bytes; both give the oracle's answers exactly, through the LDS table and
    the HBM fallback."""
    x, q = small_data
    q = (q + np.float32(qshift)).astype(np.float32)
    r, _ = small_ref
    ids, dists, cnt = r.get_knng()
    dev = hrnn_amd.Device(0)
    dev.set_base(x)
    dev.set_knng(ids, dists, cnt)
    dev.build_rknn_csr()
    dev.set_vector_precision("u8")
    off, ent = dev.get_rknn_csr()
    prox = r.proxies(q, 8, 64)
    prox[::2, 2:] = 0xFFFFFFFF
    for kib in (16, 1):
        dev.set_tuning("search_lds_kib", kib)
        for thr in (1.0, 2.0):
            goff, gids = dev.rknn_search(q, prox, 10, 48, thr)
            eoff, eids = oracle.search_bc(off, ent, dev.csr.wide, dev.csr.rank_bits, dists, cnt, r.K, x,
                                          q, prox, 10, 48, thr)
            assert np.array_equal(goff, eoff) and np.array_equal(gids, eids), (kib, thr)


@pytest.mark.parametrize("join_mode", ["1", "0"])
def test_nndescent_float_data_exact(join_mode, monkeypatch):
    """Float data (96-dim unit vectors: no BF16 Gram value is exact): one
    iteration from the reference's own seeds, then a second from the GPU's
    state with its flags.  The Gram values with their error margin choose
    the pairs whose exact FP32 distance is computed, so no pair below a
    threshold is missed: the GPU evaluates exactly the oracle's pairs and ends
    each iteration with exactly its distance lists (ids agree too: the
    distances are distinct), in the whole-sample (1) and chunked (0) BF16
    stagings."""
    if not oracle.ref_available():
        pytest.skip("compiled reference (oracle/_ref) not available")
    monkeypatch.setenv("HRNN_NND_JOIN", join_mode)
    n, K, S = 2000, 32, 12
    x = hrnn_amd.synth_sphere_mixture(31, 88, n, 96, 40, 50.0)
    r = oracle.RefIndex(n, 96, M=8, efc=40, K=K)
    r.add_points(x, nthreads=1)
    sid, sd, sc = r.get_knng()
    _first_iteration_matches_oracle(x, sid, sd, sc, K, S)
    dev = hrnn_amd.Device(0)
    dev.set_base(x)
    dev.set_knng(sid, sd, sc)
    st1 = dev.nndescent(1, S)
    g1 = dev.get_knng()
    f1 = dev.get_knng_flags()
    oid, od, oc, _, _, _ = oracle.nndescent_flags(x, sid, sd, sc, np.ones_like(f1), K, 1, S)
    m = np.arange(K) < oc[:, None]
    assert np.array_equal(np.where(m, g1[0], 0), np.where(m, oid, 0))
    # the proposals the GPU counts are those it made (exact distance below
    # the threshold), not the slots the margin reserved
    assert 0 < st1[0]["proposals"] <= 2 * st1[0]["dists"]
    st, _, _ = _resume_matches_oracle(x, g1[0], g1[1], g1[2], f1, K, S, 1)
    assert st[0]["dists"] > 0 and st[0]["updates"] > 0


def test_fetch_results_into_pinned_memory(small_data, small_ref):
    """hrnn_dev_fetch_results with page-locked host buffers (the gather
    writes the ids straight into them, one wait): the same offsets and ids as
    into pageable arrays; a capacity one short of the total is reported
    (HRNN_E_CAP, needed = the total) and writes nothing past the capacity;
    the exact capacity then succeeds."""
    from hrnn_amd import hiprt
    x, q = small_data
    r, _ = small_ref
    ids, dists, cnt = r.get_knng()
    dev = hrnn_amd.Device(0)
    dev.set_base(x)
    dev.set_knng(ids, dists, cnt)
    dev.build_rknn_csr()
    m, cap_q = 5, 256
    prox = np.ascontiguousarray(r.proxies(q, m, 64), np.uint32)
    nq = len(q)
    qd = hiprt.DeviceArray.from_host(np.ascontiguousarray(q, np.float32))
    pd = hiprt.DeviceArray.from_host(prox)
    cd = hiprt.DeviceArray((nq,), np.uint32)
    idd = hiprt.DeviceArray((nq * cap_q,), np.uint32)
    L = dev.L
    assert L.hrnn_dev_rknn_search_dptr(dev.h, qd.ptr, nq, pd.ptr, m, None, 10, 48, 1.0, cd.ptr, idd.ptr, cap_q,
                                       None) == 0
    want_off, want_ids = dev.fetch_results(nq, cd.ptr, idd.ptr, cap_q)
    total = int(want_off[-1])
    assert total > 10
    off_p = hiprt.PinnedArray((nq + 1,), np.uint64)
    ids_p = hiprt.PinnedArray((total + 8,), np.uint32)
    ids_p.a[:] = 0xDEADBEEF
    need = ctypes.c_uint64(0)
    rc = L.hrnn_dev_fetch_results(dev.h, nq, cd.ptr, idd.ptr, cap_q, off_p.ptr, ids_p.ptr, total - 1,
                                  ctypes.byref(need))
    assert rc == -5 and need.value == total
    assert (ids_p.a[total - 1:] == 0xDEADBEEF).all()
    rc = L.hrnn_dev_fetch_results(dev.h, nq, cd.ptr, idd.ptr, cap_q, off_p.ptr, ids_p.ptr, total,
                                  ctypes.byref(need))
    assert rc == 0 and need.value == total
    assert np.array_equal(off_p.a, want_off) and np.array_equal(ids_p.a[:total], want_ids)
    assert (ids_p.a[total:] == 0xDEADBEEF).all()


def test_query_reads_pinned_host_queries(small_data):
    """hrnn_dev_rknn_query_dptr with the queries in page-locked host memory
    (phase A reads them over the link and leaves a device copy that the
    search reads): the same proxies and answers as from device memory, on one
    stream and on two, and again after the pinned batch is overwritten with
    other queries (the device copy is rewritten)."""
    from hrnn_amd import hiprt
    from hrnn_amd.dev import Device
    from hrnn_amd.index import Index
    x, _ = small_data
    n, d, K = x.shape[0], x.shape[1], 32
    ix = Index(n, d, 16, 64, 64, K)
    ix.add_points_gpu(x)
    ix.refine(2, 12)
    ix.build_rknng()
    h = ix.device_handle()
    dev = Device.borrow(h, n, d, K)
    q = np.ascontiguousarray(np.concatenate([x[:1500] + np.float32(1.0), x[:1500] + np.float32(2.0)]), np.float32)
    nq, m, ef, cap = len(q), 5, 64, 512
    q_pin = hiprt.PinnedArray(q.shape, np.float32)
    q_pin.a[:] = q
    qd = hiprt.DeviceArray.from_host(q)
    out = []
    try:
        for streams in (1, 2):
            dev.set_tuning("query_streams", streams)
            for qp in (qd.ptr, q_pin.device_ptr()):
                pd = hiprt.DeviceArray((nq * m,), np.uint32)
                cd = hiprt.DeviceArray((nq,), np.uint32)
                idd = hiprt.DeviceArray((nq * cap,), np.uint32)
                assert dev.L.hrnn_dev_rknn_query_dptr(h, qp, nq, m, ef, None, 10, K, 1.0, pd.ptr, cd.ptr,
                                                      idd.ptr, cap, None) == 0
                assert dev.L.hrnn_dev_sync(h) == 0   # the library's streams, before the plain copy below
                out.append((pd.to_host(),) + dev.fetch_results(nq, cd.ptr, idd.ptr, cap))
        q2 = np.ascontiguousarray(q[::-1] + np.float32(0.5))
        q_pin.a[:] = q2
        qd2 = hiprt.DeviceArray.from_host(q2)
        again = []
        for qp in (q_pin.device_ptr(), qd2.ptr):
            pd = hiprt.DeviceArray((nq * m,), np.uint32)
            cd = hiprt.DeviceArray((nq,), np.uint32)
            idd = hiprt.DeviceArray((nq * cap,), np.uint32)
            assert dev.L.hrnn_dev_rknn_query_dptr(h, qp, nq, m, ef, None, 10, K, 1.0, pd.ptr, cd.ptr,
                                                  idd.ptr, cap, None) == 0
            assert dev.L.hrnn_dev_sync(h) == 0
            again.append((pd.to_host(),) + dev.fetch_results(nq, cd.ptr, idd.ptr, cap))
    finally:
        dev.set_tuning("query_streams", 1)
    for o in out[1:]:
        for a, b in zip(out[0], o):
            assert np.array_equal(a, b)
    assert out[0][2].size > 0
    for a, b in zip(again[0], again[1]):
        assert np.array_equal(a, b)
    assert not np.array_equal(again[0][0], out[0][0])


@pytest.mark.parametrize("thr", [1.0, 2.0])
def test_wide_csr_build_and_search(small_data, small_ref, thr):
    """The u64-entry CSR (node << rank_bits | rank in 64-bit words, the layout
    graphs past 32 bits of node and rank need), forced on a small graph by the
    "csr_wide" key: the count / fill / per-list sort build gives the oracle's
    entries as 64-bit values, and the search over it the oracle's answers, in
    presence and count modes; the u32 layout returns after the key is reset."""
    x, q = small_data
    r, _ = small_ref
    ids, dists, cnt = r.get_knng()
    dev = hrnn_amd.Device(0)
    dev.set_base(x)
    dev.set_knng(ids, dists, cnt)
    dev.set_tuning("csr_wide", 1)
    try:
        info = dev.build_rknn_csr()
        off, ent = dev.get_rknn_csr()
    finally:
        dev.set_tuning("csr_wide", 0)
    eoff, eent, erb, ewide = oracle.csr(ids, cnt, r.K)
    assert info.wide == 1 and ewide == 0 and info.rank_bits == erb and ent.dtype == np.uint64
    assert np.array_equal(off, eoff) and np.array_equal(ent, eent.astype(np.uint64))
    prox = r.proxies(q, 10, 64)
    goff, gids = dev.rknn_search(q, prox, 10, 48, thr)
    xoff, xids = oracle.search_bc(eoff, eent, 0, erb, dists, cnt, r.K, x, q, prox, 10, 48, thr)
    assert np.array_equal(goff, xoff) and np.array_equal(gids, xids) and len(gids) > 0
    assert dev.build_rknn_csr().wide == 0
