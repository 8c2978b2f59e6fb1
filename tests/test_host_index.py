"""Host HRNN<float> (include/hnswlib/hrnn.h) on the CPU: HNSW build quality,
index file compatibility with the reference (hrnn.h:861-1047) and error paths."""
import os

import numpy as np
import pytest

import hrnn_amd
import oracle
from conftest import mixture
from hrnn_amd.index import Index

needs_ref = pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built")


@pytest.fixture(scope="module")
def data():
    return mixture(2500, seed=51, centre_seed=1051), mixture(40, seed=52, centre_seed=1051)


@needs_ref
def test_hnsw_seed_matches_reference_single_thread(data):
    """addPoint KNN seeding (hrnn.h:1479-1501): same counts, same ids up to equal-distance order."""
    x, q = data
    ix = Index(x.shape[0], 128, 16, 100, 100, 40)
    ix.add_points(x, nthreads=1)
    r = oracle.RefIndex(x.shape[0], 128, M=16, efc=100, K=40)
    r.add_points(x, nthreads=1)
    ids, d, cnt = ix.get_knng()
    rid, rd, rc = r.get_knng()
    assert np.array_equal(cnt, rc)
    m = np.arange(40) < cnt[:, None]
    assert np.array_equal(np.where(m, d, 0), np.where(m, rd, 0))   # same distance profile
    assert float((np.where(m, ids, 0) == np.where(m, rid, 0)).mean()) > 0.98
    exact, _ = oracle.knn(x, q, 10)
    po, pr = ix.proxies(q, 10, 64), r.proxies(q, 10, 64)
    rec = lambda p: np.mean([len(set(p[i]) & set(exact[i])) / 10 for i in range(len(q))])
    assert abs(rec(po) - rec(pr)) <= 0.02


def test_multithreaded_build_labels(data):
    x, _ = data
    ix = Index(x.shape[0], 128, 16, 64, 100, 24)
    ix.add_points(x, nthreads=8)
    lab = ix.labels()
    assert sorted(lab.tolist()) == list(range(x.shape[0]))
    ids, d, cnt = ix.get_knng()
    exact, _ = oracle.knn(x, x, 25)
    # map internal ids to labels before comparing with label-indexed truth
    inv = np.empty_like(lab)
    inv[lab] = np.arange(len(lab))
    lab_ids = lab[ids]
    hits = [len(set(lab_ids[inv[i], :cnt[inv[i]]].tolist()) & set(exact[i, 1:].tolist()))
            for i in range(0, x.shape[0], 5)]
    assert np.mean(hits) / 24 > 0.3   # seeds see only earlier-inserted points


@needs_ref
def test_reference_file_loads_and_round_trips_byte_identical(tmp_path, data):
    x, _ = data
    r = oracle.RefIndex(x.shape[0], 128, M=12, efc=64, K=32)
    r.add_points(x, nthreads=1)
    r.refine(2, 8, nthreads=1)
    r.build_rknng()
    p_ref = str(tmp_path / "ref.idx")
    r.save(p_ref)
    ix = Index.load(p_ref, 128)
    ids, d, cnt = ix.get_knng()
    rid, rd, rc = r.get_knng()
    m = np.arange(32) < cnt[:, None]
    assert np.array_equal(cnt, rc)
    assert np.array_equal(np.where(m, ids, 0), np.where(m, rid, 0))
    off, ent, rb, wide = ix.get_rknng()
    roff, rent, rrb, rwide = r.get_csr()
    assert np.array_equal(off, roff) and np.array_equal(ent, rent) and rb == rrb
    p_ours = str(tmp_path / "ours.idx")
    ix.save(p_ours)
    assert open(p_ref, "rb").read() == open(p_ours, "rb").read()


@needs_ref
def test_our_file_loads_in_reference(tmp_path, data):
    x, q = data
    ix = Index(x.shape[0], 128, 16, 64, 100, 24)
    ix.add_points(x, nthreads=1)
    p = str(tmp_path / "ours.idx")
    ix.save(p)
    r = oracle.RefIndex.load(p, 128)
    r.K = 24
    assert r.count == x.shape[0]
    rid, rd, rc = r.get_knng()
    ids, d, cnt = ix.get_knng()
    m = np.arange(24) < cnt[:, None]
    assert np.array_equal(rc, cnt) and np.array_equal(np.where(m, rid, 0), np.where(m, ids, 0))
    links, levels = ix.level0()
    rlinks, rcnt = r.level0(32)
    assert np.array_equal(rcnt, links[:, 0])
    assert np.array_equal(np.where(rlinks == 0xFFFFFFFF, 0, rlinks),
                          np.where(links[:, 1:] == 0xFFFFFFFF, 0, links[:, 1:]))
    # the reference searches our graph and finds the same proxies (up to distance ties)
    rp, op = r.proxies(q, 5, 32), ix.proxies(q, 5, 32)
    same = np.mean([len(set(rp[i]) & set(op[i])) / 5 for i in range(len(q))])
    assert same >= 0.97


def test_loader_rejects_bad_files(tmp_path, data):
    x, _ = data
    ix = Index(600, 128, 8, 32, 100, 16)
    ix.add_points(x[:600], nthreads=1)
    p = str(tmp_path / "a.idx")
    ix.save(p)
    blob = open(p, "rb").read()
    bad = str(tmp_path / "bad.idx")
    open(bad, "wb").write(b"XXXX" + blob[4:])
    with pytest.raises(hrnn_amd.HrnnError, match="bad magic"):
        Index.load(bad, 128)
    open(bad, "wb").write(blob[: len(blob) // 2])
    with pytest.raises(hrnn_amd.HrnnError, match="truncated"):
        Index.load(bad, 128)
    with pytest.raises(hrnn_amd.HrnnError, match="layout"):
        Index.load(p, 64)   # wrong dimension
    with pytest.raises(hrnn_amd.HrnnError):
        Index.load(str(tmp_path / "missing.idx"), 128)


def test_capacity_and_preconditions(data):
    x, _ = data
    ix = Index(100, 128, 8, 32, 100, 16)
    ix.add_points(x[:100])
    with pytest.raises(hrnn_amd.HrnnError, match="exceeds"):
        ix.add_points(x[100:101])
    with pytest.raises(hrnn_amd.HrnnError, match="KNNG not built"):
        ix.compact_for_serving(8)
    with pytest.raises(hrnn_amd.HrnnError, match="buildRKNNG_mutable"):
        ix.insert_dynamic(x[:1], np.array([5000], np.uint64))
    with pytest.raises(hrnn_amd.HrnnError, match="RKNNG not built"):
        ix.search(x[:2], 5, 1.0, 5, 16, 32)
    with pytest.raises(hrnn_amd.HrnnError, match="already present"):
        ix2 = Index(10, 128, 8, 32, 100, 0)
        ix2.add_points(x[:2], labels=np.array([7, 7], np.uint64))


def test_host_query_features_consistent(data):
    """Host phase A with beam statistics (the learned-m inputs): the beam's
    top-16 distances are the exact distances of the first 16 proxies, the
    proxies are the phase-A proxies of proxies(), and every feature follows
    from them and the KNN lists."""
    x, q = data
    ix = Index(x.shape[0], 128, 16, 100, 100, 40)
    ix.add_points(x, nthreads=1)
    f, p, raw = ix.query_features(q, 10, 64, 20, nthreads=1, with_raw=True)
    assert f.shape == (len(q), 32) and np.isfinite(f).all()
    assert np.array_equal(p, ix.proxies(q, 20, 64, 1))
    lab = ix.labels()
    xi = x[lab]
    ids, kd, cnt = ix.get_knng()
    for i in range(len(q)):
        dd = np.array([oracle.l2sqr(q[i], xi[j]) for j in p[i, :16]], np.float32)
        assert np.array_equal(raw[i, 2:18], dd)                   # closest first
        assert raw[i, 19] == 64 and raw[i, 18] >= 1               # beam size, expansions
        assert raw[i, 1] <= raw[i, 0]                             # descent never moves away
        rk = np.array([kd[j, 9] if cnt[j] >= 10 else 3e38 for j in p[i, :16]], np.float32)
        assert f[i, 29] == pytest.approx(np.mean(dd <= rk))
        assert f[i, 18] == pytest.approx(np.log1p(rk[0]))
        assert f[i, 31] == pytest.approx(np.log1p(raw[i, 18]))
    with pytest.raises(hrnn_amd.HrnnError):
        ix.query_features(q, 10, 64, 8)                           # needs >= 16 proxies


@needs_ref
def test_phase_a_proxies_equal_reference_on_same_file(tmp_path):
    """Phase A (HNSW descent + searchBaseLayerST<true> beam, hrnn.h:1992-2019)
    on one index file: our proxies and the reference's are the same sets,
    distance ties included (integer data make ties frequent)."""
    x = mixture(3000, seed=61, centre_seed=1061)
    q = mixture(80, seed=62, centre_seed=1061)
    ix = Index(3000, 128, 16, 100, 100, 48)
    ix.add_points(x, nthreads=1)
    p = str(tmp_path / "a.idx")
    ix.save(p)
    r = oracle.RefIndex.load(p, 128)
    for m, ef in [(1, 10), (10, 64), (20, 200), (50, 200), (64, 64)]:
        a, b = ix.proxies(q, m, ef, 1), r.proxies(q, m, ef)
        assert np.array_equal(np.sort(a, 1), np.sort(b, 1)), (m, ef)


@needs_ref
def test_saved_index_search_matches_reference(tmp_path):
    """An index built here (host HNSW, KNN lists refined by the oracle's
    NNDescent and set through set_knng) is saved, loaded by the reference,
    and searched by it: its answers equal phases B+C of the oracle on our
    host phase-A proxies for every query (empty answers included)."""
    x = mixture(3000, seed=61, centre_seed=1061)
    q = mixture(80, seed=62, centre_seed=1061)
    ix = Index(3000, 128, 16, 100, 100, 48)
    ix.add_points(x, nthreads=1)
    sid, sd, sc = ix.get_knng()
    oid, od, oc, _, _ = oracle.nndescent(x, sid, sd, sc, 48, 3, 12)
    ix.set_knng(oid, od, oc)
    ix.build_rknng_mutable()
    p = str(tmp_path / "a.idx")
    ix.save(p)
    r = oracle.RefIndex.load(p, 128)
    rid, rd, rc = r.get_knng()
    assert np.array_equal(rc, oc)
    m = np.arange(48) < oc[:, None]
    assert np.array_equal(np.where(m, rid, 0), np.where(m, oid, 0))
    r.build_rknng()
    roff, rlab = r.search(q, 10, 1.0, 10, 48, 64)
    prox = ix.proxies(q, 10, 64, 1)
    coff, cent, rb, wide = oracle.csr(oid, oc, 48)
    eoff, eids = oracle.search_bc(coff, cent, wide, rb, od, oc, 48, x, q, prox, 10, 48)
    for i in range(len(q)):
        assert set(eids[eoff[i]:eoff[i + 1]].tolist()) == set(rlab[roff[i]:roff[i + 1]].tolist()), i


def test_parallel_build_seeds_every_node():
    """Concurrent addPoint (8 threads): a node reached through an upper layer
    while it is still being inserted is waited for (the reference holds the
    node's lock for its whole insertion), so no insertion starts its layer-0
    beam from an unlinked node.  Seed lists are then full for all but the
    first few nodes, as in a single-threaded build."""
    x = hrnn_amd.synth_mixture(1, 10_000, 128, 50, 20.0, 0.0, 255.0, True, 0.0, 1001)
    for nt in (1, 8):
        ix = Index(10_000, 128, 16, 200, 100, 100)
        ix.add_points(x, nthreads=nt)
        _, _, sc = ix.get_knng()
        late = np.arange(10_000) >= 200          # past the first few hundred insertions
        assert int((sc[late] < 50).sum()) <= 2, (nt, np.bincount(np.minimum(sc, 100))[:10])


def _reverse_invariant(ids, cnt, off, rids):
    """Every KNN entry j of node i has i in j's mutable reverse list."""
    n = len(cnt)
    rsets = [set(rids[off[j]:off[j + 1]].tolist()) for j in range(n)]
    missing = 0
    for i in range(n):
        for j in ids[i, :cnt[i]].tolist():
            missing += i not in rsets[j]
    return missing


def _insert_setup(tmp_path, n_base, x, K):
    """A saved index of the first n_base rows whose lists are the oracle's
    NNDescent result, so both implementations start from identical bytes."""
    ix = Index(n_base, 128, 16, 64, 100, K)
    ix.add_points(x[:n_base], nthreads=1)
    sid, sd, sc = ix.get_knng()
    oid, od, oc, _, _ = oracle.nndescent(x[:n_base], sid, sd, sc, K, 3, 10)
    ix.set_knng(oid, od, oc)
    p = str(tmp_path / "base.idx")
    ix.save(p)
    return p


@needs_ref
@pytest.mark.parametrize("bidir", [False, True])
def test_insert_dynamic_matches_reference_single_thread(tmp_path, data, bidir):
    """insertPointDynamic (reference hrnn.h:2360-2534, try_update_knng_evict
    :1592-1655) and the mutable reverse lists (:1942-1959): from the same
    index file, one thread inserting the same rows in the same order, both
    sides end with the same ranked lists (counts and distances; ids up to the
    order of equal distances) and the same reverse lists as multisets."""
    x, _ = data
    n_base, n_all, K = 1200, 1800, 24
    p = _insert_setup(tmp_path, n_base, x, K)
    ours = Index.load(p, 128, n_all)
    ref = oracle.RefIndex.load(p, 128, n_all)
    ref.K = K
    ours.build_rknng_mutable()
    ref.build_rknng_mutable()
    lab = np.arange(n_base, n_all, dtype=np.uint64)
    ours.insert_dynamic(x[n_base:n_all], lab, 0, bidir, nthreads=1)
    ref.insert_dynamic(x[n_base:n_all], lab, 0, bidir, nthreads=1)
    assert ours.count == ref.count == n_all
    ids, d, cnt = ours.get_knng()
    rid, rd, rc = ref.get_knng()
    assert np.array_equal(ours.labels(), ref.labels())
    assert np.array_equal(cnt, rc)
    m = np.arange(K) < cnt[:, None]
    assert np.array_equal(np.where(m, d, 0), np.where(m, rd, 0))
    assert float((np.where(m, ids, 0) == np.where(m, rid, 0)).mean()) > 0.99
    off, rl = ours.get_rknng_mutable()
    roff, rrl = ref.get_rknng_mutable()
    assert np.array_equal(np.diff(off), np.diff(roff))
    same = np.mean([np.array_equal(np.sort(rl[off[i]:off[i + 1]]), np.sort(rrl[roff[i]:roff[i + 1]]))
                    for i in range(n_all)])
    assert same > 0.99, same
    assert _reverse_invariant(ids, cnt, off, rl) == 0


def test_insert_dynamic_bidirectional_multithreaded_keeps_reverse_lists(tmp_path, data):
    """Eight threads inserting with bidirectional=True: afterwards every KNN
    entry j of every node i has i in j's reverse list, and no list holds an
    id twice (the list rewrite of the inserted node marks and clears its
    dedup buffer from the list as it is under the node's lock)."""
    x, _ = data
    n_base, n_all, K = 1000, 2500, 24
    p = _insert_setup(tmp_path, n_base, x, K)
    ix = Index.load(p, 128, n_all)
    ix.build_rknng_mutable()
    lab = np.arange(n_base, n_all, dtype=np.uint64)
    ix.insert_dynamic(x[n_base:n_all], lab, 0, True, nthreads=8)
    ids, d, cnt = ix.get_knng()
    for i in range(n_all):
        row = ids[i, :cnt[i]]
        assert len(set(row.tolist())) == len(row), i
        assert np.all(np.diff(d[i, :cnt[i]]) >= 0), i
    off, rl = ix.get_rknng_mutable()
    assert _reverse_invariant(ids, cnt, off, rl) == 0


@pytest.mark.parametrize("m_update,bidir", [(0, False), (4, False), (0, True)])
def test_host_insert_dynamic_steps_equal_oracle(m_update, bidir):
    """HRNN::insertPointDynamic on the host (include/hnswlib/hrnn.h), one
    point at a time on one thread, against the oracle's phases 2-5 (pinned to
    the compiled reference's insertPointDynamic in test_oracle_insert.py):
    after each call every list equals the oracle's step from the state
    before the call plus the new point's seed list, and the mutable reverse
    lists hold exactly the transpose of the lists (with bidirectional, a
    superset: entries of nodes the point's list dropped stay, as the
    reference leaves them).  Float data, so no equal distances."""
    n_old, nx, K = 900, 15, 32
    rng = np.random.default_rng(17 + m_update)
    x = rng.standard_normal((n_old + nx, 64)).astype(np.float32)
    ix = Index(n_old + nx, 64, 16, 64, 100, K)
    ix.add_points(x[:n_old], nthreads=1)
    exact_ids, exact_d = oracle.knn(x[:n_old], x[:n_old], K + 1)
    ix.set_knng(exact_ids[:, 1:].astype(np.uint32), exact_d[:, 1:].astype(np.float32),
                np.full(n_old, K, np.uint16))
    ix.build_rknng_mutable()
    prev = ix.get_knng()
    upd = 0
    for t in range(nx):
        xn = n_old + t
        ix.insert_dynamic(x[xn:xn + 1], np.array([xn], np.uint64), m_update, bidir)
        after = ix.get_knng()
        s = [np.zeros((xn + 1, K), np.uint32), np.zeros((xn + 1, K), np.float32), np.zeros(xn + 1, np.uint16)]
        for a, b in zip(s, prev):
            a[:xn] = b
        if bidir:
            # the seeds: the step's input row is not observable after a
            # bidirectional call; rebuild it as the oracle would see it from
            # the list the call started from (the first step only)
            if t > 0:
                break
            seeds = Index(n_old + nx, 64, 16, 64, 100, K)
            seeds.add_points(x[:n_old], nthreads=1)
            seeds.set_knng(exact_ids[:, 1:].astype(np.uint32), exact_d[:, 1:].astype(np.float32),
                           np.full(n_old, K, np.uint16))
            seeds.build_rknng_mutable()
            seeds.insert_dynamic(x[xn:xn + 1], np.array([xn], np.uint64), m_update, False)
            row = seeds.get_knng()
        else:
            row = after
        s[0][xn], s[1][xn], s[2][xn] = row[0][xn], row[1][xn], row[2][xn]
        want = oracle.insert_affect(x[:xn + 1], *s, K, xn, xn + 1, m_update, bidir)
        ids_a, d_a, c_a = after
        assert np.array_equal(c_a, want[2]), t
        live = np.arange(K)[None, :] < c_a[:, None]
        assert np.array_equal(np.where(live, ids_a, 0), np.where(live, want[0], 0)), t
        assert np.array_equal(np.where(live, d_a, 0), np.where(live, want[1], 0)), t
        upd += int(want[3][3])
        # mutable reverse lists against the transpose of the lists
        off, rids = ix.get_rknng_mutable()
        for v in range(xn + 1):
            got = sorted(rids[off[v]:off[v + 1]].tolist())
            exp = sorted(np.nonzero((ids_a == v) & live)[0].tolist())
            if bidir:
                assert set(exp) <= set(got), v
            else:
                assert got == exp, v
        prev = after
    assert upd > 0


def test_host_insert_dynamic_bidirectional_threads_keep_reverse_entries():
    """Bidirectional insertions from 8 threads at once: every list entry
    (v in u's list) has its reverse entry (u in v's mutable reverse list)
    afterwards, and each list is sorted, duplicate-free and without self."""
    n_old, nx, K = 1500, 400, 24
    rng = np.random.default_rng(29)
    x = rng.standard_normal((n_old + nx, 32)).astype(np.float32)
    ix = Index(n_old + nx, 32, 16, 64, 100, K)
    ix.add_points(x[:n_old], nthreads=8)
    exact_ids, exact_d = oracle.knn(x[:n_old], x[:n_old], K + 1)
    lab = ix.labels()
    inv = np.empty_like(lab)
    inv[lab] = np.arange(n_old)
    ix.set_knng(inv[exact_ids[lab, 1:]].astype(np.uint32), exact_d[lab, 1:].astype(np.float32),
                np.full(n_old, K, np.uint16))
    ix.build_rknng_mutable()
    ix.insert_dynamic(x[n_old:], np.arange(n_old, n_old + nx, dtype=np.uint64), 0, True, nthreads=8)
    ids, d, cnt = ix.get_knng()
    off, rids = ix.get_rknng_mutable()
    n = n_old + nx
    rev = [set(rids[off[v]:off[v + 1]].tolist()) for v in range(n)]
    for u in range(n):
        row = ids[u, :cnt[u]]
        assert len(set(row.tolist())) == len(row) and u not in row, u
        assert np.all(np.diff(d[u, :cnt[u]]) >= 0), u
        for v in row.tolist():
            assert u in rev[v], (u, v)
