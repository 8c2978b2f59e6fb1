"""Pin the CPU oracle (oracle/hrnn_oracle.c) before it is trusted as the checker.

Each oracle function is checked against the compiled reference (oracle/_ref,
the reference's own hrnn.h), against committed golden fixtures written by that
reference (tests/golden/), or against what the mathematics fixes (exact
integer arithmetic, brute force on tiny inputs, invariants).  Every check is
chosen so that a plausible slip (a dropped term, a wrong index, a swapped
operand, a wrong tie rule) fails it.
"""
import os

import numpy as np
import pytest

import oracle
from conftest import mixture

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
needs_ref = pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built")


# ---- L2 ---------------------------------------------------------------------

@needs_ref
@pytest.mark.parametrize("d", [96, 128, 960])
def test_l2_bit_exact_vs_reference(d):
    """space_l2.h:25-54: the restated 16-lane FMA order equals the compiled one bit for bit."""
    rng = np.random.default_rng(d)
    r = oracle.RefIndex(4, d, K=0)
    a = rng.standard_normal((400, d)).astype(np.float32) * 3
    b = rng.standard_normal((400, d)).astype(np.float32)
    ref = np.array([r.l2(a[i], b[i]) for i in range(400)], np.float32)
    got = np.array([oracle.l2sqr(a[i], b[i]) for i in range(400)], np.float32)
    assert np.array_equal(ref, got)
    # a plausible slip (no FMA contraction) must be caught by this check
    slip = np.array([oracle.l2sqr(a[i], b[i], fma=0) for i in range(400)], np.float32)
    assert not np.array_equal(ref, slip)


@pytest.mark.skipif(not oracle.ref_available("avx2"), reason="avx2 build absent")
def test_l2_avx_variant_bit_exact():
    """space_l2.h:60-91: 8 partials on AVX-only hosts."""
    rng = np.random.default_rng(7)
    r = oracle.RefIndex(4, 128, K=0, variant="avx2")
    a = rng.standard_normal((300, 128)).astype(np.float32)
    b = rng.standard_normal((300, 128)).astype(np.float32)
    ref = np.array([r.l2(a[i], b[i]) for i in range(300)], np.float32)
    got = np.array([oracle.l2sqr(a[i], b[i], W=8) for i in range(300)], np.float32)
    assert np.array_equal(ref, got)


def test_l2_integer_closed_form():
    """Integer data 0..255 at d=128: every partial sum is an exact integer < 2^24."""
    x = mixture(64, seed=3)
    for i in range(0, 64, 7):
        for j in range(0, 64, 5):
            exact = int(((x[i].astype(np.int64) - x[j].astype(np.int64)) ** 2).sum())
            assert oracle.l2sqr(x[i], x[j]) == float(exact)


# ---- brute force kNN / r_k / RkNN ---------------------------------------------

def _np_knn(x, q, k):
    d = ((q[:, None, :].astype(np.float64) - x[None, :, :].astype(np.float64)) ** 2).sum(-1)
    ids = np.lexsort((np.broadcast_to(np.arange(x.shape[0]), d.shape), d), axis=1)[:, :k]
    return ids, np.take_along_axis(d, ids, 1)


def test_knn_ties_to_smaller_id():
    """bruteforce.h:106-135: k smallest by (dist, id); duplicates force exact ties."""
    x = mixture(300, d=32, seed=21)
    x[100:110] = x[5]          # exact duplicates -> many equal distances
    q = np.concatenate([x[:10], mixture(10, d=32, seed=22)])
    ids, d2 = oracle.knn(x, q, 12)
    eids, ed = _np_knn(x, q, 12)
    assert np.array_equal(ids, eids.astype(np.uint32))
    assert np.array_equal(d2, ed.astype(np.float32))


def test_kth_includes_self_and_is_k_plus_1th():
    """gen_rknn_gt.py:62-75: topk(k+1) over all points including self, take index k."""
    x = mixture(400, d=32, seed=23)
    k = 5
    kth = oracle.kth(x, k)
    _, ed = _np_knn(x, x, k + 1)
    assert np.array_equal(kth, ed[:, k].astype(np.float32))
    assert not np.array_equal(kth, ed[:, k - 1].astype(np.float32))   # off-by-one slip


def test_kth_f64_matches_on_integers():
    x = mixture(200, d=32, seed=24)
    assert np.array_equal(oracle.kth_f64(x, 4).astype(np.float32), oracle.kth(x, 4))


def test_kth_f64_float_data_vs_float64_brute_force():
    """The FP64 judge on float data (the precision form's generator: |N(0,1)|
    scaled per dimension, d = 960): r_k is the (k+1)-th smallest float64
    direct-difference distance, self included.  Against a float64 numpy brute
    force (pairwise summation: differs only in the last bits) at 1e-12
    relative; the FP32 oracle misses that bar on most rows, so a slip to
    FP32 accumulation, or to the k-th instead of the (k+1)-th distance,
    fails."""
    import hrnn_amd
    x = hrnn_amd.synth_halfnormal_scaled(5, 55, 300, 960)
    k = 10
    got = oracle.kth_f64(x, k)
    xd = x.astype(np.float64)
    exp = np.empty(300)
    exp_k = np.empty(300)
    for b in range(0, 300, 20):
        dd = ((xd[b:b + 20, None, :] - xd[None, :, :]) ** 2).sum(-1)
        part = np.sort(dd, axis=1)
        exp[b:b + 20] = part[:, k]
        exp_k[b:b + 20] = part[:, k - 1]
    assert np.allclose(got, exp, rtol=1e-12, atol=0)
    assert not np.allclose(exp_k, exp, rtol=1e-12, atol=0)
    f32 = oracle.kth(x, k).astype(np.float64)
    assert np.mean(np.abs(f32 - exp) > 1e-12 * exp) > 0.5
    rows = np.array([0, 7, 299, 150], np.uint32)
    assert np.array_equal(oracle.kth_f64_rows(x, k, rows), got[rows])


def test_l2sqr_f64_is_sequential_without_fma():
    """The FP64 distance: float64 differences, squares and sums rounded one
    by one, left to right.  Pinned on a case where the order matters: a huge
    term first absorbs the small ones (a pairwise or reversed order keeps
    them), and on a random pair against the same loop written out here."""
    a = np.zeros(4, np.float32)
    b = np.array([2.0 ** 27, 1.0, 1.0, 1.0], np.float32)     # squares 2^54, 1, 1, 1
    assert oracle.l2sqr_f64(a, b) == 2.0 ** 54                 # sequential: each 1 is lost
    assert oracle.l2sqr_f64(a, b[::-1].copy()) == 2.0 ** 54 + 4.0   # 1+1+1 = 3 first, then 2^54 (ties to even)
    rng = np.random.default_rng(3)
    u, v = rng.random(960, np.float32), rng.random(960, np.float32)
    r = 0.0
    for i in range(960):
        t = float(u[i]) - float(v[i])
        r += t * t
    assert oracle.l2sqr_f64(u, v) == r


def test_rknn_gt_f64_pins():
    """The FP64 judge of compute_rknn: {x : d64(q, x) <= kth64[x]}.
    * integer data: equal to the FP32 groundtruth (every distance exact);
    * the precision form's float data at d = 960, with queries placed within
      FP32 rounding of a radius (where FP32 and FP64 decide differently):
      equal to a float64 numpy brute force; the FP32 groundtruth differs on
      exactly the planted pairs, so a slip to FP32 distances or radii fails."""
    from boundary_cases import boundary_queries
    import hrnn_amd
    x = mixture(600, d=64, seed=31)
    q = mixture(40, d=64, seed=32)
    k = 5
    o32, i32, d32 = oracle.rknn_gt(x, q, oracle.kth(x, k))
    o64, i64, d64 = oracle.rknn_gt_f64(x, q, oracle.kth_f64(x, k))
    assert np.array_equal(o32, o64) and np.array_equal(i32, i64)
    assert np.array_equal(d32.astype(np.float64), d64)

    x = hrnn_amd.synth_halfnormal_scaled(5, 55, 1200, 960)
    k = 10
    k32, k64 = oracle.kth(x, k), oracle.kth_f64(x, k)
    qb, pb, inb = boundary_queries(x, k, k32, k64, range(0, 1200, 37), want=6)
    assert len(qb) >= 4, len(qb)
    qr = hrnn_amd.synth_halfnormal_scaled(6, 55, 20, 960)
    q = np.concatenate([qb, qr])
    off, ids, dd = oracle.rknn_gt_f64(x, q, k64)
    xd = x.astype(np.float64)
    for i in range(q.shape[0]):
        e = ((xd - q[i].astype(np.float64)) ** 2).sum(1)
        exp = np.nonzero(e <= k64)[0]
        got = ids[off[i]:off[i + 1]]
        assert np.array_equal(np.sort(got), exp), i
        assert np.all(np.diff(dd[off[i]:off[i + 1]]) >= 0), i
    for i in range(len(qb)):      # the planted decision, and the FP32 groundtruth disagrees
        got = set(ids[off[i]:off[i + 1]].tolist())
        assert (int(pb[i]) in got) == bool(inb[i]), i
    f_off, f_ids, _ = oracle.rknn_gt(x, qb, k32)
    for i in range(len(qb)):
        f = set(f_ids[f_off[i]:f_off[i + 1]].tolist())
        assert (int(pb[i]) in f) != bool(inb[i]), i


def test_knng_recall_hand_built_graph():
    """orc_knng_recall: per node, the first kk exact neighbours found among
    the first min(cnt, kk) list entries, pooled over nodes.  Hand-counted:
    node 0 (cnt 1) finds 5 of exact [7, 5]; node 1 (cnt 4) finds 1 of
    [3, 1] in its first two entries [1, 2]; node 2 (cnt 0) finds nothing of
    [0, 1]: 2 hits of 6.  Counting past cnt gives 5/6 and an off-by-one in
    kk 4/9."""
    K = 4
    ids = np.array([[5, 7, 9, 1], [1, 2, 3, 4], [0, 1, 2, 3]], np.uint32)
    cnt = np.array([1, 4, 0], np.uint16)
    exact = np.array([[7, 5, 9], [3, 1, 2], [0, 1, 2]], np.uint32)
    assert oracle.knng_recall(ids, cnt, K, exact, 2) == pytest.approx(2 / 6, abs=1e-15)
    assert oracle.knng_recall(ids, cnt, K, exact, 3) == pytest.approx(4 / 9, abs=1e-15)
    full = np.full(3, 4, np.uint16)
    assert oracle.knng_recall(ids, full, K, exact, 2) == pytest.approx(5 / 6, abs=1e-15)
    # a list that is its exact neighbours, in any order, scores 1
    perm = np.array([[5, 7, 1, 9], [3, 1, 4, 2], [2, 0, 1, 3]], np.uint32)
    assert oracle.knng_recall(perm, full, K, exact, 2) < 1.0
    same = np.array([[5, 7, 9, 0], [1, 3, 2, 0], [1, 0, 2, 0]], np.uint32)
    assert oracle.knng_recall(same, np.full(3, 3, np.uint16), K, exact, 3) == pytest.approx(1.0)


def test_rknn_gt_brute_force():
    """gen_rknn_gt.py:82-151: RkNN(q) = {x : d2(q,x) <= r_k(x)}, inclusive."""
    x = mixture(500, d=32, seed=25)
    q = mixture(30, d=32, seed=26)
    q[:3] = x[:3]                       # queries on top of base points
    k = 4
    kth = oracle.kth(x, k)
    off, ids, d2 = oracle.rknn_gt(x, q, kth)
    dist = ((q[:, None, :].astype(np.int64) - x[None].astype(np.int64)) ** 2).sum(-1)
    for i in range(q.shape[0]):
        exp = np.nonzero(dist[i] <= kth)[0]
        got = ids[off[i]:off[i + 1]]
        assert sorted(got.tolist()) == exp.tolist()
        # sorted by (d2, id)
        pairs = list(zip(d2[off[i]:off[i + 1]].tolist(), got.tolist()))
        assert pairs == sorted(pairs)
    # the inclusive boundary matters: strict < would lose members on integer data
    strict = sum(int((dist[i] < kth).sum()) for i in range(q.shape[0]))
    assert strict < int(off[-1])


# ---- reverse CSR ---------------------------------------------------------------

def _random_graph(n, K, seed, full=False):
    rng = np.random.default_rng(seed)
    cnt = np.full(n, K, np.uint16) if full else rng.integers(0, K + 1, n).astype(np.uint16)
    ids = np.zeros((n, K), np.uint32)
    for i in range(n):
        ids[i] = rng.choice(n, K, replace=False)
    return ids, cnt


def test_csr_invariants_and_order():
    """Offsets = exclusive scan of in-degree; entries = transpose; (rank, src) order."""
    n, K = 700, 37
    ids, cnt = _random_graph(n, K, 3)
    off, ent, rb, wide = oracle.csr(ids, cnt, K)
    assert rb == 6 and wide == 0
    deg = np.zeros(n, np.int64)
    pairs = set()
    for i in range(n):
        for j in range(cnt[i]):
            deg[ids[i, j]] += 1
            pairs.add((int(ids[i, j]), i, j))
    assert np.array_equal(np.diff(off.astype(np.int64)), deg)
    got = set()
    for dst in range(n):
        seg = ent[off[dst]:off[dst + 1]].astype(np.int64)
        src, rank = seg >> rb, seg & ((1 << rb) - 1)
        keys = list(zip(rank.tolist(), src.tolist()))
        assert keys == sorted(keys)
        got.update((dst, s, r) for s, r in zip(src.tolist(), rank.tolist()))
    assert got == pairs


def test_csr_params_rule():
    """hrnn.h:1872-1885: rank_bits = ceil(log2 K), node_bits = ceil(log2 n), wide > 32, error > 64."""
    assert oracle.csr_params(10_000, 100)[1:] == (7, 14, 0)
    assert oracle.csr_params(1_000_000, 500)[1:] == (9, 20, 0)
    assert oracle.csr_params(10_000_000, 100)[1:] == (7, 24, 0)
    assert oracle.csr_params(1 << 25, 256)[1:] == (8, 25, 1)
    assert oracle.csr_params(1 << 24, 256)[1:] == (8, 24, 0)
    assert oracle.csr_params(1 << 24, 257)[1:] == (9, 24, 1)
    assert oracle.csr_params(1 << 40, 1 << 25)[0] == -6


def _canon(off, ent, rb):
    """Sort each equal-rank run by src (the reference's std::sort leaves it unspecified)."""
    out = ent.copy()
    mask = (1 << rb) - 1
    for dst in range(len(off) - 1):
        seg = ent[off[dst]:off[dst + 1]].astype(np.int64)
        keys = np.lexsort((seg >> rb, seg & mask))
        out[off[dst]:off[dst + 1]] = seg[keys].astype(ent.dtype)
    return out


@needs_ref
def test_csr_bit_exact_vs_reference(small_ref):
    """buildRKNNG (hrnn.h:1869-1935) on the reference's own refined graph."""
    r, _ = small_ref
    ids, dists, cnt = r.get_knng()
    roff, rent, rrb, rwide = r.get_csr()
    off, ent, rb, wide = oracle.csr(ids, cnt, r.K)
    assert (rb, wide) == (rrb, rwide)
    assert np.array_equal(off, roff)
    assert np.array_equal(ent, _canon(roff, rent, rrb))


# ---- search phases B + C ---------------------------------------------------------

@needs_ref
@pytest.mark.parametrize("m,Kp,k,thr", [(1, 48, 10, 1.0), (5, 20, 10, 1.0), (10, 48, 5, 1.0),
                                         (10, 48, 10, 2.0), (8, 48, 48, 1.0)])
def test_search_bc_vs_reference(small_data, small_ref, m, Kp, k, thr):
    """searchRknn_impl (hrnn.h:1979-2071) with the reference's own phase-A proxies."""
    x, q = small_data
    r, _ = small_ref
    ids, dists, cnt = r.get_knng()
    roff, rent, rb, wide = r.get_csr()
    prox = r.proxies(q, m, 64)
    off, res = oracle.search_bc(roff, rent, wide, rb, dists, cnt, r.K, x, q, prox, k, Kp, thr)
    eoff, elab = r.search(q, m, thr, k, Kp, 64)
    assert np.array_equal(off, eoff)
    assert np.array_equal(res.astype(np.uint64), elab)   # labels == internal ids here


def test_search_bc_brute_force_tiny():
    """Phase B/C restated from the definition on a hand-made index."""
    x = np.array([[0, 0], [1, 0], [0, 2], [5, 5], [1, 1]], np.float32)
    K = 2
    ids = np.array([[4, 1], [4, 0], [4, 0], [2, 4], [1, 0]], np.uint32)
    dists = np.array([[2, 1], [1, 1], [5, 4], [18, 32], [1, 2]], np.float32)
    dists = np.sort(dists, axis=1)
    cnt = np.array([2, 2, 2, 2, 2], np.uint16)
    off, ent, rb, wide = oracle.csr(ids, cnt, K)
    q = np.array([[0.5, 0.5]], np.float32)
    # proxies 0 and 4: reverse lists of 0 = {1 (rank1), 2 (rank1), 4 (rank1)}; of 4 = {0,1,2,3}
    o, res = oracle.search_bc(off, ent, wide, rb, dists, cnt, K, x, q, np.array([[0, 4]]), 1, 2)
    exp = []
    cands = sorted({1, 2, 4, 0, 3})
    for c in cands:
        dq = float(((q[0] - x[c]) ** 2).sum())
        if dq <= dists[c, 0]:
            exp.append(c)
    assert res.tolist() == exp
    # K' = 1 keeps only rank-0 entries: of 0 none, of 4 {0,1,2}
    o, res = oracle.search_bc(off, ent, wide, rb, dists, cnt, K, x, q, np.array([[0, 4]]), 1, 1)
    exp = [c for c in [0, 1, 2] if float(((q[0] - x[c]) ** 2).sum()) <= dists[c, 0]]
    assert res.tolist() == exp


# ---- NNDescent -------------------------------------------------------------------

@needs_ref
def test_nndescent_bit_exact_vs_reference_single_thread(small_data, small_ref):
    """refineKNNG (hrnn.h:1725-1862) with one thread is deterministic: same lists, same ranks."""
    x, _ = small_data
    r, seed = small_ref
    sid, sd, sc = seed
    ids, dists, cnt, upd, nd = oracle.nndescent(x, sid, sd, sc, r.K, 3, 12)
    rid, rd, rc = r.get_knng()
    assert np.array_equal(cnt, rc)
    for i in range(x.shape[0]):
        c = int(cnt[i])
        assert np.array_equal(ids[i, :c], rid[i, :c])
        assert np.array_equal(dists[i, :c], rd[i, :c])
    assert nd[0] > 0 and upd[0] > 0


@needs_ref
def test_nndescent_non_full_lists_vs_reference():
    """K > ef_construction: lists start partly filled (the 1M x 128, K=500 regime)."""
    x = mixture(1200, d=64, seed=31)
    r = oracle.RefIndex(1200, 64, M=8, efc=24, K=40)
    r.add_points(x, nthreads=1)
    sid, sd, sc = r.get_knng()
    assert int(sc.max()) < 40
    r.refine(2, 8, nthreads=1)
    ids, dists, cnt, _, _ = oracle.nndescent(x, sid, sd, sc, 40, 2, 8)
    rid, rd, rc = r.get_knng()
    assert np.array_equal(cnt, rc)
    assert np.array_equal(np.where(np.arange(40) < cnt[:, None], ids, 0),
                          np.where(np.arange(40) < rc[:, None], rid, 0))


# ---- golden fixtures written by the compiled reference ---------------------------

def test_nndescent_flags_resume_equals_one_run(small_data, small_ref):
    """orc_nndescent_flags: one iteration, then two more from the returned
    lists and flags, equals three iterations in one call (which is pinned
    bit-exact to the compiled refineKNNG above); all-new flags are the
    refineKNNG start state."""
    x, _ = small_data
    r, seed = small_ref
    sid, sd, sc = seed
    K, S = r.K, 12
    whole = oracle.nndescent(x, sid, sd, sc, K, 3, S)
    ones = np.ones((len(x), K), np.uint8)
    a = oracle.nndescent_flags(x, sid, sd, sc, ones, K, 1, S)
    b = oracle.nndescent_flags(x, a[0], a[1], a[2], a[3], K, 2, S)
    for i in range(3):
        assert np.array_equal(whole[i], b[i])
    assert np.concatenate([a[4], b[4]]).tolist() == whole[3].tolist()
    assert np.concatenate([a[5], b[5]]).tolist() == whole[4].tolist()
    m = np.arange(K) < a[2][:, None]
    assert a[3][m].any() and not a[3][~m].any()


def test_golden_fixtures_pin_oracle():
    """Fixtures from tests/golden/make_golden.py (the reference compiled single-threaded)."""
    path = os.path.join(GOLD, "ref_small.npz")
    if not os.path.exists(path):
        pytest.skip("golden fixture absent")
    g = np.load(path)
    x, q = g["x"], g["q"]
    K = int(g["K"])
    # NNDescent from the stored seeds reproduces the stored refined lists
    ids, dists, cnt, _, _ = oracle.nndescent(x, g["seed_ids"], g["seed_d"], g["seed_cnt"], K,
                                            int(g["iters"]), int(g["sample"]))
    assert np.array_equal(cnt, g["ref_cnt"])
    m = np.arange(K) < cnt[:, None]
    assert np.array_equal(np.where(m, ids, 0), np.where(m, g["ref_ids"], 0))
    assert np.array_equal(np.where(m, dists, 0), np.where(m, g["ref_d"], 0))
    # CSR after within-rank canonicalisation
    off, ent, rb, wide = oracle.csr(g["ref_ids"], g["ref_cnt"], K)
    assert np.array_equal(off, g["csr_off"])
    assert np.array_equal(ent, _canon(g["csr_off"], g["csr_ent"], rb))
    # search with the stored proxies reproduces the stored results
    o, res = oracle.search_bc(g["csr_off"], g["csr_ent"], 0, rb, g["ref_d"], g["ref_cnt"], K, x,
                              q, g["proxies"], int(g["k"]), int(g["Kp"]))
    assert np.array_equal(o, g["res_off"])
    assert np.array_equal(res.astype(np.uint64), g["res_lab"])
    # L2 in the compiled order
    got = np.array([oracle.l2sqr(a, b) for a, b in zip(g["l2_a"], g["l2_b"])], np.float32)
    assert np.array_equal(got, g["l2_ref"])


# ---- recall metric ---------------------------------------------------------------

def test_rknn_recall_macro_skips_empty():
    """utils.h:135-157: queries with empty GT are skipped."""
    gt_off = np.array([0, 2, 2, 5], np.uint64)
    gt_ids = np.array([1, 2, 3, 4, 5], np.uint32)
    r_off = np.array([0, 1, 3, 5], np.uint64)
    r_ids = np.array([1, 9, 9, 3, 4], np.uint32)
    mac, mic = oracle.rknn_recall(gt_off, gt_ids, r_off, r_ids, 10)
    assert mac == pytest.approx((0.5 + 2 / 3) / 2)
    assert mic == pytest.approx(3 / 5)
