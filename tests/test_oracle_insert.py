"""Pins oracle.insert_affect, the restatement of phases 2-5 of the
reference's insertPointDynamic (include/hnswlib/hrnn.h:2366-2532, with
try_update_knng_evict :1592-1655), against the compiled reference.

The reference inserts one point at a time, single-threaded.  After its call
for point x the new row holds x's seed list (phase 1 leaves it; phases 2-5
change only other rows when bidirectional is off), so the state before the
call plus that row is the input of one step, and the reference's graph after
the call is the expected output -- every row, ids, distances and counts.
Integer data makes equal distances common, so the tie rules (strict < against
a full list's last entry, a new entry after listed equal ones) are exercised.
With bidirectional on, later calls read reverse entries the reference leaves
stale (it never removes x from the lists of nodes x's list drops), so only
the first call is compared, on float data (no equal distances: the
reference's heap keeps an arbitrary one of equal keys).

The batch form (several new rows at once) has no reference counterpart; it is
pinned by its definition: the same offers applied one at a time through the
reference's evict insertion, in (distance, id) order, give the same lists."""
import numpy as np
import pytest

import hrnn_amd
import oracle

pytestmark = pytest.mark.skipif(not oracle.ref_available(), reason="compiled reference (oracle/_ref) not available")


def _ref_index(x, n_old, n_max, K):
    r = oracle.RefIndex(n_max, x.shape[1], M=16, efc=64, K=K)
    r.add_points(x[:n_old], nthreads=1)
    r.refine(2, 12, nthreads=1)
    r.build_rknng_mutable()
    return r


def _pad(state, n, K):
    ids, d, c = state
    m = ids.shape[0]
    out = (np.zeros((n, K), np.uint32), np.zeros((n, K), np.float32), np.zeros(n, np.uint16))
    out[0][:m], out[1][:m], out[2][:m] = ids, d, c
    return out


def _same(a, b):
    ids_a, d_a, c_a = a
    ids_b, d_b, c_b = b
    assert np.array_equal(c_a, c_b)
    K = ids_a.shape[1]
    live = np.arange(K)[None, :] < c_a[:, None]
    assert np.array_equal(np.where(live, ids_a, 0), np.where(live, ids_b, 0))
    assert np.array_equal(np.where(live, d_a, 0), np.where(live, d_b, 0))


@pytest.mark.parametrize("m_update", [0, 3])
def test_insert_steps_equal_reference(m_update):
    n_old, nx, K = 1200, 25, 40
    x = hrnn_amd.synth_mixture(21, n_old + nx, 128, 30, 20.0, 0.0, 255.0, True, 0.0, 1021)
    r = _ref_index(x, n_old, n_old + nx, K)
    prev = r.get_knng()
    upd = 0
    for t in range(nx):
        xn = n_old + t
        r.insert_dynamic(x[xn:xn + 1], np.array([xn], np.uint64), m_update, False)
        after = r.get_knng()
        ids, d, c = _pad(prev, xn + 1, K)
        ids[xn], d[xn], c[xn] = after[0][xn], after[1][xn], after[2][xn]   # x's seeds
        got = oracle.insert_affect(x[:xn + 1], ids, d, c, K, xn, xn + 1, m_update)
        _same(got[:3], after)
        upd += int(got[3][3])
        prev = after
    assert upd > 0   # the steps changed lists (not a vacuous comparison)


def test_insert_first_step_bidirectional_equals_reference():
    n_old, K = 1000, 32
    rng = np.random.default_rng(7)
    x = rng.standard_normal((n_old + 1, 64)).astype(np.float32)
    a = _ref_index(x, n_old, n_old + 1, K)
    b = _ref_index(x, n_old, n_old + 1, K)
    before = a.get_knng()
    a.insert_dynamic(x[n_old:], np.array([n_old], np.uint64), 0, False)   # seeds of the new row
    b.insert_dynamic(x[n_old:], np.array([n_old], np.uint64), 0, True)
    seeds, want = a.get_knng(), b.get_knng()
    ids, d, c = _pad(before, n_old + 1, K)
    ids[n_old], d[n_old], c[n_old] = seeds[0][n_old], seeds[1][n_old], seeds[2][n_old]
    got = oracle.insert_affect(x, ids, d, c, K, n_old, n_old + 1, 0, bidir=True)
    _same(got[:3], want)
    assert not np.array_equal(seeds[0][n_old], want[0][n_old])   # the point's list did improve


def _evict_insert(ids, d, c, K, tgt, dd, w):
    """try_update_knng_evict (reference hrnn.h:1592-1655) on row tgt."""
    n = int(c[tgt])
    if w in ids[tgt, :n]:
        return
    if n < K:
        pos = n
        c[tgt] = n + 1
    else:
        if not dd < d[tgt, n - 1]:
            return
        pos = n - 1
    while pos > 0 and dd < d[tgt, pos - 1]:
        ids[tgt, pos], d[tgt, pos] = ids[tgt, pos - 1], d[tgt, pos - 1]
        pos -= 1
    ids[tgt, pos], d[tgt, pos] = w, dd


def test_batch_merge_equals_one_offer_at_a_time():
    """Several new rows at once: the offers are those of each row against the
    lists as given; applying them one by one through the evict insertion in
    (distance, id) order gives the oracle's merged lists."""
    n_old, nb, K = 1500, 40, 40
    x = hrnn_amd.synth_mixture(23, n_old + nb, 128, 30, 20.0, 0.0, 255.0, True, 0.0, 1023)
    r = _ref_index(x, n_old, n_old + nb, K)
    r.insert_dynamic(x[n_old:], np.arange(n_old, n_old + nb, dtype=np.uint64), 0, False)
    ids, d, c = r.get_knng()
    # every new row back to a seed-like list: its first 24 entries
    c = c.copy()
    c[n_old:] = np.minimum(c[n_old:], 24)
    got_ids, got_d, got_c, st = oracle.insert_affect(x, ids, d, c, K, n_old, n_old + nb, 0)
    # offers, derived independently: candidates from the transpose of the given lists
    n = n_old + nb
    rev = [[] for _ in range(n)]
    for i in range(n):
        for j in range(int(c[i])):
            rev[int(ids[i, j])].append(i)
    offers = []
    for xn in range(n_old, n):
        own = set(int(v) for v in ids[xn, :c[xn]]) | {xn}
        cands = set()
        for y in ids[xn, :c[xn]]:
            cands.update(v for v in rev[int(y)] if v not in own)
        for cc in cands:
            dd = oracle.l2sqr(x[xn], x[cc])
            if c[cc] >= K and not dd < d[cc, c[cc] - 1]:
                continue
            offers.append((cc, np.float32(dd), xn))
    assert len(offers) == int(st[2]) and len(offers) > 0
    e_ids, e_d, e_c = ids.copy(), d.copy(), c.copy()
    for tgt, dd, w in sorted(offers, key=lambda o: (o[0], float(o[1]), o[2])):
        _evict_insert(e_ids, e_d, e_c, K, tgt, dd, w)
    _same((got_ids, got_d, got_c), (e_ids, e_d, e_c))
