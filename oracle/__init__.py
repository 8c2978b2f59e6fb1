"""CPU oracle for the HRNN hot path -- TEST INFRASTRUCTURE ONLY.

Two libraries are wrapped here:

* ``lib/liborc.so``: the plain-C restatement of the reference algorithms
  (``hrnn_oracle.c``; each function cites the reference file:line it follows).
* ``_ref/libhrnn_ref_{avx512,avx2}.so``: the reference's own
  ``include/hnswlib/hrnn.h`` compiled from its sources through
  ``ref_harness.cpp`` (see ``Makefile``).  It pins the restatement and is the
  CPU baseline of ``bench.py``.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package; the product (``hrnn_amd``) never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORC_PATH = os.path.join(HERE, "lib", "liborc.so")

_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u16p = np.ctypeslib.ndpointer(np.uint16, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")

# Summation order of the reference's L2Sqr on this build's CPU path:
# 16 lanes (AVX-512, space_l2.h:25-54) with FMA contraction.  Tests check the
# pair (W, FMA) bit-exactly against the compiled reference.
L2_W = 16
L2_FMA = 1


def build(quiet: bool = True) -> None:
    """Compile liborc.so (and the reference harness when its sources exist)."""
    out = subprocess.run(["make", "-C", HERE, "all"], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)
    if not quiet:
        print(out.stdout)


_orc = None


def orc() -> C.CDLL:
    global _orc
    if _orc is None:
        if not os.path.exists(ORC_PATH):
            build()
        lib = C.CDLL(ORC_PATH)
        lib.orc_l2sqr.restype = C.c_float
        lib.orc_l2sqr.argtypes = [_f32p, _f32p, C.c_uint32, C.c_int, C.c_int]
        lib.orc_l2sqr_f64.restype = C.c_double
        lib.orc_l2sqr_f64.argtypes = [_f32p, _f32p, C.c_uint32]
        lib.orc_knn.argtypes = [_f32p, C.c_uint64, _f32p, C.c_uint64, C.c_uint32, C.c_uint32,
                                C.c_int, C.c_int, _u32p, _f32p]
        lib.orc_kth.argtypes = [_f32p, C.c_uint64, C.c_uint32, C.c_uint32, C.c_int, C.c_int, _f32p]
        lib.orc_kth_f64.argtypes = [_f32p, C.c_uint64, C.c_uint32, C.c_uint32, _f64p]
        lib.orc_kth_f64_rows.argtypes = [_f32p, C.c_uint64, C.c_uint32, C.c_uint32, _u32p,
                                         C.c_uint64, _f64p]
        lib.orc_rknn_gt_f64.argtypes = [_f32p, C.c_uint64, _f32p, C.c_uint64, C.c_uint32, _f64p,
                                        _u64p, _u32p, _f64p, C.c_uint64, C.POINTER(C.c_uint64)]
        lib.orc_rknn_gt.argtypes = [_f32p, C.c_uint64, _f32p, C.c_uint64, C.c_uint32, _f32p,
                                    C.c_int, C.c_int, _u64p, _u32p, _f32p, C.c_uint64,
                                    C.POINTER(C.c_uint64)]
        lib.orc_csr.argtypes = [_u32p, _u16p, C.c_uint64, C.c_uint32, C.POINTER(C.c_uint32),
                                C.POINTER(C.c_int), _u64p, C.c_void_p]
        lib.orc_csr_params.argtypes = [C.c_uint64, C.c_uint32, C.POINTER(C.c_uint32),
                                       C.POINTER(C.c_uint32), C.POINTER(C.c_int)]
        lib.orc_search_bc.argtypes = [_u64p, C.c_void_p, C.c_int, C.c_uint32, _f32p, _u16p,
                                      C.c_uint32, _f32p, C.c_uint64, C.c_uint32, _f32p,
                                      C.c_uint64, _u32p, C.c_uint32, C.c_void_p, C.c_uint32,
                                      C.c_uint32, C.c_float, C.c_int, C.c_int, _u64p, _u32p,
                                      C.c_uint64, C.POINTER(C.c_uint64)]
        lib.orc_nndescent.argtypes = [_f32p, C.c_uint64, C.c_uint32, C.c_uint32, _u32p, _f32p,
                                      _u16p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                      C.c_void_p]
        lib.orc_nndescent_flags.argtypes = [_f32p, C.c_uint64, C.c_uint32, C.c_uint32, _u32p, _f32p,
                                            _u16p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                            C.c_void_p, C.c_void_p]
        lib.orc_knng_recall.restype = C.c_double
        lib.orc_insert_affect.argtypes = [_f32p, C.c_uint64, C.c_uint32, _u32p, _f32p, _u16p, C.c_uint32,
                                          C.c_uint64, C.c_uint64, C.c_uint32, C.c_int, C.c_int, C.c_int,
                                          C.c_void_p]
        lib.orc_knng_recall.argtypes = [_u32p, _u16p, C.c_uint32, _u32p, C.c_uint32,
                                        C.c_uint64, C.c_uint32]
        lib.orc_rknn_recall.restype = C.c_double
        lib.orc_rknn_recall.argtypes = [_u64p, _u32p, _u64p, _u32p, C.c_uint64, C.c_uint64,
                                        C.POINTER(C.c_double)]
        lib.orc_num_threads.restype = C.c_int
        lib.orc_set_threads.argtypes = [C.c_int]
        _orc = lib
    return _orc


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def l2sqr(a, b, W=L2_W, fma=L2_FMA) -> float:
    a, b = _f32(a), _f32(b)
    return orc().orc_l2sqr(a, b, a.shape[-1], W, fma)


def l2sqr_f64(a, b) -> float:
    """Squared L2 in FP64: sequential sum of (a_i - b_i)^2, no fused multiply-add."""
    a, b = _f32(a), _f32(b)
    return orc().orc_l2sqr_f64(a, b, a.shape[-1])


def knn(x, q, k, W=L2_W, fma=L2_FMA):
    x, q = _f32(x), _f32(q)
    ids = np.empty((q.shape[0], k), np.uint32)
    d2 = np.empty((q.shape[0], k), np.float32)
    rc = orc().orc_knn(x, x.shape[0], q, q.shape[0], x.shape[1], k, W, fma, ids, d2)
    assert rc == 0
    return ids, d2


def kth(x, k, W=L2_W, fma=L2_FMA):
    x = _f32(x)
    out = np.empty(x.shape[0], np.float32)
    assert orc().orc_kth(x, x.shape[0], x.shape[1], k, W, fma, out) == 0
    return out


def kth_f64(x, k):
    x = _f32(x)
    out = np.empty(x.shape[0], np.float64)
    assert orc().orc_kth_f64(x, x.shape[0], x.shape[1], k, out) == 0
    return out


def kth_f64_rows(x, k, rows):
    """FP64 r_k of the listed rows only (same definition as kth_f64)."""
    x = _f32(x)
    rows = np.ascontiguousarray(rows, dtype=np.uint32)
    out = np.empty(rows.shape[0], np.float64)
    assert orc().orc_kth_f64_rows(x, x.shape[0], x.shape[1], k, rows, rows.shape[0], out) == 0
    return out


def rknn_gt_f64(x, q, kth64):
    """FP64-judged RkNN: (offsets u64[nq+1], ids u32, d2 f64) sorted by (d2, id)."""
    x, q = _f32(x), _f32(q)
    kth64 = np.ascontiguousarray(kth64, dtype=np.float64)
    nq = q.shape[0]
    off = np.zeros(nq + 1, np.uint64)
    need = C.c_uint64(0)
    orc().orc_rknn_gt_f64(x, x.shape[0], q, nq, x.shape[1], kth64, off, np.zeros(1, np.uint32),
                          np.zeros(1, np.float64), 0, C.byref(need))
    n = int(need.value)
    ids = np.empty(max(n, 1), np.uint32)
    d2 = np.empty(max(n, 1), np.float64)
    rc = orc().orc_rknn_gt_f64(x, x.shape[0], q, nq, x.shape[1], kth64, off, ids, d2, n,
                               C.byref(need))
    assert rc == 0
    return off, ids[:n], d2[:n]


def rknn_gt(x, q, kth_arr, W=L2_W, fma=L2_FMA):
    """Returns (offsets u64[nq+1], ids u32, d2 f32) sorted by (d2, id) per query."""
    x, q, kth_arr = _f32(x), _f32(q), _f32(kth_arr)
    nq = q.shape[0]
    off = np.zeros(nq + 1, np.uint64)
    need = C.c_uint64(0)
    dummy_i = np.zeros(1, np.uint32)
    dummy_f = np.zeros(1, np.float32)
    orc().orc_rknn_gt(x, x.shape[0], q, nq, x.shape[1], kth_arr, W, fma, off, dummy_i,
                      dummy_f, 0, C.byref(need))
    n = int(need.value)
    ids = np.empty(max(n, 1), np.uint32)
    d2 = np.empty(max(n, 1), np.float32)
    rc = orc().orc_rknn_gt(x, x.shape[0], q, nq, x.shape[1], kth_arr, W, fma, off, ids, d2,
                           n, C.byref(need))
    assert rc == 0
    return off, ids[:n], d2[:n]


def csr(knn_ids, cnt, K):
    """Canonical reverse CSR: (offsets u64[n+1], entries u32|u64, rank_bits, wide)."""
    knn_ids = np.ascontiguousarray(knn_ids, np.uint32).reshape(-1)
    cnt = np.ascontiguousarray(cnt, np.uint16)
    n = cnt.shape[0]
    rb, wide = C.c_uint32(0), C.c_int(0)
    off = np.zeros(n + 1, np.uint64)
    rc = orc().orc_csr(knn_ids, cnt, n, K, C.byref(rb), C.byref(wide), off, None)
    if rc != 0:
        raise ValueError(f"orc_csr rc={rc}")
    ent = np.empty(int(off[-1]), np.uint64 if wide.value else np.uint32)
    rc = orc().orc_csr(knn_ids, cnt, n, K, C.byref(rb), C.byref(wide), off,
                       ent.ctypes.data_as(C.c_void_p))
    assert rc == 0
    return off, ent, rb.value, wide.value


def csr_params(n, K):
    rb, nb, wide = C.c_uint32(0), C.c_uint32(0), C.c_int(0)
    rc = orc().orc_csr_params(n, K, C.byref(rb), C.byref(nb), C.byref(wide))
    return rc, rb.value, nb.value, wide.value


def search_bc(offsets, entries, wide, rank_bits, knn_d, cnt, K, x, q, proxies, k, K_prime,
              threshold=1.0, k_per_q=None, W=L2_W, fma=L2_FMA):
    """Phases B+C for given proxies; returns (res_off u64[nq+1], ids u32)."""
    x, q = _f32(x), _f32(q)
    nq = q.shape[0]
    proxies = np.ascontiguousarray(proxies, np.uint32).reshape(nq, -1)
    m = proxies.shape[1]
    entries = np.ascontiguousarray(entries)
    kp = None
    if k_per_q is not None:
        k_per_q = np.ascontiguousarray(k_per_q, np.uint32)
        kp = k_per_q.ctypes.data_as(C.c_void_p)
    knn_d = _f32(knn_d).reshape(-1)
    cnt = np.ascontiguousarray(cnt, np.uint16)
    cap = max(1, nq * m * max(1, K_prime))
    offsets = np.ascontiguousarray(offsets, np.uint64)
    while True:
        res = np.empty(cap, np.uint32)
        off = np.zeros(nq + 1, np.uint64)
        need = C.c_uint64(0)
        rc = orc().orc_search_bc(offsets, entries.ctypes.data_as(C.c_void_p), int(wide),
                                 rank_bits, knn_d, cnt, K, x, x.shape[0], x.shape[1], q, nq,
                                 proxies, m, kp, k, K_prime, float(threshold), W, fma, off, res,
                                 cap, C.byref(need))
        if rc == -5 and need.value > cap:
            cap = int(need.value)
            continue
        assert rc == 0, rc
        return off, res[: int(off[-1])].copy()


def nndescent(x, ids, dists, cnt, K, iters, sample, W=L2_W, fma=L2_FMA):
    """Single-threaded refineKNNG restatement; returns (ids, dists, cnt, updates, dists_cnt)."""
    x = _f32(x)
    ids = np.ascontiguousarray(ids, np.uint32).reshape(-1).copy()
    dists = _f32(dists).reshape(-1).copy()
    cnt = np.ascontiguousarray(cnt, np.uint16).copy()
    su = np.zeros(iters, np.uint64)
    sd = np.zeros(iters, np.uint64)
    rc = orc().orc_nndescent(x, x.shape[0], x.shape[1], K, ids, dists, cnt, iters, sample, W,
                             fma, su.ctypes.data_as(C.c_void_p), sd.ctypes.data_as(C.c_void_p))
    assert rc == 0
    n = x.shape[0]
    return ids.reshape(n, K), dists.reshape(n, K), cnt, su, sd


def nndescent_flags(x, ids, dists, cnt, flags, K, iters, sample, W=L2_W, fma=L2_FMA):
    """nndescent resumed from a state with the entries' "new" flags (u8
    [n, K]); returns (ids, dists, cnt, flags, updates, dists_cnt)."""
    x = _f32(x)
    ids = np.ascontiguousarray(ids, np.uint32).reshape(-1).copy()
    dists = _f32(dists).reshape(-1).copy()
    cnt = np.ascontiguousarray(cnt, np.uint16).copy()
    fl = np.ascontiguousarray(flags, np.uint8).reshape(-1).copy()
    su = np.zeros(iters, np.uint64)
    sd = np.zeros(iters, np.uint64)
    rc = orc().orc_nndescent_flags(x, x.shape[0], x.shape[1], K, ids, dists, cnt, iters, sample, W,
                                   fma, su.ctypes.data_as(C.c_void_p), sd.ctypes.data_as(C.c_void_p),
                                   fl.ctypes.data_as(C.c_void_p))
    assert rc == 0
    n = x.shape[0]
    return ids.reshape(n, K), dists.reshape(n, K), cnt, fl.reshape(n, K), su, sd


def insert_affect(x, ids, dists, cnt, K, b0, b1, m_update=0, bidir=False, W=L2_W, fma=L2_FMA):
    """Phases 2-5 of insertPointDynamic for rows [b0, b1) holding their seeds
    (batch semantics; b1 = b0 + 1 is the reference's step for one point);
    returns (ids, dists, cnt, stats = [scanned, candidates, offers, updates])."""
    x = _f32(x)
    n = x.shape[0]
    ids = np.ascontiguousarray(ids, np.uint32).reshape(-1).copy()
    dists = _f32(dists).reshape(-1).copy()
    cnt = np.ascontiguousarray(cnt, np.uint16).copy()
    st = np.zeros(4, np.uint64)
    rc = orc().orc_insert_affect(x, n, x.shape[1], ids, dists, cnt, K, b0, b1, m_update, int(bidir), W, fma,
                                 st.ctypes.data_as(C.c_void_p))
    assert rc == 0, rc
    return ids.reshape(n, K), dists.reshape(n, K), cnt, st


def knng_recall(ids, cnt, K, exact, kk):
    ids = np.ascontiguousarray(ids, np.uint32).reshape(-1)
    exact = np.ascontiguousarray(exact, np.uint32)
    return orc().orc_knng_recall(ids, np.ascontiguousarray(cnt, np.uint16), K,
                                 exact.reshape(-1), exact.shape[1], cnt.shape[0], kk)


def rknn_recall(gt_off, gt_ids, r_off, r_ids, n):
    micro = C.c_double(0)
    gt_ids = np.ascontiguousarray(gt_ids, np.uint32)
    r_ids = np.ascontiguousarray(r_ids, np.uint32)
    if gt_ids.size == 0:
        gt_ids = np.zeros(1, np.uint32)
    if r_ids.size == 0:
        r_ids = np.zeros(1, np.uint32)
    mac = orc().orc_rknn_recall(np.ascontiguousarray(gt_off, np.uint64), gt_ids,
                                np.ascontiguousarray(r_off, np.uint64), r_ids,
                                len(gt_off) - 1, n, C.byref(micro))
    return mac, micro.value


def set_threads(t: int) -> None:
    orc().orc_set_threads(int(t))


from .ref import RefIndex, ref_available, ref_lib_path  # noqa: E402,F401
