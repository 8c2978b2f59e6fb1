"""ctypes binding of libhrnn_hip.so (the C-ABI declared in include/hrnn_dev.h).

This is the Python host mirror of the device layer: every method calls one
C-ABI entry point; nothing here computes results on the CPU.  The library must
have been built (``python -m hrnn_amd.build``); importing this module without
it raises at first use, and a missing GPU raises HrnnError from ``Device()``.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libhrnn_hip.so")

PREC_F32, PREC_BF16, PREC_F16, PREC_U8 = 0, 1, 2, 3
MLP_PARAMS = 7312              # HRNN_MLP_PARAMS
MLP_MODEL = MLP_PARAMS + 64    # + feature mean / std
_PREC = {"f32": PREC_F32, "fp32": PREC_F32, "bf16": PREC_BF16, "f16": PREC_F16, "fp16": PREC_F16,
         "u8": PREC_U8}

STATUS = {0: "OK", -1: "E_ARG", -2: "E_HIP", -3: "E_STATE", -4: "E_OOM", -5: "E_CAP",
          -6: "E_RANGE"}

# Every symbol include/hrnn_dev.h declares (checked by the tests).
EXPORTS = [
    "hrnn_dev_device_count", "hrnn_dev_open", "hrnn_dev_close", "hrnn_dev_last_error",
    "hrnn_dev_sync", "hrnn_dev_stream", "hrnn_dev_set_base", "hrnn_dev_set_base_dptr",
    "hrnn_dev_append_base", "hrnn_dev_base_is_small_int", "hrnn_dev_base_is_u8", "hrnn_dev_base_ptr",
    "hrnn_dev_set_vector_precision", "hrnn_dev_set_tuning", "hrnn_dev_get_tuning",
    "hrnn_dev_set_knng", "hrnn_dev_get_knng", "hrnn_dev_get_knng_flags", "hrnn_dev_set_knng_flags",
    "hrnn_dev_knng_ptrs", "hrnn_dev_hnsw_build", "hrnn_dev_get_hnsw", "hrnn_dev_nndescent",
    "hrnn_dev_insert_dynamic", "hrnn_dev_insert_affect",
    "hrnn_dev_build_rknn_csr", "hrnn_dev_get_rknn_csr", "hrnn_dev_set_rknn_csr",
    "hrnn_dev_rknn_search", "hrnn_dev_rknn_search_dptr", "hrnn_dev_rknn_search_noalloc_dptr", "hrnn_dev_fetch_results", "hrnn_dev_set_hnsw", "hrnn_dev_proxies",
    "hrnn_dev_proxies_dptr", "hrnn_dev_rknn_query_dptr", "hrnn_dev_bf_kth", "hrnn_dev_bf_knn",
    "hrnn_dev_bf_rknn", "hrnn_dev_bf_kth_ex", "hrnn_dev_bf_rknn_f64", "hrnn_dev_l2sqr_pairs",
    "hrnn_dev_mlp_forward",
    "hrnn_dev_mlp_grad", "hrnn_dev_mlp_train", "hrnn_dev_rknn_search_learned_dptr",
    "hrnn_dev_stage1_features_dptr", "hrnn_dev_rknn_query_adaptive_dptr",
    "hrnn_synth_mixture",
    "hrnn_synth_halfnormal_scaled", "hrnn_synth_sphere_mixture", "hrnn_synth_split_perm",
]


class HrnnError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


class HnswBuildStats(C.Structure):
    _fields_ = [("batches", C.c_uint64), ("expansions", C.c_uint64), ("evaluations", C.c_uint64),
                ("dropped_links", C.c_uint64), ("ms", C.c_double), ("seed_expansions", C.c_uint64),
                ("seed_evaluations", C.c_uint64)]


class InsertStats(C.Structure):
    _fields_ = [("batches", C.c_uint64), ("scanned", C.c_uint64), ("candidates", C.c_uint64),
                ("offers", C.c_uint64), ("updates", C.c_uint64), ("retries", C.c_uint64), ("ms", C.c_double),
                ("hnsw_ms", C.c_double), ("csr_ms", C.c_double), ("affect_ms", C.c_double),
                ("merge_ms", C.c_double), ("hnsw", HnswBuildStats)]


class NndIterStats(C.Structure):
    _fields_ = [("dists", C.c_uint64), ("updates", C.c_uint64), ("proposals", C.c_uint64),
                ("dropped", C.c_uint64), ("ms", C.c_double), ("rows", C.c_uint64)]


class BfStats(C.Structure):
    _fields_ = [("rows", C.c_uint64), ("certified", C.c_uint64), ("redo_wide", C.c_uint64),
                ("redo_f32", C.c_uint64), ("uncertified", C.c_uint64), ("ms_first", C.c_double),
                ("ms_total", C.c_double)]


_lib = None


def lib() -> C.CDLL:
    """Load libhrnn_hip.so (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(
            f"{LIB_PATH} is missing: build it with `python -m hrnn_amd.build` "
            "(there is no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    vp, u32, u64, i, f = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int, C.c_float
    P = C.POINTER
    sig = {
        "hrnn_dev_device_count": [P(i)],
        "hrnn_dev_open": [i, P(vp)],
        "hrnn_dev_close": [vp],
        "hrnn_dev_sync": [vp],
        "hrnn_dev_set_base": [vp, vp, u64, u32],
        "hrnn_dev_set_base_dptr": [vp, vp, u64, u32],
        "hrnn_dev_append_base": [vp, vp, u64],
        "hrnn_dev_base_is_small_int": [vp, P(i)],
        "hrnn_dev_base_is_u8": [vp, P(i)],
        "hrnn_dev_set_vector_precision": [vp, i],
        "hrnn_dev_set_tuning": [vp, C.c_char_p, i],
        "hrnn_dev_get_tuning": [vp, C.c_char_p, P(i)],
        "hrnn_dev_set_knng": [vp, u32, vp, vp, vp],
        "hrnn_dev_get_knng": [vp, vp, vp, vp],
        "hrnn_dev_knng_ptrs": [vp, vp, vp, vp],
        "hrnn_dev_get_knng_flags": [vp, vp],
        "hrnn_dev_hnsw_build": [vp, vp, u32, u32, u32, P(u32), P(i), vp],
        "hrnn_dev_get_hnsw": [vp, vp, vp, u64],
        "hrnn_dev_set_knng_flags": [vp, vp],
        "hrnn_dev_nndescent": [vp, i, i, u32, vp],
        "hrnn_dev_insert_dynamic": [vp, u64, vp, u32, u32, u32, i, P(u32), P(i), vp],
        "hrnn_dev_insert_affect": [vp, u64, u64, u32, i, vp],
        "hrnn_dev_build_rknn_csr": [vp, P(u32), P(i), P(u64)],
        "hrnn_dev_get_rknn_csr": [vp, vp, vp],
        "hrnn_dev_set_rknn_csr": [vp, u32, i, u64, vp, vp, u64],
        "hrnn_dev_rknn_search": [vp, vp, u32, vp, u32, vp, u32, u32, f, vp, vp, u64, P(u64)],
        "hrnn_dev_rknn_search_dptr": [vp, vp, u32, vp, u32, vp, u32, u32, f, vp, vp, u32, vp],
        "hrnn_dev_fetch_results": [vp, u32, vp, vp, u32, vp, vp, u64, P(u64)],
        "hrnn_dev_rknn_search_noalloc_dptr": [vp, vp, u32, vp, u32, vp, u32, u32, f, vp, vp, u32, vp],
        "hrnn_dev_set_hnsw": [vp, u32, u32, vp, vp, vp, vp, u64, u32, i],
        "hrnn_dev_proxies": [vp, vp, u32, u32, u32, vp],
        "hrnn_dev_proxies_dptr": [vp, vp, u32, u32, u32, vp, vp],
        "hrnn_dev_rknn_query_dptr": [vp, vp, u32, u32, u32, vp, u32, u32, f, vp, vp, vp, u32, vp],
        "hrnn_dev_bf_kth": [vp, u32, i, vp, vp],
        "hrnn_dev_bf_knn": [vp, vp, u32, u32, i, vp, vp],
        "hrnn_dev_bf_rknn": [vp, vp, u32, vp, i, vp, vp, vp, u64, P(u64)],
        "hrnn_dev_bf_kth_ex": [vp, u32, i, vp, vp, vp, vp, vp, P(BfStats)],
        "hrnn_dev_bf_rknn_f64": [vp, vp, u32, vp, vp, vp, vp, vp, u64, P(u64), P(u64), P(u64)],
        "hrnn_dev_l2sqr_pairs": [vp, vp, u64, vp, vp, u64, vp],
        "hrnn_dev_mlp_forward": [vp, vp, vp, u32, vp],
        "hrnn_dev_mlp_grad": [vp, vp, vp, vp, u32, vp, vp],
        "hrnn_dev_mlp_train": [vp, vp, vp, vp, u32, i, u32, f, f, u64, vp],
        "hrnn_dev_rknn_search_learned_dptr": [vp, vp, u32, vp, vp, vp, f, vp, u32, vp, u32, u32, f, vp,
                                              vp, u32, vp],
        "hrnn_dev_stage1_features_dptr": [vp, vp, u32, u32, u32, u32, vp, vp, vp],
        "hrnn_dev_rknn_query_adaptive_dptr": [vp, vp, u32, vp, vp, u32, f, u32, u32, u32, u32, f, vp,
                                              vp, vp, vp, u32, vp, vp],
        "hrnn_synth_mixture": [u64, u64, u32, u32, f, f, f, i, C.c_double, u64, vp, vp],
        "hrnn_synth_halfnormal_scaled": [u64, u64, u64, u32, vp],
        "hrnn_synth_sphere_mixture": [u64, u64, u64, u32, u32, f, vp, vp],
        "hrnn_synth_split_perm": [u64, u32, vp],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = C.c_int
    L.hrnn_dev_last_error.argtypes = [vp]
    L.hrnn_dev_last_error.restype = C.c_char_p
    L.hrnn_dev_stream.argtypes = [vp]
    L.hrnn_dev_stream.restype = vp
    L.hrnn_dev_base_ptr.argtypes = [vp]
    L.hrnn_dev_base_ptr.restype = vp
    _lib = L
    return L


def _p(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def device_count() -> int:
    c = C.c_int(0)
    rc = lib().hrnn_dev_device_count(C.byref(c))
    return c.value if rc == 0 else 0


# ---- synthetic data (host, deterministic) ---------------------------------

def synth_mixture(seed: int, n: int, d: int, n_centres: int, sigma: float, lo=0.0, hi=255.0,
                  round_int=True, zipf_s=0.0, centre_seed: int = 0, return_centres=False):
    out = np.empty((n, d), np.float32)
    cen = np.empty(n, np.uint32) if return_centres else None
    rc = lib().hrnn_synth_mixture(seed, n, d, n_centres, sigma, lo, hi, int(round_int),
                                  float(zipf_s), centre_seed, _p(out), _p(cen))
    if rc:
        raise HrnnError(rc, "synth_mixture")
    return (out, cen) if return_centres else out


def synth_halfnormal_scaled(seed: int, scale_seed: int, n: int, d: int):
    out = np.empty((n, d), np.float32)
    rc = lib().hrnn_synth_halfnormal_scaled(seed, scale_seed, n, d, _p(out))
    if rc:
        raise HrnnError(rc, "synth_halfnormal_scaled")
    return out


def synth_split_perm(n: int, seed: int = 42):
    """0 .. n-1 shuffled as std::shuffle with std::mt19937(seed) shuffles them
    (the reference dynamic test's split order)."""
    out = np.empty(n, np.uint32)
    rc = lib().hrnn_synth_split_perm(n, seed, _p(out))
    if rc:
        raise HrnnError(rc, "synth_split_perm")
    return out


def synth_sphere_mixture(seed: int, centre_seed: int, n: int, d: int, n_centres: int,
                         kappa: float, return_centres=False):
    out = np.empty((n, d), np.float32)
    cen = np.empty(n, np.uint32) if return_centres else None
    rc = lib().hrnn_synth_sphere_mixture(seed, centre_seed, n, d, n_centres, kappa, _p(out),
                                         _p(cen))
    if rc:
        raise HrnnError(rc, "synth_sphere_mixture")
    return (out, cen) if return_centres else out


@dataclass
class CsrInfo:
    rank_bits: int
    wide: int
    n_entries: int


class Device:
    """One device-resident HRNN index state (an hrnn_dev_t handle)."""

    def __init__(self, device: int = 0):
        self.L = lib()
        h = C.c_void_p()
        rc = self.L.hrnn_dev_open(device, C.byref(h))
        if rc:
            raise HrnnError(rc, f"hrnn_dev_open({device}) failed (no usable GPU?)")
        self.h = h
        self.n = 0
        self.d = 0
        self.K = 0

    @classmethod
    def borrow(cls, h, n: int, d: int, K: int = 0) -> "Device":
        """A view of a handle owned elsewhere (e.g. Index.device_handle()):
        never closes it; the owner must outlive the view."""
        self = cls.__new__(cls)
        self.L, self.h, self.n, self.d, self.K = lib(), h, n, d, K
        self._borrowed = True
        return self

    def close(self):
        if getattr(self, "_borrowed", False):
            self.h = None
            return
        if getattr(self, "h", None):
            self.L.hrnn_dev_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, rc: int):
        if rc != 0:
            raise HrnnError(rc, self.L.hrnn_dev_last_error(self.h).decode())

    # -- base / graph ------------------------------------------------------
    def set_base(self, x: np.ndarray):
        x = np.ascontiguousarray(x, np.float32)
        self._ck(self.L.hrnn_dev_set_base(self.h, _p(x), x.shape[0], x.shape[1]))
        self.n, self.d = x.shape

    def set_base_dptr(self, ptr: int, n: int, d: int):
        self._ck(self.L.hrnn_dev_set_base_dptr(self.h, C.c_void_p(ptr), n, d))
        self.n, self.d = n, d

    def append_base(self, x: np.ndarray):
        x = np.ascontiguousarray(x, np.float32)
        self._ck(self.L.hrnn_dev_append_base(self.h, _p(x), x.shape[0]))
        self.n += x.shape[0]

    def base_is_u8(self) -> bool:
        """Every base value an integer in [0, 255] (the u8 mirror and the
        int8 NNDescent join apply)."""
        v = C.c_int(0)
        self._ck(self.L.hrnn_dev_base_is_u8(self.h, C.byref(v)))
        return bool(v.value)

    def base_is_small_int(self) -> bool:
        v = C.c_int(0)
        self._ck(self.L.hrnn_dev_base_is_small_int(self.h, C.byref(v)))
        return bool(v.value)

    def set_vector_precision(self, prec: str):
        self._ck(self.L.hrnn_dev_set_vector_precision(self.h, _PREC[prec]))

    def set_tuning(self, key: str, value: int):
        self._ck(self.L.hrnn_dev_set_tuning(self.h, key.encode(), int(value)))

    def get_tuning(self, key: str) -> int:
        v = C.c_int(0)
        self._ck(self.L.hrnn_dev_get_tuning(self.h, key.encode(), C.byref(v)))
        return v.value

    def set_knng(self, ids, dists, cnt):
        ids = np.ascontiguousarray(ids, np.uint32)
        dists = np.ascontiguousarray(dists, np.float32)
        cnt = np.ascontiguousarray(cnt, np.uint16)
        K = ids.shape[1] if ids.ndim == 2 else ids.size // self.n
        self._ck(self.L.hrnn_dev_set_knng(self.h, K, _p(ids), _p(dists), _p(cnt)))
        self.K = K

    def get_knng(self):
        ids = np.empty((self.n, self.K), np.uint32)
        dists = np.empty((self.n, self.K), np.float32)
        cnt = np.empty(self.n, np.uint16)
        self._ck(self.L.hrnn_dev_get_knng(self.h, _p(ids), _p(dists), _p(cnt)))
        return ids, dists, cnt

    def get_knng_flags(self):
        f = np.empty((self.n, self.K), np.uint8)
        self._ck(self.L.hrnn_dev_get_knng_flags(self.h, _p(f)))
        return f

    def set_knng_flags(self, flags):
        f = np.ascontiguousarray(flags, np.uint8).reshape(self.n, self.K)
        self._ck(self.L.hrnn_dev_set_knng_flags(self.h, _p(f)))

    def nndescent(self, iters: int, sample: int, bucket: int = 0):
        st = (NndIterStats * max(iters, 1))()
        self._ck(self.L.hrnn_dev_nndescent(self.h, iters, sample, bucket, st))
        return [dict(dists=s.dists, updates=s.updates, proposals=s.proposals,
                     dropped=s.dropped, ms=s.ms, rows=s.rows) for s in st[:iters]]

    # -- incremental insertion --------------------------------------------------
    @staticmethod
    def _insert_stats(st):
        out = {k: getattr(st, k) for k, _ in InsertStats._fields_ if k != "hnsw"}
        out["hnsw"] = {k: getattr(st.hnsw, k) for k, _ in HnswBuildStats._fields_}
        return out

    def insert_affect(self, b0: int, b1: int, m_update: int = 0, bidirectional: bool = False):
        """Phases 2-5 of insertPointDynamic for rows [b0, b1) whose lists hold
        their seeds (hrnn_dev_insert_affect); returns the call's counters."""
        st = InsertStats()
        self._ck(self.L.hrnn_dev_insert_affect(self.h, b0, b1, m_update, int(bidirectional), C.byref(st)))
        return self._insert_stats(st)

    def insert_dynamic(self, n_old: int, levels, M: int, efc: int, m_update: int = 0, bidirectional: bool = False):
        """Incremental insertion of the resident rows [n_old, n) (appended) into
        the resident HNSW graph and KNN lists (hrnn_dev_insert_dynamic);
        returns (entry, maxlevel, counters)."""
        lv = np.ascontiguousarray(levels, np.int32)
        assert lv.shape[0] == self.n
        e, ml, st = C.c_uint32(0), C.c_int(0), InsertStats()
        self._ck(self.L.hrnn_dev_insert_dynamic(self.h, n_old, _p(lv), M, efc, m_update, int(bidirectional),
                                                C.byref(e), C.byref(ml), C.byref(st)))
        return e.value, ml.value, self._insert_stats(st)

    # -- reverse CSR -----------------------------------------------------------
    def build_rknn_csr(self) -> CsrInfo:
        rb, wide, ne = C.c_uint32(0), C.c_int(0), C.c_uint64(0)
        self._ck(self.L.hrnn_dev_build_rknn_csr(self.h, C.byref(rb), C.byref(wide), C.byref(ne)))
        self.csr = CsrInfo(rb.value, wide.value, ne.value)
        return self.csr

    def get_rknn_csr(self):
        off = np.empty(self.n + 1, np.uint64)
        ent = np.empty(self.csr.n_entries, np.uint64 if self.csr.wide else np.uint32)
        self._ck(self.L.hrnn_dev_get_rknn_csr(self.h, _p(off), _p(ent)))
        return off, ent

    def set_rknn_csr(self, rank_bits, wide, offsets, entries):
        offsets = np.ascontiguousarray(offsets, np.uint64)
        entries = np.ascontiguousarray(entries, np.uint64 if wide else np.uint32)
        self._ck(self.L.hrnn_dev_set_rknn_csr(self.h, rank_bits, int(wide), self.n, _p(offsets),
                                              _p(entries), entries.size))
        self.csr = CsrInfo(rank_bits, int(wide), entries.size)

    # -- search ----------------------------------------------------------------
    def rknn_search(self, q, proxies, k, K_prime, threshold=1.0, k_per_q=None):
        q = np.ascontiguousarray(q, np.float32)
        nq = q.shape[0]
        proxies = np.ascontiguousarray(proxies, np.uint32).reshape(nq, -1)
        m = proxies.shape[1]
        kp = None if k_per_q is None else np.ascontiguousarray(k_per_q, np.uint32)
        off = np.zeros(nq + 1, np.uint64)
        need = C.c_uint64(0)
        cap = max(1024, nq * 64)
        while True:
            ids = np.empty(cap, np.uint32)
            rc = self.L.hrnn_dev_rknn_search(self.h, _p(q), nq, _p(proxies), m, _p(kp), k,
                                             K_prime, threshold, _p(off), _p(ids), cap,
                                             C.byref(need))
            if rc == -5 and need.value > cap:
                cap = int(need.value)
                continue
            self._ck(rc)
            return off, ids[: int(off[-1])].copy()

    def rknn_search_resident(self, q, proxies, k, K_prime, threshold=1.0, cap_per_q=4096):
        """The resident form (hrnn_dev_rknn_search_dptr) with the answers
        downloaded by hrnn_dev_fetch_results: (off[nq+1], ids)."""
        from . import hiprt
        q = np.ascontiguousarray(q, np.float32)
        nq = q.shape[0]
        proxies = np.ascontiguousarray(proxies, np.uint32)
        m = proxies.shape[-1]
        proxies = proxies.reshape(nq, m)
        qd = hiprt.DeviceArray.from_host(q)
        pd = hiprt.DeviceArray.from_host(proxies)
        cd = hiprt.DeviceArray((max(nq, 1),), np.uint32)
        idd = hiprt.DeviceArray((max(nq, 1) * cap_per_q,), np.uint32)
        self._ck(self.L.hrnn_dev_rknn_search_dptr(self.h, qd.ptr, nq, pd.ptr, m, None, k, K_prime, threshold,
                                                  cd.ptr, idd.ptr, cap_per_q, None))
        return self.fetch_results(nq, cd.ptr, idd.ptr, cap_per_q)

    def fetch_results(self, nq, cnt_ptr, ids_ptr, cap_per_q, cap=None):
        """hrnn_dev_fetch_results into host arrays: (off[nq+1] u64, ids u32)."""
        off = np.zeros(nq + 1, np.uint64)
        need = C.c_uint64(0)
        cap = max(1024, nq * 16) if cap is None else cap
        while True:
            ids = np.empty(max(cap, 1), np.uint32)
            rc = self.L.hrnn_dev_fetch_results(self.h, nq, cnt_ptr, ids_ptr, cap_per_q, _p(off), _p(ids),
                                               cap, C.byref(need))
            if rc == -5 and need.value > cap and int(need.value) <= nq * cap_per_q:
                cap = int(need.value)
                continue
            self._ck(rc)
            return off, ids[: int(off[-1])].copy()

    # -- phase A (HNSW beam) on the GPU -------------------------------------------
    def set_hnsw(self, maxM0, maxM, l0, levels, up_off, up_pool, entry, maxlevel):
        l0 = np.ascontiguousarray(l0, np.uint32)
        levels = np.ascontiguousarray(levels, np.int32)
        up_off = np.ascontiguousarray(up_off, np.uint32)
        up_pool = np.ascontiguousarray(up_pool, np.uint32)
        self._ck(self.L.hrnn_dev_set_hnsw(self.h, maxM0, maxM, _p(l0), _p(levels), _p(up_off),
                                          _p(up_pool) if up_pool.size else None, up_pool.size,
                                          entry, maxlevel))

    def proxies(self, q, m, ef):
        q = np.ascontiguousarray(q, np.float32)
        out = np.empty((q.shape[0], m), np.uint32)
        self._ck(self.L.hrnn_dev_proxies(self.h, _p(q), q.shape[0], m, max(ef, m), _p(out)))
        return out

    # -- brute force -----------------------------------------------------------
    def bf_kth(self, k: int, prec="bf16", with_ids=False):
        kk = min(k + 1, self.n)
        kth = np.empty(self.n, np.float32)
        ids = np.empty((self.n, kk), np.uint32) if with_ids else None
        self._ck(self.L.hrnn_dev_bf_kth(self.h, k, _PREC.get(prec, prec), _p(kth), _p(ids)))
        return (kth, ids) if with_ids else kth

    def bf_kth_ex(self, k: int, prec="bf16", raw=False):
        """Certified r_k with the precision study's outputs: a dict of kth
        (FP32), kth64 (FP64), ids [n x (k+1)], stats, and with raw=True the
        first candidate pass's own r_k and ids before re-rank."""
        kk = min(k + 1, self.n)
        kth = np.empty(self.n, np.float32)
        kth64 = np.empty(self.n, np.float64)
        ids = np.empty((self.n, kk), np.uint32)
        rk = np.empty(self.n, np.float32) if raw else None
        ri = np.empty((self.n, kk), np.uint32) if raw else None
        st = BfStats()
        self._ck(self.L.hrnn_dev_bf_kth_ex(self.h, k, _PREC.get(prec, prec), _p(kth), _p(kth64), _p(ids),
                                           _p(rk), _p(ri), C.byref(st)))
        out = {"kth": kth, "kth64": kth64, "ids": ids,
               "stats": {f: getattr(st, f) for f, _ in BfStats._fields_}}
        if raw:
            out["raw_kth"], out["raw_ids"] = rk, ri
        return out

    def bf_rknn_f64(self, q, kth, kth64):
        """RkNN with the FP64 re-check of pairs within FP32 rounding of r_k:
        (offsets, ids, d2 f32, rechecked, flipped)."""
        q = np.ascontiguousarray(q, np.float32)
        kth = np.ascontiguousarray(kth, np.float32)
        kth64 = np.ascontiguousarray(kth64, np.float64)
        nq = q.shape[0]
        off = np.zeros(nq + 1, np.uint64)
        need, rc_n, fl_n = C.c_uint64(0), C.c_uint64(0), C.c_uint64(0)
        cap = max(1024, nq * 64)
        while True:
            ids = np.empty(cap, np.uint32)
            d2 = np.empty(cap, np.float32)
            rc = self.L.hrnn_dev_bf_rknn_f64(self.h, _p(q), nq, _p(kth), _p(kth64), _p(off), _p(ids), _p(d2),
                                             cap, C.byref(need), C.byref(rc_n), C.byref(fl_n))
            if rc == -5 and need.value > cap:
                cap = int(need.value)
                continue
            self._ck(rc)
            n = int(off[-1])
            return off, ids[:n].copy(), d2[:n].copy(), int(rc_n.value), int(fl_n.value)

    def bf_knn(self, q, k: int, prec="bf16"):
        q = np.ascontiguousarray(q, np.float32)
        ids = np.empty((q.shape[0], k), np.uint32)
        d2 = np.empty((q.shape[0], k), np.float32)
        self._ck(self.L.hrnn_dev_bf_knn(self.h, _p(q), q.shape[0], k, _PREC.get(prec, prec),
                                        _p(ids), _p(d2)))
        return ids, d2

    def bf_rknn(self, q, kth, prec="bf16"):
        q = np.ascontiguousarray(q, np.float32)
        kth = np.ascontiguousarray(kth, np.float32)
        nq = q.shape[0]
        off = np.zeros(nq + 1, np.uint64)
        need = C.c_uint64(0)
        cap = max(1024, nq * 64)
        while True:
            ids = np.empty(cap, np.uint32)
            d2 = np.empty(cap, np.float32)
            rc = self.L.hrnn_dev_bf_rknn(self.h, _p(q), nq, _p(kth), _PREC.get(prec, prec),
                                         _p(off), _p(ids), _p(d2), cap, C.byref(need))
            if rc == -5 and need.value > cap:
                cap = int(need.value)
                continue
            self._ck(rc)
            n = int(off[-1])
            return off, ids[:n].copy(), d2[:n].copy()

    # -- learned m ---------------------------------------------------------------
    def stage1_features(self, q, ef1: int, m: int, k: int):
        """Stage-1 phase A at ef1 with the fused 32 features:
        returns (feats[nq, 32], proxies[nq, m], expansions[nq, 2])."""
        from . import hiprt
        q = np.ascontiguousarray(q, np.float32)
        nq = q.shape[0]
        qd = hiprt.DeviceArray.from_host(q)
        pd = hiprt.DeviceArray((nq, m), np.uint32)
        fd = hiprt.DeviceArray((nq, 32), np.float32)
        ed = hiprt.DeviceArray((nq, 2), np.uint32)
        self._ck(self.L.hrnn_dev_stage1_features_dptr(self.h, qd.ptr, nq, ef1, m, k, pd.ptr, fd.ptr,
                                                      ed.ptr))
        self.sync()
        return fd.to_host(), pd.to_host(), ed.to_host()

    def query_adaptive(self, q, classes, ef1: int, m_max: int, k: int, K_prime: int, model=None,
                       tau: float = 0.0, cls=None, threshold: float = 1.0, cap_per_q: int = 4096,
                       phase_a_only: bool = False, pinned: bool = False):
        """Two-stage learned query (hrnn_dev_rknn_query_adaptive_dptr).  classes:
        [16, 2] (ef, m) rows.  With model None, cls gives each query's class.
        Returns (res_off[nq+1], res_ids, cls[nq], proxies[nq, m_max]); with
        phase_a_only (no result buffers: phases B and C are skipped) res_off
        and res_ids are None.  pinned: the queries are passed in page-locked
        host memory (read over the link by the kernels) instead of HBM."""
        from . import hiprt
        q = np.ascontiguousarray(q, np.float32)
        nq = q.shape[0]
        classes = np.asarray(classes, np.uint32).reshape(16, 2)
        cd = hiprt.DeviceArray.from_host(np.concatenate([classes[:, 0], classes[:, 1]]))
        if pinned:
            qd = hiprt.PinnedArray(q.shape, np.float32)
            qd.a[:] = q
            q_ptr = qd.device_ptr()
        else:
            qd = hiprt.DeviceArray.from_host(q)
            q_ptr = qd.ptr
        md = None if model is None else hiprt.DeviceArray.from_host(np.asarray(model, np.float32))
        cl = hiprt.DeviceArray.from_host(np.zeros(nq, np.uint32) if cls is None
                                         else np.asarray(cls, np.uint32))
        pd = hiprt.DeviceArray((nq, m_max), np.uint32)
        rc = hiprt.DeviceArray((nq,), np.uint32)
        ri = hiprt.DeviceArray((nq, cap_per_q), np.uint32)
        self._ck(self.L.hrnn_dev_rknn_query_adaptive_dptr(
            self.h, q_ptr, nq, None if md is None else md.ptr, cd.ptr, int(classes[:, 0].max()),
            tau, ef1, m_max, k, K_prime, threshold, pd.ptr, cl.ptr, None if phase_a_only else rc.ptr,
            None if phase_a_only else ri.ptr, cap_per_q, None, None))
        self.sync()
        if phase_a_only:
            return None, None, cl.to_host(), pd.to_host()
        cnt = rc.to_host().astype(np.uint64)
        if cnt.max(initial=0) > cap_per_q:
            raise HrnnError(-5, "query_adaptive: a query has more results than cap_per_q")
        ids = ri.to_host()
        off = np.zeros(nq + 1, np.uint64)
        off[1:] = np.cumsum(cnt)
        out = np.concatenate([ids[i, :cnt[i]] for i in range(nq)]) if nq else np.empty(0, np.uint32)
        return off, out.astype(np.uint32), cl.to_host(), pd.to_host()

    def mlp_forward(self, params, feats):
        params = np.ascontiguousarray(params, np.float32)
        feats = np.ascontiguousarray(feats, np.float32)
        out = np.empty((feats.shape[0], 16), np.float32)
        self._ck(self.L.hrnn_dev_mlp_forward(self.h, _p(params), _p(feats), feats.shape[0], _p(out)))
        return out

    def mlp_grad(self, params, feats, labels):
        params = np.ascontiguousarray(params, np.float32)
        feats = np.ascontiguousarray(feats, np.float32)
        labels = np.ascontiguousarray(labels, np.uint32)
        g = np.empty(MLP_PARAMS, np.float32)
        loss = C.c_float(0)
        self._ck(self.L.hrnn_dev_mlp_grad(self.h, _p(params), _p(feats), _p(labels), feats.shape[0],
                                          _p(g), C.byref(loss)))
        return g, float(loss.value)

    def mlp_train(self, params, feats, labels, epochs, batch=256, lr=1e-3, weight_decay=0.0,
                  seed=0):
        params = np.ascontiguousarray(params, np.float32).copy()
        feats = np.ascontiguousarray(feats, np.float32)
        labels = np.ascontiguousarray(labels, np.uint32)
        el = np.zeros(max(epochs, 1), np.float32)
        self._ck(self.L.hrnn_dev_mlp_train(self.h, _p(params), _p(feats), _p(labels), feats.shape[0],
                                           epochs, batch, lr, weight_decay, seed, _p(el)))
        return params, el[:epochs]

    def l2sqr_pairs(self, a, ia, ib):
        a = np.ascontiguousarray(a, np.float32)
        ia = np.ascontiguousarray(ia, np.uint32)
        ib = np.ascontiguousarray(ib, np.uint32)
        out = np.empty(ia.size, np.float32)
        self._ck(self.L.hrnn_dev_l2sqr_pairs(self.h, _p(a), a.shape[0], _p(ia), _p(ib), ia.size,
                                             _p(out)))
        return out

    def sync(self):
        self._ck(self.L.hrnn_dev_sync(self.h))
