"""ctypes wrapper of the compiled reference (oracle/_ref) -- TEST INFRASTRUCTURE ONLY.

The library is the reference's own ``include/hnswlib/hrnn.h`` compiled through
``ref_harness.cpp``.  Two builds exist: x86-64-v4 (AVX-512 distance path,
space_l2.h:25-54) and x86-64-v3 (AVX path, space_l2.h:60-91); the one this
host's CPU can run is chosen at load time.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _cpu_has_avx512() -> bool:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    fl = line.split()
                    return all(x in fl for x in ("avx512f", "avx512bw", "avx512dq", "avx512vl"))
    except OSError:
        pass
    return False


def ref_lib_path(variant: str | None = None) -> str:
    if variant is None:
        variant = "avx512" if _cpu_has_avx512() else "avx2"
    return os.path.join(HERE, "_ref", f"libhrnn_ref_{variant}.so")


def ref_available(variant: str | None = None) -> bool:
    return os.path.exists(ref_lib_path(variant))


_libs: dict = {}


def _lib(variant: str | None = None) -> C.CDLL:
    path = ref_lib_path(variant)
    if path in _libs:
        return _libs[path]
    if not os.path.exists(path):
        raise FileNotFoundError(path)
    lib = C.CDLL(path)
    vp, u64, u32, i = C.c_void_p, C.c_uint64, C.c_uint32, C.c_int
    lib.ref_last_error.restype = C.c_char_p
    lib.ref_new.restype = vp
    lib.ref_new.argtypes = [u64, u32, u32, u32, u32, u64]
    lib.ref_free.argtypes = [vp]
    lib.ref_l2.restype = C.c_float
    lib.ref_l2.argtypes = [vp, vp, vp]
    lib.ref_add_points.argtypes = [vp, vp, u64, i, vp]
    lib.ref_set_raw.argtypes = [vp, vp, u64]
    lib.ref_count.restype = u64
    lib.ref_count.argtypes = [vp]
    lib.ref_knn_depth.restype = u32
    lib.ref_knn_depth.argtypes = [vp]
    lib.ref_set_knng.argtypes = [vp, vp, vp, vp]
    lib.ref_get_knng.argtypes = [vp, vp, vp, vp]
    lib.ref_refine.argtypes = [vp, i, i, i]
    lib.ref_mark_knng_built.argtypes = [vp]
    lib.ref_build_rknng.argtypes = [vp]
    lib.ref_csr_info.argtypes = [vp, C.POINTER(u32), C.POINTER(i), C.POINTER(u64)]
    lib.ref_get_csr.argtypes = [vp, vp, vp]
    lib.ref_search.argtypes = [vp, vp, u64, u32, C.c_float, u32, u32, u32, i, vp, vp, u64,
                               C.POINTER(u64)]
    lib.ref_proxies.argtypes = [vp, vp, u64, u32, u32, vp]
    lib.ref_save.argtypes = [vp, C.c_char_p]
    lib.ref_load.restype = vp
    lib.ref_load.argtypes = [C.c_char_p, u32, u64]
    lib.ref_build_rknng_mutable.argtypes = [vp]
    lib.ref_insert_dynamic.argtypes = [vp, vp, u64, vp, u32, i, i]
    lib.ref_get_rknng_mutable.argtypes = [vp, vp, vp]
    lib.ref_get_level0.argtypes = [vp, vp, vp, u32]
    lib.ref_levels.argtypes = [vp, vp, C.POINTER(i), C.POINTER(u32)]
    lib.ref_get_labels.argtypes = [vp, vp]
    _libs[path] = lib
    return lib


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class RefIndex:
    """One reference HRNN<float> instance."""

    def __init__(self, max_elements: int, d: int, M: int = 16, efc: int = 200, K: int = 0,
                 seed: int = 100, variant: str | None = None, _handle=None):
        self.lib = _lib(variant)
        self.d = d
        self.K = K
        self.M = M
        self.h = _handle if _handle is not None else self.lib.ref_new(max_elements, d, M, efc,
                                                                      K, seed)

    def close(self):
        if self.h:
            self.lib.ref_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(self.lib.ref_last_error().decode())

    def l2(self, a, b) -> float:
        a = np.ascontiguousarray(a, np.float32)
        b = np.ascontiguousarray(b, np.float32)
        return float(self.lib.ref_l2(self.h, _p(a), _p(b)))

    def add_points(self, x, nthreads=1, labels=None):
        x = np.ascontiguousarray(x, np.float32)
        lab = None
        if labels is not None:
            labels = np.ascontiguousarray(labels, np.uint64)
            lab = _p(labels)
        self._check(self.lib.ref_add_points(self.h, _p(x), x.shape[0], nthreads, lab))

    def set_raw(self, x):
        x = np.ascontiguousarray(x, np.float32)
        self._check(self.lib.ref_set_raw(self.h, _p(x), x.shape[0]))

    @property
    def count(self) -> int:
        return int(self.lib.ref_count(self.h))

    def set_knng(self, ids, dists, cnt):
        ids = np.ascontiguousarray(ids, np.uint32)
        dists = np.ascontiguousarray(dists, np.float32)
        cnt = np.ascontiguousarray(cnt, np.uint16)
        self._check(self.lib.ref_set_knng(self.h, _p(ids), _p(dists), _p(cnt)))
        self.lib.ref_mark_knng_built(self.h)

    def get_knng(self):
        n = self.count
        ids = np.empty((n, self.K), np.uint32)
        dists = np.empty((n, self.K), np.float32)
        cnt = np.empty(n, np.uint16)
        self._check(self.lib.ref_get_knng(self.h, _p(ids), _p(dists), _p(cnt)))
        return ids, dists, cnt

    def refine(self, iters, sample, nthreads=1):
        self._check(self.lib.ref_refine(self.h, iters, sample, nthreads))

    def build_rknng(self):
        self._check(self.lib.ref_build_rknng(self.h))

    def get_csr(self):
        rb, wide, ne = C.c_uint32(0), C.c_int(0), C.c_uint64(0)
        self._check(self.lib.ref_csr_info(self.h, C.byref(rb), C.byref(wide), C.byref(ne)))
        off = np.empty(self.count + 1, np.uint64)
        ent = np.empty(int(ne.value), np.uint64 if wide.value else np.uint32)
        self.lib.ref_get_csr(self.h, _p(off), _p(ent))
        return off, ent, rb.value, wide.value

    def search(self, q, m, threshold, k, K_prime, ef, nthreads=1):
        """searchRknn per query; returns (offsets u64[nq+1], labels u64)."""
        q = np.ascontiguousarray(q, np.float32)
        nq = q.shape[0]
        off = np.zeros(nq + 1, np.uint64)
        cap = max(1 << 16, 64 * nq)
        while True:
            lab = np.empty(cap, np.uint64)
            need = C.c_uint64(0)
            rc = self.lib.ref_search(self.h, _p(q), nq, m, threshold, k, K_prime, ef, nthreads,
                                     _p(off), _p(lab), cap, C.byref(need))
            if rc == 0:
                return off, lab[: int(off[-1])].copy()
            if int(need.value) > cap:
                cap = int(need.value)
                continue
            self._check(rc)

    def proxies(self, q, m, ef):
        q = np.ascontiguousarray(q, np.float32)
        out = np.empty((q.shape[0], m), np.uint32)
        self._check(self.lib.ref_proxies(self.h, _p(q), q.shape[0], m, ef, _p(out)))
        return out

    def save(self, path: str):
        self._check(self.lib.ref_save(self.h, path.encode()))

    @classmethod
    def load(cls, path: str, d: int, max_elements: int = 0, variant: str | None = None):
        lib = _lib(variant)
        h = lib.ref_load(path.encode(), d, max_elements)
        if not h:
            raise RuntimeError(lib.ref_last_error().decode())
        obj = cls.__new__(cls)
        obj.lib, obj.d, obj.h = lib, d, h
        obj.K = int(lib.ref_knn_depth(h))
        return obj

    def build_rknng_mutable(self):
        self._check(self.lib.ref_build_rknng_mutable(self.h))

    def insert_dynamic(self, x, labels, m_update=0, bidir=False, nthreads=1):
        x = np.ascontiguousarray(x, np.float32)
        labels = np.ascontiguousarray(labels, np.uint64)
        self._check(self.lib.ref_insert_dynamic(self.h, _p(x), x.shape[0], _p(labels), m_update,
                                                int(bidir), nthreads))

    def get_rknng_mutable(self):
        """rknng_lists_ flattened: (off[n+1] u64, ids u32), each list in its own order."""
        n = self.count
        off = np.zeros(n + 1, np.uint64)
        self._check(self.lib.ref_get_rknng_mutable(self.h, _p(off), None))
        ids = np.empty(int(off[-1]), np.uint32)
        self._check(self.lib.ref_get_rknng_mutable(self.h, _p(off), _p(ids)))
        return off, ids

    def labels(self):
        out = np.empty(self.count, np.uint64)
        self.lib.ref_get_labels(self.h, _p(out))
        return out

    def level0(self, maxM0):
        n = self.count
        links = np.empty((n, maxM0), np.uint32)
        cnt = np.empty(n, np.uint32)
        self.lib.ref_get_level0(self.h, _p(links), _p(cnt), maxM0)
        return links, cnt
