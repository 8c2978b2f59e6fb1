"""ctypes binding of libhrnn_host.so: hnswlib::HRNN<float> (include/hrnn_host.h).

Host-side HNSW construction, persistence and dynamic insertion run in C++;
refine / buildRKNNG / search go to the GPU through libhrnn_hip.so.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .dev import HnswBuildStats, HrnnError, InsertStats, NndIterStats, lib as dev_lib

PKG = os.path.dirname(os.path.abspath(__file__))
HOST_LIB_PATH = os.path.join(PKG, "lib", "libhrnn_host.so")

_hl = None


def host_lib() -> C.CDLL:
    global _hl
    if _hl is not None:
        return _hl
    dev_lib()   # loads libhrnn_hip.so first (dependency)
    if not os.path.exists(HOST_LIB_PATH):
        raise FileNotFoundError(f"{HOST_LIB_PATH} is missing: run `python -m hrnn_amd.build`")
    L = C.CDLL(HOST_LIB_PATH)
    vp, u32, u64, i, f = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int, C.c_float
    P = C.POINTER
    sig = {
        "hrnn_index_create": [u64, u32, u32, u32, u64, u32, P(vp)],
        "hrnn_index_load": [C.c_char_p, u32, u64, P(vp)],
        "hrnn_index_free": [vp],
        "hrnn_index_info": [vp, P(u32), P(u32), P(u32), P(i)],
        "hrnn_index_add_points": [vp, vp, u64, vp, i],
        "hrnn_index_add_points_gpu": [vp, vp, u64, vp, i, i, vp],
        "hrnn_index_refine": [vp, i, i, vp],
        "hrnn_index_build_rknng": [vp],
        "hrnn_index_build_rknng_mutable": [vp],
        "hrnn_index_get_knng": [vp, vp, vp, vp],
        "hrnn_index_set_knng": [vp, vp, vp, vp],
        "hrnn_index_get_rknng": [vp, P(u32), P(i), P(u64), vp, vp],
        "hrnn_index_proxies": [vp, vp, u64, u32, u32, i, vp],
        "hrnn_index_proxies_gpu": [vp, vp, u64, u32, u32, vp],
        "hrnn_index_search": [vp, vp, u64, u32, f, u32, u32, u32, i, i, vp, vp, u64, P(u64)],
        "hrnn_index_search_proxies": [vp, vp, u64, vp, u32, vp, u32, u32, f, vp, vp, u64, P(u64)],
        "hrnn_index_insert_dynamic": [vp, vp, u64, vp, u32, i, i],
        "hrnn_index_insert_dynamic_gpu": [vp, vp, u64, vp, u32, i, i, vp],
        "hrnn_index_compact_for_serving": [vp, u32],
        "hrnn_index_get_rknng_mutable": [vp, vp, vp],
        "hrnn_index_save": [vp, C.c_char_p],
        "hrnn_index_device": [vp, P(vp)],
        "hrnn_index_get_level0": [vp, vp, u32, vp],
        "hrnn_index_get_labels": [vp, vp],
        "hrnn_index_export_hnsw": [vp, vp, vp, vp, vp, P(u64), P(u32), P(i)],
        "hrnn_index_query_features": [vp, vp, u64, u32, u32, u32, i, vp, vp, vp],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = C.c_int
    L.hrnn_index_last_error.argtypes = [vp]
    L.hrnn_index_last_error.restype = C.c_char_p
    L.hrnn_index_count.argtypes = [vp]
    L.hrnn_index_count.restype = u64
    _hl = L
    return L


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Index:
    """hnswlib::HRNN<float> with the reference's constructor parameters."""

    def __init__(self, max_elements: int, d: int, M: int = 16, ef_construction: int = 200,
                 random_seed: int = 100, K_knng: int = 0, _handle=None):
        self.L = host_lib()
        self.d = d
        if _handle is None:
            h = C.c_void_p()
            rc = self.L.hrnn_index_create(max_elements, d, M, ef_construction, random_seed, K_knng,
                                          C.byref(h))
            if rc:
                raise HrnnError(rc, self.L.hrnn_index_last_error(None).decode())
            self.h = h
        else:
            self.h = _handle
        M_, efc, K, ml = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_int()
        self.L.hrnn_index_info(self.h, C.byref(M_), C.byref(efc), C.byref(K), C.byref(ml))
        self.M, self.efc, self.K, self.maxlevel = M_.value, efc.value, K.value, ml.value

    @classmethod
    def load(cls, path: str, d: int, max_elements: int = 0) -> "Index":
        L = host_lib()
        h = C.c_void_p()
        rc = L.hrnn_index_load(path.encode(), d, max_elements, C.byref(h))
        if rc:
            raise HrnnError(rc, L.hrnn_index_last_error(None).decode())
        return cls(0, d, _handle=h)

    def close(self):
        if getattr(self, "h", None):
            self.L.hrnn_index_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, rc):
        if rc:
            raise HrnnError(rc, self.L.hrnn_index_last_error(self.h).decode())

    @property
    def count(self) -> int:
        return int(self.L.hrnn_index_count(self.h))

    def add_points(self, x, labels=None, nthreads=1):
        x = np.ascontiguousarray(x, np.float32)
        lab = None if labels is None else np.ascontiguousarray(labels, np.uint64)
        self._ck(self.L.hrnn_index_add_points(self.h, _p(x), x.shape[0], _p(lab), nthreads))

    def add_points_gpu(self, x, labels=None, batch_permille=0, seed_expansions=-1):
        """HNSW construction with KNN seeding on the GPU (empty index only);
        returns the build's counters."""
        x = np.ascontiguousarray(x, np.float32)
        lab = None if labels is None else np.ascontiguousarray(labels, np.uint64)
        st = HnswBuildStats()
        self._ck(self.L.hrnn_index_add_points_gpu(self.h, _p(x), x.shape[0], _p(lab), batch_permille,
                                                  seed_expansions, C.byref(st)))
        return {k: getattr(st, k) for k, _ in HnswBuildStats._fields_}

    def refine(self, iters: int, sample: int):
        st = (NndIterStats * max(iters, 1))()
        self._ck(self.L.hrnn_index_refine(self.h, iters, sample, st))
        return [dict(dists=s.dists, updates=s.updates, proposals=s.proposals, dropped=s.dropped,
                     ms=s.ms, rows=s.rows) for s in st[:iters]]

    def build_rknng(self):
        self._ck(self.L.hrnn_index_build_rknng(self.h))

    def build_rknng_mutable(self):
        self._ck(self.L.hrnn_index_build_rknng_mutable(self.h))

    def get_knng(self):
        n = self.count
        ids = np.empty((n, self.K), np.uint32)
        d = np.empty((n, self.K), np.float32)
        cnt = np.empty(n, np.uint16)
        self._ck(self.L.hrnn_index_get_knng(self.h, _p(ids), _p(d), _p(cnt)))
        return ids, d, cnt

    def set_knng(self, ids, d, cnt):
        ids = np.ascontiguousarray(ids, np.uint32)
        d = np.ascontiguousarray(d, np.float32)
        cnt = np.ascontiguousarray(cnt, np.uint16)
        self._ck(self.L.hrnn_index_set_knng(self.h, _p(ids), _p(d), _p(cnt)))

    def get_rknng(self):
        rb, wide, ne = C.c_uint32(), C.c_int(), C.c_uint64()
        self._ck(self.L.hrnn_index_get_rknng(self.h, C.byref(rb), C.byref(wide), C.byref(ne),
                                             None, None))
        off = np.empty(self.count + 1, np.uint64)
        ent = np.empty(ne.value, np.uint64 if wide.value else np.uint32)
        self._ck(self.L.hrnn_index_get_rknng(self.h, C.byref(rb), C.byref(wide), C.byref(ne),
                                             _p(off), _p(ent)))
        return off, ent, rb.value, wide.value

    def proxies(self, q, m, ef, nthreads=0):
        q = np.ascontiguousarray(q, np.float32)
        out = np.empty((q.shape[0], m), np.uint32)
        self._ck(self.L.hrnn_index_proxies(self.h, _p(q), q.shape[0], m, ef, nthreads, _p(out)))
        return out

    def query_features(self, q, k, ef, m_max=16, nthreads=0, with_raw=False):
        """Host phase A with beam statistics and the 32 learned-m features:
        returns (feats[nq, 32], proxies[nq, m_max]) (+ raw[nq, 20])."""
        q = np.ascontiguousarray(q, np.float32)
        nq = q.shape[0]
        f = np.empty((nq, 32), np.float32)
        p = np.empty((nq, m_max), np.uint32)
        raw = np.empty((nq, 20), np.float32)
        self._ck(self.L.hrnn_index_query_features(self.h, _p(q), nq, k, ef, m_max, nthreads, _p(f), _p(p),
                                                  _p(raw)))
        return (f, p, raw) if with_raw else (f, p)

    def proxies_gpu(self, q, m, ef):
        q = np.ascontiguousarray(q, np.float32)
        out = np.empty((q.shape[0], m), np.uint32)
        self._ck(self.L.hrnn_index_proxies_gpu(self.h, _p(q), q.shape[0], m, ef, _p(out)))
        return out

    def search(self, q, m, threshold, k, K_prime, ef, nthreads=0, per_query=False):
        q = np.ascontiguousarray(q, np.float32)
        nq = q.shape[0]
        off = np.zeros(nq + 1, np.uint64)
        cap = max(1024, nq * 64)
        while True:
            lab = np.empty(cap, np.uint64)
            need = C.c_uint64(0)
            rc = self.L.hrnn_index_search(self.h, _p(q), nq, m, threshold, k, K_prime, ef,
                                          nthreads, int(per_query), _p(off), _p(lab), cap,
                                          C.byref(need))
            if rc == -5:
                cap = int(need.value)
                continue
            self._ck(rc)
            return off, lab[: int(off[-1])].copy()

    def search_proxies(self, q, prox, k, K_prime, threshold=1.0, k_per_q=None):
        q = np.ascontiguousarray(q, np.float32)
        nq = q.shape[0]
        prox = np.ascontiguousarray(prox, np.uint32).reshape(nq, -1)
        kp = None if k_per_q is None else np.ascontiguousarray(k_per_q, np.uint32)
        off = np.zeros(nq + 1, np.uint64)
        cap = max(1024, nq * 64)
        while True:
            ids = np.empty(cap, np.uint32)
            need = C.c_uint64(0)
            rc = self.L.hrnn_index_search_proxies(self.h, _p(q), nq, _p(prox), prox.shape[1],
                                                  _p(kp), k, K_prime, threshold, _p(off), _p(ids),
                                                  cap, C.byref(need))
            if rc == -5:
                cap = int(need.value)
                continue
            self._ck(rc)
            return off, ids[: int(off[-1])].copy()

    def insert_dynamic(self, x, labels, m_update=0, bidirectional=False, nthreads=1):
        x = np.ascontiguousarray(x, np.float32)
        labels = np.ascontiguousarray(labels, np.uint64)
        self._ck(self.L.hrnn_index_insert_dynamic(self.h, _p(x), x.shape[0], _p(labels), m_update,
                                                  int(bidirectional), nthreads))

    def insert_dynamic_gpu(self, x, labels, m_update=0, bidirectional=False, batch_permille=0):
        """insertPointDynamic for all rows of x at once on the GPU (batched);
        returns the call's counters."""
        x = np.ascontiguousarray(x, np.float32)
        labels = np.ascontiguousarray(labels, np.uint64)
        st = InsertStats()
        self._ck(self.L.hrnn_index_insert_dynamic_gpu(self.h, _p(x), x.shape[0], _p(labels), m_update,
                                                      int(bidirectional), batch_permille, C.byref(st)))
        out = {k: getattr(st, k) for k, _ in InsertStats._fields_ if k != "hnsw"}
        out["hnsw"] = {k: getattr(st.hnsw, k) for k, _ in HnswBuildStats._fields_}
        return out

    def get_rknng_mutable(self):
        """The mutable reverse lists: (off[n+1] u64, ids u32), each list in its own order."""
        n = self.count
        off = np.zeros(n + 1, np.uint64)
        self._ck(self.L.hrnn_index_get_rknng_mutable(self.h, _p(off), None))
        ids = np.empty(int(off[-1]), np.uint32)
        self._ck(self.L.hrnn_index_get_rknng_mutable(self.h, _p(off), _p(ids)))
        return off, ids

    def compact_for_serving(self, max_k):
        self._ck(self.L.hrnn_index_compact_for_serving(self.h, max_k))

    def save(self, path):
        self._ck(self.L.hrnn_index_save(self.h, path.encode()))

    def device_handle(self):
        h = C.c_void_p()
        self._ck(self.L.hrnn_index_device(self.h, C.byref(h)))
        return h

    def labels(self):
        out = np.empty(self.count, np.uint64)
        self._ck(self.L.hrnn_index_get_labels(self.h, _p(out)))
        return out

    def level0(self):
        n = self.count
        words = 2 * self.M + 1
        links = np.empty((n, words), np.uint32)
        levels = np.empty(n, np.int32)
        self._ck(self.L.hrnn_index_get_level0(self.h, _p(links), words, _p(levels)))
        return links, levels

    def export_hnsw(self):
        """The HNSW graph in the layout of Device.set_hnsw: a dict with maxM0,
        maxM, l0, levels, up_off, up_pool, entry, maxlevel."""
        pw, ent, ml = C.c_uint64(0), C.c_uint32(0), C.c_int(0)
        self._ck(self.L.hrnn_index_export_hnsw(self.h, None, None, None, None, C.byref(pw), C.byref(ent),
                                               C.byref(ml)))
        n = self.count
        l0 = np.empty((n, 2 * self.M + 1), np.uint32)
        levels = np.empty(n, np.int32)
        up_off = np.empty(n, np.uint32)
        up_pool = np.empty(max(int(pw.value), 1), np.uint32)
        self._ck(self.L.hrnn_index_export_hnsw(self.h, _p(l0), _p(levels), _p(up_off), _p(up_pool),
                                               C.byref(pw), C.byref(ent), C.byref(ml)))
        return dict(maxM0=2 * self.M, maxM=self.M, l0=l0, levels=levels, up_off=up_off,
                    up_pool=up_pool[: int(pw.value)], entry=int(ent.value), maxlevel=int(ml.value))
