"""Autotuner for the launch shapes of the persistent search kernels.

For a resident index, query batch and proxy set it times each candidate
launch shape of the persistent phase-B/C search kernel ("search_wg_per_cu":
workgroups per CU; "search_lds_kib": KiB of LDS for the per-query
deduplication table, which trades occupancy against how many queries fall
back to the HBM visited bitmap) with HIP events on the library's stream, keeps the
fastest, and caches the choice as JSON keyed by the GPU and the workload
shape (n, d, K, m, nq).  Results are identical for every shape; only the
time changes.
"""
from __future__ import annotations

import ctypes as C
import json
import os

import numpy as np

from . import hiprt
from .dev import lib

CANDIDATES = {
    "search_lds_kib": [8, 16, 32, 48],
    "search_wg_per_cu": [1, 2, 4, 8, 16, 32],
}
CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning", "cache.json")


def _gpu_name() -> str:
    # the HIP runtime's device name: no child process (a child of a process
    # that holds the GPU must not exec another program)
    try:
        return hiprt.device_name(0)
    except Exception:
        return "gpu"


def _key(shape: dict) -> str:
    return _gpu_name() + "|" + ",".join(f"{k}={shape[k]}" for k in sorted(shape))


def load_cache() -> dict:
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def save_cache(c: dict) -> None:
    os.makedirs(os.path.dirname(CACHE), exist_ok=True)
    tmp = CACHE + ".tmp"
    with open(tmp, "w") as f:
        json.dump(c, f, indent=1, sort_keys=True)
    os.replace(tmp, CACHE)


def tune(h, qd, nq: int, m: int, k: int, K_prime: int, shape: dict, pd, cd, idd, cap,
         reps: int = 5, use_cache: bool = True) -> dict:
    """Tune on the device handle h with resident queries qd, resident proxies
    pd (nq x m) and result buffers cd/idd.
    Returns {"search_lds_kib": v, "search_wg_per_cu": v, "timings_ms": {...}}."""
    L = lib()
    key = _key(shape)
    cache = load_cache() if use_cache else {}
    if use_cache and key in cache:
        for k_, v in cache[key].items():
            if k_ in CANDIDATES:
                L.hrnn_dev_set_tuning(h, k_.encode(), int(v))
        return dict(cache[key], cached=True)
    stream = L.hrnn_dev_stream(h)
    e0, e1 = hiprt.Event(), hiprt.Event()

    def ck(rc):
        if rc:
            raise RuntimeError(L.hrnn_dev_last_error(h).decode())

    def time_it(fn):
        fn()
        hiprt.stream_sync(stream)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        e1.sync()
        return e1.ms_since(e0) / reps

    search = lambda: ck(L.hrnn_dev_rknn_search_dptr(h, qd.ptr, nq, pd.ptr, m, None, k, K_prime, 1.0,
                                                    cd.ptr, idd.ptr, cap, None))
    out, timings = {}, {}
    for name, fn in (("search_lds_kib", search), ("search_wg_per_cu", search)):
        best, best_ms = None, float("inf")
        for v in CANDIDATES[name]:
            ck(L.hrnn_dev_set_tuning(h, name.encode(), v))
            ms = time_it(fn)
            timings[f"{name}={v}"] = round(ms, 4)
            if ms < best_ms:
                best, best_ms = v, ms
        ck(L.hrnn_dev_set_tuning(h, name.encode(), best))
        out[name] = best
    out["timings_ms"] = timings
    if use_cache:
        cache[key] = out
        save_cache(cache)
    return out
