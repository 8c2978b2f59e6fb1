"""Autotuner for the launch and tile shapes of the path's kernels.

Each tuner times every candidate of its tuning keys on the workload at hand
(HIP events on the library's stream, or the wall clock around a synchronous
call), keeps the fastest, and caches the choice as JSON keyed by the GPU and
the workload shape.  Every candidate gives identical results; only the time
changes (the GPU tests check this for each key).

* tune (phases B+C, the persistent search kernel): "search_wg_per_cu"
  (workgroups per CU) and "search_lds_kib" (KiB of LDS for the per-query
  deduplication table, which trades occupancy against how many queries fall
  back to the HBM visited bitmap);
* tune_beam (phase A): "beam_vis_shift" (the LDS visited set's size against
  the ef rule: occupancy against probe runs);
* tune_bruteforce (the MFMA brute force): "bf_layout" (waves split each
  tile's rows, or its columns so that a staged tile is read from LDS once per
  block) and "bf_row_frags" (16-row fragments per wave: base-tile reuse
  against blocks per CU);
* tune_nndescent (the NNDescent join): "nnd_join" (chunked BF16, whole-sample
  BF16 or whole-sample int8 staging of each node's sample in LDS), timed on
  the first iteration from the seed graph, which is restored afterwards;
* tune_query (a resident query batch, phase A then phases B+C):
  "query_streams" (the batch on one stream, or split in two parts on two
  streams) and "query_split" (the first part's share, permille: an uneven
  split lets the smaller part's phases B+C run beside the larger part's
  phase A).
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
    "beam_vis_shift": [-1, 0, 1, 2],
    "bf_row_frags": [1, 2],
    "bf_layout": [0, 1],
    "nnd_join": [1, 2, 3],
    "query_streams": [1, 2],
    "query_split": [500, 450, 400, 350, 300],
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


def _cached(key: str, use_cache: bool, L, h):
    """The cached choice for key, applied to handle h (None when absent)."""
    if not use_cache:
        return None, {}
    cache = load_cache()
    if key in cache:
        for k_, v in cache[key].items():
            if k_ in CANDIDATES:
                ck_(L, h, L.hrnn_dev_set_tuning(h, k_.encode(), int(v)))
        return dict(cache[key], cached=True), cache
    return None, cache


def ck_(L, h, rc):
    if rc:
        raise RuntimeError(L.hrnn_dev_last_error(h).decode())


def _event_timer(L, h, reps):
    stream = L.hrnn_dev_stream(h)
    e0, e1 = hiprt.Event(), hiprt.Event()

    def time_it(fn):
        fn()
        hiprt.stream_sync(stream)
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        e1.sync()
        return e1.ms_since(e0) / reps
    return time_it


def _pick(L, h, name, cands, time_it, fn, timings):
    best, best_ms = None, float("inf")
    for v in cands:
        ck_(L, h, L.hrnn_dev_set_tuning(h, name.encode(), int(v)))
        ms = time_it(fn)
        timings[f"{name}={v}"] = round(ms, 4)
        if ms < best_ms:
            best, best_ms = v, ms
    ck_(L, h, L.hrnn_dev_set_tuning(h, name.encode(), int(best)))
    return best


def tune_beam(h, qd, nq: int, m: int, ef: int, pd, shape: dict, reps: int = 5,
              use_cache: bool = True) -> dict:
    """Phase A (hrnn_dev_proxies_dptr at this m and ef over the resident
    queries qd into pd): "beam_vis_shift"."""
    L = lib()
    key = _key(dict(shape, kernel="phase_a", m=m, ef=ef, nq=nq))
    hit, cache = _cached(key, use_cache, L, h)
    if hit:
        return hit
    time_it = _event_timer(L, h, reps)
    fn = lambda: ck_(L, h, L.hrnn_dev_proxies_dptr(h, qd.ptr, nq, m, ef, pd.ptr, None))
    timings = {}
    out = {"beam_vis_shift": _pick(L, h, "beam_vis_shift", CANDIDATES["beam_vis_shift"], time_it, fn, timings),
           "timings_ms": timings}
    if use_cache:
        cache[key] = out
        save_cache(cache)
    return out


def tune_bruteforce(dev, q, k: int, prec: str = "bf16", shape: dict | None = None,
                    reps: int = 2, use_cache: bool = True) -> dict:
    """The MFMA brute force (k+1-NN of the rows q, e.g. a sample of the base,
    over the whole base: the kernel and candidate width of the r_k pass; give
    enough rows to fill the GPU several times over, e.g. 65,536, or the
    shapes' occupancy does not show):
    "bf_layout" (waves split a tile's rows or, for d <= 128, its columns)
    and "bf_row_frags" (16-row fragments per wave in the row split; a second
    one for 32-wide lists, k + 1 <= 16, at d <= 256).  Where neither has a
    second shape the defaults stay and nothing is timed."""
    import time
    L = lib()
    h = dev.h
    shapes = [(0, 1)]
    if dev.d % 32 == 0 and dev.d <= 256 and k + 1 <= 16:
        shapes.append((0, 2))
    if dev.d % 32 == 0 and dev.d <= 128:
        shapes.append((1, 1))
    if len(shapes) == 1:
        return {"bf_layout": 0, "bf_row_frags": 1, "timings_ms": {}, "applies": False}
    key = _key(dict(shape or {}, kernel="bruteforce", n=dev.n, d=dev.d, k=k, prec=prec, rows=len(q)))
    hit, cache = _cached(key, use_cache, L, h)
    if hit:
        return hit
    q = np.ascontiguousarray(q, np.float32)
    timings, best, best_ms = {}, shapes[0], float("inf")
    for lay, fr in shapes:
        ck_(L, h, L.hrnn_dev_set_tuning(h, b"bf_layout", lay))
        ck_(L, h, L.hrnn_dev_set_tuning(h, b"bf_row_frags", fr))
        dev.bf_knn(q, k + 1, prec)
        t0 = time.perf_counter()
        for _ in range(reps):
            dev.bf_knn(q, k + 1, prec)
        ms = (time.perf_counter() - t0) / reps * 1e3
        timings[f"bf_layout={lay},bf_row_frags={fr}"] = round(ms, 4)
        if ms < best_ms:
            best, best_ms = (lay, fr), ms
    ck_(L, h, L.hrnn_dev_set_tuning(h, b"bf_layout", best[0]))
    ck_(L, h, L.hrnn_dev_set_tuning(h, b"bf_row_frags", best[1]))
    out = {"bf_layout": best[0], "bf_row_frags": best[1], "timings_ms": timings, "applies": True}
    if use_cache:
        cache[key] = out
        save_cache(cache)
    return out


def tune_nndescent(dev, sample: int, shape: dict | None = None, use_cache: bool = True) -> dict:
    """The NNDescent join's staging ("nnd_join"), timed on one iteration from
    the device's current graph (the seed graph), after one untimed iteration,
    and put back, flags included, before returning.  int8 staging is a
    candidate for 8-bit data only."""
    import time
    L = lib()
    h = dev.h
    key = _key(dict(shape or {}, kernel="nndescent", n=dev.n, d=dev.d, K=dev.K, S=sample))
    hit, cache = _cached(key, use_cache, L, h)
    if hit:
        return hit
    seed = dev.get_knng()
    flags = dev.get_knng_flags()
    cands = [1, 2, 3] if dev.base_is_u8() else [1, 2]
    timings = {}

    def one():
        dev.set_knng(*seed)
        ck_(L, h, L.hrnn_dev_set_knng_flags(h, None))
        dev.sync()
        t0 = time.perf_counter()
        dev.nndescent(1, sample)
        dev.sync()
        return (time.perf_counter() - t0) * 1e3

    try:
        one()   # first-call costs (the proposal buffer's allocation) outside the comparison
        best = _pick(L, h, "nnd_join", cands, lambda fn: fn(), one, timings)
    finally:
        dev.set_knng(*seed)
        dev.set_knng_flags(flags)
    out = {"nnd_join": best, "timings_ms": timings}
    if use_cache:
        cache[key] = out
        save_cache(cache)
    return out


def tune_query(h, qd, nq: int, m: int, ef: int, k: int, K_prime: int, pd, cd, idd, cap: int, shape: dict,
               reps: int = 5, use_cache: bool = True) -> dict:
    """A resident batch through hrnn_dev_rknn_query_dptr (phase A at this m
    and ef, then phases B+C): "query_streams" and, with two streams,
    "query_split"."""
    L = lib()
    key = _key(dict(shape, kernel="query", m=m, ef=ef, k=k, Kp=K_prime, nq=nq))
    hit, cache = _cached(key, use_cache, L, h)
    if hit:
        return hit
    time_it = _event_timer(L, h, reps)
    fn = lambda: ck_(L, h, L.hrnn_dev_rknn_query_dptr(h, qd.ptr, nq, m, ef, None, k, K_prime, 1.0, pd.ptr,
                                                      cd.ptr, idd.ptr, cap, None))
    shapes = [(1, 500)] + [(2, p) for p in CANDIDATES["query_split"]]
    timings, best, best_ms = {}, shapes[0], float("inf")
    for st, sp in shapes:
        ck_(L, h, L.hrnn_dev_set_tuning(h, b"query_streams", st))
        ck_(L, h, L.hrnn_dev_set_tuning(h, b"query_split", sp))
        ms = time_it(fn)
        timings[f"query_streams={st},query_split={sp}"] = round(ms, 4)
        if ms < best_ms:
            best, best_ms = (st, sp), ms
    ck_(L, h, L.hrnn_dev_set_tuning(h, b"query_streams", best[0]))
    ck_(L, h, L.hrnn_dev_set_tuning(h, b"query_split", best[1]))
    out = {"query_streams": best[0], "query_split": best[1], "timings_ms": timings}
    if use_cache:
        cache[key] = out
        save_cache(cache)
    return out
