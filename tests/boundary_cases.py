"""Queries placed within FP32 rounding of a base row's radius r_k, where the
FP32 decision d2(q, p) <= r_k(p) (the reference's 16-lane order against the
FP32 r_k) and the FP64 one (sequential FP64 distance against the FP64 r_k)
differ.  Built with the oracle only; used by the oracle's pins and by the
GPU parity tests of the FP64 re-check."""
import numpy as np

import oracle


def boundary_queries(x, k, kth32, kth64, rows, want=8, tries=60, seed=0):
    """For each base row p in `rows`: nb = a row whose FP64 distance to p is
    exactly r_k(p) in FP64 (p's (k+1)-th neighbour, self included), then
    q = x_nb with a random quarter of its coordinates moved by one FP32 ulp
    up or down.  That moves the distance to p by about 1e-8 (relative),
    below the FP32 rounding of the distance and of r_k (about 1e-7), so the
    two decisions often differ.  A point is kept when they do and its FP64
    distance is at least 1e-12 (relative) away from r_k(p), so that any FP64
    summation order decides it the same way; half of the points lie inside
    the FP64 radius (FP32 says outside), half outside.  Returns (q [m x d]
    float32, p [m] uint32, in64 [m] bool: the FP64 decision)."""
    x = np.ascontiguousarray(x, np.float32)
    xd = x.astype(np.float64)
    rng = np.random.default_rng(seed)
    qs, ps, dec = [], [], []
    need = {True: want // 2, False: want - want // 2}
    for p in rows:
        if need[True] == 0 and need[False] == 0:
            break
        dd = ((xd - xd[p]) ** 2).sum(1)
        order = np.argsort(dd, kind="stable")
        nb = None
        for j in order[max(0, k - 3):k + 4]:
            if oracle.l2sqr_f64(x[p], x[j]) == kth64[p]:
                nb = int(j)
                break
        if nb is None or nb == p:
            continue
        for _ in range(tries):
            q = x[nb].copy()
            sel = rng.random(x.shape[1]) < 0.25
            up = rng.random(x.shape[1]) < 0.5
            q[sel & up] = np.nextafter(q[sel & up], np.float32(np.inf))
            q[sel & ~up] = np.nextafter(q[sel & ~up], np.float32(-np.inf))
            d64 = oracle.l2sqr_f64(q, x[p])
            if abs(d64 - kth64[p]) <= 1e-12 * kth64[p]:
                continue
            in64 = bool(d64 <= kth64[p])
            if need[in64] == 0:
                continue
            in32 = bool(oracle.l2sqr(q, x[p]) <= kth32[p])
            if in64 != in32:
                qs.append(q)
                ps.append(p)
                dec.append(in64)
                need[in64] -= 1
                break
    if not qs:
        return np.zeros((0, x.shape[1]), np.float32), np.zeros(0, np.uint32), np.zeros(0, bool)
    return np.stack(qs), np.array(ps, np.uint32), np.array(dec)
