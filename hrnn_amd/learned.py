"""Learned proxy count m and early stop per query (no reference counterpart).

The reference sweeps a fixed m for all queries (src/hrnn_search.cpp:69).
Here a 32-64-64-16 MLP, trained on the GPU by the kernels of mlp.hip,
predicts per query the smallest m of a 16-value grid that reaches full
recall; plain fixed-m search stays the fallback.

Pipeline:
  1. features: hrnn_index_query_features (host phase-A beam statistics + r_k
     of the closest proxies), 32 floats per query;
  2. labels: for every grid m, phases B+C on the GPU with the first m
     proxies, per-query recall against the exact groundtruth (GPU brute
     force); label = the smallest grid index whose recall reaches `full`
     (1.0 by default), else 15;
  3. training: Adam on mean cross-entropy (hrnn_dev_mlp_train);
  4. inference: MLP forward on the GPU (hrnn_dev_mlp_forward), per-query m
     from `choose_m`, proxies beyond m masked, phases B+C on the GPU.

Early stop (`LearnedEarlyStop`, all on the GPU): the classes are 16 (ef, m)
pairs ordered by cost with ef and m non-decreasing (CLASSES).  A stage-1
beam at ef = EF1 gives every query's features (fused into the beam kernel);
the network picks the cheapest class expected to reach the query's full
recall; only the queries whose class needs ef > EF1 run phase A again, at
that ef (hrnn_dev_rknn_query_adaptive_dptr).  Labels: for every class, phase
A at its ef and phases B+C with its m; label = the cheapest class whose
recall equals the best recall any class reaches for that query.  Plain
fixed-m search (the last class is ef 200) stays the fallback.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .dev import MLP_MODEL, MLP_PARAMS, Device

GRID = np.array([1, 2, 3, 4, 5, 6, 8, 10, 12, 16, 20, 24, 32, 40, 50, 64], np.uint32)


# Stage-1 beam width and the (ef, m) classes of the early stop.
EF1 = 16
CLASSES = np.array([(16, 1), (16, 2), (16, 3), (16, 4), (24, 4), (32, 4), (48, 4), (64, 4),
                    (64, 5), (96, 5), (128, 5), (128, 6), (200, 6), (200, 8), (200, 10), (200, 16)],
                   np.uint32)


# The m-only policy: every class at the stage-1 width, so no query runs a
# second beam; the network picks m alone.
CLASSES_M = np.array([(EF1, m) for m in range(1, 17)], np.uint32)


def init_params(seed: int = 0) -> np.ndarray:
    """He-uniform weights, zero biases, in the HRNN_MLP_PARAMS layout."""
    rng = np.random.default_rng(seed)
    shapes = [(64, 32), (64,), (64, 64), (64,), (16, 64), (16,)]
    parts = []
    for s in shapes:
        if len(s) == 2:
            lim = np.sqrt(6.0 / s[1])
            parts.append(rng.uniform(-lim, lim, s).ravel())
        else:
            parts.append(np.zeros(s))
    p = np.concatenate(parts).astype(np.float32)
    assert p.size == MLP_PARAMS
    return p


def per_query_recall(gt_off, gt_ids, r_off, r_ids) -> np.ndarray:
    """|R n GT| / |GT| per query (NaN for empty groundtruth), vectorised:
    (query, id) pairs as 64-bit keys, membership by np.isin."""
    gt_off = np.asarray(gt_off, np.int64)
    r_off = np.asarray(r_off, np.int64)
    nq = len(gt_off) - 1
    gq = np.repeat(np.arange(nq, dtype=np.uint64), np.diff(gt_off))
    rq = np.repeat(np.arange(nq, dtype=np.uint64), np.diff(r_off))
    gk = (gq << np.uint64(32)) | np.asarray(gt_ids, np.uint64)[:len(gq)]
    rk = (rq << np.uint64(32)) | np.asarray(r_ids, np.uint64)[:len(rq)]
    hit = np.isin(gk, rk).astype(np.float64)
    num = np.bincount(gq.astype(np.int64), weights=hit, minlength=nq)
    den = np.diff(gt_off).astype(np.float64)
    out = np.full(nq, np.nan)
    nz = den > 0
    out[nz] = num[nz] / den[nz]
    return out


def mask_proxies(prox: np.ndarray, m_q: np.ndarray) -> np.ndarray:
    out = prox.copy()
    cols = np.arange(prox.shape[1])[None, :]
    out[cols >= np.asarray(m_q)[:, None]] = 0xFFFFFFFF
    return out


def label_queries(dev: Device, q, prox, gt_off, gt_ids, k, K_prime, grid=GRID, full=1.0,
                  id_map=None):
    """Smallest grid index reaching recall >= full per query (15 when none does).
    prox: [nq, >= grid.max()] proxies from the beam; id_map maps internal ids to labels."""
    nq = q.shape[0]
    rec = np.zeros((len(grid), nq))
    for gi, m in enumerate(grid):
        off, ids = dev.rknn_search(q, np.ascontiguousarray(prox[:, :m]), k, K_prime)
        lab = ids if id_map is None else id_map[ids]
        rec[gi] = per_query_recall(gt_off, gt_ids, off, lab.astype(np.uint32))
    rec = np.nan_to_num(rec, nan=1.0)
    ok = rec >= full - 1e-12
    labels = np.where(ok.any(0), ok.argmax(0), len(grid) - 1).astype(np.uint32)
    return labels, rec


def choose_class(probs, tau: float) -> np.ndarray:
    """The first class whose cumulative probability reaches tau (argmax when
    tau <= 0) -- the rule of the GPU pick kernels."""
    probs = np.asarray(probs)
    if tau <= 0:
        return np.argmax(probs, 1)
    cs = np.cumsum(probs, 1) >= np.float32(tau)
    return np.where(cs.any(1), cs.argmax(1), probs.shape[1] - 1)


def label_classes(dev: Device, q, gt_off, gt_ids, k, K_prime, classes=CLASSES, id_map=None,
                  ref_recall=None):
    """Per query the cheapest class reaching the best recall of any class, or,
    with ref_recall[nq] (e.g. plain fixed-m search's per-query recall), the
    cheapest class reaching that recall (the best class where none does).
    Returns (labels[nq], recall[16, nq], expansions[16, nq]); recall is NaN
    for a query with empty groundtruth (left out of macro recall, as the
    reference's compute_rknn_recall does), whose label is class 0."""
    q = np.ascontiguousarray(q, np.float32)
    nq = q.shape[0]
    classes = np.asarray(classes, np.uint32)
    mmax = int(classes[:, 1].max())
    rec = np.zeros((len(classes), nq))
    exp = np.zeros((len(classes), nq))
    for ef in np.unique(classes[:, 0]):
        _, prox, ex = dev.stage1_features(q, int(ef), mmax, k)
        for c in np.nonzero(classes[:, 0] == ef)[0]:
            off, ids = dev.rknn_search(q, np.ascontiguousarray(prox[:, :classes[c, 1]]), k, K_prime)
            lab = ids if id_map is None else id_map[ids]
            rec[c] = per_query_recall(gt_off, gt_ids, off, lab.astype(np.uint32))
            exp[c] = ex[:, 0]
    r1 = np.nan_to_num(rec, nan=1.0)
    best = r1.max(0)
    if ref_recall is not None:
        best = np.minimum(best, np.nan_to_num(np.asarray(ref_recall, np.float64), nan=1.0))
    labels = (r1 >= best[None, :] - 1e-12).argmax(0).astype(np.uint32)
    return labels, rec, exp


@dataclass
class LearnedM:
    params: np.ndarray          # HRNN_MLP_PARAMS floats
    mu: np.ndarray              # 32
    sd: np.ndarray              # 32
    grid: np.ndarray = GRID

    def model(self) -> np.ndarray:
        m = np.concatenate([self.params, self.mu, self.sd]).astype(np.float32)
        assert m.size == MLP_MODEL
        return m

    def probs(self, dev: Device, feats) -> np.ndarray:
        return dev.mlp_forward(self.params, (np.asarray(feats) - self.mu) / self.sd)

    def choose_m(self, probs, tau: float) -> np.ndarray:
        return self.grid[choose_class(probs, tau)]

    @classmethod
    def train(cls, dev: Device, feats, labels, epochs=40, batch=256, lr=2e-3, weight_decay=1e-5,
              seed=0):
        feats = np.asarray(feats, np.float32)
        mu = feats.mean(0)
        sd = feats.std(0) + 1e-6
        x = (feats - mu) / sd
        p, losses = dev.mlp_train(init_params(seed), x, labels, epochs, batch, lr, weight_decay,
                                  seed)
        return cls(p, mu.astype(np.float32), sd.astype(np.float32)), losses


@dataclass
class LearnedEarlyStop:
    """The MLP over CLASSES: model() is the HRNN_MLP_MODEL blob the adaptive
    GPU path takes (parameters, then the features' mean and std)."""
    params: np.ndarray
    mu: np.ndarray
    sd: np.ndarray
    classes: np.ndarray = CLASSES
    ef1: int = EF1

    def model(self) -> np.ndarray:
        m = np.concatenate([self.params, self.mu, self.sd]).astype(np.float32)
        assert m.size == MLP_MODEL
        return m

    def probs(self, dev: Device, feats) -> np.ndarray:
        return dev.mlp_forward(self.params, (np.asarray(feats, np.float32) - self.mu) / self.sd)

    def choose(self, probs, tau: float) -> np.ndarray:
        return choose_class(probs, tau)

    def query(self, dev: Device, q, k, K_prime, tau, threshold=1.0):
        """Two-stage query on the GPU; returns (res_off, res_ids, cls, proxies)."""
        return dev.query_adaptive(q, self.classes, self.ef1, int(self.classes[:, 1].max()), k, K_prime,
                                  model=self.model(), tau=tau, threshold=threshold)

    @classmethod
    def train(cls, dev: Device, feats, labels, epochs=60, batch=256, lr=2e-3, weight_decay=1e-5,
              seed=0, classes=CLASSES, ef1=EF1):
        feats = np.asarray(feats, np.float32)
        mu = feats.mean(0)
        sd = feats.std(0) + 1e-6
        p, losses = dev.mlp_train(init_params(seed), (feats - mu) / sd, labels, epochs, batch, lr,
                                  weight_decay, seed)
        return cls(p, mu.astype(np.float32), sd.astype(np.float32), np.asarray(classes, np.uint32),
                   ef1), losses

    def pick_tau(self, dev: Device, feats, rec, target: float,
                 taus=(0.1, 0.2, 0.3, 0.4, 0.5, 0.7, 0.8, 0.9, 0.95, 0.98, 0.995)):
        """The smallest tau whose macro recall over these queries (rec[16, nq]
        from label_classes, NaN rows of empty groundtruth left out) reaches
        target; the last tau when none does.  A smaller tau stops at an
        earlier (cheaper) class, so the first tau that passes is the
        cheapest one that does."""
        p = self.probs(dev, feats)
        cols = np.arange(p.shape[0])
        for tau in taus:
            if np.nanmean(rec[self.choose(p, tau), cols]) >= target:
                return tau
        return taus[-1]
