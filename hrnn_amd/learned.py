"""Learned proxy count m per query (no reference counterpart).

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
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .dev import MLP_MODEL, MLP_PARAMS, Device

GRID = np.array([1, 2, 3, 4, 5, 6, 8, 10, 12, 16, 20, 24, 32, 40, 50, 64], np.uint32)


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
    """|R n GT| / |GT| per query (NaN for empty groundtruth)."""
    nq = len(gt_off) - 1
    out = np.full(nq, np.nan)
    for q in range(nq):
        g = gt_ids[gt_off[q]:gt_off[q + 1]]
        if g.size == 0:
            continue
        r = r_ids[r_off[q]:r_off[q + 1]]
        out[q] = np.intersect1d(g, r).size / g.size
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
        if tau <= 0:
            return self.grid[np.argmax(probs, 1)]
        c = np.argmax(np.cumsum(probs, 1) >= tau, 1)
        return self.grid[c]

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
