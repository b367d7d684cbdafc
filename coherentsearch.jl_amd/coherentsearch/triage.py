"""Block triage: rank a band's search blocks by a learned score and search
only the best ones (no counterpart in the reference search; for callers that
trade recall for time).

Features: 32 incoherent-power statistics per search block, computed on the
device (cs_triage_features_kind; kind 1 sums the power at the nearest bins of
the harmonics of every trial, kind 0 the power-pyramid maxima over the bins
each harmonic of an integer fundamental can fall in).  Scorer: an MLP
32 -> 64 (ReLU) -> 1 logit whose forward and backward passes run on the device
in bf16 with FP32 accumulation; training runs every Adam step on the device as
well (cs_mlp_train).  Labels: whether a block holds a raw candidate of the full
plain search.  scheduled_search visits blocks most-likely first under a
wall-clock budget (every block when it allows, which is the full search).
"""
from __future__ import annotations

import math
import time
from typing import List, Optional

import numpy as np

from . import _lib

NF = _lib.CS_TRIAGE_FEATURES
NH = 64
NPARAMS = _lib.CS_TRIAGE_PARAMS


def block_labels(raw: np.ndarray, nblocks: int) -> np.ndarray:
    """1 for the blocks holding at least one raw candidate, else 0."""
    y = np.zeros(nblocks, dtype=np.float32)
    if len(raw):
        y[np.unique(raw["block"])] = 1.0
    return y


class BlockTriage:
    """The scorer's parameters (W1 [64][32], b1, w2, b2), the feature
    standardisation (mean, inverse std) fitted on the training features and
    the feature kind it was trained on."""

    def __init__(self, seed: int = 0, kind: int = 0):
        self.kind = int(kind)
        rng = np.random.default_rng(seed)
        W1 = rng.standard_normal((NH, NF)) * math.sqrt(2.0 / NF)
        w2 = rng.standard_normal(NH) * math.sqrt(1.0 / NH)
        self.params = np.concatenate([W1.ravel(), np.zeros(NH), w2, [0.0]]).astype(np.float32)
        self.norm = np.concatenate([np.zeros(NF), np.ones(NF)]).astype(np.float32)

    def fit(self, engine, X: np.ndarray, y: np.ndarray, epochs: int = 400, lr: float = 3e-3,
            pos_fraction: float = 0.25, seed: int = 0, on_device: bool = True) -> List[float]:
        """Full-batch Adam on the mean binary cross-entropy; the rare positive
        blocks are repeated until they make up pos_fraction of the rows.  With
        on_device every step runs on the GPU (cs_mlp_train); otherwise the
        device computes the gradients and the FP64 Adam step runs here."""
        X = np.asarray(X, dtype=np.float32)
        y = np.asarray(y, dtype=np.float32)
        mean = X.mean(axis=0)
        std = X.std(axis=0)
        self.norm = np.concatenate([mean, 1.0 / np.where(std > 1e-12, std, 1.0)]).astype(np.float32)
        pos = np.flatnonzero(y > 0.5)
        rows = np.arange(len(y))
        if 0 < len(pos) < pos_fraction * len(y):
            reps = int(math.ceil(pos_fraction * len(y) / len(pos)))
            rows = np.concatenate([rows] + [pos] * (reps - 1))
        Xr, yr = X[rows], y[rows]
        if on_device:
            p, losses = engine.mlp_train(self.params, self.norm, Xr, yr, epochs, lr)
            self.params = p.astype(np.float32)
            return [float(v) for v in losses]
        p = self.params.astype(np.float64)
        m1 = np.zeros_like(p)
        m2 = np.zeros_like(p)
        b1, b2, eps = 0.9, 0.999, 1e-8
        hist = []
        for it in range(1, epochs + 1):
            g, loss = engine.mlp_grad(p.astype(np.float32), self.norm, Xr, yr)
            g = g.astype(np.float64)
            m1 = b1 * m1 + (1 - b1) * g
            m2 = b2 * m2 + (1 - b2) * g * g
            p -= lr * (m1 / (1 - b1 ** it)) / (np.sqrt(m2 / (1 - b2 ** it)) + eps)
            hist.append(loss)
        self.params = p.astype(np.float32)
        return hist

    def logits(self, engine, X: np.ndarray) -> np.ndarray:
        return engine.mlp_forward(self.params, self.norm, X)

    @staticmethod
    def select(logits: np.ndarray, fraction: float) -> np.ndarray:
        """mask of the ceil(fraction * n) highest-scoring blocks"""
        n = len(logits)
        k = max(1, min(n, int(math.ceil(fraction * n))))
        mask = np.zeros(n, dtype=np.uint8)
        mask[np.argsort(-logits, kind="stable")[:k]] = 1
        return mask

    def features(self, engine, r_lo: float, r_hi: float) -> np.ndarray:
        return engine.triage_features(r_lo, r_hi, self.kind)

    def save(self, path: str):
        np.savez(path, params=self.params, norm=self.norm, kind=np.int32(self.kind))

    @classmethod
    def load(cls, path: str) -> "BlockTriage":
        d = np.load(path)
        t = cls(kind=int(d["kind"]) if "kind" in d else 0)
        t.params = d["params"].astype(np.float32)
        t.norm = d["norm"].astype(np.float32)
        return t


def scheduled_search(engine, r_lo: float, r_hi: float, threshold: float, order: np.ndarray,
                     budget_s: Optional[float] = None, first_fraction: float = 0.01,
                     growth: float = 2.0):
    """Visit the band's blocks in the given order (block indices, most likely
    first) in rounds of first_fraction, then growth times more, ... of the
    blocks, until the wall-clock budget (seconds, from the call; None: every
    block) is spent.  The first round always runs; after it, each round is cut
    to the blocks the remaining budget affords at the rate the rounds so far
    took (per-call overheads included), and none starts when not one block
    fits.  Each round is the coherent search of its blocks (cs_search_blocks),
    so every candidate is the full search's own in its block, and visiting
    every block returns the full search's raw candidates.  Returns (raw
    candidates sorted by (t, k), blocks searched, seconds)."""
    t0 = time.perf_counter()
    order = np.asarray(order, dtype=np.int64)
    n = len(order)
    parts = []
    done = 0
    spent = 0.0           # seconds in rounds
    step = max(1, int(math.ceil(first_fraction * n)))
    while done < n:
        if budget_s is not None and done > 0:
            left = budget_s - (time.perf_counter() - t0)
            fit = int(left / (spent / done)) if left > 0 else 0
            if fit < 1:
                break
            step = min(step, fit)
        end = min(n, done + step)
        mask = np.zeros(n, dtype=np.uint8)
        mask[order[done:end]] = 1
        t1 = time.perf_counter()
        parts.append(engine.search_blocks(r_lo, r_hi, threshold, mask))
        spent += time.perf_counter() - t1
        done = end
        step = max(1, int(math.ceil(step * growth)))
    raw = np.concatenate(parts) if parts else np.zeros(0, dtype=_lib.CAND_DTYPE)
    raw = raw[np.lexsort((raw["k"], raw["t"]))] if len(raw) else raw
    return raw, done, time.perf_counter() - t0


def triage_order(engine, model: BlockTriage, r_lo: float, r_hi: float) -> np.ndarray:
    """The band's blocks from the highest score down (stable among ties)."""
    logits = model.logits(engine, model.features(engine, r_lo, r_hi))
    return np.argsort(-logits, kind="stable")


def triage_search(engine, model: BlockTriage, r_lo: float, r_hi: float, threshold: float,
                  fraction: float) -> np.ndarray:
    """Raw candidates of the top `fraction` of the band's blocks by score."""
    X = model.features(engine, r_lo, r_hi)
    mask = model.select(model.logits(engine, X), fraction)
    return engine.search_blocks(r_lo, r_hi, threshold, mask)
