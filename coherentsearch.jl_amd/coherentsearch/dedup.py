"""Candidate post-processing: duplicate and harmonic-family collapse.

Host-side mirrors of ``remove_duplicates``, ``_harmonically_related`` and
``remove_harmonics`` (CoherentSearch.jl src/search.jl:2148-2214).  Python's
``sorted`` is stable and ``round`` rounds half to even, as Julia's default
``sortperm``/``sort!`` and ``round(Int, x)`` do.
"""
from __future__ import annotations

from typing import List, Sequence

from .params import Candidate


def remove_duplicates(cands: Sequence[Candidate], dr_tol: float = 1.0) -> List[Candidate]:
    """Group candidates whose consecutive (sorted) r lie within dr_tol; keep the
    maximum-metric member of each group; return sorted by freq (search.jl:2148)."""
    if not cands:
        return []
    order = sorted(range(len(cands)), key=lambda i: cands[i].r)
    kept = []
    best = cands[order[0]]
    prev_r = best.r
    for idx in order[1:]:
        c = cands[idx]
        if c.r - prev_r <= dr_tol:
            if c.metric > best.metric:
                best = c
        else:
            kept.append(best)
            best = c
        prev_r = c.r
    kept.append(best)
    kept.sort(key=lambda c: c.freq)
    return kept


def harmonically_related(r1: float, r2: float, numharm: int, tol: float) -> bool:
    """|m*hi - n*lo| <= tol*m for some 1 <= m, n <= numharm (search.jl:2179)."""
    lo, hi = min(r1, r2), max(r1, r2)
    if not lo > 0:
        return False
    for m in range(1, numharm + 1):
        n = round(m * hi / lo)
        if not (1 <= n <= numharm):
            continue
        if abs(m * hi - n * lo) <= tol * m:
            return True
    return False


def remove_harmonics(cands: Sequence[Candidate], numharm: int = 16,
                     tol: float = 1.0) -> List[Candidate]:
    """Visit strongest first; keep a candidate unless harmonically related to a
    kept one; return sorted by freq (search.jl:2202)."""
    if not cands:
        return []
    order = sorted(range(len(cands)), key=lambda i: -cands[i].metric)
    kept: List[Candidate] = []
    for idx in order:
        c = cands[idx]
        if not any(harmonically_related(c.r, k.r, numharm, tol) for k in kept):
            kept.append(c)
    kept.sort(key=lambda c: c.freq)
    return kept
