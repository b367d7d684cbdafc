"""Search parameters, the decimation ladder, boxcar banks and the Candidate
record — host-side mirrors of CoherentSearch.jl (src/search.jl)."""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Sequence

LADDER_WMAX = 6               # search.jl:401
BOXCAR_SIGMA_SAMPLES = 8192   # search.jl:466


@dataclass
class SearchParams:
    """SearchParams (search.jl:131-144); the defaults match the reference."""
    nharms: int = 32
    m: int = 16
    numbetween: int = 16
    hidr: float = 0.5
    threshold: float = 8.0
    boxcar_fsp: float = 1.5
    boxcar_maxfrac: float = 0.3
    boxcar_medmargin: float = 2.0
    decimations: List[int] = field(default_factory=lambda: [1])
    precision: str = "f64"


def proftype(params: SearchParams) -> str:
    """Profile-stage element type (search.jl:151-155)."""
    if params.precision in ("f64", "f32"):
        return params.precision
    raise ValueError(f"precision must be :f64 or :f32, got :{params.precision}")


def decimation_set(nharms: int, maxdecim: int) -> List[int]:
    """k in 1:max(1, maxdecim) with floor(nharms/k) >= 2 (search.jl:166-167)."""
    return [k for k in range(1, max(1, maxdecim) + 1) if nharms // k >= 2]


def boxcar_widths(nbins: int, fsp: float = 1.5, maxfrac: float = 0.3) -> List[int]:
    """w0=1, w_{k+1}=max(floor(fsp*w_k), w_k+1) up to floor(maxfrac*nbins)
    (search.jl:387-397)."""
    wmax = max(1, math.floor(maxfrac * nbins))
    ws, w = [], 1
    while w <= wmax:
        ws.append(w)
        w = max(math.floor(fsp * w), w + 1)
    return ws


def ladder_boxcar_widths(nbins: int, k: int, params: SearchParams) -> List[int]:
    """The decimation-k bank pruned against the rest of the ladder
    (search.jl:443-459)."""
    ws = boxcar_widths(nbins, params.boxcar_fsp, params.boxcar_maxfrac)
    ks = sorted(params.decimations)
    if len(ks) <= 1:
        return ws
    if not all(ks[i + 1] <= (LADDER_WMAX // 2) * ks[i] for i in range(len(ks) - 1)):
        return ws
    kmin, kmax = ks[0], ks[-1]
    pruned = [w for w in ws if (w >= 2 or k == kmin) and (w <= LADDER_WMAX or k == kmax)]
    return pruned if pruned else ws


@dataclass
class Candidate:
    """Candidate(freq, metric, r, nharm[, ducy]) (search.jl:179-191)."""
    freq: float
    metric: float
    r: float
    nharm: int
    ducy: float = float("nan")


def lodr_of(params: SearchParams) -> float:
    return params.hidr / params.nharms


def trial_grid(T: float, params: SearchParams, lofreq: float, hifreq: float,
               lobin: int, blocksize: int):
    """Trial grid of `search` (search.jl:2015-2027): (r_lo, r_hi, lodr, total,
    Nprof, nchunks), including the reference's `lobin != 100` precedence."""
    lodr = params.hidr / params.nharms
    r_lo = lofreq * T
    if lobin != 100:
        r_lo = float(lobin)
    r_hi = hifreq * T
    total = max(0, math.floor((r_hi - r_lo) / lodr) + 1)
    Nprof = max(1, int(blocksize))
    nchunks = -(-total // Nprof)
    return r_lo, r_hi, lodr, total, Nprof, nchunks


def candidate_r(r_lo: float, lodr: float, Nprof: int, t: int, k: int) -> float:
    """Rebuild a candidate's Fourier frequency from its integer (t, k) with the
    reference's exact FP64 expression order: rstart = r_lo + i0*lodr,
    rf = rstart + (j-1)*lodr, r_dec = k*rf (search.jl:1812, :1833, :1696)."""
    i0 = (t // Nprof) * Nprof
    rstart = r_lo + i0 * lodr
    rf = rstart + (t - i0) * lodr
    return rf if k == 1 else k * rf


def as_int_list(xs: Sequence[int]) -> List[int]:
    return [int(x) for x in xs]
