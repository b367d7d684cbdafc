"""Reference statistics of every trial's metric (test infrastructure only).

Restates what _search_region! collects with a MetricStats sink
(search.jl:1768-1894; decim_pass!, search.jl:1664-1719; _block_stats,
search.jl:1339-1360; _hist_push!, search.jl:1428-1442) from the oracle's
all-metrics array (oracle.search_region(..., medcut=-inf, metrics=...)),
with numpy's own min / max / mean / std / median as the statistics, and the
normalised detection (_normalize, search.jl:1638-1643, applied per trial as
in _search_region! and decim_pass!).
"""
import math

import numpy as np

import oracle as O


def all_metrics(params, amps, N, r_lo, total, Nprof):
    """([fold][trial] exact metrics, NaN where a fold has no valid trial; the
    raw candidates of the same pass at params.threshold)."""
    nd = 1 + sum(1 for i in range(params.ndecim) if params.decimations[i] > 1)
    M = np.full((nd, total), np.nan)
    raw = O.search_region(params, amps, N, r_lo, total, Nprof, medcut=-math.inf, metrics=M)
    return M, raw


def logedges(a, b, nwin):
    """_logedges, search.jl:1406-1411"""
    a = max(a, np.finfo(np.float64).tiny)
    b = max(b, a * (1 + 1e-9))
    la, lb = math.log10(a), math.log10(b)
    return np.array([10.0 ** (la + (lb - la) * i / nwin) for i in range(nwin + 1)])


def window(edges, f):
    """_window_index, search.jl:1415-1416 (1-based)"""
    return min(max(int(np.searchsorted(edges, f, side="right")), 1), len(edges) - 1)


def block_stats(M, ks, nharms, N, T, r_lo, r_hi, lodr, total, Nprof, lo, hi, nb, nwin):
    """(records, counts[fold asc k][win][nb], totals[fold asc k][win]).
    Records are dicts keyed like BlockMetricStats (block 0-based), ordered by
    (block, k); a fold with no valid trial in a block has none."""
    order = sorted(range(len(ks)), key=lambda f: ks[f])
    nchunks = -(-total // Nprof)
    edges = {f: logedges(ks[f] * r_lo / T, ks[f] * r_hi / T, nwin) for f in range(len(ks))}
    recs = []
    counts = np.zeros((len(ks), nwin, nb), dtype=np.int64)
    totals = np.zeros((len(ks), nwin), dtype=np.int64)
    invw = nb / (hi - lo)
    for c in range(nchunks):
        i0 = c * Nprof
        n = min(Nprof, total - i0)
        rstart = r_lo + i0 * lodr
        rmean = rstart + (n - 1) * lodr / 2
        for o, f in enumerate(order):
            k = ks[f]
            Hk = nharms // k
            v = M[f, i0:i0 + n]
            v = v[~np.isnan(v)]
            if len(v) == 0:
                continue
            w = window(edges[f], (k * rmean if k > 1 else rmean) / T) - 1
            inr = v[(v >= lo) & (v < hi)]
            b = np.minimum(np.floor((inr - lo) * invw).astype(np.int64), nb - 1)
            np.add.at(counts[o, w], b, 1)
            totals[o, w] += len(v)
            nv = len(v)
            fhi = (rstart + (n - 1) * lodr) / T if k == 1 else k * (rstart + (nv - 1) * lodr) / T
            recs.append(dict(block=c, k=k, Hk=Hk, nbins=2 * Hk, n=nv,
                             ngoodbins=min(N / 2 / (k * rmean if k > 1 else rmean), float(Hk)),
                             flo=(k * rstart if k > 1 else rstart) / T, fhi=fhi,
                             min=float(v.min()), median=float(np.median(v)),
                             mean=float(v.mean()),
                             std=float(v.std(ddof=1)) if nv > 1 else 0.0, max=float(v.max())))
    return recs, counts, totals


def normalized_candidates(M, ks, nharms, T, r_lo, lodr, total, Nprof, edges, loc, scale,
                          threshold):
    """{(t, k): z} of the trials whose z = (M - loc) / scale (window of the
    trial's own searched frequency) exceeds threshold; edges/loc/scale keyed
    by k."""
    out = {}
    t = np.arange(total)
    i0 = (t // Nprof) * Nprof
    rstart = r_lo + i0 * lodr
    rf = rstart + (t - i0) * lodr
    for f, k in enumerate(ks):
        r = rf if k == 1 else k * rf
        fq = r / T
        e = np.asarray(edges[k])
        w = np.clip(np.searchsorted(e, fq, side="right"), 1, len(e) - 1) - 1
        z = (M[f] - np.asarray(loc[k])[w]) / np.asarray(scale[k])[w]
        for tt in np.nonzero(z > threshold)[0]:
            out[(int(tt), k)] = float(z[tt])
    return out
