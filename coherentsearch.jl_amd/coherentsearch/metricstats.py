"""Metric statistics (--metricstats) and the in-situ normalisation
(--normalize): src/search.jl:1295-1645.

The device computes every trial's exact metric, the per-(block, k) records
and the per-(k, window) histograms (cs_search_stats, cs_stats_blocks,
cs_stats_hists) and runs the normalised detection pass over the resident
metrics (cs_search_normalized).  The summaries, quantiles and the
normalisation model are host code, as in the reference.
"""
import math
from dataclasses import dataclass, field
from typing import Dict, List, NamedTuple, Optional, Tuple

import numpy as np

MIN_WIN_TRIALS = 200   # _MIN_WIN_TRIALS, search.jl:1565


class BlockMetricStats(NamedTuple):
    """BlockMetricStats (search.jl:1321-1335), immutable as there; `block`
    is 1-based as there."""
    block: int
    k: int
    Hk: int
    nbins: int
    ngoodbins: float
    flo: float
    fhi: float
    n: int
    min: float
    median: float
    mean: float
    std: float
    max: float


@dataclass
class MetricHistogram:
    """MetricHistogram (search.jl:1376-1394): counts[i] covers
    [lo + i w, lo + (i + 1) w), w = (hi - lo) / len(counts); win 0 = merged."""
    k: int
    Hk: int
    nbins: int
    win: int
    flo: float
    fhi: float
    lo: float
    hi: float
    invw: float
    counts: np.ndarray
    under: int = 0
    over: int = 0
    total: int = 0
    sum: float = 0.0
    sumsq: float = 0.0
    vmin: float = math.inf
    vmax: float = -math.inf


@dataclass
class MetricStats:
    """MetricStats (search.jl:1504-1515): the sink passed to `search`."""
    hist_lo: float = 0.0
    hist_hi: float = 64.0
    hist_nb: int = 3200
    nwin: int = 16
    blocks: List[BlockMetricStats] = field(default_factory=list)
    hists: List[MetricHistogram] = field(default_factory=list)
    whists: List[MetricHistogram] = field(default_factory=list)

    def __post_init__(self):
        self.nwin = max(1, int(self.nwin))


def merge_hists(hs: List[MetricHistogram], win: int, flo: float, fhi: float) -> MetricHistogram:
    """_merge_hists (search.jl:1419-1426): the sum of histograms of one k."""
    h1 = hs[0]
    g = MetricHistogram(h1.k, h1.Hk, h1.nbins, win, flo, fhi, h1.lo, h1.hi, h1.invw,
                        np.zeros_like(h1.counts))
    for h in hs:
        g.counts = g.counts + h.counts
        g.under += h.under
        g.over += h.over
        g.total += h.total
        g.sum += h.sum
        g.sumsq += h.sumsq
        g.vmin = min(g.vmin, h.vmin)
        g.vmax = max(g.vmax, h.vmax)
    return g


def hist_quantile(h: MetricHistogram, q: float) -> float:
    """hist_quantile (search.jl:1463-1478): the q-th quantile, linear within
    the containing bin; h.hi when it falls in the overflow."""
    if h.total == 0:
        return math.nan
    target = q * h.total
    cum = h.under
    if cum >= target:
        return h.lo
    w = 1.0 / h.invw
    for i, c in enumerate(h.counts):
        c = int(c)
        if cum + c >= target:
            edge = h.lo + i * w
            return edge if c == 0 else edge + (target - cum) / c * w
        cum += c
    return h.hi


def _hist_row(h: MetricHistogram, faps) -> dict:
    """_hist_row (search.jl:1531-1540)."""
    N = h.total
    mean = h.sum / N if N > 0 else math.nan
    std = math.sqrt(max(0.0, (h.sumsq - N * mean ** 2) / (N - 1))) if N > 1 else 0.0
    return dict(k=h.k, Hk=h.Hk, nbins=h.nbins, win=h.win, flo=h.flo, fhi=h.fhi, ntrials=N,
                min=h.vmin if N > 0 else math.nan, median=hist_quantile(h, 0.5), mean=mean,
                std=std, max=h.vmax if N > 0 else math.nan,
                fap=tuple(hist_quantile(h, 1 - p) for p in faps),
                overflow=h.over / N if N > 0 else 0.0)


def metricstats_summary(ms: MetricStats, faps=(0.1, 0.01, 1e-3, 1e-4)) -> List[dict]:
    """metricstats_summary (search.jl:1542-1546): one row per k."""
    return [dict(_hist_row(h, faps), nblocks=sum(1 for s in ms.blocks if s.k == h.k and s.n > 0))
            for h in ms.hists]


def metricstats_windows(ms: MetricStats, faps=(0.1, 0.01, 1e-3, 1e-4)) -> List[dict]:
    """metricstats_windows (search.jl:1558-1559): nonempty (k, window) rows."""
    return [_hist_row(h, faps) for h in ms.whists if h.total > 0]


@dataclass
class MetricNorm:
    """MetricNorm (search.jl:1590-1594): per-k window edges, loc, scale."""
    edges: Dict[int, np.ndarray]
    loc: Dict[int, np.ndarray]
    scale: Dict[int, np.ndarray]


def _loc_scale(h: MetricHistogram) -> Tuple[float, float]:
    """_loc_scale (search.jl:1597-1601)."""
    loc = hist_quantile(h, 0.5)
    scale = hist_quantile(h, 0.8413) - loc
    return loc, (scale if scale > 0 else math.nan)


def build_metricnorm(ms: MetricStats) -> MetricNorm:
    """build_metricnorm (search.jl:1610-1635): window loc/scale, the k-global
    estimate where a window has < 200 trials or a degenerate scale."""
    edges, locs, scales = {}, {}, {}
    for g in ms.hists:
        k = g.k
        gloc, gscale = _loc_scale(g)
        if math.isnan(gscale):
            gscale = 1.0
        wins = sorted([h for h in ms.whists if h.k == k], key=lambda h: h.win)
        if not wins:
            continue
        edges[k] = np.array([h.flo for h in wins] + [wins[-1].fhi])
        lv, sv = np.empty(len(wins)), np.empty(len(wins))
        for i, h in enumerate(wins):
            wloc, wscale = _loc_scale(h)
            if h.total < MIN_WIN_TRIALS or math.isnan(wscale):
                lv[i], sv[i] = gloc, gscale
            else:
                lv[i], sv[i] = wloc, wscale
        locs[k], scales[k] = lv, sv
    return MetricNorm(edges, locs, scales)


def window_index(edges: np.ndarray, f: float) -> int:
    """_window_index (search.jl:1415-1416), 1-based, clamped to [1, nwin]."""
    return int(min(max(int(np.searchsorted(edges, f, side="right")), 1), len(edges) - 1))


def normalize(norm: MetricNorm, k: int, f: float, M: float) -> float:
    """_normalize (search.jl:1638-1643): (M - loc) / scale of f's window."""
    e = norm.edges.get(k)
    if e is None:
        return M
    w = window_index(e, f) - 1
    return (M - norm.loc[k][w]) / norm.scale[k][w]


def fill_from_device(ms: MetricStats, blocks: np.ndarray, counts: np.ndarray,
                     acc: np.ndarray, Hk_of: Dict[int, int]) -> None:
    """Append the device's records (cs_stats_blocks / cs_stats_hists, folds in
    ascending k) to `ms`, as _search_region! does after its parallel region
    (search.jl:1872-1890)."""
    cols = [blocks[f].tolist() for f in BlockMetricStats._fields]
    cols[0] = [b + 1 for b in cols[0]]
    ms.blocks.extend(map(BlockMetricStats._make, zip(*cols)))
    nf, nwin, nb = counts.shape
    invw = nb / (ms.hist_hi - ms.hist_lo)
    new_h, new_w = [], []
    for f in range(nf):
        kwins = []
        for w in range(nwin):
            a = acc[f, w]
            k = int(a["k"])
            h = MetricHistogram(k, Hk_of[k], 2 * Hk_of[k], int(a["win"]), float(a["flo"]),
                                float(a["fhi"]), ms.hist_lo, ms.hist_hi, invw,
                                counts[f, w].astype(np.int64), int(a["under"]), int(a["over"]),
                                int(a["total"]), float(a["sum"]), float(a["sumsq"]),
                                float(a["vmin"]), float(a["vmax"]))
            kwins.append(h)
        new_w.extend(kwins)
        new_h.append(merge_hists(kwins, 0, kwins[0].flo, kwins[-1].fhi))
    ms.whists.extend(new_w)
    ms.hists.extend(new_h)
    ms.hists.sort(key=lambda h: h.k)
    ms.whists.sort(key=lambda h: (h.k, h.win))
