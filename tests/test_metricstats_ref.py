"""CPU checks of the metric-statistics host code (coherentsearch/metricstats.py)
and of the oracle's every-trial metrics (oracle/metricstats_ref.py).

Pins:
* collecting statistics does not change the candidates (test_search.jl:
  591-595): the oracle's exact-everywhere pass (medcut = -inf) returns the
  gated search's candidates, and its metrics above threshold are exactly them;
* hist_quantile on a fine histogram of known samples lands within a bin
  width of numpy's quantile; empty, underflow and overflow cases;
* build_metricnorm's fallbacks (sparse window, degenerate scale) and
  window_index's clamping; normalize is monotone in M (test_search.jl:653-680).
"""
import math

import numpy as np
import pytest

import metricstats_ref as R
import oracle as O

from coherentsearch import metricstats as MS
from coherentsearch.synth import Pulsar, make_spectrum


def _hist(values, lo=0.0, hi=10.0, nb=1000, k=1, win=1):
    h = MS.MetricHistogram(k, 8, 16, win, 1.0, 2.0, lo, hi, nb / (hi - lo),
                           np.zeros(nb, dtype=np.int64))
    for x in values:
        h.total += 1
        h.sum += x
        h.sumsq += x * x
        h.vmin, h.vmax = min(h.vmin, x), max(h.vmax, x)
        if x < lo:
            h.under += 1
        elif x >= hi:
            h.over += 1
        else:
            h.counts[int(math.floor((x - lo) * h.invw))] += 1
    return h


def test_hist_quantile_against_numpy():
    rng = np.random.default_rng(1)
    v = rng.gamma(3.0, 1.0, 20000)
    h = _hist(v)
    for q in (0.01, 0.1, 0.5, 0.8413, 0.95):   # quantiles below hi: within a bin width
        assert abs(MS.hist_quantile(h, q) - np.quantile(v, q)) <= 0.011, q
    assert h.over > 0 and MS.hist_quantile(h, 1.0 - h.over / (2 * h.total)) == 10.0
    assert math.isnan(MS.hist_quantile(_hist([]), 0.5))
    assert MS.hist_quantile(_hist([-1.0, -2.0, 5.0]), 0.5) == 0.0        # in the underflow
    assert MS.hist_quantile(_hist([11.0, 12.0, 1.0]), 0.9) == 10.0       # in the overflow


def test_metricnorm_fallbacks_and_windows():
    rng = np.random.default_rng(2)
    big = rng.normal(5.0, 1.0, 5000)
    ms = MS.MetricStats(hist_lo=0.0, hist_hi=10.0, hist_nb=1000, nwin=3)
    w1 = _hist(big, win=1)
    w2 = _hist(big[:50] + 2.0, win=2)                  # < 200 trials: falls back
    w3 = _hist(np.full(300, 12.0), win=3)              # all in the overflow: scale 0
    for h, (a, b) in zip((w1, w2, w3), ((1.0, 2.0), (2.0, 4.0), (4.0, 8.0))):
        h.flo, h.fhi = a, b
    ms.whists = [w1, w2, w3]
    ms.hists = [MS.merge_hists([w1, w2, w3], 0, 1.0, 8.0)]
    norm = MS.build_metricnorm(ms)
    gl, gs = MS._loc_scale(ms.hists[0])
    assert np.array_equal(norm.edges[1], [1.0, 2.0, 4.0, 8.0])
    assert norm.loc[1][0] == MS.hist_quantile(w1, 0.5) and abs(norm.scale[1][0] - 1.0) < 0.05
    assert (norm.loc[1][1], norm.scale[1][1]) == (gl, gs)
    assert (norm.loc[1][2], norm.scale[1][2]) == (gl, gs)
    assert MS.window_index(norm.edges[1], 0.5) == 1 and MS.window_index(norm.edges[1], 99.0) == 3
    assert MS.window_index(norm.edges[1], 2.0) == 2
    assert MS.normalize(norm, 1, 1.5, 200.0) > MS.normalize(norm, 1, 1.5, 100.0)
    assert MS.normalize(norm, 5, 1.5, 3.0) == 3.0     # no model for k = 5


@pytest.fixture(scope="module")
def small_search():
    N, dt = 1 << 18, 1e-3
    T = N * dt
    amps = make_spectrum(N, dt, 77, [Pulsar(9.7, 0.05, 25.0)])
    p = O.make_params(nharms=16, m=16, numbetween=16, threshold=7.0, decimations=(1, 2, 3))
    lodr = 0.5 / 16
    r_lo, r_hi = 5.0 * T, 14.0 * T
    total = int(math.floor((r_hi - r_lo) / lodr)) + 1
    return amps, N, T, p, r_lo, r_hi, lodr, total, 1024


def test_collecting_does_not_change_candidates(small_search):
    amps, N, T, p, r_lo, r_hi, lodr, total, Nprof = small_search
    M, raw = R.all_metrics(p, amps, N, r_lo, total, Nprof)
    gated = O.search_region(p, amps, N, r_lo, total, Nprof)
    assert len(raw) > 0 and np.array_equal(raw, gated)
    ks = [1, 2, 3]
    above = sorted((int(t), ks[f]) for f in range(3) for t in np.nonzero(M[f] > 7.0)[0])
    assert above == sorted((int(r["t"]), int(r["k"])) for r in raw)
    # k > 1 leaves trials at or past Nyquist unscored; k = 1 scores every trial
    assert not np.isnan(M[0]).any()


def test_reference_block_invariants(small_search):
    amps, N, T, p, r_lo, r_hi, lodr, total, Nprof = small_search
    M, _ = R.all_metrics(p, amps, N, r_lo, total, Nprof)
    recs, counts, totals = R.block_stats(M, [1, 2, 3], 16, N, T, r_lo, r_hi, lodr, total, Nprof,
                                         0.0, 64.0, 3200, 4)
    assert {r["k"] for r in recs} == {1, 2, 3}
    assert all(r["nbins"] == 2 * r["Hk"] and r["Hk"] == 16 // r["k"] for r in recs)
    assert all(r["min"] <= r["median"] <= r["max"] and r["min"] <= r["mean"] <= r["max"]
               for r in recs)
    for o, k in enumerate([1, 2, 3]):
        assert totals[o].sum() == sum(r["n"] for r in recs if r["k"] == k)
        assert counts[o].sum() <= totals[o].sum()
