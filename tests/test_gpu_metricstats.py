"""GPU checks of the metric statistics (cs_search_stats, cs_stats_blocks,
cs_stats_hists, cs_stats_metrics) and the normalised pass
(cs_search_normalized).

Bars:
* the statistics search returns the candidates of this mode's search with the
  gate off (boxcar_medmargin = inf), exactly: the reference disables the gate
  when it collects statistics (search.jl:1779-1780);
* every-trial metrics within 1e-4 relative of the oracle's (the search's own
  bar), NaN exactly where the oracle scores no trial;
* block records and histograms exact against numpy statistics of the
  device's own metrics (counts and n equal; moments and medians to 1e-12;
  frequencies and ngoodbins equal), so the statistics kernels are pinned
  independently of the metric rounding;
* the reference's own invariants (test_search.jl:584-651);
* the normalised candidates equal the oracle's normalisation of the device
  metrics with the same model (z to 1e-12), and still detect the pulsar
  (test_search.jl:653-680).
"""
import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def setup():
    import coherentsearch as cs
    from coherentsearch.synth import Pulsar, make_spectrum
    N, dt = 1 << 20, 2.5e-4
    T = N * dt
    amps = make_spectrum(N, dt, 4711, [Pulsar(10.0123, 0.05, 30.0), Pulsar(37.3, 0.08, 18.0)])
    return cs, amps, N, dt, T


def _grid(cs, p, T, lofreq, hifreq, blocksize):
    return cs.trial_grid(T, p, lofreq, hifreq, 100, blocksize)


@pytest.mark.parametrize("mode,nh", [("plain", 32), ("fused", 32), ("plain", 18)])
def test_stats_search_against_oracle(setup, mode, nh):
    """nh 18 folds 36, 18 and 12 bins, which the per-trial exact scorer does
    not cover: the gate-free rescan writes the metrics instead."""
    import metricstats_ref as R
    import oracle as O
    cs, amps, N, dt, T = setup
    p = cs.SearchParams(nharms=nh, m=16, decimations=cs.decimation_set(nh, 3), threshold=8.0)
    r_lo, r_hi, lodr, total, Nprof, nchunks = _grid(cs, p, T, 6.0, 60.0, 1024)
    eng = cs.Engine(p, Nprof)
    eng.set_mode(mode)
    eng.set_spectrum(cs.FFTFile.from_arrays(amps, N, dt))
    _say(f"stats search mode={mode} nharms={nh} trials={total}")
    pe = cs.SearchParams(nharms=nh, m=16, decimations=cs.decimation_set(nh, 3), threshold=8.0,
                         boxcar_medmargin=math.inf)
    ee = cs.Engine(pe, Nprof)
    ee.set_mode(mode)
    ee.set_spectrum(cs.FFTFile.from_arrays(amps, N, dt))
    exact = ee.search_raw(r_lo, r_hi, 8.0)
    ms = cs.MetricStats(hist_lo=0.0, hist_hi=64.0, hist_nb=3200, nwin=8)
    raw = eng.search_stats(r_lo, r_hi, 8.0, ms)
    assert len(raw) > 0 and np.array_equal(raw, exact)
    ks = sorted({h.k for h in ms.hists})
    assert ks == [1, 2, 3]
    # every-trial metrics against the oracle
    op = O.make_params(nharms=nh, m=16, numbetween=16, threshold=8.0,
                       decimations=list(p.decimations))
    Mo, _ = R.all_metrics(op, amps, N, r_lo, total, Nprof)
    order = [1] + [k for k in p.decimations if k > 1]
    Md = np.stack([eng.stats_metrics(k, total) for k in order])
    assert np.array_equal(np.isnan(Md), np.isnan(Mo))
    ok = ~np.isnan(Mo)
    rel = np.abs(Md[ok] - Mo[ok]) / np.maximum(np.abs(Mo[ok]), 1.0)
    assert rel.max() <= 1e-4, rel.max()
    # block records and histograms against numpy statistics of the device metrics
    recs, counts, totals = R.block_stats(Md, order, nh, N, T, r_lo, r_hi, lodr, total, Nprof,
                                         ms.hist_lo, ms.hist_hi, ms.hist_nb, ms.nwin)
    assert len(recs) == len(ms.blocks)
    for a, b in zip(ms.blocks, recs):
        assert (a.block - 1, a.k, a.Hk, a.nbins, a.n) == (b["block"], b["k"], b["Hk"],
                                                          b["nbins"], b["n"])
        assert (a.flo, a.fhi, a.ngoodbins) == (b["flo"], b["fhi"], b["ngoodbins"])
        for key in ("min", "max", "median"):
            assert getattr(a, key) == b[key], key
        for key in ("mean", "std"):
            assert abs(getattr(a, key) - b[key]) <= 1e-12 * max(1.0, abs(b[key])), key
    wh = sorted(ms.whists, key=lambda h: (h.k, h.win))
    for o, k in enumerate(ks):
        for w in range(ms.nwin):
            h = wh[o * ms.nwin + w]
            assert (h.k, h.win) == (k, w + 1)
            assert np.array_equal(h.counts, counts[o, w]), (k, w)
            assert h.total == totals[o, w]
            assert int(h.counts.sum()) + h.under + h.over == h.total


def test_reference_invariants_and_normalize(setup):
    import metricstats_ref as R
    cs, amps, N, dt, T = setup
    nh = 60
    p = cs.SearchParams(nharms=nh, m=16, decimations=cs.decimation_set(nh, 3))
    r_lo, r_hi, lodr, total, Nprof, nchunks = _grid(cs, p, T, 8.0, 12.0, 1024)
    eng = cs.Engine(p, Nprof)
    eng.set_mode("fused")
    ft = cs.FFTFile.from_arrays(amps, N, dt)
    eng.set_spectrum(ft)
    _say("metricstats invariants")
    ms = cs.MetricStats(hist_hi=30000.0, hist_nb=6000)
    eng.search_stats(r_lo, r_hi, 8.0, ms)
    assert {s.k for s in ms.blocks} == {1, 2, 3}
    assert all(s.nbins == 2 * s.Hk and s.Hk == nh // s.k for s in ms.blocks)
    assert all(s.min <= s.median <= s.max and s.min <= s.mean <= s.max for s in ms.blocks)
    assert [h.k for h in ms.hists] == [1, 2, 3]
    for h in ms.hists:
        assert int(h.counts.sum()) + h.under + h.over == h.total
        assert h.total == sum(s.n for s in ms.blocks if s.k == h.k)
        assert h.vmin <= cs.hist_quantile(h, 0.5) <= h.vmax
        assert cs.hist_quantile(h, 0.1) <= cs.hist_quantile(h, 0.9)
    for g in ms.hists:
        wk = sorted([h for h in ms.whists if h.k == g.k], key=lambda h: h.win)
        assert len(wk) == ms.nwin and all(h.flo < h.fhi for h in wk)
        assert sum(h.total for h in wk) == g.total
        assert math.isclose(sum(h.sum for h in wk), g.sum, rel_tol=1e-12)
        assert np.array_equal(sum(h.counts for h in wk), g.counts)
        assert all(math.isclose(wk[i].fhi, wk[i + 1].flo, rel_tol=1e-12) for i in range(len(wk) - 1))
    summ = cs.metricstats_summary(ms, faps=(0.1, 0.01, 1e-3))
    assert [r["k"] for r in summ] == [1, 2, 3]
    assert summ[0]["mean"] > summ[1]["mean"] > summ[2]["mean"]
    assert summ[0]["fap"][0] > summ[2]["fap"][0]
    assert all(r["ntrials"] > 0 for r in cs.metricstats_windows(ms, faps=(0.1, 0.01)))

    _say("normalize")
    r_lo, r_hi, lodr, total, Nprof, nchunks = _grid(cs, p, T, 9.5, 10.5, 1024)
    ms = cs.MetricStats(hist_hi=30000.0, hist_nb=6000, nwin=8)
    eng.search_stats(r_lo, r_hi, math.inf, ms)
    norm = cs.build_metricnorm(ms)
    assert set(norm.loc) == {1, 2, 3}
    for g in ms.hists:
        assert len(norm.loc[g.k]) == ms.nwin and np.all(norm.scale[g.k] > 0)
        assert np.all((g.vmin <= norm.loc[g.k]) & (norm.loc[g.k] <= g.vmax))
    raw = eng.search_normalized(5.0, norm)
    order = [1, 2, 3]
    Md = np.stack([eng.stats_metrics(k, total) for k in order])
    want = R.normalized_candidates(Md, order, nh, T, r_lo, lodr, total, Nprof, norm.edges,
                                   norm.loc, norm.scale, 5.0)
    got = {(int(r["t"]), int(r["k"])): float(r["metric"]) for r in raw}
    assert set(got) == set(want) and len(got) > 0
    assert all(abs(got[key] - want[key]) <= 1e-12 * max(1.0, abs(want[key])) for key in got)
    cands = cs.search(ft, p, lofreq=9.5, hifreq=10.5, blocksize=1024, threshold=5.0,
                      metricstats=cs.MetricStats(hist_hi=30000.0, hist_nb=6000, nwin=8),
                      normalize=True, mode="fused")
    assert any(abs(c.freq - 10.0123) < 1e-2 for c in cands)
