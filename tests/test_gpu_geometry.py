"""GPU parity of the search over trial-grid geometries other than the default
hidr = 0.5: other periods q of the rational trial step (lodr = hidr/nharms =
p/q, directinterp.jl:217-225), the 128-bin fold (nharms 64), and a trial step
that is not a small rational (the per-trial FP64 interpolation,
directinterp.jl:331-377), against the oracle on the same spectrum.

Bars as in the other search parity tests: identical (t, k) sets away from
threshold, identical harmonic counts and blocks, S/N within 1e-4 relative.
"""
import math
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _compare(got, want, thr, tol=1e-4, band=1e-3):
    g = {(int(r["t"]), int(r["k"])): r for r in got}
    w = {(int(r["t"]), int(r["k"])): r for r in want}
    for key in set(g) ^ set(w):
        r = g.get(key, w.get(key))
        assert abs(float(r["metric"]) - thr) <= band * thr, (key, float(r["metric"]))
    worst = 0.0
    for key in set(g) & set(w):
        assert g[key]["nharm"] == w[key]["nharm"] and g[key]["block"] == w[key]["block"], key
        worst = max(worst, abs(g[key]["metric"] - w[key]["metric"]) / abs(w[key]["metric"]))
    assert worst <= tol, worst
    return worst, len(set(g) & set(w))


# (nharms, hidr, maxdecim, blocksize): q = 240 (S = 240), q = 200 (S = 400),
# q = 90 (S = 720), and the 128-bin fold at q = 128
GEOMS = [(60, 0.25, 6, 2048), (60, 0.3, 4, 1500), (30, 1.0 / 3.0, 3, 1024), (64, 0.5, 2, 2048)]


@pytest.mark.parametrize("mode", ["plain", "fused", "fused_valu"])
@pytest.mark.parametrize("nharms,hidr,maxdecim,blocksize", GEOMS)
def test_search_geometry_vs_oracle(small_spectrum, nharms, hidr, maxdecim, blocksize, mode):
    import coherentsearch as cs
    import oracle as O
    amps, N, dt = small_spectrum
    T = N * dt
    dec = cs.decimation_set(nharms, maxdecim)
    _say(f"geometry mode={mode} nharms={nharms} hidr={hidr:.6g} decims={dec} blocksize={blocksize}")
    p = cs.SearchParams(nharms=nharms, m=16, hidr=hidr, decimations=dec)
    eng = cs.Engine(p, blocksize)
    eng.set_mode(mode)
    eng.set_spectrum(cs.FFTFile.from_arrays(amps, N, dt))
    r_lo, r_hi = 3.0 * T, 30.0 * T
    lodr = hidr / nharms
    total = math.floor((r_hi - r_lo) / lodr) + 1
    got = eng.search_raw(r_lo, r_hi, 6.5)
    want = O.search_region(O.make_params(nharms=nharms, m=16, hidr=hidr, decimations=dec),
                           amps, N, r_lo, total, blocksize, threshold=6.5)
    assert len(want) > 0
    worst, ncommon = _compare(got, want, 6.5)
    _say(f"  {len(got)} vs {len(want)} candidates, common {ncommon}, worst rel {worst:.2e}")


def test_irrational_step_plain_vs_oracle(small_spectrum):
    """hidr = sqrt(2)/4: no period q <= 2^16, so the plain mode interpolates
    each trial directly in FP64 and the fused modes refuse the grid."""
    import coherentsearch as cs
    import oracle as O
    amps, N, dt = small_spectrum
    T = N * dt
    nharms, hidr = 16, math.sqrt(2.0) / 4.0
    dec = cs.decimation_set(nharms, 3)
    _say(f"irrational step nharms={nharms} hidr={hidr!r}")
    assert O.trial_grid_rational(hidr / nharms) is None
    p = cs.SearchParams(nharms=nharms, m=16, hidr=hidr, decimations=dec)
    eng = cs.Engine(p, 1024)
    eng.set_mode("plain")
    eng.set_spectrum(cs.FFTFile.from_arrays(amps, N, dt))
    r_lo, r_hi = 8.0 * T, 20.0 * T
    total = math.floor((r_hi - r_lo) / (hidr / nharms)) + 1
    got = eng.search_raw(r_lo, r_hi, 6.0)
    want = O.search_region(O.make_params(nharms=nharms, m=16, hidr=hidr, decimations=dec),
                           amps, N, r_lo, total, 1024, threshold=6.0)
    assert len(want) > 0
    worst, ncommon = _compare(got, want, 6.0)
    _say(f"  {len(got)} vs {len(want)} candidates, common {ncommon}, worst rel {worst:.2e}")
    for mode in ("fused", "fused_valu"):
        e = cs.Engine(p, 1024)
        e.set_mode(mode)
        e.set_spectrum(cs.FFTFile.from_arrays(amps, N, dt))
        with pytest.raises(ValueError):
            e.search_raw(r_lo, r_hi, 6.0)
