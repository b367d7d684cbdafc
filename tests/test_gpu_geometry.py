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


@pytest.fixture(scope="module")
def big_spectrum():
    """2^28 samples (2^27 bins, 1 GiB): unit-power noise and a comb of 60
    harmonics of r0 = 2e6 + 37.25 (amplitude 3 at the nearest bins)."""
    N = 1 << 28
    rng = np.random.default_rng(2028)
    nb = N // 2
    amps = np.empty(nb, dtype=np.complex64)
    for i0 in range(0, nb, 1 << 24):   # in 16M-bin slabs: bounded host memory
        n = min(1 << 24, nb - i0)
        amps[i0:i0 + n] = ((rng.standard_normal(n) + 1j * rng.standard_normal(n))
                           / np.sqrt(2)).astype(np.complex64)
    amps[0] = 0
    r0 = 2.0e6 + 37.25
    for h in range(1, 61):
        amps[int(round(h * r0))] += np.complex64(3.0 * np.exp(1j * 0.37 * h))
    return amps, N, 64e-6, r0


@pytest.mark.parametrize("mode", ["plain", "fused", "fused_valu"])
def test_harmonic_bins_near_1e8_vs_oracle(big_spectrum, mode):
    """Fundamentals near r = 2e6, so the 60 harmonics reach h r = 1.2e8 bins:
    the integer harmonic-bin arithmetic (directinterp.jl:414-418, 470-496)
    must stay exact there, and the comb at r0 is the strongest detection."""
    import coherentsearch as cs
    import oracle as O
    amps, N, dt, r0 = big_spectrum
    nharms = 60
    dec = cs.decimation_set(nharms, 2)
    p = cs.SearchParams(nharms=nharms, m=16, decimations=dec)
    r_lo, r_hi = r0 - 40.0, r0 + 40.0
    total = math.floor((r_hi - r_lo) / (0.5 / nharms)) + 1
    _say(f"large r mode={mode} N=2^28 r0={r0} trials={total} max h r={nharms * r_hi:.4g}")
    eng = cs.Engine(p, 2048)
    eng.set_mode(mode)
    eng.set_spectrum(cs.FFTFile.from_arrays(amps, N, dt))
    got = eng.search_raw(r_lo, r_hi, 6.0)
    want = O.search_region(O.make_params(nharms=nharms, m=16, decimations=dec), amps, N, r_lo,
                           total, 2048, threshold=6.0)
    assert len(want) > 0
    worst, ncommon = _compare(got, want, 6.0)
    best = want[np.argmax(want["metric"])]
    rbest = r_lo + int(best["t"]) * (0.5 / nharms)
    assert abs(rbest - r0) < 0.5 and int(best["k"]) == 1
    _say(f"  {len(got)} vs {len(want)} candidates, common {ncommon}, worst rel {worst:.2e}, "
         f"best r {rbest:.3f} S/N {float(best['metric']):.1f}")
    eng.close()


def test_fold_factor_8_vs_oracle(small_spectrum):
    """Folds 1, 2 and 8 at nharms 64 (128, 64 and 16 bins) in the split-BF16
    mode: no fold-set kernel matches, so the run-time kernel gathers the 16-bin
    fold's stacked rows at the K-slot spacing of fold factor 8 (one harmonic
    between lane groups) and the 64-bin fold's at that of factor 2."""
    import coherentsearch as cs
    import oracle as O
    amps, N, dt = small_spectrum
    T = N * dt
    nharms, dec, blocksize = 64, [1, 2, 8], 2048
    _say(f"fold factor 8: nharms={nharms} decims={dec} blocksize={blocksize}")
    p = cs.SearchParams(nharms=nharms, m=16, decimations=dec)
    eng = cs.Engine(p, blocksize)
    eng.set_mode("fused")
    eng.set_spectrum(cs.FFTFile.from_arrays(amps, N, dt))
    r_lo, r_hi = 3.0 * T, 30.0 * T
    total = math.floor((r_hi - r_lo) / (0.5 / nharms)) + 1
    got = eng.search_raw(r_lo, r_hi, 6.5)
    want = O.search_region(O.make_params(nharms=nharms, m=16, decimations=dec),
                           amps, N, r_lo, total, blocksize, threshold=6.5)
    assert len(want) > 0
    assert {int(r["k"]) for r in want} >= {8}, "no fold-8 candidate to compare"
    worst, ncommon = _compare(got, want, 6.5)
    _say(f"  {len(got)} vs {len(want)} candidates, common {ncommon}, worst rel {worst:.2e}")
    eng.close()
