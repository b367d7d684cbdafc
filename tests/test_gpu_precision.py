"""The profile-stage precision of SearchParams (src/search.jl:143,151-155) and
ties in the split-BF16 pass's median select.

precision :f64 (the default): every reported metric is scaled by the block
sigma of FP64 profiles over the reference's sampled columns (_block_sigma,
src/search.jl:490-566), recomputed for each (block, fold) that holds gate
survivors; so the three execution modes return the same raw candidates bit for
bit, and they equal the oracle's (a Float64 profile stage) to FP64 rounding.
precision :f32: the sigma is the pass's own (FP32 profiles in the plain mode,
split-BF16 in the fused mode).

The tie tests give k_select (the fused modes' exact median of FP32 sample
keys) inputs with many equal keys: a constant-amplitude spectrum, whose
trials with equal residues of the trial grid have identical profiles, and a
noise spectrum zeroed above 20 Hz, whose blocks there hold only zeros (sigma
0) and whose blocks below hold a varying share of all-zero columns; each
(block, fold) sigma against the oracle's.
"""
import math
import sys

import numpy as np
import pytest

from conftest import noise_amps

pytestmark = pytest.mark.gpu

MODES = ("plain", "fused", "fused_valu")


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def cs():
    import coherentsearch
    return coherentsearch


@pytest.fixture(scope="module")
def O():
    import oracle
    return oracle


def _search(cs, amps, N, dt, nharms, dec, blocksize, prec, mode, lo, hi, thr):
    p = cs.SearchParams(nharms=nharms, m=16, decimations=dec, precision=prec)
    eng = cs.Engine(p, blocksize)
    try:
        eng.set_mode(mode)
        eng.set_spectrum(cs.FFTFile.from_arrays(amps, N, dt))
        raw = eng.search_raw(lo, hi, thr)
        sig = eng.last_sigmas()
    finally:
        eng.close()
    return raw, sig


@pytest.mark.parametrize("nharms,maxdecim,blocksize,thr", [(60, 6, 2048, 6.5), (32, 4, 512, 6.0),
                                                           (8, 1, 2048, 6.0)])
def test_f64_modes_identical_and_equal_to_oracle(cs, O, small_spectrum, nharms, maxdecim,
                                                 blocksize, thr):
    amps, N, dt = small_spectrum
    T = N * dt
    dec = cs.decimation_set(nharms, maxdecim)
    lo, hi = 1.0 * T, 60.0 * T
    total = math.floor((hi - lo) / (0.5 / nharms)) + 1
    nch = -(-total // blocksize)
    wsig = np.zeros((nch, len(dec)))
    want = O.search_region(O.make_params(nharms=nharms, m=16, decimations=dec), amps, N, lo,
                           total, blocksize, threshold=thr, sigmas=wsig)
    assert len(want) > 0
    out = {m: _search(cs, amps, N, dt, nharms, dec, blocksize, "f64", m, lo, hi, thr)
           for m in MODES}
    for m in MODES[1:]:
        assert np.array_equal(out[m][0], out["plain"][0]), m
    got, sig = out["fused"]
    g = {(int(c["t"]), int(c["k"])): c for c in got}
    w = {(int(c["t"]), int(c["k"])): c for c in want}
    for key in set(g) ^ set(w):
        rec = g.get(key, w.get(key))
        assert abs(float(rec["metric"]) - thr) <= 1e-9 * thr, key
    worst = 0.0
    for key in set(g) & set(w):
        assert (g[key]["nharm"], g[key]["block"]) == (w[key]["nharm"], w[key]["block"])
        worst = max(worst, abs(g[key]["metric"] - w[key]["metric"]) / w[key]["metric"])
    # the pairs that hold a candidate carry the FP64 sigma
    fidx = {k: i for i, k in enumerate([1] + [k for k in dec if k > 1])}
    pairs = {(int(c["block"]), fidx[int(c["k"])]) for c in got}
    srel = max(abs(sig[b, f] - wsig[b, f]) / wsig[b, f] for b, f in pairs)
    _say(f"  f64: {len(got)} candidates in every mode, worst S/N rel {worst:.2e}, "
         f"{len(pairs)} (block, fold) sigmas worst rel {srel:.2e}")
    assert worst <= 1e-10
    assert srel <= 1e-12


def test_f64_refine_past_the_sample_scratch(cs, O, small_spectrum):
    """Blocks of 16 trials and a threshold of 3 flag about 15,800 (block, fold)
    pairs, far more than the refinement's sample scratch holds (2048): the
    first 2048 listed pairs go through the column-chunk and select kernels, the
    rest through the one-workgroup-per-pair kernel.  Every mode returns the
    same raw candidates, equal to the oracle's, and every flagged pair's sigma
    is the oracle's FP64 sigma."""
    amps, N, dt = small_spectrum
    T = N * dt
    nharms, blocksize, thr = 32, 16, 3.0
    dec = cs.decimation_set(nharms, 4)
    lo, hi = 1.0 * T, 60.0 * T
    total = math.floor((hi - lo) / (0.5 / nharms)) + 1
    nch = -(-total // blocksize)
    wsig = np.zeros((nch, len(dec)))
    want = O.search_region(O.make_params(nharms=nharms, m=16, decimations=dec), amps, N, lo,
                           total, blocksize, threshold=thr, sigmas=wsig)
    out = {m: _search(cs, amps, N, dt, nharms, dec, blocksize, "f64", m, lo, hi, thr)
           for m in MODES}
    for m in MODES[1:]:
        assert np.array_equal(out[m][0], out["plain"][0]), m
    got, sig = out["fused"]
    fidx = {k: i for i, k in enumerate([1] + [k for k in dec if k > 1])}
    pairs = {(int(c["block"]), fidx[int(c["k"])]) for c in got}
    _say(f"  {len(got)} candidates, {len(pairs)} (block, fold) pairs with candidates, "
         f"{nch * len(dec)} pairs")
    assert len(pairs) > 4096
    g = {(int(c["t"]), int(c["k"])): c for c in got}
    w = {(int(c["t"]), int(c["k"])): c for c in want}
    for key in set(g) ^ set(w):
        rec = g.get(key, w.get(key))
        assert abs(float(rec["metric"]) - thr) <= 1e-9 * thr, key
    worst = max(abs(g[k]["metric"] - w[k]["metric"]) / w[k]["metric"] for k in set(g) & set(w))
    srel = max(abs(sig[b, f] - wsig[b, f]) / wsig[b, f] for b, f in pairs)
    _say(f"  worst S/N rel {worst:.2e}, sigma rel {srel:.2e}")
    assert worst <= 1e-10
    assert srel <= 1e-12


def test_f32_keeps_the_pass_sigma(cs, O, small_spectrum):
    """:f32 reports with the pass's own sigma: the split-BF16 mode's metrics
    move from the :f64 ones by ~1e-5 (not 0), still within 1e-4 of the oracle."""
    amps, N, dt = small_spectrum
    T = N * dt
    dec = cs.decimation_set(60, 6)
    lo, hi, thr = 1.0 * T, 60.0 * T, 6.5
    a, _ = _search(cs, amps, N, dt, 60, dec, 2048, "f64", "fused", lo, hi, thr)
    b, _ = _search(cs, amps, N, dt, 60, dec, 2048, "f32", "fused", lo, hi, thr)
    ka = {(int(c["t"]), int(c["k"])): float(c["metric"]) for c in a}
    kb = {(int(c["t"]), int(c["k"])): float(c["metric"]) for c in b}
    common = set(ka) & set(kb)
    assert len(common) > 0.9 * len(ka)
    d = max(abs(ka[k] - kb[k]) / ka[k] for k in common)
    _say(f"  :f32 fused vs :f64: worst rel {d:.2e} over {len(common)} candidates")
    assert 0 < d <= 1e-4


def _tie_spectra():
    N, dt = 1 << 18, 64e-6
    const = np.full(N // 2, 1.0 + 0.5j, dtype=np.complex64)
    const[0] = 0
    band = noise_amps(N // 2, seed=77)
    T = N * dt
    band[int(20.0 * T):] = 0       # fundamentals above 20 Hz read zeros at every harmonic
    return {"constant": (const, N, dt), "zero_band": (band, N, dt)}


@pytest.mark.parametrize("kind", ["constant", "zero_band"])
@pytest.mark.parametrize("mode", ["fused", "fused_valu", "plain"])
def test_sigma_select_with_ties(cs, O, kind, mode):
    amps, N, dt = _tie_spectra()[kind]
    T = N * dt
    nharms, maxdecim = (32, 4) if kind == "zero_band" else (60, 6)
    dec = cs.decimation_set(nharms, maxdecim)
    blocksize = 1000
    lo, hi = 2.0 * T, 40.0 * T
    total = math.floor((hi - lo) / (0.5 / nharms)) + 1
    nch = -(-total // blocksize)
    wsig = np.zeros((nch, len(dec)))
    O.search_region(O.make_params(nharms=nharms, m=16, decimations=dec), amps, N, lo, total,
                    blocksize, threshold=1e9, sigmas=wsig)
    _, sig = _search(cs, amps, N, dt, nharms, dec, blocksize, "f32", mode, lo, hi, 1e9)
    assert sig.shape == wsig.shape
    zero = wsig == 0
    # sigma 0 exactly where more than half of the samples are exact zeros
    assert np.array_equal(sig == 0, zero), np.argwhere((sig == 0) != zero)[:5]
    rel = np.abs(sig[~zero] - wsig[~zero]) / wsig[~zero]
    tol = 5e-5 if mode == "fused" else 1e-5
    _say(f"  {kind} {mode}: {sig.size} sigmas ({zero.sum()} zero), worst rel {rel.max():.2e}")
    assert rel.max() <= tol
