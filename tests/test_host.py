"""Host rows without a GPU: candidate dedup, the PRESTO file reader and the
synthetic-spectrum generator.

* remove_duplicates / remove_harmonics (src/search.jl:2148-2214) against the
  reference's own fixtures (test/test_search.jl:113-150) and against a brute
  force over exact rationals;
* SimpleInf / FFTFile (src/fileio.jl:12-85): write-then-read of .fft/.inf
  pairs, the N-and-dt-only header, the missing-key and missing-file errors,
  and the reference's data-file checks (test/test_fileio.jl:9-26) re-hosted on
  a synthetic file of the same shape;
* the generator's definitions: unit mean power, von Mises harmonic ratios
  I_h(kappa)/I_0(kappa) against scipy's Bessel functions, the injected S/N as a
  matched-filter S/N in noise units, and the red-noise normalisation.
"""
import json
import math
import os
from fractions import Fraction

import numpy as np
import pytest

from coherentsearch import synth
from coherentsearch.dedup import harmonically_related, remove_duplicates, remove_harmonics
from coherentsearch.fileio import FFTFile, SimpleInf, freqs, write_fft
from coherentsearch.params import Candidate

FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                  "reference_fixtures.json")))


# ---- dedup ----------------------------------------------------------------------

def _mk(r, s, nharm, T=1000.0):
    return Candidate(r / T, s, r, nharm)


def test_remove_duplicates_reference_fixture():
    fx = FIX["remove_duplicates"]
    cands = [_mk(r, s, 32) for r, s in fx["cands"]]
    kept = remove_duplicates(cands, dr_tol=1.0)
    assert len(kept) == 3
    assert [c.freq for c in kept] == sorted(c.freq for c in kept)
    assert sorted(c.metric for c in kept) == fx["kept_metrics_sorted"]
    # the strongest member of each cluster, at its own r
    assert [(c.r, c.metric) for c in kept] == [(10000.02, 12.0), (20000.0, 7.0), (30000.1, 15.0)]
    assert len(remove_duplicates(cands, dr_tol=1e9)) == 1
    assert remove_duplicates([]) == []


def test_remove_duplicates_chains_and_order():
    """Grouping is by consecutive sorted r (a chain of close neighbours merges
    even when its ends are further apart than dr_tol), the input order does not
    matter, and a tie in metric keeps the first member by r."""
    cands = [_mk(100.0, 5.0, 8), _mk(100.9, 6.0, 8), _mk(101.8, 4.0, 8), _mk(103.0, 9.0, 8),
             _mk(103.0 + 1.0, 9.0, 8)]
    kept = remove_duplicates(cands[::-1], dr_tol=1.0)
    assert [(c.r, c.metric) for c in kept] == [(100.9, 6.0), (103.0, 9.0)]
    kept = remove_duplicates(cands, dr_tol=0.5)
    assert [c.r for c in kept] == [100.0, 100.9, 101.8, 103.0, 104.0]


def test_remove_harmonics_reference_fixture():
    fx = FIX["remove_harmonics"]
    fam = [_mk(r, s, 16) for r, s in fx["family"]]
    kept = remove_harmonics(fam, numharm=16, tol=1.0)
    assert len(kept) == 1 and kept[0].r == 10000.0 and kept[0].metric == 20.0
    mixed = fam + [_mk(fx["unrelated"][0], fx["unrelated"][1], 16)]
    kept2 = remove_harmonics(mixed, numharm=16, tol=1.0)
    assert len(kept2) == 2
    assert [c.freq for c in kept2] == sorted(c.freq for c in kept2)
    assert {round(c.r) for c in kept2} == {10000, 13337}
    assert len(remove_harmonics(fam, numharm=1, tol=1.0)) == 4
    assert remove_harmonics([]) == []


def _related_brute(r1, r2, numharm, tol):
    """|m*hi - n*lo| <= tol*m for some integers 1 <= m, n <= numharm, every
    pair tried, in exact rational arithmetic."""
    lo, hi = sorted((Fraction(r1), Fraction(r2)))
    if lo <= 0:
        return False
    return any(abs(m * hi - n * lo) <= Fraction(tol) * m
               for m in range(1, numharm + 1) for n in range(1, numharm + 1))


def test_harmonically_related_vs_brute_force():
    """The nearest-ratio search (src/search.jl:2179-2188) finds a related pair
    exactly when some (m, n) in the full grid does; r on a quarter-bin grid so
    the FP64 products are exact."""
    rng = np.random.default_rng(5)
    agree = 0
    for _ in range(3000):
        lo = float(rng.integers(4, 400)) / 4.0
        num, den = int(rng.integers(1, 9)), int(rng.integers(1, 9))
        hi = round(4.0 * lo * num / den) / 4.0 + float(rng.integers(-12, 13)) / 4.0
        hi = max(hi, 0.25)
        numharm = int(rng.integers(1, 9))
        tol = float(rng.choice([0.25, 1.0, 2.0]))
        want = _related_brute(lo, hi, numharm, tol)
        # any n other than the nearest one leaves |m*hi - n*lo| >= lo/2, so the
        # nearest-ratio search is complete when tol*m < lo/2 for every m
        if tol * numharm < min(lo, hi) / 2:
            assert harmonically_related(lo, hi, numharm, tol) == want, (lo, hi, numharm, tol)
            agree += 1
    assert agree > 1000
    assert not harmonically_related(0.0, 5.0, 16, 1.0)


def test_remove_harmonics_keeps_strongest_first():
    """Visiting by descending metric: a weaker fundamental is dropped against
    a stronger harmonic of it, and equal metrics keep input order."""
    a = _mk(1000.0, 10.0, 16)
    b = _mk(3000.0, 12.0, 16)
    c = _mk(1700.0, 11.0, 16)
    kept = remove_harmonics([a, b, c], numharm=4, tol=1.0)
    assert [x.r for x in kept] == [1700.0, 3000.0]
    d = _mk(2000.0, 10.0, 16)
    kept = remove_harmonics([d, a], numharm=4, tol=1.0)
    assert [x.r for x in kept] == [2000.0]
    kept = remove_harmonics([a, d], numharm=4, tol=1.0)
    assert [x.r for x in kept] == [1000.0]


# ---- PRESTO files ---------------------------------------------------------------

def test_write_read_full_inf(tmp_path):
    amps = synth.make_spectrum(1 << 12, 1e-3, seed=3)
    amps[0] = np.complex64(3.5 - 1.25j)           # DC and Nyquist packed in element 0
    stem = str(tmp_path / "obs_DM12.50")
    write_fft(stem, amps, 1 << 12, 1e-3, full_inf=True, obj="J0000+0000", epoch=58000.25, dm=12.5)
    ft = FFTFile(stem + ".fft")
    assert ft.inf.object == "J0000+0000"
    assert ft.inf.epoch == 58000.25 and ft.inf.DM == 12.5
    assert ft.N == 4096 and ft.inf.dt == 1e-3
    assert ft.T == 4096 * 1e-3 and ft.df == 1.0 / ft.T
    assert ft.amps.dtype == np.complex64 and len(ft.amps) == 2048
    assert np.array_equal(np.asarray(ft.amps), amps)
    assert ft.DC == np.float32(3.5) and ft.Nyquist == np.float32(-1.25)
    assert not ft.dereddened and not ft.detrended
    f = freqs(ft)
    assert len(f) == ft.N // 2 and f[0] == 0.0 and f[1] == ft.df


def test_reference_data_file_shape(tmp_path):
    """test/test_fileio.jl:9-26 on a file of the same shape: N = 1,000,000,
    dt = 0.001, so T = 1000 s, df = 1e-3 Hz and 500,000 complex64 bins."""
    amps = np.zeros(500_000, dtype=np.complex64)
    amps[1:] = 1.0
    stem = str(tmp_path / "harmonics_hi")
    write_fft(stem, amps, 1_000_000, 0.001)
    ft = FFTFile(stem + ".fft")
    assert ft.N == 1_000_000
    assert math.isclose(ft.inf.dt, 0.001)
    assert math.isclose(ft.T, 1000.0) and math.isclose(ft.df, 1.0 / 1000.0)
    assert len(ft.amps) == ft.N // 2 and ft.amps.dtype == np.complex64
    assert ft.dereddened is False
    f = freqs(ft)
    assert len(f) == ft.N // 2 and f[0] == 0.0


def test_inf_with_n_and_dt_only(tmp_path):
    stem = str(tmp_path / "c1")
    write_fft(stem, np.ones(8, np.complex64), 16, 64e-6, full_inf=False)
    assert len(open(stem + ".inf").read().splitlines()) == 2
    inf = SimpleInf.read(stem + ".inf")
    assert (inf.object, inf.epoch, inf.DM) == (None, None, None)
    assert inf.N == 16 and inf.dt == 64e-6
    ft = FFTFile(stem + ".fft")
    assert ft.T == 16 * 64e-6


def test_dereddened_from_name(tmp_path):
    stem = str(tmp_path / "obs_red")
    write_fft(stem, np.ones(8, np.complex64), 16, 1e-3)
    ft = FFTFile(stem + ".fft")
    assert ft.dereddened and ft.detrended


@pytest.mark.parametrize("drop,msg", [("Number of bins", "Missing 'Number of bins'"),
                                      ("Width of each time series bin",
                                       "Missing 'Width of each time series bin'")])
def test_missing_inf_key_raises(tmp_path, drop, msg):
    """src/fileio.jl:67-68: a header without N or dt is an error naming the key."""
    stem = str(tmp_path / "bad")
    write_fft(stem, np.ones(8, np.complex64), 16, 1e-3, full_inf=True)
    lines = [ln for ln in open(stem + ".inf").read().splitlines() if drop not in ln]
    open(stem + ".inf", "w").write("\n".join(lines) + "\n")
    with pytest.raises(ValueError, match=msg):
        FFTFile(stem + ".fft")


def test_missing_inf_file_raises(tmp_path):
    """src/fileio.jl:25: no .inf beside the .fft is an error naming the path."""
    p = str(tmp_path / "lonely.fft")
    np.ones(4, np.complex64).tofile(p)
    with pytest.raises(FileNotFoundError, match="lonely.inf"):
        FFTFile(p)


def test_inf_keys_need_leading_space(tmp_path):
    """Keys match only at the PRESTO column (a leading space), as startswith
    in the reference; other lines are ignored."""
    p = str(tmp_path / "x.inf")
    open(p, "w").write("Number of bins in the time series = 99\n"
                       " Number of bins in the time series          =  64\n"
                       " Width of each time series bin (sec)        =  0.5\n"
                       " Any other line = 7\n")
    inf = SimpleInf.read(p)
    assert inf.N == 64 and inf.dt == 0.5


# ---- the generator --------------------------------------------------------------

def test_noise_unit_mean_power_and_dc():
    amps = synth.make_spectrum(1 << 21, 64e-6, seed=17)
    assert amps.dtype == np.complex64 and len(amps) == 1 << 20
    assert amps[0] == 0
    p = np.abs(amps[1:].astype(np.complex128)) ** 2
    # exponential power with mean 1: the sample mean's sd is 1/sqrt(n) ~ 1e-3
    assert abs(p.mean() - 1.0) < 6e-3
    # Re and Im independent N(0, 1/2)
    assert abs(np.var(amps.real[1:]) - 0.5) < 5e-3 and abs(np.var(amps.imag[1:]) - 0.5) < 5e-3
    assert abs(np.corrcoef(amps.real[1:], amps.imag[1:])[0, 1]) < 5e-3
    # the same seed gives the same spectrum
    assert np.array_equal(amps, synth.make_spectrum(1 << 21, 64e-6, seed=17))


@pytest.mark.parametrize("duty", [0.02, 0.08, 0.2])
def test_vonmises_harmonics_closed_form(duty):
    """The generator's time-domain pulse train, sampled with a whole number of
    periods in the series, has DFT lines at multiples of its fundamental bin
    whose ratios are I_h(kappa)/I_1(kappa) (the Fourier series of
    exp(kappa cos phi) is I_0 + 2 sum I_h cos(h phi)); kappa from the FWHM."""
    from scipy.special import iv
    N, dt = 1 << 16, 1e-3
    T = N * dt
    fbin = 64                                   # 1024 samples per period
    psr = synth.Pulsar(fbin / T, duty, 10.0)
    ts = synth.pulsar_timeseries(N, dt, psr)
    assert abs(ts.mean()) < 1e-12 * N           # exact zero mean removed
    S = np.fft.rfft(ts)
    k = synth.vonmises_kappa(duty)
    # the FWHM definition: the profile is 1/2 of its peak at +-duty/2 cycles
    assert math.isclose(math.exp(k * (math.cos(math.pi * duty) - 1.0)), 0.5, rel_tol=1e-12)
    H = 20
    got = np.abs(S[fbin * np.arange(1, H + 1)])
    want = iv(np.arange(1, H + 1), k) / iv(0, k)
    closed = synth.vonmises_harmonics(duty, H)
    assert np.allclose(closed, want, rtol=1e-12, atol=0)
    rel = got / got[0]
    ref = want / want[0]
    keep = ref > 1e-6
    assert np.allclose(rel[keep], ref[keep], rtol=1e-6, atol=1e-12)
    # the absolute scale: N * e^-kappa * I_h(kappa) per line
    assert np.allclose(got[keep], N * math.exp(-k) * iv(np.arange(1, H + 1), k)[keep],
                       rtol=1e-6)
    # no power between the lines
    off = np.ones(len(S), bool)
    off[::fbin] = False
    assert np.max(np.abs(S[off])) < 1e-6 * got[0]


def test_injected_snr_definition():
    """The injection S (the spectrum with the pulsar minus the same seed's noise)
    has sqrt(2 sum |S|^2) = snr, and the matched filter
    Z = Re(sum conj(S) X) / sqrt(sum |S|^2 / 2) has mean snr and unit sd over
    noise realisations, i.e. snr is in units of the noise rms."""
    N, dt = 1 << 14, 1e-3
    psr = synth.Pulsar(37.3, 0.05, 12.0, phase=0.3)
    base = synth.make_spectrum(N, dt, seed=101)
    withp = synth.make_spectrum(N, dt, seed=101, pulsars=[psr])
    S = withp.astype(np.complex128) - base.astype(np.complex128)
    assert S[0] == 0
    assert math.isclose(math.sqrt(2.0 * np.sum(np.abs(S) ** 2)), 12.0, rel_tol=1e-5)
    norm = math.sqrt(np.sum(np.abs(S) ** 2) / 2.0)
    zs, zn = [], []
    for seed in range(400):
        noise = synth.make_spectrum(N, dt, seed=1000 + seed).astype(np.complex128)
        zn.append(np.real(np.sum(np.conj(S) * noise)) / norm)
        zs.append(np.real(np.sum(np.conj(S) * (noise + S))) / norm)
    zn, zs = np.array(zn), np.array(zs)
    # 400 draws: the sd estimate is good to ~4 %, the mean to ~0.05
    assert abs(np.std(zn) - 1.0) < 0.15
    assert abs(np.mean(zn)) < 0.2
    assert abs(np.mean(zs) - 12.0) < 0.2


def test_red_noise_normalised():
    """Power ~ f^-1 below red_fc before normalisation; afterwards every
    1000-bin block's median power is ln 2 (an exponential's median at unit
    mean), so blocks whose gain is flat have unit mean power; inside a block
    the f^-1 slope is left as it was (the block holding the lowest bins keeps
    its excess toward DC)."""
    N, dt = 1 << 24, 64e-6
    T = N * dt                                   # 1 Hz = bin 1074
    amps = synth.make_spectrum(N, dt, seed=23, red_fc=1.0)
    p = np.abs(amps.astype(np.complex128)) ** 2
    for s in (1, 1001, 2001, 300001, 8387001):
        med = np.median(p[s:s + 1000])
        assert math.isclose(med, math.log(2.0), rel_tol=1e-5), (s, med)
    assert abs(p[1001:5001].mean() - 1.0) < 0.05     # f > 0.93 Hz: gain within 7 %
    assert abs(p[5001:].mean() - 1.0) < 2e-3
    first = p[1:1001]
    assert first[:100].mean() > 3.0 * first[500:].mean()
    # before the normalisation the low band carries the f^-1 excess
    rng = np.random.Generator(np.random.PCG64(23))
    nb = N // 2
    raw = (rng.standard_normal(nb) + 1j * rng.standard_normal(nb)) / math.sqrt(2)
    f = np.arange(nb) / T
    low = (f > 0) & (f < 1.0)
    raw[low] *= np.sqrt(1.0 / f[low])
    lo_band = (f > 0.1) & (f < 0.2)
    hi_band = (f > 0.4) & (f < 0.8)
    ratio = (np.abs(raw[lo_band]) ** 2).mean() / (np.abs(raw[hi_band]) ** 2).mean()
    want = (np.log(0.2 / 0.1) / 0.1) / (np.log(0.8 / 0.4) / 0.4)
    assert abs(ratio / want - 1.0) < 0.15


def test_parallel_generation_equals_serial():
    cfgs = [dict(N=1 << 14, dt=1e-3, seed=s, pulsars=[synth.Pulsar(7.0 + s, 0.05, 10.0)])
            for s in range(3)]
    serial = [synth.make_config_spectrum(c) for c in cfgs]
    par = synth.make_config_spectra(cfgs, workers=2)
    assert all(np.array_equal(a, b) for a, b in zip(serial, par))


# ---- the time-budget block scheduler (host logic) ---------------------------

class _Blocks:
    """A stand-in for Engine.search_blocks over nb blocks: one record per
    selected block (returned in reverse order), each call taking `wait` s."""

    def __init__(self, nb, wait=0.0):
        self.nb, self.wait, self.calls = nb, wait, []

    def search_blocks(self, r_lo, r_hi, threshold, mask):
        import time
        from coherentsearch import _lib
        assert len(mask) == self.nb
        sel = np.flatnonzero(mask)
        self.calls.append(sel)
        if self.wait:
            time.sleep(self.wait)
        out = np.zeros(len(sel), dtype=_lib.CAND_DTYPE)
        out["t"], out["k"], out["block"], out["metric"] = sel * 100 + 7, 1, sel, 9.0
        return out[::-1]


def test_scheduled_search_rounds_cover_the_order():
    """Rounds of 1 %, then twice as many blocks each time, in the given order,
    every block once; the merged records come back sorted by (t, k)."""
    from coherentsearch.triage import scheduled_search
    nb = 1000
    order = np.random.default_rng(1).permutation(nb)
    eng = _Blocks(nb)
    raw, done, _ = scheduled_search(eng, 0.0, 1.0, 8.0, order)
    assert done == nb
    sizes = [len(c) for c in eng.calls]
    assert sizes == [10, 20, 40, 80, 160, 320, 370]
    for call, want in zip(eng.calls, np.split(order, np.cumsum(sizes)[:-1])):
        assert np.array_equal(call, np.sort(want))
    assert np.array_equal(raw["block"], np.arange(nb))          # sorted by t = 100 b + 7


def test_scheduled_search_budget():
    """A zero budget runs the first round only; a budget of a few calls stops
    before the band is done and overshoots it by about one call at most (the
    bound leaves room for a loaded host)."""
    import time
    from coherentsearch.triage import scheduled_search
    nb = 1000
    order = np.arange(nb)[::-1]
    eng = _Blocks(nb)
    raw, done, _ = scheduled_search(eng, 0.0, 1.0, 8.0, order, budget_s=0.0)
    assert done == 10 and len(eng.calls) == 1
    assert np.array_equal(raw["block"], np.arange(990, 1000))
    eng = _Blocks(nb, wait=0.02)
    t0 = time.perf_counter()
    raw, done, secs = scheduled_search(eng, 0.0, 1.0, 8.0, order, budget_s=0.07)
    assert 10 < done < nb and len(raw) == done
    assert secs < 0.07 + 0.06 and time.perf_counter() - t0 >= secs
