"""Pins of the CPU oracle (oracle/cs_oracle.c) — the checker every GPU parity
test trusts.  Each check is independent of the restatement's own formula: the
reference tests' golden values, closed forms, numpy/scipy routines, exact
rational arithmetic and brute force on small inputs.
"""
import json
import math
import os
from fractions import Fraction

import numpy as np
import pytest

import oracle as O
from conftest import noise_amps

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                   "reference_fixtures.json")))


# ---- fourierinterp.jl ---------------------------------------------------------

def test_finterp_coeffs_golden():
    g = GOLD["finterp_coeffs_dr0.1_m10"]
    want = np.array(g["re"]) + 1j * np.array(g["im"])
    got = O.finterp_coeffs(g["dr"], g["m"])
    assert np.allclose(got, want, rtol=1e-5, atol=1e-6)
    g0 = GOLD["finterp_coeffs_dr0_m6"]
    assert np.allclose(O.finterp_coeffs(0.0, 6), np.array(g0["re"]) + 1j * np.array(g0["im"]),
                       atol=1e-12)
    with pytest.raises(ValueError):
        O.finterp_coeffs(0.1, 5)
    with pytest.raises(ValueError):
        O.finterp_coeffs(1.0, 6)


def test_finterp_coeffs_closed_form():
    """sinc(dr-j) exp(i pi (dr-j)) with numpy's own sinc/exp."""
    for dr in (0.013, 0.25, 0.5, 0.73, 0.999):
        for m in (8, 16, 32):
            off = dr - np.arange(-(m // 2) + 1, m // 2 + 1)
            want = np.sinc(off) * np.exp(1j * np.pi * off)
            # numpy forms sin(pi*x) from the rounded product pi*x (~1e-13 near
            # integers); the oracle reduces the argument exactly first.
            assert np.max(np.abs(O.finterp_coeffs(dr, m) - want)) < 1e-12


def test_nearby_bins_fixtures():
    for r, m, want in GOLD["nearby_bins"]["cases"]:
        lo, hi = O.nearby_fourier_bin_range(r, m)
        assert list(range(lo, hi)) == want
    lo, hi = O.nearby_fourier_bin_range(100.7, 32)
    assert hi - lo == 32 and (lo, hi) == O.nearby_fourier_bin_range(100.2, 32)


def _cosine_ft(r, N=32768):
    sig = np.cos(2 * np.pi * r * np.arange(N) / N + np.pi / 4)
    return np.fft.rfft(sig)


def test_fourier_interp_analytic_cosine():
    g = GOLD["analytic_cosine"]
    N = g["N"]
    want = N / 2 / math.sqrt(2) * (1 + 1j)
    for r in g["r"]:
        ft = _cosine_ft(r, N)
        iv = O.fourier_interp(r, ft, g["m"])
        assert abs(iv - want) <= g["atol"] + g["rtol"] * abs(want)
        # finterp_multi consistency: the same points on a 20-point sub-bin grid
        rs = math.floor(r) + np.arange(20) / 20
        idx = int(round((r - math.floor(r)) * 20))
        assert abs(O.fourier_interp(rs[idx], ft, g["m"]) - iv) < 1e-9


def test_fourier_interp_numpy_dot():
    """sum conj(c) * bins with numpy, on random data and random r."""
    rng = np.random.default_rng(3)
    ft = rng.standard_normal(4000) + 1j * rng.standard_normal(4000)
    for _ in range(50):
        r = rng.uniform(40, 3900)
        m = int(rng.choice([8, 16, 32]))
        lo = math.floor(r + 1e-15) + 1 - m // 2
        c = np.sinc((r % 1) - np.arange(-(m // 2) + 1, m // 2 + 1)) * \
            np.exp(1j * np.pi * ((r % 1) - np.arange(-(m // 2) + 1, m // 2 + 1)))
        want = np.sum(np.conj(c) * ft[lo:lo + m])
        assert abs(O.fourier_interp(r, ft, m) - want) < 1e-12 * max(1, abs(want))


def test_finterp_fft_matches_multi():
    ft = _cosine_ft(12400.0)
    rs = 12400 + np.arange(20) / 20
    v1 = np.array([O.fourier_interp(r, ft, 16) for r in rs])
    v2 = O.finterp_fft(12400, 1, 20, ft, 16)
    assert np.allclose(v1, v2, rtol=1e-5, atol=1e-7)


def test_pow2_and_smooth():
    for n, w in GOLD["next_pow_of_2"]["cases"]:
        assert O.next_pow_of_2(n) == w
    with pytest.raises(ValueError):
        O.next_pow_of_2(0)
    assert O.is_smooth(1) and O.is_smooth(2) and O.is_smooth(210)
    assert not O.is_smooth(11) and not O.is_smooth(22) and not O.is_smooth(13)
    for n, w in GOLD["next_smooth"]["cases"]:
        assert O.next_smooth(n) == w
    assert O.is_smooth(O.next_smooth(12345)) and O.next_smooth(12345) >= 12345
    needs = GOLD["next_smooth"]["needs"]
    assert all(O.next_smooth(n) / n < 1.02 for n in needs)
    assert max(O.next_pow_of_2(n) / n for n in needs) > 1.9


def test_finterp_fft_padding_independent():
    ft = _cosine_ft(12400.0)
    need = (4 + 16) * 8
    a = O.finterp_fft(12400, 4, 8, ft, 16, fftlen=O.next_pow_of_2(need))
    b = O.finterp_fft(12400, 4, 8, ft, 16, fftlen=O.next_smooth(need))
    assert O.next_smooth(need) != O.next_pow_of_2(need)
    assert np.max(np.abs(a - b)) / np.max(np.abs(a)) < 1e-13
    with pytest.raises(ValueError):
        O.finterp_fft(12400, 4, 8, ft, 16, fftlen=need - 1)


def test_fft_vs_numpy():
    rng = np.random.default_rng(0)
    for n in (1, 2, 3, 5, 7, 12, 30, 64, 77, 210, 1000, 1024):
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        assert np.allclose(O.fft(x, -1), np.fft.fft(x), atol=1e-10 * n)
        assert np.allclose(O.fft(x, +1), np.fft.ifft(x) * n, atol=1e-10 * n)


def test_finterp_direct_is_exact_kernel():
    ft = _cosine_ft(12400.0)
    for m in (8, 16, 32):
        got = O.finterp_direct(12395.137, 41, 0.25, ft, m)
        want = np.array([O.fourier_interp(12395.137 + k * 0.25, ft, m) for k in range(41)])
        assert np.max(np.abs(got - want)) / np.max(np.abs(want)) < 1e-13
    ex = O.finterp_direct(12400.0, 1, 1.0, ft, 16)[0]
    assert np.isfinite(ex.real) and np.isfinite(ex.imag)
    assert abs(ex - O.fourier_interp(12400.0, ft, 16)) <= 1e-12 * abs(ex)
    with pytest.raises(ValueError):
        O.finterp_direct(12400.0, 4, 0.25, ft, 15)


@pytest.mark.parametrize("nb", [16, 20, 24, 30, 40, 60, 64, 120, 2, 6])
def test_irfft_convention_numpy(nb):
    """brfft == nbins * numpy.fft.irfft, which ignores Im of DC and Nyquist;
    and == the explicitly Hermitian-extended inverse DFT."""
    rng = np.random.default_rng(nb)
    H = nb // 2
    X = rng.standard_normal((H + 1, 7)) + 1j * rng.standard_normal((H + 1, 7))
    got = O.brfft_rows(X, nb)
    assert np.allclose(got, np.fft.irfft(X, nb, axis=0) * nb, atol=1e-12)
    Xc = X.copy()
    Xc[0] = Xc[0].real
    Xc[-1] = Xc[-1].real
    full = np.concatenate([Xc, np.conj(Xc[1:-1][::-1])], axis=0)
    assert np.allclose(got, np.fft.ifft(full, axis=0).real * nb, atol=1e-12)


def test_brfft_decimated_rows():
    """Rows 0, k, 2k, .. of the stack (the stride-k view) transform like a gather."""
    rng = np.random.default_rng(5)
    X = rng.standard_normal((61, 9)) + 1j * rng.standard_normal((61, 9))
    X[0] = 0
    for k in (2, 3, 4, 5, 6):
        Hk = 60 // k
        got = O.brfft_rows(X, 2 * Hk, k)
        want = np.fft.irfft(X[0:Hk * k + 1:k], 2 * Hk, axis=0) * (2 * Hk)
        assert np.allclose(got, want, atol=1e-12)


# ---- directinterp.jl ----------------------------------------------------------

def test_trial_grid_rational():
    assert O.trial_grid_rational(0.5 / 60) == (1, 120)
    assert O.trial_grid_rational(0.5 / 32) == (1, 64)
    assert O.trial_grid_rational(0.5 / 8) == (1, 16)
    assert O.trial_grid_rational(0.3 / 7) == (3, 70)
    assert O.trial_grid_rational(math.pi / 100) is None
    assert O.trial_grid_rational(0.0) is None


@pytest.mark.parametrize("r_lo", [10010.0, 214.748364, 5000.25, 4000.0, 67.108864])
def test_plan_floor_exact_rational(r_lo):
    """floor(h*r_t) from the plan's exact integer state equals exact rational
    arithmetic: Fraction(fl(h*r_lo)) + t*h*pnum/q (fl = the FP64 product the
    reference forms), and dr matches the tabulated phase."""
    nh, m = 12, 16
    pl = O.Plans(nh, m, 0.5, r_lo, False)
    lodr = Fraction(1, 2 * nh)
    for h in (1, 2, 5, 7, 12):
        info = pl.info(h)
        row, carry, W, A = pl.tables(h)
        hr = Fraction(float(h) * r_lo)
        for t in list(range(0, 300, 7)) + [10 ** 6 + 3, 123456789]:
            res, qint = pl.chunk_state(h, t)
            exact = hr + t * h * lodr
            p = row[res] - 1
            fl = info["rfloor0"] + qint + int(carry[p])
            assert fl == math.floor(exact), (h, t)
            dr = exact - math.floor(exact)
            want = 1.0 / (float(dr) - (np.arange(1, m + 1) - m // 2)) if dr != 0 else None
            if want is not None:
                assert np.allclose(W[p], np.float32(want), rtol=2e-7)


def test_group_form_equals_pertrial_bitexact():
    amps = noise_amps(1 << 16, 7)
    for f64 in (False, True):
        for (t0, n) in ((0, 256), (37, 100), (5, 51), (1000, 1)):
            pl = O.Plans(32, 32, 0.5, 2010.0, f64)
            a = pl.fill_chunk(amps, 1 << 17, t0, n)
            b = pl.fill_chunk(amps, 1 << 17, t0, n, pertrial=True)
            assert np.array_equal(a, b)
            assert np.count_nonzero(a) > 0


def test_interp_tabulated_vs_exact():
    """Tabulated interpolation vs fourier_interp: F64 < 1e-7, F32 < 1e-6
    (test/test_search.jl:201-232), re-hosted on synthetic noise."""
    N = 1 << 20
    amps = noise_amps(N // 2, 8)
    ft = amps.astype(np.complex128)
    nh, m, rstart, n = 60, 32, 5000.0, 256
    for f64, tol in ((True, 1e-7), (False, 1e-6)):
        X = O.Plans(nh, m, 0.5, rstart, f64).fill_chunk(amps, N, 0, n)
        worst = 0.0
        for h in (1, 2, 7, 15, 30, 59, 60):
            for j in (0, 1, 96, n - 1):
                r = h * (rstart + j * (0.5 / nh))
                v = O.fourier_interp(r, ft, m)
                worst = max(worst, abs(X[h, j] - v) / abs(v))
        assert worst < tol, (f64, worst)


def test_interp_chunk_invariance_bitexact():
    """Global-trial anchoring: chunks 128/100/37/51 reproduce one long chunk
    exactly (test/test_search.jl:234-276)."""
    N = 1 << 20
    amps = noise_amps(N // 2, 9)
    pl = O.Plans(32, 32, 0.5, 10010.0, False)
    big = pl.fill_chunk(amps, N, 0, 384)
    mis = 0
    for nsm in (128, 100, 37, 51):
        mis += nsm % 32 != 0
        t0 = 0
        while t0 + nsm <= 384:
            small = pl.fill_chunk(amps, N, t0, nsm)
            for h in (1, 5, 32):
                assert np.array_equal(small[h], big[h, t0:t0 + nsm])
            t0 += nsm
    assert mis == 3


def test_row_guard_zeroes_whole_rows():
    """A harmonic that runs off the amplitudes is zero for the whole chunk."""
    N = 1 << 16
    amps = noise_amps(N // 2, 2)
    X = O.Plans(8, 16, 0.5, 4000.0, False).fill_chunk(amps, N, 0, 512)
    for h in range(1, 9):
        top = h * (4000.0 + 511 * 0.5 / 8) + 8
        if top >= N // 2:
            assert not np.any(X[h])
        else:
            assert np.all(X[h] != 0)


# ---- metric ----------------------------------------------------------------------

def test_boxcar_widths_and_ladder():
    g = GOLD["boxcar_widths"]
    assert O.boxcar_widths(64) == g["nbins64"]
    assert O.boxcar_widths(4) == g["nbins4"]
    w = O.boxcar_widths(64)
    assert all(w[i + 1] == max(math.floor(1.5 * w[i]), w[i] + 1) for i in range(len(w) - 1))
    assert O.boxcar_widths(64, maxfrac=0.5)[-1] <= 32
    solo = O.make_params(nharms=60)
    assert O.ladder_boxcar_widths(120, 1, solo) == O.boxcar_widths(120)
    p = O.make_params(nharms=60, decimations=list(range(1, 7)))
    banks = {k: O.ladder_boxcar_widths(2 * (60 // k), k, p) for k in range(1, 7)}
    assert banks[1] == GOLD["ladder"]["bank_k1"] and banks[6] == GOLD["ladder"]["bank_k6"]
    assert all(1 not in banks[k] for k in range(2, 7))
    assert all(max(banks[k]) <= 6 for k in range(1, 6))
    assert max(banks[6]) == max(O.boxcar_widths(20))
    sparse = O.make_params(nharms=60, decimations=[1, 6])
    assert O.ladder_boxcar_widths(120, 1, sparse) == O.boxcar_widths(120)
    assert O.ladder_boxcar_widths(120, 1, O.make_params(nharms=60, decimations=[1, 2, 4])) != \
        O.boxcar_widths(120)


def test_median_vs_numpy():
    rng = np.random.default_rng(1)
    for n in (1, 2, 3, 16, 20, 119, 120, 8160, 8192):
        v = rng.standard_normal(n)
        assert O.median(v) == np.median(v)
        vi = np.round(v * 3)            # heavy ties
        assert O.median(vi) == np.median(vi)


def _numpy_block_sigma(M, cap, center="zero"):
    nbins, n = M.shape
    if nbins * n <= cap:
        s = M.T.reshape(-1)
    else:
        ncol = min(n, max(1, cap // nbins))
        parts, ns = [], 0
        for c in range(1, ncol + 1):
            j = 1 + ((c - 1) * n) // ncol
            take = min(nbins, cap - ns)
            if take == 0:
                break
            parts.append(M[:take, j - 1])
            ns += take
        s = np.concatenate(parts)
    if center == "median":
        s = np.abs(s - np.median(s))
    else:
        s = np.abs(s)
    return 1.4826 * np.median(s)


@pytest.mark.parametrize("nbins,n", [(120, 2048), (60, 2048), (40, 1000), (30, 2048), (24, 341),
                                     (20, 2048), (16, 100), (120, 37)])
def test_block_sigma_vs_numpy(nbins, n):
    rng = np.random.default_rng(nbins * 7 + n)
    M = rng.standard_normal((nbins, n))
    for center in ("zero", "median"):
        assert O.block_sigma(M, 8192, center) == _numpy_block_sigma(M, 8192, center)
    assert O.block_sigma(np.zeros((nbins, n))) == 0.0


def _naive_metric(prof, widths, invsigma):
    nb = len(prof)
    z = prof - np.median(prof)
    best = -np.inf
    for w in widths:
        for p in range(nb):
            s = sum(z[(p + i) % nb] for i in range(w)) * invsigma / math.sqrt(w)
            best = max(best, s)
    return best


def test_boxcar_exact_vs_naive():
    rng = np.random.default_rng(20260811)
    for nb in (16, 20, 64, 120):
        ws = O.boxcar_widths(nb)
        for _ in range(5):
            prof = rng.standard_normal(nb)
            prof[5:9] += 3
            got = O.boxcar_exact(prof, ws, 0.7)
            assert abs(got - _naive_metric(prof, ws, 0.7)) < 1e-12 * abs(got)


def test_boxcar_best_width():
    nb = 64
    for w in O.boxcar_widths(nb):
        prof = np.zeros(nb)
        prof[9:9 + w] = 1.0
        bw, ducy = O.boxcar_best_width(prof)
        assert bw == w and ducy == pytest.approx(w / nb)
    prof = np.zeros(nb)
    prof[62:64] = 1
    prof[0:2] = 1
    assert O.boxcar_best_width(prof)[0] == GOLD["boxcar_best_width"]["wrap_expected"]
    prof = np.zeros(nb)
    prof[19:25] = 1
    w0 = O.boxcar_best_width(prof)[0]
    assert O.boxcar_best_width(1e6 * prof)[0] == w0 and O.boxcar_best_width(prof + 37)[0] == w0

    def naive_bw(prof):
        z = prof - np.median(prof)
        best, bw = -np.inf, 0
        for w in O.boxcar_widths(nb):
            for p in range(nb):
                s = sum(z[(p + i) % nb] for i in range(w)) / math.sqrt(w)
                if s > best:
                    best, bw = s, w
        return bw

    rng = np.random.default_rng(20260811)
    for w in (1, 3, 9, 13):
        for _ in range(5):
            prof = 0.05 * rng.standard_normal(nb)
            prof[29:29 + w] += 3.0
            assert O.boxcar_best_width(prof)[0] == naive_bw(prof)
    for _ in range(20):
        prof = rng.standard_normal(nb)
        assert O.boxcar_best_width(prof)[0] == naive_bw(prof)


def test_metric_invariants():
    nb = 64
    ramp = np.arange(1.0, nb + 1).reshape(nb, 1)
    m1 = O.snr_metrics(ramp)[0]
    assert m1 == pytest.approx(O.snr_metrics(ramp * 7.0)[0], rel=1e-12) and m1 > 0
    assert O.snr_metrics(np.full((nb, 1), 3.0))[0] == 0.0
    rng = np.random.default_rng(1234)
    noise = rng.standard_normal(nb)
    sn = O.snr_metrics(noise.reshape(nb, 1))[0]
    sig = noise.copy()
    sig[29] += 20.0
    ss = O.snr_metrics(sig.reshape(nb, 1))[0]
    assert ss > sn + 10 and 0 < sn < 8 and ss > 15
    small = rng.standard_normal((nb, 4))
    assert np.array_equal(O.snr_metrics(small), O.snr_metrics(small, sigma_samples=8192))


def test_gate_batched_close_to_scalar():
    """F32 batched gate within 1e-3*medmargin of the scalar zero-baseline scan;
    the < 128-column tail takes the scalar path exactly (test_search.jl:345-373)."""
    rng = np.random.default_rng(4)
    nb, n = 120, 500
    profs = rng.standard_normal((nb, n))
    ws = O.boxcar_widths(nb)
    sig = O.block_sigma(profs)
    got = O.boxcar_metrics(profs, ws, 1 / sig, medcut=1e300)   # gate only, nothing rescanned
    want = np.array([O.boxcar_scan(profs[:, j], ws, 1 / sig, 0.0) for j in range(n)])
    assert np.max(np.abs(got - want)) < 1e-3 * 2.0
    assert np.array_equal(got[384:], want[384:])


# ---- drivers ----------------------------------------------------------------------

def test_chunk_metrics_vs_block_metrics_direct():
    """Optimised path vs the exact-kernel reference: < 1e-8 (F64 weights), < 1e-6
    (F32), and the :fft reference much further away (test_search.jl:155-199)."""
    N = 1 << 19
    amps = noise_amps(N // 2, 11)
    p = O.make_params(nharms=32, m=32)
    rstart, n = 10010.0, 256
    rf = rstart + np.arange(n) * (0.5 / 32)
    ref = O.block_metrics(p, amps, N, rf, "direct")
    opt = O.chunk_metrics(p, amps, N, rstart, n, True)
    rel = np.max(np.abs(opt - ref)) / np.max(np.abs(ref))
    assert rel < 1e-8
    opt32 = O.chunk_metrics(p, amps, N, rstart, n, False)
    assert np.max(np.abs(opt32 - ref)) / np.max(np.abs(ref)) < 1e-6
    reffft = O.block_metrics(p, amps, N, rf, "fft")
    assert np.max(np.abs(opt - reffft)) / np.max(np.abs(reffft)) > 1e3 * rel


def test_decimation_equals_native_fold():
    """Fold k of the 60-harmonic stack == a native Hk-harmonic search at k*rf
    (test_search.jl:445-485)."""
    N = 1 << 19
    amps = noise_amps(N // 2, 12)
    for k in (2, 3, 4):
        Hk = 60 // k
        p = O.make_params(nharms=60, m=32, decimations=[1, k])
        rf = 5000.0 + np.arange(64) * (0.5 / 60)
        pn = O.make_params(nharms=Hk, m=32)
        ref = O.block_metrics(pn, amps, N, k * rf, "direct",
                              widths=O.ladder_boxcar_widths(2 * Hk, k, p))
        got = O.decim_metrics(p, amps, N, 5000.0, 64, k, True)
        assert len(got) == 64
        assert np.max(np.abs(got - ref)) / np.max(np.abs(ref)) < 1e-8


def test_search_gate_equals_exact(small_spectrum):
    amps, N, dt = small_spectrum
    T = N * dt
    dec = [1, 2, 3, 4]
    total = math.floor((15 * T - 12 * T) / (0.5 / 60)) + 1
    a = O.search_region(O.make_params(nharms=60, m=32, decimations=dec), amps, N, 12 * T, total,
                        512, threshold=6.0)
    b = O.search_region(O.make_params(nharms=60, m=32, decimations=dec,
                                      boxcar_medmargin=math.inf), amps, N, 12 * T, total, 512,
                        threshold=6.0)
    assert len(a) > 0 and np.array_equal(a, b)


def test_search_detects_and_thread_invariant(small_spectrum):
    amps, N, dt = small_spectrum
    T = N * dt
    p = O.make_params(nharms=32, m=16, decimations=[1, 2, 3, 4])
    total = math.floor((30 * T - 5 * T) / (0.5 / 32)) + 1
    a = O.search_region(p, amps, N, 5 * T, total, 2048, threshold=8.0, nthreads=1)
    b = O.search_region(p, amps, N, 5 * T, total, 2048, threshold=8.0, nthreads=4)
    assert np.array_equal(a, b)
    best = a[np.argmax(a["metric"])]
    t, k = int(best["t"]), int(best["k"])
    i0 = (t // 2048) * 2048
    r = k * (5 * T + i0 * (0.5 / 32) + (t - i0) * (0.5 / 32))
    assert abs(r / T - 13.7) < 1 / T


def test_search_region_sigmas_layout():
    """The sigmas output of the oracle's search ([chunk][fold], fold 1 then the
    decimations > 1): every chunk's row equals that chunk searched alone, and
    each entry is 1.4826 x numpy's median |x| of that chunk's fold profiles
    built from the oracle's own interpolation rows and c2r."""
    N = 1 << 16
    amps = noise_amps(N // 2, 41)
    p = O.make_params(nharms=8, m=16, numbetween=16, threshold=5.0, decimations=(1, 2, 3))
    r_lo, Nprof = 900.0, 256
    total = 5 * Nprof - 17                      # a ragged last chunk
    s_all = np.zeros((5, 3))
    O.search_region(p, amps, N, r_lo, total, Nprof, sigmas=s_all)
    assert np.all(s_all > 0)
    for c in range(5):
        s_one = np.zeros((5, 3))
        O.search_region(p, amps, N, r_lo, total, Nprof, c_begin=c, c_end=c + 1, sigmas=s_one)
        assert np.array_equal(s_one[c], s_all[c]) and not s_one[np.arange(5) != c].any()
    # every fold's profiles hold at most 16 x 256 = 4096 <= 8192 samples, so
    # sigma is 1.4826 x numpy's median |x| over all of them (search.jl:490-566)
    pl = O.Plans(8, 16, 0.5, r_lo)
    lodr = 0.5 / 8
    for c in range(5):
        n = min(Nprof, total - c * Nprof)
        X = pl.fill_chunk(amps, N, c * Nprof, n)
        for d, k in enumerate((1, 2, 3)):
            P = O.brfft_rows(X, 2 * (8 // k), k)
            rstart = r_lo + c * Nprof * lodr
            nv = sum(1 for j in range(n) if k * (rstart + j * lodr) < N / 2)
            want = 1.4826 * float(np.median(np.abs(P[:, :nv])))
            assert s_all[c, d] == pytest.approx(want, rel=1e-14), (c, k)
