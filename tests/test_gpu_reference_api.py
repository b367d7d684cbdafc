"""GPU checks of the reference-path building blocks exposed by the engine:
reference_profiles (cs_reference_profiles), snr_metrics (cs_snr_metrics) and
fourier_interp (cs_fourier_interp), each against the oracle's restatement.

Bars: FP64 on both sides; profiles within 1e-10 (:direct) and 1e-9 (:fft) of
the largest value, the same bars as block_metrics; metrics within 1e-12
relative (exact medians, the same prefix-sum order); interpolated amplitudes
within 1e-12 of the largest.
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
def spec():
    import coherentsearch as cs
    from coherentsearch.synth import Pulsar, make_spectrum
    N, dt = 1 << 18, 1e-3
    amps = make_spectrum(N, dt, 321, [Pulsar(7.3, 0.06, 25.0)])
    return cs, amps, N, dt


@pytest.mark.parametrize("kernel,tol", [("direct", 1e-10), ("fft", 1e-9)])
def test_reference_profiles(spec, kernel, tol):
    import oracle as O
    cs, amps, N, dt = spec
    T = N * dt
    _say(f"reference_profiles {kernel}")
    p = cs.SearchParams(nharms=16, m=16, numbetween=16)
    op = O.make_params(nharms=16, m=16, numbetween=16)
    rf = 7.3 * T + np.linspace(-2.0, 2.0, 57)
    got = cs.reference_profiles(cs.FFTFile.from_arrays(amps, N, dt), rf, p, kernel=kernel)
    want = O.reference_profiles(op, amps, N, rf, kernel)
    assert got.shape == want.shape == (32, 57)
    assert np.max(np.abs(got - want)) <= tol * np.max(np.abs(want))


@pytest.mark.parametrize("center", ["zero", "median"])
@pytest.mark.parametrize("samples", [None, 8192, 1000])
def test_snr_metrics(spec, center, samples):
    import oracle as O
    cs, amps, N, dt = spec
    _say(f"snr_metrics center={center} samples={samples}")
    rng = np.random.default_rng(5)
    nbins, L = 40, 300                      # 12000 samples: more than one LDS sort holds
    P = rng.standard_normal((nbins, L))
    P[10:14, 7] += 6.0                      # one pulse
    eng = cs.Engine(cs.SearchParams(), 2048)
    got = eng.snr_metrics(P, sigma_samples=samples, sigma_center=center)
    want = O.snr_metrics(P, sigma_samples=samples, sigma_center=center)
    assert np.max(np.abs(got - want) / np.maximum(np.abs(want), 1e-300)) <= 1e-12
    assert int(np.argmax(got)) == 7
    w = [1, 3, 5, 8]
    assert np.allclose(eng.snr_metrics(P, widths=w), O.snr_metrics(P, widths=w), rtol=1e-12, atol=0)
    with pytest.raises(ValueError):
        eng.snr_metrics(P, sigma_center="mean")
    with pytest.raises(ValueError):
        eng.snr_metrics(np.zeros((200, 3)))        # more than 128 bins


def test_fourier_interp(spec):
    import oracle as O
    cs, amps, N, dt = spec
    _say("fourier_interp")
    ft = cs.FFTFile.from_arrays(amps, N, dt)
    rng = np.random.default_rng(9)
    r = np.concatenate([[100.0, 2000.0, 57.999999999999999, 300.25],
                        rng.uniform(40.0, 1.2e5, 200)])
    for m in (16, 32, 64):
        got = cs.fourier_interp(r, ft, m)
        want = np.array([O.fourier_interp(float(x), amps, m) for x in r])
        assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))
    assert isinstance(cs.fourier_interp(300.25, ft, 16), complex)
    with pytest.raises(ValueError):
        cs.fourier_interp(-1.0, ft, 16)
    with pytest.raises(ValueError):
        cs.fourier_interp(10.0, ft, 15)
    from coherentsearch._lib import DimensionMismatch
    with pytest.raises(DimensionMismatch):
        cs.fourier_interp(3.0, ft, 16)              # window starts below bin 0
