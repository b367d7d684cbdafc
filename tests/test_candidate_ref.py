"""CPU checks of the post-detection steps: the oracle's candidate_profile /
measure_ducy (oracle/oracle.py), and the host code of rotate_to_peak, the
candidate ranking and the `.cohout` table (coherentsearch/candidate.py).

Pins (test_candidate.jl:13-61, re-hosted on a synthetic spectrum because the
reference's example file is not in the tree):
* rotate_to_peak centres the maximum (the reference's own cases);
* at an integer r every Eqn-30 weight but the bin's own vanishes, so the
  profile is exactly numpy's irfft of the bins r h;
* with a matched kernel (m = 32) and a fine grid (numbetween = 256) the
  profile agrees with reference_profiles to 1e-3 (the reference's pin);
* Nyquist capping: the fold stops at the first harmonic at or above N/2.
"""
import math
import os

import numpy as np
import pytest

import oracle as O

from coherentsearch import Candidate, candidate
from coherentsearch.synth import Pulsar, make_spectrum


@pytest.mark.parametrize("n", [8, 9, 64, 65])
def test_rotate_to_peak_centres_the_maximum(n):
    prof = np.arange(1.0, n + 1)
    rp = candidate.rotate_to_peak(prof)
    assert len(rp) == n
    assert int(np.argmax(rp)) == n // 2          # Julia's n ÷ 2 + 1, 1-based
    assert np.array_equal(np.sort(rp), np.sort(prof))
    assert len(candidate.rotate_to_peak([])) == 0


@pytest.fixture(scope="module")
def psr_spectrum():
    N, dt = 1 << 21, 5e-4
    T = N * dt
    rbest = 10012.33125
    amps = make_spectrum(N, dt, 4242, [Pulsar(rbest / T, 0.06, 40.0)])
    return amps, N, rbest


def test_integer_r_is_the_irfft_of_the_bins():
    rng = np.random.default_rng(8)
    amps = ((rng.standard_normal(4096) + 1j * rng.standard_normal(4096)) / math.sqrt(2)).astype(np.complex64)
    r, nh = 97.0, 12
    got = O.candidate_profile(amps, 8192, r, nh, m=32)
    stack = np.zeros(nh + 1, dtype=np.complex128)
    stack[1:] = amps[(np.arange(1, nh + 1) * 97)].astype(np.complex128)
    want = np.fft.irfft(stack, 2 * nh)
    assert np.max(np.abs(got - want)) <= 1e-15 * np.max(np.abs(want))


def test_profile_pinned_to_reference_profiles(psr_spectrum):
    amps, N, rbest = psr_spectrum
    nh = 32
    p = O.make_params(nharms=nh, m=32, numbetween=256)
    ref = O.reference_profiles(p, amps, N, [rbest], "fft")[:, 0]
    cand = O.candidate_profile(amps, N, rbest, nh, m=32)
    assert len(cand) == 2 * nh
    assert np.max(np.abs(cand - ref)) / np.max(np.abs(ref)) < 1e-3
    wide = O.candidate_profile(amps, N, rbest, nh)
    assert len(wide) == 2 * nh
    assert int(np.argmax(candidate.rotate_to_peak(wide))) == nh


def test_nyquist_capping(psr_spectrum):
    amps, N, rbest = psr_spectrum
    capped = O.candidate_profile(amps, N, rbest, 200)
    H = len(capped) // 2
    assert len(capped) % 2 == 0 and H < 200
    assert rbest * H < N // 2 and rbest * (H + 2) > N // 2
    # a window that starts below bin 0 stops the fold at once: two zero bins
    z = O.candidate_profile(amps, N, 10.5, 4, m=64)
    assert np.array_equal(z, np.zeros(2))
    assert math.isnan(O.measure_ducy(amps, N, 0.0, 4))
    assert math.isnan(O.measure_ducy(amps, N, 100.0, 0))
    assert O.measure_ducy(amps, N, 10.5, 4) == 0.5   # widths of 2 bins start at 1
    with pytest.raises(ValueError):
        O.candidate_profile(amps, N, rbest, 0)
    with pytest.raises(ValueError):
        O.candidate_profile(amps, N, rbest, 4, m=33)


def test_cohout_table_and_paths(tmp_path):
    cands = [Candidate(10.0123, 25.5, 10012.3, 32, 0.0625),
             Candidate(123.456789, 9.123, 123456.0, 8, float("nan"))]
    lines = candidate.candidate_lines(cands)
    assert lines[0] == "#       'S/N'      Frequency (Hz)        Period (ms)    #Harm  Ducy(%)"
    assert lines[1] == ("1        25.50     10.012300000000     99.877151104142    32     6.25")
    assert lines[2] == ("2         9.12    123.456789000000      8.100000073710     8        -")
    out = tmp_path / "x.cohout"
    candidate.write_candidates(cands, str(out), 8.0)
    assert out.read_text().splitlines() == lines
    assert candidate.output_path("d/a.fft") == ""
    assert candidate.output_path("d/a.fft", "o.txt") == "o.txt"
    assert candidate.output_path("d/a.fft", "", "", 2) == os.path.join("d", "a.cohout")
    assert candidate.output_path("d/a.fft", "", "out", 1) == os.path.join("out", "a.cohout")


def test_ranking_is_stable_and_trims():
    cs = [Candidate(1.0, 9.0, 1.0, 8), Candidate(2.0, 12.0, 2.0, 8), Candidate(3.0, 9.0, 3.0, 8),
          Candidate(4.0, 9.0, 4.0, 8)]
    assert [c.freq for c in candidate.rank_candidates(cs, 3)] == [2.0, 1.0, 3.0]
    assert [c.freq for c in candidate.rank_candidates(cs, 10)] == [2.0, 1.0, 3.0, 4.0]
    assert candidate.rank_candidates(cs, 0) == []
