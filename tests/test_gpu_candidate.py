"""GPU checks of the post-detection steps (cs_candidate_profiles,
cs_measure_ducy) against the oracle's candidate_profile / measure_ducy.

Bars: profile lengths equal; profiles within 1e-12 of the largest bin (FP64
Eqn-30 and DFT on both sides, different summation orders); duty cycles equal
(the best width is an index).
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
    N, dt = 1 << 21, 5e-4
    T = N * dt
    amps = make_spectrum(N, dt, 4242, [Pulsar(10012.33125 / T, 0.06, 40.0),
                                       Pulsar(27.7, 0.03, 30.0)])
    eng = cs.Engine(cs.SearchParams(nharms=32, m=16), 2048)
    eng.set_spectrum(cs.FFTFile.from_arrays(amps, N, dt))
    return cs, eng, amps, N, T


def test_profiles_match_oracle(setup):
    import oracle as O
    cs, eng, amps, N, T = setup
    rng = np.random.default_rng(3)
    r = np.concatenate([[10012.33125, 10012.33125, 27.7 * T, 97.0, 10.5, 40.0],
                        rng.uniform(50.0, 9.0e5, 40)])
    nh = np.concatenate([[32, 200, 60, 12, 4, 1], rng.integers(1, 80, 40)]).astype(np.int32)
    for m in (64, 32):
        _say(f"candidate profiles m={m} n={len(r)}")
        got = eng.candidate_profiles(r, nh, m)
        for i in range(len(r)):
            want = O.candidate_profile(amps, N, float(r[i]), int(nh[i]), m)
            assert len(got[i]) == len(want), (i, r[i], nh[i])
            scale = max(np.max(np.abs(want)), 1e-300)
            assert np.max(np.abs(got[i] - want)) <= 1e-12 * scale, (i, r[i], nh[i])
    assert len(got[1]) < 400 and len(got[4]) == 2


def test_measure_ducy_matches_oracle(setup):
    import oracle as O
    cs, eng, amps, N, T = setup
    p = cs.SearchParams(nharms=32, m=16)
    rng = np.random.default_rng(4)
    r = np.concatenate([[10012.33125, 27.7 * T, 0.0, -5.0, 300.0, 10.5],
                        rng.uniform(50.0, 9.0e5, 60)])
    nh = np.concatenate([[32, 60, 8, 8, 0, 4], rng.integers(1, 70, 60)]).astype(np.int32)
    _say(f"measure_ducy n={len(r)}")
    got = eng.measure_ducy(r, nh)
    for i in range(len(r)):
        want = O.measure_ducy(amps, N, float(r[i]), int(nh[i]), p.boxcar_fsp, p.boxcar_maxfrac)
        if math.isnan(want):
            assert math.isnan(got[i]), i
        else:
            assert got[i] == want, (i, r[i], nh[i], got[i], want)


def test_errors_and_report(setup, tmp_path):
    cs, eng, amps, N, T = setup
    _say("candidate errors and report")
    with pytest.raises(ValueError):
        eng.candidate_profiles([100.0], [0])
    with pytest.raises(ValueError):
        eng.candidate_profiles([100.0], [4], m=33)
    with pytest.raises(ValueError):
        eng.candidate_profiles([100.0], [600], m=4)   # more than 512 harmonics
    ft = cs.FFTFile.from_arrays(amps, N, 5e-4)
    p = cs.SearchParams(nharms=32, m=16)
    cands = [cs.Candidate(27.7, 20.0, 27.7 * T, 32), cs.Candidate(1.0, 30.0, T, 16),
             cs.Candidate(2.0, 9.0, 0.0, 8)]
    rep = cs.report_candidates(ft, cands, p, ncands=2, engine=eng)
    assert [c.freq for c in rep] == [1.0, 27.7]
    assert all(0.0 < c.ducy <= 0.5 for c in rep)
    # the 3 % duty pulsar: its best width is a few bins of 64
    assert rep[1].ducy <= 0.1
    out = tmp_path / "a.cohout"
    cs.write_candidates(rep, str(out), 8.0)
    lines = out.read_text().splitlines()
    assert len(lines) == 3 and lines[2].startswith("2 ")
