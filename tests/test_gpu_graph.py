"""GPU checks of the graph replay of a search's launch sequence
(cs_set_tuning(CS_TUNE_GRAPHS, 1)).

The graph is replayed several times with other eager searches and
synchronisations between the replays; every replay must return exactly the
eager search's raw candidates, also after an argument change forces a new
capture and after the change is undone.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("mode", ["fused", "plain", "fused_valu"])
def test_graph_replays_equal_eager(small_spectrum, mode):
    import coherentsearch as cs
    amps, N, dt = small_spectrum
    T = N * dt
    p = cs.SearchParams(nharms=32, m=16, decimations=cs.decimation_set(32, 4))
    ft = cs.FFTFile.from_arrays(amps, N, dt)
    eager, replay = cs.Engine(p, 512), cs.Engine(p, 512)
    for e in (eager, replay):
        e.set_mode(mode)
        e.set_spectrum(ft)
    replay.set_tuning(graphs=True)
    r_lo, r_hi = 5.0 * T, 30.0 * T
    base = {thr: eager.search_raw(r_lo, r_hi, thr) for thr in (6.0, 7.0)}
    assert len(base[6.0]) > len(base[7.0]) > 0
    for i, thr in enumerate((6.0, 6.0, 6.0, 7.0, 7.0, 6.0, 6.0)):
        _say(f"graph replay mode={mode} call={i} threshold={thr}")
        got = replay.search_raw(r_lo, r_hi, thr)
        assert np.array_equal(got, base[thr]), (i, thr)
        assert replay.timings()["total"] > 0.0
        # eager work and a synchronisation between replays
        assert np.array_equal(eager.search_raw(r_lo, r_hi, thr), base[thr])
    replay.set_tuning(graphs=False)
    assert np.array_equal(replay.search_raw(r_lo, r_hi, 6.0), base[6.0])
    assert replay.timings()["rescan"] > 0.0
