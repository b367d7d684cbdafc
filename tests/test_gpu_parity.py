"""GPU parity: libcsgpu.so (through its C ABI) against the CPU oracle on the
same seeded inputs.

Bars: integer work (trial indices, folds, blocks, row-zeroing) bit-exact; the
FP32-weight interpolation bit-exact (same FMA chain); FP64 bridge metrics
(chunk_metrics / decim / block_metrics :direct) within 1e-12 relative;
block_metrics :fft within 1e-9 relative of the oracle's FFT correlation (the
device evaluates the same grid values as direct FP64 sums);
search candidates: identical (t, k) sets away from threshold and S/N within
1e-4 relative (the FP32 profile stage moves the block sigma by ~1e-7).
"""
import math
import sys

import numpy as np
import pytest

from conftest import noise_amps

pytestmark = pytest.mark.gpu


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


def _engine(cs, **kw):
    bs = kw.pop("blocksize", 2048)
    p = cs.SearchParams(**kw)
    return cs.Engine(p, bs), p


def _ft(cs, amps, N, dt=64e-6):
    return cs.FFTFile.from_arrays(amps, N, dt)


@pytest.mark.parametrize("nharms,m,r_lo,t0,n", [
    (32, 32, 10010.0, 0, 256),
    (32, 32, 10010.0, 37, 100),      # misaligned with the 32-trial groups
    (60, 16, 5000.0, 0, 300),
    (60, 16, 5000.25, 1000, 77),
    (8, 16, 67.108864, 0, 2048),      # C1-like grid
    (60, 16, 4000.0, 0, 129),         # integer r_lo: dr == 0 delta phases
    (32, 32, 30000.0, 0, 200),        # top harmonics run off the amplitudes (rows zeroed)
])
@pytest.mark.parametrize("f64", [False, True])
def test_interp_rows_bit_exact(cs, O, nharms, m, r_lo, t0, n, f64):
    _say(f"interp nharms={nharms} m={m} r_lo={r_lo} t0={t0} n={n} f64={f64}")
    N = 1 << 17
    amps = noise_amps(N // 2, seed=nharms + m + n)
    eng, p = _engine(cs, nharms=nharms, m=m)
    eng.set_spectrum(_ft(cs, amps, N))
    got = eng.interp_rows(r_lo, t0, n, weights_f64=f64)
    want = O.Plans(nharms, m, 0.5, r_lo, f64).fill_chunk(amps, N, t0, n)
    assert got.shape == want.shape
    zero_rows_g = [h for h in range(nharms + 1) if not np.any(got[h])]
    zero_rows_w = [h for h in range(nharms + 1) if not np.any(want[h])]
    assert zero_rows_g == zero_rows_w
    assert np.array_equal(got, want), np.max(np.abs(got - want))


@pytest.mark.parametrize("f64", [False, True])
@pytest.mark.parametrize("nharms,m,rstart,n", [(32, 32, 10010.0, 256), (60, 16, 5000.0, 512),
                                               (8, 16, 1000.5, 100)])
def test_chunk_metrics_vs_oracle(cs, O, nharms, m, rstart, n, f64):
    _say(f"chunk_metrics nharms={nharms} m={m} rstart={rstart} n={n} f64={f64}")
    N = 1 << 18
    amps = noise_amps(N // 2, seed=5)
    eng, p = _engine(cs, nharms=nharms, m=m, blocksize=n)
    eng.set_spectrum(_ft(cs, amps, N))
    got = eng.chunk_metrics(rstart, n, weights_f64=f64)
    want = O.chunk_metrics(O.make_params(nharms=nharms, m=m), amps, N, rstart, n, f64)
    rel = np.max(np.abs(got - want)) / np.max(np.abs(want))
    _say(f"  rel={rel:.3e}")
    assert rel < 1e-12


@pytest.mark.parametrize("k", [2, 3, 4, 6])
def test_decim_metrics_vs_oracle(cs, O, k):
    _say(f"decim k={k}")
    N = 1 << 18
    amps = noise_amps(N // 2, seed=9)
    eng, p = _engine(cs, nharms=60, m=32, decimations=[1, k], blocksize=64)
    eng.set_spectrum(_ft(cs, amps, N))
    got = eng.decim_metrics(5000.0, 64, k, weights_f64=True)
    want = O.decim_metrics(O.make_params(nharms=60, m=32, decimations=[1, k]), amps, N, 5000.0,
                           64, k, True)
    assert len(got) == len(want) == 64
    assert np.max(np.abs(got - want)) / np.max(np.abs(want)) < 1e-12


def test_decim_nvalid_near_nyquist(cs, O):
    """Decimated trials past Nyquist are cut at the reference's exact FP64 prefix."""
    N = 1 << 16
    amps = noise_amps(N // 2, seed=3)
    # k * r crosses N/2 = 32768 inside the block
    rstart = 32768 / 3 - 0.7
    eng, p = _engine(cs, nharms=6, m=16, decimations=[1, 3], blocksize=128)
    eng.set_spectrum(_ft(cs, amps, N))
    got = eng.decim_metrics(rstart, 128, 3)
    want = O.decim_metrics(O.make_params(nharms=6, m=16, decimations=[1, 3]), amps, N, rstart, 128, 3)
    assert len(got) == len(want) and 0 < len(want) < 128
    assert np.allclose(got, want, rtol=1e-12, atol=0)


@pytest.mark.parametrize("nharms,m,widths", [(32, 32, None), (16, 16, [1, 2, 3, 4, 6, 9])])
def test_block_metrics_direct_vs_oracle(cs, O, nharms, m, widths):
    N = 1 << 18
    amps = noise_amps(N // 2, seed=21)
    eng, p = _engine(cs, nharms=nharms, m=m)
    eng.set_spectrum(_ft(cs, amps, N))
    rf = 10010.0 + np.arange(256) * (0.5 / nharms)
    got = eng.block_metrics(rf, "direct", widths)
    want = O.block_metrics(O.make_params(nharms=nharms, m=m), amps, N, rf, "direct", widths)
    assert np.max(np.abs(got - want)) / np.max(np.abs(want)) < 1e-10


@pytest.mark.parametrize("nharms,m,widths", [(8, 16, None), (16, 32, None), (32, 16, [1, 2, 3, 4, 6])])
def test_block_metrics_fft_vs_oracle(cs, O, nharms, m, widths):
    """kernel=:fft (the reference's default, search.jl:1036-1040): the device's
    FP64 Eqn. 30 grid values with clamped linear interpolation against the
    oracle's FFT correlation (finterp_fft, fourierinterp.jl:195-214).  The two
    differ only by FFT rounding."""
    N = 1 << 18
    amps = noise_amps(N // 2, seed=23)
    eng, p = _engine(cs, nharms=nharms, m=m)
    eng.set_spectrum(_ft(cs, amps, N))
    rf = 10010.3 + np.arange(200) * (0.37 / nharms)
    got = eng.block_metrics(rf, "fft", widths)
    want = O.block_metrics(O.make_params(nharms=nharms, m=m), amps, N, rf, "fft", widths)
    assert np.max(np.abs(got - want)) / np.max(np.abs(want)) < 1e-9


def test_block_metrics_fft_on_grid_equals_direct(cs):
    """When every harmonic frequency r*h lies on the 1/numbetween grid the
    linear interpolation weight is zero and :fft is Eqn. 30 itself, so it equals
    :direct (a property of the method, independent of the oracle)."""
    N = 1 << 18
    amps = noise_amps(N // 2, seed=29)
    eng, p = _engine(cs, nharms=16, m=16)
    eng.set_spectrum(_ft(cs, amps, N))
    rf = 9000.0 + np.arange(128) / 16.0
    a = eng.block_metrics(rf, "fft")
    d = eng.block_metrics(rf, "direct")
    assert np.max(np.abs(a - d)) / np.max(np.abs(d)) < 1e-12
    # off the grid the two methods differ (linear interpolation error ~1e-2)
    rf2 = rf + 0.5 / 16.0
    assert np.max(np.abs(eng.block_metrics(rf2, "fft") - eng.block_metrics(rf2, "direct"))) > 1e-6


def test_search_block_default_fft(cs, O, small_spectrum):
    """search_block keeps the trials whose :fft metric exceeds the threshold
    (search.jl:1087-1097), on an injected pulsar."""
    amps, N, dt = small_spectrum
    T = N * dt
    p = cs.SearchParams(nharms=16, m=16)
    ft = _ft(cs, amps, N, dt)
    rf = 13.7 * T + (np.arange(-64, 64) * (0.5 / 16))
    cands = cs.search_block(ft, rf, p, threshold=6.0)
    want = O.block_metrics(O.make_params(nharms=16, m=16), amps, N, rf, "fft")
    keep = [float(r) for r, v in zip(rf, want) if v > 6.0]
    assert len(keep) > 0
    assert [c.r for c in cands] == keep
    got = np.array([c.metric for c in cands])
    assert np.allclose(got, want[want > 6.0], rtol=1e-9, atol=0)


@pytest.mark.parametrize("batch,wg,order", [(3, 0, 0), (1, 37, 0), (5, 8, 1), (2, 1, 1), (0, 0, 1)])
def test_fused_launch_shape_invariance(cs, small_spectrum, batch, wg, order):
    """The split-BF16 search returns bit-identical raw candidates whatever the
    batch size and persistent grid: batches of 500-trial blocks cut 512-trial
    super-tiles, whose trials outside the batch are left to the neighbouring
    batch (the chunk-invariance property, test/test_search.jl:234-276)."""
    _say(f"launch shape batch={batch} wg={wg}")
    amps, N, dt = small_spectrum
    T = N * dt
    eng, p = _engine(cs, nharms=32, m=16, decimations=cs.decimation_set(32, 4), blocksize=500)
    eng.set_spectrum(_ft(cs, amps, N, dt))
    eng.set_mode("fused")
    ref = eng.search_raw(5.0 * T, 30.0 * T, 6.0)
    eng.set_tuning(batch_blocks=batch, workgroups=wg, tile_order=order)
    got = eng.search_raw(5.0 * T, 30.0 * T, 6.0)
    nb = eng.timings()["batches"]
    eng.set_tuning(batch_blocks=0, workgroups=0, tile_order=0)
    assert len(ref) > 0
    assert nb > 1 or batch == 0
    assert np.array_equal(ref, got)
    # the plain mode's batches (runs of blocks) give the same list too
    if batch:
        eng.set_mode("plain")
        p0 = eng.search_raw(5.0 * T, 30.0 * T, 6.0)
        eng.set_tuning(batch_blocks=batch)
        p1 = eng.search_raw(5.0 * T, 30.0 * T, 6.0)
        eng.set_tuning(batch_blocks=0)
        assert np.array_equal(p0, p1)


def test_autotune_cache_and_invariance(cs, small_spectrum, tmp_path):
    """cs_autotune picks a launch shape by timing, the tuned search returns the
    default search's raw candidates bit for bit, and the tuning cache keeps one
    line per geometry that a later call reads instead of measuring."""
    _say("autotune")
    amps, N, dt = small_spectrum
    T = N * dt
    eng, p = _engine(cs, nharms=32, m=16, decimations=cs.decimation_set(32, 4), blocksize=500)
    eng.set_spectrum(_ft(cs, amps, N, dt))
    eng.set_mode("fused")
    ref = eng.search_raw(5.0 * T, 30.0 * T, 6.0)
    path = str(tmp_path / "tune.txt")
    b, w = eng.autotune(5.0 * T, 30.0 * T, 6.0, cache_path=path, reps=1)
    assert b >= 0 and 0 <= w <= 65536
    assert np.array_equal(eng.search_raw(5.0 * T, 30.0 * T, 6.0), ref)
    lines = open(path).read().splitlines()
    assert len(lines) == 1 and "\t" in lines[0]
    eng.set_tuning(batch_blocks=0, workgroups=0)
    assert eng.autotune(5.0 * T, 30.0 * T, 6.0, cache_path=path) == (b, w)
    assert open(path).read().splitlines() == lines          # read, not re-measured
    eng.autotune(5.0 * T, 20.0 * T, 6.0, cache_path=path, reps=1)
    assert len(open(path).read().splitlines()) == 2
    assert eng.tuning()["tile_order"] in (0, 1)
    # the plain mode tunes its blocks per batch, with the same results
    eng.set_mode("plain")
    eng.set_tuning(batch_blocks=0, workgroups=0, tile_order=0)
    pref = eng.search_raw(5.0 * T, 30.0 * T, 6.0)
    b, w = eng.autotune(5.0 * T, 30.0 * T, 6.0, cache_path=path, reps=1)
    assert w == 0 and b >= 0
    assert np.array_equal(eng.search_raw(5.0 * T, 30.0 * T, 6.0), pref)
    assert len(open(path).read().splitlines()) == 3


@pytest.mark.parametrize("mode", ["fused", "plain"])
def test_search_files_equals_single_file_runs(cs, mode):
    """Several spectra (three lengths, one repeated, one with a pulsar) through
    one cs_search_files call give, file by file, the raw candidates of
    cs_set_spectrum + cs_search on that file (the multi-file .cohout equals
    separate runs, test/test_cli.jl:65-85)."""
    _say(f"search_files mode={mode}")
    from coherentsearch.synth import Pulsar, make_spectrum
    dt = 64e-6
    specs = [(1 << 17, 41), (1 << 18, 42), (1 << 16, 43), (1 << 18, 44), (1 << 17, 41)]
    fts = []
    for N, seed in specs:
        amps = make_spectrum(N, dt, seed=seed,
                             pulsars=[Pulsar(21.3, 0.06, 20.0)] if seed == 42 else [])
        fts.append(_ft(cs, amps, N, dt))
    p = cs.SearchParams(nharms=32, m=16, decimations=cs.decimation_set(32, 4))
    eng = cs.Engine(p, 1024)
    eng.set_mode(mode)
    jobs = [(ft, 5.0 * ft.T, 40.0 * ft.T) for ft in fts]
    multi = eng.search_files(jobs, 6.5)
    assert len(multi) == len(fts)
    for (ft, r_lo, r_hi), got in zip(jobs, multi):
        eng.set_spectrum(ft)
        want = eng.search_raw(r_lo, r_hi, 6.5)
        assert np.array_equal(got, want)
    assert len(multi[1]) > 0          # the file with the pulsar
    assert np.array_equal(multi[0], multi[4])
    # and through the host mirror, per file as search() does
    lists = cs.search_files(fts, p, lofreq=5.0, hifreq=40.0, threshold=6.5, blocksize=1024,
                            mode=mode)
    one = cs.search(fts[1], p, lofreq=5.0, hifreq=40.0, threshold=6.5, blocksize=1024, mode=mode)
    assert [(c.r, c.metric) for c in lists[1]] == [(c.r, c.metric) for c in one]
    assert any(abs(c.freq - 21.3) < 2.0 / fts[1].T for c in lists[1])


@pytest.mark.parametrize("nharms,maxdecim,blocksize", [(60, 6, 2048), (32, 4, 512), (8, 1, 2048)])
def test_fold_set_kernel_equals_generic(cs, small_spectrum, nharms, maxdecim, blocksize):
    """The split-BF16 kernels specialised on a fold set (every pass inlined,
    unclamped constant-stride row gathers) compute exactly what the kernel
    that dispatches each fold on its length does: raw candidates and every
    (block, fold) pass sigma (precision :f32, so the sigmas are the pass's own,
    from the kernel's sample columns) bit for bit; the split-BF16 probe's
    stacked amplitudes and per-fold bounds too."""
    _say(f"fold-set kernel vs generic nharms={nharms}")
    amps, N, dt = small_spectrum
    T_ = N * dt
    p = cs.SearchParams(nharms=nharms, m=16, decimations=cs.decimation_set(nharms, maxdecim),
                        precision="f32")
    eng = cs.Engine(p, blocksize)
    eng.set_mode("fused")
    eng.set_spectrum(_ft(cs, amps, N, dt))
    r_lo, r_hi = 1.0 * T_, 60.0 * T_
    out = {}
    for generic in (False, True):
        eng.set_tuning(generic=generic)
        assert eng.tuning()["generic"] == int(generic)
        raw = eng.search_raw(r_lo, r_hi, 6.0)
        out[generic] = (raw, eng.last_sigmas(), eng.probe_split_bf16(r_lo, r_hi, 1000, 256))
    assert len(out[False][0]) > 0
    assert np.array_equal(out[False][0], out[True][0])
    assert np.array_equal(out[False][1], out[True][1])
    for a, b in zip(out[False][2], out[True][2]):
        assert np.array_equal(a, b)
    eng.close()


@pytest.mark.parametrize("generic", [False, True])
def test_lds_prefetch_off_equals_on(cs, small_spectrum, generic):
    """The split-BF16 kernel's LDS prefetch of each tile's first conj(A)
    operands (direct-to-LDS loads during the previous tile) changes no result:
    raw candidates, every pass sigma (:f32) and the probe's stacked amplitudes
    and bounds equal those of the same kernel loading them at the
    interpolation (CS_TUNE_NO_PREFETCH), for the fold-set and the generic
    kernel."""
    _say(f"LDS prefetch off vs on, generic={generic}")
    amps, N, dt = small_spectrum
    T_ = N * dt
    p = cs.SearchParams(nharms=60, m=16, decimations=cs.decimation_set(60, 6), precision="f32")
    eng = cs.Engine(p, 2048)
    eng.set_mode("fused")
    eng.set_spectrum(_ft(cs, amps, N, dt))
    eng.set_tuning(generic=generic)
    r_lo, r_hi = 1.0 * T_, 60.0 * T_
    out = {}
    for off in (False, True):
        eng.set_tuning(no_prefetch=off)
        assert eng.tuning()["no_prefetch"] == int(off)
        raw = eng.search_raw(r_lo, r_hi, 6.0)
        out[off] = (raw, eng.last_sigmas(), eng.probe_split_bf16(r_lo, r_hi, 1000, 256))
    assert len(out[False][0]) > 0
    assert np.array_equal(out[False][0], out[True][0])
    assert np.array_equal(out[False][1], out[True][1])
    for a, b in zip(out[False][2], out[True][2]):
        assert np.array_equal(a, b)
    eng.close()


def test_search_files_with_tuning_cache(cs, tmp_path):
    """A sweep whose length classes take their launch shapes from cs_autotune's
    cache (cs_set_tuning_cache) returns exactly the raw candidates of the sweep
    with default shapes; a cache holding non-default shapes for every class
    (one block per batch, a quarter of the grid, round-robin tiles) does too;
    and the context's own knobs are as before the sweep."""
    _say("search_files with a tuning cache")
    from coherentsearch.synth import Pulsar, make_spectrum
    dt = 64e-6
    specs = [(1 << 17, 51), (1 << 18, 52), (1 << 17, 53), (1 << 18, 54)]
    fts = [_ft(cs, make_spectrum(N, dt, seed=sd,
                                 pulsars=[Pulsar(17.9, 0.05, 18.0)] if sd == 52 else []), N, dt)
           for N, sd in specs]
    p = cs.SearchParams(nharms=32, m=16, decimations=cs.decimation_set(32, 4))
    eng = cs.Engine(p, 1024)
    eng.set_mode("fused")
    jobs = [(ft, 5.0 * ft.T, 40.0 * ft.T) for ft in fts]
    want = eng.search_files(jobs, 6.5)
    assert len(want[1]) > 0
    cache = str(tmp_path / "tuning.txt")
    for ft, r_lo, r_hi in jobs[:2]:           # one file of each length
        eng.set_spectrum(ft)
        eng.autotune(r_lo, r_hi, 6.5, cache_path=cache, reps=1)
    lines = open(cache).read().splitlines()
    assert len(lines) == 2
    eng.set_tuning(batch_blocks=0, workgroups=0, tile_order=0)
    eng.set_tuning_cache(cache)
    got = eng.search_files(jobs, 6.5)
    assert all(np.array_equal(a, b) for a, b in zip(got, want))
    # every class forced to an unusual shape through the cache
    odd = [l.split("\t")[0] + "\t1 64 1.0 1" for l in lines]
    open(cache, "w").write("\n".join(odd) + "\n")
    eng.set_tuning(batch_blocks=3, workgroups=0, tile_order=0)
    before = eng.tuning()
    got = eng.search_files(jobs, 6.5)
    assert all(np.array_equal(a, b) for a, b in zip(got, want))
    assert eng.tuning() == before
    eng.set_tuning_cache(None)
    eng.close()


def _compare_search(cands, want, threshold, tol=1e-4):
    """(t, k) sets equal apart from trials within tol of threshold; metrics agree."""
    g = {(int(c["t"]), int(c["k"])): c for c in cands}
    w = {(int(c["t"]), int(c["k"])): c for c in want}
    only_g = [key for key in g if key not in w]
    only_w = [key for key in w if key not in g]
    for key in only_g:
        assert abs(g[key]["metric"] - threshold) <= tol * threshold, ("extra", key, g[key])
    for key in only_w:
        assert abs(w[key]["metric"] - threshold) <= tol * threshold, ("missing", key, w[key])
    worst = 0.0
    for key in set(g) & set(w):
        assert g[key]["nharm"] == w[key]["nharm"] and g[key]["block"] == w[key]["block"]
        worst = max(worst, abs(g[key]["metric"] - w[key]["metric"]) / abs(w[key]["metric"]))
    assert worst <= tol, worst
    return worst, len(set(g) & set(w))


@pytest.mark.parametrize("mode", ["plain", "fused", "fused_valu"])
@pytest.mark.parametrize("nharms,maxdecim,blocksize,thr", [(8, 1, 2048, 6.0), (32, 4, 512, 6.0),
                                                           (60, 6, 2048, 6.5), (60, 6, 1000, 6.5)])
def test_search_vs_oracle(cs, O, small_spectrum, nharms, maxdecim, blocksize, thr, mode):
    amps, N, dt = small_spectrum
    T = N * dt
    dec = cs.decimation_set(nharms, maxdecim)
    _say(f"search mode={mode} nharms={nharms} decims={dec} blocksize={blocksize}")
    eng, p = _engine(cs, nharms=nharms, m=16, decimations=dec, blocksize=blocksize)
    eng.set_mode(mode)
    eng.set_spectrum(_ft(cs, amps, N, dt))
    r_lo, r_hi = 1.0 * T, 60.0 * T
    lodr = 0.5 / nharms
    total = math.floor((r_hi - r_lo) / lodr) + 1
    got = eng.search_raw(r_lo, r_hi, thr)
    want = O.search_region(O.make_params(nharms=nharms, m=16, decimations=dec), amps, N, r_lo,
                           total, blocksize, threshold=thr)
    assert len(want) > 0
    worst, ncommon = _compare_search(got, want, thr)
    _say(f"  {len(got)} vs {len(want)} candidates, common {ncommon}, worst rel {worst:.2e}")


@pytest.mark.parametrize("mode", ["plain", "fused", "fused_valu"])
@pytest.mark.parametrize("blocksize,lo,hi", [(1000, 1.0, 60.0), (64, 10.0, 14.0)])
def test_block_sigmas_vs_oracle(cs, O, small_spectrum, mode, blocksize, lo, hi):
    """Every (block, fold) sigma of the search against _block_sigma of the
    oracle's FP64 profiles (search.jl:490-566): 1000-trial blocks subsample
    8192 of 20-120 bins x 1000 columns; 64-trial blocks of 120 bins take all
    7680 samples (an even count: the mean of the two central values).  The
    bar (1e-5 relative; 5e-5 for the split-BF16 profiles) is well inside the
    ~3e-4 relative gap between neighbouring order statistics of 8192 samples
    at their median, so a wrong rank or a dropped average fails it."""
    amps, N, dt = small_spectrum
    T = N * dt
    dec = cs.decimation_set(60, 6)
    _say(f"block sigmas mode={mode} blocksize={blocksize}")
    eng, p = _engine(cs, nharms=60, m=16, decimations=dec, blocksize=blocksize)
    eng.set_mode(mode)
    eng.set_spectrum(_ft(cs, amps, N, dt))
    r_lo, r_hi = lo * T, hi * T
    total = math.floor((r_hi - r_lo) / (0.5 / 60)) + 1
    nchunks = -(-total // blocksize)
    eng.search_raw(r_lo, r_hi, 6.5)
    got = eng.last_sigmas()
    want = np.zeros((nchunks, len(dec)))
    O.search_region(O.make_params(nharms=60, m=16, decimations=dec), amps, N, r_lo, total,
                    blocksize, threshold=6.5, sigmas=want)
    assert got.shape == want.shape and np.all(want > 0)
    rel = np.abs(got - want) / want
    tol = 5e-5 if mode == "fused" else 1e-5
    _say(f"  {got.size} sigmas, worst rel {rel.max():.2e}")
    assert rel.max() <= tol, (np.unravel_index(rel.argmax(), rel.shape), rel.max())


@pytest.mark.parametrize("mode", ["plain", "fused", "fused_valu"])
def test_search_gate_equals_exact(cs, small_spectrum, mode):
    """boxcar_medmargin = Inf (exact scan everywhere) gives the identical list
    (test/test_search.jl:287-309)."""
    amps, N, dt = small_spectrum
    T = N * dt
    dec = cs.decimation_set(60, 4)
    e1, _ = _engine(cs, nharms=60, m=32, decimations=dec, blocksize=512)
    e2, _ = _engine(cs, nharms=60, m=32, decimations=dec, blocksize=512, boxcar_medmargin=math.inf)
    for e in (e1, e2):
        e.set_mode(mode)
        e.set_spectrum(_ft(cs, amps, N, dt))
    a = e1.search_raw(12.0 * T, 15.0 * T, 6.0)
    b = e2.search_raw(12.0 * T, 15.0 * T, 6.0)
    assert len(a) > 0
    assert np.array_equal(a, b)


def test_search_detects_injection(cs, small_spectrum):
    amps, N, dt = small_spectrum
    ft = _ft(cs, amps, N, dt)
    p = cs.SearchParams(nharms=32, m=16, decimations=cs.decimation_set(32, 4))
    cands = cs.search(ft, p, lofreq=5.0, hifreq=30.0, threshold=8.0)
    assert cands
    best = max(cands, key=lambda c: c.metric)
    assert abs(best.freq - 13.7) < 1.0 / (N * dt)


def test_invalid_params_raise(cs):
    with pytest.raises(ValueError):
        cs.Engine(cs.SearchParams(m=15))
    with pytest.raises(ValueError):
        cs.Engine(cs.SearchParams(precision="f16"))


@pytest.mark.parametrize("nharms,maxdecim,blocksize,m", [(60, 6, 2048, 16), (32, 4, 777, 32),
                                                         (8, 1, 2048, 16)])
def test_fused_equals_plain(cs, small_spectrum, nharms, maxdecim, blocksize, m):
    """The VALU fused mode feeds the same exact rescan from the same FP32
    profiles, so its raw candidate list is identical to the plain mode's; the
    split-BF16 MFMA mode moves sigma by ~1e-5, so it agrees to that tolerance."""
    amps, N, dt = small_spectrum
    T = N * dt
    dec = cs.decimation_set(nharms, maxdecim)
    out = {}
    for mode in ("plain", "fused_valu", "fused"):
        e, _ = _engine(cs, nharms=nharms, m=m, decimations=dec, blocksize=blocksize)
        e.set_mode(mode)
        e.set_spectrum(_ft(cs, amps, N, dt))
        out[mode] = e.search_raw(2.0 * T, 40.0 * T, 6.0)
    assert len(out["plain"]) > 0
    assert np.array_equal(out["plain"], out["fused_valu"])
    worst, _ = _compare_search(out["fused"], out["plain"], 6.0)
    _say(f"  mfma vs plain worst S/N rel {worst:.2e}")
