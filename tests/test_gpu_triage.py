"""GPU checks of the block triage and the block-subset search.

Bars: features within FP32 rounding of the FP64 reference (oracle/triage_ref.py);
MLP logits and gradients against the reference with the device's bf16
roundings emulated (FP32 versus FP64 accumulation only); gradients
deterministic; cs_search_blocks bit-identical to cs_search filtered to the
selected blocks.
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
def cs():
    import coherentsearch
    return coherentsearch


@pytest.fixture(scope="module")
def T():
    import triage_ref
    return triage_ref


def _spectrum(cs, N, seed, pulsars=()):
    from coherentsearch.synth import make_spectrum
    return make_spectrum(N, 64e-6, seed=seed, pulsars=list(pulsars))


@pytest.mark.parametrize("nharms,decims,nprof", [(32, [1, 2, 3, 4], 256),
                                                 (60, [1, 2, 3, 4, 5, 6], 256), (8, [1], 256),
                                                 (8, [1], 16384)])
def test_features_vs_reference(cs, T, nharms, decims, nprof):
    """nprof 16384 at nharms 8: 1024 fundamentals per block, so the medians
    sample every second one"""
    _say(f"triage features nharms={nharms} nprof={nprof}")
    from coherentsearch.synth import Pulsar
    N = 1 << 16
    amps = _spectrum(cs, N, 5, [Pulsar(31.3, 0.05, 30.0)])
    p = cs.SearchParams(nharms=nharms, m=16, decimations=decims)
    eng = cs.Engine(p, nprof)
    eng.set_spectrum(cs.FFTFile.from_arrays(amps, N, 64e-6))
    lodr = 0.5 / nharms
    r_lo, r_hi = 60.3, 60.3 + nprof * 5.5 * lodr
    X = eng.triage_features(r_lo, r_hi)
    total = int(math.floor((r_hi - r_lo) / lodr)) + 1
    want = T.block_features(amps, r_lo, lodr, nprof, total, nharms, decims)
    assert X.shape == want.shape == (6, 32)
    np.testing.assert_allclose(X, want, rtol=2e-5, atol=2e-5)


def test_mlp_forward_and_grad_vs_reference(cs, T):
    _say("mlp forward / grad")
    rng = np.random.default_rng(7)
    n = 1000   # not a multiple of the 32-row wave tile
    params = (rng.standard_normal(T.NPARAMS) * 0.2).astype(np.float32)
    norm = np.concatenate([rng.standard_normal(32) * 0.1, 0.5 + rng.random(32)]).astype(np.float32)
    X = rng.standard_normal((n, 32)).astype(np.float32)
    y = (rng.random(n) > 0.8).astype(np.float32)
    eng = cs.Engine(cs.SearchParams(nharms=8, m=16), 256)
    z = eng.mlp_forward(params, norm, X)
    zr, _ = T.mlp_forward(params, norm, X, round_bf16=True)
    assert np.max(np.abs(z - zr)) <= 2e-3 * np.max(np.abs(zr))
    g, loss = eng.mlp_grad(params, norm, X, y)
    gr, lr_ = T.mlp_grad(params, norm, X, y, round_bf16=True)
    assert abs(loss - lr_) <= 1e-4 * abs(lr_)
    assert np.linalg.norm(g - gr) <= 1e-2 * np.linalg.norm(gr)
    g2, loss2 = eng.mlp_grad(params, norm, X, y)
    assert np.array_equal(g, g2) and loss == loss2          # fixed-order reductions


@pytest.mark.parametrize("mode", ["fused", "plain"])
def test_search_blocks_equals_filtered_search(cs, small_spectrum, mode):
    _say(f"search_blocks mode={mode}")
    amps, N, dt = small_spectrum
    T_ = N * dt
    p = cs.SearchParams(nharms=32, m=16, decimations=cs.decimation_set(32, 4))
    eng = cs.Engine(p, 500)
    eng.set_spectrum(cs.FFTFile.from_arrays(amps, N, dt))
    eng.set_mode(mode)
    r_lo, r_hi = 5.0 * T_, 30.0 * T_
    full = eng.search_raw(r_lo, r_hi, 6.0)
    nb = len(eng.triage_features(r_lo, r_hi))
    assert len(full) > 0 and nb > 4
    assert np.array_equal(eng.search_blocks(r_lo, r_hi, 6.0, np.ones(nb, np.uint8)), full)
    rng = np.random.default_rng(11)
    mask = (rng.random(nb) < 0.4).astype(np.uint8)
    mask[full["block"][0]] = 1
    got = eng.search_blocks(r_lo, r_hi, 6.0, mask)
    want = full[mask[full["block"]] == 1]
    assert len(got) > 0 and np.array_equal(got, want)
    assert len(eng.search_blocks(r_lo, r_hi, 6.0, np.zeros(nb, np.uint8))) == 0
    with pytest.raises(ValueError):
        eng.search_blocks(r_lo, r_hi, 6.0, np.ones(nb + 1, np.uint8))


@pytest.mark.parametrize("mode", ["fused", "plain"])
def test_scheduled_search_rounds(cs, small_spectrum, mode):
    """The budgeted scheduler: with no budget it visits every block and returns
    exactly the full search's raw candidates (in any block order); with a zero
    budget it stops after its first round, whose candidates are the full
    search's in exactly the first ceil(1 %) of the given order."""
    _say(f"scheduled search mode={mode}")
    from coherentsearch.triage import scheduled_search
    amps, N, dt = small_spectrum
    T_ = N * dt
    p = cs.SearchParams(nharms=32, m=16, decimations=cs.decimation_set(32, 4))
    eng = cs.Engine(p, 500)
    eng.set_spectrum(cs.FFTFile.from_arrays(amps, N, dt))
    eng.set_mode(mode)
    r_lo, r_hi = 5.0 * T_, 30.0 * T_
    full = eng.search_raw(r_lo, r_hi, 6.0)
    nb = len(eng.triage_features(r_lo, r_hi))
    order = np.random.default_rng(3).permutation(nb)
    got, done, _ = scheduled_search(eng, r_lo, r_hi, 6.0, order)
    assert done == nb and np.array_equal(got, full)
    # put a block holding candidates first, then stop after one round
    order = np.concatenate([[full["block"][0]], np.setdiff1d(order, [full["block"][0]])])
    got, done, _ = scheduled_search(eng, r_lo, r_hi, 6.0, order, budget_s=0.0, first_fraction=0.01)
    first = max(1, int(np.ceil(0.01 * nb)))
    assert done == first
    assert len(got) > 0 and np.array_equal(got, full[np.isin(full["block"], order[:first])])
    eng.close()


def test_triage_training_separates_blocks(cs):
    """Full-batch training on one spectrum lowers the loss and ranks the blocks
    holding candidates above the others."""
    _say("triage training")
    from coherentsearch.synth import Pulsar
    from coherentsearch.triage import BlockTriage, block_labels
    N, dt = 1 << 19, 64e-6
    rng = np.random.default_rng(5)
    psrs = [Pulsar(float(f), float(d), float(s)) for f, d, s in
            zip(rng.uniform(3.0, 60.0, 12), rng.uniform(0.02, 0.12, 12), rng.uniform(10, 25, 12))]
    amps = _spectrum(cs, N, 21, psrs)
    p = cs.SearchParams(nharms=32, m=16, decimations=cs.decimation_set(32, 4))
    eng = cs.Engine(p, 512)
    eng.set_mode("fused")
    eng.set_spectrum(cs.FFTFile.from_arrays(amps, N, dt))
    T_ = N * dt
    r_lo, r_hi = 2.0 * T_, 70.0 * T_
    raw = eng.search_raw(r_lo, r_hi, 8.0)
    X = eng.triage_features(r_lo, r_hi)
    y = block_labels(raw, len(X))
    assert 0 < y.sum() < len(y)
    model = BlockTriage(seed=1)
    hist = model.fit(eng, X, y, epochs=150, lr=1e-2)
    assert hist[-1] < 0.7 * hist[0]
    z = model.logits(eng, X)
    assert z[y > 0].mean() > z[y == 0].mean() + 1.0


@pytest.mark.parametrize("nharms,decims,nprof", [(32, [1, 2, 3, 4], 256),
                                                 (60, [1, 2, 3, 4, 5, 6], 512), (8, [1], 2048)])
def test_trial_features_vs_reference(cs, T, nharms, decims, nprof):
    """Feature kind 1 (nearest-bin power at the harmonics of every trial)
    against the FP64 reference: FP32 sums against FP64 ones."""
    _say(f"triage trial features nharms={nharms} nprof={nprof}")
    from coherentsearch.synth import Pulsar
    N = 1 << 18
    amps = _spectrum(cs, N, 6, [Pulsar(31.3, 0.05, 30.0)])
    p = cs.SearchParams(nharms=nharms, m=16, decimations=decims)
    eng = cs.Engine(p, nprof)
    eng.set_spectrum(cs.FFTFile.from_arrays(amps, N, 64e-6))
    lodr = 0.5 / nharms
    T_ = N * 64e-6
    r_lo = 31.3 * T_ - 2.5 * nprof * lodr
    r_hi = r_lo + nprof * 5.5 * lodr
    X = eng.triage_features(r_lo, r_hi, kind=1)
    total = int(math.floor((r_hi - r_lo) / lodr)) + 1
    want = T.block_features_trials(amps, r_lo, lodr, nprof, total, nharms, decims)
    assert X.shape == want.shape == (6, 32)
    np.testing.assert_allclose(X, want, rtol=2e-5, atol=2e-5)
    assert np.argmax(X[:, 6]) == 2          # the block holding 31.3 Hz


def test_device_training_equals_fp32_adam(cs, T):
    """cs_mlp_train (forward, backward and the Adam step on the device) equals
    FP32 Adam applied on the host to the device's own gradients, step by step,
    and its reported losses are those gradients' losses."""
    _say("device training")
    from coherentsearch.triage import BlockTriage
    rng = np.random.default_rng(8)
    X = rng.standard_normal((3000, 32)).astype(np.float32)
    y = (X[:, 3] + 0.5 * X[:, 7] > 0.8).astype(np.float32)
    m = BlockTriage(seed=2)
    m.norm = np.concatenate([X.mean(0), 1.0 / X.std(0)]).astype(np.float32)
    eng = cs.Engine(cs.SearchParams(), 2048)
    steps, lr = 6, 5e-3
    got, losses = eng.mlp_train(m.params, m.norm, X, y, steps, lr)
    p = m.params.copy()
    grads, want_losses = [], []
    for _ in range(steps):
        g, loss = eng.mlp_grad(p, m.norm, X, y)
        grads.append(g)
        want_losses.append(loss)
        p = T.adam_steps(m.params, grads, lr)
    assert np.allclose(losses, want_losses, rtol=1e-6, atol=0)
    assert np.max(np.abs(got - p)) <= 1e-6 * max(1.0, np.max(np.abs(p)))
    # and it learns: many steps separate the classes
    m.params, hist = eng.mlp_train(m.params, m.norm, X, y, 300, 1e-2)
    z = eng.mlp_forward(m.params, m.norm, X)
    assert hist[-1] < 0.5 * hist[0]
    assert z[y > 0].mean() > z[y == 0].mean() + 2.0


@pytest.mark.timeout(600)
def test_c4_triage_on_held_out_files(cs):
    """A reduced run of the triage workload: six 2^24-sample files of eight
    injected pulsars each (seeds 4000..4005), labels from the full plain
    search, both feature kinds trained on the device on four files; on the two
    held-out files the top-scored 10 % of the blocks hold more of the full
    search's candidates than the first 10 % in band order (the plain search
    under the same block budget)."""
    from coherentsearch.synth import config_c4_file, make_config_spectra
    from coherentsearch.triage import BlockTriage, block_labels
    cfgs = [config_c4_file(i) for i in range(6)]
    _say("C4 reduced: generating 6 files")
    amps = make_config_spectra(cfgs, workers=6)
    p = cs.SearchParams(nharms=60, m=16, decimations=cs.decimation_set(60, 6), threshold=8.0)
    eng = cs.Engine(p, 2048)
    eng.set_mode("plain")
    feats = {0: [], 1: []}
    labels, raws, grids = [], [], []
    for a, c in zip(amps, cfgs):
        ft = cs.FFTFile.from_arrays(a, c["N"], c["dt"])
        eng.set_spectrum(ft)
        g = cs.trial_grid(ft.T, p, 0.1, 125.0, 100, 2048)
        raw = eng.search_raw(g[0], g[1], 8.0)
        for k in (0, 1):
            feats[k].append(eng.triage_features(g[0], g[1], kind=k))
        labels.append(block_labels(raw, g[5]))
        raws.append(raw)
        grids.append(g)
    ytr = np.concatenate(labels[:4])
    _say(f"C4 reduced: {len(ytr)} training blocks, {int(ytr.sum())} positive")
    for k in (0, 1):
        model = BlockTriage(seed=0, kind=k)
        hist = model.fit(eng, np.concatenate(feats[k][:4]), ytr, epochs=300, lr=3e-3)
        assert hist[-1] < hist[0]
        rec_t = rec_b = 0
        tot = 0
        for X, raw, g in zip(feats[k][4:], raws[4:], grids[4:]):
            mask = model.select(model.logits(eng, X), 0.10)
            nb = g[5]
            band = np.zeros(nb, np.uint8)
            band[:int(mask.sum())] = 1
            blk = raw["block"]
            rec_t += int(np.sum(mask[blk] == 1))
            rec_b += int(np.sum(band[blk] == 1))
            tot += len(raw)
        _say(f"C4 reduced kind {k}: loss {hist[0]:.3f} -> {hist[-1]:.3f}; held-out candidates in "
             f"the top 10 % of blocks {rec_t}/{tot}, in the first 10 % in band order {rec_b}/{tot}")
        assert tot > 0 and rec_t > rec_b
    eng.close()
