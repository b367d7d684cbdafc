"""CPU checks of the block-triage reference (oracle/triage_ref.py).

The triage (features + MLP scorer) is this project's own addition, so its
reference is pinned by mathematics rather than by reference fixtures:
* the FP64 gradient equals central finite differences of the loss;
* bf16 rounding is round-to-nearest-even on known values;
* the cell maxima equal an independently built max pyramid;
* power features scale with |X|^2 and the ratio features do not.
"""
import math

import numpy as np
import pytest

import triage_ref as T


def _rand_model(rng):
    params = rng.standard_normal(T.NPARAMS) * 0.3
    norm = np.concatenate([rng.standard_normal(T.NF) * 0.1, 1.0 + rng.random(T.NF)])
    return params, norm


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    params, norm = _rand_model(rng)
    X = rng.standard_normal((40, T.NF))
    y = (rng.random(40) > 0.5).astype(float)
    g, loss = T.mlp_grad(params, norm, X, y)

    def L(p):
        z, _ = T.mlp_forward(p, norm, X)
        return np.mean(np.maximum(z, 0) - z * y + np.log1p(np.exp(-np.abs(z))))

    assert abs(L(params) - loss) < 1e-12
    idx = np.concatenate([rng.choice(T.NH * T.NF, 40, replace=False),
                          T.NH * T.NF + np.arange(0, 2 * T.NH + 1, 7), [T.NPARAMS - 1]])
    eps = 1e-6
    for i in idx:
        e = np.zeros(T.NPARAMS)
        e[i] = eps
        fd = (L(params + e) - L(params - e)) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(fd), (i, fd, g[i])


def test_bf16_rounding():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 1.0 + 2 ** -7 + 2 ** -8],
                    dtype=np.float32)
    want = [1.0, 1.0, 1.0 + 2 ** -7, -2.5, 1.0 + 2 ** -6]   # ties to even
    assert np.array_equal(T.bf16(vals), np.array(want))


def test_cell_maxima_equal_a_numpy_pyramid():
    rng = np.random.default_rng(2)
    P = rng.exponential(size=5000).astype(np.float64)
    levels = {}
    for s in range(1, 7):
        w = 1 << s
        pad = np.concatenate([P, np.zeros((-len(P)) % w)])
        levels[s] = pad.reshape(-1, w).max(axis=1)
    for _ in range(400):
        g = int(rng.integers(2, 65))
        f = int(rng.integers(0, 5000 // g + 2))
        s = int(math.floor(math.log2(g)))
        c0, c1 = (g * f) >> s, (g * f + g - 1) >> s
        L = levels[s]
        want = max([L[c] for c in range(c0, c1 + 1) if 0 <= c < len(L)], default=0.0)
        assert T._q(P, g, f) == want


def test_feature_scaling():
    rng = np.random.default_rng(3)
    n = 1 << 14
    amps = ((rng.standard_normal(n) + 1j * rng.standard_normal(n)) / np.sqrt(2)).astype(np.complex64)
    args = dict(r_lo=60.0, lodr=1 / 64, Nprof=256, total=256 * 5 + 17, nharms=32,
                decimations=[1, 2, 3, 4])
    F1 = T.block_features(amps, **args)
    F2 = T.block_features(amps * np.float32(2.0), **args)
    assert F1.shape == (6, 32)
    power = list(range(0, 28)) + [30]
    assert np.allclose(F2[:, power], 4.0 * F1[:, power], rtol=1e-12, atol=0)
    assert np.allclose(F2[:, [28, 31]], F1[:, [28, 31]], rtol=1e-12, atol=0)
    assert np.all(F1[:, 25] == 0.0)   # fold 6 is not searched


def test_lower_median_and_sampling():
    rng = np.random.default_rng(4)
    for n in (1, 2, 7, 64, 513):
        v = rng.exponential(size=n)
        want = np.median(v) if n % 2 else np.sort(v)[n // 2 - 1]
        assert T.lower_median(v) == want
    assert T.lower_median([]) == 0.0
    # a block of 1024 fundamentals samples every second one for its medians
    n = 1 << 15
    amps = ((rng.standard_normal(n) + 1j * rng.standard_normal(n)) / np.sqrt(2)).astype(np.complex64)
    lodr, Nprof = 1 / 16, 16384
    F = T.block_features(amps, r_lo=40.0, lodr=lodr, Nprof=Nprof, total=Nprof, nharms=8,
                         decimations=[1])
    P = (amps.real.astype(np.float32) ** 2 + amps.imag.astype(np.float32) ** 2).astype(np.float64)
    fs = np.arange(40, 40 + 1024)
    assert F[0, 30] == np.sort(P[fs[::2]])[255]
    assert F[0, 30] != np.sort(P[fs])[511]   # the full median differs on this seed


def test_trial_features_against_brute_force():
    """Kind 1 against plain Python loops over trials and harmonics."""
    rng = np.random.default_rng(3)
    nb = 3000
    amps = (rng.standard_normal(nb) + 1j * rng.standard_normal(nb)).astype(np.complex64)
    amps[0] = 0
    P = [float(np.float32(a.real) * np.float32(a.real) + np.float32(a.imag) * np.float32(a.imag))
         for a in amps]
    r_lo, lodr, Nprof, total, nh = 40.3, 0.5 / 8, 64, 150, 8
    got = T.block_features_trials(amps, r_lo, lodr, Nprof, total, nh, [1, 2, 3, 4])
    for c in range(-(-total // Nprof)):
        i0 = c * Nprof
        n = min(Nprof, total - i0)
        rs = r_lo + i0 * lodr
        E = []
        for j in range(n):
            r = rs + j * lodr
            E.append([P[int(math.floor(h * r + 0.5))] if math.floor(h * r + 0.5) < nb else 0.0
                      for h in range(1, nh + 1)])
        E = np.array(E)
        s8 = E.sum(axis=1) / 8
        s2 = E[:, :2].sum(axis=1) / 2
        k2 = E[:, [1, 3, 5, 7]].sum(axis=1) / 4
        assert math.isclose(got[c, 6], s8.max(), rel_tol=1e-12)
        assert math.isclose(got[c, 8], s2.mean(), rel_tol=1e-12)
        assert math.isclose(got[c, 21], k2.max(), rel_tol=1e-12)
        assert math.isclose(got[c, 26], E[:, 0].mean(), rel_tol=1e-12)
        assert math.isclose(got[c, 30], sorted(E[:, 0])[(n - 1) // 2], rel_tol=1e-12)


def test_trial_features_find_a_comb():
    """A comb of power at nint(h r0) lights up the block holding r0 and its
    harmonic-sum features grow with the harmonics summed."""
    nb = 20000
    amps = np.zeros(nb, dtype=np.complex64) + 0.1
    r0 = 1234.56
    for h in range(1, 9):
        amps[int(math.floor(h * r0 + 0.5))] = 4.0
    lodr = 0.5 / 8
    r_lo = 1200.0
    F = T.block_features_trials(amps, r_lo, lodr, 256, 1200, 8, [1])
    b = int((r0 - r_lo) / lodr) // 256
    assert np.argmax(F[:, 6]) == b
    assert math.isclose(F[b, 6], 16.0, rel_tol=1e-6)        # every harmonic on the comb
    assert F[b, 0] == pytest.approx(16.0)


def test_adam_steps_closed_form_and_fp64():
    """The device's FP32 Adam: the first step moves each parameter by lr
    sign(g) |g| / (|g| + eps); later steps follow FP64 Adam to FP32 rounding."""
    rng = np.random.default_rng(4)
    p0 = rng.standard_normal(50).astype(np.float32)
    gs = [rng.standard_normal(50).astype(np.float32) for _ in range(5)]
    p1 = T.adam_steps(p0, gs[:1], 1e-2)
    want1 = p0.astype(np.float64) - 1e-2 * np.sign(gs[0]) * np.abs(gs[0]) / (np.abs(gs[0]) + 1e-8)
    assert np.allclose(p1, want1, rtol=0, atol=2e-7)
    p = p0.astype(np.float64)
    m1 = np.zeros(50)
    m2 = np.zeros(50)
    for t, g in enumerate(gs, start=1):
        g = g.astype(np.float64)
        m1 = 0.9 * m1 + 0.1 * g
        m2 = 0.999 * m2 + 0.001 * g * g
        p = p - 1e-2 * (m1 / (1 - 0.9 ** t)) / (np.sqrt(m2 / (1 - 0.999 ** t)) + 1e-8)
    assert np.allclose(T.adam_steps(p0, gs, 1e-2), p, rtol=0, atol=1e-6)
