"""FP64 reference of the block triage (test infrastructure only).

The triage has no counterpart in the reference search: its features and
scorer are this project's own, so this module restates their definitions
(coherentsearch.jl_amd/csrc/cs_triage.hip) for the GPU tests:

* block_features: the 32 per-block features of the trial grid
  r_lo + t * lodr (blocks of Nprof trials), from |X|^2 and the max over the
  2^s-bin cells (s = floor(log2 g)) covering harmonic g's bin range
  [g f, g f + g - 1] of each integer fundamental f, evaluated here as a
  plain max over the bins of those cells; features 30 and 31 are lower
  medians over every stride-th fundamental, stride = ceil(nfund / 512);
* block_features_trials: kind 1, the same features from the nearest-bin
  power at the harmonics of every trial of a block;
* adam_steps: FP32 Adam, the device's training update;
* mlp_forward / mlp_grad: 32 -> 64 (ReLU) -> 1 logit and the gradient of
  the mean binary cross-entropy, in FP64, optionally with the operands
  rounded to bf16 where the device rounds them (inputs after
  standardisation, W1, ReLU outputs, w2).
"""
import math

import numpy as np

NF, NH = 32, 64
MED_SAMPLES = 512   # fundamentals sampled (evenly strided) for the block medians
NPARAMS = NH * NF + 2 * NH + 1


def unpack(params):
    p = np.asarray(params, dtype=np.float64)
    W1 = p[:NH * NF].reshape(NH, NF)
    b1 = p[NH * NF:NH * NF + NH]
    w2 = p[NH * NF + NH:NH * NF + 2 * NH]
    b2 = p[NH * NF + 2 * NH]
    return W1, b1, w2, b2


def pack(W1, b1, w2, b2):
    return np.concatenate([np.ravel(W1), b1, w2, [b2]])


def bf16(x):
    """Round float32 values to bf16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def _q(P, g, f):
    """max |X|^2 over the level-s cells covering bins [g f, g f + g - 1]."""
    if g == 1:
        return float(P[f]) if 0 <= f < len(P) else 0.0
    s = int(math.floor(math.log2(g)))
    lo, hi = g * f, g * f + g - 1
    c0, c1 = lo >> s, hi >> s
    a, b = max(c0 << s, 0), min((c1 + 1) << s, len(P))
    return float(np.max(P[a:b])) if b > a else 0.0


def block_features(amps, r_lo, lodr, Nprof, total, nharms, decimations):
    P = (np.asarray(amps, dtype=np.complex64).real.astype(np.float32) ** 2 +
         np.asarray(amps, dtype=np.complex64).imag.astype(np.float32) ** 2).astype(np.float64)
    nblocks = -(-total // Nprof)
    Hs = [min(h, nharms) for h in (1, 2, 4, 8, 16, 32, nharms)]
    folds = [k for k in range(2, 7) if k in decimations]
    out = np.zeros((nblocks, NF))
    for c in range(nblocks):
        i0 = c * Nprof
        n = min(Nprof, total - i0)
        rs = r_lo + i0 * lodr
        re = rs + (n - 1) * lodr
        fs = range(int(math.floor(rs)), int(math.floor(re)) + 1)
        S = np.zeros((len(fs), 7))
        Sk = np.zeros((len(fs), 5))
        p1 = np.zeros(len(fs))
        qtop = np.zeros(len(fs))
        for i, f in enumerate(fs):
            q = np.array([_q(P, g, f) for g in range(1, nharms + 1)])
            cs = np.cumsum(q)
            S[i] = [cs[H - 1] / H for H in Hs]
            for k in folds:
                Hk = nharms // k
                Sk[i, k - 2] = q[np.arange(1, Hk + 1) * k - 1].sum() / Hk
            p1[i] = q[0]
            qtop[i] = q[nharms - 1]
        F = out[c]
        F[0:7] = S.max(axis=0)
        F[7:14] = S.mean(axis=0)
        F[14:21] = S.std(axis=0)
        F[21:26] = Sk.max(axis=0)
        F[26] = p1.mean()
        F[27] = qtop.mean()
        F[28] = math.log2(max(rs, 1.0)) / 20.0
        F[29] = np.mean(S[:, 6] > 1.5)
        stride = -(-len(fs) // MED_SAMPLES)
        F[30] = lower_median(p1[::stride])
        F[31] = F[6] / max(lower_median(S[::stride, 6]), 1e-30)
    return out


def block_features_trials(amps, r_lo, lodr, Nprof, total, nharms, decimations):
    """Kind 1: the same 32 features over every trial r_t = (r_lo + i0 lodr) +
    j lodr of a block, with the harmonic power P(nint(h r_t)), nint(x) =
    floor(x + 1/2) in FP64, in place of the pyramid maxima; medians over
    every stride-th trial, stride = ceil(n / 512)."""
    a = np.asarray(amps, dtype=np.complex64)
    P = (a.real.astype(np.float32) ** 2 + a.imag.astype(np.float32) ** 2).astype(np.float64)
    nb = len(P)
    nblocks = -(-total // Nprof)
    Hs = [min(h, nharms) for h in (1, 2, 4, 8, 16, 32, nharms)]
    folds = [k for k in range(2, 7) if k in decimations]
    h = np.arange(1, nharms + 1, dtype=np.float64)
    out = np.zeros((nblocks, NF))
    for c in range(nblocks):
        i0 = c * Nprof
        n = min(Nprof, total - i0)
        rs = r_lo + i0 * lodr
        rt = rs + np.arange(n, dtype=np.float64) * lodr
        b = np.floor(h[None, :] * rt[:, None] + 0.5).astype(np.int64)   # [trial][h]
        q = np.where((b >= 0) & (b < nb), P[np.clip(b, 0, nb - 1)], 0.0)
        cs = np.cumsum(q, axis=1)
        S = np.stack([cs[:, H - 1] / H for H in Hs], axis=1)
        Sk = np.zeros((n, 5))
        for k in folds:
            Hk = nharms // k
            Sk[:, k - 2] = q[:, np.arange(1, Hk + 1) * k - 1].sum(axis=1) / Hk
        F = out[c]
        F[0:7] = S.max(axis=0)
        F[7:14] = S.mean(axis=0)
        F[14:21] = S.std(axis=0)
        F[21:26] = Sk.max(axis=0)
        F[26] = q[:, 0].mean()
        F[27] = q[:, nharms - 1].mean()
        F[28] = math.log2(max(rs, 1.0)) / 20.0
        F[29] = np.mean(S[:, 6] > 1.5)
        stride = -(-n // MED_SAMPLES)
        F[30] = lower_median(q[::stride, 0])
        F[31] = F[6] / max(lower_median(S[::stride, 6]), 1e-30)
    return out


def adam_steps(params, grads, lr, b1=0.9, b2=0.999, eps=1e-8):
    """FP32 Adam, as the device applies it, over a sequence of gradients."""
    p = np.asarray(params, dtype=np.float32).copy()
    m1 = np.zeros_like(p)
    m2 = np.zeros_like(p)
    f = np.float32
    for t, g in enumerate(grads, start=1):
        g = np.asarray(g, dtype=np.float32)
        c1, c2 = f(1.0 - b1 ** t), f(1.0 - b2 ** t)
        m1 = f(b1) * m1 + f(1.0 - b1) * g
        m2 = f(b2) * m2 + f(1.0 - b2) * (g * g)
        p = p - f(lr) * (m1 / c1) / (np.sqrt(m2 / c2) + f(eps))
    return p


def lower_median(v):
    """the ((n - 1) // 2)-th smallest value (0 for no values)"""
    v = np.sort(np.asarray(v, dtype=np.float64))
    return float(v[(len(v) - 1) // 2]) if len(v) else 0.0


def _std_inputs(X, norm, round_bf16):
    X = np.asarray(X, dtype=np.float64)
    mean = np.asarray(norm[:NF], dtype=np.float64)
    inv = np.asarray(norm[NF:], dtype=np.float64)
    if round_bf16:   # FP32 standardisation, then bf16, as the device does
        x = (np.asarray(X, np.float32) - np.asarray(mean, np.float32)) * np.asarray(inv, np.float32)
        return bf16(x)
    return (X - mean) * inv


def mlp_forward(params, norm, X, round_bf16=False):
    W1, b1, w2, b2 = unpack(params)
    x = _std_inputs(X, norm, round_bf16)
    W = bf16(W1) if round_bf16 else W1
    h = x @ W.T + b1
    a = np.maximum(h, 0.0)
    if round_bf16:
        a = bf16(a)
        w2 = bf16(w2)
    return a @ w2 + b2, (x, h, a, w2)


def mlp_grad(params, norm, X, y, round_bf16=False):
    """(gradient in the params layout, mean BCE loss)"""
    z, (x, h, a, w2e) = mlp_forward(params, norm, X, round_bf16)
    y = np.asarray(y, dtype=np.float64)
    n = len(y)
    p = 1.0 / (1.0 + np.exp(-z))
    loss = np.mean(np.maximum(z, 0) - z * y + np.log1p(np.exp(-np.abs(z))))
    dz = (p - y) / n
    gb2 = dz.sum()
    gw2 = a.T @ dz
    dh = np.outer(dz, w2e) * (h > 0)
    if round_bf16:
        dh = bf16(dh)
    gb1 = dh.sum(axis=0) if not round_bf16 else (np.outer(dz, w2e) * (h > 0)).sum(axis=0)
    gW1 = dh.T @ x
    return pack(gW1, gb1, gw2, gb2), loss
