"""ctypes front end to the CPU restatement in ``cs_oracle.c``.

TEST INFRASTRUCTURE ONLY.  This module is the parity checker for the HIP
library and the CPU baseline leg of ``bench.py``; the product path never
imports it.  Every wrapper names the reference function (CoherentSearch.jl
file:line) that the C code restates.
"""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class OrParams(C.Structure):
    _fields_ = [
        ("nharms", C.c_int32),
        ("m", C.c_int32),
        ("numbetween", C.c_int32),
        ("ndecim", C.c_int32),
        ("hidr", C.c_double),
        ("threshold", C.c_double),
        ("boxcar_fsp", C.c_double),
        ("boxcar_maxfrac", C.c_double),
        ("boxcar_medmargin", C.c_double),
        ("decimations", C.POINTER(C.c_int32)),
    ]


class OrCand(C.Structure):
    _fields_ = [
        ("t", C.c_int64),
        ("k", C.c_int32),
        ("nharm", C.c_int32),
        ("block", C.c_int64),
        ("metric", C.c_double),
    ]


CAND_DTYPE = np.dtype([("t", "<i8"), ("k", "<i4"), ("nharm", "<i4"),
                       ("block", "<i8"), ("metric", "<f8")])


class OracleError(RuntimeError):
    pass


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libcsoracle.so")
        if not os.path.exists(path):
            raise OracleError(f"{path} is not built (run `make -C oracle`)")
        L = C.CDLL(path)
        dp = C.POINTER(C.c_double)
        fp = C.POINTER(C.c_float)
        i32p = C.POINTER(C.c_int32)
        i64p = C.POINTER(C.c_int64)
        L.or_finterp_coeffs.argtypes = [C.c_double, C.c_int32, dp]
        L.or_nearby_bin_range.argtypes = [C.c_double, C.c_int32, i64p, i64p]
        L.or_fourier_interp.argtypes = [C.c_double, dp, C.c_int64, C.c_int32, dp]
        L.or_next_pow_of_2.argtypes = [C.c_int64]
        L.or_next_pow_of_2.restype = C.c_int64
        L.or_is_smooth.argtypes = [C.c_int64]
        L.or_is_smooth.restype = C.c_int32
        L.or_next_smooth.argtypes = [C.c_int64]
        L.or_next_smooth.restype = C.c_int64
        L.or_finterp_fft.argtypes = [C.c_int64, C.c_int64, C.c_int32, dp, C.c_int64,
                                     C.c_int32, C.c_int64, dp]
        L.or_fft.argtypes = [dp, C.c_int64, C.c_int32]
        L.or_finterp_direct.argtypes = [C.c_double, C.c_int64, C.c_double, dp,
                                        C.c_int64, C.c_int32, dp]
        L.or_trial_grid_rational.argtypes = [C.c_double, i64p, i64p]
        L.or_plans_build.argtypes = [C.c_int32, C.c_int32, C.c_double, C.c_double,
                                     C.c_int32]
        L.or_plans_build.restype = C.c_void_p
        L.or_plans_free.argtypes = [C.c_void_p]
        L.or_plan_info.argtypes = [C.c_void_p, C.c_int32, i64p, dp]
        L.or_plan_tables.argtypes = [C.c_void_p, C.c_int32, i32p,
                                     C.POINTER(C.c_uint8), dp, dp]
        L.or_chunk_state.argtypes = [C.c_void_p, C.c_int32, C.c_int64, i64p, i64p]
        for nm in ("or_fill_chunk", "or_fill_chunk_pertrial"):
            getattr(L, nm).argtypes = [C.c_void_p, fp, C.c_int64, C.c_int64,
                                       C.c_int64, C.c_int64, C.c_int64, dp, dp]
        L.or_brfft_rows.argtypes = [dp, dp, C.c_int64, C.c_int32, C.c_int32,
                                    C.c_int64, dp]
        L.or_median.argtypes = [dp, C.c_int64]
        L.or_median.restype = C.c_double
        L.or_boxcar_widths.argtypes = [C.c_int32, C.c_double, C.c_double, i32p]
        L.or_boxcar_widths.restype = C.c_int32
        L.or_ladder_boxcar_widths.argtypes = [C.c_int32, C.c_int32,
                                              C.POINTER(OrParams), i32p]
        L.or_ladder_boxcar_widths.restype = C.c_int32
        L.or_block_sigma.argtypes = [dp, C.c_int64, C.c_int64, C.c_int32, C.c_int64,
                                     C.c_int64, C.c_int32]
        L.or_block_sigma.restype = C.c_double
        L.or_boxcar_scan_col.argtypes = [dp, C.c_int64, C.c_int32, i32p, C.c_int32,
                                         C.c_double, C.c_double]
        L.or_boxcar_scan_col.restype = C.c_double
        L.or_boxcar_exact_col.argtypes = [dp, C.c_int64, C.c_int32, i32p, C.c_int32,
                                          C.c_double]
        L.or_boxcar_exact_col.restype = C.c_double
        L.or_boxcar_best_width.argtypes = [dp, C.c_int32, C.c_double, C.c_double]
        L.or_boxcar_best_width.restype = C.c_int32
        L.or_boxcar_metrics.argtypes = [dp, C.c_int64, C.c_int64, i32p, C.c_int32,
                                        C.c_int32, C.c_double, C.c_double, dp]
        L.or_snr_metrics.argtypes = [dp, C.c_int32, C.c_int64, C.c_int64, C.c_int32,
                                     i32p, C.c_int32, dp]
        L.or_chunk_metrics.argtypes = [C.POINTER(OrParams), fp, C.c_int64, C.c_int64,
                                       C.c_double, C.c_int64, C.c_int32, dp]
        L.or_block_metrics.argtypes = [C.POINTER(OrParams), fp, C.c_int64, C.c_int64,
                                       dp, C.c_int64, C.c_int32, i32p, C.c_int32, dp]
        L.or_reference_profiles.argtypes = [C.POINTER(OrParams), fp, C.c_int64,
                                            C.c_int64, dp, C.c_int64, C.c_int32, dp]
        L.or_decim_metrics.argtypes = [C.POINTER(OrParams), fp, C.c_int64, C.c_int64,
                                       C.c_double, C.c_int64, C.c_int32, C.c_int32,
                                       dp, i64p]
        L.or_search_region.argtypes = [C.POINTER(OrParams), fp, C.c_int64, C.c_int64,
                                       C.c_double, C.c_int64, C.c_int64, C.c_int64,
                                       C.c_int64, C.c_double, C.c_double, C.c_int32,
                                       C.POINTER(OrCand), C.c_int64, i64p, dp, dp]
        _LIB = L
    return _LIB


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _i32p(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _check(st, what):
    if st == 1:
        raise ValueError(f"{what}: invalid argument")
    if st == 2:
        raise IndexError(f"{what}: dimension mismatch")
    if st != 0:
        raise OracleError(f"{what}: status {st}")


def _c128(ft):
    a = np.ascontiguousarray(np.asarray(ft, dtype=np.complex128))
    return a.view(np.float64), a


def _amps64(amps):
    a = np.ascontiguousarray(np.asarray(amps, dtype=np.complex64))
    return a.view(np.float32), a


def make_params(nharms=32, m=16, numbetween=16, hidr=0.5, threshold=8.0,
                boxcar_fsp=1.5, boxcar_maxfrac=0.3, boxcar_medmargin=2.0,
                decimations=(1,)):
    """SearchParams (search.jl:131-144) as the C struct; keeps its arrays alive."""
    dec = np.ascontiguousarray(np.asarray(decimations, dtype=np.int32))
    p = OrParams(int(nharms), int(m), int(numbetween), len(dec), float(hidr),
                 float(threshold), float(boxcar_fsp), float(boxcar_maxfrac),
                 float(boxcar_medmargin), _i32p(dec))
    p._keep = dec
    return p


# ---- fourierinterp.jl -------------------------------------------------------

def finterp_coeffs(dr, m):
    """fourierinterp.jl:23"""
    out = np.zeros(2 * m)
    _check(lib().or_finterp_coeffs(float(dr), int(m), _dp(out)), "finterp_coeffs")
    return out.view(np.complex128)


def nearby_fourier_bin_range(r, m):
    """fourierinterp.jl:44 — returned as a 0-based half-open (lo, hi)."""
    lo, hi = C.c_int64(), C.c_int64()
    _check(lib().or_nearby_bin_range(float(r), int(m), C.byref(lo), C.byref(hi)),
           "nearby_fourier_bin_range")
    return lo.value, hi.value


def fourier_interp(r, ft, m):
    """fourierinterp.jl:66"""
    f, keep = _c128(ft)
    out = np.zeros(2)
    _check(lib().or_fourier_interp(float(r), _dp(f), len(keep), int(m), _dp(out)),
           "fourier_interp")
    return complex(out[0], out[1])


def next_pow_of_2(n):
    if n <= 0:
        raise ValueError("n must be a positive integer")
    return lib().or_next_pow_of_2(int(n))


def is_smooth(n):
    return bool(lib().or_is_smooth(int(n)))


def next_smooth(n):
    if n <= 0:
        raise ValueError("n must be a positive integer")
    return lib().or_next_smooth(int(n))


def finterp_fft(lobin, numbins, numbetween, ft, m, fftlen=0):
    """fourierinterp.jl:190"""
    f, keep = _c128(ft)
    out = np.zeros(2 * numbins * numbetween)
    _check(lib().or_finterp_fft(int(lobin), int(numbins), int(numbetween), _dp(f),
                                len(keep), int(m), int(fftlen), _dp(out)), "finterp_fft")
    return out.view(np.complex128)


def fft(x, sign=-1):
    a = np.ascontiguousarray(np.asarray(x, dtype=np.complex128)).copy()
    _check(lib().or_fft(_dp(a.view(np.float64)), len(a), int(sign)), "fft")
    return a


# ---- directinterp.jl --------------------------------------------------------

def finterp_direct(r0, n, step, ft, m):
    """directinterp.jl:71"""
    f, keep = _c128(ft)
    out = np.zeros(2 * n)
    _check(lib().or_finterp_direct(float(r0), int(n), float(step), _dp(f), len(keep),
                                   int(m), _dp(out)), "finterp_direct")
    return out.view(np.complex128)


def trial_grid_rational(lodr):
    """directinterp.jl:217 — (pnum, q) or None."""
    p, q = C.c_int64(), C.c_int64()
    ok = lib().or_trial_grid_rational(float(lodr), C.byref(p), C.byref(q))
    return (p.value, q.value) if ok else None


class Plans:
    """build_direct_plans(WT, params, r_lo), directinterp.jl:345."""

    def __init__(self, nharms, m, hidr, r_lo, weights_f64=False):
        self.nharms, self.m = nharms, m
        self._h = lib().or_plans_build(int(nharms), int(m), float(hidr), float(r_lo),
                                       int(bool(weights_f64)))
        if not self._h:
            raise ValueError("build_direct_plans: invalid arguments")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().or_plans_free(h)
            self._h = None

    def info(self, h):
        info = np.zeros(8, dtype=np.int64)
        dr0 = C.c_double()
        _check(lib().or_plan_info(self._h, int(h), info.ctypes.data_as(C.POINTER(C.c_int64)),
                                  C.byref(dr0)), "plan_info")
        keys = ("q", "s", "base_adv", "pnum", "rfloor0", "P", "ngrp", "gstep")
        d = {k: int(v) for k, v in zip(keys, info)}
        d["dr0"] = dr0.value
        return d

    def tables(self, h):
        d = self.info(h)
        row = np.zeros(d["q"], dtype=np.int32)
        carry = np.zeros(d["P"], dtype=np.uint8)
        W = np.zeros(self.m * d["P"])
        A = np.zeros(2 * d["P"])
        _check(lib().or_plan_tables(self._h, int(h), _i32p(row),
                                    carry.ctypes.data_as(C.POINTER(C.c_uint8)), _dp(W),
                                    _dp(A)), "plan_tables")
        return row, carry.astype(bool), W.reshape(d["P"], self.m), A.view(np.complex128)

    def chunk_state(self, h, t0):
        res, qint = C.c_int64(), C.c_int64()
        _check(lib().or_chunk_state(self._h, int(h), int(t0), C.byref(res), C.byref(qint)),
               "chunk_state")
        return res.value, qint.value

    def fill_chunk(self, amps, N, t0, n, pertrial=False):
        """fill_harmonic_row_direct! for every harmonic; returns (nharms+1, n) complex."""
        a, keep = _amps64(amps)
        re = np.zeros((self.nharms + 1) * n)
        im = np.zeros((self.nharms + 1) * n)
        fn = lib().or_fill_chunk_pertrial if pertrial else lib().or_fill_chunk
        _check(fn(self._h, _fp(a), len(keep), int(N), int(t0), int(n), int(n), _dp(re),
                  _dp(im)), "fill_chunk")
        return (re + 1j * im).reshape(self.nharms + 1, n)


# ---- c2r --------------------------------------------------------------------

def brfft_rows(X, nbins, k=1):
    """Unnormalised c2r of rows 0,k,..,H*k of X (rows, ncol) -> (nbins, ncol)."""
    X = np.asarray(X, dtype=np.complex128)
    re = np.ascontiguousarray(X.real)
    im = np.ascontiguousarray(X.imag)
    ncol = X.shape[1]
    out = np.zeros(nbins * ncol)
    _check(lib().or_brfft_rows(_dp(re), _dp(im), ncol, int(k), int(nbins), ncol,
                               _dp(out)), "brfft_rows")
    return out.reshape(nbins, ncol)


# ---- search.jl: metric ------------------------------------------------------

def median(v):
    a = np.array(v, dtype=np.float64)
    return lib().or_median(_dp(a), len(a))


def boxcar_widths(nbins, fsp=1.5, maxfrac=0.3):
    """search.jl:387"""
    out = np.zeros(64, dtype=np.int32)
    n = lib().or_boxcar_widths(int(nbins), float(fsp), float(maxfrac), _i32p(out))
    return [int(x) for x in out[:n]]


def ladder_boxcar_widths(nbins, k, params):
    """search.jl:443"""
    out = np.zeros(64, dtype=np.int32)
    n = lib().or_ladder_boxcar_widths(int(nbins), int(k), C.byref(params), _i32p(out))
    return [int(x) for x in out[:n]]


def block_sigma(M, cap=8192, center="zero"):
    """_block_sigma(M, nbins, n, buf), search.jl:490; M is (nbins, n)."""
    M = np.ascontiguousarray(np.asarray(M, dtype=np.float64))
    nbins, n = M.shape
    return lib().or_block_sigma(_dp(M), n, 1, nbins, n, int(cap),
                                1 if center == "median" else 0)


def boxcar_exact(col, widths, invsigma):
    c = np.ascontiguousarray(np.asarray(col, dtype=np.float64))
    w = np.ascontiguousarray(np.asarray(widths, dtype=np.int32))
    return lib().or_boxcar_exact_col(_dp(c), 1, len(c), _i32p(w), len(w), float(invsigma))


def boxcar_scan(col, widths, invsigma, b=0.0):
    c = np.ascontiguousarray(np.asarray(col, dtype=np.float64))
    w = np.ascontiguousarray(np.asarray(widths, dtype=np.int32))
    return lib().or_boxcar_scan_col(_dp(c), 1, len(c), _i32p(w), len(w), float(invsigma),
                                    float(b))


def boxcar_best_width(prof, fsp=1.5, maxfrac=0.3):
    """search.jl:618 — (w, ducy)."""
    p = np.ascontiguousarray(np.asarray(prof, dtype=np.float64))
    if len(p) < 1:
        raise ValueError("profile must be non-empty")
    w = lib().or_boxcar_best_width(_dp(p), len(p), float(fsp), float(maxfrac))
    return w, w / len(p)


def boxcar_metrics(profs, widths, invsigma, medcut):
    """boxcar_metrics!, search.jl:825; profs (nbins, n)."""
    P = np.ascontiguousarray(np.asarray(profs, dtype=np.float64))
    nbins, n = P.shape
    w = np.ascontiguousarray(np.asarray(widths, dtype=np.int32))
    out = np.zeros(n)
    _check(lib().or_boxcar_metrics(_dp(P), n, n, _i32p(w), len(w), nbins, float(invsigma),
                                   float(medcut), _dp(out)), "boxcar_metrics")
    return out


def snr_metrics(profs, sigma_samples=None, sigma_center="zero", widths=None):
    """search.jl:933; profs (nbins, L)."""
    P = np.asarray(profs, dtype=np.float64)
    nbins, L = P.shape
    Pc = np.ascontiguousarray(P.T)       # column-major (nbins, L) storage
    out = np.zeros(L)
    ss = -1 if sigma_samples is None else int(sigma_samples)
    if widths is None:
        wp, nw = None, 0
    else:
        w = np.ascontiguousarray(np.asarray(widths, dtype=np.int32))
        wp, nw = _i32p(w), len(w)
    _check(lib().or_snr_metrics(_dp(Pc), nbins, L, ss, 1 if sigma_center == "median" else 0,
                                wp, nw, _dp(out)), "snr_metrics")
    return out


# ---- drivers ------------------------------------------------------------------

def chunk_metrics(params, amps, N, rstart, n, weights_f64=False):
    """search.jl:1740"""
    a, keep = _amps64(amps)
    out = np.zeros(n)
    _check(lib().or_chunk_metrics(C.byref(params), _fp(a), len(keep), int(N),
                                  float(rstart), int(n), int(bool(weights_f64)), _dp(out)),
           "chunk_metrics")
    return out


def block_metrics(params, amps, N, rfund, kernel="fft", widths=None):
    """search.jl:1061"""
    if kernel not in ("fft", "direct"):
        raise ValueError(f"kernel must be :fft or :direct, got :{kernel}")
    a, keep = _amps64(amps)
    r = np.ascontiguousarray(np.asarray(rfund, dtype=np.float64))
    out = np.zeros(len(r))
    if widths is None:
        wp, nw = None, 0
    else:
        w = np.ascontiguousarray(np.asarray(widths, dtype=np.int32))
        wp, nw = _i32p(w), len(w)
    _check(lib().or_block_metrics(C.byref(params), _fp(a), len(keep), int(N), _dp(r), len(r),
                                  1 if kernel == "direct" else 0, wp, nw, _dp(out)),
           "block_metrics")
    return out


def reference_profiles(params, amps, N, rfund, kernel="fft"):
    """search.jl:1011 — (2*nharms, L) normalised profiles."""
    a, keep = _amps64(amps)
    r = np.ascontiguousarray(np.asarray(rfund, dtype=np.float64))
    nb = 2 * params.nharms
    out = np.zeros(nb * len(r))
    _check(lib().or_reference_profiles(C.byref(params), _fp(a), len(keep), int(N), _dp(r),
                                       len(r), 1 if kernel == "direct" else 0, _dp(out)),
           "reference_profiles")
    return out.reshape(len(r), nb).T


def decim_metrics(params, amps, N, rstart, n, k, weights_f64=False):
    """decim_pass! metrics of fold k over one chunk at t0=0 (search.jl:1664)."""
    a, keep = _amps64(amps)
    out = np.zeros(n)
    nv = C.c_int64()
    _check(lib().or_decim_metrics(C.byref(params), _fp(a), len(keep), int(N), float(rstart),
                                  int(n), int(k), int(bool(weights_f64)), _dp(out),
                                  C.byref(nv)), "decim_metrics")
    return out[:nv.value]


def search_region(params, amps, N, r_lo, total, Nprof=2048, c_begin=0, c_end=None,
                  threshold=None, medcut=None, nthreads=0, metrics=None, sigmas=None):
    """_search_region! (search.jl:1768): raw candidates as a CAND_DTYPE array sorted
    by (t, k).  `metrics`, a float64 array [1 + decimations > 1][total] (NaN
    filled by the caller), receives every scored metric; `sigmas`, a float64
    array [chunk][1 + decimations > 1], each block's _block_sigma."""
    a, keep = _amps64(amps)
    thr = params.threshold if threshold is None else float(threshold)
    mc = math.nan if medcut is None else float(medcut)
    nchunks = -(-int(total) // int(Nprof))
    ce = nchunks if c_end is None else int(c_end)
    mp = None
    if metrics is not None:
        assert metrics.dtype == np.float64 and metrics.flags.c_contiguous
        mp = _dp(metrics)
    sp = None
    if sigmas is not None:
        assert sigmas.dtype == np.float64 and sigmas.flags.c_contiguous
        assert sigmas.shape[0] >= ce
        sp = _dp(sigmas)
    cap = 1 << 16
    while True:
        out = (OrCand * cap)()
        n = C.c_int64()
        st = lib().or_search_region(C.byref(params), _fp(a), len(keep), int(N), float(r_lo),
                                    int(total), int(Nprof), int(c_begin), ce, thr, mc,
                                    int(nthreads), out, cap, C.byref(n), mp, sp)
        if st == 3:
            cap = n.value + 16
            continue
        _check(st, "search_region")
        break
    arr = np.frombuffer(out, dtype=CAND_DTYPE, count=n.value).copy()
    return np.sort(arr, order=("t", "k"))


def candidate_profile(amps, N, r, nharm, m=64):
    """candidate_profile, candidate.jl:37-61: harmonics h = 1, 2, ... of r by
    fourier_interp (m-bin kernel) until the first whose window [rint - m/2,
    rint + m/2] leaves the amplitudes or whose r h reaches N/2 (rint =
    floor(r h + 1e-15) + 1); irfft of [0, X_1 .. X_H] into 2H bins, or two zero
    bins when H = 0."""
    if nharm < 1:
        raise ValueError("nharm must be >= 1")
    if m % 2:
        raise ValueError("m must be even")
    a = np.asarray(amps, dtype=np.complex64).astype(np.complex128)
    m2, nhalf = m // 2, int(N) // 2
    X = []
    for h in range(1, nharm + 1):
        rh = r * h
        rint = math.floor(rh + 1e-15) + 1
        if not (rint - m2 >= 0 and rint + m2 <= len(a) and rh < nhalf):
            break
        X.append(fourier_interp(rh, a, m))
    if not X:
        return np.zeros(2)
    stack = np.zeros(len(X) + 1, dtype=np.complex128)
    stack[1:] = X
    return np.fft.irfft(stack, 2 * len(X))


def measure_ducy(amps, N, r, nharm, fsp=1.5, maxfrac=0.3, m=64):
    """measure_ducy's duty cycle of one candidate, candidate.jl:82-100: NaN
    unless nharm >= 1 and r > 0, else boxcar_best_width of its profile."""
    if not (nharm >= 1 and r > 0):
        return float("nan")
    prof = candidate_profile(amps, N, r, nharm, m)
    return boxcar_best_width(prof, fsp, maxfrac)[1]
