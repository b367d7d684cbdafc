"""The device engine and the reference's public search API on top of it.

``search``, ``search_block``, ``block_metrics``, ``chunk_metrics`` and
``SearchCache`` mirror CoherentSearch.jl's functions of the same names
(src/search.jl); the arithmetic runs in libcsgpu.so on the GPU.  Candidate
frequencies are rebuilt on the host from the integer (trial, fold) pairs the
device reports, with the reference's FP64 expression order.
"""
from __future__ import annotations

import ctypes as C
import math
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _lib
from .dedup import remove_duplicates, remove_harmonics
from .fileio import FFTFile
from .params import (Candidate, SearchParams, candidate_r, ladder_boxcar_widths,
                     proftype, trial_grid)


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _i32p(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


class Engine:
    """One device context: the reference's SearchCache plus the spectrum held in
    HBM.  Not thread-safe (one per host thread)."""

    def __init__(self, params: SearchParams, blocksize: int = 2048, device: int = 0):
        proftype(params)
        if params.m % 2 != 0:
            raise ValueError("m must be even")
        self.params = params
        self.blocksize = int(blocksize)
        self._cp = _lib.make_cparams(params)
        h = C.c_void_p()
        st = _lib.lib().cs_ctx_create(C.byref(self._cp), self.blocksize, int(device), C.byref(h))
        if st != 0:
            _lib.raise_for(st, None, "cs_ctx_create (invalid parameters or no GPU)")
        self._h = h
        self._spec_key = None
        self.N = None
        self.T = None
        self._cbuf = None

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().cs_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, st, what):
        _lib.raise_for(st, self._h, what)

    # -- spectrum --------------------------------------------------------------
    def set_spectrum(self, ft: FFTFile):
        key = (id(ft), ft.N, ft.T, len(ft.amps))
        if key == self._spec_key:
            return
        a = np.ascontiguousarray(np.asarray(ft.amps, dtype=np.complex64))
        self._chk(_lib.lib().cs_set_spectrum(self._h, a.view(np.float32).ctypes.data_as(
            C.POINTER(C.c_float)), len(a), int(ft.N), float(ft.T)), "cs_set_spectrum")
        self._spec_key = key
        self.N, self.T = ft.N, ft.T

    def set_spectrum_device(self, ptr: int, namps: int, N: int, T: float):
        """Amplitudes already resident in HBM (e.g. a torch tensor's data_ptr())."""
        self._chk(_lib.lib().cs_set_spectrum_device(self._h, C.c_void_p(ptr), int(namps), int(N),
                                                    float(T)), "cs_set_spectrum_device")
        self._spec_key = ("device", ptr, namps, N, T)
        self.N, self.T = N, T

    def set_mode(self, mode: str):
        m = {"plain": _lib.CS_MODE_PLAIN, "fused": _lib.CS_MODE_FUSED,
             "fused_valu": _lib.CS_MODE_FUSED_VALU}[mode]
        self._chk(_lib.lib().cs_set_mode(self._h, m), "cs_set_mode")

    def set_tuning(self, batch_blocks: Optional[int] = None, workgroups: Optional[int] = None,
                   graphs: Optional[bool] = None, tile_order: Optional[int] = None,
                   generic: Optional[bool] = None, no_prefetch: Optional[bool] = None):
        """Launch shape of the search (0 leaves the default, None leaves the
        knob as it is), graph replay of the search's launches, (generic) the
        split-BF16 kernel that dispatches on profile lengths at run time
        instead of one specialised on the fold set, and (no_prefetch) that
        kernel without its LDS prefetch of each tile's first conj(A)
        operands; results do not depend on them."""
        if no_prefetch is not None:
            self._chk(_lib.lib().cs_set_tuning(self._h, _lib.CS_TUNE_NO_PREFETCH,
                                               int(bool(no_prefetch))), "cs_set_tuning")
        if generic is not None:
            self._chk(_lib.lib().cs_set_tuning(self._h, _lib.CS_TUNE_GENERIC, int(bool(generic))),
                      "cs_set_tuning")
        if tile_order is not None:
            self._chk(_lib.lib().cs_set_tuning(self._h, _lib.CS_TUNE_TILE_ORDER, int(tile_order)),
                      "cs_set_tuning")
        if graphs is not None:
            self._chk(_lib.lib().cs_set_tuning(self._h, _lib.CS_TUNE_GRAPHS, int(bool(graphs))),
                      "cs_set_tuning")
        if batch_blocks is not None:
            self._chk(_lib.lib().cs_set_tuning(self._h, _lib.CS_TUNE_BATCH_BLOCKS,
                                               int(batch_blocks)), "cs_set_tuning")
        if workgroups is not None:
            self._chk(_lib.lib().cs_set_tuning(self._h, _lib.CS_TUNE_WORKGROUPS, int(workgroups)),
                      "cs_set_tuning")

    def tuning(self) -> dict:
        """Every CS_TUNE_* knob's current value (cs_get_tuning)."""
        out = {}
        for name, knob in (("batch_blocks", _lib.CS_TUNE_BATCH_BLOCKS),
                           ("workgroups", _lib.CS_TUNE_WORKGROUPS),
                           ("graphs", _lib.CS_TUNE_GRAPHS), ("tile_order", _lib.CS_TUNE_TILE_ORDER),
                           ("generic", _lib.CS_TUNE_GENERIC),
                           ("no_prefetch", _lib.CS_TUNE_NO_PREFETCH)):
            v = C.c_int64()
            self._chk(_lib.lib().cs_get_tuning(self._h, knob, C.byref(v)), "cs_get_tuning")
            out[name] = int(v.value)
        return out

    def autotune(self, r_lo: float, r_hi: float, threshold: float,
                 cache_path: Optional[str] = None, reps: int = 2) -> Tuple[int, int]:
        """Time the split-BF16 search's launch shapes for this band and keep the
        fastest (cs_autotune); returns (blocks per batch, workgroups), 0 = default."""
        b = C.c_int64()
        w = C.c_int32()
        self._chk(_lib.lib().cs_autotune(self._h, float(r_lo), float(r_hi), float(threshold),
                                         cache_path.encode() if cache_path else None, int(reps),
                                         C.byref(b), C.byref(w)), "cs_autotune")
        return int(b.value), int(w.value)

    def set_tuning_cache(self, path: Optional[str]):
        """The tuning cache search_files reads: each length class of a sweep
        runs with the launch shape autotune stored for its geometry there (no
        timing during the sweep); None clears it (cs_set_tuning_cache)."""
        self._chk(_lib.lib().cs_set_tuning_cache(self._h, path.encode() if path else None),
                  "cs_set_tuning_cache")

    # -- entry points ------------------------------------------------------------
    def _out(self, cap: int):
        """The engine's candidate output buffer, grown to `cap` records (kept
        between calls: a fresh zero-filled 2 MB array per search would cost
        more host time than a small search takes on the device)."""
        if self._cbuf is None or len(self._cbuf) < cap:
            self._cbuf = np.empty(cap, dtype=_lib.CAND_DTYPE)
        return self._cbuf.ctypes.data_as(C.POINTER(_lib.CsCand))

    def search_raw(self, r_lo: float, r_hi: float, threshold: float) -> np.ndarray:
        cap = 1 << 16
        while True:
            out = self._out(cap)
            n = C.c_int64()
            st = _lib.lib().cs_search(self._h, float(r_lo), float(r_hi), float(threshold), out,
                                      cap, C.byref(n))
            if st == _lib.CS_ENOSPC and n.value > cap:
                cap = n.value
                continue
            self._chk(st, "cs_search")
            return self._cbuf[:n.value].copy()

    def _cands(self, call, what):
        cap = 1 << 16
        while True:
            out = self._out(cap)
            n = C.c_int64()
            st = call(out, cap, n)
            if st == _lib.CS_ENOSPC and n.value > cap:
                cap = n.value
                continue
            self._chk(st, what)
            return self._cbuf[:n.value].copy()

    def search_stats(self, r_lo: float, r_hi: float, threshold: float, ms) -> np.ndarray:
        """cs_search_stats: raw candidates as search_raw, with every trial's
        metric exact; appends the block records and histograms to the
        MetricStats `ms` (the collecting _search_region!, search.jl:1768-1894)."""
        from .metricstats import fill_from_device
        cfg = _lib.CsStatsCfg(float(ms.hist_lo), float(ms.hist_hi), int(ms.hist_nb), int(ms.nwin))
        raw = self._cands(lambda out, cap, n: _lib.lib().cs_search_stats(
            self._h, float(r_lo), float(r_hi), float(threshold), C.byref(cfg), out, cap,
            C.byref(n)), "cs_search_stats")
        nb = C.c_int64()
        self._chk(_lib.lib().cs_stats_blocks(self._h, None, 0, C.byref(nb)), "cs_stats_blocks")
        blocks = np.zeros(nb.value, dtype=_lib.BLOCK_STAT_DTYPE)
        self._chk(_lib.lib().cs_stats_blocks(self._h, blocks.ctypes.data, nb.value, C.byref(nb)),
                  "cs_stats_blocks")
        nf = C.c_int32()
        self._chk(_lib.lib().cs_stats_hists(self._h, None, None, C.byref(nf)), "cs_stats_hists")
        counts = np.zeros((nf.value, ms.nwin, int(ms.hist_nb)), dtype=np.int64)
        acc = np.zeros((nf.value, ms.nwin), dtype=_lib.HIST_ACC_DTYPE)
        self._chk(_lib.lib().cs_stats_hists(self._h, counts.ctypes.data_as(C.POINTER(C.c_int64)),
                                            acc.ctypes.data, C.byref(nf)), "cs_stats_hists")
        nh = self.params.nharms
        fill_from_device(ms, blocks, counts, acc, {int(a["k"]): nh // int(a["k"]) for a in acc[:, 0]})
        return raw

    def stats_metrics(self, k: int, total: int) -> np.ndarray:
        """The last search_stats call's metric of every trial of fold k (NaN
        where the fold has no valid trial)."""
        out = np.zeros(int(total))
        self._chk(_lib.lib().cs_stats_metrics(self._h, int(k), _dp(out), int(total)),
                  "cs_stats_metrics")
        return out

    def search_normalized(self, threshold: float, norm) -> np.ndarray:
        """cs_search_normalized: the --normalize detection pass over the last
        search_stats call's metrics, candidates carrying z."""
        ks = sorted(norm.loc.keys())
        loc = np.ascontiguousarray(np.concatenate([norm.loc[k] for k in ks]), dtype=np.float64)
        scale = np.ascontiguousarray(np.concatenate([norm.scale[k] for k in ks]), dtype=np.float64)
        return self._cands(lambda out, cap, n: _lib.lib().cs_search_normalized(
            self._h, float(threshold), _dp(loc), _dp(scale), out, cap, C.byref(n)),
            "cs_search_normalized")

    def search_files(self, files: Sequence[Tuple[FFTFile, float, float]],
                     threshold: float) -> List[np.ndarray]:
        """Raw candidates of several spectra, each (ft, r_lo, r_hi), from one
        cs_search_files call; the next file uploads while one is searched."""
        keep, recs = [], (_lib.CsFile * len(files))()
        for i, (ft, r_lo, r_hi) in enumerate(files):
            a = np.ascontiguousarray(np.asarray(ft.amps, dtype=np.complex64))
            keep.append(a)
            recs[i] = _lib.CsFile(a.view(np.float32).ctypes.data_as(C.POINTER(C.c_float)), len(a),
                                  int(ft.N), float(ft.T), float(r_lo), float(r_hi))
        counts = np.zeros(len(files), dtype=np.int64)
        cap = 1 << 16
        while True:
            out = self._out(cap)
            n = C.c_int64()
            st = _lib.lib().cs_search_files(self._h, len(files), recs, float(threshold), out, cap,
                                            counts.ctypes.data_as(C.POINTER(C.c_int64)),
                                            C.byref(n))
            self._spec_key = None   # the context holds no spectrum afterwards
            if st == _lib.CS_ENOSPC and n.value > cap:
                cap = n.value
                continue
            self._chk(st, "cs_search_files")
            allc = self._cbuf[:n.value].copy()
            offs = np.concatenate([[0], np.cumsum(counts)])
            return [allc[offs[i]:offs[i + 1]] for i in range(len(files))]

    # -- block triage ------------------------------------------------------------
    def triage_features(self, r_lo: float, r_hi: float, kind: int = 0) -> np.ndarray:
        """[block][32] features of the band's search blocks (cs_triage_features_kind):
        kind 0 from power-pyramid cell maxima per integer fundamental, kind 1
        from the nearest-bin power at the harmonics of every trial."""
        L = _lib.lib()
        nb = C.c_int64()
        st = L.cs_triage_features_kind(self._h, int(kind), float(r_lo), float(r_hi), None, 0,
                                       C.byref(nb))
        if st not in (_lib.CS_OK, _lib.CS_ENOSPC):
            self._chk(st, "cs_triage_features")
        out = np.zeros((nb.value, _lib.CS_TRIAGE_FEATURES), dtype=np.float32)
        if nb.value:
            self._chk(L.cs_triage_features_kind(self._h, int(kind), float(r_lo), float(r_hi),
                                                _fp(out), nb.value, C.byref(nb)),
                      "cs_triage_features")
        return out

    def mlp_forward(self, params: np.ndarray, norm: np.ndarray, X: np.ndarray) -> np.ndarray:
        p, nm, x = _f32(params), _f32(norm), _f32(X)
        out = np.zeros(len(x), dtype=np.float32)
        self._chk(_lib.lib().cs_mlp_forward(self._h, _fp(p), _fp(nm), _fp(x), len(x), _fp(out)),
                  "cs_mlp_forward")
        return out

    def mlp_grad(self, params: np.ndarray, norm: np.ndarray, X: np.ndarray, y: np.ndarray):
        """(gradient in the params layout, mean binary cross-entropy)"""
        p, nm, x, yy = _f32(params), _f32(norm), _f32(X), _f32(y)
        g = np.zeros(_lib.CS_TRIAGE_PARAMS + 1, dtype=np.float32)
        self._chk(_lib.lib().cs_mlp_grad(self._h, _fp(p), _fp(nm), _fp(x), _fp(yy), len(x),
                                         _fp(g)), "cs_mlp_grad")
        return g[:-1], float(g[-1])

    def mlp_train(self, params: np.ndarray, norm: np.ndarray, X: np.ndarray, y: np.ndarray,
                  epochs: int, lr: float):
        """(trained parameters, per-step losses): cs_mlp_train, every step on the device"""
        p, nm, x, yy = _f32(params).copy(), _f32(norm), _f32(X), _f32(y)
        losses = np.zeros(int(epochs), dtype=np.float32)
        self._chk(_lib.lib().cs_mlp_train(self._h, _fp(p), _fp(nm), _fp(x), _fp(yy), len(x),
                                          int(epochs), float(lr), _fp(losses)), "cs_mlp_train")
        return p, losses

    def search_blocks(self, r_lo: float, r_hi: float, threshold: float,
                      mask: np.ndarray) -> np.ndarray:
        """Raw candidates of the blocks with mask != 0 only (cs_search_blocks)."""
        m = np.ascontiguousarray(np.asarray(mask, dtype=np.uint8))
        cap = 1 << 16
        while True:
            out = self._out(cap)
            n = C.c_int64()
            st = _lib.lib().cs_search_blocks(self._h, float(r_lo), float(r_hi), float(threshold),
                                             m.ctypes.data_as(C.POINTER(C.c_uint8)), len(m), out,
                                             cap, C.byref(n))
            if st == _lib.CS_ENOSPC and n.value > cap:
                cap = n.value
                continue
            self._chk(st, "cs_search_blocks")
            return self._cbuf[:n.value].copy()

    def timings(self) -> dict:
        a = np.zeros(8)
        self._chk(_lib.lib().cs_last_timings(self._h, _dp(a)), "cs_last_timings")
        keys = ("interp", "c2r", "sigma", "gate", "rescan", "fused", "total", "batches")
        return dict(zip(keys, a.tolist()))

    def last_sigmas(self) -> np.ndarray:
        """(nblocks, nfolds) sigma of every (block, fold) of the last search
        over a whole range (_block_sigma inside _search_region!, search.jl:
        1817, and decim_pass!, :1688); folds: 1, then the decimations > 1."""
        n = C.c_int64()
        self._chk(_lib.lib().cs_last_sigmas(self._h, None, 0, C.byref(n)), "cs_last_sigmas")
        a = np.zeros(n.value)
        self._chk(_lib.lib().cs_last_sigmas(self._h, _dp(a), n.value, C.byref(n)),
                  "cs_last_sigmas")
        nf = 1 + sum(1 for k in self.params.decimations if k > 1)
        return a.reshape(-1, nf)

    def probe_split_bf16(self, r_lo: float, r_hi: float, t0: int, n: int):
        """The split-BF16 pass's stacked amplitudes ((nharms, n) complex64, as its
        DFT reads them) and per-fold bounds before sigma ((folds, n) float32) of
        the band's trials t0 .. t0 + n - 1 (cs_probe_split_bf16); fused mode."""
        nh = self.params.nharms
        nf = 1 + sum(1 for k in self.params.decimations if k > 1)
        amps = np.zeros(2 * nh * n, dtype=np.float32)
        bnd = np.zeros(nf * n, dtype=np.float32)
        self._chk(_lib.lib().cs_probe_split_bf16(self._h, float(r_lo), float(r_hi), int(t0),
                                                 int(n), _fp(amps), _fp(bnd)),
                  "probe_split_bf16")
        return amps.view(np.complex64).reshape(nh, n), bnd.reshape(nf, n)

    def chunk_metrics(self, rstart: float, n: int, lodr: Optional[float] = None,
                      weights_f64: bool = False) -> np.ndarray:
        out = np.zeros(max(n, 0))
        lodr = self.params.hidr / self.params.nharms if lodr is None else lodr
        self._chk(_lib.lib().cs_chunk_metrics(self._h, float(rstart), int(n), float(lodr),
                                              int(bool(weights_f64)), _dp(out)), "chunk_metrics")
        return out

    def decim_metrics(self, rstart: float, n: int, k: int, weights_f64: bool = False):
        out = np.zeros(max(n, 0))
        nv = C.c_int64()
        self._chk(_lib.lib().cs_decim_metrics(self._h, float(rstart), int(n), int(k),
                                              int(bool(weights_f64)), _dp(out), C.byref(nv)),
                  "decim_metrics")
        return out[:nv.value]

    def block_metrics(self, rfund: Sequence[float], kernel: str = "fft",
                      widths: Optional[Sequence[int]] = None) -> np.ndarray:
        if kernel not in ("fft", "direct"):
            raise ValueError(f"kernel must be :fft or :direct, got :{kernel}")
        r = np.ascontiguousarray(np.asarray(rfund, dtype=np.float64))
        out = np.zeros(len(r))
        if widths is None:
            wp, nw = None, 0
        else:
            w = np.ascontiguousarray(np.asarray(widths, dtype=np.int32))
            wp, nw = w.ctypes.data_as(C.POINTER(C.c_int32)), len(w)
        self._chk(_lib.lib().cs_block_metrics(self._h, _dp(r), len(r),
                                              1 if kernel == "direct" else 0, wp, nw, _dp(out)),
                  "block_metrics")
        return out

    def candidate_profiles(self, r: Sequence[float], nharm: Sequence[int], m: int = 64):
        """candidate_profile (src/candidate.jl:37-61) for every (r, nharm):
        list of unrotated FP64 profiles of 2H bins (H harmonics below Nyquist)."""
        rv = np.ascontiguousarray(np.asarray(r, dtype=np.float64))
        hv = np.ascontiguousarray(np.asarray(nharm, dtype=np.int32))
        if len(rv) != len(hv):
            raise ValueError("r and nharm differ in length")
        n = len(rv)
        if n == 0:
            return []
        ld = max(2, 2 * int(hv.max()))
        prof = np.zeros((n, ld))
        ln = np.zeros(n, dtype=np.int32)
        self._chk(_lib.lib().cs_candidate_profiles(self._h, _dp(rv), _i32p(hv), n, int(m),
                                                   _dp(prof), ld, _i32p(ln)),
                  "candidate_profile")
        return [prof[i, :ln[i]].copy() for i in range(n)]

    def measure_ducy(self, r: Sequence[float], nharm: Sequence[int], m: int = 64) -> np.ndarray:
        """measure_ducy's duty cycles (src/candidate.jl:82-100): best boxcar
        width / profile bins, NaN unless nharm >= 1 and r > 0."""
        rv = np.ascontiguousarray(np.asarray(r, dtype=np.float64))
        hv = np.ascontiguousarray(np.asarray(nharm, dtype=np.int32))
        if len(rv) != len(hv):
            raise ValueError("r and nharm differ in length")
        out = np.full(len(rv), np.nan)
        if len(rv):
            self._chk(_lib.lib().cs_measure_ducy(self._h, _dp(rv), _i32p(hv), len(rv), int(m),
                                                 _dp(out)), "measure_ducy")
        return out

    def reference_profiles(self, rfund: Sequence[float], kernel: str = "fft") -> np.ndarray:
        """reference_profiles (src/search.jl:1011-1051): (2 nharms, L) FP64."""
        if kernel not in ("fft", "direct"):
            raise ValueError(f"kernel must be :fft or :direct, got :{kernel}")
        r = np.ascontiguousarray(np.asarray(rfund, dtype=np.float64))
        nb = 2 * self.params.nharms
        out = np.zeros((len(r), nb))
        if len(r):
            self._chk(_lib.lib().cs_reference_profiles(self._h, _dp(r), len(r),
                                                       1 if kernel == "direct" else 0, _dp(out)),
                      "reference_profiles")
        return out.T.copy()

    def snr_metrics(self, profs, boxcar_fsp: float = 1.5, boxcar_maxfrac: float = 0.3,
                    sigma_samples: Optional[int] = None, sigma_center: str = "zero",
                    widths: Optional[Sequence[int]] = None) -> np.ndarray:
        """snr_metrics (src/search.jl:933-952) of the columns of an (nbins, L)
        profile matrix."""
        if sigma_center not in ("zero", "median"):
            raise ValueError(f"sigma_center must be :zero or :median, got :{sigma_center}")
        P = np.asarray(profs, dtype=np.float64)
        if P.ndim != 2:
            raise ValueError("profs must be an (nbins, L) matrix")
        nbins, L = P.shape
        cols = np.ascontiguousarray(P.T)            # column j contiguous
        out = np.zeros(L)
        if widths is None:
            wp, nw = None, 0
        else:
            w = np.ascontiguousarray(np.asarray(widths, dtype=np.int32))
            wp, nw = _i32p(w), len(w)
        ss = 0 if sigma_samples is None else int(sigma_samples)
        self._chk(_lib.lib().cs_snr_metrics(self._h, _dp(cols), int(nbins), int(L),
                                            float(boxcar_fsp), float(boxcar_maxfrac), ss,
                                            1 if sigma_center == "median" else 0, wp, nw,
                                            _dp(out)), "snr_metrics")
        return out

    def snr_metrics_device(self, d_profs: int, nbins: int, L: int, d_out: int,
                           boxcar_fsp: float = 1.5, boxcar_maxfrac: float = 0.3,
                           sigma_samples: Optional[int] = None, sigma_center: str = "zero",
                           widths: Optional[Sequence[int]] = None) -> None:
        """snr_metrics on device memory with no allocation (cs_snr_metrics_noalloc):
        d_profs is the address of an nbins x L FP64 matrix, column j at
        d_profs + 8 j nbins; the L metrics are written at d_out."""
        if sigma_center not in ("zero", "median"):
            raise ValueError(f"sigma_center must be :zero or :median, got :{sigma_center}")
        if widths is None:
            wp, nw = None, 0
        else:
            w = np.ascontiguousarray(np.asarray(widths, dtype=np.int32))
            wp, nw = _i32p(w), len(w)
        ss = 0 if sigma_samples is None else int(sigma_samples)
        self._chk(_lib.lib().cs_snr_metrics_noalloc(
            self._h, C.c_void_p(int(d_profs)), int(nbins), int(L), float(boxcar_fsp),
            float(boxcar_maxfrac), ss, 1 if sigma_center == "median" else 0, wp, nw,
            C.c_void_p(int(d_out))), "snr_metrics_noalloc")

    def fourier_interp(self, r, m: int) -> np.ndarray:
        """fourier_interp (src/fourierinterp.jl:66-74) at each r: complex128."""
        rv = np.ascontiguousarray(np.atleast_1d(np.asarray(r, dtype=np.float64)))
        out = np.zeros(2 * len(rv))
        if len(rv):
            self._chk(_lib.lib().cs_fourier_interp(self._h, _dp(rv), len(rv), int(m), _dp(out)),
                      "fourier_interp")
        return out.view(np.complex128)

    def interp_rows(self, r_lo: float, t0: int, n: int, weights_f64: bool = False) -> np.ndarray:
        nh = self.params.nharms
        out = np.zeros(2 * (nh + 1) * max(n, 0))
        self._chk(_lib.lib().cs_interp_rows(self._h, float(r_lo), int(t0), int(n),
                                            int(bool(weights_f64)), _dp(out)), "interp_rows")
        return out.view(np.complex128).reshape(nh + 1, max(n, 0))


class SearchCache:
    """Reusable per-SearchParams device context (src/search.jl:1914-1961).
    Reuse is keyed on the params object's identity and the block size."""

    def __init__(self):
        self.params = None
        self.Nprof = 0
        self.engine: Optional[Engine] = None

    def get(self, params: SearchParams, Nprof: int, device: int = 0) -> Engine:
        if self.engine is not None and self.params is params and self.Nprof == Nprof:
            return self.engine
        if self.engine is not None:
            self.engine.close()
        self.engine = Engine(params, Nprof, device)
        self.params = params
        self.Nprof = Nprof
        return self.engine


def raw_to_candidates(raw: np.ndarray, ft: FFTFile, r_lo: float, lodr: float,
                      Nprof: int) -> List[Candidate]:
    """Candidate(freq, metric, r, nharm) from the device's integer records."""
    out = []
    for rec in raw:
        r = candidate_r(r_lo, lodr, Nprof, int(rec["t"]), int(rec["k"]))
        out.append(Candidate(r / ft.T, float(rec["metric"]), r, int(rec["nharm"])))
    return out


def search(ft: FFTFile, params: Optional[SearchParams] = None, lofreq: float = 0.1,
           hifreq: float = 100.0, lobin: int = 100, blocksize: int = 2048,
           threshold: Optional[float] = None, remove: bool = True, dr_tol: float = 1.0,
           harm_remove: bool = True, numharm: int = 16, harm_tol: float = 1.0,
           progress: str = "none", metricstats=None, normalize: bool = False,
           verbose: bool = False, wisdom: bool = True, wisdom_file=None,
           cache: Optional[SearchCache] = None, device: int = 0,
           mode: str = "auto") -> List[Candidate]:
    """search(ft, params; ...) (src/search.jl:2002-2111) on the GPU.

    ``wisdom``/``wisdom_file`` are FFTW plan caches with no device counterpart
    and are accepted for signature parity.  ``metricstats`` (a MetricStats)
    collects the every-trial statistics; ``normalize`` runs the two-pass
    search (search.jl:2076-2095): statistics, the MetricNorm model, then the
    detection over the normalised metrics (kept on the device from pass 1).
    ``mode`` "auto" runs the split-BF16 fused pipeline and, for parameters it
    does not cover (profile lengths, boxcar banks, irrational trial steps),
    the plain FP32 one; "fused", "fused_valu" and "plain" force a mode.  Under
    the default precision :f64 every mode reports the same candidates."""
    from .metricstats import MetricStats, build_metricnorm
    params = SearchParams() if params is None else params
    if progress not in ("none", "text", "bar"):
        raise ValueError(f"progress must be :none, :text or :bar, got :{progress}")
    thr = params.threshold if threshold is None else float(threshold)
    r_lo, r_hi, lodr, total, Nprof, nchunks = trial_grid(ft.T, params, lofreq, hifreq, lobin,
                                                         blocksize)
    if total == 0:
        return []
    own = cache is None
    eng = Engine(params, Nprof, device) if own else cache.get(params, Nprof, device)
    try:
        eng.set_spectrum(ft)

        def run():
            if normalize:
                ms = metricstats if metricstats is not None else MetricStats()
                eng.search_stats(r_lo, r_hi, math.inf, ms)
                return eng.search_normalized(thr, build_metricnorm(ms))
            if metricstats is not None:
                return eng.search_stats(r_lo, r_hi, thr, metricstats)
            return eng.search_raw(r_lo, r_hi, thr)

        if mode == "auto":
            eng.set_mode("fused")
            try:
                raw = run()
            except ValueError as e:
                if "fused" not in str(e):
                    raise
                eng.set_mode("plain")
                raw = run()
        else:
            eng.set_mode(mode)
            raw = run()
    finally:
        if own:
            eng.close()
    cands = raw_to_candidates(raw, ft, r_lo, lodr, Nprof)
    if verbose:
        import sys
        print(f"Search: {total} trials in {nchunks} chunks of {Nprof} over r = {r_lo:.1f} … "
              f"{r_hi:.1f} bins, {len(cands)} above threshold", file=sys.stderr)
    if remove:
        cands = remove_duplicates(cands, dr_tol=dr_tol)
    else:
        cands = sorted(cands, key=lambda c: c.freq)
    if harm_remove:
        cands = remove_harmonics(cands, numharm=numharm, tol=harm_tol)
    return cands


def search_files(fts: Sequence[FFTFile], params: Optional[SearchParams] = None,
                 lofreq: float = 0.1, hifreq: float = 100.0, lobin: int = 100,
                 blocksize: int = 2048, threshold: Optional[float] = None, remove: bool = True,
                 dr_tol: float = 1.0, harm_remove: bool = True, numharm: int = 16,
                 harm_tol: float = 1.0, device: int = 0,
                 mode: str = "auto") -> List[List[Candidate]]:
    """search over several files with one device context, as the CLI's
    search_one loop shares one SearchCache (src/cli.jl:231-289); each file's
    list equals search(ft, params; ...) on that file.  The next file's upload
    runs while a file is searched.  ``mode`` as in search()."""
    params = SearchParams() if params is None else params
    thr = params.threshold if threshold is None else float(threshold)
    grids = [trial_grid(ft.T, params, lofreq, hifreq, lobin, blocksize) for ft in fts]
    out: List[List[Candidate]] = [[] for _ in fts]
    todo = [i for i, g in enumerate(grids) if g[3] > 0]
    if not todo:
        return out
    eng = Engine(params, grids[todo[0]][4], device)
    jobs = [(fts[i], grids[i][0], grids[i][1]) for i in todo]
    try:
        if mode == "auto":
            eng.set_mode("fused")
            try:
                raws = eng.search_files(jobs, thr)
            except ValueError as e:
                if "fused" not in str(e):
                    raise
                eng.set_mode("plain")
                raws = eng.search_files(jobs, thr)
        else:
            eng.set_mode(mode)
            raws = eng.search_files(jobs, thr)
    finally:
        eng.close()
    for i, raw in zip(todo, raws):
        r_lo, r_hi, lodr, total, Nprof, nchunks = grids[i]
        cands = raw_to_candidates(raw, fts[i], r_lo, lodr, Nprof)
        cands = remove_duplicates(cands, dr_tol=dr_tol) if remove else \
            sorted(cands, key=lambda c: c.freq)
        if harm_remove:
            cands = remove_harmonics(cands, numharm=numharm, tol=harm_tol)
        out[i] = cands
    return out


def block_metrics(ft: FFTFile, rfund, params: SearchParams, kernel: str = "fft",
                  widths=None, engine: Optional[Engine] = None) -> np.ndarray:
    """block_metrics (src/search.jl:1061-1079): kernel :fft (the reference's
    default) or :direct, both on the device in FP64."""
    eng = engine or Engine(params, 2048)
    eng.set_spectrum(ft)
    return eng.block_metrics(rfund, kernel=kernel, widths=widths)


def search_block(ft: FFTFile, rfund, params: SearchParams, threshold: Optional[float] = None,
                 kernel: str = "fft", engine: Optional[Engine] = None) -> List[Candidate]:
    """search_block (src/search.jl:1087-1097)."""
    thr = params.threshold if threshold is None else threshold
    m = block_metrics(ft, rfund, params, kernel=kernel, engine=engine)
    return [Candidate(float(r) / ft.T, float(v), float(r), params.nharms)
            for r, v in zip(rfund, m) if v > thr]


def reference_profiles(ft: FFTFile, rfund, params: SearchParams, kernel: str = "fft",
                       engine: Optional[Engine] = None) -> np.ndarray:
    """reference_profiles (src/search.jl:1011-1051) on the device."""
    eng = engine or Engine(params, 2048)
    eng.set_spectrum(ft)
    return eng.reference_profiles(rfund, kernel=kernel)


def snr_metrics(profs, boxcar_fsp: float = 1.5, boxcar_maxfrac: float = 0.3,
                sigma_samples: Optional[int] = None, sigma_center: str = "zero", widths=None,
                engine: Optional[Engine] = None) -> np.ndarray:
    """snr_metrics (src/search.jl:933-952) on the device."""
    eng = engine or Engine(SearchParams(), 2048)
    return eng.snr_metrics(profs, boxcar_fsp, boxcar_maxfrac, sigma_samples, sigma_center, widths)


def fourier_interp(r, ft: FFTFile, m: int, engine: Optional[Engine] = None):
    """fourier_interp (src/fourierinterp.jl:66-74) on the device; a scalar r
    gives a complex, an array of r an array."""
    eng = engine or Engine(SearchParams(), 2048)
    eng.set_spectrum(ft)
    out = eng.fourier_interp(r, m)
    return complex(out[0]) if np.ndim(r) == 0 else out


def coherent_profiles(ftprofs, nbins: int) -> np.ndarray:
    """coherent_profiles (src/search.jl:964-965): the irfft of stacked harmonic
    amplitudes along the first axis, as the reference's one-line wrapper."""
    return np.fft.irfft(np.asarray(ftprofs), int(nbins), axis=0)


def chunk_ngoodbins(ft: FFTFile, nharms: int, rmean: float) -> float:
    """chunk_ngoodbins (src/search.jl:206-207)."""
    return min(ft.N / 2 / rmean, float(nharms))


def chunk_metrics(ft: FFTFile, params: SearchParams, rstart: float, n: int,
                  lodr: Optional[float] = None, weights: str = "f32",
                  engine: Optional[Engine] = None) -> np.ndarray:
    """chunk_metrics (src/search.jl:1740-1755)."""
    eng = engine or Engine(params, max(1, n))
    eng.set_spectrum(ft)
    return eng.chunk_metrics(rstart, n, lodr, weights_f64=(weights == "f64"))
