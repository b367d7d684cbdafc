"""ctypes binding of libcsgpu.so (include/coherent_search.h).

The product path has no CPU fallback: if the library is missing or no GPU is
visible, every call raises.
"""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcsgpu.so")
HEADER = os.path.join(_HERE, "..", "..", "include", "coherent_search.h")

CS_OK, CS_EINVAL, CS_EDIM, CS_ENOSPC, CS_EDEVICE, CS_ENOMEM = range(6)
CS_MODE_PLAIN, CS_MODE_FUSED, CS_MODE_FUSED_VALU = 0, 1, 2
CS_TUNE_BATCH_BLOCKS, CS_TUNE_WORKGROUPS, CS_TUNE_GRAPHS, CS_TUNE_TILE_ORDER = 1, 2, 3, 4
CS_TUNE_GENERIC = 5
CS_TUNE_NO_PREFETCH = 6
CS_TRIAGE_FEATURES, CS_TRIAGE_PARAMS = 32, 2177


class DimensionMismatch(ValueError):
    """The reference's DimensionMismatch."""


class CsParams(C.Structure):
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
        ("precision", C.c_int32),
        ("reserved", C.c_int32),
    ]


class CsCand(C.Structure):
    _fields_ = [
        ("t", C.c_int64),
        ("k", C.c_int32),
        ("nharm", C.c_int32),
        ("block", C.c_int64),
        ("metric", C.c_double),
    ]


class CsFile(C.Structure):
    _fields_ = [
        ("amps", C.POINTER(C.c_float)),
        ("namps", C.c_int64),
        ("N", C.c_int64),
        ("T", C.c_double),
        ("r_lo", C.c_double),
        ("r_hi", C.c_double),
    ]


CAND_DTYPE = np.dtype([("t", "<i8"), ("k", "<i4"), ("nharm", "<i4"),
                       ("block", "<i8"), ("metric", "<f8")])


class CsStatsCfg(C.Structure):
    _fields_ = [("hist_lo", C.c_double), ("hist_hi", C.c_double), ("hist_nb", C.c_int32),
                ("nwin", C.c_int32)]


BLOCK_STAT_DTYPE = np.dtype([("block", "<i8"), ("k", "<i4"), ("Hk", "<i4"), ("nbins", "<i4"),
                             ("pad", "<i4"), ("n", "<i8"), ("ngoodbins", "<f8"), ("flo", "<f8"),
                             ("fhi", "<f8"), ("min", "<f8"), ("median", "<f8"), ("mean", "<f8"),
                             ("std", "<f8"), ("max", "<f8")])
HIST_ACC_DTYPE = np.dtype([("k", "<i4"), ("win", "<i4"), ("under", "<i8"), ("over", "<i8"),
                           ("total", "<i8"), ("flo", "<f8"), ("fhi", "<f8"), ("sum", "<f8"),
                           ("sumsq", "<f8"), ("vmin", "<f8"), ("vmax", "<f8")])

_LIB = None


def lib():
    """Load libcsgpu.so, raising loudly if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built: run `make -C coherentsearch.jl_amd` "
                           "(or __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    dp = C.POINTER(C.c_double)
    fp = C.POINTER(C.c_float)
    i32p = C.POINTER(C.c_int32)
    i64p = C.POINTER(C.c_int64)
    vp = C.c_void_p
    L.cs_version.restype = C.c_char_p
    L.cs_last_error.argtypes = [vp]
    L.cs_last_error.restype = C.c_char_p
    L.cs_ctx_create.argtypes = [C.POINTER(CsParams), C.c_int32, C.c_int32, C.POINTER(vp)]
    L.cs_ctx_destroy.argtypes = [vp]
    L.cs_set_spectrum.argtypes = [vp, fp, C.c_int64, C.c_int64, C.c_double]
    L.cs_set_spectrum_device.argtypes = [vp, vp, C.c_int64, C.c_int64, C.c_double]
    L.cs_search.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.POINTER(CsCand),
                            C.c_int64, i64p]
    L.cs_block_metrics.argtypes = [vp, dp, C.c_int64, C.c_int32, i32p, C.c_int32, dp]
    L.cs_search_files.argtypes = [vp, C.c_int32, C.POINTER(CsFile), C.c_double,
                                  C.POINTER(CsCand), C.c_int64, i64p, i64p]
    L.cs_triage_features.argtypes = [vp, C.c_double, C.c_double, fp, C.c_int64, i64p]
    L.cs_mlp_forward.argtypes = [vp, fp, fp, fp, C.c_int64, fp]
    L.cs_mlp_grad.argtypes = [vp, fp, fp, fp, fp, C.c_int64, fp]
    L.cs_mlp_train.argtypes = [vp, fp, fp, fp, fp, C.c_int64, C.c_int32, C.c_float, fp]
    L.cs_triage_features_kind.argtypes = [vp, C.c_int32, C.c_double, C.c_double, fp, C.c_int64,
                                          i64p]
    L.cs_search_blocks.argtypes = [vp, C.c_double, C.c_double, C.c_double,
                                   C.POINTER(C.c_uint8), C.c_int64, C.POINTER(CsCand), C.c_int64,
                                   i64p]
    L.cs_chunk_metrics.argtypes = [vp, C.c_double, C.c_int64, C.c_double, C.c_int32, dp]
    L.cs_search_stats.argtypes = [vp, C.c_double, C.c_double, C.c_double,
                                  C.POINTER(CsStatsCfg), C.POINTER(CsCand), C.c_int64, i64p]
    L.cs_stats_blocks.argtypes = [vp, vp, C.c_int64, i64p]
    L.cs_stats_hists.argtypes = [vp, i64p, vp, C.POINTER(C.c_int32)]
    L.cs_stats_metrics.argtypes = [vp, C.c_int32, dp, C.c_int64]
    L.cs_search_normalized.argtypes = [vp, C.c_double, dp, dp, C.POINTER(CsCand), C.c_int64,
                                       i64p]
    L.cs_reference_profiles.argtypes = [vp, dp, C.c_int64, C.c_int32, dp]
    L.cs_snr_metrics.argtypes = [vp, dp, C.c_int32, C.c_int64, C.c_double, C.c_double,
                                 C.c_int64, C.c_int32, i32p, C.c_int32, dp]
    L.cs_snr_metrics_noalloc.argtypes = [vp, vp, C.c_int32, C.c_int64, C.c_double, C.c_double,
                                         C.c_int64, C.c_int32, i32p, C.c_int32, vp]
    L.cs_fourier_interp.argtypes = [vp, dp, C.c_int64, C.c_int32, dp]
    L.cs_candidate_profiles.argtypes = [vp, dp, i32p, C.c_int64, C.c_int32, dp, C.c_int64, i32p]
    L.cs_measure_ducy.argtypes = [vp, dp, i32p, C.c_int64, C.c_int32, dp]
    L.cs_decim_metrics.argtypes = [vp, C.c_double, C.c_int64, C.c_int32, C.c_int32, dp, i64p]
    L.cs_interp_rows.argtypes = [vp, C.c_double, C.c_int64, C.c_int64, C.c_int32, dp]
    L.cs_set_mode.argtypes = [vp, C.c_int32]
    L.cs_set_tuning.argtypes = [vp, C.c_int32, C.c_int64]
    L.cs_get_tuning.argtypes = [vp, C.c_int32, i64p]
    L.cs_autotune.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_char_p, C.c_int32,
                              i64p, C.POINTER(C.c_int32)]
    L.cs_set_tuning_cache.argtypes = [vp, C.c_char_p]
    L.cs_last_timings.argtypes = [vp, dp]
    L.cs_last_sigmas.argtypes = [vp, dp, C.c_int64, i64p]
    L.cs_probe_split_bf16.argtypes = [vp, C.c_double, C.c_double, C.c_int64, C.c_int64, fp, fp]
    L.cs_plan_tables.argtypes = [C.POINTER(CsParams), C.c_double, C.c_int32, i64p, i32p,
                                 C.POINTER(C.c_uint8), fp, fp]
    _LIB = L
    return L


def declared_symbols(header: str = HEADER):
    """Every function the public header declares."""
    with open(header) as fh:
        txt = fh.read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(cs_\w+)\s*\(", txt, re.M)))


def raise_for(status: int, ctx=None, what: str = ""):
    if status == CS_OK:
        return
    msg = ""
    if ctx is not None:
        try:
            m = lib().cs_last_error(ctx)
            msg = m.decode() if m else ""
        except Exception:  # pragma: no cover
            msg = ""
    text = f"{what}: {msg}" if msg else what
    if status == CS_EINVAL:
        raise ValueError(text)
    if status == CS_EDIM:
        raise DimensionMismatch(text)
    if status == CS_ENOMEM:
        raise MemoryError(text)
    if status == CS_ENOSPC:
        raise OverflowError(text)
    raise RuntimeError(f"{text} (status {status})")


def make_cparams(params):
    dec = np.ascontiguousarray(np.asarray(params.decimations, dtype=np.int32))
    prec = {"f64": 0, "f32": 1}.get(params.precision)
    if prec is None:
        raise ValueError(f"precision must be :f64 or :f32, got :{params.precision}")
    cp = CsParams(int(params.nharms), int(params.m), int(params.numbetween), len(dec),
                  float(params.hidr), float(params.threshold), float(params.boxcar_fsp),
                  float(params.boxcar_maxfrac), float(params.boxcar_medmargin),
                  dec.ctypes.data_as(C.POINTER(C.c_int32)), prec, 0)
    cp._keep = dec
    return cp
