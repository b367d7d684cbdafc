"""CPU tests of the C-ABI library and its host-side logic (no GPU needed).

* libcsgpu.so loads and exports every function include/coherent_search.h
  declares;
* parameter validation happens before any device call;
* the host-only entry point cs_plan_tables reproduces the oracle's
  build_direct_plans tables (directinterp.jl:345-405);
* the native host checks of the split-BF16 kernel's register geometry and
  operand tables compile and pass.
"""
import ctypes as C
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT

import coherentsearch as cs
from coherentsearch import _lib

import oracle as O

CSRC = os.path.join(ROOT, "coherentsearch.jl_amd", "csrc")
NATIVE = os.path.join(ROOT, "tests", "native")


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    declared = _lib.declared_symbols()
    assert len(declared) >= 14
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, missing


def test_version_string():
    v = _lib.lib().cs_version()
    assert b"gfx950" in v


@pytest.mark.parametrize("kw", [dict(m=15), dict(nharms=0), dict(nharms=65)])
def test_ctx_create_rejects_invalid_params(kw):
    """Invalid SearchParams fail with CS_EINVAL before any device is touched."""
    base = dict(nharms=16, m=16, numbetween=16, hidr=0.5, threshold=8.0, boxcar_fsp=1.5,
                boxcar_maxfrac=0.3, boxcar_medmargin=2.0)
    base.update(kw)
    dec = np.array([1], dtype=np.int32)
    p = _lib.CsParams(base["nharms"], base["m"], base["numbetween"], 1, base["hidr"],
                      base["threshold"], base["boxcar_fsp"], base["boxcar_maxfrac"],
                      base["boxcar_medmargin"], dec.ctypes.data_as(C.POINTER(C.c_int32)), 0, 0)
    h = C.c_void_p()
    st = _lib.lib().cs_ctx_create(C.byref(p), 2048, 0, C.byref(h))
    assert st == _lib.CS_EINVAL
    assert not h.value


def test_ctx_create_bad_decimation():
    dec = np.array([1, 0], dtype=np.int32)
    p = _lib.CsParams(16, 16, 16, 2, 0.5, 8.0, 1.5, 0.3, 2.0,
                      dec.ctypes.data_as(C.POINTER(C.c_int32)), 0, 0)
    h = C.c_void_p()
    assert _lib.lib().cs_ctx_create(C.byref(p), 2048, 0, C.byref(h)) == _lib.CS_EINVAL


def test_engine_rejects_bad_precision_on_host():
    with pytest.raises(ValueError):
        cs.Engine(cs.SearchParams(precision="f16"))


def _plan_tables(params, r_lo, h):
    cp = _lib.make_cparams(params)
    info = np.zeros(6, dtype=np.int64)
    st = _lib.lib().cs_plan_tables(C.byref(cp), float(r_lo), int(h),
                                   info.ctypes.data_as(C.POINTER(C.c_int64)), None, None, None,
                                   None)
    assert st == 0
    q, P = int(info[0]), int(info[5])
    row = np.zeros(q, dtype=np.int32)
    carry = np.zeros(P, dtype=np.uint8)
    W = np.zeros(P * params.m, dtype=np.float32)
    A = np.zeros(2 * P, dtype=np.float32)
    st = _lib.lib().cs_plan_tables(C.byref(cp), float(r_lo), int(h),
                                   info.ctypes.data_as(C.POINTER(C.c_int64)),
                                   row.ctypes.data_as(C.POINTER(C.c_int32)),
                                   carry.ctypes.data_as(C.POINTER(C.c_uint8)),
                                   W.ctypes.data_as(C.POINTER(C.c_float)),
                                   A.ctypes.data_as(C.POINTER(C.c_float)))
    assert st == 0
    return info, row, carry, W.reshape(P, params.m), A


@pytest.mark.parametrize("nharms,r_lo", [(60, 0.1 * 2 ** 25 * 64e-6), (8, 2 ** 20 * 64e-6),
                                         (32, 1234.5678)])
def test_plan_tables_match_oracle(nharms, r_lo):
    """The library's host plans equal the oracle's: exact integer tables, and
    FP32 weights / conj(A) equal to the oracle's FP64 values rounded to FP32."""
    params = cs.SearchParams(nharms=nharms, m=16)
    orp = O.Plans(nharms, 16, 0.5, r_lo)
    for h in sorted({1, 2, 3, nharms // 2, nharms - 1, nharms}):
        info, row, carry, W, A = _plan_tables(params, r_lo, h)
        d = orp.info(h)
        assert (int(info[0]), int(info[1]), int(info[2]), int(info[3]), int(info[4]),
                int(info[5])) == (d["q"], d["s"], d["base_adv"], d["pnum"], d["rfloor0"], d["P"])
        orow, ocarry, oW, oA = orp.tables(h)
        # the oracle keeps the reference's 1-based phase index (0 = residue never visited)
        assert np.array_equal(row, np.where(orow > 0, orow - 1, -1))
        assert np.array_equal(carry.astype(bool), ocarry)
        assert np.array_equal(W, oW.astype(np.float32))
        # conj(A) goes through sinpi/cospi products; allow one FP32 ulp
        Af = A.view(np.complex64)
        assert np.allclose(Af.real, oA.real.astype(np.float32), rtol=2 ** -23, atol=1e-30)
        assert np.allclose(Af.imag, oA.imag.astype(np.float32), rtol=2 ** -23, atol=1e-30)


def _native(src, extra=()):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = os.path.join("/tmp", "cs_native_" + os.path.splitext(os.path.basename(src))[0])
    cmd = [gxx, "-std=c++17", "-O2", "-I", CSRC, src, *extra, "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True)
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    return out


def test_register_geometry_host_check():
    """Every boxcar window the split-BF16 kernel forms from lane registers equals
    the direct circular boxcar, each bin starts exactly one window, for every
    supported profile length and bank (tests/native/geo_check.cpp)."""
    out = _native(os.path.join(NATIVE, "geo_check.cpp"))
    assert out.returncode == 0, out.stdout[-2000:]
    assert out.stdout.strip().endswith("OK")


def test_mfma_operand_tables_host_check():
    """Window starts, weight tiles, conj(A), the permuted even/odd c2r tiles and
    the direct c2r tiles in their per-fold-factor K-slot order (against the
    gather's row formula, with the two lane groups of each half-wave in
    disjoint LDS banks) of the split-BF16 kernel, recomputed from the plans
    (tests/native/mfma_tables_check.cpp)."""
    out = _native(os.path.join(NATIVE, "mfma_tables_check.cpp"),
                  [os.path.join(CSRC, "cs_plan.cpp")])
    assert out.returncode == 0, out.stdout[-2000:]
    assert out.stdout.strip().endswith("OK")


def test_trial_grid_and_candidate_r_host():
    """Host trial grid and the FP64 candidate-frequency reconstruction order
    (search.jl:2015-2027, 1812, 1833, 1696)."""
    p = cs.SearchParams(nharms=60, m=16, decimations=cs.decimation_set(60, 6))
    T = 2 ** 25 * 64e-6
    r_lo, r_hi, lodr, total, Nprof, nch = cs.trial_grid(T, p, 0.1, 125.0, 100, 2048)
    assert total == 32186485 and nch == 15717 and Nprof == 2048
    assert lodr == 0.5 / 60
    t = 123457
    i0 = (t // Nprof) * Nprof
    rstart = r_lo + i0 * lodr
    rf = rstart + (t - i0) * lodr
    assert cs.candidate_r(r_lo, lodr, Nprof, t, 1) == rf
    assert cs.candidate_r(r_lo, lodr, Nprof, t, 3) == 3 * rf
    assert math.isfinite(rf)
