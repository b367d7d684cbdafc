"""cs_snr_metrics_noalloc (device buffers, scratch in a fixed device array)
against cs_snr_metrics (the allocating entry point it is a variant of) on the
same profiles, through the C ABI: bit-identical metrics on two calls in a row,
the second with a smaller input; the scratch limit is an error."""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def eng():
    import coherentsearch as cs
    e = cs.Engine(cs.SearchParams(nharms=32), 2048, device=0)
    yield e
    e.close()


class _Hip:
    """Device buffers through the HIP runtime libcsgpu.so itself links (the
    test's own allocations; the entry point under test makes none)."""

    def __init__(self):
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipDeviceSynchronize.argtypes = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), nbytes) == 0
        return p.value

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0      # host to device
        return p

    def get(self, p, n, dtype=np.float64):
        out = np.empty(n, dtype=dtype)
        assert self.L.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def free(self, p):
        assert self.L.hipFree(p) == 0


def test_noalloc_equals_allocating_entry_point(eng):
    hip = _Hip()
    rng = np.random.default_rng(31)
    cases = [  # (nbins, L, sigma_samples, center, widths)
        (120, 4096, 8192, "zero", None),      # subsampled: 68 whole columns
        (64, 300, None, "median", [1, 2, 3, 5, 8]),   # every sample, smaller input
        (16, 70000, None, "zero", None),      # 1,120,000 > 2^20 samples: refused
        (20, 1000, 8192, "zero", [2, 3, 4, 6]),
    ]
    for nbins, L, ss, center, widths in cases:
        _say(f"noalloc nbins={nbins} L={L} sigma_samples={ss} center={center} widths={widths}")
        P = rng.standard_normal((nbins, L))
        P[:, ::7] *= 3.0
        d_cols = hip.put(P.T.copy())                   # column j contiguous
        d_out = hip.put(np.full(L, np.nan))
        try:
            if min(nbins * L, ss or nbins * L) > (1 << 20):
                with pytest.raises(ValueError, match="2\\^20"):
                    eng.snr_metrics_device(d_cols, nbins, L, d_out, sigma_samples=ss,
                                           sigma_center=center, widths=widths)
                continue
            eng.snr_metrics_device(d_cols, nbins, L, d_out, sigma_samples=ss,
                                   sigma_center=center, widths=widths)
            got = hip.get(d_out, L)
        finally:
            hip.free(d_cols)
            hip.free(d_out)
        want = eng.snr_metrics(P, sigma_samples=ss, sigma_center=center, widths=widths)
        assert np.all(np.isfinite(got))
        assert np.array_equal(got, want), np.max(np.abs(got - want))
        _say(f"  equal, max metric {got.max():.4f}")


def test_noalloc_null_arguments(eng):
    import coherentsearch._lib as L
    st = L.lib().cs_snr_metrics_noalloc(eng._h, None, 16, 4, 1.5, 0.3, 0, 0, None, 0, None)
    assert st == L.CS_EINVAL
