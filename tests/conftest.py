"""Shared test setup: import paths, the `gpu` marker and seeded synthetic data."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for sub in ("coherentsearch.jl_amd", "oracle"):
    p = os.path.join(ROOT, sub)
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and libcsgpu.so")


@pytest.fixture(scope="session")
def small_spectrum():
    """2^18-sample spectrum (dt = 64 us) with one 13.7 Hz, 8 % duty pulsar."""
    from coherentsearch.synth import Pulsar, make_spectrum
    N, dt = 1 << 18, 64e-6
    amps = make_spectrum(N, dt, seed=11, pulsars=[Pulsar(13.7, 0.08, 25.0)])
    return amps, N, dt


def noise_amps(n, seed):
    rng = np.random.default_rng(seed)
    a = (rng.standard_normal(n) + 1j * rng.standard_normal(n)) / np.sqrt(2)
    a[0] = 0
    return a.astype(np.complex64)
