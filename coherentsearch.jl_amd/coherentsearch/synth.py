"""Seeded synthetic PRESTO spectra: complex-Gaussian noise plus von Mises pulsars.

* Noise bins are X = (a + i b)/sqrt(2) with a, b ~ N(0, 1) i.i.d. from
  numpy's ``Generator(PCG64(seed))``, so E|X|^2 = 1 (unit mean power).
  Element 0 (DC / Nyquist slot) is 0.
* A pulsar is a von Mises pulse train p(phi) = exp(kappa (cos 2 pi phi - 1)),
  with kappa fixed by the FWHM duty cycle delta: kappa = ln 2 / (1 - cos(pi delta)).
  Its harmonic amplitudes are proportional to I_h(kappa)/I_0(kappa).
* Injection is the exact DFT response: the noiseless time series is built and
  real-FFT'd, then added to the noise bins.
* Injected S/N is the ideal coherent matched-filter S/N of the noiseless
  spectrum S in units of the noise: snr = sqrt(2 * sum_k |S_k|^2) over bins
  1 .. N/2-1 (Re(sum conj(S) X) has noise variance sum |S|^2 / 2).
* Optional red noise: power proportional to f^-1 below ``red_fc`` Hz, then a
  running-median normalisation of the power spectrum (as PRESTO's
  rednoise/accelsearch de-reddening leaves it).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np


@dataclass
class Pulsar:
    freq: float          # Hz
    duty: float          # FWHM duty cycle (fraction of a period)
    snr: float           # injected coherent S/N (definition above)
    phase: float = 0.0   # pulse phase at t=0, cycles


def vonmises_kappa(duty: float) -> float:
    """kappa with FWHM duty cycle `duty`: exp(kappa (cos(pi duty) - 1)) = 1/2."""
    return math.log(2.0) / (1.0 - math.cos(math.pi * duty))


def vonmises_harmonics(duty: float, nh: int) -> np.ndarray:
    """Closed-form relative harmonic amplitudes I_h(kappa)/I_0(kappa), h=1..nh."""
    from scipy.special import ive
    k = vonmises_kappa(duty)
    h = np.arange(1, nh + 1)
    return ive(h, k) / ive(0, k)


def pulsar_timeseries(N: int, dt: float, psr: Pulsar, out: Optional[np.ndarray] = None,
                      chunk: int = 1 << 22) -> np.ndarray:
    """Noiseless, zero-mean von Mises pulse train (unit peak before scaling)."""
    if out is None:
        out = np.zeros(N, dtype=np.float64)
    k = vonmises_kappa(psr.duty)
    # exact zero mean of exp(k(cos x - 1)) over a cycle: I0(k) e^{-k}
    from scipy.special import ive
    mean = ive(0, k)
    for s in range(0, N, chunk):
        n = np.arange(s, min(N, s + chunk), dtype=np.float64)
        ph = psr.freq * dt * n + psr.phase
        ph -= np.floor(ph)
        out[s:s + len(n)] += np.exp(k * (np.cos(2.0 * np.pi * ph) - 1.0)) - mean
    return out


def make_spectrum(N: int, dt: float, seed: int, pulsars: Sequence[Pulsar] = (),
                  red_fc: Optional[float] = None, red_nmed: int = 1000) -> np.ndarray:
    """Return the N/2 complex64 PRESTO amplitudes (element 0 = 0)."""
    nb = N // 2
    rng = np.random.Generator(np.random.PCG64(seed))
    amps = np.empty(nb, dtype=np.complex128)
    amps.real = rng.standard_normal(nb)
    amps.imag = rng.standard_normal(nb)
    amps *= 1.0 / math.sqrt(2.0)
    if red_fc is not None:
        f = np.arange(nb) / (N * dt)
        gain = np.ones(nb)
        low = (f > 0) & (f < red_fc)
        gain[low] = np.sqrt(red_fc / f[low])       # power ~ f^-1 below red_fc
        amps *= gain
    for psr in pulsars:
        ts = pulsar_timeseries(N, dt, psr)
        S = np.fft.rfft(ts)[:nb]
        del ts
        S[0] = 0.0
        e = math.sqrt(2.0 * float(np.sum(np.abs(S[1:]) ** 2)))
        if e > 0:
            amps += S * (psr.snr / e)
        del S
    if red_fc is not None:
        amps = _median_normalise(amps, red_nmed)
    amps[0] = 0.0
    return amps.astype(np.complex64)


def _median_normalise(amps: np.ndarray, nmed: int) -> np.ndarray:
    """Divide each block of `nmed` bins by sqrt(median power / ln 2) so the local
    mean power is 1 (the median of an exponential is ln 2 times its mean)."""
    out = amps.copy()
    nb = len(amps)
    for s in range(1, nb, nmed):
        e = min(nb, s + nmed)
        med = np.median(np.abs(amps[s:e]) ** 2)
        if med > 0:
            out[s:e] /= math.sqrt(med / math.log(2.0))
    return out


# ---- the named workloads --------------------------------------------------

DT = 64e-6


def config_c1():
    """2^20 samples, nharms=8, 1-100 Hz, one 13.7 Hz pulsar (8 % duty, S/N 15)."""
    return dict(N=1 << 20, dt=DT, seed=1001,
                pulsars=[Pulsar(13.7, 0.08, 15.0)],
                nharms=8, lofreq=1.0, hifreq=100.0, threshold=8.0, maxdecim=1)


def config_c2():
    """2^25 samples, nharms=60, decimation_set(60, 6), 0.1-125 Hz, 3 pulsars."""
    return dict(N=1 << 25, dt=DT, seed=2001,
                pulsars=[Pulsar(0.7, 0.03, 10.0), Pulsar(61.3, 0.10, 15.0),
                         Pulsar(412.9, 0.05, 20.0)],
                nharms=60, lofreq=0.1, hifreq=125.0, threshold=8.0, maxdecim=6)


def config_c3_file(i: int):
    """File i of the 64-file DM sweep: length from {2^22..2^25}, one pulsar."""
    rng = np.random.Generator(np.random.PCG64(3000 + i))
    N = int(1 << int(rng.integers(22, 26)))
    f = float(math.exp(rng.uniform(math.log(0.5), math.log(500.0))))
    duty = float(rng.uniform(0.02, 0.15))
    snr = float(rng.uniform(8.0, 25.0))
    return dict(N=N, dt=DT, seed=3000 + i, pulsars=[Pulsar(f, duty, snr)],
                nharms=32, lofreq=0.1, hifreq=125.0, threshold=8.0, maxdecim=4)


def config_c4_file(i: int, npulsars: int = 8):
    """File i of the block-triage training and evaluation set: 2^24 samples,
    seed 4000 + i, npulsars von Mises pulsars with f log-uniform on
    [0.1, 750] Hz, duty uniform on [2, 20] %, S/N uniform on [6, 20]; searched
    like the 2^25-sample workload (nharms 60, folds 1..6, 0.1-125 Hz).  The
    pulsars are drawn from the substream (4000 + i, 1) of numpy's SeedSequence,
    apart from the noise stream of seed 4000 + i."""
    rng = np.random.default_rng([4000 + i, 1])
    psrs = []
    for _ in range(npulsars):
        f = float(math.exp(rng.uniform(math.log(0.1), math.log(750.0))))
        psrs.append(Pulsar(f, float(rng.uniform(0.02, 0.20)), float(rng.uniform(6.0, 20.0)),
                           float(rng.uniform(0.0, 1.0))))
    return dict(N=1 << 24, dt=DT, seed=4000 + i, pulsars=psrs,
                nharms=60, lofreq=0.1, hifreq=125.0, threshold=8.0, maxdecim=6)


def config_c5():
    """2^27 samples with red noise below 1 Hz and five pulsars (incl. 716 Hz)."""
    return dict(N=1 << 27, dt=DT, seed=5001, red_fc=1.0,
                pulsars=[Pulsar(0.9, 0.05, 12.0), Pulsar(7.3, 0.02, 10.0),
                         Pulsar(88.1, 0.10, 15.0), Pulsar(301.7, 0.04, 9.0),
                         Pulsar(716.0, 0.01, 8.0)],
                nharms=60, lofreq=0.1, hifreq=125.0, threshold=8.0, maxdecim=6)


def make_config_spectrum(cfg: dict) -> np.ndarray:
    return make_spectrum(cfg["N"], cfg["dt"], cfg["seed"], cfg.get("pulsars", ()),
                         red_fc=cfg.get("red_fc"))


def make_config_spectra(cfgs: Sequence[dict], workers: int = 8) -> List[np.ndarray]:
    """make_config_spectrum of every config, in order, on `workers` processes
    (spawned, so nothing the parent initialised is shared).  Each spectrum
    depends on its own seed only, so the result equals the serial one."""
    cfgs = list(cfgs)
    if workers <= 1 or len(cfgs) <= 1:
        return [make_config_spectrum(c) for c in cfgs]
    import multiprocessing as mp
    from concurrent.futures import ProcessPoolExecutor
    with ProcessPoolExecutor(max_workers=min(workers, len(cfgs)),
                             mp_context=mp.get_context("spawn")) as ex:
        return list(ex.map(make_config_spectrum, cfgs))
