"""PRESTO ``.inf`` / ``.fft`` reading and writing.

Mirrors ``SimpleInf``/``FFTFile``/``freqs`` of CoherentSearch.jl
(src/fileio.jl:12-85): the ``.fft`` file is raw little-endian complex64 whose
element 0 packs DC (real part) and Nyquist (imaginary part); the ``.inf`` file
supplies the number of time-series samples ``N`` and the sample width ``dt``.
"""
from __future__ import annotations

import os
from dataclasses import dataclass
from typing import Optional

import numpy as np


def _inf_value(line: str) -> str:
    return line.split("=")[-1].strip()


@dataclass
class SimpleInf:
    """The five ``.inf`` keys the search reads (src/fileio.jl:12-46)."""
    path: str
    object: Optional[str] = None
    epoch: Optional[float] = None
    N: Optional[int] = None
    dt: Optional[float] = None
    DM: Optional[float] = None

    @classmethod
    def read(cls, path: str) -> "SimpleInf":
        if not os.path.isfile(path):
            raise FileNotFoundError(f"The .inf file '{path}' was not found.")
        inf = cls(path=str(path))
        with open(path, "r") as fh:
            for line in fh:
                line = line.rstrip("\n")
                if line.startswith(" Object being observed"):
                    inf.object = _inf_value(line)
                elif line.startswith(" Epoch"):
                    inf.epoch = float(_inf_value(line))
                elif line.startswith(" Number of bins"):
                    inf.N = int(_inf_value(line))
                elif line.startswith(" Width of each time series bin"):
                    inf.dt = float(_inf_value(line))
                elif line.startswith(" Dispersion measure"):
                    inf.DM = float(_inf_value(line))
        return inf


class FFTFile:
    """A memory-mapped PRESTO ``.fft`` plus its ``.inf`` metadata
    (src/fileio.jl:48-78).  ``amps`` is a read-only complex64 memmap."""

    def __init__(self, path: str):
        path = str(path)
        inf_path = path[:-4] + ".inf"
        inf = SimpleInf.read(inf_path)
        if inf.N is None:
            raise ValueError(f"Missing 'Number of bins' in {inf_path}")
        if inf.dt is None:
            raise ValueError(f"Missing 'Width of each time series bin' in {inf_path}")
        nbytes = os.path.getsize(path)
        count = nbytes // 8
        self.path = path
        self.inf = inf
        self.amps = (np.memmap(path, dtype=np.complex64, mode="r", shape=(count,))
                     if count else np.zeros(0, dtype=np.complex64))
        self.N = int(inf.N)
        self.T = self.N * float(inf.dt)
        self.df = 1.0 / self.T
        self.dereddened = "_red.fft" in path
        self.detrended = self.dereddened
        a0 = self.amps[0] if count else np.complex64(0)
        self.DC = np.float32(a0.real)
        self.Nyquist = np.float32(a0.imag)

    @classmethod
    def from_arrays(cls, amps: np.ndarray, N: int, dt: float, path: str = "<memory>"):
        """An in-memory FFTFile (the reference's bench/precompile 'fake' file)."""
        self = cls.__new__(cls)
        self.path = path
        self.inf = SimpleInf(path=path, N=int(N), dt=float(dt))
        self.amps = np.ascontiguousarray(np.asarray(amps, dtype=np.complex64))
        self.N = int(N)
        self.T = self.N * float(dt)
        self.df = 1.0 / self.T
        self.dereddened = "_red.fft" in path
        self.detrended = self.dereddened
        self.DC = np.float32(self.amps[0].real) if len(self.amps) else np.float32(0)
        self.Nyquist = np.float32(self.amps[0].imag) if len(self.amps) else np.float32(0)
        return self


def freqs(ft: FFTFile) -> np.ndarray:
    """Fourier frequencies (Hz) of bins 0 .. N//2-1 (src/fileio.jl:85)."""
    return np.arange(ft.N // 2) * ft.df


def write_fft(stem: str, amps: np.ndarray, N: int, dt: float, full_inf: bool = False,
              obj: str = "SYNTH", epoch: float = 50000.0, dm: float = 0.0) -> str:
    """Write ``stem.fft`` (complex64) and ``stem.inf``.  With ``full_inf=False``
    the header carries only N and dt."""
    a = np.ascontiguousarray(np.asarray(amps, dtype=np.complex64))
    fft_path = stem + ".fft"
    a.tofile(fft_path)
    lines = []
    if full_inf:
        lines.append(f" Object being observed                      =  {obj}")
        lines.append(f" Epoch of observation (MJD)                 =  {epoch}")
    lines.append(f" Number of bins in the time series          =  {int(N)}")
    lines.append(f" Width of each time series bin (sec)        =  {float(dt)!r}")
    if full_inf:
        lines.append(f" Dispersion measure (cm-3 pc)               =  {dm}")
    with open(stem + ".inf", "w") as fh:
        fh.write("\n".join(lines) + "\n")
    return fft_path
