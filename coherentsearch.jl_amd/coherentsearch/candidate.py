"""Post-detection steps and the candidate table (src/candidate.jl, src/cli.jl).

* candidate_profile / measure_ducy run on the device (cs_candidate_profiles,
  cs_measure_ducy): FP64 fourier_interp of each harmonic, FP64 inverse real DFT,
  best boxcar width.
* rotate_to_peak, report_candidates (search_one's ranking, cli.jl:296-310) and
  the `.cohout` table (write_candidates, cli.jl:317-340; output_path,
  cli.jl:260-267) are host code, as in the reference.
"""
import math
import os
import sys
from dataclasses import replace
from typing import List, Optional, Sequence

import numpy as np

from .engine import Engine
from .fileio import FFTFile
from .params import Candidate, SearchParams


def _engine(ft: FFTFile, params: Optional[SearchParams], engine: Optional[Engine]) -> Engine:
    eng = engine or Engine(params or SearchParams(), 2048)
    eng.set_spectrum(ft)
    return eng


def candidate_profile(ft: FFTFile, r: float, nharm: int, m: int = 64,
                      engine: Optional[Engine] = None) -> np.ndarray:
    """candidate_profile(ft, r, nharm; m=64) (src/candidate.jl:37-61): the
    unrotated 2H-bin profile of the first H harmonics of r that fit below
    Nyquist (two zero bins if none)."""
    if nharm < 1:
        raise ValueError("nharm must be >= 1")
    if m % 2:
        raise ValueError("m must be even")
    return _engine(ft, None, engine).candidate_profiles([r], [nharm], m)[0]


def measure_ducy(ft: FFTFile, cands: Sequence[Candidate], params: SearchParams, m: int = 64,
                 engine: Optional[Engine] = None) -> List[Candidate]:
    """measure_ducy(ft, cands, params; m=64) (src/candidate.jl:82-100): the
    candidates with `ducy` = best boxcar width / profile bins of their own
    (r, nharm) profile; NaN where it cannot be measured."""
    if not cands:
        return []
    eng = _engine(ft, params, engine)
    d = eng.measure_ducy([c.r for c in cands], [c.nharm for c in cands], m)
    return [replace(c, ducy=float(x)) for c, x in zip(cands, d)]


def rotate_to_peak(prof) -> np.ndarray:
    """rotate_to_peak(prof) (src/candidate.jl:108-113): circular shift that
    puts the (first) maximum at bin n // 2 (0-based)."""
    p = np.asarray(prof, dtype=np.float64)
    n = len(p)
    if n == 0:
        return p.copy()
    return np.roll(p, n // 2 - int(np.argmax(p)))


def rank_candidates(cands: Sequence[Candidate], ncands: int = 100) -> List[Candidate]:
    """search_one's ranking (src/cli.jl:296-300): best metric first, equal
    metrics in their incoming order (a stable sort), at most `ncands`."""
    return sorted(cands, key=lambda c: c.metric, reverse=True)[:max(0, ncands)]


def report_candidates(ft: FFTFile, cands: Sequence[Candidate], params: SearchParams,
                      ncands: int = 100, engine: Optional[Engine] = None) -> List[Candidate]:
    """search_one's reporting (src/cli.jl:296-305): rank_candidates, then
    measure_ducy on the reported ones."""
    out = rank_candidates(cands, ncands)
    return measure_ducy(ft, out, params, engine=engine) if out else out


def output_path(fftfile: str, outputfilenm: str = "", outdir: str = "", nfiles: int = 1) -> str:
    """output_path (src/cli.jl:260-267): "" (stdout) for a lone file with
    neither -o nor --outdir, else <fftfile without .fft>.cohout beside the input
    or in outdir."""
    if nfiles == 1 and not outdir:
        return outputfilenm
    stem = os.path.splitext(fftfile)[0]
    if outdir:
        stem = os.path.join(outdir, os.path.basename(stem))
    return stem + ".cohout"


def candidate_lines(cands: Sequence[Candidate]) -> List[str]:
    """The `.cohout` table (src/cli.jl:324-330): header, then one fixed-width
    row per candidate."""
    lines = ["#       'S/N'      Frequency (Hz)        Period (ms)    #Harm  Ducy(%)"]
    for i, c in enumerate(cands, start=1):
        ducy = "-" if math.isnan(c.ducy) else "%.2f" % (100 * c.ducy)
        lines.append("%-4d  %8.2f  %18.12f  %18.12f   %3d   %6s" %
                     (i, c.metric, c.freq, 1000.0 / c.freq, c.nharm, ducy))
    return lines


def write_candidates(cands: Sequence[Candidate], outfile: str, threshold: float) -> None:
    """write_candidates (src/cli.jl:317-340): the table to `outfile`, or to
    stdout (with a count line on stderr) when `outfile` is empty."""
    lines = candidate_lines(cands)
    if not outfile:
        for ln in lines:
            print(ln)
        print(f"# {len(cands)} candidates above threshold {float(threshold)!r}", file=sys.stderr)
    else:
        with open(outfile, "w") as io:
            for ln in lines:
                io.write(ln + "\n")
