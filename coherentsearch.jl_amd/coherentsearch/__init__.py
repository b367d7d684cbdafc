"""MI355X coherent-search engine: the coherent harmonic-summing search path of
CoherentSearch.jl (Eqn-30 interpolation, harmonic stack with decimation
folding, batched c2r, boxcar matched-filter S/N) on gfx950, behind the C ABI of
include/coherent_search.h."""
from .candidate import (candidate_profile, measure_ducy, output_path, report_candidates,
                        rotate_to_peak, write_candidates)
from .dedup import harmonically_related, remove_duplicates, remove_harmonics
from .engine import (Engine, SearchCache, block_metrics, chunk_metrics, chunk_ngoodbins,
                     coherent_profiles, fourier_interp, raw_to_candidates, reference_profiles,
                     search, search_block, search_files, snr_metrics)
from .fileio import FFTFile, SimpleInf, freqs, write_fft
from .metricstats import (BlockMetricStats, MetricHistogram, MetricNorm, MetricStats,
                          build_metricnorm, hist_quantile, metricstats_summary,
                          metricstats_windows)
from .params import (Candidate, SearchParams, boxcar_widths, candidate_r, decimation_set,
                     ladder_boxcar_widths, proftype, trial_grid)

__all__ = [
    "BlockMetricStats", "Candidate", "Engine", "FFTFile", "MetricHistogram", "MetricNorm",
    "MetricStats", "SearchCache", "SearchParams", "SimpleInf", "build_metricnorm",
    "hist_quantile", "metricstats_summary", "metricstats_windows", "chunk_ngoodbins",
    "coherent_profiles", "fourier_interp", "reference_profiles", "snr_metrics",
    "block_metrics", "boxcar_widths", "candidate_profile", "candidate_r", "chunk_metrics",
    "decimation_set", "measure_ducy", "output_path", "report_candidates", "rotate_to_peak",
    "write_candidates",
    "freqs", "harmonically_related", "ladder_boxcar_widths", "proftype", "raw_to_candidates",
    "remove_duplicates", "remove_harmonics", "search", "search_block", "search_files",
    "trial_grid",
    "write_fft",
]
