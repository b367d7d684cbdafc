/*
 * coherent_search.h — C ABI of libcsgpu.so, the MI355X (gfx950) engine for the
 * coherent-search path of CoherentSearch.jl.
 *
 * Plain C types only, so Julia `ccall`, Python `ctypes` and C callers bind it
 * directly.  The caller owns every input/output buffer (pointers are never
 * retained); the library owns the opaque context, its HIP stream and its
 * device buffers.  One context per host thread, as with the reference's
 * SearchCache; calls are synchronous with respect to their outputs.
 *
 * Citations are CoherentSearch.jl file:line of the interface each entry point
 * replaces.
 */
#ifndef COHERENT_SEARCH_H
#define COHERENT_SEARCH_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  CS_EINVAL maps to the reference's ArgumentError, CS_EDIM to
 * DimensionMismatch; the message is available from cs_last_error. */
#define CS_OK 0
#define CS_EINVAL 1
#define CS_EDIM 2
#define CS_ENOSPC 3
#define CS_EDEVICE 4
#define CS_ENOMEM 5

/* Execution modes for cs_search. */
#define CS_MODE_PLAIN 0       /* FP32 VALU interpolation, one launch per stage, HBM intermediates */
#define CS_MODE_FUSED 1       /* split-BF16 MFMA interpolation and c2r, fused pipeline */
#define CS_MODE_FUSED_VALU 2  /* fused pipeline with FP32 VALU interpolation and c2r */

/* Tuning knobs of cs_set_tuning (0 restores the default). */
#define CS_TUNE_BATCH_BLOCKS 1  /* blocks per batch (split-BF16 default 16384; plain and
                                   VALU-fused: 2^21 and 2^25 trials' worth) */
#define CS_TUNE_WORKGROUPS 2    /* split-BF16 mode: persistent workgroups (default: one per CU) */
#define CS_TUNE_GRAPHS 3        /* 1: cs_search replays its launch sequence as a HIP graph,
                                   captured again whenever an argument changes (stage
                                   timers then report only the total); default 0 */
#define CS_TUNE_TILE_ORDER 4    /* split-BF16 mode: 0 deals the tile list to the 8 XCDs in
                                   contiguous ranges (neighbouring windows share an L2), 1
                                   round-robin; default 0 */
#define CS_TUNE_GENERIC 5       /* split-BF16 mode: 1 runs the kernel that dispatches each
                                   fold on its profile length at run time even where a kernel
                                   specialised on the search's fold set exists (same results);
                                   default 0 */
#define CS_TUNE_NO_PREFETCH 6   /* split-BF16 mode: 1 loads each tile's first conj(A) operands
                                   at its interpolation instead of into LDS during the previous
                                   tile (same results; the kernel does so by itself when its
                                   LDS plan has no room); default 0 */

/* SearchParams (src/search.jl:131-144).  `decimations` is pointer + length
 * (the reference's Vector{Int}).  precision is the reference's profile-stage
 * type (src/search.jl:143,151-155): 0 = :f64 (the default), 1 = :f32.  Every
 * mode runs its gate on FP32 (plain) or split-BF16 (fused) profiles and its
 * exact rescan in FP64.  With :f64 the block sigma that scales each reported
 * metric is recomputed from FP64 profiles of the reference's sampled columns
 * (stack with FP32 weights, FP64 real DFT) for every (block, fold) holding a
 * gate survivor, so all modes report the same metrics, those of a Float64
 * profile stage; with :f32 it is the pass's own sigma (FP32 or split-BF16
 * samples, ~1e-7 or ~2e-5 relative from the FP64 one). */
typedef struct cs_params {
    int32_t nharms;
    int32_t m;
    int32_t numbetween;
    int32_t ndecim;
    double hidr;
    double threshold;
    double boxcar_fsp;
    double boxcar_maxfrac;
    double boxcar_medmargin;
    const int32_t *decimations;
    int32_t precision;
    int32_t reserved;
} cs_params;

/* Raw candidate before dedup (Candidate, src/search.jl:179-191, in integer
 * form): global trial index t, decimation factor k, harmonics summed
 * (nharms or floor(nharms/k)), 0-based block (chunk) and the metric.  The
 * caller rebuilds r = k * (r_lo + i0*lodr + (t-i0)*lodr), i0 = block*blocksize,
 * with the reference's FP64 expression order (src/search.jl:1812,1833,1696). */
typedef struct cs_cand {
    int64_t t;
    int32_t k;
    int32_t nharm;
    int64_t block;
    double metric;
} cs_cand;

typedef struct cs_ctx cs_ctx;

/* SearchCache() + _plans! (src/search.jl:1914-1961): validates params
 * (odd m, bad precision, empty decimation set -> CS_EINVAL), selects `device`
 * and creates the context's stream. blocksize = Nprof (src/search.jl:2024). */
int cs_ctx_create(const cs_params *params, int32_t blocksize, int32_t device,
                  cs_ctx **out);
int cs_ctx_destroy(cs_ctx *ctx);
const char *cs_last_error(const cs_ctx *ctx);

/* FFTFile amplitudes (src/fileio.jl:64-78): namps interleaved complex64
 * values copied host->device; N = time-series length, T = N*dt. */
int cs_set_spectrum(cs_ctx *ctx, const float *amps_re_im, int64_t namps,
                    int64_t N, double T);
/* Same, from a device pointer already resident in HBM (copied device->device
 * into the context). */
int cs_set_spectrum_device(cs_ctx *ctx, const void *d_amps, int64_t namps,
                           int64_t N, double T);

/* search() minus dedup (src/search.jl:2002-2111 -> _search_region! :1768-1894):
 * every trial fundamental r_lo + t*lodr, t < floor((r_hi-r_lo)/lodr)+1, in
 * blocks of `blocksize`, all decimation folds.  Writes up to `cap` raw
 * candidates sorted by (t, k); *n_out is the total found (CS_ENOSPC if > cap). */
int cs_search(cs_ctx *ctx, double r_lo, double r_hi, double threshold,
              cs_cand *out, int64_t cap, int64_t *n_out);

/* Block triage (no counterpart in the reference; for callers that trade
 * recall for time): 32 incoherent-power features per search block and an MLP
 * scorer 32 -> 64 (ReLU) -> 1 logit whose forward and backward passes run on
 * bf16 operands with FP32 accumulation.  params: W1[64][32] (row = hidden
 * unit), b1[64], w2[64], b2 (FP32); norm: feature mean[32], inverse std[32]. */
#define CS_TRIAGE_FEATURES 32
#define CS_TRIAGE_PARAMS 2177

/* Features of every search block of the band r_lo .. r_hi of the current
 * spectrum, [block][CS_TRIAGE_FEATURES] (*nblocks = block count; CS_ENOSPC if
 * cap_blocks is smaller). */
int cs_triage_features(cs_ctx *ctx, double r_lo, double r_hi, float *out, int64_t cap_blocks,
                       int64_t *nblocks);
/* The same features computed two ways: kind 0 (cs_triage_features) takes, for
 * each integer fundamental f of the block and harmonic g, the power-pyramid
 * maximum over the bins g f .. g f + g - 1 the harmonic can fall in; kind 1
 * takes the incoherent summed power at the harmonics of every trial of the
 * block, P(nint(h r_t)) with P = |X|^2, nint(x) = floor(x + 1/2) in FP64.
 * Both fill the same 32 slots (max, mean and std over the block of the sums at
 * 1, 2, 4, ..., 32 and nharms harmonics, the fold sums, local power levels and
 * medians). */
int cs_triage_features_kind(cs_ctx *ctx, int32_t kind, double r_lo, double r_hi, float *out,
                            int64_t cap_blocks, int64_t *nblocks);
/* logits of n feature rows */
int cs_mlp_forward(cs_ctx *ctx, const float *params, const float *norm, const float *X,
                   int64_t n, float *logits);
/* gradient of the mean binary cross-entropy of n rows with labels y (0/1):
 * grad[0 .. CS_TRIAGE_PARAMS-1] in the params layout, grad[CS_TRIAGE_PARAMS] =
 * the loss.  Reductions run in a fixed order (deterministic). */
int cs_mlp_grad(cs_ctx *ctx, const float *params, const float *norm, const float *X,
                const float *y, int64_t n, float *grad);
/* Training on the device: `epochs` full-batch steps, each the forward and
 * backward pass of cs_mlp_grad over the n rows followed by an FP32 Adam step
 * (beta1 0.9, beta2 0.999, eps 1e-8, step size lr) on the device; params is
 * read as the start and receives the result, losses[epochs] (optional) each
 * step's mean binary cross-entropy before its update. */
int cs_mlp_train(cs_ctx *ctx, float *params, const float *norm, const float *X, const float *y,
                 int64_t n, int32_t epochs, float lr, float *losses);
/* cs_search over the blocks with mask[block] != 0 only (nmask = the band's
 * block count, CS_EDIM otherwise); each selected block's candidates equal
 * those cs_search reports for it. */
int cs_search_blocks(cs_ctx *ctx, double r_lo, double r_hi, double threshold,
                     const uint8_t *mask, int64_t nmask, cs_cand *out, int64_t cap,
                     int64_t *n_out);

/* One spectrum of a multi-file search: host complex64 amplitudes (namps
 * (re, im) pairs, as cs_set_spectrum), the .inf N and T, and the band. */
typedef struct cs_file {
    const float *amps;
    int64_t namps;
    int64_t N;
    double T;
    double r_lo;
    double r_hi;
} cs_file;

/* search over several files with one context, as the CLI's search_one loop
 * shares one SearchCache (src/cli.jl:231-289, search.jl:2002-2111 per file).
 * Each file's raw candidates equal those of cs_set_spectrum + cs_search on it;
 * they are written in file order, counts[i] per file, *n_out in total
 * (CS_ENOSPC if > cap).  The next file's upload (pinned staging, second
 * stream) runs while the current file is searched. */
int cs_search_files(cs_ctx *ctx, int32_t nfiles, const cs_file *files, double threshold,
                    cs_cand *out, int64_t cap, int64_t *counts, int64_t *n_out);

/* Metric statistics (--metricstats: MetricStats, BlockMetricStats and
 * MetricHistogram, src/search.jl:1295-1549, filled by _search_region! and
 * decim_pass!, src/search.jl:1664-1719, 1768-1894).  cs_search_stats runs
 * cs_search with every valid trial's metric computed exactly (the gate
 * disabled, as the reference does when collecting) and returns the same
 * candidates; it keeps every metric on the device and forms, per (block, fold),
 * the metrics' count, extremes, mean, standard deviation and median, and per
 * (fold, searched-frequency window) a histogram of hist_nb bins over
 * [hist_lo, hist_hi).  Windows are nwin log-spaced bands of k * [r_lo, r_hi] / T;
 * a block's metrics go to the window of k * (its mean trial r) / T. */
typedef struct cs_stats_cfg {
    double hist_lo, hist_hi;   /* histogram range (MetricStats defaults 0, 64) */
    int32_t hist_nb;           /* bins, 1 .. 12288 (default 3200) */
    int32_t nwin;              /* windows per fold (default 16) */
} cs_stats_cfg;

typedef struct cs_block_stat {   /* BlockMetricStats (block 0-based) */
    int64_t block;
    int32_t k, Hk, nbins, pad;
    int64_t n;
    double ngoodbins, flo, fhi, min, median, mean, std, max;
} cs_block_stat;

typedef struct cs_hist_acc {     /* MetricHistogram accumulators (win 1-based) */
    int32_t k, win;
    int64_t under, over, total;
    double flo, fhi, sum, sumsq, vmin, vmax;
} cs_hist_acc;

int cs_search_stats(cs_ctx *ctx, double r_lo, double r_hi, double threshold,
                    const cs_stats_cfg *cfg, cs_cand *out, int64_t cap, int64_t *n_out);
/* The last statistics search's block records, ordered by (block, k); blocks
 * of a fold k > 1 with no trial below Nyquist have none.  *n = count
 * (CS_ENOSPC if out is given and cap is smaller). */
int cs_stats_blocks(cs_ctx *ctx, cs_block_stat *out, int64_t cap, int64_t *n);
/* Its histograms, folds in ascending k: counts[fold][win][hist_nb] and
 * acc[fold][win] (either may be NULL); *nfolds = number of folds. */
int cs_stats_hists(cs_ctx *ctx, int64_t *counts, cs_hist_acc *acc, int32_t *nfolds);
/* Its every-trial metrics of fold k (n = the search's trial count; NaN where
 * the fold has no valid trial). */
int cs_stats_metrics(cs_ctx *ctx, int32_t k, double *out, int64_t n);
/* The normalised detection pass over the last statistics search's metrics
 * (the --normalize second pass, src/search.jl:1629-1634, 1848-1852, 1707-1709):
 * z = (M - loc) / scale with loc/scale[fold in ascending k][window] for the
 * window of the trial's own searched frequency; candidates with z > threshold,
 * their metric z. */
int cs_search_normalized(cs_ctx *ctx, double threshold, const double *loc,
                         const double *scale, cs_cand *out, int64_t cap, int64_t *n_out);

/* block_metrics(ft, rfund, params; kernel, widths) (src/search.jl:1061-1079):
 * kernel 1 = :direct (Eqn. 30 point by point, FP64); kernel 0 = :fft
 * (finterp_fft onto the numbetween grid, then clamped linear interpolation,
 * search.jl:974-986, 1036-1040; the grid values are FP64 Eqn. 30, which the FFT
 * correlation reproduces up to its rounding).  widths == NULL means
 * ladder_boxcar_widths(2*nharms, 1, params). */
int cs_block_metrics(cs_ctx *ctx, const double *rfund, int64_t L,
                     int32_t kernel, const int32_t *widths, int32_t nwidths,
                     double *out);

/* reference_profiles(ft, rfund, params; kernel) (src/search.jl:1011-1051): the
 * (2 nharms) x L FP64 profiles, column j of the matrix at out[j * 2 nharms]
 * (the irfft of the stacked harmonics, DC zero; a harmonic whose band leaves
 * the amplitudes or reaches N/2 stays zero). */
int cs_reference_profiles(cs_ctx *ctx, const double *rfund, int64_t L, int32_t kernel,
                          double *out);
/* snr_metrics(profs; boxcar_fsp, boxcar_maxfrac, sigma_samples, sigma_center,
 * widths) (src/search.jl:933-952) of an nbins x L FP64 profile matrix (column
 * j at profs[j * nbins], nbins <= 128): one sigma for the set (every sample,
 * or sigma_samples > 0 of them as whole columns; sigma_center 0 = :zero,
 * 1 = :median), then each profile's exact median-baseline boxcar metric over
 * widths (NULL: boxcar_widths(nbins; fsp, maxfrac)). */
int cs_snr_metrics(cs_ctx *ctx, const double *profs, int32_t nbins, int64_t L, double fsp,
                   double maxfrac, int64_t sigma_samples, int32_t sigma_center,
                   const int32_t *widths, int32_t nwidths, double *out);
/* cs_snr_metrics on device buffers, allocating nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * d_profs: the nbins x L FP64 profile matrix in device memory, column j at
 * d_profs[j * nbins] (the layout cs_snr_metrics takes on the host); d_out: L
 * doubles in device memory; widths: a host array (or NULL), as there.  The
 * sigma, its samples and the width bank live in a fixed scratch array of
 * 8 MiB + 4 KiB per GPU, held until the process ends: at most 2^20 sigma
 * samples (min(nbins * L, sigma_samples), CS_EINVAL beyond).  Every call on
 * one GPU uses that GPU's one scratch array, so calls that overlap on a GPU
 * (from several contexts or threads) race.  Runs on the context's stream and
 * returns after it has drained; no hipMalloc, hipFree or host-to-device copy. */
int cs_snr_metrics_noalloc(cs_ctx *ctx, const double *d_profs, int32_t nbins, int64_t L,
                           double fsp, double maxfrac, int64_t sigma_samples,
                           int32_t sigma_center, const int32_t *widths, int32_t nwidths,
                           double *d_out);
/* fourier_interp(r, ft, m) (src/fourierinterp.jl:66-74) at n frequencies of
 * the current spectrum: out[2i], out[2i+1] = re, im.  CS_EINVAL for r < 0 or
 * odd m, CS_EDIM when a window leaves the amplitudes. */
int cs_fourier_interp(cs_ctx *ctx, const double *r, int64_t n, int32_t m, double *out);

/* candidate_profile(ft, r, nharm; m) (src/candidate.jl:37-61) for n
 * candidates: harmonics h = 1, 2, ... of r[i] by FP64 fourier_interp with an
 * m-bin kernel, stopping at the first whose window leaves the amplitudes or
 * whose r h reaches N/2; the H used give a 2H-bin irfft profile (2 zero bins
 * when H = 0), not rotated.  prof (optional) receives row i at prof[i * ld],
 * len[i] its length (ld >= 2 * nharm[i]).  CS_EINVAL for nharm < 1, odd m,
 * or more than 512 harmonics. */
int cs_candidate_profiles(cs_ctx *ctx, const double *r, const int32_t *nharm, int64_t n,
                          int32_t m, double *prof, int64_t ld, int32_t *len);
/* measure_ducy (src/candidate.jl:82-100): ducy[i] = best boxcar width /
 * profile bins of candidate_profile(r[i], nharm[i]; m), over
 * boxcar_widths(bins; params' fsp and maxfrac) (boxcar_best_width,
 * src/search.jl:617-638); NaN unless nharm[i] >= 1 and r[i] > 0. */
int cs_measure_ducy(cs_ctx *ctx, const double *r, const int32_t *nharm, int64_t n, int32_t m,
                    double *ducy);

/* chunk_metrics(ft, params, rstart, n; lodr, weights) (src/search.jl:1740-1755):
 * the production machinery over one block at t0 = 0 with the plans anchored
 * at rstart, exact scan, k = 1.  `lodr` is accepted for signature parity and,
 * as in the reference, does not enter the interpolation. */
int cs_chunk_metrics(cs_ctx *ctx, double rstart, int64_t n, double lodr,
                     int32_t weights_f64, double *out);

/* decim_pass! metrics for fold k over one block at t0 = 0
 * (src/search.jl:1664-1691; the bridge test/test_search.jl:465-477 calls):
 * *nvalid valid trials, exact scan. */
int cs_decim_metrics(cs_ctx *ctx, double rstart, int64_t n, int32_t k,
                     int32_t weights_f64, double *out, int64_t *nvalid);

/* fill_chunk_profiles! rows (src/search.jl:1280-1287 ->
 * fill_harmonic_row_direct!, src/directinterp.jl:470-572): the (nharms+1) x n
 * stacked amplitudes for trials t0..t0+n-1 of a grid anchored at r_lo, as
 * interleaved complex128, row-major (row h, column j).  Row 0 (DC) is zero. */
int cs_interp_rows(cs_ctx *ctx, double r_lo, int64_t t0, int64_t n,
                   int32_t weights_f64, double *out);

/* Execution mode for cs_search (CS_MODE_PLAIN / CS_MODE_FUSED / CS_MODE_FUSED_VALU). */
int cs_set_mode(cs_ctx *ctx, int32_t mode);

/* Launch-shape knob of cs_search (CS_TUNE_*), value >= 0 (0: default).  The
 * reference has no such knob; its counterpart is the chunk-invariance property
 * the search keeps (test/test_search.jl:234-276): every knob changes only how
 * the trials are scheduled, never a result bit. */
int cs_set_tuning(cs_ctx *ctx, int32_t knob, int64_t value);
/* The current value of a CS_TUNE_* knob. */
int cs_get_tuning(const cs_ctx *ctx, int32_t knob, int64_t *value);

/* Autotuner of the search's launch shape for one search geometry (so per
 * spectrum length): times cs_search (device time, median of reps after one
 * warm search) over the knobs of the context's mode and keeps the fastest,
 * then reports blocks per batch and workgroups (either output may be NULL;
 * cs_get_tuning reads every knob).  Split-BF16 mode: persistent grid sizes,
 * batch sizes (1, 2 or 4 batches) and the tile order; plain and VALU-fused
 * modes: blocks per batch.  With cache_path, a line per geometry (parameters,
 * band, spectrum length, mode, device) is looked up first and written after
 * measuring, by temp file + rename as the reference's FFTW wisdom export does
 * (src/wisdom.jl:46-61).  The tile shape (16 trials x 8 occurrences, the
 * v_mfma_f32_16x16x32_bf16 output) and the block size Nprof (the reference's
 * blocksize, which the results depend on) are not knobs. */
int cs_autotune(cs_ctx *ctx, double r_lo, double r_hi, double threshold,
                const char *cache_path, int32_t reps, int64_t *best_batch,
                int32_t *best_workgroups);

/* The tuning cache cs_search_files reads (NULL or "" for none): each length
 * class of a sweep (files of one spectrum length and band) runs with the
 * launch shape cs_autotune stored for its geometry in that file, looked up
 * without timing; a class with no line runs with the context's knobs, which
 * the sweep leaves as it found them.  Like cs_autotune, it has no counterpart
 * in the reference beyond the FFTW wisdom file it mirrors (src/wisdom.jl:18-61,
 * the plan cache one SearchCache carries across the CLI's files,
 * src/cli.jl:231-289). */
int cs_set_tuning_cache(cs_ctx *ctx, const char *path);

/* Per-stage device time of the last cs_search, milliseconds, HIP events on the
 * context's stream: [0] interp, [1] c2r, [2] sigma, [3] gate, [4] rescan (with
 * the FP64 sigma of the survivors' blocks under precision :f64), [5] fused,
 * [6] total device, [7] number of batches. */
int cs_last_timings(const cs_ctx *ctx, double *ms8);

/* The sigma estimate of every (block, fold) of the last cs_search (or
 * statistics search) over a whole range: out[block * nfolds + fold], fold 0
 * the undecimated one, then the decimations > 1 in parameter order, 0 where a
 * fold has no valid trial in the block.  These are the sigmas each block's
 * _block_sigma returns inside _search_region! (src/search.jl:490-566,
 * :1817; decim_pass! :1688); the reference keeps them internal.
 * *n_out = nblocks * nfolds; out may be NULL to query it.  CS_EINVAL if no
 * search has completed or the last one covered selected blocks only. */
int cs_last_sigmas(cs_ctx *ctx, double *out, int64_t cap, int64_t *n_out);

/* Diagnostic of the split-BF16 mode, for its per-result error statement (the
 * reference states the error of its own Float32-weight interpolation,
 * src/directinterp.jl:310-319): runs the CS_MODE_FUSED pass of cs_search over
 * r_lo .. r_hi (in one batch: at most 16384 blocks) with no candidates, and
 * returns for the band's trials t0 .. t0 + n - 1
 *  - amps[2 ((h - 1) n + j)] (re, then im at + 1), h = 1 .. nharms: the stacked
 *    amplitude of harmonic h (fill_harmonic_row_direct!, directinterp.jl:470-619)
 *    as the split-BF16 DFT reads it, the hi + lo bf16 pair of the interpolation's
 *    FP32-accumulated split-precision product (zero where the block's row
 *    guard drops the row);
 *  - bounds[f n + j]: fold f's zero-baseline boxcar bound before sigma (the
 *    quantity _boxcar_gate! tests, search.jl:676-807: the maximum over the
 *    fold's widths w and circular window starts of the window sum of the
 *    unnormalised profile over sqrt(w)), f = 0 the undecimated fold, then the
 *    decimations > 1 in parameter order; -inf where the trial is past the
 *    fold's valid prefix.
 * Either output may be NULL.  CS_EINVAL unless the context is in
 * CS_MODE_FUSED and the trials lie in the band. */
int cs_probe_split_bf16(cs_ctx *ctx, double r_lo, double r_hi, int64_t t0, int64_t n,
                        float *amps, float *bounds);

/* Host-only (no device): the per-harmonic interpolation tables the library
 * builds for a grid anchored at r_lo (build_direct_plans,
 * src/directinterp.jl:345-405).  info[0..5] = q, s, base_adv, pnum, rfloor0, P;
 * row (q entries, 0-based phase or -1), carry (P), W (P*m, FP32 weights),
 * A (2P, FP32).  Any output pointer may be NULL. */
int cs_plan_tables(const cs_params *params, double r_lo, int32_t h,
                   int64_t *info, int32_t *row, uint8_t *carry, float *W,
                   float *A);

/* Library version string. */
const char *cs_version(void);

#ifdef __cplusplus
}
#endif
#endif
