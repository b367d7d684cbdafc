/*
 * cs_oracle.h — CPU restatement of the coherent-search path of
 * scottransom/CoherentSearch.jl, used ONLY as test infrastructure: the parity
 * checker for the HIP library and the CPU baseline timed by bench.py.  Nothing
 * in the product path links or calls it.
 *
 * Every function names the reference file:line it follows.  Complex arrays are
 * interleaved (re, im) doubles unless stated otherwise.  The .fft amplitudes
 * are interleaved float (complex64), as PRESTO stores them.
 *
 * Layouts chosen for this restatement (the reference is column-major Julia):
 *   ftprofs : [row h = 0..nharms][col j = 0..Nprof-1] complex128, split into
 *             separate re/im planes (row-major, columns contiguous).
 *   profs   : [bin i = 0..nbins-1][col j] double.
 * Only the storage order differs; every arithmetic step keeps the reference's
 * order of operations.
 */
#ifndef CS_ORACLE_H
#define CS_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes (mirror ArgumentError / DimensionMismatch in the reference). */
#define OR_OK 0
#define OR_EINVAL 1
#define OR_EDIM 2
#define OR_ENOSPC 3

/* SearchParams, src/search.jl:131-144 */
typedef struct or_params {
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
} or_params;

/* Raw candidate (before dedup): global trial index, decimation factor, the
 * number of harmonics summed, the 0-based block (chunk) and the metric. */
typedef struct or_cand {
    int64_t t;
    int32_t k;
    int32_t nharm;
    int64_t block;
    double metric;
} or_cand;

/* ---- src/fourierinterp.jl ------------------------------------------------ */
int or_finterp_coeffs(double dr, int32_t m, double *out /* 2m */);          /* :23 */
int or_nearby_bin_range(double r, int32_t m, int64_t *lo0, int64_t *hi0);   /* :44 */
int or_fourier_interp(double r, const double *ft, int64_t nft, int32_t m,
                      double *out2);                                         /* :66 */
int64_t or_next_pow_of_2(int64_t n);                                         /* :104 */
int32_t or_is_smooth(int64_t n);                                             /* :115 */
int64_t or_next_smooth(int64_t n);                                           /* :141 */
int or_finterp_fft(int64_t lobin, int64_t numbins, int32_t numbetween,
                   const double *ft, int64_t nft, int32_t m, int64_t fftlen,
                   double *out /* 2*numbins*numbetween */);                 /* :161-212 */
/* In-place complex DFT of any length (mixed radix); sign -1 forward, +1 inverse
 * (unnormalised).  Used by or_finterp_fft. */
int or_fft(double *data, int64_t n, int32_t sign);

/* ---- src/directinterp.jl ------------------------------------------------- */
int or_finterp_direct(double r0, int64_t n, double step, const double *ft,
                      int64_t nft, int32_t m, double *out);                  /* :71-101 */
int or_trial_grid_rational(double lodr, int64_t *pnum, int64_t *q);          /* :217-225 */

typedef struct or_plans or_plans;
/* build_direct_plans(WT, params, r_lo), directinterp.jl:345-405 (with
 * build_group_table :237-282).  weights_f64 selects WT = Float64, else Float32. */
or_plans *or_plans_build(int32_t nharms, int32_t m, double hidr, double r_lo,
                         int32_t weights_f64);
void or_plans_free(or_plans *pl);
/* info[0..7] = q, s, base_adv, pnum, rfloor0, P, ngrp, gstep; *dr0 */
int or_plan_info(const or_plans *pl, int32_t h, int64_t *info, double *dr0);
/* row (q int32), carry (P uint8), W (m*P, column p at W[p*m]), A (2P) —
 * values as stored in WT, widened to double. */
int or_plan_tables(const or_plans *pl, int32_t h, int32_t *row, uint8_t *carry,
                   double *W, double *A);
/* direct_chunk_state, directinterp.jl:414-418 (exact, 128-bit). */
int or_chunk_state(const or_plans *pl, int32_t h, int64_t t0, int64_t *res,
                   int64_t *qint);

/* fill_chunk_profiles! minus the transform (search.jl:1280-1287): zero the
 * stack and fill rows 1..nharms for trials t0..t0+n-1 via
 * fill_harmonic_row_direct! (directinterp.jl:470-572), trials-axis group form.
 * re/im are [(nharms+1) * ncol] planes, ncol >= n. */
int or_fill_chunk(const or_plans *pl, const float *amps, int64_t namps,
                  int64_t N, int64_t t0, int64_t n, int64_t ncol, double *re,
                  double *im);
/* Same rows, evaluated trial by trial as an m-term FMA chain (no group
 * tables).  Bit-identical to or_fill_chunk; tests prove it. */
int or_fill_chunk_pertrial(const or_plans *pl, const float *amps, int64_t namps,
                           int64_t N, int64_t t0, int64_t n, int64_t ncol,
                           double *re, double *im);

/* ---- batched c2r (brfft) with numpy's DC/Nyquist convention --------------
 * search.jl:1291 (base) and :1675 (decimated, rows 0,k,..,Hk*k of the stack).
 * out[i*ncol + j] = sum over h of the Hermitian-extended spectrum, unnormalised
 * (= nbins * irfft).  Rows read: row index h*k for h = 0..H, H = nbins/2. */
int or_brfft_rows(const double *re, const double *im, int64_t ncol_stride,
                  int32_t k, int32_t nbins, int64_t ncol, double *out);

/* ---- src/search.jl: metric ------------------------------------------------ */
double or_median(double *v, int64_t n);                                       /* :296-305 */
int32_t or_boxcar_widths(int32_t nbins, double fsp, double maxfrac,
                         int32_t *out /* >= 64 */);                           /* :387-397 */
int32_t or_ladder_boxcar_widths(int32_t nbins, int32_t k, const or_params *p,
                                int32_t *out);                                /* :443-459 */
/* _block_sigma (:490-566).  M element (bin i, col j) at M[i*bin_stride + j*col_stride].
 * cap = length of the reference's sample buffer.  center: 0 = :zero, 1 = :median. */
double or_block_sigma(const double *M, int64_t bin_stride, int64_t col_stride,
                      int32_t nbins, int64_t n, int64_t cap, int32_t center);
/* _boxcar_psum! + _boxcar_scan against baseline b (:571-598), one column. */
double or_boxcar_scan_col(const double *col, int64_t stride, int32_t nbins,
                          const int32_t *widths, int32_t nw, double invsigma,
                          double b);
/* _boxcar_exact (:893-904): median baseline then scan. */
double or_boxcar_exact_col(const double *col, int64_t stride, int32_t nbins,
                           const int32_t *widths, int32_t nw, double invsigma);
/* boxcar_best_width (:618-639) */
int32_t or_boxcar_best_width(const double *prof, int32_t nbins, double fsp,
                             double maxfrac);
/* boxcar_metrics! (:825-845) with the batched Float32 gate (:729-807).
 * profs as [bin][col] with column stride 1 and bin stride ncol_stride. */
int or_boxcar_metrics(const double *profs, int64_t ncol_stride, int64_t n,
                      const int32_t *widths, int32_t nw, int32_t nbins,
                      double invsigma, double medcut, double *mvals);
/* snr_metrics (:933-951); profs column-major (nbins, L) like the reference. */
int or_snr_metrics(const double *profs, int32_t nbins, int64_t L,
                   int64_t sigma_samples, int32_t center, const int32_t *widths,
                   int32_t nw, double *out);

/* ---- drivers ---------------------------------------------------------------- */
/* chunk_metrics (search.jl:1740-1755). */
int or_chunk_metrics(const or_params *p, const float *amps, int64_t namps,
                     int64_t N, double rstart, int64_t n, int32_t weights_f64,
                     double *out);
/* reference_profiles + block_metrics (search.jl:1011-1079); kernel 0 = :fft,
 * 1 = :direct.  widths NULL => ladder_boxcar_widths(nbins, 1, params). */
int or_block_metrics(const or_params *p, const float *amps, int64_t namps,
                     int64_t N, const double *rfund, int64_t L, int32_t kernel,
                     const int32_t *widths, int32_t nw, double *out);
/* reference_profiles only: out is (2*nharms, L) column-major, normalised irfft. */
int or_reference_profiles(const or_params *p, const float *amps, int64_t namps,
                          int64_t N, const double *rfund, int64_t L,
                          int32_t kernel, double *out);
/* decim_pass! metrics for one fold k of one chunk (search.jl:1664-1719) after a
 * fill: returns nvalid and the metric of every valid trial (exact: medcut=-Inf). */
int or_decim_metrics(const or_params *p, const float *amps, int64_t namps,
                     int64_t N, double rstart, int64_t n, int32_t k,
                     int32_t weights_f64, double *out, int64_t *nvalid);
/* _search_region! (search.jl:1768-1894) over chunks [c_begin, c_end) of the
 * band: raw candidates (unsorted: per-thread lists concatenated). nthreads <= 0
 * means all OpenMP threads.  Returns OR_ENOSPC if cap is too small (*n_out then
 * holds the count needed).  medcut: pass NaN for the production rule
 * threshold - boxcar_medmargin, or -INFINITY for the exact path everywhere.
 * mall (NULL or [1 + decimations > 1][total]) receives every scored metric. */
int or_search_region(const or_params *p, const float *amps, int64_t namps,
                     int64_t N, double r_lo, int64_t total, int64_t Nprof,
                     int64_t c_begin, int64_t c_end, double threshold,
                     double medcut, int32_t nthreads, or_cand *out, int64_t cap,
                     int64_t *n_out, double *mall, double *sigmas);

#ifdef __cplusplus
}
#endif
#endif
