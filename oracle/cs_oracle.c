/*
 * cs_oracle.c — CPU restatement of CoherentSearch.jl's coherent-search path.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker for the HIP library and the CPU
 * baseline leg of bench.py.  Compiled with -ffp-contract=off so that every
 * fused multiply-add is an explicit fma() call, exactly where the reference
 * writes `muladd`, and nowhere else.
 *
 * Each function cites the reference file:line it restates (paths relative to
 * the CoherentSearch.jl repository root).
 */
#include "cs_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_PI 3.14159265358979323846
#define DIRECT_QMAX (1 << 16)       /* directinterp.jl:111 */
#define DIRECT_GROUP_V 32           /* directinterp.jl:151 */
#define BOXCAR_SIGMA_SAMPLES 8192   /* search.jl:466 */
#define BC_BATCH 128                /* search.jl:681 */
#define LADDER_WMAX 6               /* search.jl:401 */

/* ------------------------------------------------------------------------ */
/* Elementary functions with exact argument reduction (Julia Base sinpi,     */
/* cospi, sinc semantics; agreement is to ~1 ulp, which is below every        */
/* tolerance this restatement is held to).                                    */
/* ------------------------------------------------------------------------ */
static double or_sinpi(double x) {
    double r = fmod(x, 2.0);
    if (r < 0) r += 2.0;
    int neg = 0;
    if (r >= 1.0) { r -= 1.0; neg = 1; }
    double v;
    if (r <= 0.25) v = sin(OR_PI * r);
    else if (r <= 0.75) v = cos(OR_PI * (r - 0.5));
    else v = sin(OR_PI * (1.0 - r));
    return neg ? -v : v;
}

static double or_cospi(double x) {
    double r = fabs(fmod(x, 2.0));
    if (r > 1.0) r = 2.0 - r;
    if (r <= 0.25) return cos(OR_PI * r);
    if (r <= 0.75) return -sin(OR_PI * (r - 0.5));
    return -cos(OR_PI * (1.0 - r));
}

/* Julia's normalised sinc: Taylor branch near 0, sinpi(x)/(pi*x) elsewhere. */
static double or_sinc(double x) {
    if (x == 0.0) return 1.0;
    if (fabs(x) < 4.5 * sqrt(DBL_EPSILON)) {
        double x2 = x * x, p2 = OR_PI * OR_PI;
        return 1.0 + x2 * (-p2 / 6.0 + x2 * (p2 * p2 / 120.0));
    }
    return or_sinpi(x) / (OR_PI * x);
}

static int64_t fld64(int64_t a, int64_t b) {   /* floor division, b > 0 */
    int64_t q = a / b;
    if ((a % b) != 0 && ((a < 0) != (b < 0))) q -= 1;
    return q;
}
static int64_t mod64(int64_t a, int64_t b) { return a - fld64(a, b) * b; }
static int64_t gcd64(int64_t a, int64_t b) {
    if (a < 0) a = -a;
    if (b < 0) b = -b;
    while (b) { int64_t t = a % b; a = b; b = t; }
    return a;
}

/* ------------------------------------------------------------------------ */
/* fourierinterp.jl                                                           */
/* ------------------------------------------------------------------------ */

/* finterp_coeffs, fourierinterp.jl:23-31 */
int or_finterp_coeffs(double dr, int32_t m, double *out) {
    if (m % 2 != 0) return OR_EINVAL;
    if (!(0.0 <= dr && dr < 1.0)) return OR_EINVAL;
    for (int32_t i = 0; i < m; i++) {
        double off = dr - (double)(-(m / 2) + 1 + i);
        double s = or_sinc(off);
        out[2 * i] = s * or_cospi(off);
        out[2 * i + 1] = s * or_sinpi(off);
    }
    return OR_OK;
}

/* nearby_fourier_bin_range, fourierinterp.jl:44-50 (returned 0-based, half-open) */
int or_nearby_bin_range(double r, int32_t m, int64_t *lo0, int64_t *hi0) {
    if (m % 2 != 0) return OR_EINVAL;
    int64_t r_int = (int64_t)floor(r + 1e-15) + 1;
    *lo0 = r_int - m / 2;
    *hi0 = r_int + m / 2;
    return OR_OK;
}

/* fourier_interp, fourierinterp.jl:66-74: sum conj(coeff_j) * bin_j */
int or_fourier_interp(double r, const double *ft, int64_t nft, int32_t m,
                      double *out2) {
    if (!(r >= 0.0)) return OR_EINVAL;
    if (m % 2 != 0) return OR_EINVAL;
    double c[2 * 512];
    if (m > 512) return OR_EINVAL;
    double dr = r - floor(r);                   /* mod(r, 1.0) for r >= 0 */
    int st = or_finterp_coeffs(dr, m, c);
    if (st) return st;
    int64_t lo0, hi0;
    or_nearby_bin_range(r, m, &lo0, &hi0);
    if (lo0 < 0 || hi0 > nft) return OR_EDIM;
    double sr = 0.0, si = 0.0;
    for (int32_t i = 0; i < m; i++) {
        double cr = c[2 * i], ci = c[2 * i + 1];
        double br = ft[2 * (lo0 + i)], bi = ft[2 * (lo0 + i) + 1];
        sr += cr * br + ci * bi;                /* conj(c) * b */
        si += cr * bi - ci * br;
    }
    out2[0] = sr;
    out2[1] = si;
    return OR_OK;
}

/* next_pow_of_2 / is_smooth / next_smooth, fourierinterp.jl:104-149 */
int64_t or_next_pow_of_2(int64_t n) {
    if (n <= 0) return -1;
    int64_t p = 1;
    while (p < n) p <<= 1;
    return p;
}
int32_t or_is_smooth(int64_t n) {
    if (n <= 0) return 0;
    const int64_t f[4] = {2, 3, 5, 7};
    for (int i = 0; i < 4; i++)
        while (n % f[i] == 0) n /= f[i];
    return n == 1;
}
int64_t or_next_smooth(int64_t n) {
    if (n <= 0) return -1;
    if (n <= 4) return n;
    int64_t k = n;
    while (!or_is_smooth(k)) k++;
    return k;
}

/* Generic out-of-place mixed-radix decimation-in-time DFT, exp(sign*2*pi*i*jk/n),
 * unnormalised.  Twiddles come from cospi/sinpi of exact rationals. */
static void twiddle(int64_t e, int64_t n, int32_t sign, double *wr, double *wi) {
    e %= n;
    if (e < 0) e += n;
    double a = 2.0 * (double)e / (double)n;
    *wr = or_cospi(a);
    *wi = sign * or_sinpi(a);
}

static void fft_rec(const double *in, int64_t n, int64_t s, double *out,
                    int32_t sign) {
    if (n == 1) { out[0] = in[0]; out[1] = in[1]; return; }
    int64_t r = 0;
    const int64_t pref[5] = {4, 2, 3, 5, 7};
    for (int i = 0; i < 5 && !r; i++) if (n % pref[i] == 0) r = pref[i];
    if (!r) {
        for (int64_t f = 11; f * f <= n; f += 2) if (n % f == 0) { r = f; break; }
        if (!r) r = n;
    }
    int64_t m = n / r;
    for (int64_t q = 0; q < r; q++)
        fft_rec(in + 2 * q * s, m, s * r, out + 2 * q * m, sign);
    double *t = (double *)malloc(sizeof(double) * 4 * r);
    double *y = t + 2 * r;
    for (int64_t k = 0; k < m; k++) {
        for (int64_t q = 0; q < r; q++) {
            double wr, wi;
            twiddle(q * k, n, sign, &wr, &wi);
            double xr = out[2 * (q * m + k)], xi = out[2 * (q * m + k) + 1];
            t[2 * q] = xr * wr - xi * wi;
            t[2 * q + 1] = xr * wi + xi * wr;
        }
        for (int64_t u = 0; u < r; u++) {
            double ar = 0.0, ai = 0.0;
            for (int64_t q = 0; q < r; q++) {
                double wr, wi;
                twiddle(q * u, r, sign, &wr, &wi);
                ar += t[2 * q] * wr - t[2 * q + 1] * wi;
                ai += t[2 * q] * wi + t[2 * q + 1] * wr;
            }
            y[2 * u] = ar;
            y[2 * u + 1] = ai;
        }
        for (int64_t u = 0; u < r; u++) {
            out[2 * (k + m * u)] = y[2 * u];
            out[2 * (k + m * u) + 1] = y[2 * u + 1];
        }
    }
    free(t);
}

int or_fft(double *data, int64_t n, int32_t sign) {
    if (n <= 0) return OR_EINVAL;
    double *tmp = (double *)malloc(sizeof(double) * 2 * n);
    if (!tmp) return OR_ENOSPC;
    fft_rec(data, n, 1, tmp, sign < 0 ? -1 : 1);
    memcpy(data, tmp, sizeof(double) * 2 * n);
    free(tmp);
    return OR_OK;
}

/* finterp_fft_coeffs, fourierinterp.jl:161-174 */
static int finterp_fft_coeffs(int32_t nb, int32_t m, int64_t fftlen, double *c) {
    if (m % 2 != 0) return OR_EINVAL;
    if (fftlen < (int64_t)nb * m) return OR_EINVAL;
    memset(c, 0, sizeof(double) * 2 * fftlen);
    int64_t n = ((int64_t)nb * m) / 2;
    for (int64_t i = 0; i < n; i++) {
        double off = (double)i / nb;
        double s = or_sinc(off);
        c[2 * i] = s * or_cospi(-off);
        c[2 * i + 1] = s * or_sinpi(-off);
    }
    /* offsets2 = reverse(-(offsets .+ 1/nb)) placed in the last n slots */
    for (int64_t i = 0; i < n; i++) {
        double off = -((double)(n - 1 - i) / nb + 1.0 / nb);
        double s = or_sinc(off);
        int64_t d = fftlen - n + i;
        c[2 * d] = s * or_cospi(-off);
        c[2 * d + 1] = s * or_sinpi(-off);
    }
    or_fft(c, fftlen, -1);
    for (int64_t i = 0; i < fftlen; i++) c[2 * i + 1] = -c[2 * i + 1];
    return OR_OK;
}

/* finterp_fft, fourierinterp.jl:190-212 */
int or_finterp_fft(int64_t lobin, int64_t numbins, int32_t numbetween,
                   const double *ft, int64_t nft, int32_t m, int64_t fftlen,
                   double *out) {
    if (m % 2 != 0) return OR_EINVAL;
    int64_t m2 = m / 2;
    int64_t numftbins = (numbins + m) * numbetween;
    if (fftlen <= 0) fftlen = or_next_pow_of_2(numftbins);
    if (fftlen < numftbins) return OR_EINVAL;
    int64_t tmplobin = lobin - m2, tmphibin = lobin + numbins + m2;
    if (tmplobin < 0 || tmphibin > nft) return OR_EDIM;
    double *c = (double *)malloc(sizeof(double) * 2 * fftlen);
    double *a = (double *)calloc(2 * fftlen, sizeof(double));
    int st = finterp_fft_coeffs(numbetween, m, fftlen, c);
    if (st) { free(c); free(a); return st; }
    for (int64_t i = 0; i < numbins + m; i++) {
        a[2 * i * numbetween] = ft[2 * (tmplobin + i)];
        a[2 * i * numbetween + 1] = ft[2 * (tmplobin + i) + 1];
    }
    or_fft(a, fftlen, -1);
    for (int64_t i = 0; i < fftlen; i++) {
        double xr = a[2 * i], xi = a[2 * i + 1], cr = c[2 * i], ci = c[2 * i + 1];
        a[2 * i] = xr * cr - xi * ci;
        a[2 * i + 1] = xr * ci + xi * cr;
    }
    or_fft(a, fftlen, +1);
    double sc = 1.0 / (double)fftlen;
    int64_t o = m2 * numbetween, cnt = numbins * numbetween;
    for (int64_t i = 0; i < cnt; i++) {
        out[2 * i] = a[2 * (o + i)] * sc;
        out[2 * i + 1] = a[2 * (o + i) + 1] * sc;
    }
    free(c);
    free(a);
    return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* directinterp.jl                                                            */
/* ------------------------------------------------------------------------ */

/* conj(sin(pi dr) cispi(dr) / pi) evaluated in Float64 in the reference's order. */
static void kernel_A(double dr, double *ar, double *ai) {
    double s = or_sinpi(dr), c = or_cospi(dr);
    *ar = (s * c) / OR_PI;
    *ai = -((s * s) / OR_PI);
}

/* finterp_direct!, directinterp.jl:76-101 */
int or_finterp_direct(double r0, int64_t n, double step, const double *ft,
                      int64_t nft, int32_t m, double *out) {
    if (m % 2 != 0) return OR_EINVAL;
    int64_t m2 = m / 2;
    for (int64_t k = 0; k < n; k++) {
        double r = r0 + (double)k * step;
        int64_t rint = (int64_t)floor(r + 1e-15) + 1;
        double dr = r - floor(r);
        int64_t base = rint - m2;                /* 1-based: ft[base + i] */
        if (dr == 0.0) {
            int64_t ix = base + m2 - 1;
            if (ix < 0 || ix >= nft) return OR_EDIM;
            out[2 * k] = ft[2 * ix];
            out[2 * k + 1] = ft[2 * ix + 1];
            continue;
        }
        if (base < 0 || base - 1 + m >= nft) return OR_EDIM;
        double sre = 0.0, sim = 0.0;
        for (int64_t i = 1; i <= m; i++) {
            double w = 1.0 / (dr - (double)(i - m2));
            sre = fma(w, ft[2 * (base + i - 1)], sre);
            sim = fma(w, ft[2 * (base + i - 1) + 1], sim);
        }
        double Ar, Ai;
        kernel_A(dr, &Ar, &Ai);
        out[2 * k] = Ar * sre - Ai * sim;
        out[2 * k + 1] = Ar * sim + Ai * sre;
    }
    return OR_OK;
}

/* rationalize(Int, x; tol) by continued-fraction convergents, then the
 * acceptance test of trial_grid_rational, directinterp.jl:217-225. */
int or_trial_grid_rational(double lodr, int64_t *pnum, int64_t *q) {
    if (!(lodr > 0)) return 0;
    const double tol = 1e-13;
    double x = lodr;
    int64_t p0 = 0, q0 = 1, p1 = 1, q1 = 0;
    double y = x;
    int64_t P = 0, Q = 1;
    for (int it = 0; it < 64; it++) {
        double a = floor(y);
        if (a > 4e18) break;
        int64_t ai = (int64_t)a;
        int64_t p2 = ai * p1 + p0, q2 = ai * q1 + q0;
        if (q2 <= 0 || q2 > (int64_t)1 << 40) break;
        p0 = p1; q0 = q1; p1 = p2; q1 = q2;
        P = p1; Q = q1;
        if (fabs((double)P / (double)Q - x) <= tol) break;
        double fr = y - a;
        if (fr == 0.0) break;
        y = 1.0 / fr;
    }
    int64_t g = gcd64(P, Q);
    if (g > 1) { P /= g; Q /= g; }
    if (!(Q <= DIRECT_QMAX && P >= 1)) return 0;
    double epsx = nextafter(lodr, INFINITY) - lodr;
    if (!(fabs((double)P / (double)Q - lodr) <= 8.0 * epsx)) return 0;
    *pnum = P;
    *q = Q;
    return 1;
}

typedef struct or_plan {
    int64_t h, m, q, s, base_adv, pnum, rfloor0, P;
    double dr0;
    int32_t *row;       /* q, values 1..P (0 = unvisited) */
    uint8_t *carry;     /* P */
    double *W;          /* m*P : value of WT, column p at W[(p-1)*m] */
    double *Ar, *Ai;    /* P */
    /* group tables */
    int64_t ngrp;
    int32_t *grow;      /* q */
    int64_t *goff;      /* ngrp+1 */
    int32_t *gnj;       /* ngrp */
    double *gW;         /* flat (V x gnj) column-major per group */
    double *gAr, *gAi;  /* V x ngrp */
} or_plan;

struct or_plans {
    int32_t nharms, m, weights_f64;
    double hidr, r_lo;
    int fallback;       /* 1 when lodr is not a small rational */
    or_plan *p;         /* nharms */
};

static double wt_round(double v, int32_t f64) { return f64 ? v : (double)(float)v; }

/* build_group_table, directinterp.jl:237-282 */
static void build_group_table(or_plan *d, int32_t f64) {
    const int64_t V = DIRECT_GROUP_V;
    int64_t q = d->q, s = d->s, m = d->m;
    int64_t gstep = mod64(V * s, q);
    int64_t ngrp = gstep == 0 ? 1 : q / gcd64(gstep, q);
    d->ngrp = ngrp;
    d->grow = (int32_t *)calloc(q, sizeof(int32_t));
    d->goff = (int64_t *)malloc(sizeof(int64_t) * (ngrp + 1));
    d->gnj = (int32_t *)malloc(sizeof(int32_t) * ngrp);
    d->gAr = (double *)malloc(sizeof(double) * V * ngrp);
    d->gAi = (double *)malloc(sizeof(double) * V * ngrp);
    int64_t *ps = (int64_t *)malloc(sizeof(int64_t) * V);
    int64_t *ds = (int64_t *)malloc(sizeof(int64_t) * V);
    /* first pass: sizes */
    int64_t total = 0, res0 = 0;
    int64_t cap = 0;
    for (int64_t g = 0; g < ngrp; g++) {
        int64_t res = res0, B = 0, B0 = 0;
        for (int64_t k = 0; k < V; k++) {
            int64_t p = d->row[res];
            int64_t Bk = B + d->carry[p - 1];
            if (k == 0) B0 = Bk;
            ds[k] = Bk - B0;
            res += s; B += d->base_adv;
            if (res >= q) { res -= q; B += 1; }
        }
        cap += V * (m + ds[V - 1]);
        res0 = mod64(res0 + gstep, q);
    }
    d->gW = (double *)calloc(cap > 0 ? cap : 1, sizeof(double));
    res0 = 0;
    for (int64_t g = 0; g < ngrp; g++) {
        d->grow[res0] = (int32_t)(g + 1);
        int64_t res = res0, B = 0, B0 = 0;
        for (int64_t k = 0; k < V; k++) {
            int64_t p = d->row[res];
            int64_t Bk = B + d->carry[p - 1];
            if (k == 0) B0 = Bk;
            ps[k] = p;
            ds[k] = Bk - B0;
            d->gAr[g * V + k] = d->Ar[p - 1];
            d->gAi[g * V + k] = d->Ai[p - 1];
            res += s; B += d->base_adv;
            if (res >= q) { res -= q; B += 1; }
        }
        int64_t nj = m + ds[V - 1];
        d->gnj[g] = (int32_t)nj;
        d->goff[g] = total;
        for (int64_t k = 0; k < V; k++)
            for (int64_t i = 1; i <= m; i++)
                d->gW[total + (ds[k] + i - 1) * V + k] = d->W[(ps[k] - 1) * m + (i - 1)];
        total += V * nj;
        res0 = mod64(res0 + gstep, q);
    }
    d->goff[ngrp] = total;
    (void)f64;
    free(ps);
    free(ds);
}

/* build_direct_plans, directinterp.jl:345-405 */
or_plans *or_plans_build(int32_t nharms, int32_t m, double hidr, double r_lo,
                         int32_t weights_f64) {
    if (m % 2 != 0 || nharms < 1) return NULL;
    or_plans *pl = (or_plans *)calloc(1, sizeof(or_plans));
    pl->nharms = nharms; pl->m = m; pl->weights_f64 = weights_f64;
    pl->hidr = hidr; pl->r_lo = r_lo;
    pl->p = (or_plan *)calloc(nharms, sizeof(or_plan));
    double lodr = hidr / nharms;
    int64_t pnum = 0, q = 0;
    int rq = or_trial_grid_rational(lodr, &pnum, &q);
    pl->fallback = !rq;
    int64_t m2 = m / 2;
    for (int32_t h = 1; h <= nharms; h++) {
        or_plan *d = &pl->p[h - 1];
        double hr = (double)h * r_lo;
        d->h = h; d->m = m;
        d->rfloor0 = (int64_t)floor(hr);
        d->dr0 = hr - (double)d->rfloor0;
        if (!rq) { d->q = 1; d->P = 0; d->pnum = 1; continue; }
        int64_t step = (int64_t)h * pnum;
        d->q = q; d->pnum = pnum;
        d->s = mod64(step, q);
        d->base_adv = fld64(step, q);
        d->P = q / gcd64(step, q);
        d->row = (int32_t *)calloc(q, sizeof(int32_t));
        d->carry = (uint8_t *)malloc(d->P);
        d->W = (double *)malloc(sizeof(double) * m * d->P);
        d->Ar = (double *)malloc(sizeof(double) * d->P);
        d->Ai = (double *)malloc(sizeof(double) * d->P);
        int64_t res = 0;
        for (int64_t p = 1; p <= d->P; p++) {
            d->row[res] = (int32_t)p;
            double u = d->dr0 + (double)res / (double)q;
            d->carry[p - 1] = u >= 1.0;
            double dr = d->carry[p - 1] ? u - 1.0 : u;
            double *W = d->W + (p - 1) * m;
            if (dr == 0.0) {
                d->Ar[p - 1] = 1.0; d->Ai[p - 1] = 0.0;
                for (int64_t i = 1; i <= m; i++) W[i - 1] = (i == m2) ? 1.0 : 0.0;
            } else {
                double ar, ai;
                kernel_A(dr, &ar, &ai);
                d->Ar[p - 1] = wt_round(ar, weights_f64);
                d->Ai[p - 1] = wt_round(ai, weights_f64);
                for (int64_t i = 1; i <= m; i++)
                    W[i - 1] = wt_round(1.0 / (dr - (double)(i - m2)), weights_f64);
            }
            res += d->s;
            if (res >= q) res -= q;
        }
        build_group_table(d, weights_f64);
    }
    return pl;
}

void or_plans_free(or_plans *pl) {
    if (!pl) return;
    for (int32_t h = 0; h < pl->nharms; h++) {
        or_plan *d = &pl->p[h];
        free(d->row); free(d->carry); free(d->W); free(d->Ar); free(d->Ai);
        free(d->grow); free(d->goff); free(d->gnj); free(d->gW);
        free(d->gAr); free(d->gAi);
    }
    free(pl->p);
    free(pl);
}

int or_plan_info(const or_plans *pl, int32_t h, int64_t *info, double *dr0) {
    if (!pl || h < 1 || h > pl->nharms) return OR_EINVAL;
    const or_plan *d = &pl->p[h - 1];
    info[0] = d->q; info[1] = d->s; info[2] = d->base_adv; info[3] = d->pnum;
    info[4] = d->rfloor0; info[5] = d->P; info[6] = d->ngrp;
    info[7] = d->q ? mod64(DIRECT_GROUP_V * d->s, d->q) : 0;
    *dr0 = d->dr0;
    return OR_OK;
}

int or_plan_tables(const or_plans *pl, int32_t h, int32_t *row, uint8_t *carry,
                   double *W, double *A) {
    if (!pl || h < 1 || h > pl->nharms) return OR_EINVAL;
    const or_plan *d = &pl->p[h - 1];
    if (d->P == 0) return OR_EINVAL;
    if (row) memcpy(row, d->row, sizeof(int32_t) * d->q);
    if (carry) memcpy(carry, d->carry, d->P);
    if (W) memcpy(W, d->W, sizeof(double) * d->m * d->P);
    if (A) for (int64_t p = 0; p < d->P; p++) { A[2 * p] = d->Ar[p]; A[2 * p + 1] = d->Ai[p]; }
    return OR_OK;
}

/* direct_chunk_state, directinterp.jl:414-418 (Int128 widemul) */
static void chunk_state(const or_plan *d, int64_t t0, int64_t *res, int64_t *qint) {
    __int128 T = (__int128)t0 * (__int128)(d->h * d->pnum);
    __int128 q = d->q;
    __int128 f = T / q;
    if ((T % q) != 0 && (T < 0)) f -= 1;
    *qint = (int64_t)f;
    *res = (int64_t)(T - f * q);
}

int or_chunk_state(const or_plans *pl, int32_t h, int64_t t0, int64_t *res,
                   int64_t *qint) {
    if (!pl || h < 1 || h > pl->nharms || pl->fallback) return OR_EINVAL;
    chunk_state(&pl->p[h - 1], t0, res, qint);
    return OR_OK;
}

/* _fill_row_direct_slow!, directinterp.jl:623-661 (non-rational lodr) */
static void fill_row_slow(const or_plans *pl, const or_plan *d, const float *amps,
                          int64_t namps, int64_t N, int64_t t0, int64_t n,
                          double *re, double *im) {
    int64_t m = d->m, m2 = m / 2, h = d->h;
    double lodr = pl->hidr / pl->nharms;
    double hr_lo = (double)d->rfloor0 + d->dr0;
    double dh = (double)h * lodr;
    int64_t Nhalf = N / 2;
    double rfirst = hr_lo + (double)t0 * dh;
    double rlast = hr_lo + (double)(t0 + n - 1) * dh;
    int64_t lo = (int64_t)floor(rfirst + 1e-15) + 2 - m2;
    int64_t hi = (int64_t)floor(rlast + 1e-15) + 1 + m2;
    if (!(lo >= 1 && hi <= namps && (int64_t)floor(rlast) < Nhalf)) return;
    for (int64_t k = 0; k < n; k++) {
        double r = hr_lo + (double)(t0 + k) * dh;
        int64_t rint = (int64_t)floor(r + 1e-15) + 1;
        double dr = r - floor(r);
        int64_t base = rint - m2;
        if (dr == 0.0) {
            re[k] = amps[2 * (base + m2 - 1)];
            im[k] = amps[2 * (base + m2 - 1) + 1];
            continue;
        }
        double sre = 0.0, sim = 0.0;
        for (int64_t i = 1; i <= m; i++) {
            double w = 1.0 / (dr - (double)(i - m2));
            sre = fma(w, (double)amps[2 * (base + i - 1)], sre);
            sim = fma(w, (double)amps[2 * (base + i - 1) + 1], sim);
        }
        double Ar, Ai;
        kernel_A(dr, &Ar, &Ai);
        re[k] = Ar * sre - Ai * sim;
        im[k] = Ar * sim + Ai * sre;
    }
}

/* Row guard on the TRUE trial range, directinterp.jl:490-496 */
static int row_guard(const or_plan *d, int64_t namps, int64_t N, int64_t t0,
                     int64_t n) {
    int64_t m2 = d->m / 2, Nhalf = N / 2, res, qint, resl, qintl;
    chunk_state(d, t0, &res, &qint);
    int64_t p0 = d->row[res];
    int64_t lo_trial = d->rfloor0 + qint + d->carry[p0 - 1] + 2 - m2;
    chunk_state(d, t0 + n - 1, &resl, &qintl);
    int64_t pl = d->row[resl];
    int64_t hi_trial = d->rfloor0 + qintl + d->carry[pl - 1] + 1 + m2;
    return lo_trial >= 1 && hi_trial <= namps && (d->rfloor0 + qintl) < Nhalf;
}

/* One complex product in WT = Float32 or Float64 with no contraction, as
 * `gA[k, g] * complex(ar[k], ai[k])` evaluates (directinterp.jl:603). */
static void cmul_wt(double ar, double ai, double xr, double xi, int32_t f64,
                    double *yr, double *yi) {
    if (f64) {
        *yr = ar * xr - ai * xi;
        *yi = ar * xi + ai * xr;
    } else {
        float a = (float)ar, b = (float)ai, c = (float)xr, d = (float)xi;
        float p1 = a * c, p2 = b * d, p3 = a * d, p4 = b * c;
        *yr = (double)(p1 - p2);
        *yi = (double)(p3 + p4);
    }
}

/* fill_harmonic_row_direct!, directinterp.jl:470-572, trials-axis groups. */
static void fill_row_group(const or_plans *pl, const or_plan *d,
                           const float *amps, int64_t namps, int64_t N,
                           int64_t t0, int64_t n, double *re, double *im) {
    const int64_t V = DIRECT_GROUP_V;
    int64_t m = d->m, m2 = m / 2, q = d->q, s = d->s;
    int32_t f64 = pl->weights_f64;
    if (!row_guard(d, namps, N, t0, n)) return;
    int64_t T_first = fld64(t0, V) * V;
    int64_t T_last = fld64(t0 + n - 1, V) * V + V - 1;
    int64_t res_e, qint_e, resx, qintx;
    chunk_state(d, T_first, &res_e, &qint_e);
    int64_t lo = d->rfloor0 + qint_e + d->carry[d->row[res_e] - 1] + 2 - m2;
    chunk_state(d, T_last, &resx, &qintx);
    int64_t hi = d->rfloor0 + qintx + d->carry[d->row[resx] - 1] + 1 + m2;
    int64_t nw = hi - lo + 1;
    float *pr = (float *)malloc(sizeof(float) * nw);
    float *pi = (float *)malloc(sizeof(float) * nw);
    for (int64_t i = 1; i <= nw; i++) {
        int64_t b = lo + i - 1;                 /* 1-based bin */
        if (b >= 1 && b <= namps) {
            pr[i - 1] = amps[2 * (b - 1)];
            pi[i - 1] = amps[2 * (b - 1) + 1];
        } else {
            pr[i - 1] = 0.0f;
            pi[i - 1] = 0.0f;
        }
    }
    int64_t res = res_e, qint = qint_e, T = T_first;
    float arf[DIRECT_GROUP_V], aif[DIRECT_GROUP_V];
    double ard[DIRECT_GROUP_V], aid[DIRECT_GROUP_V];
    while (T <= T_last) {
        int64_t g = d->grow[res] - 1;
        int64_t b0 = d->rfloor0 + qint + d->carry[d->row[res] - 1] + 2 - m2 - lo;
        int64_t o = d->goff[g];
        int64_t nj = d->gnj[g];
        int64_t j0 = T - t0;
        const double *gW = d->gW + o;
        if (f64) {
            for (int64_t k = 0; k < V; k++) { ard[k] = 0.0; aid[k] = 0.0; }
            for (int64_t j = 1; j <= nj; j++) {
                double rv = pr[b0 + j - 1], iv = pi[b0 + j - 1];
                const double *w = gW + (j - 1) * V;
                for (int64_t k = 0; k < V; k++) {
                    ard[k] = fma(w[k], rv, ard[k]);
                    aid[k] = fma(w[k], iv, aid[k]);
                }
            }
        } else {
            for (int64_t k = 0; k < V; k++) { arf[k] = 0.0f; aif[k] = 0.0f; }
            for (int64_t j = 1; j <= nj; j++) {
                float rv = pr[b0 + j - 1], iv = pi[b0 + j - 1];
                const double *w = gW + (j - 1) * V;
                for (int64_t k = 0; k < V; k++) {
                    float wk = (float)w[k];
                    arf[k] = fmaf(wk, rv, arf[k]);
                    aif[k] = fmaf(wk, iv, aif[k]);
                }
            }
            for (int64_t k = 0; k < V; k++) { ard[k] = arf[k]; aid[k] = aif[k]; }
        }
        for (int64_t k = 0; k < V; k++) {
            int64_t j = j0 + k;                     /* 0-based local column */
            if (j >= 0 && j < n) {
                cmul_wt(d->gAr[g * V + k], d->gAi[g * V + k], ard[k], aid[k], f64,
                        &re[j], &im[j]);
            }
        }
        int64_t tot = res + V * s;
        qint += V * d->base_adv + fld64(tot, q);
        res = mod64(tot, q);
        T += V;
    }
    free(pr);
    free(pi);
}

/* The same row, trial by trial: an m-term FMA chain against each trial's own
 * weight column (no group tables). */
static void fill_row_pertrial(const or_plans *pl, const or_plan *d,
                              const float *amps, int64_t namps, int64_t N,
                              int64_t t0, int64_t n, double *re, double *im) {
    int64_t m = d->m, m2 = m / 2;
    int32_t f64 = pl->weights_f64;
    if (!row_guard(d, namps, N, t0, n)) return;
    for (int64_t j = 0; j < n; j++) {
        int64_t res, qint;
        chunk_state(d, t0 + j, &res, &qint);
        int64_t p = d->row[res];
        int64_t b = d->rfloor0 + qint + d->carry[p - 1] + 1 - m2;   /* 0-based */
        const double *W = d->W + (p - 1) * m;
        double xr, xi;
        if (f64) {
            double sr = 0.0, si = 0.0;
            for (int64_t i = 0; i < m; i++) {
                sr = fma(W[i], (double)amps[2 * (b + i)], sr);
                si = fma(W[i], (double)amps[2 * (b + i) + 1], si);
            }
            xr = sr; xi = si;
        } else {
            float sr = 0.0f, si = 0.0f;
            for (int64_t i = 0; i < m; i++) {
                sr = fmaf((float)W[i], amps[2 * (b + i)], sr);
                si = fmaf((float)W[i], amps[2 * (b + i) + 1], si);
            }
            xr = sr; xi = si;
        }
        cmul_wt(d->Ar[p - 1], d->Ai[p - 1], xr, xi, f64, &re[j], &im[j]);
    }
}

static int fill_chunk_impl(const or_plans *pl, const float *amps, int64_t namps,
                           int64_t N, int64_t t0, int64_t n, int64_t ncol,
                           double *re, double *im, int pertrial) {
    if (!pl || n < 0 || ncol < n) return OR_EINVAL;
    int32_t nh = pl->nharms;
    memset(re, 0, sizeof(double) * (nh + 1) * ncol);
    memset(im, 0, sizeof(double) * (nh + 1) * ncol);
    if (n < 1) return OR_OK;
    for (int32_t h = 1; h <= nh; h++) {
        const or_plan *d = &pl->p[h - 1];
        double *rr = re + (int64_t)h * ncol, *ii = im + (int64_t)h * ncol;
        if (d->P == 0) fill_row_slow(pl, d, amps, namps, N, t0, n, rr, ii);
        else if (pertrial) fill_row_pertrial(pl, d, amps, namps, N, t0, n, rr, ii);
        else fill_row_group(pl, d, amps, namps, N, t0, n, rr, ii);
    }
    return OR_OK;
}

int or_fill_chunk(const or_plans *pl, const float *amps, int64_t namps,
                  int64_t N, int64_t t0, int64_t n, int64_t ncol, double *re,
                  double *im) {
    return fill_chunk_impl(pl, amps, namps, N, t0, n, ncol, re, im, 0);
}

int or_fill_chunk_pertrial(const or_plans *pl, const float *amps, int64_t namps,
                           int64_t N, int64_t t0, int64_t n, int64_t ncol,
                           double *re, double *im) {
    return fill_chunk_impl(pl, amps, namps, N, t0, n, ncol, re, im, 1);
}

/* ------------------------------------------------------------------------ */
/* Batched c2r: x[n] = X0.re + (-1)^n XH.re + 2 sum_{h=1}^{H-1}               */
/*   (Re X_h cos(2 pi h n/N) - Im X_h sin(2 pi h n/N)), computed as an even    */
/*   part E and odd part O with x[n] = E+O, x[N-n] = E-O.                       */
/* ------------------------------------------------------------------------ */
int or_brfft_rows(const double *re, const double *im, int64_t ncol_stride,
                  int32_t k, int32_t nbins, int64_t ncol, double *out) {
    if (nbins < 2 || nbins % 2 != 0 || k < 1) return OR_EINVAL;
    int32_t N = nbins, H = nbins / 2;
    double *cr = (double *)malloc(sizeof(double) * (H + 1) * (H + 1));
    double *ci = (double *)malloc(sizeof(double) * (H + 1) * (H + 1));
    for (int32_t n = 0; n <= H; n++)
        for (int32_t h = 0; h <= H; h++) {
            int64_t e = ((int64_t)h * n) % N;
            double a = 2.0 * (double)e / (double)N;
            double c = or_cospi(a), s = or_sinpi(a);
            if (h == 0) { c = 1.0; s = 0.0; }
            else if (h == H) { c = (n % 2) ? -1.0 : 1.0; s = 0.0; }
            else { c *= 2.0; s *= 2.0; }
            cr[n * (H + 1) + h] = c;
            ci[n * (H + 1) + h] = -s;
        }
    const int64_t CB = 256;
    double E[256], O[256];
    for (int64_t j0 = 0; j0 < ncol; j0 += CB) {
        int64_t nb = ncol - j0 < CB ? ncol - j0 : CB;
        for (int32_t n = 0; n <= H; n++) {
            for (int64_t j = 0; j < nb; j++) { E[j] = 0.0; O[j] = 0.0; }
            for (int32_t h = 0; h <= H; h++) {
                double c = cr[n * (H + 1) + h], s = ci[n * (H + 1) + h];
                const double *xr = re + (int64_t)h * k * ncol_stride + j0;
                const double *xi = im + (int64_t)h * k * ncol_stride + j0;
                for (int64_t j = 0; j < nb; j++) {
                    E[j] = fma(c, xr[j], E[j]);
                    O[j] = fma(s, xi[j], O[j]);
                }
            }
            double *o1 = out + (int64_t)n * ncol + j0;
            if (n == 0 || n == H) {
                for (int64_t j = 0; j < nb; j++) o1[j] = E[j];
            } else {
                double *o2 = out + (int64_t)(N - n) * ncol + j0;
                for (int64_t j = 0; j < nb; j++) {
                    o1[j] = E[j] + O[j];
                    o2[j] = E[j] - O[j];
                }
            }
        }
    }
    free(cr);
    free(ci);
    return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* search.jl: medians, widths, sigma, boxcar                                  */
/* ------------------------------------------------------------------------ */

/* k-th smallest (0-based) of v[0..n-1] in place, quickselect (median of 3). */
static void select_k(double *v, int64_t n, int64_t kth) {
    int64_t lo = 0, hi = n - 1;
    while (hi > lo) {
        if (hi - lo < 16) {
            for (int64_t i = lo + 1; i <= hi; i++) {
                double x = v[i];
                int64_t j = i - 1;
                while (j >= lo && v[j] > x) { v[j + 1] = v[j]; j--; }
                v[j + 1] = x;
            }
            return;
        }
        int64_t mid = lo + (hi - lo) / 2;
        double a = v[lo], b = v[mid], c = v[hi], t;
        if (b < a) { t = a; a = b; b = t; }
        if (c < b) { t = b; b = c; c = t; if (b < a) { t = a; a = b; b = t; } }
        double pivot = b;
        int64_t i = lo, j = hi;
        while (i <= j) {
            while (v[i] < pivot) i++;
            while (v[j] > pivot) j--;
            if (i <= j) { t = v[i]; v[i] = v[j]; v[j] = t; i++; j--; }
        }
        if (kth <= j) hi = j;
        else if (kth >= i) lo = i;
        else return;
    }
}

/* _median!, search.jl:296-305: mean of the two central order statistics for
 * even n (computed as T(0.5) * (lower + upper)), middle one for odd n. */
double or_median(double *v, int64_t n) {
    if (n <= 0) return NAN;
    int64_t half = n / 2;
    select_k(v, n, half);                     /* upper median at v[half] */
    double upper = v[half];
    if (n % 2) return upper;
    double lower = v[0];
    for (int64_t i = 1; i < half; i++) if (v[i] > lower) lower = v[i];
    return 0.5 * (lower + upper);
}

/* boxcar_widths, search.jl:387-397 */
int32_t or_boxcar_widths(int32_t nbins, double fsp, double maxfrac, int32_t *out) {
    int64_t wmax = (int64_t)floor(maxfrac * nbins);
    if (wmax < 1) wmax = 1;
    int32_t n = 0;
    int64_t w = 1;
    while (w <= wmax && n < 64) {
        out[n++] = (int32_t)w;
        int64_t nw = (int64_t)floor(fsp * (double)w);
        w = nw > w + 1 ? nw : w + 1;
    }
    return n;
}

static int cmp_i32(const void *a, const void *b) {
    int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
    return (x > y) - (x < y);
}

/* ladder_boxcar_widths, search.jl:443-459 */
int32_t or_ladder_boxcar_widths(int32_t nbins, int32_t k, const or_params *p,
                                int32_t *out) {
    int32_t ws[64];
    int32_t nw = or_boxcar_widths(nbins, p->boxcar_fsp, p->boxcar_maxfrac, ws);
    int32_t nd = p->ndecim;
    int32_t ks[64];
    if (nd > 64) nd = 64;
    for (int32_t i = 0; i < nd; i++) ks[i] = p->decimations[i];
    qsort(ks, nd, sizeof(int32_t), cmp_i32);
    int ok = nd > 1;
    for (int32_t i = 0; ok && i + 1 < nd; i++)
        if (!(ks[i + 1] <= (LADDER_WMAX / 2) * ks[i])) ok = 0;
    if (!ok) { memcpy(out, ws, sizeof(int32_t) * nw); return nw; }
    int32_t kmin = ks[0], kmax = ks[nd - 1];
    int32_t n = 0;
    for (int32_t i = 0; i < nw; i++) {
        int32_t w = ws[i];
        if ((w >= 2 || k == kmin) && (w <= LADDER_WMAX || k == kmax)) out[n++] = w;
    }
    if (n == 0) { memcpy(out, ws, sizeof(int32_t) * nw); return nw; }
    return n;
}

/* _block_sigma, search.jl:490-566 */
double or_block_sigma(const double *M, int64_t bin_stride, int64_t col_stride,
                      int32_t nbins, int64_t n, int64_t cap, int32_t center) {
    int64_t NN = (int64_t)nbins * n;
    if (NN == 0) return 0.0;
    double *buf = (double *)malloc(sizeof(double) * (cap > 0 ? cap : 1));
    int64_t ns = 0;
    if (NN <= cap) {
        for (int64_t j = 0; j < n; j++)
            for (int32_t i = 0; i < nbins; i++)
                buf[ns++] = M[i * bin_stride + j * col_stride];
    } else {
        int64_t ncol = cap / nbins;
        if (ncol < 1) ncol = 1;
        if (ncol > n) ncol = n;
        for (int64_t c = 1; c <= ncol; c++) {
            int64_t j = 1 + ((c - 1) * n) / ncol;
            int64_t take = cap - ns < nbins ? cap - ns : nbins;
            if (take == 0) break;
            for (int64_t i = 0; i < take; i++)
                buf[ns++] = M[i * bin_stride + (j - 1) * col_stride];
        }
    }
    if (center == 1) {
        double *tmp = (double *)malloc(sizeof(double) * ns);
        memcpy(tmp, buf, sizeof(double) * ns);
        double med = or_median(tmp, ns);
        free(tmp);
        for (int64_t t = 0; t < ns; t++) buf[t] = fabs(buf[t] - med);
    } else {
        for (int64_t t = 0; t < ns; t++) buf[t] = fabs(buf[t]);
    }
    double s = 1.4826 * or_median(buf, ns);
    free(buf);
    return s;
}

/* _boxcar_psum! + _boxcar_scan, search.jl:571-598 (P = Float64) */
double or_boxcar_scan_col(const double *col, int64_t stride, int32_t nbins,
                          const int32_t *widths, int32_t nw, double invsigma,
                          double b) {
    int32_t wmax = widths[nw - 1];
    double ps[1024];
    if (nbins + wmax + 1 > 1024) return NAN;
    ps[0] = 0.0;
    for (int32_t i = 1; i <= nbins + wmax; i++) {
        int32_t idx = i > nbins ? i - nbins : i;
        ps[i] = ps[i - 1] + (col[(idx - 1) * stride] - b);
    }
    double best = -INFINITY;
    for (int32_t a = 0; a < nw; a++) {
        int32_t w = widths[a];
        double invsw = invsigma / sqrt((double)w);
        double m = ps[w] - ps[0];
        for (int32_t p = 2; p <= nbins; p++) {
            double d = ps[p - 1 + w] - ps[p - 1];
            m = d > m ? d : m;
        }
        double cand = m * invsw;
        if (cand > best) best = cand;
    }
    return best;
}

/* _boxcar_exact, search.jl:893-904 (baseline median by any exact selection:
 * quickselect and the sorting network return the same order statistics). */
double or_boxcar_exact_col(const double *col, int64_t stride, int32_t nbins,
                           const int32_t *widths, int32_t nw, double invsigma) {
    double mb[1024];
    if (nbins > 1024) return NAN;
    for (int32_t i = 0; i < nbins; i++) mb[i] = col[i * stride];
    double med = or_median(mb, nbins);
    return or_boxcar_scan_col(col, stride, nbins, widths, nw, invsigma, med);
}

/* boxcar_best_width, search.jl:618-639 */
int32_t or_boxcar_best_width(const double *prof, int32_t nbins, double fsp,
                             double maxfrac) {
    int32_t ws[64];
    int32_t nw = or_boxcar_widths(nbins, fsp, maxfrac, ws);
    double mb[1024];
    for (int32_t i = 0; i < nbins; i++) mb[i] = prof[i];
    double med = or_median(mb, nbins);
    int32_t wmax = ws[nw - 1];
    double ps[1024];
    ps[0] = 0.0;
    for (int32_t i = 1; i <= nbins + wmax; i++) {
        int32_t idx = i > nbins ? i - nbins : i;
        ps[i] = ps[i - 1] + (prof[idx - 1] - med);
    }
    double best = -INFINITY;
    int32_t bestw = ws[0];
    for (int32_t a = 0; a < nw; a++) {
        int32_t w = ws[a];
        double invsw = 1.0 / sqrt((double)w);
        double m = ps[w] - ps[0];
        for (int32_t p = 2; p <= nbins; p++) {
            double d = ps[p - 1 + w] - ps[p - 1];
            m = d > m ? d : m;
        }
        double cand = m * invsw;
        if (cand > best) { best = cand; bestw = w; }
    }
    return bestw;
}

/* _bc_scan_batch! for one lane, Float32 tile (search.jl:743-779): the batched
 * kernel's per-profile arithmetic is independent of the batch width. */
static float gate_f32_col(const double *col, int64_t stride, int32_t nbins,
                          const int32_t *widths, int32_t nw, float invsig_t) {
    float ps[1024];
    int32_t wmax = widths[nw - 1];
    ps[0] = 0.0f;
    for (int32_t i = 1; i <= nbins + wmax; i++) {
        int32_t idx = i > nbins ? i - nbins : i;
        ps[i] = ps[i - 1] + (float)col[(idx - 1) * stride];
    }
    float res = -INFINITY;
    for (int32_t a = 0; a < nw; a++) {
        int32_t w = widths[a];
        float invsw = invsig_t / sqrtf((float)w);
        float mb = ps[w] - ps[0];
        for (int32_t p = 2; p <= nbins; p++) {
            float d = ps[p - 1 + w] - ps[p - 1];
            mb = d > mb ? d : mb;
        }
        float c = mb * invsw;
        res = c > res ? c : res;
    }
    return res;
}

/* boxcar_metrics!, search.jl:825-845, with _boxcar_gate! :786-807 */
int or_boxcar_metrics(const double *profs, int64_t ncol_stride, int64_t n,
                      const int32_t *widths, int32_t nw, int32_t nbins,
                      double invsigma, double medcut, double *mvals) {
    if (medcut == -INFINITY || invsigma <= 0) {
        for (int64_t j = 0; j < n; j++)
            mvals[j] = invsigma > 0
                ? or_boxcar_exact_col(profs + j, ncol_stride, nbins, widths, nw, invsigma)
                : 0.0;
        return OR_OK;
    }
    float invsig_t = (float)invsigma;
    int64_t j0 = 0;
    while (j0 + BC_BATCH <= n) {
        for (int64_t b = 0; b < BC_BATCH; b++)
            mvals[j0 + b] = (double)gate_f32_col(profs + j0 + b, ncol_stride, nbins,
                                                 widths, nw, invsig_t);
        j0 += BC_BATCH;
    }
    for (int64_t j = j0; j < n; j++)
        mvals[j] = or_boxcar_scan_col(profs + j, ncol_stride, nbins, widths, nw,
                                      invsigma, 0.0);
    for (int64_t j = 0; j < n; j++) {
        if (mvals[j] < medcut) continue;
        mvals[j] = or_boxcar_exact_col(profs + j, ncol_stride, nbins, widths, nw,
                                       invsigma);
    }
    return OR_OK;
}

/* snr_metrics, search.jl:933-951 (profs column-major (nbins, L)) */
int or_snr_metrics(const double *profs, int32_t nbins, int64_t L,
                   int64_t sigma_samples, int32_t center, const int32_t *widths,
                   int32_t nw, double *out) {
    int32_t wb[64];
    if (!widths) {
        nw = or_boxcar_widths(nbins, 1.5, 0.3, wb);
        widths = wb;
    }
    int64_t cap = (int64_t)nbins * L;
    if (sigma_samples > 0 && sigma_samples < cap) cap = sigma_samples;
    double sigma = or_block_sigma(profs, 1, nbins, nbins, L, cap, center);
    double invsigma = sigma > 0 ? 1.0 / sigma : 0.0;
    for (int64_t j = 0; j < L; j++)
        out[j] = invsigma > 0
            ? or_boxcar_exact_col(profs + j * nbins, 1, nbins, widths, nw, invsigma)
            : 0.0;
    return OR_OK;
}

/* ------------------------------------------------------------------------ */
/* Drivers                                                                    */
/* ------------------------------------------------------------------------ */

/* chunk_metrics, search.jl:1740-1755 */
int or_chunk_metrics(const or_params *p, const float *amps, int64_t namps,
                     int64_t N, double rstart, int64_t n, int32_t weights_f64,
                     double *out) {
    if (p->m % 2 != 0) return OR_EINVAL;
    if (n < 1) return OR_OK;
    int32_t nh = p->nharms, nbins = 2 * nh;
    or_plans *pl = or_plans_build(nh, p->m, p->hidr, rstart, weights_f64);
    if (!pl) return OR_EINVAL;
    double *re = (double *)malloc(sizeof(double) * (nh + 1) * n);
    double *im = (double *)malloc(sizeof(double) * (nh + 1) * n);
    double *pr = (double *)malloc(sizeof(double) * nbins * n);
    or_fill_chunk(pl, amps, namps, N, 0, n, n, re, im);
    or_brfft_rows(re, im, n, 1, nbins, n, pr);
    int64_t cap = (int64_t)nbins * n < BOXCAR_SIGMA_SAMPLES ? (int64_t)nbins * n
                                                            : BOXCAR_SIGMA_SAMPLES;
    double sigma = or_block_sigma(pr, n, 1, nbins, n, cap, 0);
    double invsigma = sigma > 0 ? 1.0 / sigma : 0.0;
    int32_t ws[64];
    int32_t nw = or_ladder_boxcar_widths(nbins, 1, p, ws);
    for (int64_t j = 0; j < n; j++)
        out[j] = invsigma > 0
            ? or_boxcar_exact_col(pr + j, n, nbins, ws, nw, invsigma) : 0.0;
    free(re); free(im); free(pr);
    or_plans_free(pl);
    return OR_OK;
}

/* uniform_linear_interp, search.jl:974-986 */
static void uniform_linear_interp(double r, int64_t lobin, int32_t nb,
                                  const double *amps, int64_t K, double *o) {
    double pp = (r - (double)lobin) * nb;
    if (pp <= 0) { o[0] = amps[0]; o[1] = amps[1]; return; }
    if (pp >= (double)(K - 1)) { o[0] = amps[2 * (K - 1)]; o[1] = amps[2 * (K - 1) + 1]; return; }
    int64_t i0 = (int64_t)floor(pp);
    double f = pp - (double)i0;
    o[0] = amps[2 * i0] * (1 - f) + amps[2 * (i0 + 1)] * f;
    o[1] = amps[2 * i0 + 1] * (1 - f) + amps[2 * (i0 + 1) + 1] * f;
}

/* reference_profiles, search.jl:1011-1049 (+ coherent_profiles :964) */
int or_reference_profiles(const or_params *p, const float *amps, int64_t namps,
                          int64_t N, const double *rfund, int64_t L,
                          int32_t kernel, double *out) {
    if (kernel != 0 && kernel != 1) return OR_EINVAL;
    if (p->m % 2 != 0) return OR_EINVAL;
    if (L < 1) return OR_OK;
    int32_t nh = p->nharms, m = p->m, nb = p->numbetween, m2 = m / 2;
    int64_t Nhalf = N / 2;
    double *ft = (double *)malloc(sizeof(double) * 2 * namps);
    for (int64_t i = 0; i < 2 * namps; i++) ft[i] = amps[i];
    double *re = (double *)calloc((nh + 1) * L, sizeof(double));
    double *im = (double *)calloc((nh + 1) * L, sizeof(double));
    int st = OR_OK;
    for (int32_t h = 1; h <= nh && st == OR_OK; h++) {
        double rmin = rfund[0] * h, rmax = rfund[L - 1] * h;
        int64_t lobin = (int64_t)floor(rmin);
        int64_t hibin = (int64_t)ceil(rmax) + 1;
        int64_t numbins = hibin - lobin;
        if (!(lobin >= m2 && (lobin + numbins + m2) <= namps && hibin < Nhalf)) continue;
        if (kernel == 0) {
            int64_t K = numbins * nb;
            double *g = (double *)malloc(sizeof(double) * 2 * K);
            st = or_finterp_fft(lobin, numbins, nb, ft, namps, m, 0, g);
            for (int64_t j = 0; j < L && st == OR_OK; j++) {
                double o[2];
                uniform_linear_interp(rfund[j] * h, lobin, nb, g, K, o);
                re[h * L + j] = o[0];
                im[h * L + j] = o[1];
            }
            free(g);
        } else {
            for (int64_t j = 0; j < L && st == OR_OK; j++) {
                double o[2];
                st = or_fourier_interp(rfund[j] * h, ft, namps, m, o);
                re[h * L + j] = o[0];
                im[h * L + j] = o[1];
            }
        }
    }
    if (st == OR_OK) {
        int32_t nbins = 2 * nh;
        double *tmp = (double *)malloc(sizeof(double) * nbins * L);
        or_brfft_rows(re, im, L, 1, nbins, L, tmp);
        double sc = 1.0 / (double)nbins;
        for (int32_t i = 0; i < nbins; i++)
            for (int64_t j = 0; j < L; j++) out[j * nbins + i] = tmp[i * L + j] * sc;
        free(tmp);
    }
    free(ft); free(re); free(im);
    return st;
}

/* block_metrics, search.jl:1061-1079 */
int or_block_metrics(const or_params *p, const float *amps, int64_t namps,
                     int64_t N, const double *rfund, int64_t L, int32_t kernel,
                     const int32_t *widths, int32_t nw, double *out) {
    int32_t nbins = 2 * p->nharms;
    double *profs = (double *)malloc(sizeof(double) * nbins * (L > 0 ? L : 1));
    int st = or_reference_profiles(p, amps, namps, N, rfund, L, kernel, profs);
    if (st == OR_OK && L > 0) {
        int32_t ws[64];
        if (!widths) {
            nw = or_ladder_boxcar_widths(nbins, 1, p, ws);
            widths = ws;
        }
        st = or_snr_metrics(profs, nbins, L, BOXCAR_SIGMA_SAMPLES, 0, widths, nw, out);
    }
    free(profs);
    return st;
}

/* decim_pass! metrics for one fold (search.jl:1664-1691), exact scan. */
int or_decim_metrics(const or_params *p, const float *amps, int64_t namps,
                     int64_t N, double rstart, int64_t n, int32_t k,
                     int32_t weights_f64, double *out, int64_t *nvalid) {
    int32_t nh = p->nharms;
    if (k < 1 || nh / k < 1) return OR_EINVAL;
    int32_t Hk = nh / k, nbins = 2 * Hk;
    double lodr = p->hidr / nh;
    or_plans *pl = or_plans_build(nh, p->m, p->hidr, rstart, weights_f64);
    if (!pl) return OR_EINVAL;
    double *re = (double *)malloc(sizeof(double) * (nh + 1) * n);
    double *im = (double *)malloc(sizeof(double) * (nh + 1) * n);
    double *pr = (double *)malloc(sizeof(double) * nbins * n);
    or_fill_chunk(pl, amps, namps, N, 0, n, n, re, im);
    or_brfft_rows(re, im, n, k, nbins, n, pr);
    double nyq = (double)N / 2.0;
    int64_t nv = 0;
    while (nv < n && (double)k * (rstart + (double)nv * lodr) < nyq) nv++;
    int64_t cap = (int64_t)nbins * n < BOXCAR_SIGMA_SAMPLES ? (int64_t)nbins * n
                                                            : BOXCAR_SIGMA_SAMPLES;
    double sigma = or_block_sigma(pr, n, 1, nbins, nv, cap, 0);
    double invsigma = sigma > 0 ? 1.0 / sigma : 0.0;
    int32_t ws[64];
    int32_t nw = or_ladder_boxcar_widths(nbins, k, p, ws);
    for (int64_t j = 0; j < nv; j++)
        out[j] = invsigma > 0 ? or_boxcar_exact_col(pr + j, n, nbins, ws, nw, invsigma) : 0.0;
    *nvalid = nv;
    free(re); free(im); free(pr);
    or_plans_free(pl);
    return OR_OK;
}

typedef struct cand_vec { or_cand *v; int64_t n, cap; } cand_vec;
static void cv_push(cand_vec *cv, or_cand c) {
    if (cv->n == cv->cap) {
        cv->cap = cv->cap ? 2 * cv->cap : 256;
        cv->v = (or_cand *)realloc(cv->v, sizeof(or_cand) * cv->cap);
    }
    cv->v[cv->n++] = c;
}

/* _search_region!, search.jl:1768-1894.  mall (optional, [fold][total], fold
 * order 1 then the decimations as listed) receives every scored metric: the
 * metricstats branch's raw metrics when called with medcut = -inf.  sigmas
 * (optional, [chunk][fold], same fold order) receives each block's
 * _block_sigma (search.jl:1817; decim_pass!, :1688). */
int or_search_region(const or_params *p, const float *amps, int64_t namps,
                     int64_t N, double r_lo, int64_t total, int64_t Nprof,
                     int64_t c_begin, int64_t c_end, double threshold,
                     double medcut, int32_t nthreads, or_cand *out, int64_t cap,
                     int64_t *n_out, double *mall, double *sigmas) {
    if (p->m % 2 != 0 || Nprof < 1) return OR_EINVAL;
    int32_t nh = p->nharms, nbins = 2 * nh;
    double lodr = p->hidr / nh;
    if (medcut != medcut) medcut = threshold - p->boxcar_medmargin;
    int64_t nchunks = (total + Nprof - 1) / Nprof;
    if (c_end > nchunks) c_end = nchunks;
    if (c_begin < 0) c_begin = 0;
    *n_out = 0;
    if (c_end <= c_begin) return OR_OK;
    or_plans *pl = or_plans_build(nh, p->m, p->hidr, r_lo, 0);
    if (!pl) return OR_EINVAL;
    int32_t nt = 1;
#ifdef _OPENMP
    nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#endif
    int64_t nwork = c_end - c_begin;
    if (nt > nwork) nt = (int32_t)nwork;
    if (nt < 1) nt = 1;
    cand_vec *parts = (cand_vec *)calloc(nt, sizeof(cand_vec));
    /* decimation factors k > 1 in the order params lists them (search.jl:1239) */
    int32_t dk[64], nd = 0;
    for (int32_t i = 0; i < p->ndecim && nd < 64; i++)
        if (p->decimations[i] > 1) dk[nd++] = p->decimations[i];
    int32_t wbase[64], nwbase = or_ladder_boxcar_widths(nbins, 1, p, wbase);
    int32_t wdec[64][64], nwdec[64];
    for (int32_t d = 0; d < nd; d++)
        nwdec[d] = or_ladder_boxcar_widths(2 * (nh / dk[d]), dk[d], p, wdec[d]);
    double nyq = (double)N / 2.0;
#ifdef _OPENMP
#pragma omp parallel num_threads(nt)
#endif
    {
        int32_t tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double *re = (double *)malloc(sizeof(double) * (nh + 1) * Nprof);
        double *im = (double *)malloc(sizeof(double) * (nh + 1) * Nprof);
        double *pr = (double *)malloc(sizeof(double) * nbins * Nprof);
        double *mv = (double *)malloc(sizeof(double) * Nprof);
        cand_vec *cv = &parts[tid];
        for (int64_t c = c_begin + tid; c < c_end; c += nt) {
            int64_t i0 = c * Nprof;
            int64_t n = total - i0 < Nprof ? total - i0 : Nprof;
            double rstart = r_lo + (double)i0 * lodr;
            or_fill_chunk(pl, amps, namps, N, i0, n, Nprof, re, im);
            or_brfft_rows(re, im, Nprof, 1, nbins, n, pr);
            int64_t capb = (int64_t)nbins * Nprof < BOXCAR_SIGMA_SAMPLES
                               ? (int64_t)nbins * Nprof : BOXCAR_SIGMA_SAMPLES;
            double sig = or_block_sigma(pr, n, 1, nbins, n, capb, 0);
            if (sigmas) sigmas[c * (1 + nd)] = sig;
            double invs = sig > 0 ? 1.0 / sig : 0.0;
            or_boxcar_metrics(pr, n, n, wbase, nwbase, nbins, invs, medcut, mv);
            if (mall)
                for (int64_t j = 0; j < n; j++) mall[i0 + j] = mv[j];
            for (int64_t j = 0; j < n; j++)
                if (mv[j] > threshold) {
                    or_cand cc = {i0 + j, 1, nh, c, mv[j]};
                    cv_push(cv, cc);
                }
            for (int32_t d = 0; d < nd; d++) {
                int32_t k = dk[d], Hk = nh / k, nbk = 2 * Hk;
                or_brfft_rows(re, im, Nprof, k, nbk, n, pr);
                int64_t nv = 0;
                while (nv < n && (double)k * (rstart + (double)nv * lodr) < nyq) nv++;
                int64_t capk = (int64_t)nbk * Nprof < BOXCAR_SIGMA_SAMPLES
                                   ? (int64_t)nbk * Nprof : BOXCAR_SIGMA_SAMPLES;
                double sk = or_block_sigma(pr, n, 1, nbk, nv, capk, 0);
                if (sigmas) sigmas[c * (1 + nd) + 1 + d] = sk;
                double ik = sk > 0 ? 1.0 / sk : 0.0;
                or_boxcar_metrics(pr, n, nv, wdec[d], nwdec[d], nbk, ik, medcut, mv);
                for (int64_t j = 0; j < n; j++) {
                    double r_dec = (double)k * (rstart + (double)j * lodr);
                    if (!(r_dec < nyq)) continue;
                    if (mall) mall[(int64_t)(d + 1) * total + i0 + j] = mv[j];
                    if (mv[j] > threshold) {
                        or_cand cc = {i0 + j, k, Hk, c, mv[j]};
                        cv_push(cv, cc);
                    }
                }
            }
        }
        free(re); free(im); free(pr); free(mv);
    }
    int64_t tot = 0;
    for (int32_t t = 0; t < nt; t++) tot += parts[t].n;
    *n_out = tot;
    int st = OR_OK;
    if (tot > cap) st = OR_ENOSPC;
    else {
        int64_t o = 0;
        for (int32_t t = 0; t < nt; t++) {
            if (parts[t].n) memcpy(out + o, parts[t].v, sizeof(or_cand) * parts[t].n);
            o += parts[t].n;
        }
    }
    for (int32_t t = 0; t < nt; t++) free(parts[t].v);
    free(parts);
    or_plans_free(pl);
    return st;
}
