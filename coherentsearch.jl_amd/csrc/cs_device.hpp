// Device helpers shared by the plain and fused kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "cs_geo.hpp"
#include "cs_gpu.hpp"

namespace cs {

// ---------------------------------------------------------------------------
// Exact trial-grid state: T = t*hp, (qint, res) = (T div q, T mod q), for
// 0 <= t < 2^32 and q*hp < 2^32 (the host checks both).  Two 32-bit divisions
// instead of a 64-bit one (direct_chunk_state, directinterp.jl:414-418).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void cstate(int64_t t, uint32_t hp, uint32_t q, int64_t& qint,
                                       uint32_t& res) {
    uint32_t tt = (uint32_t)t;
    uint32_t tq = tt / q;
    uint32_t tr = tt - tq * q;
    uint32_t u = tr * hp;
    uint32_t uq = u / q;
    res = u - uq * q;
    qint = (int64_t)tq * hp + uq;
}

// 0-based first bin of trial t for harmonic plan d: floor(h r_t) + 1 - m/2.
__device__ __forceinline__ int64_t bin0(const HarmDev& d, const TabDev& tb, int64_t t,
                                        uint32_t& p) {
    int64_t qint;
    uint32_t res;
    cstate(t, (uint32_t)d.hp, (uint32_t)d.q, qint, res);
    p = (uint32_t)tb.row[d.row_off + res];
    return d.rfloor0 + qint + tb.carry[d.tab_off + p] + 1 - tb.m / 2;
}

// Row guard on the block's TRUE trial range (directinterp.jl:490-496): the
// row is kept for the whole block or zeroed for the whole block.
__device__ __forceinline__ bool row_guard(const HarmDev& d, const TabDev& tb,
                                          const SpecDev& sp, int64_t first,
                                          int64_t last) {
    uint32_t p0, pl;
    int64_t b_first = bin0(d, tb, first, p0);
    int64_t qint;
    uint32_t res;
    cstate(last, (uint32_t)d.hp, (uint32_t)d.q, qint, res);
    pl = (uint32_t)tb.row[d.row_off + res];
    int64_t b_last = d.rfloor0 + qint + tb.carry[d.tab_off + pl] + 1 - tb.m / 2;
    return b_first >= 0 && b_last + tb.m <= sp.namps && (d.rfloor0 + qint) < sp.Nhalf;
}

// Per-trial FP64 evaluation for a trial step that is not a small rational
// (_fill_row_direct_slow!, directinterp.jl:623-661).
__device__ inline double2 interp_slow(const HarmDev& d, const TabDev& tb, const SpecDev& sp,
                               int h, int64_t t, int64_t first, int64_t last) {
    const int m = tb.m, m2 = m / 2;
    double hr_lo = (double)d.rfloor0 + d.dr0;
    double dh = (double)h * tb.lodr;
    double rfirst = hr_lo + (double)first * dh;
    double rlast = hr_lo + (double)last * dh;
    int64_t lo = (int64_t)floor(rfirst + 1e-15) + 2 - m2;
    int64_t hi = (int64_t)floor(rlast + 1e-15) + 1 + m2;
    if (!(lo >= 1 && hi <= sp.namps && (int64_t)floor(rlast) < sp.Nhalf))
        return make_double2(0.0, 0.0);
    double r = hr_lo + (double)t * dh;
    int64_t rint = (int64_t)floor(r + 1e-15) + 1;
    double dr = r - floor(r);
    int64_t base = rint - m2;
    if (dr == 0.0) {
        float2 a = sp.amps[base + m2 - 1];
        return make_double2(a.x, a.y);
    }
    double sre = 0.0, sim = 0.0;
    for (int i = 1; i <= m; ++i) {
        double w = 1.0 / (dr - (double)(i - m2));
        float2 a = sp.amps[base + i - 1];
        sre = fma(w, (double)a.x, sre);
        sim = fma(w, (double)a.y, sim);
    }
    double s = sinpi(dr), c = cospi(dr);
    double ar = (s * c) / M_PI, ai = -((s * s) / M_PI);
    return make_double2(ar * sre - ai * sim, ar * sim + ai * sre);
}

// One interpolated amplitude (harmonic h, global trial t of block [first, last]).
// FP32 weights: an m-term FP32 FMA chain, then the complex FP32 product with
// conj(A) evaluated without contraction, exactly as the reference stores
// `gA[k, g] * complex(ar[k], ai[k])`.
// interp_with takes harmonic h's plan d from the caller (who may hold it in
// LDS); interp_one reads it from the plan table.
template <int MT>
__device__ __forceinline__ double2 interp_with(const HarmDev& d, const SpecDev& sp,
                                               const TabDev& tb, int h, int64_t t,
                                               int64_t first, int64_t last, bool guard_known,
                                               bool guard_ok) {
    if (d.P == 0) return interp_slow(d, tb, sp, h, t, first, last);
    bool ok = guard_known ? guard_ok : row_guard(d, tb, sp, first, last);
    if (!ok) return make_double2(0.0, 0.0);
    uint32_t p;
    int64_t b = bin0(d, tb, t, p);
    const int m = MT > 0 ? MT : tb.m;
    const float2* x = sp.amps + b;
    if (!tb.weights_f64) {
        const float* w = tb.W32 + (size_t)(d.tab_off + p) * m;
        float ar = 0.0f, ai = 0.0f;
#pragma unroll
        for (int i = 0; i < (MT > 0 ? MT : 1); ++i) {
            if (MT > 0) {
                float2 v = x[i];
                ar = __builtin_fmaf(w[i], v.x, ar);
                ai = __builtin_fmaf(w[i], v.y, ai);
            }
        }
        if (MT == 0) {
            for (int i = 0; i < m; ++i) {
                float2 v = x[i];
                ar = __builtin_fmaf(w[i], v.x, ar);
                ai = __builtin_fmaf(w[i], v.y, ai);
            }
        }
        float2 A = tb.A32[d.tab_off + p];
        float p1 = A.x * ar, p2 = A.y * ai, p3 = A.x * ai, p4 = A.y * ar;
        return make_double2((double)(p1 - p2), (double)(p3 + p4));
    } else {
        const double* w = tb.W64 + (size_t)(d.tab_off + p) * m;
        double ar = 0.0, ai = 0.0;
        for (int i = 0; i < m; ++i) {
            float2 v = x[i];
            ar = fma(w[i], (double)v.x, ar);
            ai = fma(w[i], (double)v.y, ai);
        }
        double2 A = tb.A64[d.tab_off + p];
        return make_double2(A.x * ar - A.y * ai, A.x * ai + A.y * ar);
    }
}
template <int MT>
__device__ __forceinline__ double2 interp_one(const SpecDev& sp, const TabDev& tb, int h,
                                              int64_t t, int64_t first, int64_t last,
                                              bool guard_known, bool guard_ok) {
    const HarmDev d = tb.harm[h - 1];
    return interp_with<MT>(d, sp, tb, h, t, first, last, guard_known, guard_ok);
}

// Sample slot column of block trial j (_block_sigma's whole-column subsample,
// search.jl:504-538): j itself when all nv columns fit in cap samples, else
// c when j = floor(c * nv / ncol) for some c < ncol = clamp(cap / nbins, 1, nv);
// -1 if the column is not sampled.
__host__ __device__ inline int sample_column(int64_t j, int64_t nv, int nbins, int64_t cap) {
    if ((int64_t)nbins * nv <= cap) return (int)j;
    int64_t ncol = cap / nbins;
    if (ncol < 1) ncol = 1;
    if (ncol > nv) ncol = nv;
    const int64_t c = (j * ncol + nv - 1) / nv;
    return (c < ncol && (c * nv) / ncol == j) ? (int)c : -1;
}
// The same in 32-bit arithmetic, exact while nv * cap / nbins < 2^32 (block
// sizes below 2^19 with cap 8192; the split-BF16 path checks this on the host).
__host__ __device__ inline int sample_column32(uint32_t j, uint32_t nv, uint32_t nbins,
                                               uint32_t cap) {
    if (nbins * nv <= cap) return (int)j;
    uint32_t ncol = cap / nbins;
    if (ncol < 1) ncol = 1;
    if (ncol > nv) ncol = nv;
    const uint32_t c = (j * ncol + nv - 1) / nv;
    return (c < ncol && (c * nv) / ncol == j) ? (int)c : -1;
}

__device__ __forceinline__ void block_of(const GridDev& g, int64_t t, int64_t& cg,
                                         int64_t& first, int64_t& last) {
    cg = (t - g.origin) / g.Nprof;
    first = g.origin + cg * g.Nprof;
    last = min(first + g.Nprof, g.total) - 1;
}

template <int NB>
struct GateCT {
    static constexpr int lim = (3 * NB) / 10 < 1 ? 1 : (3 * NB) / 10;
    static constexpr int nbank = bank_count(lim);
    static constexpr int wmax = bank_w(nbank - 1);
};

__device__ __forceinline__ void push_survivor(bool pred, int64_t t, int fidx, const OutDev& o) {
    const int lane = threadIdx.x & 63;
    unsigned long long bal = __ballot(pred);
    if (bal == 0) return;
    int leader = __ffsll((long long)bal) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(o.surv_count, (uint32_t)__popcll(bal));
    base = (uint32_t)__shfl((int)base, leader, 64);
    if (pred) {
        uint32_t idx = base + (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
        if (idx < o.surv_cap) o.surv[idx] = ((uint64_t)t << 4) | (uint64_t)fidx;
    }
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        double u = __shfl_xor(v, o, 64);
        v = u > v ? u : v;
    }
    return v;
}


}  // namespace cs
