// Metric statistics of a search over every trial (--metricstats,
// search.jl:1295-1549 and the collecting branch of _search_region!,
// search.jl:1768-1894, decim_pass!, search.jl:1664-1719) and the normalised
// detection pass (--normalize, search.jl:1565-1643, 2081-2095).
//
// The exact metric of every valid (fold, trial) is resident in HBM
// ([fold][trial] doubles, written by the exact rescan).  One workgroup per
// (block, fold) then forms the block's moments, extremes and exact median
// (radix select on order-preserving 64-bit keys) and its histogram in LDS,
// whose nonzero bins are added to the (fold, window) histogram in HBM.  The
// normalised pass reads the same resident metrics.
#include <hip/hip_runtime.h>

#include "cs_device.hpp"
#include "cs_stats.hpp"

namespace cs {

namespace {

constexpr int kThreads = 256;

// order-preserving map of a double to an unsigned key, and back
__device__ __forceinline__ uint64_t okey(double x) {
    const uint64_t u = (uint64_t)__double_as_longlong(x);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}
__device__ __forceinline__ double okey_inv(uint64_t k) {
    const uint64_t u = (k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k;
    return __longlong_as_double((long long)u);
}

// workgroup reductions in a fixed order (wave shuffles, then the 4 waves)
__device__ double wg_sum(double v, double* red) {
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    const int wv = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[wv] = v;
    __syncthreads();
    return (red[0] + red[1]) + (red[2] + red[3]);
}
__device__ double wg_min(double v, double* red) {
    for (int o = 32; o >= 1; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    const int wv = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[wv] = v;
    __syncthreads();
    return fmin(fmin(red[0], red[1]), fmin(red[2], red[3]));
}
__device__ double wg_max(double v, double* red) {
    for (int o = 32; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    const int wv = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[wv] = v;
    __syncthreads();
    return fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

// the rank-th smallest (0-based) of v[0, n): eight passes of 8 key bits
__device__ double select_rank(const double* v, int64_t n, int64_t rank, uint32_t* hist,
                              uint64_t* shared_prefix) {
    uint64_t prefix = 0, mask = 0;
    for (int pass = 0; pass < 8; ++pass) {
        const int sh = 56 - 8 * pass;
        for (int i = threadIdx.x; i < 256; i += kThreads) hist[i] = 0;
        __syncthreads();
        for (int64_t i = threadIdx.x; i < n; i += kThreads) {
            const uint64_t key = okey(v[i]);
            if ((key & mask) == prefix) atomicAdd(&hist[(key >> sh) & 255u], 1u);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int64_t r = rank;
            int d = 0;
            for (; d < 255; ++d) {
                if (r < (int64_t)hist[d]) break;
                r -= hist[d];
            }
            shared_prefix[0] = prefix | ((uint64_t)d << sh);
            shared_prefix[1] = (uint64_t)r;
        }
        __syncthreads();
        prefix = shared_prefix[0];
        rank = (int64_t)shared_prefix[1];
        mask |= (uint64_t)255 << sh;
        __syncthreads();
    }
    return okey_inv(prefix);
}

__global__ __launch_bounds__(kThreads) void k_block_stats(const double* __restrict__ metrics,
                                                          int64_t ld,
                                                          const int32_t* __restrict__ nvalid,
                                                          int nf, int64_t Nprof,
                                                          int64_t nblocks,
                                                          const int32_t* __restrict__ win,
                                                          int nwin, double lo, double hi, int nb,
                                                          uint32_t* __restrict__ counts,
                                                          BlockStatDev* __restrict__ out) {
    extern __shared__ uint32_t hcount[];   // nb bins
    __shared__ double red[4];
    __shared__ uint32_t sel[256];
    __shared__ uint64_t selp[2];
    const int64_t blk = blockIdx.x;
    const int f = blockIdx.y;
    const int64_t n = nvalid[blk * nf + f];
    const double* v = metrics + (int64_t)f * ld + blk * Nprof;
    for (int i = threadIdx.x; i < nb; i += kThreads) hcount[i] = 0;
    __syncthreads();
    const double invw = (double)nb / (hi - lo);
    double s = 0.0, s2 = 0.0, mn = INFINITY, mx = -INFINITY;
    uint32_t under = 0, over = 0;
    for (int64_t i = threadIdx.x; i < n; i += kThreads) {
        const double x = v[i];
        s += x;
        s2 += x * x;
        mn = fmin(mn, x);
        mx = fmax(mx, x);
        if (x < lo) under++;
        else if (x >= hi) over++;
        else {
            int64_t b = (int64_t)floor((x - lo) * invw);
            b = b < nb ? b : nb - 1;
            atomicAdd(&hcount[b], 1u);
        }
    }
    const double sum = wg_sum(s, red);
    const double sumsq = wg_sum(s2, red);
    const double vmin = wg_min(mn, red), vmax = wg_max(mx, red);
    const double nu = wg_sum((double)under, red), no = wg_sum((double)over, red);
    const double mean = n > 0 ? sum / (double)n : NAN;
    double ss = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += kThreads) {
        const double d = v[i] - mean;
        ss += d * d;
    }
    const double ssum = wg_sum(ss, red);
    double med = NAN;
    if (n > 0) {
        const int64_t half = n >> 1;
        const double upper = select_rank(v, n, half, sel, selp);
        if (n & 1) med = upper;
        else med = 0.5 * (select_rank(v, n, half - 1, sel, selp) + upper);
    }
    // the block's histogram into its (fold, window) histogram
    if (n > 0) {
        uint32_t* dst = counts + ((int64_t)f * nwin + win[(int64_t)f * nblocks + blk]) * nb;
        for (int i = threadIdx.x; i < nb; i += kThreads)
            if (hcount[i]) atomicAdd(&dst[i], hcount[i]);
    }
    if (threadIdx.x == 0) {
        BlockStatDev r;
        r.n = n;
        r.under = (int64_t)nu;
        r.over = (int64_t)no;
        r.pad = 0;
        r.vmin = n > 0 ? vmin : NAN;
        r.vmax = n > 0 ? vmax : NAN;
        r.sum = sum;
        r.sumsq = sumsq;
        r.mean = mean;
        r.median = med;
        r.stdv = n > 1 ? sqrt(ssum / (double)(n - 1)) : (n == 1 ? 0.0 : NAN);
        out[(int64_t)f * nblocks + blk] = r;
    }
}

__global__ __launch_bounds__(256) void k_normalize(NormDev nd, cs_cand_dev* __restrict__ cands,
                                                   uint32_t* __restrict__ count, uint32_t cap) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int f = blockIdx.y;
    if (t >= nd.total) return;
    const double M = nd.metrics[(int64_t)f * nd.ld + t];
    if (M != M) return;   // not a valid trial of this fold
    const int64_t blk = t / nd.Nprof, i0 = blk * nd.Nprof, j = t - i0;
    const int k = nd.k[f];
    const double rstart = nd.r_lo + (double)i0 * nd.lodr;
    const double rf = rstart + (double)j * nd.lodr;
    const double r = k == 1 ? rf : (double)k * rf;
    const double fq = r / nd.T;
    const double* e = nd.edges + (int64_t)f * (nd.nwin + 1);
    int le = 0;   // edges <= fq
    while (le <= nd.nwin && e[le] <= fq) ++le;
    int w = le - 1;
    w = w < 0 ? 0 : (w > nd.nwin - 1 ? nd.nwin - 1 : w);
    const double z = (M - nd.loc[f * nd.nwin + w]) / nd.scale[f * nd.nwin + w];
    if (z > nd.threshold) {
        const uint32_t idx = atomicAdd(count, 1u);
        if (idx < cap) {
            cs_cand_dev c;
            c.t = t;
            c.k = k;
            c.nharm = nd.Hk[f];
            c.block = blk;
            c.metric = z;
            cands[idx] = c;
        }
    }
}

constexpr int pow2ceil(int n) {
    int p = 1;
    while (p < n) p <<= 1;
    return p;
}

// per bank width (ascending, compile-time index A) the largest prefix-sum
// difference times invsigma / sqrt(w); widths outside the fold's bank skipped
template <int NB, int WX, int A>
__device__ __forceinline__ void exact_widths(const float* ps, uint32_t bankmask, float invs,
                                             float& best) {
    if constexpr (A < bank_count(WX)) {
        constexpr int w = bank_w(A);
        if (bankmask & (1u << A)) {
            float m = ps[w] - ps[0];
#pragma unroll
            for (int p = 1; p < NB; ++p) {
                const float d = ps[p + w] - ps[p];
                m = d > m ? d : m;
            }
            const float cand = m * (invs / __fsqrt_rn((float)w));
            if (cand > best) best = cand;
        }
        exact_widths<NB, WX, A + 1>(ps, bankmask, invs, best);
    }
}

// Exact metric of every valid trial of fold fidx from the batch's FP32
// profiles (prof[bin][trial]), in the search's element type as the reference
// scores it (_boxcar_exact / _boxcar_psum! / _boxcar_scan, search.jl:571-596,
// 893-904): the exact median (a bitonic network over the column padded with
// +inf), the circular prefix sums in order, per ascending width the largest
// difference times invsigma / sqrt(w), strict >.
template <int NB>
__global__ __launch_bounds__(256) void k_exact_all(const float* __restrict__ prof, int64_t ldp,
                                                   uint32_t bankmask, int fidx, int nf, GridDev g,
                                                   BatchDev b, double* __restrict__ mall,
                                                   int64_t ld) {
    constexpr int NP = pow2ceil(NB);
    constexpr int WX = gate_wmax(NB);
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= b.nt) return;
    const int64_t t = b.t_base + j;
    int64_t cg, first, last;
    block_of(g, t, cg, first, last);
    if (t - first >= b.nvalid[cg * nf + fidx]) return;
    const double sig = b.sigma[(cg - b.c_first) * nf + fidx];
    const float invs = sig > 0 ? 1.0f / (float)sig : 0.0f;
    double metric = 0.0;
    if (invs > 0.0f) {
        float med;
        if constexpr (NP > 64) {
            // lower and upper middle order statistics by bisection on
            // order-preserving keys (a 128-entry network would not fit in
            // registers): upper = smallest key with #(keys <= it) > NB / 2
            uint32_t kv[NB];
#pragma unroll
            for (int i = 0; i < NB; ++i) {
                const uint32_t u = __float_as_uint(prof[(int64_t)i * ldp + j]);
                kv[i] = (u >> 31) ? ~u : (u | 0x80000000u);
            }
            uint32_t lo = 0u, hi = 0xFFFFFFFFu;
            while (lo < hi) {
                const uint32_t mid = lo + (hi - lo) / 2;
                int c = 0;
#pragma unroll
                for (int i = 0; i < NB; ++i) c += kv[i] <= mid ? 1 : 0;
                if (c > NB / 2) hi = mid;
                else lo = mid + 1;
            }
            const uint32_t up = lo;
            auto fl = [](uint32_t k) {
                return __uint_as_float((k >> 31) ? (k & 0x7FFFFFFFu) : ~k);
            };
            if (NB & 1) {
                med = fl(up);
            } else {
                // lower middle: `up` if at least NB/2 keys lie at or below it
                // besides one copy of itself, else the largest key below it
                int below = 0;
                uint32_t mx = 0u;
#pragma unroll
                for (int i = 0; i < NB; ++i) {
                    const bool lt = kv[i] < up;
                    below += lt ? 1 : 0;
                    mx = lt && kv[i] > mx ? kv[i] : mx;
                }
                med = 0.5f * ((below >= NB / 2 ? fl(mx) : fl(up)) + fl(up));
            }
        } else {
            float v[NP];
#pragma unroll
            for (int i = 0; i < NP; ++i) v[i] = i < NB ? prof[(int64_t)i * ldp + j] : INFINITY;
#pragma unroll
            for (int k = 2; k <= NP; k <<= 1)
#pragma unroll
                for (int q = k >> 1; q > 0; q >>= 1)
#pragma unroll
                    for (int i = 0; i < NP; ++i) {
                        const int l = i ^ q;
                        if (l > i) {
                            const float a = v[i], c = v[l];
                            const bool up = (i & k) == 0;
                            v[i] = up ? fminf(a, c) : fmaxf(a, c);
                            v[l] = up ? fmaxf(a, c) : fminf(a, c);
                        }
                    }
            med = (NB & 1) ? v[NB / 2] : 0.5f * (v[NB / 2 - 1] + v[NB / 2]);
        }
        float ps[NB + WX + 1];
        ps[0] = 0.0f;
#pragma unroll
        for (int i = 1; i <= NB + WX; ++i) {
            const int idx = i > NB ? i - NB : i;
            ps[i] = ps[i - 1] + (prof[(int64_t)(idx - 1) * ldp + j] - med);
        }
        float best = -INFINITY;
        exact_widths<NB, WX, 0>(ps, bankmask, invs, best);
        metric = (double)best;
    }
    mall[(int64_t)fidx * ld + t] = metric;
}

// sigma of a profile set (_block_sigma, search.jl:490-566) for any sample
// count: the samples (every bin, or whole columns spread evenly) gathered to
// buf, then 1.4826 x the exact median of |x| (center 0) or of |x - median(x)|
// (center 1), medians by radix selection over buf.
__global__ __launch_bounds__(kThreads) void k_sigma_any(const double* __restrict__ prof,
                                                        int64_t ldp, int64_t ldt, int nbins,
                                                        int64_t n,
                                                        int64_t cap, int center_median,
                                                        double* __restrict__ buf,
                                                        double* __restrict__ sigma_out) {
    __shared__ uint32_t sel[256];
    __shared__ uint64_t selp[2];
    const int64_t NN = (int64_t)nbins * n;
    if (NN == 0) {
        if (threadIdx.x == 0) *sigma_out = 0.0;
        return;
    }
    const bool all = NN <= cap;
    int64_t ncol, ns;
    if (all) {
        ncol = n;
        ns = NN;
    } else {
        ncol = cap / nbins;
        if (ncol < 1) ncol = 1;
        if (ncol > n) ncol = n;
        ns = ncol * nbins < cap ? ncol * nbins : cap;
    }
    for (int64_t s = threadIdx.x; s < ns; s += kThreads) {
        const int64_t cc = s / nbins, i = s - cc * nbins;
        const int64_t j = all ? cc : (cc * n) / ncol;
        const double x = prof[i * ldp + j * ldt];
        buf[s] = center_median ? x : fabs(x);
    }
    __threadfence_block();
    __syncthreads();
    auto median = [&]() {
        const int64_t half = ns >> 1;
        const double up = select_rank(buf, ns, half, sel, selp);
        return (ns & 1) ? up : 0.5 * (select_rank(buf, ns, half - 1, sel, selp) + up);
    };
    if (center_median) {
        const double m0 = median();
        for (int64_t s = threadIdx.x; s < ns; s += kThreads) buf[s] = fabs(buf[s] - m0);
        __threadfence_block();
        __syncthreads();
    }
    const double med = median();
    if (threadIdx.x == 0) *sigma_out = 1.4826 * med;
}

}  // namespace

hipError_t launch_sigma_any(hipStream_t st, const double* prof, int64_t ldp, int nbins, int64_t n,
                            int64_t cap, int center_median, double* buf, double* sigma_out,
                            int64_t ldt) {
    hipLaunchKernelGGL(k_sigma_any, dim3(1), dim3(kThreads), 0, st, prof, ldp, ldt, nbins, n, cap,
                       center_median, buf, sigma_out);
    return hipGetLastError();
}

bool exact_all_supported(int nbins) {
    switch (nbins) {
        case 16: case 20: case 24: case 30: case 32: case 40: case 60: case 64: case 120:
        case 128:
            return true;
        default:
            return false;
    }
}

hipError_t launch_exact_all(hipStream_t st, const float* prof, int64_t ldp, int nbins,
                            uint32_t bankmask, int fidx, int nf, const GridDev& g,
                            const BatchDev& b, double* mall, int64_t ld) {
    if (b.nt <= 0) return hipSuccess;
    const dim3 grid((unsigned)((b.nt + 255) / 256));
    switch (nbins) {
#define X(S)                                                                                  \
    case S:                                                                                   \
        hipLaunchKernelGGL(k_exact_all<S>, grid, dim3(256), 0, st, prof, ldp, bankmask, fidx, \
                           nf, g, b, mall, ld);                                               \
        break;
        X(16) X(20) X(24) X(30) X(32) X(40) X(60) X(64) X(120) X(128)
#undef X
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_block_stats(hipStream_t st, const double* metrics, int64_t ld,
                              const int32_t* nvalid, int nf, int64_t Nprof, int64_t nblocks,
                              const int32_t* win, int nwin, double lo, double hi, int nb,
                              uint32_t* counts, BlockStatDev* out) {
    if (nblocks <= 0 || nf <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_block_stats, dim3((unsigned)nblocks, (unsigned)nf), dim3(kThreads),
                       (size_t)nb * sizeof(uint32_t), st, metrics, ld, nvalid, nf, Nprof, nblocks,
                       win, nwin, lo, hi, nb, counts, out);
    return hipGetLastError();
}

hipError_t launch_normalize(hipStream_t st, const NormDev& nd, cs_cand_dev* cands,
                            uint32_t* count, uint32_t cap) {
    if (nd.total <= 0 || nd.nf <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_normalize, dim3((unsigned)((nd.total + 255) / 256), (unsigned)nd.nf),
                       dim3(256), 0, st, nd, cands, count, cap);
    return hipGetLastError();
}

}  // namespace cs
