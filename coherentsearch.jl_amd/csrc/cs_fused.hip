// Fused search pipeline for gfx950.
//
// One workgroup owns a tile of 64 consecutive trials.  It interpolates every
// harmonic of the tile into LDS, transforms each decimation fold in registers
// (one profile per lane), and leaves only two things in HBM: the zero-baseline
// boxcar bound of every (trial, fold) before sigma scaling, and the sigma
// samples of the block's sampled columns.  A second kernel then selects each
// (block, fold)'s exact sigma, thresholds the bounds and queues survivors for
// the exact FP64 rescan (k_rescan, shared with the plain path).
#include <hip/hip_runtime.h>

#define CS_HD __host__ __device__
#include "c2r_gen.hpp"
#include "cs_device.hpp"
#include "cs_fused.hpp"

namespace cs {

constexpr int kTile = 64;

// Row guard of every (block, harmonic) of the batch (directinterp.jl:490-496).
__global__ void k_guard(SpecDev sp, TabDev tb, GridDev g, int64_t c_first, int nblocks,
                        uint8_t* rowok) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nblocks * tb.nharms) return;
    int cl = i / tb.nharms, h = i - cl * tb.nharms + 1;
    int64_t cg = c_first + cl;
    int64_t first = g.origin + cg * g.Nprof;
    int64_t last = min(first + g.Nprof, g.total) - 1;
    const HarmDev d = tb.harm[h - 1];
    bool ok;
    if (d.P == 0) ok = true;   // the per-trial fallback guards itself
    else ok = row_guard(d, tb, sp, first, last);
    rowok[i] = ok ? 1 : 0;
}

// Raw zero-baseline bound: max over the fold's bank of max_p(boxsum)/sqrt(w),
// prefix sums in the reference's sequential FP32 order.
template <int NB, int WM, int A>
__device__ __forceinline__ void raw_widths(const float* ps, uint32_t bankmask, float& res) {
    if constexpr (A < bank_count(WM)) {
        constexpr int w = bank_w(A);
        if (bankmask & (1u << A)) {
            float mb = ps[w] - ps[0];
#pragma unroll
            for (int p = 1; p < NB; ++p) {
                float d = ps[p + w] - ps[p];
                mb = d > mb ? d : mb;
            }
            float c = mb * (1.0f / sqrtf((float)w));
            res = c > res ? c : res;
        }
        raw_widths<NB, WM, A + 1>(ps, bankmask, res);
    }
}

template <int NB, int WM>
__device__ __forceinline__ void fold_tile(const float2 (*ftp)[kTile], int lane, int k,
                                          uint32_t bankmask, bool valid, int64_t j, int64_t nv,
                                          int64_t cap, float* samp, float* mraw_out) {
    constexpr int H = NB / 2;
    if (!valid) return;
    if (j >= nv) {
        *mraw_out = -INFINITY;
        return;
    }
    float xr[H + 1], xi[H + 1];
    xr[0] = 0.0f;
    xi[0] = 0.0f;
#pragma unroll
    for (int h = 1; h <= H; ++h) {
        float2 v = ftp[h * k - 1][lane];
        xr[h] = v.x;
        xi[h] = v.y;
    }
    float out[NB];
    cs_c2r::Fixed<NB>::run(xr, xi, out);
    // sigma samples (_block_sigma, search.jl:504-538): whole columns
    const int64_t NN = (int64_t)NB * nv;
    int64_t slot = -1, take = NB;
    if (NN <= cap) {
        slot = j * NB;
    } else {
        int64_t ncol = cap / NB;
        if (ncol < 1) ncol = 1;
        if (ncol > nv) ncol = nv;
        int64_t c = (j * ncol + nv - 1) / nv;
        if (c < ncol && (c * nv) / ncol == j) {
            slot = c * NB;
            take = cap - slot < NB ? cap - slot : NB;
        }
    }
    if (slot >= 0) {
#pragma unroll
        for (int i = 0; i < NB; ++i)
            if (i < take) samp[slot + i] = fabsf(out[i]);
    }
    float ps[NB + WM + 1];
    ps[0] = 0.0f;
#pragma unroll
    for (int i = 1; i <= NB + WM; ++i) ps[i] = ps[i - 1] + out[(i - 1) % NB];
    float res = -INFINITY;
    raw_widths<NB, WM, 0>(ps, bankmask, res);
    *mraw_out = res;
}

#define CS_FUSED_SIZES(X) X(16) X(20) X(24) X(30) X(32) X(40) X(60) X(64) X(120) X(128)

__host__ __device__ constexpr bool fused_size(int nb) {
    return nb == 16 || nb == 20 || nb == 24 || nb == 30 || nb == 32 || nb == 40 || nb == 60 ||
           nb == 64 || nb == 120 || nb == 128;
}

template <int MT>
__global__ __launch_bounds__(256) void k_fused(SpecDev sp, TabDev tb, GridDev g, FoldSet fs,
                                               BatchDev b, FusedDev fu) {
    __shared__ float2 ftp[64][kTile];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t jl = (int64_t)blockIdx.x * kTile + lane;     // batch column
    const bool valid = jl < b.nt;
    const int64_t t = b.t_base + jl;
    int64_t cg = 0, first = 0, last = 0;
    if (valid) block_of(g, t, cg, first, last);
    const int cl = (int)(cg - b.c_first);
    const int nh = tb.nharms;
    for (int h = wv + 1; h <= nh; h += 4) {
        double2 a = make_double2(0.0, 0.0);
        if (valid && fu.rowok[cl * nh + h - 1]) a = interp_one<MT>(sp, tb, h, t, first, last, true, true);
        ftp[h - 1][lane] = make_float2((float)a.x, (float)a.y);
    }
    __syncthreads();
    const int64_t j = t - first;
    for (int f = wv; f < fs.nf; f += 4) {
        const FoldDev& fd = fs.f[f];
        const int64_t nv = valid ? b.nvalid[cg * fs.nf + f] : 0;
        float* samp = fu.samples + ((int64_t)cl * fs.nf + f) * fu.sld;
        float* mo = fu.mraw + (int64_t)f * fu.ldm + jl;
        const bool small = fd.widths[fd.nw - 1] <= 6;
        switch (fd.nbins) {
#define X(S)                                                                                  \
    case S:                                                                                   \
        if (small && GateCT<S>::wmax >= 6)                                                    \
            fold_tile<S, 6>(ftp, lane, fd.k, fs.bankmask[f], valid, j, nv, fs.cap[f], samp, mo); \
        else                                                                                  \
            fold_tile<S, GateCT<S>::wmax>(ftp, lane, fd.k, fs.bankmask[f], valid, j, nv,       \
                                          fs.cap[f], samp, mo);                               \
        break;
            CS_FUSED_SIZES(X)
#undef X
            default:
                break;
        }
    }
}

// ---------------------------------------------------------------------------
// Exact selection of the median |x| of one (block, fold)'s samples: radix
// select on the FP32 bit patterns of the magnitudes (bits 30..19, 18..9,
// 8..0; the last pass is skipped when the second leaves a single key) with
// plain LDS histogram atomics, then for an even count the lower central value
// from the final rank and a max.  The sign bit of every sample key is zero,
// so the first digit starts below it: exponent and four mantissa bits spread
// a wave's keys over enough bins that same-address atomics stay few.  Padding
// keys (0xFFFFFFFF) fail the first pass's prefix test and are never counted.
// ---------------------------------------------------------------------------
constexpr int kST = 256;
constexpr int kSPer = 32;
constexpr int kHist = 4096;
constexpr int kHist2 = 1024;   // the second pass's histogram
static_assert(kST * kSPer == 8192, "the select's division by nbins is exact below 8192 samples");

// Inclusive scan over the 64 lanes of a wave (whole wave active) with DPP
// moves, no LDS: row_shr 1, 2, 4, 8 within each 16-lane row, then
// row_bcast 15 into rows 1 and 3 and row_bcast 31 into rows 2 and 3.  Lanes
// a move does not reach read 0, the identity of both ops used here.
template <class Op>
__device__ __forceinline__ uint32_t wave_scan(uint32_t v, Op op) {
    v = op(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false));
    v = op(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false));
    v = op(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false));
    v = op(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false));
    v = op(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false));
    v = op(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false));
    return v;
}
struct AddU32 {
    __device__ uint32_t operator()(uint32_t a, uint32_t b) const { return a + b; }
};
struct MaxU32 {
    __device__ uint32_t operator()(uint32_t a, uint32_t b) const { return a > b ? a : b; }
};

// Exclusive prefix sum over the workgroup.  One barrier: every thread adds
// the totals of the waves before its own.  `tot` (kST/64 words) is rewritten
// by the next call only after the caller's next barrier.
__device__ __forceinline__ uint32_t blk_excl_scan(uint32_t v, uint32_t* tot) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const uint32_t x = wave_scan(v, AddU32{});
    if (lane == 63) tot[wid] = x;
    __syncthreads();
    uint32_t pre = 0;
#pragma unroll
    for (int w = 0; w < kST / 64; ++w) pre += w < wid ? tot[w] : 0u;
    return pre + x - v;
}

// One radix pass: the histogram of digit (key >> SH) & (2^BITS - 1) over the
// keys whose bits above SH + BITS equal `prefix` (under `mask`); returns the
// digit that holds rank `kth` and leaves in `kth` the rank inside it.
// ZERO: clear `hist` first (one barrier); otherwise an earlier pass cleared
// it.  The NZ words at `znext` (the next pass's histogram, whose readers are
// past this call's first barrier) are cleared alongside the counting, so the
// next pass needs no clearing barrier of its own.
template <int SH, int BITS, bool ZERO, int NZ>
__device__ __forceinline__ uint32_t radix_pass(const uint32_t (&keys)[kSPer], uint32_t prefix,
                                               uint32_t mask, uint32_t& kth, uint32_t& cnt,
                                               uint32_t* hist, uint32_t* sh, uint32_t* znext) {
    constexpr int NBIN = 1 << BITS, PQ = NBIN / kST;
    static_assert(PQ >= 1 && NBIN <= kHist, "digit width");
    static_assert(NZ % kST == 0, "whole rows of the next histogram");
    constexpr uint32_t dmask = (uint32_t)NBIN - 1u;
    if constexpr (ZERO) {
#pragma unroll
        for (int q = 0; q < PQ; ++q) hist[threadIdx.x + kST * q] = 0;
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < NZ / kST; ++q) znext[threadIdx.x + kST * q] = 0;
#pragma unroll
    for (int r = 0; r < kSPer; ++r) {
        const uint32_t key = keys[r];
        if ((key & mask) == prefix) atomicAdd(&hist[(key >> SH) & dmask], 1u);
    }
    __syncthreads();
    uint32_t a[PQ], loc = 0;
#pragma unroll
    for (int q = 0; q < PQ; ++q) { a[q] = hist[PQ * threadIdx.x + q]; loc += a[q]; }
    const uint32_t ex = blk_excl_scan(loc, sh + 16);
    if (ex <= kth && kth < ex + loc) {
        uint32_t c = ex, n = 0;
        int d = 0;
#pragma unroll
        for (int q = 0; q < PQ; ++q) {
            if (kth >= c && kth < c + a[q]) { d = q; n = a[q]; break; }
            c += a[q];
        }
        sh[0] = PQ * threadIdx.x + d;
        sh[1] = kth - c;
        sh[2] = n;
    }
    __syncthreads();
    // sh[0..2] are next written after the next pass's two barriers, and hist
    // was read before this pass's scan barrier: no trailing barrier needed
    kth = sh[1];
    cnt = sh[2];
    return sh[0];
}

// The key of rank kth (0-based); kth is left as its rank among the keys equal
// to it (the last digit's bin holds exactly those keys).  The second pass
// counts into hist2 (cleared during the first); the third into hist again,
// cleared during the second: the first pass's reads of hist come before its
// scan barrier.
__device__ uint32_t select_kth(const uint32_t (&keys)[kSPer], uint32_t& kth, uint32_t* hist,
                               uint32_t* hist2, uint32_t* sh) {
    uint32_t cnt;
    const uint32_t d0 =
        radix_pass<19, 12, true, kHist2>(keys, 0u, 0x80000000u, kth, cnt, hist, sh, hist2);
    const uint32_t p1 = d0 << 19;
    const uint32_t d1 =
        radix_pass<9, 10, false, 512>(keys, p1, 0xFFF80000u, kth, cnt, hist2, sh, hist);
    const uint32_t p2 = p1 | (d1 << 9);
    if (cnt == 1) {
        // one key has the selected 23-bit prefix: it is the answer (rank 0
        // among its equals); the maximum over the keys with that prefix (the
        // others count as 0) finds it without the third pass.  cnt is
        // block-uniform, so is this branch.
        uint32_t m = 0;
#pragma unroll
        for (int r = 0; r < kSPer; ++r)
            m = (keys[r] & 0xFFFFFE00u) == p2 && keys[r] > m ? keys[r] : m;
        m = wave_scan(m, MaxU32{});
        if ((threadIdx.x & 63) == 63) sh[32 + (threadIdx.x >> 6)] = m;
        __syncthreads();
        uint32_t up = 0;
#pragma unroll
        for (int w = 0; w < kST / 64; ++w) up = sh[32 + w] > up ? sh[32 + w] : up;
        kth = 0;
        return up;
    }
    const uint32_t d2 = radix_pass<0, 9, false, 0>(keys, p2, 0xFFFFFE00u, kth, cnt, hist, sh, hist);
    return p2 | d2;
}

__global__ __launch_bounds__(kST) void k_select(FoldSet fs, GridDev g, BatchDev b, FusedDev fu,
                                                OutDev o) {
    __shared__ uint32_t hist[kHist];
    __shared__ uint32_t hist2[kHist2];
    __shared__ uint32_t sh[64];
    const int cl = blockIdx.x, f = blockIdx.y, nf = fs.nf;
    const int64_t cg = batch_block(b, cl);
    const int64_t first = g.origin + cg * g.Nprof;
    const int64_t nv = b.nvalid[cg * nf + f];
    const int nbins = fs.f[f].nbins;
    const int64_t cap = fs.cap[f];
    const int64_t NN = (int64_t)nbins * nv;
    double sigma = 0.0;
    if (NN > 0) {
        int64_t ns;
        if (NN <= cap) ns = NN;
        else {
            int64_t ncol = cap / nbins;
            if (ncol < 1) ncol = 1;
            if (ncol > nv) ncol = nv;
            ns = ncol * nbins < cap ? ncol * nbins : cap;
        }
        const float* S = fu.samples + ((int64_t)cl * nf + f) * fu.sld;
        // sample s is bin s % nbins of sampled column s / nbins; columns start at
        // multiples of nbins rounded up to salign
        const int nbs = (nbins + fu.salign - 1) / fu.salign * fu.salign;
        // s / nbins as (s * M) >> 20, M = ceil(2^20 / nbins): exact for every
        // s < 8192 and nbins <= 128 (checked exhaustively)
        const uint32_t M = ((1u << 20) + (uint32_t)nbins - 1u) / (uint32_t)nbins;
        uint32_t keys[kSPer];
        const int slast = (int)ns - 1;
#pragma unroll
        for (int r = 0; r < kSPer; ++r) {
            const int s = threadIdx.x + kST * r;
            // unconditional load (past the samples: the last one's address)
            const int sc = s < slast ? s : slast;
            int pos = sc;
            if (nbs != nbins) {
                const int cc = (int)(((uint32_t)sc * M) >> 20);
                pos = sc + cc * (nbs - nbins);
            }
            // the sample's magnitude (the split-BF16 kernel stores signed values)
            const uint32_t v = __float_as_uint(S[pos]) & 0x7FFFFFFFu;
            keys[r] = s < ns ? v : 0xFFFFFFFFu;
        }
        const uint32_t half = (uint32_t)(ns / 2);
        uint32_t rk = half;   // becomes the rank of the selected key among its equals
        const uint32_t up = select_kth(keys, rk, hist, hist2, sh);
        double med;
        if ((ns & 1) || rk > 0) {
            // odd count; or the key of rank half - 1 equals `up` as well
            med = (double)__uint_as_float(up);
        } else {
            // exactly `half` keys lie below `up`: the lower central value is
            // the largest of them (rk is block-uniform, so is this branch)
            uint32_t mx = 0;
#pragma unroll
            for (int r = 0; r < kSPer; ++r) {
                const uint32_t key = keys[r];
                mx = key < up && key > mx ? key : mx;
            }
            mx = wave_scan(mx, MaxU32{});   // the wave's maximum in lane 63
            const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
            if (lane == 63) hist[16 + wid] = mx;
            __syncthreads();
            uint32_t tm = 0;
            for (int w = 0; w < kST / 64; ++w) tm = hist[16 + w] > tm ? hist[16 + w] : tm;
            med = 0.5 * ((double)__uint_as_float(tm) + (double)__uint_as_float(up));
        }
        sigma = 1.4826 * med;
    }
    if (threadIdx.x == 0) b.sigma[cl * nf + f] = sigma;
    if (!fu.mraw) return;   // sigma only (the plain pipeline gates from its profiles)
    // threshold the block's bounds: m0 = bound * invsigma against medcut
    const double invs = sigma > 0 ? 1.0 / sigma : 0.0;
    const bool all = !(o.medcut > -INFINITY) || invs <= 0;
    // bound columns are [local block][trial of the block] (for a contiguous
    // batch: the trial minus the batch's first trial)
    const int64_t col0 = (int64_t)cl * g.Nprof;
    for (int64_t j0 = 0; j0 < nv; j0 += kST) {
        int64_t j = j0 + threadIdx.x;
        bool pred = false;
        if (j < nv) {
            if (all) pred = true;
            else {
                float m = fu.mraw[(int64_t)f * fu.ldm + col0 + j];
                pred = (double)m * invs >= o.medcut;
            }
        }
        if (o.gm) pred = j < nv && o.gm[(int64_t)f * o.gm_ld + first + j] > o.gmcut;
        push_survivor(pred, first + j, f, o);
    }
}

hipError_t launch_guard(hipStream_t st, const SpecDev& sp, const TabDev& tb, const GridDev& g,
                        const BatchDev& b, uint8_t* rowok) {
    int n = b.nblocks * tb.nharms;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_guard, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, sp, tb, g,
                       b.c_first, b.nblocks, rowok);
    return hipGetLastError();
}

bool fused_supported(const FoldSet& fs, int nharms) {
    if (nharms > 64) return false;
    for (int f = 0; f < fs.nf; ++f) {
        if (!fused_size(fs.f[f].nbins) || fs.bankmask[f] == 0) return false;
    }
    return true;
}

hipError_t launch_fused(hipStream_t st, const SpecDev& sp, const TabDev& tb, const GridDev& g,
                        const FoldSet& fs, const BatchDev& b, const FusedDev& fu) {
    if (b.nt <= 0) return hipSuccess;
    dim3 grid((unsigned)((b.nt + kTile - 1) / kTile));
    if (tb.m == 16 && !tb.weights_f64)
        hipLaunchKernelGGL(k_fused<16>, grid, dim3(256), 0, st, sp, tb, g, fs, b, fu);
    else
        hipLaunchKernelGGL(k_fused<0>, grid, dim3(256), 0, st, sp, tb, g, fs, b, fu);
    return hipGetLastError();
}

hipError_t launch_select(hipStream_t st, const FoldSet& fs, const GridDev& g, const BatchDev& b,
                         const FusedDev& fu, const OutDev& o) {
    if (b.nblocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_select, dim3((unsigned)b.nblocks, (unsigned)fs.nf), dim3(kST), 0, st, fs,
                       g, b, fu, o);
    return hipGetLastError();
}

}  // namespace cs
