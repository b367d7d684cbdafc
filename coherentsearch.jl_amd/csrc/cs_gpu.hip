// Coherent-search kernels for gfx950 (MI355X).
//
// Built with -ffp-contract=off: every fused multiply-add below is an explicit
// __builtin_fmaf / fma, placed where the reference's `muladd` sits, so the FP32
// interpolation reproduces the reference's Float32-weight arithmetic bit for
// bit (src/directinterp.jl:584-605).
#include <hip/hip_runtime.h>

#define CS_HD __host__ __device__
#include "c2r_gen.hpp"
#include "cs_gpu.hpp"
#include "cs_device.hpp"

namespace cs {

// K1: plain FP32 (or FP64-weight bridge) interpolation, one thread per
// (trial, harmonic); output row h-1 of the stack, columns = batch trials.
template <int MT, bool OUT64>
__global__ __launch_bounds__(256) void k_interp(SpecDev sp, TabDev tb, GridDev g,
                                                int64_t t_base, int64_t nt, void* out,
                                                int64_t ld, const int32_t* __restrict__ sel,
                                                int32_t slot0) {
    const int h = blockIdx.y + 1;
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nt) return;
    int64_t t = t_base + j;
    if (sel) {
        const int64_t cl = j / g.Nprof;
        t = g.origin + (int64_t)sel[slot0 + cl] * g.Nprof + (j - cl * g.Nprof);
    }
    double2 v = make_double2(0.0, 0.0);
    if (t < g.total) {
        int64_t cg, first, last;
        block_of(g, t, cg, first, last);
        v = interp_one<MT>(sp, tb, h, t, first, last, false, false);
    }
    if (OUT64) {
        ((double2*)out)[(int64_t)(h - 1) * ld + j] = v;
    } else {
        ((float2*)out)[(int64_t)(h - 1) * ld + j] = make_float2((float)v.x, (float)v.y);
    }
}

hipError_t launch_interp(hipStream_t st, const SpecDev& sp, const TabDev& tb,
                         const GridDev& g, int64_t t_base, int64_t nt, bool out64,
                         void* out, int64_t ld, const int32_t* sel, int32_t slot0) {
    if (nt <= 0) return hipSuccess;
    dim3 grid((unsigned)((nt + 255) / 256), (unsigned)tb.nharms);
    bool f32 = !tb.weights_f64;
#define CS_INTERP(MT, O64) \
    hipLaunchKernelGGL((k_interp<MT, O64>), grid, dim3(256), 0, st, sp, tb, g, t_base, nt, out, ld, sel, slot0)
    if (out64) {
        if (f32 && tb.m == 16) CS_INTERP(16, true);
        else if (f32 && tb.m == 32) CS_INTERP(32, true);
        else CS_INTERP(0, true);
    } else {
        if (f32 && tb.m == 16) CS_INTERP(16, false);
        else if (f32 && tb.m == 32) CS_INTERP(32, false);
        else CS_INTERP(0, false);
    }
#undef CS_INTERP
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// K2: c2r of one fold, one profile per thread.  Rows h*k (h = 1..H) of the
// stack; DC (row 0) is zero by construction; the Nyquist imaginary part is
// ignored (numpy convention).  Generated straight-line FFTs for the common
// lengths, an O(H^2) fallback otherwise.
// ---------------------------------------------------------------------------
template <int NB>
__global__ __launch_bounds__(256) void k_c2r(const float2* __restrict__ ftp, int64_t ldf,
                                             int k, int64_t nt, float* __restrict__ prof,
                                             int64_t ldp) {
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nt) return;
    constexpr int H = NB / 2;
    float xr[H + 1], xi[H + 1];
    xr[0] = 0.0f;
    xi[0] = 0.0f;
#pragma unroll
    for (int h = 1; h <= H; ++h) {
        float2 v = ftp[(int64_t)(h * k - 1) * ldf + j];
        xr[h] = v.x;
        xi[h] = v.y;
    }
    float out[NB];
    cs_c2r::Fixed<NB>::run(xr, xi, out);
#pragma unroll
    for (int i = 0; i < NB; ++i) prof[(int64_t)i * ldp + j] = out[i];
}

__global__ __launch_bounds__(256) void k_c2r_generic(const float2* __restrict__ ftp,
                                                     int64_t ldf, int k, int NB, int64_t nt,
                                                     float* __restrict__ prof, int64_t ldp,
                                                     const double2* __restrict__ twid) {
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nt) return;
    const int H = NB / 2;
    for (int n = 0; n <= H; ++n) {
        float E = 0.0f, O = 0.0f;
        for (int h = 1; h <= H; ++h) {
            float2 v = ftp[(int64_t)(h * k - 1) * ldf + j];
            float c, s;
            if (h == H) { c = (n & 1) ? -1.0f : 1.0f; s = 0.0f; }
            else {
                double2 w = twid[(h * n) % NB];
                c = (float)(2.0 * w.x);
                s = (float)(-2.0 * w.y);
            }
            E = __builtin_fmaf(c, v.x, E);
            O = __builtin_fmaf(s, v.y, O);
        }
        if (n == 0 || n == H) prof[(int64_t)n * ldp + j] = E;
        else {
            prof[(int64_t)n * ldp + j] = E + O;
            prof[(int64_t)(NB - n) * ldp + j] = E - O;
        }
    }
}

// The same c2r for the search's batches, which also writes |x| of every
// sampled column (_block_sigma's whole-column subsample, search.jl:504-538) to
// its block's sample region for the median select: column slot sc of (local
// block cl, fold f) at samples[(cl * nf + f) * sld + sc * NB].
template <int NB>
__global__ __launch_bounds__(256) void k_c2r_samples(const float2* __restrict__ ftp, int64_t ldf,
                                                     int k, int64_t nt, float* __restrict__ prof,
                                                     int64_t ldp, SampleSink sk) {
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nt) return;
    constexpr int H = NB / 2;
    float xr[H + 1], xi[H + 1];
    xr[0] = 0.0f;
    xi[0] = 0.0f;
#pragma unroll
    for (int h = 1; h <= H; ++h) {
        float2 v = ftp[(int64_t)(h * k - 1) * ldf + j];
        xr[h] = v.x;
        xi[h] = v.y;
    }
    float out[NB];
    cs_c2r::Fixed<NB>::run(xr, xi, out);
#pragma unroll
    for (int i = 0; i < NB; ++i) prof[(int64_t)i * ldp + j] = out[i];
    int64_t t = sk.t_base + j, cl;
    if (sk.sel) {
        cl = j / sk.g.Nprof;
        t = sk.g.origin + (int64_t)sk.sel[sk.slot0 + cl] * sk.g.Nprof + (j - cl * sk.g.Nprof);
        if (t >= sk.g.total) return;
    }
    int64_t cg, first, last;
    block_of(sk.g, t, cg, first, last);
    if (!sk.sel) cl = cg - sk.c_first;
    const int64_t jj = t - first;
    const int64_t nv = sk.nvalid[cg * sk.nf + sk.f];
    if (jj >= nv) return;
    const int sc = sample_column(jj, nv, NB, sk.cap);
    if (sc < 0) return;
    float* d = sk.samples + (cl * sk.nf + sk.f) * sk.sld + (int64_t)sc * NB;
#pragma unroll
    for (int i = 0; i < NB; ++i) d[i] = fabsf(out[i]);
}

bool c2r_samples_supported(int nbins) {
    switch (nbins) {
#define X(S) case S:
        CS_C2R_SIZES(X)
#undef X
        return true;
        default:
            return false;
    }
}

hipError_t launch_c2r_samples(hipStream_t st, const float2* ftp, int64_t ldf, int k, int nbins,
                              int64_t nt, float* prof, int64_t ldp, const SampleSink& sk) {
    if (nt <= 0) return hipSuccess;
    dim3 grid((unsigned)((nt + 255) / 256));
    switch (nbins) {
#define X(S)                                                                                      \
    case S:                                                                                       \
        hipLaunchKernelGGL(k_c2r_samples<S>, grid, dim3(256), 0, st, ftp, ldf, k, nt, prof, ldp,   \
                           sk);                                                                   \
        break;
        CS_C2R_SIZES(X)
#undef X
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_c2r_f32(hipStream_t st, const float2* ftp, int64_t ldf, int k,
                          int nbins, int64_t nt, float* prof, int64_t ldp,
                          const double2* twid) {
    if (nt <= 0) return hipSuccess;
    dim3 grid((unsigned)((nt + 255) / 256));
    switch (nbins) {
#define X(S)                                                                              \
    case S:                                                                               \
        hipLaunchKernelGGL(k_c2r<S>, grid, dim3(256), 0, st, ftp, ldf, k, nt, prof, ldp); \
        break;
        CS_C2R_SIZES(X)
#undef X
        default:
            hipLaunchKernelGGL(k_c2r_generic, grid, dim3(256), 0, st, ftp, ldf, k, nbins, nt,
                               prof, ldp, twid);
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// K3: block sigma (_block_sigma, search.jl:490-566, center = :zero):
// 1.4826 * median(|x|) over the reference's whole-column subsample of the
// block's valid profiles, by exact radix selection on the FP32 bit patterns
// (|x| >= 0, so the unsigned order of the bits is the numeric order).
// One 512-thread workgroup per (block, fold); up to 8192 samples in registers.
// ---------------------------------------------------------------------------
constexpr int kSelThreads = 512;
constexpr int kSelPer = 16;   // 512 * 16 = 8192 = _BOXCAR_SIGMA_SAMPLES

__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* sh, uint32_t& total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) sh[wid] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (int w = 0; w < kSelThreads / 64; ++w) {
            uint32_t tmp = sh[w];
            sh[w] = acc;
            acc += tmp;
        }
        sh[kSelThreads / 64] = acc;
    }
    __syncthreads();
    uint32_t res = x - v + sh[wid];
    total = sh[kSelThreads / 64];
    __syncthreads();
    return res;
}

// kth (0-based) smallest key of the block.
__device__ uint32_t block_select(const uint32_t (&keys)[kSelPer], uint32_t kth, uint32_t* hist,
                                 uint32_t* sh) {
    uint32_t prefix = 0, mask = 0;
    const int shifts[3] = {21, 10, 0};
    const int bits[3] = {11, 11, 10};
    for (int pass = 0; pass < 3; ++pass) {
        const int sh_ = shifts[pass];
        const uint32_t dmask = (1u << bits[pass]) - 1u;
        for (int i = threadIdx.x; i < 2048; i += kSelThreads) hist[i] = 0;
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kSelPer; ++r) {
            uint32_t key = keys[r];
            if ((key & mask) == prefix) atomicAdd(&hist[(key >> sh_) & dmask], 1u);
        }
        __syncthreads();
        uint32_t a[4], loc = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) { a[q] = hist[4 * threadIdx.x + q]; loc += a[q]; }
        uint32_t tot;
        uint32_t ex = block_excl_scan(loc, sh + 16, tot);
        if (ex <= kth && kth < ex + loc) {
            uint32_t c = ex;
            int d = 0;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (kth >= c && kth < c + a[q]) { d = q; break; }
                c += a[q];
            }
            sh[0] = 4 * threadIdx.x + d;
            sh[1] = kth - c;
        }
        __syncthreads();
        uint32_t dsel = sh[0];
        kth = sh[1];
        prefix |= dsel << sh_;
        mask |= dmask << sh_;
        __syncthreads();
    }
    return prefix;
}

__global__ __launch_bounds__(kSelThreads) void k_sigma32(FoldSet fs, GridDev g, BatchDev b) {
    const int64_t ldp = fs.ldp;
    __shared__ uint32_t hist[2048];
    __shared__ uint32_t sh[64];
    const int cl = blockIdx.x, f = blockIdx.y, nf = fs.nf;
    const int64_t cg = batch_block(b, cl);
    const int64_t first = g.origin + cg * g.Nprof;
    const int64_t nv = b.nvalid[cg * nf + f];
    const int nbins = fs.f[f].nbins;
    const int64_t cap = fs.cap[f];
    const int64_t NN = (int64_t)nbins * nv;
    if (NN == 0) {
        if (threadIdx.x == 0) b.sigma[cl * nf + f] = 0.0;
        return;
    }
    const bool all = NN <= cap;
    int64_t ncol, ns;
    if (all) { ncol = nv; ns = NN; }
    else {
        ncol = cap / nbins;
        if (ncol < 1) ncol = 1;
        if (ncol > nv) ncol = nv;
        ns = ncol * nbins < cap ? ncol * nbins : cap;
    }
    const float* P = fs.prof[f];
    const int64_t col0 = b.sel ? (int64_t)cl * g.Nprof : first - b.t_base;
    uint32_t keys[kSelPer];
#pragma unroll
    for (int r = 0; r < kSelPer; ++r) {
        int64_t s = threadIdx.x + (int64_t)kSelThreads * r;
        uint32_t key = 0xFFFFFFFFu;
        if (s < ns) {
            int64_t cc = s / nbins, i = s - cc * nbins;
            int64_t j = all ? cc : (cc * nv) / ncol;
            float v = P[i * ldp + col0 + j];
            key = __float_as_uint(fabsf(v));
        }
        keys[r] = key;
    }
    uint32_t half = (uint32_t)(ns / 2);
    uint32_t up = block_select(keys, half, hist, sh);
    double med;
    if (ns & 1) med = (double)__uint_as_float(up);
    else {
        uint32_t lo = block_select(keys, half - 1, hist, sh);
        med = 0.5 * ((double)__uint_as_float(lo) + (double)__uint_as_float(up));
    }
    if (threadIdx.x == 0) b.sigma[cl * nf + f] = 1.4826 * med;
}

hipError_t launch_sigma_f32(hipStream_t st, const FoldSet& fs, const GridDev& g,
                            const BatchDev& b) {
    if (b.nblocks <= 0) return hipSuccess;
    dim3 grid((unsigned)b.nblocks, (unsigned)fs.nf);
    hipLaunchKernelGGL(k_sigma32, grid, dim3(kSelThreads), 0, st, fs, g, b);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// K4: zero-baseline boxcar gate in FP32 (_bc_scan_batch!, search.jl:743-779,
// same per-profile operation order), then wave-aggregated compaction of the
// (trial, fold) pairs whose bound reaches medcut into the rescan queue.
// ---------------------------------------------------------------------------
// One bank width (compile-time w = bank_w(A)) of the zero-baseline scan, then
// the next; every prefix-sum index is a constant, so ps stays in registers.
template <int NB, int WM, int A>
__device__ __forceinline__ void gate_widths(const float* ps, uint32_t bankmask, float invsig_t,
                                            float& res) {
    if constexpr (A < bank_count(WM)) {
        constexpr int w = bank_w(A);
        if (bankmask & (1u << A)) {
            float invsw = invsig_t / sqrtf((float)w);
            float mb = ps[w] - ps[0];
#pragma unroll
            for (int p = 1; p < NB; ++p) {
                float d = ps[p + w] - ps[p];
                mb = d > mb ? d : mb;
            }
            float c = mb * invsw;
            res = c > res ? c : res;
        }
        gate_widths<NB, WM, A + 1>(ps, bankmask, invsig_t, res);
    }
}

// WM: compile-time bound on the fold's widest boxcar (6, the ladder cap, or the
// full default bank of this length), so the prefix sums stay in registers.
template <int NB, int WM>
__global__ __launch_bounds__(256) void k_gate(const float* __restrict__ prof, int64_t ldp,
                                              int fidx, uint32_t bankmask, GridDev g,
                                              BatchDev b, int nf, OutDev o) {
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool pred = false;
    const int64_t t = j < b.nt ? batch_trial(b, g, j) : -1;
    if (t >= 0) {
        int64_t cg, first, last;
        block_of(g, t, cg, first, last);
        int cl = batch_slot(b, cg);
        if (t - first < b.nvalid[cg * nf + fidx]) {
            double sig = b.sigma[cl * nf + fidx];
            double invs = sig > 0 ? 1.0 / sig : 0.0;
            if (!(o.medcut > -INFINITY) || invs <= 0) {
                pred = true;
            } else {
                float ps[NB + WM + 1];
                ps[0] = 0.0f;
#pragma unroll
                for (int i = 1; i <= NB; ++i) ps[i] = ps[i - 1] + prof[(int64_t)(i - 1) * ldp + j];
#pragma unroll
                for (int i = NB + 1; i <= NB + WM; ++i)
                    ps[i] = ps[i - 1] + prof[(int64_t)(i - NB - 1) * ldp + j];
                float res = -INFINITY;
                gate_widths<NB, WM, 0>(ps, bankmask, (float)invs, res);
                pred = (double)res >= o.medcut;
            }
        }
    }
    if (o.gm) pred = t >= 0 && o.gm[(int64_t)fidx * o.gm_ld + t] > o.gmcut;
    push_survivor(pred, t, fidx, o);
}

// Generic gate: runtime length and bank, prefix sums transposed in LDS (one
// column per lane, as the reference's batched tile), 64 profiles per group.
__global__ __launch_bounds__(64) void k_gate_generic(const float* __restrict__ prof,
                                                     int64_t ldp, int fidx, FoldDev fd,
                                                     GridDev g, BatchDev b, int nf, OutDev o) {
    extern __shared__ float psT[];
    const int NB = fd.nbins;
    const int wmax = fd.widths[fd.nw - 1];
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool pred = false;
    const int64_t t = j < b.nt ? batch_trial(b, g, j) : -1;
    const int L = threadIdx.x;
    if (t >= 0) {
        int64_t cg, first, last;
        block_of(g, t, cg, first, last);
        int cl = batch_slot(b, cg);
        if (t - first < b.nvalid[cg * nf + fidx]) {
            double sig = b.sigma[cl * nf + fidx];
            double invs = sig > 0 ? 1.0 / sig : 0.0;
            if (!(o.medcut > -INFINITY) || invs <= 0) {
                pred = true;
            } else {
                psT[L] = 0.0f;
                for (int i = 1; i <= NB + wmax; ++i) {
                    int idx = i > NB ? i - NB : i;
                    psT[i * 64 + L] = psT[(i - 1) * 64 + L] + prof[(int64_t)(idx - 1) * ldp + j];
                }
                const float invsig_t = (float)invs;
                float res = -INFINITY;
                for (int a = 0; a < fd.nw; ++a) {
                    int w = fd.widths[a];
                    float invsw = invsig_t / sqrtf((float)w);
                    float mb = psT[w * 64 + L] - psT[L];
                    for (int p = 1; p < NB; ++p) {
                        float d = psT[(p + w) * 64 + L] - psT[p * 64 + L];
                        mb = d > mb ? d : mb;
                    }
                    float c = mb * invsw;
                    res = c > res ? c : res;
                }
                pred = (double)res >= o.medcut;
            }
        }
    }
    if (o.gm) pred = t >= 0 && o.gm[(int64_t)fidx * o.gm_ld + t] > o.gmcut;
    push_survivor(pred, t, fidx, o);
}

hipError_t launch_gate_f32(hipStream_t st, const FoldSet& fs, const GridDev& g,
                           const BatchDev& b, const OutDev& o) {
    if (b.nt <= 0) return hipSuccess;
    const int64_t ldp = fs.ldp;
    dim3 grid((unsigned)((b.nt + 255) / 256));
    for (int f = 0; f < fs.nf; ++f) {
        const FoldDev& fd = fs.f[f];
        bool fast = fs.bankmask[f] != 0;
        bool launched = false;
        if (fast) {
            switch (fd.nbins) {
#define X(S)                                                                                 \
    case S:                                                                                  \
        if (fd.widths[fd.nw - 1] <= 6 && GateCT<S>::wmax >= 6) {                             \
            hipLaunchKernelGGL((k_gate<S, 6>), grid, dim3(256), 0, st, fs.prof[f], ldp, f,     \
                               fs.bankmask[f], g, b, fs.nf, o);                               \
            launched = true;                                                                 \
        } else if (fd.widths[fd.nw - 1] <= GateCT<S>::wmax) {                                \
            hipLaunchKernelGGL((k_gate<S, GateCT<S>::wmax>), grid, dim3(256), 0, st,          \
                               fs.prof[f], ldp, f, fs.bankmask[f], g, b, fs.nf, o);          \
            launched = true;                                                                 \
        }                                                                                    \
        break;
                CS_C2R_SIZES(X)
#undef X
                default:
                    break;
            }
        }
        if (!launched) {
            int wmax = fd.widths[fd.nw - 1];
            size_t shm = (size_t)(fd.nbins + wmax + 1) * 64 * sizeof(float);
            dim3 g64((unsigned)((b.nt + 63) / 64));
            hipLaunchKernelGGL(k_gate_generic, g64, dim3(64), shm, st, fs.prof[f], ldp, f, fd, g,
                               b, fs.nf, o);
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// ---------------------------------------------------------------------------
// Exact metric of one FP64 profile held in LDS by one 64-lane group
// (_boxcar_exact, search.jl:893-904): exact median baseline (rank counting;
// the same order statistics quickselect or a sorting network returns), the
// wrapped prefix sum in the reference's sequential order, then the
// width x phase maximum.
// ---------------------------------------------------------------------------
__device__ double group_exact_metric(const double* prof, int NB, const int32_t* widths,
                                     int nw, double invs, double* srt, double* ps) {
    const int lane = threadIdx.x;
    for (int i = lane; i < NB; i += 64) {
        double v = prof[i];
        int r = 0;
        for (int q = 0; q < NB; ++q) {
            double u = prof[q];
            r += (u < v) || (u == v && q < i);
        }
        srt[r] = v;
    }
    __syncthreads();
    double med = (NB & 1) ? srt[NB / 2] : 0.5 * (srt[NB / 2 - 1] + srt[NB / 2]);
    const int wmax = widths[nw - 1];
    if (lane == 0) {
        ps[0] = 0.0;
        for (int i = 1; i <= NB + wmax; ++i) {
            int idx = i > NB ? i - NB : i;
            ps[i] = ps[i - 1] + (prof[idx - 1] - med);
        }
    }
    __syncthreads();
    double best = -INFINITY;
    for (int a = 0; a < nw; ++a) {
        int w = widths[a];
        double invsw = invs / sqrt((double)w);
        double m = -INFINITY;
        for (int p = lane; p < NB; p += 64) {
            double d = ps[p + w] - ps[p];
            m = d > m ? d : m;
        }
        m = wave_max(m);
        double cand = m * invsw;
        if (cand > best) best = cand;
    }
    __syncthreads();
    return best;
}

// FP64 real DFT of X[0..H] (DC and Nyquist real parts only) into prof[0..NB-1],
// unnormalised, even/odd split x[n] = E + O, x[NB-n] = E - O.
__device__ void group_c2r_f64(const double2* X, int NB, const double2* twid, double* prof) {
    const int H = NB / 2;
    for (int n = threadIdx.x; n <= H; n += 64) {
        double E = 0.0, O = 0.0;
        for (int h = 0; h <= H; ++h) {
            double c, s;
            if (h == 0) { c = 1.0; s = 0.0; }
            else if (h == H) { c = (n & 1) ? -1.0 : 1.0; s = 0.0; }
            else {
                double2 w = twid[(h * n) % NB];
                c = 2.0 * w.x;
                s = -(2.0 * w.y);
            }
            E = fma(c, X[h].x, E);
            O = fma(s, X[h].y, O);
        }
        if (n == 0 || n == H) prof[n] = E;
        else {
            prof[n] = E + O;
            prof[NB - n] = E - O;
        }
    }
    __syncthreads();
}

constexpr int kMaxNB = 128;

// K5: exact rescan of the queued (trial, fold) pairs.  Each 64-lane group
// rebuilds its trial's fold stack from the spectrum (the same FP32-weight
// arithmetic as K1), transforms it in FP64 and scores it exactly; candidates
// above threshold are appended (the host sorts them by (t, k)).
__global__ __launch_bounds__(64) void k_rescan(SpecDev sp, TabDev tb, GridDev g, FoldSet fs,
                                               BatchDev b, OutDev o) {
    __shared__ double2 X[kMaxNB / 2 + 1];
    __shared__ double prof[kMaxNB];
    __shared__ double srt[kMaxNB];
    __shared__ double ps[2 * kMaxNB + 2];
    const int lane = threadIdx.x;
    uint32_t nsurv = *o.surv_count;
    if (nsurv > o.surv_cap) nsurv = o.surv_cap;
    for (uint32_t i = blockIdx.x; i < nsurv; i += gridDim.x) {
        uint64_t pk = o.surv[i];
        int64_t t = (int64_t)(pk >> 4);
        int f = (int)(pk & 15u);
        const FoldDev& fd = fs.f[f];
        const int k = fd.k, Hk = fd.Hk, NB = fd.nbins;
        int64_t cg, first, last;
        block_of(g, t, cg, first, last);
        int cl = batch_slot(b, cg);
        double sig = b.sigma[cl * fs.nf + f];
        double invs = sig > 0 ? 1.0 / sig : 0.0;
        double metric = 0.0;
        if (invs > 0) {
            for (int l = lane; l <= Hk; l += 64) {
                double2 v = make_double2(0.0, 0.0);
                if (l >= 1) {
                    if (tb.m == 16 && !tb.weights_f64) v = interp_one<16>(sp, tb, l * k, t, first, last, false, false);
                    else if (tb.m == 32 && !tb.weights_f64) v = interp_one<32>(sp, tb, l * k, t, first, last, false, false);
                    else v = interp_one<0>(sp, tb, l * k, t, first, last, false, false);
                }
                X[l] = v;
            }
            __syncthreads();
            group_c2r_f64(X, NB, fs.twid[f], prof);
            metric = group_exact_metric(prof, NB, fd.widths, fd.nw, invs, srt, ps);
        }
        if (lane == 0) {
            if (o.mall) o.mall[(int64_t)f * o.mall_ld + t] = metric;
            if (metric > o.threshold) {
                uint32_t idx = atomicAdd(o.cand_count, 1u);
                if (idx < o.cand_cap) {
                    cs_cand_dev c;
                    c.t = t; c.k = k; c.nharm = Hk; c.block = cg; c.metric = metric;
                    o.cands[idx] = c;
                }
            }
        }
        __syncthreads();
    }
}

hipError_t launch_rescan(hipStream_t st, const SpecDev& sp, const TabDev& tb,
                         const GridDev& g, const FoldSet& fs, const BatchDev& b,
                         const OutDev& o, int nblocks_grid) {
    hipLaunchKernelGGL(k_rescan, dim3((unsigned)nblocks_grid), dim3(64), 0, st, sp, tb, g, fs,
                       b, o);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// FP64 block sigma of the (block, fold) pairs that hold gate survivors: the
// reference's default profile stage is Float64 (SearchParams precision :f64,
// search.jl:143,151-155), so its _block_sigma (search.jl:490-566) takes the
// median of |x| over FP64 profiles.  The pass-1 sigma of this build comes from
// FP32 (plain) or split-BF16 (fused) profiles; it only has to decide which
// trials the gate keeps (the 2-sigma medmargin dwarfs its ~1e-5 relative
// error).  The sigma that scales a reported metric is recomputed here from the
// same whole-column subsample: each sampled column's stack with the FP32
// weights (as the reference fills ftprofs), an FP64 real DFT, |x|, then the
// exact median by radix selection over the sample keys held in registers.
//
// The flagged pairs are listed once each; eight workgroups per listed pair
// compute its columns into a scratch of FP64 samples (so a search with few
// flagged pairs still spreads over the GPU), then one workgroup per pair
// selects its median.  Pairs past the scratch's capacity are done whole by
// one workgroup each, columns in LDS.
// ---------------------------------------------------------------------------
__global__ void k_reset_u32(uint32_t* p, int n);

// Flag every (block, fold) pair holding a survivor; the flag's first setter
// appends the pair to the list (count: the word after the flags).
__global__ void k_refine_mark(const OutDev o, GridDev g, BatchDev b, int nf,
                              uint32_t* __restrict__ flags, uint32_t* __restrict__ list,
                              uint32_t* __restrict__ count) {
    uint32_t nsurv = *o.surv_count;
    if (nsurv > o.surv_cap) nsurv = o.surv_cap;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nsurv;
         i += gridDim.x * blockDim.x) {
        uint64_t pk = o.surv[i];
        int64_t t = (int64_t)(pk >> 4);
        int f = (int)(pk & 15u);
        int64_t cg, first, last;
        block_of(g, t, cg, first, last);
        int cl = batch_slot(b, cg);
        if (cl >= 0) {
            const uint32_t pair = (uint32_t)(cl * nf + f);
            if (atomicCAS(&flags[pair], 0u, 1u) == 0u) list[atomicAdd(count, 1u)] = pair;
        }
    }
}

constexpr int kRefineThreads = 1024;
constexpr int kRefinePer = kSigmaSamples / kRefineThreads;   // sample keys per thread
constexpr int kRefineX = kRefineThreads + 64;                // stack entries of one round
constexpr int kRefineColThreads = 256;                       // a column chunk's workgroup
constexpr int kRefineColX = 1024 + 64;                       // its stack entries per round

// One refined pair: its block, fold and sampled columns (_block_sigma's
// whole-column subsample, search.jl:504-538).  False when the fold has no
// valid trial in the block (sigma 0).
struct RefinePair {
    int f, k, Hk, NB;
    int64_t first, last, nv, ncol, ns;
    bool all;
};
__device__ __forceinline__ bool refine_pair(const GridDev& g, const FoldSet& fs,
                                            const BatchDev& b, uint32_t pair, RefinePair& r) {
    const int nf = fs.nf;
    const int cl = (int)pair / nf;
    r.f = (int)pair - cl * nf;
    const int64_t cg = batch_block(b, cl);
    r.first = g.origin + cg * g.Nprof;
    r.last = min(r.first + g.Nprof, g.total) - 1;
    const FoldDev& fd = fs.f[r.f];
    r.k = fd.k;
    r.Hk = fd.Hk;
    r.NB = fd.nbins;
    r.nv = b.nvalid[cg * nf + r.f];
    const int64_t NN = (int64_t)r.NB * r.nv;
    const int64_t cap = fs.cap[r.f];
    r.all = NN <= cap;
    if (r.all) {
        r.ncol = r.nv;
        r.ns = NN;
    } else {
        r.ncol = cap / r.NB;
        if (r.ncol < 1) r.ncol = 1;
        if (r.ncol > r.nv) r.ncol = r.nv;
        r.ns = r.ncol * r.NB < cap ? r.ncol * r.NB : cap;
    }
    return NN > 0;
}

// The pair's twiddles, and the plans and row guards of its harmonics l*k
// (l = 1..Hk), once for all its columns.
template <int NT>
__device__ __forceinline__ void refine_setup(const SpecDev& sp, const TabDev& tb,
                                             const FoldSet& fs, const RefinePair& r, double2* tw,
                                             HarmDev* hd, uint8_t* hok) {
    for (int e = threadIdx.x; e < r.NB; e += NT) tw[e] = fs.twid[r.f][e];
    for (int l = threadIdx.x; l < r.Hk; l += NT) {
        const HarmDev d = tb.harm[(l + 1) * r.k - 1];
        hd[l] = d;
        hok[l] = d.P == 0 || row_guard(d, tb, sp, r.first, r.last) ? 1 : 0;
    }
    __syncthreads();
}

// Sampled columns c_lo .. c_hi - 1 of pair r, in rounds of NX stack entries:
// every (column, harmonic) stack entry, then every (column, bin pair) of the
// FP64 real DFT; put(s, |x|) receives sample s = column * NB + bin.  Called
// by the whole workgroup of NT threads; ends at a barrier.
template <int NT, int NX, class Put>
__device__ __forceinline__ void refine_columns(const SpecDev& sp, const TabDev& tb,
                                               const RefinePair& r, int64_t c_lo, int64_t c_hi,
                                               double2* X, const double2* tw, const HarmDev* hd,
                                               const uint8_t* hok, Put put) {
    const int tid = threadIdx.x;
    const int H = r.Hk, W = H + 1, NB = r.NB;
    const int CR = NX / W;
    for (int64_t c0 = c_lo; c0 < c_hi; c0 += CR) {
        const int nc = (int)min((int64_t)CR, c_hi - c0);
        for (int task = tid; task < nc * W; task += NT) {
            const int cr = task / W, l = task - cr * W;
            const int64_t cc = c0 + cr;
            const int64_t t = r.first + (r.all ? cc : (cc * r.nv) / r.ncol);
            double2 x = make_double2(0.0, 0.0);
            if (l >= 1 && hok[l - 1]) {
                const HarmDev& d = hd[l - 1];
                if (tb.m == 16 && !tb.weights_f64) x = interp_with<16>(d, sp, tb, l * r.k, t, r.first, r.last, true, true);
                else if (tb.m == 32 && !tb.weights_f64) x = interp_with<32>(d, sp, tb, l * r.k, t, r.first, r.last, true, true);
                else x = interp_with<0>(d, sp, tb, l * r.k, t, r.first, r.last, true, true);
            }
            X[task] = x;
        }
        __syncthreads();
        for (int task = tid; task < nc * (H + 1); task += NT) {
            const int cr = task / (H + 1), n = task - cr * (H + 1);
            const double2* Xc = X + cr * W;
            const int64_t s0 = (c0 + cr) * NB;
            // the column gives min(NB, ns - s0) samples: its bins 0 .. that - 1
            const int64_t take = r.ns - s0 < NB ? r.ns - s0 : NB;
            double E = 0.0, O = 0.0;
            int idx = 0;                                 // (h n) mod NB
            for (int h = 1; h <= H; ++h) {
                idx += n;
                if (idx >= NB) idx -= NB;
                double c, sn;
                if (h == H) { c = (n & 1) ? -1.0 : 1.0; sn = 0.0; }
                else {
                    const double2 w = tw[idx];
                    c = 2.0 * w.x;
                    sn = -(2.0 * w.y);
                }
                E = fma(c, Xc[h].x, E);
                O = fma(sn, Xc[h].y, O);
            }
            if (n == 0 || n == H) {
                if (n < take) put(s0 + n, fabs(E));
            } else {
                if (n < take) put(s0 + n, fabs(E + O));
                if (NB - n < take) put(s0 + NB - n, fabs(E - O));
            }
        }
        __syncthreads();
    }
}

// Exclusive prefix sum over the workgroup; `tot` holds kRefineThreads/64 words,
// rewritten by the next call only after the caller's next two barriers.
__device__ __forceinline__ uint32_t refine_excl_scan(uint32_t v, uint32_t* tot) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) tot[wid] = x;
    __syncthreads();
    uint32_t pre = 0;
    for (int w = 0; w < wid; ++w) pre += tot[w];
    return pre + x - v;
}

// Largest key of the workgroup among those for which `take` holds (others 0).
__device__ __forceinline__ uint64_t refine_max(uint64_t m, uint64_t* red) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        const uint64_t u = __shfl_xor(m, o, 64);
        m = u > m ? u : m;
    }
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    uint64_t r = 0;
    for (int w = 0; w < kRefineThreads / 64; ++w) r = red[w] > r ? red[w] : r;
    __syncthreads();
    return r;
}

// The key of rank kth (0-based) among the workgroup's keys (bit patterns of
// doubles >= 0, whose unsigned order is the numeric order): radix digits of
// 11 bits from the top; kth is left as the key's rank among its equals.  A
// digit whose bin holds one key ends the search: that key is the maximum of
// the keys with the prefix so far.
__device__ uint64_t refine_select(const uint64_t (&keys)[kRefinePer], uint32_t& kth,
                                  uint32_t* hist, uint32_t* shv, uint64_t* red) {
    uint64_t prefix = 0, mask = 0;
    for (int pass = 0; pass < 6; ++pass) {
        const int sh = pass < 5 ? 53 - 11 * pass : 0;
        const int bits = pass < 5 ? 11 : 9;
        const uint64_t dmask = ((uint64_t)1 << bits) - 1u;
        hist[2 * threadIdx.x] = 0;
        hist[2 * threadIdx.x + 1] = 0;
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kRefinePer; ++r)
            if ((keys[r] & mask) == prefix) atomicAdd(&hist[(keys[r] >> sh) & dmask], 1u);
        __syncthreads();
        const uint32_t a0 = hist[2 * threadIdx.x], a1 = hist[2 * threadIdx.x + 1];
        const uint32_t ex = refine_excl_scan(a0 + a1, shv + 16);
        if (ex <= kth && kth < ex + a0 + a1) {
            const bool second = kth >= ex + a0;
            shv[0] = 2 * threadIdx.x + (second ? 1u : 0u);
            shv[1] = kth - ex - (second ? a0 : 0u);
            shv[2] = second ? a1 : a0;
        }
        __syncthreads();
        prefix |= (uint64_t)shv[0] << sh;
        mask |= dmask << sh;
        kth = shv[1];
        const uint32_t cnt = shv[2];
        __syncthreads();
        if (cnt == 1u && pass < 5) {    // block-uniform
            uint64_t m = 0;
#pragma unroll
            for (int r = 0; r < kRefinePer; ++r)
                m = (keys[r] & mask) == prefix && keys[r] > m ? keys[r] : m;
            kth = 0;
            return refine_max(m, red);
        }
    }
    return prefix;
}

// sigma = 1.4826 * median of the ns sample keys (bit patterns of doubles >= 0)
// the workgroup holds, kRefinePer per thread (keys past ns are all ones)
__device__ __forceinline__ double refine_sigma(const uint64_t (&keys)[kRefinePer], int64_t ns,
                                               uint32_t* hist, uint32_t* shv, uint64_t* red) {
    const uint32_t half = (uint32_t)(ns / 2);
    uint32_t rk = half;
    const uint64_t up = refine_select(keys, rk, hist, shv, red);
    double med;
    if ((ns & 1) || rk > 0) {
        med = __longlong_as_double((long long)up);
    } else {
        // exactly `half` keys lie below `up`: the lower central value is the
        // largest of them
        uint64_t m = 0;
#pragma unroll
        for (int r = 0; r < kRefinePer; ++r) m = keys[r] < up && keys[r] > m ? keys[r] : m;
        const uint64_t lo = refine_max(m, red);
        med = 0.5 * (__longlong_as_double((long long)lo) + __longlong_as_double((long long)up));
    }
    return 1.4826 * med;
}

// Column chunks of the listed pairs 0 .. min(count, cap) - 1: eight chunks per
// pair, each workgroup one chunk at a time; samples of list entry e go to
// samples[e * kSigmaSamples + s].
__global__ __launch_bounds__(kRefineColThreads) void k_refine_cols(
    SpecDev sp, TabDev tb, GridDev g, FoldSet fs, BatchDev b, const uint32_t* __restrict__ list,
    const uint32_t* __restrict__ count, uint32_t cap, double* __restrict__ samples) {
    __shared__ double2 X[kRefineColX];
    __shared__ double2 tw[kMaxNB];
    __shared__ HarmDev hd[kMaxNB / 2];
    __shared__ uint8_t hok[kMaxNB / 2];
    const uint32_t cnt = *count, n = cnt < cap ? cnt : cap;
    for (uint32_t item = blockIdx.x; item < n * kRefineChunks; item += gridDim.x) {
        const uint32_t e = item / kRefineChunks, j = item - e * kRefineChunks;
        RefinePair r;
        if (!refine_pair(g, fs, b, list[e], r)) continue;   // uniform over the workgroup
        const int64_t per = (r.ncol + kRefineChunks - 1) / kRefineChunks;
        const int64_t c_lo = (int64_t)j * per, c_hi = min(r.ncol, c_lo + per);
        if (c_lo >= c_hi) continue;
        refine_setup<kRefineColThreads>(sp, tb, fs, r, tw, hd, hok);
        double* out = samples + (size_t)e * kSigmaSamples;
        refine_columns<kRefineColThreads, kRefineColX>(sp, tb, r, c_lo, c_hi, X, tw, hd, hok,
                                                       [&](int64_t s, double v) { out[s] = v; });
    }
}

// The median of each listed pair below cap from its scratch samples.
__global__ __launch_bounds__(kRefineThreads) void k_refine_select(
    GridDev g, FoldSet fs, BatchDev b, const uint32_t* __restrict__ list,
    const uint32_t* __restrict__ count, uint32_t cap, const double* __restrict__ samples) {
    __shared__ uint32_t hist[2 * kRefineThreads];
    __shared__ uint32_t shv[32];
    __shared__ uint64_t red[kRefineThreads / 64];
    const uint32_t cnt = *count, n = cnt < cap ? cnt : cap;
    for (uint32_t e = blockIdx.x; e < n; e += gridDim.x) {
        const uint32_t pair = list[e];
        RefinePair r;
        if (!refine_pair(g, fs, b, pair, r)) {
            if (threadIdx.x == 0) b.sigma[pair] = 0.0;
            continue;
        }
        const double* in = samples + (size_t)e * kSigmaSamples;
        uint64_t keys[kRefinePer];
#pragma unroll
        for (int q = 0; q < kRefinePer; ++q) {
            const int64_t s = threadIdx.x + (int64_t)kRefineThreads * q;
            keys[q] = s < r.ns ? (uint64_t)__double_as_longlong(in[s]) : ~(uint64_t)0;
        }
        const double sig = refine_sigma(keys, r.ns, hist, shv, red);
        if (threadIdx.x == 0) b.sigma[pair] = sig;
        __syncthreads();
    }
}

// Listed pairs from cap on (more than the scratch holds), each whole on one
// workgroup: its columns into LDS, then the median.
__global__ __launch_bounds__(kRefineThreads) void k_sigma_refine(
    SpecDev sp, TabDev tb, GridDev g, FoldSet fs, BatchDev b, const uint32_t* __restrict__ list,
    const uint32_t* __restrict__ count, uint32_t cap) {
    __shared__ double v[kSigmaSamples];
    __shared__ double2 X[kRefineX];
    __shared__ double2 tw[kMaxNB];
    __shared__ uint32_t hist[2 * kRefineThreads];
    __shared__ uint32_t shv[32];
    __shared__ uint64_t red[kRefineThreads / 64];
    __shared__ HarmDev hd[kMaxNB / 2];
    __shared__ uint8_t hok[kMaxNB / 2];
    const uint32_t cnt = *count;
    for (uint32_t e = cap + blockIdx.x; e < cnt; e += gridDim.x) {
        const uint32_t pair = list[e];
        RefinePair r;
        if (!refine_pair(g, fs, b, pair, r)) {
            if (threadIdx.x == 0) b.sigma[pair] = 0.0;
            continue;
        }
        refine_setup<kRefineThreads>(sp, tb, fs, r, tw, hd, hok);
        refine_columns<kRefineThreads, kRefineX>(sp, tb, r, 0, r.ncol, X, tw, hd, hok,
                                                 [&](int64_t s, double x) { v[s] = x; });
        uint64_t keys[kRefinePer];
#pragma unroll
        for (int q = 0; q < kRefinePer; ++q) {
            const int64_t s = threadIdx.x + (int64_t)kRefineThreads * q;
            keys[q] = s < r.ns ? (uint64_t)__double_as_longlong(v[s]) : ~(uint64_t)0;
        }
        const double sig = refine_sigma(keys, r.ns, hist, shv, red);
        if (threadIdx.x == 0) b.sigma[pair] = sig;
        __syncthreads();
    }
}

hipError_t launch_sigma_refine(hipStream_t st, const SpecDev& sp, const TabDev& tb,
                               const GridDev& g, const FoldSet& fs, const BatchDev& b,
                               const OutDev& o, const RefineBufs& rb) {
    const int npair = b.nblocks * fs.nf;
    if (npair <= 0) return hipSuccess;
    uint32_t* count = rb.flags + npair;
    if (!rb.flags_clear)
        hipLaunchKernelGGL(k_reset_u32, dim3((unsigned)((npair + 1 + 63) / 64)), dim3(64), 0, st,
                           rb.flags, npair + 1);
    hipLaunchKernelGGL(k_refine_mark, dim3(64), dim3(256), 0, st, o, g, b, fs.nf, rb.flags, rb.list,
                       count);
    const uint32_t held = (uint32_t)npair < rb.cap ? (uint32_t)npair : rb.cap;
    const unsigned gcols = held * kRefineChunks < 4096u ? held * kRefineChunks : 4096u;
    hipLaunchKernelGGL(k_refine_cols, dim3(gcols), dim3(kRefineColThreads), 0, st, sp, tb, g, fs, b,
                       (const uint32_t*)rb.list, (const uint32_t*)count, rb.cap, rb.samples);
    hipLaunchKernelGGL(k_refine_select, dim3(held < 1024u ? held : 1024u), dim3(kRefineThreads), 0,
                       st, g, fs, b, (const uint32_t*)rb.list, (const uint32_t*)count, rb.cap,
                       (const double*)rb.samples);
    if ((uint32_t)npair > rb.cap) {   // more pairs than the scratch holds may be listed
        const uint32_t rest = (uint32_t)npair - rb.cap;
        hipLaunchKernelGGL(k_sigma_refine, dim3(rest < 256u ? rest : 256u), dim3(kRefineThreads), 0,
                           st, sp, tb, g, fs, b, (const uint32_t*)rb.list, (const uint32_t*)count,
                           rb.cap);
    }
    return hipGetLastError();
}

__global__ void k_reset_u32(uint32_t* p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

__global__ void k_reset2_u32(uint32_t* p, int n, uint32_t* q, int m) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
    else if (q && i - n < m) q[i - n] = 0u;
}

__global__ void k_copy_u32(const uint32_t* src, uint32_t* dst) {
    if (threadIdx.x == 0) *dst = *src;
}

hipError_t launch_copy_u32(hipStream_t st, const uint32_t* src, uint32_t* dst) {
    hipLaunchKernelGGL(k_copy_u32, dim3(1), dim3(64), 0, st, src, dst);
    return hipGetLastError();
}

hipError_t launch_reset_u32(hipStream_t st, uint32_t* p, int n) {
    hipLaunchKernelGGL(k_reset_u32, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, st, p, n);
    return hipGetLastError();
}

hipError_t launch_reset2_u32(hipStream_t st, uint32_t* p, int n, uint32_t* q, int m) {
    const int tot = n + (q ? m : 0);
    hipLaunchKernelGGL(k_reset2_u32, dim3((unsigned)((tot + 63) / 64)), dim3(64), 0, st, p, n, q, m);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// FP64 bridge path: chunk_metrics / decim_pass! / block_metrics.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_c2r_f64(const double2* __restrict__ ftp, int64_t ldf,
                                                 int k, int NB, int64_t n,
                                                 double* __restrict__ prof, int64_t ldp,
                                                 const double2* __restrict__ twid,
                                                 double scale) {
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int H = NB / 2;
    for (int nn = 0; nn <= H; ++nn) {
        double E = 0.0, O = 0.0;
        for (int h = 1; h <= H; ++h) {
            double2 v = ftp[(int64_t)(h * k - 1) * ldf + j];
            double c, s;
            if (h == H) { c = (nn & 1) ? -1.0 : 1.0; s = 0.0; }
            else {
                double2 w = twid[(h * nn) % NB];
                c = 2.0 * w.x;
                s = -(2.0 * w.y);
            }
            E = fma(c, v.x, E);
            O = fma(s, v.y, O);
        }
        double a, bb;
        if (nn == 0 || nn == H) { a = E; bb = E; }
        else { a = E + O; bb = E - O; }
        if (scale != 1.0) { a *= scale; bb *= scale; }
        prof[(int64_t)nn * ldp + j] = a;
        if (nn != 0 && nn != H) prof[(int64_t)(NB - nn) * ldp + j] = bb;
    }
}

hipError_t launch_c2r_f64(hipStream_t st, const double2* ftp, int64_t ldf, int k, int nbins,
                          int64_t n, double* prof, int64_t ldp, const double2* twid,
                          double scale) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_c2r_f64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, ftp, ldf,
                       k, nbins, n, prof, ldp, twid, scale);
    return hipGetLastError();
}

// Exact median of |x| over the subsample, by a bitonic sort of up to 8192
// doubles in LDS (one workgroup: the bridge handles one block at a time).
__global__ __launch_bounds__(1024) void k_sigma64(const double* __restrict__ prof, int64_t ldp,
                                                  int nbins, int64_t n, int64_t cap,
                                                  double* sigma_out) {
    __shared__ double v[8192];
    const int64_t NN = (int64_t)nbins * n;
    if (NN == 0) {
        if (threadIdx.x == 0) *sigma_out = 0.0;
        return;
    }
    const bool all = NN <= cap;
    int64_t ncol, ns;
    if (all) { ncol = n; ns = NN; }
    else {
        ncol = cap / nbins;
        if (ncol < 1) ncol = 1;
        if (ncol > n) ncol = n;
        ns = ncol * nbins < cap ? ncol * nbins : cap;
    }
    int S = 1;
    while (S < ns) S <<= 1;
    for (int s = threadIdx.x; s < S; s += blockDim.x) {
        double x = INFINITY;
        if (s < ns) {
            int64_t cc = s / nbins, i = s - cc * nbins;
            int64_t j = all ? cc : (cc * n) / ncol;
            x = fabs(prof[i * ldp + j]);
        }
        v[s] = x;
    }
    __syncthreads();
    for (int kk = 2; kk <= S; kk <<= 1) {
        for (int jj = kk >> 1; jj > 0; jj >>= 1) {
            for (int i = threadIdx.x; i < S; i += blockDim.x) {
                int ixj = i ^ jj;
                if (ixj > i) {
                    bool up = (i & kk) == 0;
                    double a = v[i], c = v[ixj];
                    if ((a > c) == up) { v[i] = c; v[ixj] = a; }
                }
            }
            __syncthreads();
        }
    }
    if (threadIdx.x == 0) {
        int64_t half = ns / 2;
        double med = (ns & 1) ? v[half] : 0.5 * (v[half - 1] + v[half]);
        *sigma_out = 1.4826 * med;
    }
}

hipError_t launch_sigma_f64(hipStream_t st, const double* prof, int64_t ldp, int nbins, int64_t n,
                            int64_t cap, double* sigma_out) {
    hipLaunchKernelGGL(k_sigma64, dim3(1), dim3(1024), 0, st, prof, ldp, nbins, n, cap, sigma_out);
    return hipGetLastError();
}

__global__ __launch_bounds__(64) void k_exact_f64(const double* __restrict__ profs, int64_t ldp,
                                                  int64_t ldt, int NB, int64_t n,
                                                  const int32_t* widths, int nw,
                                                  const double* sigma, double* out) {
    __shared__ double prof[kMaxNB];
    __shared__ double srt[kMaxNB];
    __shared__ double ps[2 * kMaxNB + 2];
    __shared__ int32_t wsh[64];
    for (int a = threadIdx.x; a < nw; a += 64) wsh[a] = widths[a];
    __syncthreads();
    double sig = *sigma;
    double invs = sig > 0 ? 1.0 / sig : 0.0;
    for (int64_t j = blockIdx.x; j < n; j += gridDim.x) {
        double metric = 0.0;
        if (invs > 0) {
            for (int i = threadIdx.x; i < NB; i += 64) prof[i] = profs[(int64_t)i * ldp + j * ldt];
            __syncthreads();
            metric = group_exact_metric(prof, NB, wsh, nw, invs, srt, ps);
        }
        if (threadIdx.x == 0) out[j] = metric;
        __syncthreads();
    }
}

hipError_t launch_exact_f64(hipStream_t st, const double* prof, int64_t ldp, int nbins, int64_t n,
                            const int32_t* widths, int nw, const double* sigma, double* out,
                            int64_t ldt) {
    if (n <= 0) return hipSuccess;
    int64_t gsz = n < 65536 ? n : 65536;
    hipLaunchKernelGGL(k_exact_f64, dim3((unsigned)gsz), dim3(64), 0, st, prof, ldp, ldt, nbins, n,
                       widths, nw, sigma, out);
    return hipGetLastError();
}

// fourier_interp (fourierinterp.jl:66-74) point by point in FP64: coefficients
// sinc(dr-j) * cispi(dr-j), j = -m/2+1..m/2, applied as sum conj(c) * bin.
__device__ double sinc_d(double x) {
    if (x == 0.0) return 1.0;
    if (fabs(x) < 4.5 * 1.4901161193847656e-08) {
        double x2 = x * x, p2 = M_PI * M_PI;
        return 1.0 + x2 * (-p2 / 6.0 + x2 * (p2 * p2 / 120.0));
    }
    return sinpi(x) / (M_PI * x);
}

__global__ __launch_bounds__(256) void k_point_interp(SpecDev sp, const double* __restrict__ rfund,
                                                      int64_t L, int m, const uint8_t* rowok,
                                                      double2* out, int64_t ld) {
    const int h = blockIdx.y + 1;
    int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= L) return;
    double2 res = make_double2(0.0, 0.0);
    if (rowok[h - 1]) {
        double r = rfund[j] * (double)h;
        double dr = r - floor(r);
        int64_t lo0 = (int64_t)floor(r + 1e-15) + 1 - m / 2;
        if (lo0 >= 0 && lo0 + m <= sp.namps) {
            double sr = 0.0, si = 0.0;
            for (int i = 0; i < m; ++i) {
                double off = dr - (double)(-(m / 2) + 1 + i);
                double s = sinc_d(off);
                double cr = s * cospi(off), ci = s * sinpi(off);
                float2 bv = sp.amps[lo0 + i];
                sr += cr * (double)bv.x + ci * (double)bv.y;
                si += cr * (double)bv.y - ci * (double)bv.x;
            }
            res = make_double2(sr, si);
        } else {
            res = make_double2(NAN, NAN);
        }
    }
    out[(int64_t)(h - 1) * ld + j] = res;
}

// FP64 Eqn. 30 at fine-grid point lobin + i/nb (fourierinterp.jl:23-74): the
// FFT correlation of finterp_fft (fourierinterp.jl:195-214) produces exactly
// this sum at each grid point, the kernel taps b - r in [-m/2, m/2 - 1/nb]
// reducing to Eqn. 30's j = -m/2+1 .. m/2 (and to the bin itself at offset 0).
__device__ double2 e30_grid(const SpecDev& sp, int64_t lobin, int64_t i, int nb, int m) {
    const int64_t q = i / nb, rem = i - q * nb;
    const int64_t b0 = lobin + q;
    if (rem == 0) {
        if (b0 < 0 || b0 >= sp.namps) return make_double2(NAN, NAN);
        const float2 v = sp.amps[b0];
        return make_double2((double)v.x, (double)v.y);
    }
    const double dr = (double)rem / (double)nb;
    const int64_t lo0 = b0 + 1 - m / 2;
    if (lo0 < 0 || lo0 + m > sp.namps) return make_double2(NAN, NAN);
    double sr = 0.0, si = 0.0;
    for (int k = 0; k < m; ++k) {
        const double off = dr - (double)(-(m / 2) + 1 + k);
        const double s = sinc_d(off);
        const double cr = s * cospi(off), ci = s * sinpi(off);
        const float2 bv = sp.amps[lo0 + k];
        sr += cr * (double)bv.x + ci * (double)bv.y;
        si += cr * (double)bv.y - ci * (double)bv.x;
    }
    return make_double2(sr, si);
}

// block_metrics(kernel = :fft): harmonic h of trial j is the fine-grid
// interpolation of finterp_fft(lobin_h, numbins_h, nb) at r = rfund[j] * h,
// linear between the two neighbouring grid points and clamped to the grid ends
// (uniform_linear_interp, search.jl:974-986); rows failing the guard stay 0.
__global__ __launch_bounds__(256) void k_point_interp_fft(SpecDev sp, const double* __restrict__ rfund,
                                                          int64_t L, int m, int nb,
                                                          const int64_t* __restrict__ lob,
                                                          const int64_t* __restrict__ kgrid,
                                                          const uint8_t* rowok, double2* out,
                                                          int64_t ld) {
    const int h = blockIdx.y + 1;
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= L) return;
    double2 res = make_double2(0.0, 0.0);
    if (rowok[h - 1]) {
        const int64_t lobin = lob[h - 1], K = kgrid[h - 1];
        const double r = rfund[j] * (double)h;
        const double p = (r - (double)lobin) * (double)nb;
        if (p <= 0.0) {
            res = e30_grid(sp, lobin, 0, nb, m);
        } else if (p >= (double)(K - 1)) {
            res = e30_grid(sp, lobin, K - 1, nb, m);
        } else {
            const int64_t i0 = (int64_t)floor(p);
            const double f = p - (double)i0;
            const double2 a = e30_grid(sp, lobin, i0, nb, m), c = e30_grid(sp, lobin, i0 + 1, nb, m);
            res = make_double2(a.x * (1.0 - f) + c.x * f, a.y * (1.0 - f) + c.y * f);
        }
    }
    out[(int64_t)(h - 1) * ld + j] = res;
}

hipError_t launch_point_interp_fft(hipStream_t st, const SpecDev& sp, const double* rfund,
                                   int64_t L, int nharms, int m, int nb, const int64_t* lob,
                                   const int64_t* kgrid, const uint8_t* rowok, double2* out,
                                   int64_t ld) {
    if (L <= 0) return hipSuccess;
    dim3 grid((unsigned)((L + 255) / 256), (unsigned)nharms);
    hipLaunchKernelGGL(k_point_interp_fft, grid, dim3(256), 0, st, sp, rfund, L, m, nb, lob, kgrid,
                       rowok, out, ld);
    return hipGetLastError();
}

hipError_t launch_point_interp_f64(hipStream_t st, const SpecDev& sp, const double* rfund,
                                   int64_t L, int nharms, int m, const uint8_t* rowok,
                                   double2* out, int64_t ld) {
    if (L <= 0) return hipSuccess;
    dim3 grid((unsigned)((L + 255) / 256), (unsigned)nharms);
    hipLaunchKernelGGL(k_point_interp, grid, dim3(256), 0, st, sp, rfund, L, m, rowok, out, ld);
    return hipGetLastError();
}

// Candidate profiles (candidate_profile, candidate.jl:37-61) and their best
// boxcar width (boxcar_best_width, search.jl:617-638), one 64-lane group per
// candidate: the first H harmonics at r h by FP64 fourier_interp with an m-bin
// kernel, a zero DC term, the FP64 real inverse DFT into 2H bins scaled by
// 1/(2H); H = 0 gives the two-bin zero profile.  Then, over the fold's width
// bank, the median-baseline circular boxcar maximum, the narrowest width
// winning ties (strict >, ascending widths).
__global__ __launch_bounds__(64) void k_cand_profiles(SpecDev sp, const double* __restrict__ r,
                                                      const int32_t* __restrict__ H, int64_t n,
                                                      int m, const int32_t* __restrict__ wtab,
                                                      const int32_t* __restrict__ woff,
                                                      double* __restrict__ prof_out, int64_t ld,
                                                      int32_t* __restrict__ best_w) {
    __shared__ double2 X[kMaxCandH + 1];
    __shared__ double2 tw[2 * kMaxCandH];
    __shared__ double prof[2 * kMaxCandH];
    __shared__ double srt[2 * kMaxCandH];
    __shared__ double ps[4 * kMaxCandH + 2];
    const int lane = threadIdx.x;
    for (int64_t c = blockIdx.x; c < n; c += gridDim.x) {
        const int Hc = H[c];
        if (Hc < 0) {   // not measured
            if (lane == 0) best_w[c] = 0;
            continue;
        }
        const int NB = Hc >= 1 ? 2 * Hc : 2;
        const double rc = r[c];
        for (int h = lane; h <= Hc; h += 64) {
            double2 v = make_double2(0.0, 0.0);
            if (h >= 1) {   // fourier_interp(r h, amps, m), fourierinterp.jl:66-74
                const double rh = rc * (double)h;
                const double dr = rh - floor(rh);
                const int64_t lo0 = (int64_t)floor(rh + 1e-15) + 1 - m / 2;
                double sr = 0.0, si = 0.0;
                for (int i = 0; i < m; ++i) {
                    const double off = dr - (double)(-(m / 2) + 1 + i);
                    const double s = sinc_d(off);
                    const double cr = s * cospi(off), ci = s * sinpi(off);
                    const float2 bv = sp.amps[lo0 + i];
                    sr += cr * (double)bv.x + ci * (double)bv.y;
                    si += cr * (double)bv.y - ci * (double)bv.x;
                }
                v = make_double2(sr, si);
            }
            X[h] = v;
        }
        for (int e = lane; e < NB; e += 64) {
            double sn, cs_;
            sincospi(2.0 * (double)e / (double)NB, &sn, &cs_);
            tw[e] = make_double2(cs_, sn);
        }
        __syncthreads();
        if (Hc >= 1) {
            group_c2r_f64(X, NB, tw, prof);
            const double sc = 1.0 / (double)NB;
            for (int i = lane; i < NB; i += 64) prof[i] *= sc;
        } else {
            for (int i = lane; i < NB; i += 64) prof[i] = 0.0;
        }
        __syncthreads();
        if (prof_out)
            for (int i = lane; i < NB; i += 64) prof_out[c * ld + i] = prof[i];
        // boxcar_best_width: median baseline, circular prefix sums
        for (int i = lane; i < NB; i += 64) {
            const double v = prof[i];
            int rk = 0;
            for (int q = 0; q < NB; ++q) {
                const double u = prof[q];
                rk += (u < v) || (u == v && q < i);
            }
            srt[rk] = v;
        }
        __syncthreads();
        const double med = (NB & 1) ? srt[NB / 2] : 0.5 * (srt[NB / 2 - 1] + srt[NB / 2]);
        const int32_t* ws = wtab + woff[c];
        const int nw = woff[c + 1] - woff[c];
        const int wmax = ws[nw - 1];
        if (lane == 0) {
            ps[0] = 0.0;
            for (int i = 1; i <= NB + wmax; ++i) {
                const int idx = i > NB ? i - NB : i;
                ps[i] = ps[i - 1] + (prof[idx - 1] - med);
            }
        }
        __syncthreads();
        double best = -INFINITY;
        int bw = ws[0];
        for (int a = 0; a < nw; ++a) {
            const int w = ws[a];
            const double invsw = 1.0 / sqrt((double)w);
            double mx = -INFINITY;
            for (int p = lane; p < NB; p += 64) {
                const double d = ps[p + w] - ps[p];
                mx = d > mx ? d : mx;
            }
            mx = wave_max(mx);
            const double cand = mx * invsw;
            if (cand > best) {
                best = cand;
                bw = w;
            }
        }
        if (lane == 0) best_w[c] = bw;
        __syncthreads();
    }
}

hipError_t launch_cand_profiles(hipStream_t st, const SpecDev& sp, const double* r,
                                const int32_t* H, int64_t n, int m, const int32_t* wtab,
                                const int32_t* woff, double* prof_out, int64_t ld,
                                int32_t* best_w) {
    if (n <= 0) return hipSuccess;
    const unsigned grid = (unsigned)(n < 65536 ? n : 65536);
    hipLaunchKernelGGL(k_cand_profiles, dim3(grid), dim3(64), 0, st, sp, r, H, n, m, wtab, woff,
                       prof_out, ld, best_w);
    return hipGetLastError();
}

}  // namespace cs
