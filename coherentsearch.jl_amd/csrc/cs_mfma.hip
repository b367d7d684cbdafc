// Split-BF16 MFMA fused search kernel for gfx950.
//
// A workgroup owns 128 trials: 8 occurrences, S = lcm(16, q) trials apart, of
// one 16-trial group.  Because S is a multiple of the trial-grid period q,
// every occurrence has the same residue for every harmonic, so one 16 x K
// weight block serves all of them:
//
//   interpolation  D[16 trials][8 occ x (re, im)] = Wx[16][K] . Win[K][16]
//   c2r (per fold) E[H+1][128] = Ccos[H+1][H] . Re F[H][128]
//                  O[H+1][128] = Csin[H+1][H] . Im F[H][128]
//                  x[n] = E + O, x[nbins - n] = E - O
//
// each evaluated as v_mfma_f32_16x16x32_bf16 on bf16 hi/lo splits
// (hi.hi + hi.lo + lo.hi, FP32 accumulation).  The stacked amplitudes live in
// LDS together with one fold's profiles and its DFT tiles (copied once per
// workgroup); four lanes per trial then form the zero-baseline bound and write
// the sigma samples.  Index and phase bookkeeping stay exact integer arithmetic.
#include <hip/hip_runtime.h>

#define CS_HD __host__ __device__
#include "c2r_gen.hpp"
#include "cs_device.hpp"
#include "cs_mfma.hpp"

namespace cs {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));


__device__ __forceinline__ bf16x8 as_bf16x8(uint4 v) { return __builtin_bit_cast(bf16x8, v); }

__device__ __forceinline__ void split8(const float* x, bf16x8& hi, bf16x8& lo) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        __bf16 h = (__bf16)x[e];
        hi[e] = h;
        lo[e] = (__bf16)(x[e] - (float)h);
    }
}

__device__ __forceinline__ f32x4 mfma3(bf16x8 ah, bf16x8 al, bf16x8 bh, bf16x8 bl, f32x4 acc) {
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc, 0, 0, 0);
    return acc;
}

__host__ __device__ constexpr int nbp_of(int nb) {
    // row stride of the LDS profile tile: a multiple of 4 floats with an odd
    // number of 16-byte units, so 16 lanes' b128 accesses hit distinct banks
    return (((nb + 3) / 4) % 2 == 1) ? ((nb + 3) / 4) * 4 : ((nb + 3) / 4) * 4 + 4;
}

constexpr int kOcc8 = 8;          // occurrences per workgroup
constexpr int kWT8 = 128;         // trials per workgroup
constexpr int kThreads = 512;     // 8 waves
constexpr int kRound = 16;        // harmonics staged per round (32-bin windows)

// Zero-baseline bound of one (trial, fold) profile by four lanes: lane q owns
// bins [q*L, q*L + L) and carries the running prefix sum WM bins past its end
// (wrapping), so every boxcar that starts in its segment is formed locally.
template <int NB, int WM, int A>
__device__ __forceinline__ void seg_widths(const float* ps, int seglen, uint32_t bankmask,
                                           float& res) {
    if constexpr (A < bank_count(WM)) {
        constexpr int w = bank_w(A);
        if (bankmask & (1u << A)) {
            constexpr int L = (NB + 3) / 4;
            float mb = -INFINITY;
#pragma unroll
            for (int p = 0; p < L; ++p) {
                float d = ps[p + w] - ps[p];
                if (p < seglen) mb = d > mb ? d : mb;
            }
            float c = mb * (1.0f / sqrtf((float)w));
            res = c > res ? c : res;
        }
        __builtin_amdgcn_sched_barrier(0);   // one width at a time: bounds live registers
        seg_widths<NB, WM, A + 1>(ps, seglen, bankmask, res);
    }
}

template <int NB, int WM>
__device__ __noinline__ void seg_bound(const float* row, int q, bool active, uint32_t bankmask,
                                          int64_t j, int64_t nv, int64_t cap, float* samp,
                                          float* mo) {
    constexpr int L = (NB + 3) / 4;
    const int s0 = q * L;
    const int seglen = s0 < NB ? (NB - s0 < L ? NB - s0 : L) : 0;
    float x[L + WM];
#pragma unroll
    for (int i = 0; i < L + WM; ++i) {
        int idx = s0 + i;
        idx = idx >= NB ? idx - NB : idx;
        idx = idx >= NB ? idx - NB : idx;
        x[i] = row[idx];
    }
    // sigma samples (_block_sigma, search.jl:504-538): whole columns
    if (active) {
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
            for (int i = 0; i < L; ++i) {
                int bb = s0 + i;
                if (i < seglen && bb < take) samp[slot + bb] = fabsf(x[i]);
            }
        }
    }
    float tot = 0.0f;
#pragma unroll
    for (int i = 0; i < L; ++i)
        if (i < seglen) tot += x[i];
    // totals of the four segments -> this lane's starting offset
    const int base = (threadIdx.x & 63) & ~3;
    float t0 = __shfl(tot, base + 0, 64), t1 = __shfl(tot, base + 1, 64);
    float t2 = __shfl(tot, base + 2, 64);
    float off = q == 0 ? 0.0f : q == 1 ? t0 : q == 2 ? t0 + t1 : (t0 + t1) + t2;
    float ps[L + WM + 1];
    ps[0] = off;
#pragma unroll
    for (int i = 0; i < L + WM; ++i) ps[i + 1] = ps[i] + x[i];
    __builtin_amdgcn_sched_barrier(0);
    float res = -INFINITY;
    seg_widths<NB, WM, 0>(ps, seglen, bankmask, res);
    float r1 = __shfl_xor(res, 1, 64);
    res = r1 > res ? r1 : res;
    float r2 = __shfl_xor(res, 2, 64);
    res = r2 > res ? r2 : res;
    if (active && q == 0) *mo = res;
}

#define CS_MFMA_SIZES(X) X(16) X(20) X(24) X(30) X(32) X(40) X(60) X(64) X(120) X(128)

__host__ __device__ constexpr bool mfma_size(int nb) {
    return nb == 16 || nb == 20 || nb == 24 || nb == 30 || nb == 32 || nb == 40 || nb == 60 ||
           nb == 64 || nb == 120 || nb == 128;
}

struct MfmaLds {
    size_t ftp, region, dft, b0s, total;
};

// LDS plan: stacked amplitudes | staged windows or one fold's profiles | the
// fold's DFT tiles (hi then lo) | window starts.
__host__ __device__ inline MfmaLds mfma_lds_layout(int nharms, int nbp_max, int kwin,
                                                   int dft_tiles_max) {
    MfmaLds L;
    L.ftp = (size_t)nharms * kWT8 * 8;
    const int R = kRound / (kwin / 32);
    size_t stage = (size_t)R * 2 * kOcc8 * kwin * 4;
    size_t prof = (size_t)kWT8 * nbp_max * 4;
    L.region = stage > prof ? stage : prof;
    L.dft = (size_t)dft_tiles_max * 64 * 32;
    L.b0s = (size_t)kRound * kOcc8 * 8;
    L.total = L.ftp + L.region + L.dft + L.b0s;
    return L;
}

__global__ __launch_bounds__(kThreads, 1) void k_fused_mfma(SpecDev sp, TabDev tb, GridDev g,
                                                            const FoldSet* __restrict__ fsd,
                                                            BatchDev b, FusedDev fu,
                                                            const MfmaDev* __restrict__ mdd) {
    // fold and table descriptors are read from device memory: indexing a
    // by-value kernel argument with a runtime fold index would copy it to scratch
    const FoldSet& fs = *fsd;
    const MfmaDev& md = *mdd;
    extern __shared__ __align__(16) char smem[];
    const int nh = tb.nharms;
    const MfmaLds lay = mfma_lds_layout(nh, md.nbp_max, md.kwin, md.dft_tiles_max);
    float2* ftp = (float2*)smem;                          // [h][128 trials]
    float* region = (float*)(smem + lay.ftp);             // staged windows | fold profiles
    uint4* dft = (uint4*)(smem + lay.ftp + lay.region);   // fold DFT tiles: hi then lo
    int64_t* b0s = (int64_t*)(smem + lay.ftp + lay.region + lay.dft);
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int64_t st = blockIdx.x / md.U;
    const int u = blockIdx.x - (int)(st * md.U);
    const int64_t tbase = st * kOcc8 * md.S + (int64_t)u * 16;
    const int kwin = md.kwin;
    const int R = kRound / (kwin / 32);

    // ---------------- interpolation (split-BF16 MFMA), R harmonics per round --------
    for (int h0 = 1; h0 <= nh; h0 += R) {
        const int nr = nh - h0 + 1 < R ? nh - h0 + 1 : R;
        if (tid < nr * kOcc8) {
            const int hh = tid >> 3, o = tid & 7;
            const HarmDev d = tb.harm[h0 + hh - 1];
            uint32_t p;
            b0s[tid] = bin0(d, tb, tbase + o * md.S, p);
        }
        __syncthreads();
        const int nitems = nr * kOcc8 * kwin;
        for (int it = tid; it < nitems; it += kThreads) {
            const int k = it % kwin, ho = it / kwin;            // ho = hh*8 + o
            const int hh = ho >> 3, o = ho & 7;
            const int64_t bin = b0s[ho] + k;
            float2 v = make_float2(0.0f, 0.0f);
            if (bin >= 0 && bin < sp.namps) v = sp.amps[bin];
            region[((hh * 2 + 0) * kOcc8 + o) * kwin + k] = v.x;
            region[((hh * 2 + 1) * kOcc8 + o) * kwin + k] = v.y;
        }
        // weight tiles of this wave's harmonics (looked up before the barrier)
        int tcx[2] = {-1, -1}, tcy[2] = {-1, -1};
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const int hh = wv + 8 * r;
            if (hh < nr) {
                const int h = h0 + hh;
                tcx[r] = md.clsu[2 * ((size_t)(h - 1) * md.U + u)];
                tcy[r] = md.clsu[2 * ((size_t)(h - 1) * md.U + u) + 1];
            }
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            const int hh = wv + 8 * r;
            if (hh < nr && tcx[r] >= 0) {
                const int h = h0 + hh;
                const int ks = md.ks[h - 1];
                const int j = lane & 15, o = j >> 1, c = j & 1;
                f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
                for (int kk = 0; kk < ks; ++kk) {
                    const size_t ti = ((size_t)(tcx[r] + kk)) * 64 + lane;
                    bf16x8 ah = as_bf16x8(md.wA_hi[ti]);
                    bf16x8 al = as_bf16x8(md.wA_lo[ti]);
                    const float* src = region + ((hh * 2 + c) * kOcc8 + o) * kwin + kk * 32 + (lane >> 4) * 8;
                    float xv[8];
#pragma unroll
                    for (int e = 0; e < 8; ++e) xv[e] = src[e];
                    bf16x8 bh, bl;
                    split8(xv, bh, bl);
                    acc = mfma3(ah, al, bh, bl, acc);
                }
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    float other = __shfl_xor(acc[v], 1, 64);
                    if (c == 0) {
                        const int i = 4 * (lane >> 4) + v;
                        const int64_t t = tbase + o * md.S + i;
                        float2 a = make_float2(0.0f, 0.0f);
                        if (t < g.total) {
                            const int64_t cg = (t - g.origin) / g.Nprof;
                            if (fu.rowok[cg * nh + h - 1]) {
                                const float2 A = md.gA[tcy[r] * 16 + i];
                                const float ar = acc[v], ai = other;
                                float p1 = A.x * ar, p2 = A.y * ai, p3 = A.x * ai, p4 = A.y * ar;
                                a = make_float2(p1 - p2, p3 + p4);
                            }
                        }
                        ftp[(size_t)(h - 1) * kWT8 + o * 16 + i] = a;
                    }
                }
            }
        }
        __syncthreads();
    }

    // ---------------- c2r (even/odd split-BF16 DFT) and bounds, fold by fold --------
    {
        const int nt0 = md.c_mt[0] * md.c_ks[0] * 2;
        const size_t base0 = (size_t)md.c_base[0] * 64;
        for (int i = tid; i < nt0 * 64; i += kThreads) {
            dft[i] = md.cA_hi[base0 + i];
            dft[nt0 * 64 + i] = md.cA_lo[base0 + i];
        }
    }
    __syncthreads();
    for (int f = 0; f < fs.nf; ++f) {
        const FoldDev& fd = fs.f[f];
        const int k = fd.k, H = fd.Hk, NB = fd.nbins;
        const int NBP = nbp_of(NB);
        const int mt = md.c_mt[f], ksn = md.c_ks[f], nt = mt * ksn * 2;
        {   // wave wv: trials 16*wv .. 16*wv+15, every row tile
            const int col = lane & 15, kg = lane >> 4;
            const int tl = wv * 16 + col;
            bf16x8 rh[2], rl[2], ih[2], il[2];
#pragma unroll
            for (int kk = 0; kk < 2; ++kk) {
                if (kk < ksn) {
                    float xr[8], xi[8];
#pragma unroll
                    for (int e = 0; e < 8; ++e) {
                        const int h = kk * 32 + kg * 8 + e + 1;
                        float2 v = make_float2(0.0f, 0.0f);
                        if (h <= H) v = ftp[(size_t)(h * k - 1) * kWT8 + tl];
                        xr[e] = v.x;
                        xi[e] = v.y;
                    }
                    split8(xr, rh[kk], rl[kk]);
                    split8(xi, ih[kk], il[kk]);
                }
            }
            for (int a = 0; a < mt; ++a) {
                f32x4 E = {0.0f, 0.0f, 0.0f, 0.0f}, O = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int kk = 0; kk < 2; ++kk) {
                    if (kk < ksn) {
                        const int tcos = (a * ksn + kk) * 2, tsin = tcos + 1;
                        E = mfma3(as_bf16x8(dft[tcos * 64 + lane]), as_bf16x8(dft[(nt + tcos) * 64 + lane]),
                                  rh[kk], rl[kk], E);
                        O = mfma3(as_bf16x8(dft[tsin * 64 + lane]), as_bf16x8(dft[(nt + tsin) * 64 + lane]),
                                  ih[kk], il[kk], O);
                    }
                }
                float* dst = region + (size_t)tl * NBP;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int n = a * 16 + 4 * kg + v;
                    if (n <= H) {
                        dst[n] = E[v] + O[v];
                        if (n > 0 && n < H) dst[NB - n] = E[v] - O[v];
                    }
                }
            }
        }
        __syncthreads();
        if (f + 1 < fs.nf) {   // next fold's tiles: overlaps the bound phase
            const int nt1 = md.c_mt[f + 1] * md.c_ks[f + 1] * 2;
            const size_t base1 = (size_t)md.c_base[f + 1] * 64;
            for (int i = tid; i < nt1 * 64; i += kThreads) {
                dft[i] = md.cA_hi[base1 + i];
                dft[nt1 * 64 + i] = md.cA_lo[base1 + i];
            }
        }
        {   // bounds: four lanes per trial
            const int jq = lane >> 2, q = lane & 3;
            const int tl = wv * 16 + jq;
            const int64_t t = tbase + (tl >> 4) * md.S + (tl & 15);
            bool active = t < g.total;
            int64_t j = 0, nv = 0;
            int cl = 0;
            if (active) {
                const int64_t cg = (t - g.origin) / g.Nprof;
                const int64_t first = g.origin + cg * g.Nprof;
                cl = (int)(cg - b.c_first);
                j = t - first;
                nv = b.nvalid[cl * fs.nf + f];
                if (j >= nv) {
                    if (q == 0) fu.mraw[(int64_t)f * fu.ldm + (t - b.t_base)] = -INFINITY;
                    active = false;
                }
            }
            float* mo = fu.mraw + (int64_t)f * fu.ldm + (active ? (t - b.t_base) : 0);
            float* samp = fu.samples + ((int64_t)cl * fs.nf + f) * kSigmaSamples;
            const float* row = region + (size_t)tl * NBP;
            const bool small = fd.widths[fd.nw - 1] <= 6;
            switch (NB) {
#define X(SZ)                                                                                    \
    case SZ:                                                                                     \
        if (small && GateCT<SZ>::wmax >= 6)                                                      \
            seg_bound<SZ, 6>(row, q, active, fs.bankmask[f], j, nv, fs.cap[f], samp, mo);       \
        else                                                                                     \
            seg_bound<SZ, GateCT<SZ>::wmax>(row, q, active, fs.bankmask[f], j, nv, fs.cap[f],   \
                                            samp, mo);                                           \
        break;
                CS_MFMA_SIZES(X)
#undef X
                default:
                    break;
            }
        }
        __syncthreads();
    }
}

bool mfma_supported(const FoldSet& fs, int nharms) {
    if (nharms > 64) return false;
    for (int f = 0; f < fs.nf; ++f)
        if (!mfma_size(fs.f[f].nbins) || fs.bankmask[f] == 0) return false;
    return true;
}

size_t mfma_lds_bytes(int nharms, const FoldSet& fs, int kwin, int* nbp_max, int* dft_max,
                      const MfmaDev& md) {
    int nbp = 0, dmax = 0;
    for (int f = 0; f < fs.nf; ++f) {
        nbp = nbp_of(fs.f[f].nbins) > nbp ? nbp_of(fs.f[f].nbins) : nbp;
        int nt = md.c_mt[f] * md.c_ks[f] * 2;
        dmax = nt > dmax ? nt : dmax;
    }
    *nbp_max = nbp;
    *dft_max = dmax;
    return mfma_lds_layout(nharms, nbp, kwin, dmax).total;
}

hipError_t launch_fused_mfma(hipStream_t st, const SpecDev& sp, const TabDev& tb, const GridDev& g,
                             const FoldSet& fs, const BatchDev& b, const FusedDev& fu,
                             const MfmaDev& md, const FoldSet* fs_dev, const MfmaDev* md_dev,
                             int64_t nwg) {
    if (nwg <= 0) return hipSuccess;
    size_t shm = mfma_lds_layout(tb.nharms, md.nbp_max, md.kwin, md.dft_tiles_max).total;
    hipError_t e = hipFuncSetAttribute((const void*)k_fused_mfma,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_fused_mfma, dim3((unsigned)nwg), dim3(kThreads), shm, st, sp, tb, g, fs_dev,
                       b, fu, md_dev);
    return hipGetLastError();
}

}  // namespace cs
