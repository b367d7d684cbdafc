// Split-BF16 MFMA fused search kernel for gfx950.
//
// A workgroup owns 64 trials: 4 occurrences, S = lcm(16, q) trials apart, of
// one 16-trial group.  Because S is a multiple of the trial-grid period q,
// every occurrence has the same residue for every harmonic, so one 16 x K
// weight block serves all of them:
//
//   interpolation  D[16 trials][4 occ x (re, im)] = Wx[16][K] . Win[K][8]
//   c2r (per fold) P[nbins][64 trials]             = C[nbins][2H] . F[2H][64]
//
// each evaluated as v_mfma_f32_16x16x32_bf16 on bf16 hi/lo splits
// (hi.hi + hi.lo + lo.hi, FP32 accumulation).  The stacked amplitudes live in
// LDS; one fold's profiles at a time are written to LDS for the zero-baseline
// bound and the sigma samples (the same per-trial logic as the VALU tile
// kernel).  Index and phase bookkeeping stay exact integer arithmetic.
#include <hip/hip_runtime.h>

#define CS_HD __host__ __device__
#include "c2r_gen.hpp"
#include "cs_device.hpp"
#include "cs_mfma.hpp"

namespace cs {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kOcc = 4;
constexpr int kWT = 64;          // trials per workgroup
constexpr int kWin = 64;         // staged bins per (harmonic, occurrence): <= 2 K-steps

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

// Zero-baseline bound and sigma samples of one valid (trial, fold) profile
// held in registers (same logic as the VALU tile kernel).
template <int NB, int WM, int A>
__device__ __forceinline__ void bound_widths(const float* ps, uint32_t bankmask, float& res) {
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
        bound_widths<NB, WM, A + 1>(ps, bankmask, res);
    }
}

template <int NB, int WM>
__device__ __forceinline__ void bound_from_lds(const float* row, uint32_t bankmask, int64_t j,
                                               int64_t nv, int64_t cap, float* samp,
                                               float* mo) {
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
        for (int i = 0; i < take; ++i) samp[slot + i] = fabsf(row[i]);
    }
    // prefix sums straight from LDS; only ps lives in registers
    float ps[NB + WM + 1];
    ps[0] = 0.0f;
#pragma unroll
    for (int i = 0; i < NB; i += 4) {
        float4 v = *(const float4*)(row + i);
        if (i + 0 < NB) ps[i + 1] = ps[i] + v.x;
        if (i + 1 < NB) ps[i + 2] = ps[i + 1] + v.y;
        if (i + 2 < NB) ps[i + 3] = ps[i + 2] + v.z;
        if (i + 3 < NB) ps[i + 4] = ps[i + 3] + v.w;
    }
#pragma unroll
    for (int i = NB + 1; i <= NB + WM; ++i) ps[i] = ps[i - 1] + row[i - NB - 1];
    float res = -INFINITY;
    bound_widths<NB, WM, 0>(ps, bankmask, res);
    *mo = res;
}

#define CS_MFMA_SIZES(X) X(16) X(20) X(24) X(30) X(32) X(40) X(60) X(64) X(120) X(128)

__global__ __launch_bounds__(256, 2) void k_fused_mfma(SpecDev sp, TabDev tb, GridDev g, FoldSet fs,
                                                       BatchDev b, FusedDev fu, MfmaDev md) {
    extern __shared__ __align__(16) char smem[];
    const int nh = tb.nharms;
    float2* ftp = (float2*)smem;                                   // [h][64 trials]
    float* stage = (float*)(smem + (size_t)nh * kWT * sizeof(float2));  // windows / profiles
    float* prof = stage;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int64_t st = blockIdx.x / md.U;
    const int u = blockIdx.x - (int)(st * md.U);
    const int64_t tbase = st * kOcc * md.S + (int64_t)u * 16;

    // ---------------- interpolation: 4 harmonics per pass, one per wave ----------
    for (int h0 = 1; h0 <= nh; h0 += 4) {
        {   // stage the windows: wave wv -> harmonic h0 + wv, lane -> (occurrence, 8-bin segment)
            const int h = h0 + wv;
            if (h <= nh) {
                const HarmDev d = tb.harm[h - 1];
                const int o = lane >> 4, seg = lane & 15;     // 4 occ x 16 segments of 4 bins
                const int64_t t0 = tbase + o * md.S;
                uint32_t p;
                const int64_t b0 = bin0(d, tb, t0, p);
                float* re = stage + ((wv * 2 + 0) * kOcc + o) * kWin;
                float* im = stage + ((wv * 2 + 1) * kOcc + o) * kWin;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int k = seg * 4 + e;
                    const int64_t bin = b0 + k;
                    float2 v = make_float2(0.0f, 0.0f);
                    if (bin >= 0 && bin < sp.namps) v = sp.amps[bin];
                    re[k] = v.x;
                    im[k] = v.y;
                }
            }
        }
        __syncthreads();
        {
            const int h = h0 + wv;
            if (h <= nh) {
                const HarmDev d = tb.harm[h - 1];
                int64_t qint;
                uint32_t res;
                cstate(tbase, (uint32_t)d.hp, (uint32_t)d.q, qint, res);
                const int cls = md.clsmap[(size_t)(h - 1) * md.q + res];
                const int ks = cls >= 0 ? md.ks[h - 1] : 0;
                const int j = lane & 15, o = (j >> 1), c = j & 1;
                f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
                for (int kk = 0; kk < ks; ++kk) {
                    const size_t ti = ((size_t)(md.cls_tile[cls] + kk)) * 64 + lane;
                    bf16x8 ah = as_bf16x8(md.wA_hi[ti]);
                    bf16x8 al = as_bf16x8(md.wA_lo[ti]);
                    bf16x8 bh, bl;
                    float xv[8];
                    if (o < kOcc) {
                        const float* src = stage + ((wv * 2 + c) * kOcc + o) * kWin + kk * 32 + (lane >> 4) * 8;
#pragma unroll
                        for (int e = 0; e < 8; ++e) xv[e] = src[e];
                    } else {
#pragma unroll
                        for (int e = 0; e < 8; ++e) xv[e] = 0.0f;
                    }
                    split8(xv, bh, bl);
                    acc = mfma3(ah, al, bh, bl, acc);
                }
                // lane holds D[i = 4*(lane/16) + v][j]; pair re (even j) with im (odd j)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    float other = __shfl_xor(acc[v], 1, 64);
                    if (c == 0 && o < kOcc && cls >= 0) {
                        const int i = 4 * (lane >> 4) + v;
                        const int64_t t = tbase + o * md.S + i;
                        float2 a = make_float2(0.0f, 0.0f);
                        if (t < g.total) {
                            const int64_t cg = (t - g.origin) / g.Nprof;
                            if (fu.rowok[cg * nh + h - 1]) {
                                const float2 A = md.gA[cls * 16 + i];
                                const float ar = acc[v], ai = other;
                                float p1 = A.x * ar, p2 = A.y * ai, p3 = A.x * ai, p4 = A.y * ar;
                                a = make_float2(p1 - p2, p3 + p4);
                            }
                        }
                        ftp[(size_t)(h - 1) * kWT + o * 16 + i] = a;
                    }
                }
            }
        }
        __syncthreads();
    }

    // ---------------- c2r and bounds, one fold at a time ----------------------------
    for (int f = 0; f < fs.nf; ++f) {
        const FoldDev& fd = fs.f[f];
        const int k = fd.k, H = fd.Hk, NB = fd.nbins;
        const int NBP = nbp_of(NB);
        const int mt = md.c_mt[f], ksn = md.c_ks[f];
        {   // wave wv: trials 16*wv .. 16*wv+15 (one N tile), every M tile
            const int col = lane & 15, kg = lane >> 4;
            const int tl = wv * 16 + col;
            bf16x8 bh[4], bl[4];
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                if (kk < ksn) {
                    float xv[8];
#pragma unroll
                    for (int e2 = 0; e2 < 4; ++e2) {
                        const int h = kk * 16 + kg * 4 + e2 + 1;
                        float2 v = make_float2(0.0f, 0.0f);
                        if (h <= H) v = ftp[(size_t)(h * k - 1) * kWT + tl];
                        xv[2 * e2] = v.x;
                        xv[2 * e2 + 1] = v.y;
                    }
                    split8(xv, bh[kk], bl[kk]);
                }
            }
            for (int a = 0; a < mt; ++a) {
                f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int kk = 0; kk < 4; ++kk) {
                    if (kk < ksn) {
                        const size_t ti = ((size_t)(md.c_base[f] + a * ksn + kk)) * 64 + lane;
                        acc = mfma3(as_bf16x8(md.cA_hi[ti]), as_bf16x8(md.cA_lo[ti]), bh[kk], bl[kk], acc);
                    }
                }
                const int n0 = a * 16 + 4 * kg;
                float* dst = prof + (size_t)tl * NBP + n0;
                if (n0 + 4 <= NB) {
                    *(float4*)dst = make_float4(acc[0], acc[1], acc[2], acc[3]);
                } else {
#pragma unroll
                    for (int v = 0; v < 4; ++v)
                        if (n0 + v < NB) dst[v] = acc[v];
                }
            }
        }
        __syncthreads();
        if (wv == (f & 3)) {    // one wave scores the fold's 64 trials (lane = trial)
            const int tl = lane;
            const int64_t t = tbase + (tl >> 4) * md.S + (tl & 15);
            if (t < g.total) {
                const int64_t cg = (t - g.origin) / g.Nprof;
                const int64_t first = g.origin + cg * g.Nprof;
                const int cl = (int)(cg - b.c_first);
                const int64_t j = t - first;
                const int64_t nv = b.nvalid[cl * fs.nf + f];
                float* mo = fu.mraw + (int64_t)f * fu.ldm + (t - b.t_base);
                if (j >= nv) {
                    *mo = -INFINITY;
                } else {
                    float* samp = fu.samples + ((int64_t)cl * fs.nf + f) * kSigmaSamples;
                    const float* row = prof + (size_t)tl * NBP;
                    const bool small = fd.widths[fd.nw - 1] <= 6;
                    switch (NB) {
#define X(SZ)                                                                                  \
    case SZ:                                                                                   \
        if (small && GateCT<SZ>::wmax >= 6)                                                    \
            bound_from_lds<SZ, 6>(row, fs.bankmask[f], j, nv, fs.cap[f], samp, mo);            \
        else                                                                                   \
            bound_from_lds<SZ, GateCT<SZ>::wmax>(row, fs.bankmask[f], j, nv, fs.cap[f], samp, mo); \
        break;
                        CS_MFMA_SIZES(X)
#undef X
                        default:
                            break;
                    }
                }
            }
        }
        __syncthreads();
    }
}

bool mfma_supported(const FoldSet& fs, int nharms) {
    if (nharms > 64) return false;
    for (int f = 0; f < fs.nf; ++f) {
        const int nb = fs.f[f].nbins;
        bool okz = nb == 16 || nb == 20 || nb == 24 || nb == 30 || nb == 32 || nb == 40 ||
                   nb == 60 || nb == 64 || nb == 120 || nb == 128;
        if (!okz || fs.bankmask[f] == 0) return false;
    }
    return true;
}

size_t mfma_lds_bytes(int nharms, const FoldSet& fs) {
    size_t prof = 0;
    for (int f = 0; f < fs.nf; ++f) {
        size_t p = (size_t)kWT * nbp_of(fs.f[f].nbins) * sizeof(float);
        prof = p > prof ? p : prof;
    }
    size_t stage = (size_t)4 * 2 * kOcc * kWin * sizeof(float);
    return (size_t)nharms * kWT * sizeof(float2) + (prof > stage ? prof : stage);
}

hipError_t launch_fused_mfma(hipStream_t st, const SpecDev& sp, const TabDev& tb, const GridDev& g,
                             const FoldSet& fs, const BatchDev& b, const FusedDev& fu,
                             const MfmaDev& md, int64_t nwg) {
    if (nwg <= 0) return hipSuccess;
    size_t shm = mfma_lds_bytes(tb.nharms, fs);
    hipError_t e = hipFuncSetAttribute((const void*)k_fused_mfma,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_fused_mfma, dim3((unsigned)nwg), dim3(256), shm, st, sp, tb, g, fs, b, fu,
                       md);
    return hipGetLastError();
}

}  // namespace cs
