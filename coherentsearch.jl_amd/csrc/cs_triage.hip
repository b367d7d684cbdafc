// Block triage: per-block incoherent-power features and a small MLP scorer
// (32 features -> 64 hidden (ReLU) -> 1 logit) whose forward and backward
// passes run on bf16 operands with FP32 accumulation.  A search can then visit
// only the blocks the scorer ranks highest (cs_search_blocks).
//
// Features of block c (trial fundamentals rs .. re, integer fundamentals
// f = floor(rs) .. floor(re)), with P = |X|^2 and
//   Q(g, f) = max of the power-pyramid cells of level s = floor(log2 g)
//             covering bins [g f, g f + g - 1]   (cells of 2^s bins),
//   S_H(f)  = sum_{g=1..H} Q(g, f),   S^(k)(f) = sum_{h=1..nharms/k} Q(h k, f):
//   [0..6]   max_f S_H / H      for H in {1, 2, 4, 8, 16, 32, nharms} (capped)
//   [7..13]  mean_f S_H / H
//   [14..20] std_f S_H / H      (population)
//   [21..25] max_f S^(k) / (nharms/k) for k = 2..6 when k is a searched fold, else 0
//   [26]     mean_f P(f)
//   [27]     mean_f Q(nharms, f)
//   [28]     log2(max(rs, 1)) / 20
//   [29]     fraction of f with S_nharms / nharms > 1.5
//   [30]     median_f P(f)                       (local noise level)
//   [31]     [6] / max(median_f S_nharms / nharms, 1e-30)   (peak contrast)
// The medians are the lower medians over the sampled fundamentals
// f = floor(rs) + i * stride, stride = ceil(nfund / 512) (every fundamental
// when a block has at most 512).
//
// Kind 1 keeps that layout with incoherent summed power at the harmonics of
// every trial of the block instead: the units are the block's trials
// r_t = (r_lo + i0 lodr) + j lodr (j = 0 .. n - 1, FP64, the search's own
// expression order) and Q(g, t) = P(nint(g r_t)), nint(x) = floor(x + 1/2) in
// FP64, so S_H(t) = sum_{h=1..H} P(nint(h r_t)); P(f) becomes P(nint(r_t)),
// and the medians sample every stride-th trial.
#include <hip/hip_runtime.h>

#include <cmath>

#include "cs_triage.hpp"

namespace cs {

namespace {

constexpr int kPyrLevels = 7;   // levels 0 (|X|^2) .. 6 (cells of 64 bins)

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf16r(float x) { return (float)(__bf16)x; }

// level 1 of the pyramid from the amplitudes, then each level from the last
__global__ void k_pyr_first(const float2* __restrict__ amps, int64_t namps, float* __restrict__ l1,
                            int64_t n1) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n1) return;
    float m = 0.0f;
    for (int e = 0; e < 2; ++e) {
        const int64_t b = 2 * i + e;
        if (b < namps) {
            const float2 v = amps[b];
            m = fmaxf(m, v.x * v.x + v.y * v.y);
        }
    }
    l1[i] = m;
}

__global__ void k_pyr_next(const float* __restrict__ prev, int64_t nprev, float* __restrict__ next,
                           int64_t nnext) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nnext) return;
    const float a = 2 * i < nprev ? prev[2 * i] : 0.0f;
    const float b = 2 * i + 1 < nprev ? prev[2 * i + 1] : 0.0f;
    next[i] = fmaxf(a, b);
}

struct PyrDev {
    const float2* amps;
    int64_t namps;
    const float* lev[kPyrLevels];   // lev[0] unused (|X|^2 read from amps)
    int64_t len[kPyrLevels];
};

__device__ __forceinline__ float power_at(const PyrDev& p, int64_t b) {
    if (b < 0 || b >= p.namps) return 0.0f;
    const float2 v = p.amps[b];
    return v.x * v.x + v.y * v.y;
}

__device__ __forceinline__ float q_of(const PyrDev& p, int g, int64_t f) {
    if (g == 1) return power_at(p, f);
    const int s = 31 - __builtin_clz((unsigned)g);
    const int64_t lo = (int64_t)g * f, hi = lo + g - 1;
    const int64_t c0 = lo >> s, c1 = hi >> s;
    const float* L = p.lev[s];
    const int64_t n = p.len[s];
    float m = 0.0f;
    for (int64_t cc = c0; cc <= c1; ++cc)
        if (cc >= 0 && cc < n) m = fmaxf(m, L[cc]);
    return m;
}

__device__ __forceinline__ float wave_sum(float v) {
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
    for (int o = 32; o >= 1; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

constexpr int kMedSamples = 512;   // fundamentals sampled for the medians

// lower median ((n - 1) / 2-th smallest) of n <= kMedSamples non-negative
// floats in LDS, one wave; n wave-uniform.  Up to 64 values: the lane whose
// rank (ties broken by position) is (n - 1) / 2.  More: the smallest key x with
// #(keys <= x) > (n - 1) / 2, keys ordered as their bits.
__device__ float wave_lower_median(const float* v, int n, int lane) {
    if (n <= 0) return 0.0f;
    const int k = (n - 1) / 2;
    if (n <= 64) {
        const float mine = lane < n ? v[lane] : 0.0f;
        int rank = 0;
        for (int j = 0; j < n; ++j) {
            const float o = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(mine), j));
            rank += (o < mine || (o == mine && j < lane)) ? 1 : 0;
        }
        const uint64_t hit = __ballot(lane < n && rank == k);
        return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(mine), __ffsll((long long)hit) - 1));
    }
    uint32_t lo = 0, hi = 0x7F800000u;   // +inf bounds every non-negative float
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2;
        float cnt = 0.0f;
        for (int i = lane; i < n; i += 64) cnt += (__float_as_uint(v[i]) <= mid) ? 1.0f : 0.0f;
        if (wave_sum(cnt) > (float)k) hi = mid;
        else lo = mid + 1;
    }
    return __uint_as_float(lo);
}

// one wave per search block; lanes over its integer fundamentals (KIND 0) or
// its trials (KIND 1)
template <int KIND>
__global__ __launch_bounds__(256) void k_triage_features(PyrDev p, double r_lo, double lodr,
                                                         int64_t Nprof, int64_t total, int nh,
                                                         uint32_t foldmask, int64_t nblocks,
                                                         float* __restrict__ out) {
    __shared__ float med_p[4][kMedSamples], med_s[4][kMedSamples];
    const int64_t blk = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (blk >= nblocks) return;
    const int64_t i0 = blk * Nprof;
    const int64_t n = min(Nprof, total - i0);
    const double rs = r_lo + (double)i0 * lodr;
    const double re = rs + (double)(n - 1) * lodr;
    const int64_t f0 = (int64_t)floor(rs), f1 = (int64_t)floor(re);
    const int64_t nfund = KIND == 0 ? f1 - f0 + 1 : n;   // units: fundamentals or trials
    const int64_t stride = (nfund + kMedSamples - 1) / kMedSamples;
    const int nsamp = (int)((nfund + stride - 1) / stride);
    int Hs[7] = {1, 2, 4, 8, 16, 32, nh};
    for (int j = 0; j < 7; ++j) Hs[j] = min(Hs[j], nh);
    float smax[7], ssum[7], ssq[7], kmax[5];
    for (int j = 0; j < 7; ++j) smax[j] = ssum[j] = ssq[j] = 0.0f;
    for (int j = 0; j < 5; ++j) kmax[j] = 0.0f;
    float pmean = 0.0f, qtop = 0.0f, over = 0.0f;
    for (int64_t base = 0; base < nfund; base += 64) {
        const int64_t idx = base + lane;
        const bool on = idx < nfund;
        const int64_t f = f0 + (on ? idx : 0);
        const double rt = rs + (double)(on ? idx : 0) * lodr;
        float S[7], Sk[5];
        for (int j = 0; j < 5; ++j) Sk[j] = 0.0f;
        float acc = 0.0f, qlast = 0.0f, q1 = 0.0f;
        int hj = 0;
        for (int g = 1; g <= nh; ++g) {
            const float q = KIND == 0 ? q_of(p, g, f)
                                      : power_at(p, (int64_t)floor((double)g * rt + 0.5));
            if (g == 1) q1 = q;
            acc += q;
            qlast = q;
            for (int k = 2; k <= 6; ++k)
                if (((foldmask >> k) & 1u) && g % k == 0 && g / k <= nh / k) Sk[k - 2] += q;
            while (hj < 7 && Hs[hj] == g) S[hj++] = acc;
        }
        if (on) {
            for (int j = 0; j < 7; ++j) {
                const float v = S[j] / (float)Hs[j];
                smax[j] = fmaxf(smax[j], v);
                ssum[j] += v;
                ssq[j] += v * v;
            }
            for (int k = 2; k <= 6; ++k)
                if ((foldmask >> k) & 1u) kmax[k - 2] = fmaxf(kmax[k - 2], Sk[k - 2] / (float)(nh / k));
            const float p1 = q1;   // P(f), or P(nint(r_t))
            pmean += p1;
            qtop += qlast;
            over += (S[6] / (float)Hs[6] > 1.5f) ? 1.0f : 0.0f;
            if (idx % stride == 0) {
                med_p[wv][idx / stride] = p1;
                med_s[wv][idx / stride] = S[6] / (float)Hs[6];
            }
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const float pmed = wave_lower_median(med_p[wv], nsamp, lane);
    const float smed = wave_lower_median(med_s[wv], nsamp, lane);
    const float inv = 1.0f / (float)nfund;
    float F[32];
    for (int j = 0; j < 7; ++j) {
        const float mx = wave_max(smax[j]), mean = wave_sum(ssum[j]) * inv;
        const float var = fmaxf(wave_sum(ssq[j]) * inv - mean * mean, 0.0f);
        F[j] = mx;
        F[7 + j] = mean;
        F[14 + j] = sqrtf(var);
    }
    for (int j = 0; j < 5; ++j) F[21 + j] = wave_max(kmax[j]);
    F[26] = wave_sum(pmean) * inv;
    F[27] = wave_sum(qtop) * inv;
    F[28] = (float)(log2(fmax(rs, 1.0)) / 20.0);
    F[29] = wave_sum(over) * inv;
    F[30] = pmed;
    F[31] = F[6] / fmaxf(smed, 1e-30f);
    if (lane < 32) {
        float v = 0.0f;
        for (int j = 0; j < 32; ++j) v = (lane == j) ? F[j] : v;
        out[blk * kTriageFeatures + lane] = v;
    }
}

// ---------------------------------------------------------------------------
// MLP.  params: W1 [64 hidden][32 in], b1 [64], w2 [64], b2 [1] (FP32 master
// copy); norm: mean [32], inverse std [32].  Each wave takes 32 blocks (two
// 16-row MFMA tiles).  Layer 1 on v_mfma_f32_16x16x32_bf16 (bf16 inputs and
// weights, FP32 accumulation); layer 2 and the gradients multiply bf16-rounded
// operands and accumulate in FP32, dW1 again on the MFMA (K = the wave's 32
// blocks).  Per-wave gradient partials are reduced in a fixed order.
// ---------------------------------------------------------------------------
constexpr int kIn = kTriageFeatures, kHid = kTriageHidden;
constexpr int kWaveBlocks = 32;

__device__ __forceinline__ bf16x8 pack8(const float* v) {
    bf16x8 r;
    for (int e = 0; e < 8; ++e) r[e] = (__bf16)v[e];
    return r;
}

__global__ __launch_bounds__(256) void k_mlp(const float* __restrict__ params,
                                             const float* __restrict__ norm,
                                             const float* __restrict__ X,
                                             const float* __restrict__ y, int64_t n,
                                             float* __restrict__ logits,
                                             float* __restrict__ partial) {
    __shared__ float sx[4][kWaveBlocks][kIn + 1];     // standardised, bf16-rounded
    __shared__ float sdh[4][kWaveBlocks][kHid + 1];   // dL/dh (bf16-rounded)
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int col = lane & 15, g = lane >> 4;
    const int64_t wave_id = (int64_t)blockIdx.x * 4 + wv;
    const int64_t b0 = wave_id * kWaveBlocks;
    const float* W1 = params;
    const float* b1 = params + kHid * kIn;
    const float* w2 = b1 + kHid;
    const float b2 = w2[kHid];
    // standardised inputs
    for (int i = lane; i < kWaveBlocks * kIn; i += 64) {
        const int r = i / kIn, k = i % kIn;
        const int64_t row = b0 + r;
        float v = 0.0f;
        if (row < n) v = (X[row * kIn + k] - norm[k]) * norm[kIn + k];
        sx[wv][r][k] = bf16r(v);
    }
    __syncthreads();
    // layer 1: for M tile m (blocks 16m..), hidden tile t: D[i][j] = sum_k x[i][k] W1[16t+j][k]
    float h[2][4][4];   // [m][t][v]: block 16m + 4g + v, hidden 16t + col
    for (int m = 0; m < 2; ++m) {
        float xa[8];
        for (int e = 0; e < 8; ++e) xa[e] = sx[wv][16 * m + col][8 * g + e];
        const bf16x8 A = pack8(xa);
        for (int t = 0; t < 4; ++t) {
            float wb[8];
            for (int e = 0; e < 8; ++e) wb[e] = W1[(16 * t + col) * kIn + 8 * g + e];
            const bf16x8 B = pack8(wb);
            f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B, acc, 0, 0, 0);
            for (int v = 0; v < 4; ++v) h[m][t][v] = acc[v] + b1[16 * t + col];
        }
    }
    // layer 2: z = sum_j bf16(w2_j) bf16(relu(h_j)) + b2, lanes of one g hold hidden j = col
    float z[2][4];
    for (int m = 0; m < 2; ++m)
        for (int v = 0; v < 4; ++v) {
            float s = 0.0f;
            for (int t = 0; t < 4; ++t)
                s += bf16r(w2[16 * t + col]) * bf16r(fmaxf(h[m][t][v], 0.0f));
            for (int o = 8; o >= 1; o >>= 1) s += __shfl_xor(s, o, 64);
            z[m][v] = s + b2;
        }
    if (logits && col == 0)
        for (int m = 0; m < 2; ++m)
            for (int v = 0; v < 4; ++v) {
                const int64_t row = b0 + 16 * m + 4 * g + v;
                if (row < n) logits[row] = z[m][v];
            }
    if (!partial) return;
    // backward of the mean binary cross-entropy
    float gw2[4] = {0.0f, 0.0f, 0.0f, 0.0f}, gb1[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    float gb2 = 0.0f, loss = 0.0f;
    const float invn = 1.0f / (float)n;
    for (int m = 0; m < 2; ++m)
        for (int v = 0; v < 4; ++v) {
            const int r = 16 * m + 4 * g + v;
            const int64_t row = b0 + r;
            const bool ok = row < n;
            const float zz = z[m][v];
            const float yy = ok ? y[row] : 0.0f;
            const float pr = 1.0f / (1.0f + expf(-zz));
            const float dz = ok ? (pr - yy) * invn : 0.0f;
            // log(1 + e^-|z|) form of the loss
            const float l = fmaxf(zz, 0.0f) - zz * yy + log1pf(expf(-fabsf(zz)));
            if (ok && col == 0) loss += l * invn;
            if (col == 0) gb2 += dz;
            for (int t = 0; t < 4; ++t) {
                const float a = bf16r(fmaxf(h[m][t][v], 0.0f));
                gw2[t] += dz * a;
                const float dh = h[m][t][v] > 0.0f ? dz * bf16r(w2[16 * t + col]) : 0.0f;
                gb1[t] += dh;
                sdh[wv][r][16 * t + col] = bf16r(dh);
            }
        }
    __syncthreads();
    // dW1[j][k] = sum_r dh[r][j] x[r][k]: A = dH^T tile (hidden rows, K = blocks),
    // B = X tile (K = blocks, feature columns)
    float gw1[4][2][4];   // [hidden tile][feature tile][v]
    for (int t = 0; t < 4; ++t) {
        float av[8];
        for (int e = 0; e < 8; ++e) av[e] = sdh[wv][8 * g + e][16 * t + col];
        const bf16x8 A = pack8(av);
        for (int u = 0; u < 2; ++u) {
            float bv[8];
            for (int e = 0; e < 8; ++e) bv[e] = sx[wv][8 * g + e][16 * u + col];
            const bf16x8 B = pack8(bv);
            f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(A, B, acc, 0, 0, 0);
            for (int v = 0; v < 4; ++v) gw1[t][u][v] = acc[v];
        }
    }
    // partial gradients of this wave, in the params layout, then the loss
    float* P = partial + wave_id * (kTriageParams + 1);
    for (int t = 0; t < 4; ++t)
        for (int u = 0; u < 2; ++u)
            for (int v = 0; v < 4; ++v)   // D[i][j]: hidden 16t + 4g + v, feature 16u + col
                P[(16 * t + 4 * g + v) * kIn + 16 * u + col] = gw1[t][u][v];
    for (int t = 0; t < 4; ++t) {
        float s1 = gb1[t], s2 = gw2[t];
        s1 += __shfl_xor(s1, 16, 64);
        s1 += __shfl_xor(s1, 32, 64);
        s2 += __shfl_xor(s2, 16, 64);
        s2 += __shfl_xor(s2, 32, 64);
        if (g == 0) {
            P[kHid * kIn + 16 * t + col] = s1;
            P[kHid * kIn + kHid + 16 * t + col] = s2;
        }
    }
    gb2 = wave_sum(gb2);
    loss = wave_sum(loss);
    if (lane == 0) {
        P[kTriageParams - 1] = gb2;
        P[kTriageParams] = loss;
    }
}

__global__ void k_mlp_reduce(const float* __restrict__ partial, int64_t nwaves,
                             float* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > kTriageParams) return;
    double s = 0.0;
    for (int64_t w = 0; w < nwaves; ++w) s += (double)partial[w * (kTriageParams + 1) + i];
    out[i] = (float)s;
}

}  // namespace

int64_t triage_pyramid_floats(int64_t namps) {
    int64_t tot = 0, len = namps;
    for (int s = 1; s < kPyrLevels; ++s) {
        len = (len + 1) / 2;
        tot += len;
    }
    return tot;
}

hipError_t launch_triage_features(hipStream_t st, int kind, const float2* amps, int64_t namps,
                                  float* pyr, double r_lo, double lodr, int64_t Nprof,
                                  int64_t total, int nh, uint32_t foldmask, int64_t nblocks,
                                  float* out) {
    PyrDev p{};
    p.amps = amps;
    p.namps = namps;
    if (kind == 1) {   // harmonic sums at nint(h r_t): no pyramid
        if (nblocks > 0)
            hipLaunchKernelGGL(k_triage_features<1>, dim3((unsigned)((nblocks + 3) / 4)), dim3(256),
                               0, st, p, r_lo, lodr, Nprof, total, nh, foldmask, nblocks, out);
        return hipGetLastError();
    }
    int64_t len = namps, off = 0;
    for (int s = 1; s < kPyrLevels; ++s) {
        const int64_t nl = (len + 1) / 2;
        p.lev[s] = pyr + off;
        p.len[s] = nl;
        if (s == 1)
            hipLaunchKernelGGL(k_pyr_first, dim3((unsigned)((nl + 255) / 256)), dim3(256), 0, st,
                               amps, namps, pyr + off, nl);
        else
            hipLaunchKernelGGL(k_pyr_next, dim3((unsigned)((nl + 255) / 256)), dim3(256), 0, st,
                               p.lev[s - 1], len, pyr + off, nl);
        off += nl;
        len = nl;
    }
    if (nblocks > 0)
        hipLaunchKernelGGL(k_triage_features<0>, dim3((unsigned)((nblocks + 3) / 4)), dim3(256), 0,
                           st, p, r_lo, lodr, Nprof, total, nh, foldmask, nblocks, out);
    return hipGetLastError();
}

int64_t mlp_waves(int64_t n) { return (n + kWaveBlocks - 1) / kWaveBlocks; }

hipError_t launch_mlp(hipStream_t st, const float* params, const float* norm, const float* X,
                      const float* y, int64_t n, float* logits, float* partial, float* grad) {
    const int64_t nw = mlp_waves(n);
    if (nw <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_mlp, dim3((unsigned)((nw + 3) / 4)), dim3(256), 0, st, params, norm, X, y,
                       n, logits, partial);
    if (partial && grad)
        hipLaunchKernelGGL(k_mlp_reduce, dim3((kTriageParams + 1 + 255) / 256), dim3(256), 0, st,
                           partial, mlp_partial_rows(n), grad);
    return hipGetLastError();
}

// One Adam step (FP32) of the scorer's parameters from the gradient the MLP
// pass left in grad[0 .. n) (grad[n] = that pass's loss, kept in losses[it]).
// c1, c2 = 1 - beta1^t, 1 - beta2^t of step t.
__global__ void k_adam(float* __restrict__ p, float* __restrict__ m1, float* __restrict__ m2,
                       const float* __restrict__ grad, int n, float lr, float b1, float b2,
                       float eps, float c1, float c2, float* __restrict__ losses, int it) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0 && losses) losses[it] = grad[n];
    if (i >= n) return;
    const float g = grad[i];
    const float a = b1 * m1[i] + (1.0f - b1) * g;
    const float v = b2 * m2[i] + (1.0f - b2) * (g * g);
    m1[i] = a;
    m2[i] = v;
    p[i] = p[i] - lr * (a / c1) / (sqrtf(v / c2) + eps);
}

hipError_t launch_adam(hipStream_t st, float* p, float* m1, float* m2, const float* grad, int n,
                       float lr, float b1, float b2, float eps, int step, float* losses, int it) {
    const float c1 = (float)(1.0 - std::pow((double)b1, step));
    const float c2 = (float)(1.0 - std::pow((double)b2, step));
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, m1, m2, grad,
                       n, lr, b1, b2, eps, c1, c2, losses, it);
    return hipGetLastError();
}

}  // namespace cs
