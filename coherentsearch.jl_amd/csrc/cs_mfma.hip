// Split-BF16 MFMA fused search kernel for gfx950 (persistent workgroups).
//
// A tile is 128 trials: 8 occurrences, S = lcm(16, q) trials apart, of one
// 16-trial group u.  Because S is a multiple of the trial-grid period q, every
// occurrence has the same residue for every harmonic, so one 16 x K weight
// block (per harmonic and group) serves all eight, and the window start moves
// by an exact integer number of bins per occurrence:
//
//   interpolation  D[16 trials][8 occ x (re, im)] = Wx[16][K] . Win[K][16]
//   c2r (per fold) E[n][128] = Ccos[n][h] . Re F[h][128]
//                  O[n][128] = Csin[n][h] . Im F[h][128]
//                  bin n = E + O, bin NB - n = E - O
//
// each a v_mfma_f32_16x16x32_bf16 chain on bf16 hi/lo splits (hi.hi + hi.lo +
// lo.hi, FP32 accumulation).  The spectrum is kept pre-split as (hi | lo << 16)
// words in two planes, so a lane's window is two 16-byte loads and eight
// byte-permutes; the stacked amplitudes go to LDS pre-split the same way.  The
// c2r rows are permuted so each of the four lanes of a trial column holds a
// contiguous run of the profile, and the boxcar bound is formed from registers
// (cs_geo.hpp).  LDS holds the stacked amplitudes and every fold's DFT tiles;
// a workgroup loops over tiles (XCD-contiguous ranges), two barriers per tile.
// Index and phase bookkeeping stay exact integer arithmetic.
#include <hip/hip_runtime.h>

#include "cs_device.hpp"
#include "cs_geo.hpp"
#include "cs_mfma.hpp"

namespace cs {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint4 uint4_a4 __attribute__((aligned(4)));

constexpr int kOcc8 = 8;          // occurrences per tile
constexpr int kWT = 128;          // trials per tile
constexpr int kThreads = 512;     // 8 waves

__device__ __forceinline__ bf16x8 as_bf16x8(uint4 v) { return __builtin_bit_cast(bf16x8, v); }

// (hi | lo << 16) words -> hi and lo operand vectors (element e from word e)
__device__ __forceinline__ void unpack8(const uint32_t* w, bf16x8& hi, bf16x8& lo) {
    uint4 h, l;
    h.x = __builtin_amdgcn_perm(w[1], w[0], 0x05040100u);
    h.y = __builtin_amdgcn_perm(w[3], w[2], 0x05040100u);
    h.z = __builtin_amdgcn_perm(w[5], w[4], 0x05040100u);
    h.w = __builtin_amdgcn_perm(w[7], w[6], 0x05040100u);
    l.x = __builtin_amdgcn_perm(w[1], w[0], 0x07060302u);
    l.y = __builtin_amdgcn_perm(w[3], w[2], 0x07060302u);
    l.z = __builtin_amdgcn_perm(w[5], w[4], 0x07060302u);
    l.w = __builtin_amdgcn_perm(w[7], w[6], 0x07060302u);
    hi = as_bf16x8(h);
    lo = as_bf16x8(l);
}

__device__ __forceinline__ uint32_t split_word(float x) {
    const __bf16 h = (__bf16)x;
    const __bf16 l = (__bf16)(x - (float)h);
    return (uint32_t)__builtin_bit_cast(uint16_t, h) |
           ((uint32_t)__builtin_bit_cast(uint16_t, l) << 16);
}

__device__ __forceinline__ f32x4 mfma3(bf16x8 ah, bf16x8 al, bf16x8 bh, bf16x8 bl, f32x4 acc) {
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc, 0, 0, 0);
    return acc;
}

// value of the neighbouring lane of an (even, odd) lane pair
__device__ __forceinline__ float pair_swap(float x) {
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, x), 0xB1,
                                                              0xF, 0xF, true));
}

// column swizzle of the stacked-amplitude planes: spreads the rows one c2r
// operand read touches over distinct LDS banks
__device__ __forceinline__ int swz(int r) { return ((r >> 3) & 3) << 4; }

// ---------------------------------------------------------------------------
// Profile sequences from registers (cs_geo.hpp)
// ---------------------------------------------------------------------------
template <int KIND, int REG, int R>
__device__ __forceinline__ float reg_of(const float (&P)[R], const float (&M)[R]) {
    if constexpr (KIND == 0) return P[REG];
    else return M[REG];
}

template <class G, int W, int I, int RD>
__device__ __forceinline__ float fetch_round(const float (&P)[G::R], const float (&M)[G::R], int g,
                                             int col) {
    constexpr typename G::Prov pv = G::prov(W, I);
    constexpr geo::Src s0 = G::src(W, 0, I), s1 = G::src(W, 1, I), s2 = G::src(W, 2, I),
                       s3 = G::src(W, 3, I);
    // what this lane offers when it is a provider in this round
    const float o0 = reg_of<pv.kind[RD][0], pv.reg[RD][0]>(P, M);
    const float o1 = reg_of<pv.kind[RD][1], pv.reg[RD][1]>(P, M);
    const float o2 = reg_of<pv.kind[RD][2], pv.reg[RD][2]>(P, M);
    const float o3 = reg_of<pv.kind[RD][3], pv.reg[RD][3]>(P, M);
    const float offer = g == 0 ? o0 : g == 1 ? o1 : g == 2 ? o2 : o3;
    constexpr int l0 = pv.round[0] == RD ? s0.lane : 0;
    constexpr int l1 = pv.round[1] == RD ? s1.lane : 1;
    constexpr int l2 = pv.round[2] == RD ? s2.lane : 2;
    constexpr int l3 = pv.round[3] == RD ? s3.lane : 3;
    const int from = g == 0 ? l0 : g == 1 ? l1 : g == 2 ? l2 : l3;
    return __shfl(offer, col + 16 * from, 64);
}

template <class G, int W, int I>
__device__ __forceinline__ float seq_entry(const float (&P)[G::R], const float (&M)[G::R], int g,
                                           int col) {
    constexpr typename G::Prov pv = G::prov(W, I);
    static_assert(pv.nround <= 2, "more than two fetch rounds");
    constexpr geo::Src s0 = G::src(W, 0, I), s1 = G::src(W, 1, I), s2 = G::src(W, 2, I),
                       s3 = G::src(W, 3, I);
    float r0 = 0.0f, r1 = 0.0f;
    if constexpr (pv.nround > 0) r0 = fetch_round<G, W, I, 0>(P, M, g, col);
    if constexpr (pv.nround > 1) r1 = fetch_round<G, W, I, 1>(P, M, g, col);
    const float v0 = pv.round[0] < 0 ? reg_of<s0.kind, s0.reg>(P, M) : (pv.round[0] == 0 ? r0 : r1);
    const float v1 = pv.round[1] < 0 ? reg_of<s1.kind, s1.reg>(P, M) : (pv.round[1] == 0 ? r0 : r1);
    const float v2 = pv.round[2] < 0 ? reg_of<s2.kind, s2.reg>(P, M) : (pv.round[2] == 0 ? r0 : r1);
    const float v3 = pv.round[3] < 0 ? reg_of<s3.kind, s3.reg>(P, M) : (pv.round[3] == 0 ? r0 : r1);
    return g == 0 ? v0 : g == 1 ? v1 : g == 2 ? v2 : v3;
}

template <class G, int W, int I>
struct Fill {
    static __device__ __forceinline__ void run(float (&Xs)[G::X], const float (&P)[G::R],
                                               const float (&M)[G::R], int g, int col) {
        if constexpr (I < G::X) {
            Xs[I] = seq_entry<G, W, I>(P, M, g, col);
            Fill<G, W, I + 1>::run(Xs, P, M, g, col);
        }
    }
};

// per-width maxima (widths 1, 2, 3, 4, 6) of the windows starting in this
// lane's run: running pair sums
template <class G, int W, int I>
struct WinMax {
    static __device__ __forceinline__ void run(const float (&Xs)[G::X], int g, float (&m)[12]) {
        if constexpr (I < G::R) {
            constexpr bool a0 = G::start(W, 0, I), a1 = G::start(W, 1, I), a2 = G::start(W, 2, I),
                           a3 = G::start(W, 3, I);
            if constexpr (a0 || a1 || a2 || a3) {
                float s1 = Xs[I];
                float s2 = Xs[I] + Xs[I + 1];
                float s3 = s2 + Xs[I + 2];
                float s4 = s2 + (Xs[I + 2] + Xs[I + 3]);
                float s6 = s4 + (Xs[I + 4] + Xs[I + 5]);
                if constexpr (!(a0 && a1 && a2 && a3)) {
                    const bool ok = g == 0 ? a0 : g == 1 ? a1 : g == 2 ? a2 : a3;
                    s1 = ok ? s1 : -INFINITY;
                    s2 = ok ? s2 : -INFINITY;
                    s3 = ok ? s3 : -INFINITY;
                    s4 = ok ? s4 : -INFINITY;
                    s6 = ok ? s6 : -INFINITY;
                }
                m[0] = fmaxf(m[0], s1);
                m[1] = fmaxf(m[1], s2);
                m[2] = fmaxf(m[2], s3);
                m[3] = fmaxf(m[3], s4);
                m[4] = fmaxf(m[4], s6);
            }
            WinMax<G, W, I + 1>::run(Xs, g, m);
        }
    }
};

// per-width maxima over the bank widths <= WM: prefix-sum differences
template <class G, int W, int WM>
__device__ __forceinline__ void bank_max(const float (&Xs)[G::X], int g, float (&m)[12]) {
    constexpr int R = G::R, X = G::X;
    float ps[X + 1];
    ps[0] = 0.0f;
#pragma unroll
    for (int i = 0; i < X; ++i) ps[i + 1] = ps[i] + Xs[i];
#pragma unroll
    for (int a = 0; a < bank_count(WM); ++a) {
        const int w = bank_w(a);
#pragma unroll
        for (int i = 0; i < R; ++i) {
            const bool a0 = G::start(W, 0, i), a1 = G::start(W, 1, i), a2 = G::start(W, 2, i),
                       a3 = G::start(W, 3, i);
            if (a0 || a1 || a2 || a3) {
                float d = ps[i + w] - ps[i];
                if (!(a0 && a1 && a2 && a3)) {
                    const bool ok = g == 0 ? a0 : g == 1 ? a1 : g == 2 ? a2 : a3;
                    d = ok ? d : -INFINITY;
                }
                m[a] = fmaxf(m[a], d);
            }
        }
    }
}

// One fold of one 16-trial column block: split-BF16 c2r into registers, sigma
// samples, zero-baseline bound max_w max_p boxsum / sqrt(w) over the fold's
// bank (all widths <= WM).  Called by the whole wave; `active` gates the
// outputs of the lane's trial column.
template <int NB, int WM>
__device__ __noinline__ void fold_pass(const uint32_t* __restrict__ fre,
                                       const uint32_t* __restrict__ fim,
                                       const uint4* __restrict__ thi, const uint4* __restrict__ tlo,
                                       int k, int tcol, uint32_t bankmask, bool active,
                                       int64_t jcol, int64_t nv, int32_t cap, float* samp,
                                       float* mo) {
    constexpr int EXT = WM <= 6 ? 5 : WM - 1;
    using G = geo::Geo<NB, EXT>;
    constexpr int H = G::H, MT = G::MT, R = G::R, KS = G::KS, X = G::X;
    const int lane = threadIdx.x & 63, col = lane & 15, g = lane >> 4;

    // B operands: harmonic rows h*k (h = 1..H) of this lane's trial column
    bf16x8 rh[KS], rl[KS], ih[KS], il[KS];
#pragma unroll
    for (int kk = 0; kk < KS; ++kk) {
        uint32_t wr[8], wi[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int hh = kk * 32 + g * 8 + e + 1;
            uint32_t a = 0u, bq = 0u;
            if (hh <= H) {
                const int r = hh * k - 1;
                const int idx = r * kWT + (tcol ^ swz(r));
                a = fre[idx];
                bq = fim[idx];
            }
            wr[e] = a;
            wi[e] = bq;
        }
        unpack8(wr, rh[kk], rl[kk]);
        unpack8(wi, ih[kk], il[kk]);
    }
    // c2r: lane group g holds profile entries n = g*R + 4a + v
    float P[R], M[R];
#pragma unroll
    for (int a = 0; a < MT; ++a) {
        f32x4 E = {0.0f, 0.0f, 0.0f, 0.0f}, O = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kk = 0; kk < KS; ++kk) {
            const int tc = (a * KS + kk) * 2, ts = tc + 1;
            E = mfma3(as_bf16x8(thi[tc * 64 + lane]), as_bf16x8(tlo[tc * 64 + lane]), rh[kk],
                      rl[kk], E);
            O = mfma3(as_bf16x8(thi[ts * 64 + lane]), as_bf16x8(tlo[ts * 64 + lane]), ih[kk],
                      il[kk], O);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            P[4 * a + v] = E[v] + O[v];
            M[4 * a + v] = E[v] - O[v];
        }
    }
    // sigma samples (_block_sigma, search.jl:504-538): whole columns
    if (active) {
        const int64_t NN = (int64_t)NB * nv;
        int64_t slot = -1, take = NB;
        if (NN <= cap) {
            slot = jcol * NB;
        } else {
            int64_t ncol = cap / NB;
            if (ncol < 1) ncol = 1;
            if (ncol > nv) ncol = nv;
            const int64_t c = (jcol * ncol + nv - 1) / nv;
            if (c < ncol && (c * nv) / ncol == jcol) {
                slot = c * NB;
                take = cap - slot < NB ? cap - slot : NB;
            }
        }
        if (slot >= 0) {
#pragma unroll
            for (int i = 0; i < R; ++i) {
                const int n = g * R + i;
                if (n <= H && n < take) samp[slot + n] = fabsf(P[i]);
                if (n >= 1 && n <= H - 1 && NB - n < take) samp[slot + NB - n] = fabsf(M[i]);
            }
        }
    }
    // boxcar bound
    float m[12];
#pragma unroll
    for (int a = 0; a < 12; ++a) m[a] = -INFINITY;
    if constexpr (WM <= 6) {
        float XP[X], XM[X];
        Fill<G, 0, 0>::run(XP, P, M, g, col);
        Fill<G, 1, 0>::run(XM, P, M, g, col);
        WinMax<G, 0, 0>::run(XP, g, m);
        WinMax<G, 1, 0>::run(XM, g, m);
    } else {
        {
            float XP[X];
            Fill<G, 0, 0>::run(XP, P, M, g, col);
            bank_max<G, 0, WM>(XP, g, m);
        }
        {
            float XM[X];
            Fill<G, 1, 0>::run(XM, P, M, g, col);
            bank_max<G, 1, WM>(XM, g, m);
        }
    }
    float res = -INFINITY;
#pragma unroll
    for (int a = 0; a < bank_count(WM <= 6 ? 6 : WM); ++a)
        if (bankmask & (1u << a)) res = fmaxf(res, m[a] * (1.0f / sqrtf((float)bank_w(a))));
    res = fmaxf(res, __shfl_xor(res, 16, 64));
    res = fmaxf(res, __shfl_xor(res, 32, 64));
    if (active && g == 0) *mo = res;
}

#define CS_MFMA_SIZES(X) X(16) X(20) X(24) X(30) X(32) X(40) X(60) X(64) X(120) X(128)

__host__ __device__ constexpr bool mfma_size(int nb) {
    return nb == 16 || nb == 20 || nb == 24 || nb == 30 || nb == 32 || nb == 40 || nb == 60 ||
           nb == 64 || nb == 120 || nb == 128;
}

__host__ __device__ inline MfmaLds mfma_lds_layout(int nharms, int dft_tiles) {
    MfmaLds L;
    L.ftp = (size_t)2 * nharms * kWT * 4;
    L.region = 0;
    L.dft = (size_t)dft_tiles * 2 * 64 * 16;
    L.total = L.ftp + L.region + L.dft;
    return L;
}

// KS: interpolation K steps (1 or 2); WIDE: some fold's bank has widths > 6
// (kept apart so the narrow kernel does not reserve the wide bounds' registers)
template <int KS, bool WIDE>
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
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    uint32_t* fre = (uint32_t*)smem;   // [h][128 trials], (hi | lo << 16)
    uint32_t* fim = fre + nh * kWT;
    const int ntc = md.dft_tiles_max;  // DFT tiles of all folds
    uint4* dhi = (uint4*)(smem + (size_t)2 * nh * kWT * 4);
    uint4* dlo = dhi + (size_t)ntc * 64;
    for (int i = tid; i < ntc * 64; i += kThreads) {
        dhi[i] = md.cA_hi[i];
        dlo[i] = md.cA_lo[i];
    }

    // tile schedule: contiguous tile ranges per XCD (workgroups are dealt to
    // XCDs round-robin), so neighbouring tiles share that XCD's L2
    int64_t t_first, t_end, t_step;
    {
        const int G = gridDim.x, nx = md.nxcd;
        if (nx > 1 && G % nx == 0) {
            const int x = blockIdx.x % nx, l = blockIdx.x / nx;
            t_first = md.ntiles * x / nx + l;
            t_end = md.ntiles * (x + 1) / nx;
            t_step = G / nx;
        } else {
            t_first = blockIdx.x;
            t_end = md.ntiles;
            t_step = G;
        }
    }

    const int j16 = lane & 15, o = j16 >> 1, c = j16 & 1, g4 = lane >> 4;
    const int64_t S = md.S;
    const int U = md.U;
    const uint32_t* plane = md.spl + (size_t)c * md.spl_ld + kSpecPad;
    const int64_t bmin = -kSpecPad, bmax = sp.namps + kSpecPad - 32 * KS;
    const uint32_t np = (uint32_t)g.Nprof;

    for (int64_t tile = t_first; tile < t_end; tile += t_step) {
        const int64_t st = tile / U;
        const int u = (int)(tile - st * U);
        const int64_t tbase = st * kOcc8 * S + 16 * (int64_t)u;
        const int64_t occ = (int64_t)kOcc8 * st + o;

        // ---------------- interpolation: harmonics wv+1, wv+9, ... of this wave ----------
        int4 hd[8];
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const int h = wv + 1 + 8 * s;
            if (h <= nh) hd[s] = md.hdesc[(size_t)(h - 1) * U + u];
        }
        uint64_t rmv[4];
        {
            const int64_t t0 = tbase + o * S + 4 * g4;
            const uint32_t rel = (uint32_t)(t0 - g.origin);
            const uint32_t cg = rel / np, rem = rel - cg * np;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const uint32_t cgv = cg + (rem + v >= np ? 1u : 0u);
                rmv[v] = (t0 + v < g.total) ? md.rowmask[cgv] : 0ull;
            }
        }
        // operands of one harmonic in flight: window words, weight tiles, conj(A)
        struct Op {
            uint4 wa, wb, ah, al, q0, q1;
        };
        Op ring[3];
        auto issue = [&](int s, Op& x) {
            const int4 d = hd[s];
            int64_t b0 = (int64_t)d.x + (int64_t)d.y * occ;
            b0 = b0 < bmin ? bmin : (b0 > bmax ? bmax : b0);
            const uint32_t* p = plane + b0 + 8 * g4;
            x.wa = *(const uint4_a4*)p;
            x.wb = *(const uint4_a4*)(p + 4);
            x.ah = md.wA_hi[(size_t)d.z * 64 + lane];
            x.al = md.wA_lo[(size_t)d.z * 64 + lane];
            const uint4* gp = (const uint4*)(md.gA + (size_t)(d.w & 0x0fffffff) * 16 + 4 * g4);
            x.q0 = gp[0];
            x.q1 = gp[1];
        };
        if (wv + 1 <= nh) issue(0, ring[0]);
        if (wv + 9 <= nh) issue(1, ring[1]);
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const int h = wv + 1 + 8 * s;
            if (h > nh) break;
            if (s + 2 < 8 && h + 16 <= nh) issue(s + 2, ring[(s + 2) % 3]);
            const Op& x = ring[s % 3];
            bf16x8 bh, bl;
            {
                const uint32_t w[8] = {x.wa.x, x.wa.y, x.wa.z, x.wa.w, x.wb.x, x.wb.y, x.wb.z, x.wb.w};
                unpack8(w, bh, bl);
            }
            f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
            acc = mfma3(as_bf16x8(x.ah), as_bf16x8(x.al), bh, bl, acc);
            if constexpr (KS == 2) {
                if ((hd[s].w >> 28) == 2) {   // second 32-bin K step
                    int64_t b0 = (int64_t)hd[s].x + (int64_t)hd[s].y * occ;
                    b0 = b0 < bmin ? bmin : (b0 > bmax ? bmax : b0);
                    const uint32_t* p = plane + b0 + 32 + 8 * g4;
                    const uint4 x0 = *(const uint4_a4*)p, x1 = *(const uint4_a4*)(p + 4);
                    const uint32_t w[8] = {x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w};
                    bf16x8 ch, cl;
                    unpack8(w, ch, cl);
                    const size_t ti = ((size_t)hd[s].z + 1) * 64 + lane;
                    acc = mfma3(as_bf16x8(md.wA_hi[ti]), as_bf16x8(md.wA_lo[ti]), ch, cl, acc);
                }
            }
            const float gA[8] = {__builtin_bit_cast(float, x.q0.x), __builtin_bit_cast(float, x.q0.y),
                                 __builtin_bit_cast(float, x.q0.z), __builtin_bit_cast(float, x.q0.w),
                                 __builtin_bit_cast(float, x.q1.x), __builtin_bit_cast(float, x.q1.y),
                                 __builtin_bit_cast(float, x.q1.z), __builtin_bit_cast(float, x.q1.w)};
            uint32_t ow[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const float mine = acc[v], oth = pair_swap(mine);
                const float re = c ? oth : mine, im = c ? mine : oth;
                const float ar = gA[2 * v], ai = gA[2 * v + 1];
                // gA[k, g] * complex(ar, ai), product order of the reference
                const float p1 = ar * re, p2 = ai * im, p3 = ar * im, p4 = ai * re;
                float val = c ? (p3 + p4) : (p1 - p2);
                val = ((rmv[v] >> (h - 1)) & 1ull) ? val : 0.0f;
                ow[v] = split_word(val);
            }
            const int r = h - 1;
            uint32_t* dst = (c ? fim : fre) + r * kWT + ((o * 16 + 4 * g4) ^ swz(r));
            *(uint4*)dst = make_uint4(ow[0], ow[1], ow[2], ow[3]);
        }
        __syncthreads();

        // ---------------- c2r + bound, fold by fold, wave-local --------------------------
        {
            const int col = lane & 15;
            const int tcol = wv * 16 + col;
            const int64_t t = tbase + (int64_t)wv * S + col;
            const bool inb = t < g.total;
            int64_t jj = 0;
            int cl = 0;
            if (inb) {
                const uint32_t rel = (uint32_t)(t - g.origin);
                const uint32_t cg = rel / np;
                jj = rel - cg * np;
                cl = (int)((int64_t)cg - b.c_first);
            }
            const int nf = fs.nf;
            for (int f = 0; f < nf; ++f) {
                const int NB = fs.f[f].nbins;
                const int64_t nv = inb ? (int64_t)b.nvalid[cl * nf + f] : 0;
                const bool act = inb && jj < nv;
                float* mo = fu.mraw + (int64_t)f * fu.ldm + (inb ? (t - b.t_base) : 0);
                if (inb && !act && g4 == 0) *mo = -INFINITY;
                float* samp = fu.samples + ((int64_t)cl * nf + f) * kSigmaSamples;
                const uint4* th = dhi + (size_t)md.c_base[f] * 64;
                const uint4* tl = dlo + (size_t)md.c_base[f] * 64;
                const int k = fs.f[f].k;
                const uint32_t bm = fs.bankmask[f];
                const int32_t cap = fs.cap[f];
                const bool narrow = (bm >> 5) == 0;   // bank within widths 1, 2, 3, 4, 6
                switch (NB) {
#define X(SZ)                                                                                   \
    case SZ:                                                                                    \
        if constexpr (GateCT<SZ>::wmax <= 6 || !WIDE)                                           \
            fold_pass<SZ, 6>(fre, fim, th, tl, k, tcol, bm, act, jj, nv, cap, samp, mo);        \
        else if (narrow)                                                                        \
            fold_pass<SZ, 6>(fre, fim, th, tl, k, tcol, bm, act, jj, nv, cap, samp, mo);        \
        else                                                                                    \
            fold_pass<SZ, GateCT<SZ>::wmax>(fre, fim, th, tl, k, tcol, bm, act, jj, nv, cap,    \
                                            samp, mo);                                          \
        break;
                    CS_MFMA_SIZES(X)
#undef X
                    default:
                        break;
                }
            }
        }
        __syncthreads();
    }
}

__global__ void k_split_spectrum(const float2* __restrict__ amps, int64_t namps, int64_t ld,
                                 uint32_t* __restrict__ spl) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ld) return;
    const int64_t bin = i - kSpecPad;
    const float2 v = (bin >= 0 && bin < namps) ? amps[bin] : make_float2(0.0f, 0.0f);
    spl[i] = split_word(v.x);
    spl[ld + i] = split_word(v.y);
}

// one wave per block, lane h-1 evaluates the row guard of harmonic h
__global__ void k_rowmask(SpecDev sp, TabDev tb, GridDev g, int64_t nblocks,
                          uint64_t* __restrict__ out) {
    const int64_t blk = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (blk >= nblocks) return;
    const int64_t first = g.origin + blk * g.Nprof;
    const int64_t last = min(first + g.Nprof, g.total) - 1;
    bool ok = false;
    if (lane < tb.nharms) {
        const HarmDev d = tb.harm[lane];
        ok = d.P == 0 ? true : row_guard(d, tb, sp, first, last);
    }
    const uint64_t m = __ballot(ok);
    if (lane == 0) out[blk] = m;
}

bool mfma_supported(const FoldSet& fs, int nharms) {
    if (nharms > 64) return false;
    for (int f = 0; f < fs.nf; ++f) {
        const int nb = fs.f[f].nbins;
        if (!mfma_size(nb) || fs.bankmask[f] == 0) return false;
        // the kernel's bank for this profile length: widths up to 0.3 * nbins
        const int lim = (3 * nb) / 10 < 1 ? 1 : (3 * nb) / 10;
        if ((fs.bankmask[f] >> bank_count(lim)) != 0) return false;
    }
    return true;
}

MfmaLds mfma_lds_plan(int nharms, const FoldSet& fs, MfmaDev& md) {
    int tot = 0;
    for (int f = 0; f < fs.nf; ++f) tot += md.c_mt[f] * md.c_ks[f] * 2;
    md.dft_tiles_max = tot;
    md.nbp = 0;
    return mfma_lds_layout(nharms, tot);
}

hipError_t launch_split_spectrum(hipStream_t st, const SpecDev& sp, uint32_t* spl) {
    const int64_t ld = sp.namps + 2 * kSpecPad;
    hipLaunchKernelGGL(k_split_spectrum, dim3((unsigned)((ld + 255) / 256)), dim3(256), 0, st,
                       sp.amps, sp.namps, ld, spl);
    return hipGetLastError();
}

hipError_t launch_rowmask(hipStream_t st, const SpecDev& sp, const TabDev& tb, const GridDev& g,
                          int64_t nblocks, uint64_t* rowmask) {
    if (nblocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_rowmask, dim3((unsigned)((nblocks + 3) / 4)), dim3(256), 0, st, sp, tb, g,
                       nblocks, rowmask);
    return hipGetLastError();
}

hipError_t launch_fused_mfma(hipStream_t st, const SpecDev& sp, const TabDev& tb, const GridDev& g,
                             const FoldSet& fs, const BatchDev& b, const FusedDev& fu,
                             const MfmaDev& md, const FoldSet* fs_dev, const MfmaDev* md_dev,
                             int nwg) {
    if (nwg <= 0) return hipSuccess;
    const size_t shm = mfma_lds_layout(tb.nharms, md.dft_tiles_max).total;
    bool wide = false;
    for (int f = 0; f < fs.nf; ++f) wide = wide || (fs.bankmask[f] >> 5) != 0;
    auto go = [&](auto kern) -> hipError_t {
        hipError_t e = hipFuncSetAttribute((const void*)kern,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(kern, dim3((unsigned)nwg), dim3(kThreads), shm, st, sp, tb, g, fs_dev, b,
                           fu, md_dev);
        return hipGetLastError();
    };
    if (md.ksmax == 2) return wide ? go(k_fused_mfma<2, true>) : go(k_fused_mfma<2, false>);
    return wide ? go(k_fused_mfma<1, true>) : go(k_fused_mfma<1, false>);
}

}  // namespace cs
