// Split-BF16 MFMA fused search kernel for gfx950 (persistent workgroups).
//
// A tile is 128 trials: 8 occurrences, S = lcm(16, q) trials apart, of one
// 16-trial group u.  Because S is a multiple of the trial-grid period q, every
// occurrence has the same residue for every harmonic, so one 16 x K weight
// block (per harmonic and group) serves all eight, and the window start moves
// by an exact integer number of bins per occurrence:
//
//   interpolation  D[16 trials][8 occ x (re, im)] = Wx[16][K] . Win[K][16]
//   c2r, NB <= 64  x[n][128] = C[n][Re h | Im h] . F[h][128]   (direct)
//   c2r, NB > 64   E[n][128] = Ccos[n][h] . Re F[h][128]
//                  O[n][128] = Csin[n][h] . Im F[h][128]
//                  bin n = E + O, bin NB - n = E - O             (even/odd)
//
// each a v_mfma_f32_16x16x32_bf16 chain on bf16 hi/lo splits (hi.hi + hi.lo +
// lo.hi, FP32 accumulation).  The spectrum is kept pre-split as (hi | lo << 16)
// words in two planes, so a lane's window is two 16-byte loads and eight
// byte-permutes; the stacked amplitudes go to LDS pre-split the same way.  The
// c2r rows are permuted so each of the four lanes of a trial column holds a
// contiguous run of the profile, and the boxcar bound is formed from registers
// (cs_geo.hpp): in the direct layout a lane's run already carries most of its
// window extension, in the even/odd layout runs are stitched across lanes.
// LDS holds the stacked amplitudes and every fold's DFT tiles; a workgroup
// loops over tiles (XCD-contiguous ranges), two barriers per tile.  Index and
// phase bookkeeping stay exact integer arithmetic.
#include <hip/hip_runtime.h>

#include <utility>

#include "cs_device.hpp"
#include "cs_geo.hpp"
#include "cs_mfma.hpp"

namespace cs {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint4 uint4_a4 __attribute__((aligned(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// explicit address spaces for pointers handed to out-of-line device functions
// (a generic pointer would turn their LDS reads into FLAT accesses)
#define CS_LDS __attribute__((address_space(3)))
#define CS_GLB __attribute__((address_space(1)))
#define CS_CONST __attribute__((address_space(4)))

template <bool B>
struct BoolC {
    static constexpr bool value = B;
};

constexpr int kOcc8 = 8;          // occurrences per tile
constexpr int kThreads = 512;     // 8 waves


__device__ __forceinline__ bf16x8 as_bf16x8(uint4 v) { return __builtin_bit_cast(bf16x8, v); }
__device__ __forceinline__ bf16x8 as_bf16x8(u32x4 v) { return __builtin_bit_cast(bf16x8, v); }

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

// x -> (hi | lo << 16): hi = bf16(x), lo = bf16(x - hi), both round to
// nearest even; two packed conversions, the second re-rounding x to hi
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t split_word(float x) {
    const uint32_t hh = __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{x, x}, bf16x2));
    const float hf = __builtin_bit_cast(float, hh << 16);
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{x, x - hf}, bf16x2));
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

// Maximum over the four 16-lane rows, in every lane (the whole wave active):
// v_permlane16_swap of x with itself yields rows (0,0,2,2) and (1,1,3,3),
// v_permlane32_swap rows (0,1,0,1) and (2,3,2,3); no LDS round trip.
__device__ __forceinline__ float max_over_rows(float x) {
    const uint32_t u = __builtin_bit_cast(uint32_t, x);
    const auto a = __builtin_amdgcn_permlane16_swap(u, u, false, false);
    const float y = fmaxf(__builtin_bit_cast(float, (uint32_t)a[0]),
                          __builtin_bit_cast(float, (uint32_t)a[1]));
    const uint32_t v = __builtin_bit_cast(uint32_t, y);
    const auto b = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return fmaxf(__builtin_bit_cast(float, (uint32_t)b[0]), __builtin_bit_cast(float, (uint32_t)b[1]));
}

// Row SRC (16 lanes) of x in every row (whole wave active): rows (0,1,0,1)
// or (2,3,2,3) by v_permlane32_swap, then the wanted one of each pair by
// v_permlane16_swap.
template <int SRC>
__device__ __forceinline__ float row_broadcast(float x) {
    const uint32_t u = __builtin_bit_cast(uint32_t, x);
    const auto h = __builtin_amdgcn_permlane32_swap(u, u, false, false);
    const uint32_t v = SRC < 2 ? (uint32_t)h[0] : (uint32_t)h[1];
    const auto q = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    return __builtin_bit_cast(float, (SRC & 1) ? (uint32_t)q[1] : (uint32_t)q[0]);
}

// Workgroup barrier that orders LDS only.  No wave of the fused kernel reads
// global memory another wave wrote, so the barrier need not wait for this
// wave's outstanding global stores (bounds, sigma samples), which a full
// __syncthreads() release fence would.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// A read-only descriptor block through a pointer the compiler cannot see
// through, so loads through it are not hoisted out of a loop into registers
// held for the whole kernel; the first-lane reads keep it uniform and the
// constant address space makes its loads scalar (the kernel never writes it).
template <class T>
__device__ __forceinline__ const CS_CONST T* uniform_opaque(const T* p) {
    uint64_t v = (uint64_t)(uintptr_t)p;
    asm volatile("" : "+v"(v));
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
    return (const CS_CONST T*)(uintptr_t)(((uint64_t)hi << 32) | lo);
}

// An LDS pointer the optimiser cannot see through: keeps a constant base
// from being propagated into out-of-line callees, which would then fetch the
// dynamic-LDS base from its offset table at every access.
template <class T>
__device__ __forceinline__ CS_LDS T* opaque_lds(CS_LDS T* p) {
    uint32_t v = (uint32_t)(uintptr_t)p;
    asm volatile("" : "+s"(v));
    return (CS_LDS T*)(uintptr_t)v;
}

// Global-memory view of a pointer read from a device-side descriptor.  The
// compiler cannot infer the address space of such a pointer and would emit
// FLAT accesses, which also count against the LDS/scalar wait counter, so a
// later scalar-load wait would wait for them too.
template <class T>
__device__ __forceinline__ const CS_GLB T* glb(const T* p) {
    return (const CS_GLB T*)p;
}
// 16-byte global loads through such a pointer (p 16-byte aligned, or 4-byte
// aligned for ldg4_a4)
__device__ __forceinline__ uint4 ldg4(const void* p) {
    const u32x4 v = *(const CS_GLB u32x4*)p;
    return make_uint4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ uint4 ldg4_a4(const void* p) {
    typedef u32x4 u32x4_a4 __attribute__((aligned(4)));
    const u32x4_a4 v = *(const CS_GLB u32x4_a4*)p;
    return make_uint4(v.x, v.y, v.z, v.w);
}

// Read-only table entry at a wave-uniform index, read through the constant
// address space so it becomes a scalar load (the kernel's global stores
// cannot alias it).
__device__ __forceinline__ int4 ld_const4(const int4* p, size_t i) {
    typedef int i32x4v __attribute__((ext_vector_type(4)));
    const i32x4v v = ((const CS_CONST i32x4v*)p)[i];
    return make_int4(v.x, v.y, v.z, v.w);
}

// n / d (and n % d in r) for 32-bit n and 2 <= d with m = floor(2^32 / d):
// the high product is the quotient or one less (n < 2^32), one correction
__device__ __forceinline__ uint32_t udiv_m(uint32_t n, uint32_t d, uint32_t m, uint32_t& r) {
    uint32_t q = __umulhi(n, m);
    r = n - q * d;
    const bool c = r >= d;
    q += c ? 1u : 0u;
    r -= c ? d : 0u;
    return q;
}

// Half-word f of a 16-byte value: unsigned, signed, and replaced by the low
// half of v (f a compile-time constant where these are used)
__device__ __forceinline__ uint32_t half_u(const uint4& q, int f) {
    const uint32_t w = f < 2 ? q.x : f < 4 ? q.y : f < 6 ? q.z : q.w;
    return (f & 1) ? w >> 16 : w & 0xFFFFu;
}
__device__ __forceinline__ int half_s(const uint4& q, int f) {
    return (int)(int16_t)(uint16_t)half_u(q, f);
}
__device__ __forceinline__ void set_half(uint4& q, int f, uint32_t v) {
    uint32_t& w = f < 2 ? q.x : f < 4 ? q.y : f < 6 ? q.z : q.w;
    w = (f & 1) ? ((w & 0x0000FFFFu) | (v << 16)) : ((w & 0xFFFF0000u) | (v & 0xFFFFu));
}

// Row stride (words) of the stacked-amplitude planes.  An even/odd fold's c2r
// operand read touches rows 8k apart in the four lane groups of a column;
// with 130-word rows the two groups of a half-wave land 16k (mod 32) banks
// apart, disjoint for odd k.  The direct folds space the groups' rows by the
// fold factor (geo::dperm), disjoint for every k.
constexpr int kFtpLd = 130;

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

// per-width maxima (widths 1, 2, 3, 4, 6) of the windows at the R positions
// from I of a sequence Xs (each a real boxcar, cs_geo.hpp): running pair
// sums.  BM: the fold's bank as a compile-time mask (bit a: width bank_w(a)),
// whose absent widths are neither summed nor maximised; 0 when the bank is
// only known at run time (all five computed).
template <uint32_t BM, int N, int I, int R>
__device__ __forceinline__ void win_max(const float (&Xs)[N], float (&m)[12]) {
    constexpr bool all = BM == 0;
#pragma unroll
    for (int i = I; i < I + R; ++i) {
        const float s2 = Xs[i] + Xs[i + 1];
        if constexpr (all || (BM & 1u)) m[0] = fmaxf(m[0], Xs[i]);
        if constexpr (all || (BM & 2u)) m[1] = fmaxf(m[1], s2);
        if constexpr (all || (BM & 4u)) m[2] = fmaxf(m[2], s2 + Xs[i + 2]);
        if constexpr (all || (BM & 8u) || (BM & 16u)) {
            const float s4 = s2 + (Xs[i + 2] + Xs[i + 3]);
            if constexpr (all || (BM & 8u)) m[3] = fmaxf(m[3], s4);
            if constexpr (all || (BM & 16u)) m[4] = fmaxf(m[4], s4 + (Xs[i + 4] + Xs[i + 5]));
        }
    }
}

// per-width maxima over the bank widths <= WM: prefix-sum differences
template <class G, int W, int WM>
__device__ __forceinline__ void bank_max(const float (&Xs)[G::X], float (&m)[12]) {
    constexpr int R = G::R, X = G::X;
    float ps[X + 1];
    ps[0] = 0.0f;
#pragma unroll
    for (int i = 0; i < X; ++i) ps[i + 1] = ps[i] + Xs[i];
#pragma unroll
    for (int a = 0; a < bank_count(WM); ++a) {
        const int w = bank_w(a);
#pragma unroll
        for (int i = 0; i < R; ++i) m[a] = fmaxf(m[a], ps[i + w] - ps[i]);
    }
}

// max over the fold's bank of (largest window sum of width w) / sqrt(w).  With
// a compile-time bank (BM != 0) one straight max chain over its widths; else
// the mask is wave-uniform: read into a scalar register, its bits choose the
// widths by scalar branches, and the bank of widths 1, 2, 3, 4, 6 takes one
// straight chain with no per-width select.
template <uint32_t BM>
__device__ __forceinline__ float bank_bound_ct(const float (&m)[12]) {
    float res = -INFINITY;
#pragma unroll
    for (int a = 0; a < 5; ++a)
        if ((BM >> a) & 1u) res = fmaxf(res, m[a] * (1.0f / sqrtf((float)bank_w(a))));
    return res;
}
template <int NA>
__device__ __forceinline__ float bank_bound(const float (&m)[12], uint32_t bankmask) {
    const uint32_t bm = __builtin_amdgcn_readfirstlane(bankmask);
    if (NA == 5 && bm == 0x1Fu) {
        const float r0 = fmaxf(m[0], m[1] * (1.0f / sqrtf(2.0f)));
        const float r1 = fmaxf(m[2] * (1.0f / sqrtf(3.0f)), m[3] * 0.5f);
        return fmaxf(fmaxf(r0, r1), m[4] * (1.0f / sqrtf(6.0f)));
    }
    float res = -INFINITY;
#pragma unroll
    for (int a = 0; a < NA; ++a)
        if (bm & (1u << a)) res = fmaxf(res, m[a] * (1.0f / sqrtf((float)bank_w(a))));
    return res;
}

// Stores through a buffer resource with a per-lane byte offset.  An offset at
// or past the resource's size drops the store, so a lane that has nothing to
// store passes kNoStore (plus the chunk's own small offset, which stays below
// 2^32) and no branch is needed: the folds stay one straight-line region that
// the compiler can interleave.  The buffers these resources cover are kept
// below kNoStore bytes by the host.
constexpr uint32_t kNoStore = kMfmaNoStore;
// the same offset for a load: past every buffer the kernel reads through a
// resource, so the load returns zeros and reads no memory
constexpr uint32_t kNoLoad = kMfmaNoStore;
typedef unsigned int u32x3 __attribute__((ext_vector_type(3)));
template <int CH>
__device__ __forceinline__ void buf_store(const float* v, __amdgpu_buffer_rsrc_t r, uint32_t off) {
    if constexpr (CH == 4) {
        const u32x4 w = {__float_as_uint(v[0]), __float_as_uint(v[1]), __float_as_uint(v[2]),
                         __float_as_uint(v[3])};
        __builtin_amdgcn_raw_buffer_store_b128(w, r, off, 0, 0);
    } else if constexpr (CH == 3) {
        const u32x3 w = {__float_as_uint(v[0]), __float_as_uint(v[1]), __float_as_uint(v[2])};
        __builtin_amdgcn_raw_buffer_store_b96(w, r, off, 0, 0);
    } else if constexpr (CH == 2) {
        typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
        const u32x2 w = {__float_as_uint(v[0]), __float_as_uint(v[1])};
        __builtin_amdgcn_raw_buffer_store_b64(w, r, off, 0, 0);
    } else {
        __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v[0]), r, off, 0, 0);
    }
}

// Where a lane's sampled column goes: the batch's sample regions and the
// byte offset of the column's slot block, kNoStore when the lane's trial is
// not a sampled column of the fold (sigma samples, _block_sigma,
// search.jl:504-538: the whole profile of a sampled column goes to its NB
// slots, columns NBS slots apart; the median only sees the multiset, so the
// order of a column's slots is free).
struct ColumnSink {
    __amdgpu_buffer_rsrc_t rsrc;
    uint32_t off;
};

// One sampled column's share of this lane group (runs n = g*R .. g*R + R - 1
// of P and M) into its slot block: P[n] at n, M[n] at H + n, P[H] in place of
// M[0]; whole CH-word chunks, the chunks past H dropped.  Signed values: the
// select takes each sample's magnitude as it reads it, so no lane spends an
// instruction on the columns that are not stored.
template <int H, int R>
__device__ __forceinline__ void put_column(const ColumnSink& sk, int g, const float (&P)[R],
                                           const float (&M)[R], float pH) {
    constexpr int CH = (H % 4 == 0) ? 4 : (H % 2 == 0) ? 2 : 1;
#pragma unroll
    for (int c0 = 0; c0 < R; c0 += CH) {
        const int n0 = g * R + c0;
        const uint32_t off = n0 + CH <= H ? sk.off : kNoStore;
        float pv[CH], mv[CH];
#pragma unroll
        for (int e = 0; e < CH; ++e) {
            pv[e] = P[c0 + e];
            mv[e] = M[c0 + e];
        }
        if (c0 == 0) mv[0] = n0 == 0 ? pH : mv[0];
        buf_store<CH>(pv, sk.rsrc, off + 4u * (uint32_t)n0);
        buf_store<CH>(mv, sk.rsrc, off + 4u * (uint32_t)(H + n0));
    }
}

// Words of eight stacked-amplitude rows h0 + e (e < 8) of both LDS planes
// (im = re + dplane bytes), stride rstride words from rbase, one address
// chain; rows past H read row H (its c2r coefficients beyond H are exactly
// zero, so any finite row serves).  The byte addresses advance by one add and
// are clamped by one min each.
template <int H>
__device__ __forceinline__ void gather_rows2(const CS_LDS uint32_t* re, const CS_LDS uint32_t* im,
                                             int h0, int rstride, int rbase, uint32_t* wr,
                                             uint32_t* wi) {
    const uint32_t b = (uint32_t)(uintptr_t)re;
    const uint32_t dp = (uint32_t)(uintptr_t)im - b;
    const uint32_t st = 4u * (uint32_t)rstride;
    uint32_t a = b + 4u * (uint32_t)((int)__umul24((unsigned)h0, (unsigned)rstride) + rbase);
    const uint32_t cap = b + 4u * (uint32_t)((int)__umul24((unsigned)H, (unsigned)rstride) + rbase);
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const uint32_t x = a < cap ? a : cap;
        wr[e] = *(const CS_LDS uint32_t*)(uintptr_t)x;
        wi[e] = *(const CS_LDS uint32_t*)(uintptr_t)(x + dp);
        a += st;
    }
}

// The same gathers at a compile-time fold factor KF (row stride KF rows), with
// no clamp: a row past the fold's last harmonic meets c2r coefficients that
// are exactly zero, so any finite word serves, and the kernels that use these
// (fold sets checked by set_reads_in_tables) keep every such row inside the
// stacked planes and the DFT tiles.  The row offsets become immediate LDS
// offsets.  Stack row h*KF sits at plane row h*KF - 1.
template <int KF>
__device__ __forceinline__ void gather_rows2_k(const CS_LDS uint32_t* re, const CS_LDS uint32_t* im,
                                               int h0, int tcol, uint32_t* wr, uint32_t* wi) {
    const int off = ((int)__umul24((unsigned)h0, KF) - 1) * kFtpLd + tcol;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        wr[e] = re[off + e * KF * kFtpLd];
        wi[e] = im[off + e * KF * kFtpLd];
    }
}

// Direct-layout gathers in the tiles' K-slot order (geo::dperm): operand e
// reads row h0 + (e mod d) + ng*d*(e div d), d = 2^s, for the ng lane groups
// that share a plane (2 for one K chunk of Re | Im, 4 for a chunk per plane).
// Run-time fold factor: rows past H read row H, as in gather_rows2.
template <int H, int NG>
__device__ __forceinline__ void gather_drows(const CS_LDS uint32_t* pl, int h0, int s,
                                             int rstride, int rbase, uint32_t* w) {
    const uint32_t b = (uint32_t)(uintptr_t)pl;
    const uint32_t st = 4u * (uint32_t)rstride;
    const uint32_t a = b + 4u * (uint32_t)((int)__umul24((unsigned)h0, (unsigned)rstride) + rbase);
    const uint32_t cap = b + 4u * (uint32_t)((int)__umul24((unsigned)H, (unsigned)rstride) + rbase);
    const int dm = (1 << s) - 1;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const uint32_t x = a + st * (uint32_t)((e & dm) + NG * ((e >> s) << s));
        w[e] = *(const CS_LDS uint32_t*)(uintptr_t)(x < cap ? x : cap);
    }
}
// Compile-time fold factor KF (immediate offsets, no clamp, as gather_rows2_k)
template <int KF, int NG>
__device__ __forceinline__ void gather_drows_k(const CS_LDS uint32_t* pl, int h0, int tcol,
                                               uint32_t* w) {
    constexpr int S = geo::dshift(KF), D = 1 << S;
    const CS_LDS uint32_t* p = pl + ((int)__umul24((unsigned)h0, KF) - 1) * kFtpLd + tcol;
#pragma unroll
    for (int e = 0; e < 8; ++e) w[e] = p[(e % D + NG * D * (e / D)) * KF * kFtpLd];
}

// BM: the fold's bank as a compile-time mask (0: read from bankmask at run time)
template <int NB, int WM, int KF = 0, uint32_t BM = 0>
__device__ __forceinline__ float fold_pass_body(const CS_LDS uint32_t* __restrict__ fre,
                                                const CS_LDS uint32_t* __restrict__ fim,
                                                const CS_LDS u32x4* __restrict__ thi,
                                                const CS_LDS u32x4* __restrict__ tlo, int k,
                                                int tcol, uint32_t bankmask, ColumnSink sk) {
    constexpr int EXT = WM <= 6 ? 5 : WM - 1;
    using G = geo::Geo<NB, EXT>;
    constexpr int H = G::H, MT = G::MT, R = G::R, KS = G::KS, X = G::X;
    const int lane = threadIdx.x & 63, col = lane & 15, g = lane >> 4;

    // B operands: harmonic rows h*k (h = 1..H) of this lane's trial column;
    // stack row h*k sits at plane row h*k - 1
    const int rstride = k * kFtpLd, rbase = tcol - kFtpLd;
    bf16x8 rh[KS], rl[KS], ih[KS], il[KS];
#pragma unroll
    for (int kk = 0; kk < KS; ++kk) {
        uint32_t wr[8], wi[8];
        if constexpr (KF > 0) gather_rows2_k<KF>(fre, fim, kk * 32 + g * 8 + 1, tcol, wr, wi);
        else gather_rows2<H>(fre, fim, kk * 32 + g * 8 + 1, rstride, rbase, wr, wi);
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
    // sigma samples: P[n] at slot n (n < H), the Nyquist bin P[H] at H and
    // M[n] at H + n
    {
        const float pH = row_broadcast<(H / R) & 3>(P[H % R]);
        put_column<H, R>(sk, g, P, M, pH);
    }
    // boxcar bound
    float m[12];
#pragma unroll
    for (int a = 0; a < 12; ++a) m[a] = -INFINITY;
    if constexpr (WM <= 6) {
        // one sequence at a time: its window registers die before the next
        {
            float XP[X];
            Fill<G, 0, 0>::run(XP, P, M, g, col);
            win_max<BM, X, 0, R>(XP, m);
        }
        {
            float XM[X];
            Fill<G, 1, 0>::run(XM, P, M, g, col);
            win_max<BM, X, 0, R>(XM, m);
        }
    } else {
        {
            float XP[X];
            Fill<G, 0, 0>::run(XP, P, M, g, col);
            bank_max<G, 0, WM>(XP, m);
        }
        {
            float XM[X];
            Fill<G, 1, 0>::run(XM, P, M, g, col);
            bank_max<G, 1, WM>(XM, m);
        }
    }
    if constexpr (BM != 0) return max_over_rows(bank_bound_ct<BM>(m));
    else return max_over_rows(bank_bound<bank_count(WM <= 6 ? 6 : WM)>(m, bankmask));
}

// One sampled column's share of lane group g in the direct layout: registers
// X[r] = bin g*RD + r at slots g*RD + r, in 4-word chunks (4-byte aligned) and
// a shorter tail.  Group 3's last SH registers repeat group 0's first bins;
// they land in the column's padding slots NB .. round_up(NB, 4) - 1, which
// the median never reads.  Signed values, as in put_column.
template <int RD>
__device__ __forceinline__ void put_run(const ColumnSink& sk, int g, const float* Xs) {
    const uint32_t base = sk.off + 4u * (uint32_t)(g * RD);
#pragma unroll
    for (int c0 = 0; c0 < RD; c0 += 4) {
        float v[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) v[e] = c0 + e < RD ? Xs[c0 + e] : 0.0f;
        if (c0 + 4 <= RD) buf_store<4>(v, sk.rsrc, base + 4u * c0);
        else if (RD - c0 == 3) buf_store<3>(v, sk.rsrc, base + 4u * c0);
        else if (RD - c0 == 2) buf_store<2>(v, sk.rsrc, base + 4u * c0);
        else buf_store<1>(v, sk.rsrc, base + 4u * c0);
    }
}

// One fold of one 16-trial column block in the direct layout (NB <= 64,
// cs_geo.hpp DGeo): split-BF16 c2r straight to bins, lane group g holding
// bins g*RD .. g*RD + 4*MT - 1 (mod NB), the rest of its window extension
// read from the next lane group; sigma samples; the bound
// max_w max_p boxsum / sqrt(w) over the fold's bank (widths <= WM).
template <int NB, int WM, int KF = 0, uint32_t BM = 0>
__device__ __forceinline__ float fold_direct_body(const CS_LDS uint32_t* __restrict__ fre,
                                                  const CS_LDS uint32_t* __restrict__ fim,
                                                  const CS_LDS u32x4* __restrict__ thi,
                                                  const CS_LDS u32x4* __restrict__ tlo, int k,
                                                  int tcol, uint32_t bankmask, ColumnSink sk) {
    using D = geo::DGeo<NB, WM>;
    constexpr int H = D::H, RD = D::RD, SH = D::SH, MT = D::MT, NR = D::NR, X = D::X,
                  NF = D::NF, KD = D::KD;
    static_assert(KD <= 2, "the direct layout covers H <= 32");
    const int lane = threadIdx.x & 63, g = lane >> 4;
    // stack row h*k at plane row h*k - 1 (rows < 64, strides < 2^24: 24-bit products)
    const int rstride = k * kFtpLd, rbase = tcol - kFtpLd;

    // B operands: K slot kk*32 + geo::dperm(g, e) holds Re h (first half) or
    // Im h (second half); slots past the last harmonic read row H (zero
    // coefficients)
    bf16x8 bh[KD], bl[KD];
    const int s = KF > 0 ? geo::dshift(KF) : geo::dshift(k);
    if constexpr (KD == 1) {
        uint32_t w[8];
        const int h0 = ((g & 1) << s) + 1;
        if constexpr (KF > 0) gather_drows_k<KF, 2>(g < 2 ? fre : fim, h0, tcol, w);
        else gather_drows<H, 2>(g < 2 ? fre : fim, h0, s, rstride, rbase, w);
        unpack8(w, bh[0], bl[0]);
    } else {
        uint32_t wr[8], wi[8];
        const int h0 = (g << s) + 1;
        if constexpr (KF > 0) {
            gather_drows_k<KF, 4>(fre, h0, tcol, wr);
            gather_drows_k<KF, 4>(fim, h0, tcol, wi);
        } else {
            gather_drows<H, 4>(fre, h0, s, rstride, rbase, wr);
            gather_drows<H, 4>(fim, h0, s, rstride, rbase, wi);
        }
        unpack8(wr, bh[0], bl[0]);
        unpack8(wi, bh[1], bl[1]);
    }
    float Xs[X];
#pragma unroll
    for (int a = 0; a < MT; ++a) {
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int kk = 0; kk < KD; ++kk) {
            const int tt = a * KD + kk;
            acc = mfma3(as_bf16x8(thi[tt * 64 + lane]), as_bf16x8(tlo[tt * 64 + lane]), bh[kk],
                        bl[kk], acc);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) Xs[4 * a + v] = acc[v];
    }
    // window extension past the MFMA registers: the next lane group's
    // registers (group 0 serves group 3 across the wrap)
#pragma unroll
    for (int i = 0; i < NF; ++i) {
        const int r = NR + i;
        float offer = Xs[r - RD];
        if constexpr (SH != 0) offer = g == 0 ? Xs[r - RD + SH] : offer;
        Xs[r] = __shfl(offer, (lane + 16) & 63, 64);
    }

    put_run<RD>(sk, g, Xs);

    // boxcar bound over the windows at the lane group's RD positions (group 3's
    // last SH repeat group 0's first windows: no mask needed)
    float m[12];
#pragma unroll
    for (int a = 0; a < 12; ++a) m[a] = -INFINITY;
    if constexpr (WM <= 6) {
        win_max<BM, X, 0, RD>(Xs, m);
    } else {
        float ps[X + 1];
        ps[0] = 0.0f;
#pragma unroll
        for (int i = 0; i < X; ++i) ps[i + 1] = ps[i] + Xs[i];
#pragma unroll
        for (int a = 0; a < bank_count(WM); ++a) {
            const int w = bank_w(a);
#pragma unroll
            for (int r = 0; r < RD; ++r) m[a] = fmaxf(m[a], ps[r + w] - ps[r]);
        }
    }
    if constexpr (BM != 0) return max_over_rows(bank_bound_ct<BM>(m));
    else return max_over_rows(bank_bound<bank_count(WM <= 6 ? 6 : WM)>(m, bankmask));
}

// Out-of-line copies of the fold passes, for the kernel that dispatches on
// the profile length at run time (one call site per length and fold).
template <int NB, int WM>
__device__ __noinline__ float fold_pass(const CS_LDS uint32_t* __restrict__ fre,
                                        const CS_LDS uint32_t* __restrict__ fim,
                                        const CS_LDS u32x4* __restrict__ thi,
                                        const CS_LDS u32x4* __restrict__ tlo, int k, int tcol,
                                        uint32_t bankmask, ColumnSink sk) {
    return fold_pass_body<NB, WM>(fre, fim, thi, tlo, k, tcol, bankmask, sk);
}
template <int NB, int WM>
__device__ __noinline__ float fold_direct(const CS_LDS uint32_t* __restrict__ fre,
                                          const CS_LDS uint32_t* __restrict__ fim,
                                          const CS_LDS u32x4* __restrict__ thi,
                                          const CS_LDS u32x4* __restrict__ tlo, int k, int tcol,
                                          uint32_t bankmask, ColumnSink sk) {
    return fold_direct_body<NB, WM>(fre, fim, thi, tlo, k, tcol, bankmask, sk);
}

// The pass for a fold of NB bins: the direct layout up to 64 bins, the
// even/odd layout above; widths up to 6 unless the fold's bank is wider.
// INL: the pass body inlined into the caller (the kernels specialised on a
// fold set) rather than called.
template <int NB, int WM, bool INL, int KF, uint32_t BM>
__device__ __forceinline__ float fold_of(const CS_LDS uint32_t* fre, const CS_LDS uint32_t* fim,
                                         const CS_LDS u32x4* thi, const CS_LDS u32x4* tlo, int k,
                                         int tcol, uint32_t bankmask, const ColumnSink& sk) {
    if constexpr (geo::c2r_direct(NB)) {
        if constexpr (INL)
            return fold_direct_body<NB, WM, KF, BM>(fre, fim, thi, tlo, k, tcol, bankmask, sk);
        else
            return fold_direct<NB, WM>(fre, fim, thi, tlo, k, tcol, bankmask, sk);
    } else {
        if constexpr (INL)
            return fold_pass_body<NB, WM, KF, BM>(fre, fim, thi, tlo, k, tcol, bankmask, sk);
        else
            return fold_pass<NB, WM>(fre, fim, thi, tlo, k, tcol, bankmask, sk);
    }
}

template <int NB, bool WIDE, bool INL = false, int KF = 0, uint32_t BM = 0>
__device__ __forceinline__ float run_fold(bool narrow, const CS_LDS uint32_t* fre,
                                          const CS_LDS uint32_t* fim, const CS_LDS u32x4* thi,
                                          const CS_LDS u32x4* tlo, int k, int tcol,
                                          uint32_t bankmask, const ColumnSink& sk) {
    constexpr int WW = gate_wmax(NB);
    if constexpr (WW <= 6 || !WIDE)
        return fold_of<NB, 6, INL, KF, BM>(fre, fim, thi, tlo, k, tcol, bankmask, sk);
    else if (narrow)
        return fold_of<NB, 6, INL, KF, BM>(fre, fim, thi, tlo, k, tcol, bankmask, sk);
    else
        return fold_of<NB, WW, INL, 0, 0>(fre, fim, thi, tlo, k, tcol, bankmask, sk);
}

// Fold sets with a kernel of their own (profile length of fold f, in order):
// set 1 is decimation_set(60, 6), nharms 60 folded by 1..6; set 2 is
// decimation_set(32, 4); set 3 is nharms 8 unfolded.  Set 0 is the kernel that
// takes any supported lengths.
constexpr int kFoldSets = 4;
constexpr int kFoldSetNB[kFoldSets][kMaxFolds] = {
    {0}, {120, 60, 40, 30, 24, 20}, {64, 32, 20, 16}, {16}};
// and their boxcar banks as width masks (bit a: width 1, 2, 3, 4, 6): the
// default bank (boxcar_fsp 1.5, boxcar_maxfrac 0.3), width 1 pruned from the
// folds above the first by the ladder (ladder_boxcar_widths, search.jl:443-459)
constexpr uint32_t kFoldSetBM[kFoldSets][kMaxFolds] = {
    {0}, {0x1F, 0x1E, 0x1E, 0x1E, 0x1E, 0x1E}, {0x1F, 0x1E, 0x1E, 0x0E}, {0x0F}};
constexpr int fold_set_size(int set) {
    int n = 0;
    for (int f = 0; f < kMaxFolds; ++f) n += kFoldSetNB[set][f] > 0 ? 1 : 0;
    return n;
}
// c2r DFT tiles (hi or lo) of a fold of NB bins with a bank within widths 1..6
// (the tables cs_plan.cpp builds for it)
constexpr int fold_tiles(int NB) {
    const int H = NB / 2;
    if (geo::c2r_direct(NB)) {
        const int RD = (NB + 3) / 4, SH = 4 * RD - NB, ext = 5;
        const int mt = (RD + 3) / 4 > (ext + SH + 3) / 4 ? (RD + 3) / 4 : (ext + SH + 3) / 4;
        return mt * (H <= 16 ? 1 : H <= 32 ? 2 : 4);
    }
    return 2 * ((H + 1 + 15) / 16) * ((H + 31) / 32);
}
constexpr int set_tiles(int set) {
    int n = 0;
    for (int f = 0; f < fold_set_size(set); ++f) n += fold_tiles(kFoldSetNB[set][f]);
    return n;
}
// first DFT tile of fold f of the set (the tables cs_plan.cpp lays out in fold order)
constexpr int set_cbase(int set, int f) {
    int n = 0;
    for (int i = 0; i < f && i < fold_set_size(set); ++i) n += fold_tiles(kFoldSetNB[set][i]);
    return n;
}
// highest stack row index (in units of the fold factor) a fold pass gathers:
// 32 per K step of the even/odd layout, 16 per K step of the direct one
constexpr int fold_hmax(int NB) {
    const int H = NB / 2;
    if (geo::c2r_direct(NB)) return 16 * (H <= 16 ? 1 : H <= 32 ? 2 : 4);
    return 32 * ((H + 31) / 32);
}
// Whether the unclamped gathers of every fold of the set (fold f has factor
// f + 1) stay inside the LDS allocation, i.e. read only the two stacked
// planes and the DFT tiles behind them, all finite words.
constexpr bool set_reads_in_tables(int set) {
    const int nh = kFoldSetNB[set][0] / 2;
    const int stage = 4 * (2 * nh * kFtpLd) + set_tiles(set) * 2 * 64 * 16;   // end of the tiles
    for (int f = 0; f < fold_set_size(set); ++f) {
        const int last = 4 * (nh * kFtpLd + (fold_hmax(kFoldSetNB[set][f]) * (f + 1) - 1) * kFtpLd +
                              (kThreads / 64) * 16 - 1);
        if (last + 4 > stage) return false;
    }
    return true;
}
static_assert(set_reads_in_tables(1) && set_reads_in_tables(2), "fold set reads past the tables");

template <class Fn, int... I>
__device__ __forceinline__ void static_for_impl(Fn& fn, std::integer_sequence<int, I...>) {
    (fn(std::integral_constant<int, I>{}), ...);
}
// fn(std::integral_constant<int, i>) for i = 0 .. N-1, in order
template <int N, class Fn>
__device__ __forceinline__ void static_for(Fn fn) {
    static_for_impl(fn, std::make_integer_sequence<int, N>{});
}

#define CS_MFMA_SIZES(X) X(16) X(20) X(24) X(30) X(32) X(40) X(60) X(64) X(120) X(128)

__host__ __device__ constexpr bool mfma_size(int nb) {
    return nb == 16 || nb == 20 || nb == 24 || nb == 30 || nb == 32 || nb == 40 || nb == 60 ||
           nb == 64 || nb == 120 || nb == 128;
}

// LDS of the split-BF16 kernel: the stacked amplitudes, the DFT tiles, then
// (when the plan still fits in 160 KB) each wave's 2 KB landing area for the
// next tile's first conj(A) operands (32 bytes per lane)
constexpr size_t kWpfBytes = (size_t)(kThreads / 64) * 2 * 64 * 16;
__host__ __device__ inline MfmaLds mfma_lds_layout(int nharms, int dft_tiles) {
    MfmaLds L;
    L.ftp = (size_t)2 * nharms * kFtpLd * 4;
    L.dft = (size_t)dft_tiles * 2 * 64 * 16;
    L.region = L.ftp + L.dft + kWpfBytes <= (size_t)160 * 1024 ? kWpfBytes : 0;
    L.total = L.ftp + L.dft + L.region;
    return L;
}

// The interpolation operand tables as a by-value kernel argument: their
// addresses live in the kernel-argument segment, so wherever the compiler
// re-reads them under scalar-register pressure it does so with scalar loads.
// Read through the descriptor block they would be vector loads counted with
// the window prefetch and this wave's stores, whose wait drains all of those.
struct MfmaOps {
    const uint4* wA_hi;
    const uint4* wA_lo;
    const float2* gA;
    const int4* hdesc;
    const uint64_t* rowmask;
    const uint4* scol;   // [block trial]: eight int16 sampled columns (fold f: half-word f)
    const uint4* nv16;   // [global block]: eight uint16 valid counts
};

// KS: interpolation K steps (1 or 2); WIDE: some fold's bank has widths > 6
// (kept apart so the narrow kernel does not reserve the wide bounds' registers);
// SET: 0 dispatches each fold on its profile length at run time (an
// out-of-line pass per length), a fold set above 0 has its lengths fixed at
// compile time and every pass inlined, so no call boundary separates one
// fold's cross-lane and LDS latency from the next fold's work.
template <int KS, bool WIDE, bool DUMP, int SET>
__global__ __launch_bounds__(kThreads, 1) void k_fused_mfma(SpecDev sp, TabDev tb, GridDev g,
                                                            const FoldSet* __restrict__ fsd,
                                                            BatchDev b, FusedDev fu,
                                                            const MfmaDev* __restrict__ mdd,
                                                            MfmaOps ops) {
    // fold and table descriptors are read from device memory: indexing a
    // by-value kernel argument with a runtime fold index would copy it to scratch
    const FoldSet& fs = *fsd;
    const MfmaDev& md = *mdd;
    extern __shared__ __align__(16) char smem[];
    const int nh = tb.nharms;
    // nharms, a compile-time constant in the fold-set kernels
    const int nhc = SET > 0 ? kFoldSetNB[SET][0] / 2 : nh;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6) & 7;   // wave-uniform (scalar branches)
    uint32_t* fre = (uint32_t*)smem;   // [h][128 trials], (hi | lo << 16)
    uint32_t* fim = fre + nh * kFtpLd;
    const int ntc = md.dft_tiles_max;  // DFT tiles of all folds
    uint4* dhi = (uint4*)(smem + (size_t)2 * nh * kFtpLd * 4);
    uint4* dlo = dhi + (size_t)ntc * 64;
    // this wave's landing area for the next tile's first conj(A) operands
    // (q0 at 0, q1 at 64; 16 bytes per lane), used when the plan made room
    uint4* pfw = dlo + (size_t)ntc * 64 + (size_t)wv * 128;
    const bool wpf = md.wpf != 0;
    for (int i = tid; i < ntc * 64; i += kThreads) {
        dhi[i] = ldg4(md.cA_hi + i);
        dlo[i] = ldg4(md.cA_lo + i);
    }

    // tile schedule over the batch's tile list (ascending): contiguous list
    // ranges per XCD (workgroups are dealt to XCDs round-robin), so tiles of
    // neighbouring 16-trial groups, whose windows overlap, run together and
    // share that XCD's L2.  `tile` below is a list position.
    // (32-bit: a batch has far fewer than 2^31 tiles)
    int t_first, t_end, t_step;
    {
        const int G = gridDim.x, nx = (G % 8 == 0 && md.tile_order == 0) ? 8 : 1;
        if (nx > 1) {
            const int x = blockIdx.x % nx, l = blockIdx.x / nx;
            t_first = (int)(b.ntile * x / nx) + l;
            t_end = (int)(b.ntile * (x + 1) / nx);
            t_step = G / nx;
        } else {
            t_first = blockIdx.x;
            t_end = (int)b.ntile;
            t_step = G;
        }
    }

    const int j16 = lane & 15, o = j16 >> 1, c = j16 & 1, g4 = lane >> 4;
    // trial indices stay 32-bit: the host limits a search to total + 8 S < 2^32
    const uint32_t S = (uint32_t)md.S;
    const uint32_t gtotal = (uint32_t)g.total, gorigin = (uint32_t)g.origin;
    const int U = md.U;
    const CS_GLB uint32_t* plane = glb(md.spl) + (size_t)c * md.spl_ld + kSpecPad;
    const int64_t bmin = -kSpecPad, bmax = sp.namps + kSpecPad - 32 * KS;
    // the pre-split planes as one buffer resource (2 spl_ld words, under 2^31
    // bytes): a window is a 32-bit byte offset from this lane's plane column.
    // Lane group g4 reads window taps 4 g4 .. 4 g4 + 3 and 16 + 4 g4 ..
    // 16 + 4 g4 + 3 (interp_tap): the first load takes one contiguous 64-byte
    // run per occurrence and plane.
    const __amdgpu_buffer_rsrc_t srsc = __builtin_amdgcn_make_buffer_rsrc(
        (void*)md.spl, (short)0, (int)(2 * md.spl_ld * 4), 0x00020000);
    const uint32_t lane_off = 4u * (uint32_t)((int64_t)c * md.spl_ld + kSpecPad + 4 * g4);
    const int bmin32 = (int)bmin, bmax32 = (int)bmax;
    const uint32_t np = (uint32_t)g.Nprof;
    const uint32_t npm = (uint32_t)(0x100000000ull / np);   // floor(2^32 / np), np >= 16

    // Everything a tile reads from global memory before its compute: the
    // harmonic descriptors' window starts, the c2r trial's valid counts and
    // sampled columns for all folds, the row masks and every interpolation
    // window.  It is fetched for the next tile before this tile's outputs are
    // stored, so the loads are not ordered behind those stores.  Only the
    // descriptors are waited for here (the window addresses need them); every
    // other load is unconditional, from a clamped index, and its validity is
    // applied where it is used.  The table bases are kernel arguments, so
    // wherever they are re-read under register pressure that is a scalar load
    // from the argument segment (a reload from scratch would be a vector load,
    // and its wait would drain the tile's outstanding stores).
    const uint4* nvb = ops.nv16;
    const uint4* scb = ops.scol;
    const uint64_t* rmb = ops.rowmask;
    const int4* hdb = ops.hdesc;
    const int32_t* slb = b.slot;
    const uint32_t* tlb = b.tiles;
    const int nfv = fs.nf;
    // every fold's bank mask, lane f holding fold f's
    const uint32_t bmv = fs.bankmask[lane < nfv ? lane : nfv - 1];
    // the batch's sigma sample regions as one buffer (the host keeps it below
    // kNoStore bytes, so every offset is 32-bit and kNoStore drops a store)
    const __amdgpu_buffer_rsrc_t srs = __builtin_amdgcn_make_buffer_rsrc(
        (void*)fu.samples, (short)0, (int)(uint32_t)((int64_t)b.nblocks * nfv * fu.sld * 4),
        0x00020000);
    const uint32_t cg_last = (uint32_t)((g.total - 1 - g.origin) / g.Nprof);
    uint32_t st = 0, tbase = 0, t = 0, jj = 0;
    int occ = 0;
    int u = 0;
    bool tin = false;   // trial inside the search
    int32_t slotv = 0;  // selection position of the trial's block (-1: not selected)
    // the c2r trial's valid count and sampled column of every fold, packed
    // as uint16 / int16 (fold f in half-word f)
    uint4 nvq = make_uint4(0u, 0u, 0u, 0u), scq = make_uint4(0u, 0u, 0u, 0u);
    // the row masks of the interpolation lane's four trials 4 g4 + v: the
    // blocks of its first and last trial (the four are consecutive, so each
    // lies in one of the two); trial v is in the last one's when v >= rsplit
    uint64_t rm0 = 0, rm3 = 0;
    int rsplit = 4;
    uint4 wa[8], wb[8];
    // the harmonic descriptors of this wave's harmonics (window start x,
    // advance y, weight tile z, conj(A) row w), one value per lane (lane
    // 4s + c holds component c of harmonic slot s; slots past nharms repeat
    // the last harmonic), loaded one tile ahead of their use
    auto load_desc = [&](uint32_t tile) {
        const int uu = (int)((uint32_t)tile % (uint32_t)U);   // tile counts stay below 2^31
        const int s = (lane >> 2) & 7, h = wv + 1 + 8 * s;
        // a 32-bit element index (the table is far below 2^32 entries), so the
        // load takes the scalar base plus a per-lane offset
        const uint32_t idx = ((uint32_t)((h <= nh ? h : nh) - 1) * (uint32_t)U + (uint32_t)uu) * 4u +
                             (uint32_t)(lane & 3);
        return glb((const int*)hdb)[idx];
    };
    // tile number at list position k (clamped to the last one: past the
    // range a fetch reloads it), the same address in every lane
    auto tile_at = [&](int k) { return glb(tlb)[k < t_end ? k : t_end - 1]; };
    // idc: tile number of the next position to fetch; idn: of the one after
    // (in flight since the previous fetch)
    uint32_t idc = 0, idn = 0;
    int dnext = 0;
    int dcur = 0;   // the descriptors of the tile last fetched (its interpolation reads them)
    if (t_first < t_end) {
        idc = __builtin_amdgcn_readfirstlane(tile_at(t_first));
        dnext = load_desc(idc);
        idn = tile_at(t_first + t_step);
    }
    // Interpolation operands of one harmonic slot: the weight tiles (L2
    // resident) run two harmonics ahead in a ring of three; conj(A), eight
    // distinct 32-byte runs per load (L1 resident), one harmonic ahead in a
    // ring of two, which keeps eight registers free through the loop.
    struct Op {
        uint4 ah, al;
        int4 d;
    };
    struct OpA {
        uint4 q0, q1;
    };
    Op ring[3];
    OpA ringa[2];
    // this wave's descriptors of the tile, lane 4s + component, from the
    // fetch (v_readlane: no scalar load, so no wait that would also drain
    // this wave's LDS stores)
    auto desc = [&](int s) {
        return make_int4(__builtin_amdgcn_readlane(dcur, 4 * s), __builtin_amdgcn_readlane(dcur, 4 * s + 1),
                         __builtin_amdgcn_readlane(dcur, 4 * s + 2),
                         __builtin_amdgcn_readlane(dcur, 4 * s + 3));
    };
    auto issue = [&](int s, Op& x) {
        const int4 d = desc(s);
        x.d = d;
        x.ah = ldg4(ops.wA_hi + (size_t)d.z * 64 + lane);
        x.al = ldg4(ops.wA_lo + (size_t)d.z * 64 + lane);
    };
    auto issue_a = [&](int s, OpA& x) {
        const int cls = __builtin_amdgcn_readlane(dcur, 4 * s + 3) & kDescClass;
        const float2* gp = ops.gA + ((size_t)cls * 2 + c) * 16 + 4 * g4;
        x.q0 = ldg4(gp);
        x.q1 = ldg4(gp + 2);
    };
    auto fetch = [&](int tile) {
        const uint32_t tl = idc;
        const uint32_t stu = tl / (uint32_t)U;
        st = stu;
        u = (int)(tl - stu * (uint32_t)U);
        const int dreg = dnext;
        dcur = dreg;
        // the tile's first conj(A) operands straight into LDS (no registers
        // held through the folds): its interpolation's first complex product
        // does not wait on them
        if (wpf) {
            const int cls0 = __builtin_amdgcn_readlane(dreg, 3) & kDescClass;
            const float2* gp = ops.gA + ((size_t)cls0 * 2 + c) * 16 + 4 * g4;
            __builtin_amdgcn_global_load_lds((const CS_GLB void*)gp, (CS_LDS void*)pfw, 16, 0, 0);
            __builtin_amdgcn_global_load_lds((const CS_GLB void*)(gp + 2), (CS_LDS void*)(pfw + 64), 16, 0, 0);
        }
        // unconditional: a load under a branch leaves the merged wait state
        // pending at once
        idc = __builtin_amdgcn_readfirstlane(idn);
        dnext = load_desc(idc);
        idn = tile_at(tile + 2 * t_step);
        tbase = st * kOcc8 * S + 16u * (uint32_t)u;
        occ = (int)(kOcc8 * st) + o;
        t = tbase + (uint32_t)wv * S + (uint32_t)(lane & 15);
        tin = t < gtotal;
        // block and block trial of t: the wave's first trial divided on the
        // scalar unit, then at most one carry (the 16 trials of a wave span at
        // most two blocks, blocks holding at least 16 trials)
        uint32_t cgr;
        {
            uint32_t rb;
            const uint32_t qb = udiv_m(tbase + (uint32_t)wv * S - gorigin, np, npm, rb);
            const uint32_t rem = rb + (uint32_t)(lane & 15);
            const bool carry = rem >= np;
            cgr = qb + (carry ? 1u : 0u);
            jj = carry ? rem - np : rem;
        }
        const uint32_t cgc = cgr < cg_last ? cgr : cg_last;
        // whether the block is in this batch is decided where it is used
        slotv = glb(slb)[cgc];
        nvq = ldg4(nvb + cgc);
        scq = ldg4(scb + (uint32_t)jj);
        {
            uint32_t r0, rem;
            const uint32_t q0 = udiv_m(tbase - gorigin, np, npm, r0);
            const uint32_t cg = q0 + udiv_m(r0 + (uint32_t)o * S + 4u * (uint32_t)g4, np, npm, rem);
            const uint32_t cg3 = cg + (rem + 3 >= np ? 1u : 0u);
            rm0 = glb(rmb)[cg < cg_last ? cg : cg_last];
            rm3 = glb(rmb)[cg3 < cg_last ? cg3 : cg_last];
            rsplit = rem + 3 >= np ? (int)(np - rem) : 4;
        }
        // every window of this wave's harmonics (L2 / MALL / HBM)
        // (32-bit: tile occurrences stay within the planes' range, whose
        // product with an advance fits; the clamp keeps tail tiles inside)
        const int occ32 = (int)occ;
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            int b0 = __builtin_amdgcn_readlane(dreg, 4 * s) +
                     __builtin_amdgcn_readlane(dreg, 4 * s + 1) * occ32;
            b0 = b0 < bmin32 ? bmin32 : (b0 > bmax32 ? bmax32 : b0);
            const uint32_t vo = lane_off + 4u * (uint32_t)b0;
            // the second run: lane groups past the class's last nonzero weight
            // load past the buffer (zeros, no memory read)
            const int run2 = (__builtin_amdgcn_readlane(dreg, 4 * s + 3) >> kDescRun2Shift) & 7;
            wa[s] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(srsc, vo, 0, 0));
            wb[s] = __builtin_bit_cast(
                uint4, __builtin_amdgcn_raw_buffer_load_b128(srsc, g4 < run2 ? vo + 64u : kNoLoad, 0, 0));
        }
    };
    if (t_first < t_end) {
        fetch(t_first);
    }

    for (int tile = t_first; tile < t_end; tile += t_step) {
        // ---------------- interpolation: harmonics wv+1, wv+9, ... of this wave ----------
        // two copies of the loop: with the row guard applied, and without it
        // for tiles whose rows all pass (the usual case)
        // this lane's word in row wv of its plane; re-derived every tile (an
        // opaque copy) so the compiler keeps one base and adds each slot's
        // constant row offset instead of holding eight addresses
        CS_LDS uint32_t* wrow;
        {
            uint32_t a = (uint32_t)(uintptr_t)opaque_lds((CS_LDS uint32_t*)fre) +
                         4u * (uint32_t)((c ? nh * kFtpLd : 0) + wv * kFtpLd + o * 16 + 4 * g4);
            asm volatile("" : "+v"(a));
            wrow = (CS_LDS uint32_t*)(uintptr_t)a;
        }
        auto interp = [&](auto masked) {
        issue(0, ring[0]);
        issue(1, ring[1]);
        if (wpf) {
            ringa[0].q0 = pfw[lane];
            ringa[0].q1 = pfw[64 + lane];
        } else {
            issue_a(0, ringa[0]);
        }
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const int h = wv + 1 + 8 * s;
            if (s + 2 < 8) issue(s + 2, ring[(s + 2) % 3]);
            if (s + 1 < 8) issue_a(s + 1, ringa[(s + 1) % 2]);
            if (h > nhc) continue;
            const Op& x = ring[s % 3];
            const OpA& xa = ringa[s % 2];
            bf16x8 bh, bl;
            {
                const uint32_t w[8] = {wa[s].x, wa[s].y, wa[s].z, wa[s].w,
                                       wb[s].x, wb[s].y, wb[s].z, wb[s].w};
                unpack8(w, bh, bl);
            }
            f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
            acc = mfma3(as_bf16x8(x.ah), as_bf16x8(x.al), bh, bl, acc);
            if constexpr (KS == 2) {
                if ((x.d.w >> 28) == 2) {   // second 32-bin K step
                    int64_t b0 = (int64_t)x.d.x + (int64_t)x.d.y * occ;
                    b0 = b0 < bmin ? bmin : (b0 > bmax ? bmax : b0);
                    const CS_GLB uint32_t* p = plane + b0 + 32 + 4 * g4;
                    const uint4 x0 = ldg4_a4((const void*)p), x1 = ldg4_a4((const void*)(p + 16));
                    const uint32_t w[8] = {x0.x, x0.y, x0.z, x0.w, x1.x, x1.y, x1.z, x1.w};
                    bf16x8 ch, cl;
                    unpack8(w, ch, cl);
                    const size_t ti = ((size_t)x.d.z + 1) * 64 + lane;
                    acc = mfma3(as_bf16x8(ldg4(ops.wA_hi + ti)), as_bf16x8(ldg4(ops.wA_lo + ti)), ch, cl,
                                acc);
                }
            }
            const float gA[8] = {__builtin_bit_cast(float, xa.q0.x), __builtin_bit_cast(float, xa.q0.y),
                                 __builtin_bit_cast(float, xa.q0.z), __builtin_bit_cast(float, xa.q0.w),
                                 __builtin_bit_cast(float, xa.q1.x), __builtin_bit_cast(float, xa.q1.y),
                                 __builtin_bit_cast(float, xa.q1.z), __builtin_bit_cast(float, xa.q1.w)};
            uint32_t ow[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                // gA[k, g] * complex(ar, ai): the even lane holds re and forms
                // ar re - ai im, the odd lane holds im and forms ar im + ai re,
                // as one product and one fused multiply-add (the split-BF16
                // amplitudes carry ~1e-5 relative error, against which the
                // reference's separate rounding of the second product is nothing)
                const float mine = acc[v], oth = pair_swap(mine);
                float val = __builtin_fmaf(mine, gA[2 * v], oth * gA[2 * v + 1]);   // Im A pre-signed
                if constexpr (decltype(masked)::value)
                    val = ((((v >= rsplit) ? rm3 : rm0) >> (h - 1)) & 1ull) ? val : 0.0f;
                ow[v] = split_word(val);
            }
            CS_LDS uint32_t* dst = wrow + 8 * s * kFtpLd;   // row h - 1
            typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
            *(CS_LDS u32x2*)dst = u32x2{ow[0], ow[1]};
            *(CS_LDS u32x2*)(dst + 2) = u32x2{ow[2], ow[3]};
        }
        };
        // every harmonic row of the tile passes its guard (wave-uniform); the
        // row masks were loaded with the previous tile's prefetch
        const uint64_t need = nh >= 64 ? ~0ull : (1ull << nh) - 1ull;
        if (__all(((rm0 & rm3) & need) == need)) interp(BoolC<false>{});
        else interp(BoolC<true>{});
        lds_barrier();
        if constexpr (DUMP) {
            // diagnostic copy of the stacked amplitudes this wave's DFT reads:
            // occurrence wv, trial lane & 15 of the group, each word hi + lo
            const int i = lane & 15;
            const int64_t tw = (int64_t)tbase + (int64_t)wv * S + i;
            const int64_t rel = tw - fu.dump_t0;
            if (rel >= 0 && rel < fu.dump_n && tw < g.total) {
                for (int h = 1 + (lane >> 4); h <= nh; h += 4) {
                    const uint32_t wr = fre[(h - 1) * kFtpLd + wv * 16 + i];
                    const uint32_t wi = fim[(h - 1) * kFtpLd + wv * 16 + i];
                    const float re = __uint_as_float(wr << 16) + __uint_as_float(wr & 0xFFFF0000u);
                    const float im = __uint_as_float(wi << 16) + __uint_as_float(wi & 0xFFFF0000u);
                    fu.dump[(int64_t)(h - 1) * fu.dump_n + rel] = make_float2(re, im);
                }
            }
        }

        // ---------------- c2r + bound, fold by fold, wave-local --------------------------
        {
            const int tcol = wv * 16 + (lane & 15);
            const int nf = fs.nf;
            const CS_LDS uint32_t* lre = opaque_lds((CS_LDS uint32_t*)fre);
            const CS_LDS uint32_t* lim = opaque_lds((CS_LDS uint32_t*)fim);
            // the fold descriptors through pointers the compiler cannot follow
            // across tiles: their fields are re-read (scalar cache) where used
            // instead of being held in scalar registers for the whole kernel
            const CS_CONST FoldSet* fsx = uniform_opaque(fsd);
            const CS_CONST MfmaDev* mdx = uniform_opaque(mdd);
            // the trial's block in this batch (local block cl)
            const int clv = slotv - b.slot0;
            const bool inb = tin && (uint32_t)clv < (uint32_t)b.nblocks;
            const int cl = inb ? clv : 0;
            // bound column [local block][trial of the block], stored by lane
            // group 0 (byte offset in a fold's bound row; kNoStore: no store)
            const uint32_t moff = (inb && g4 == 0) ? 4u * ((uint32_t)cl * np + (uint32_t)jj) : kNoStore;
            // sampled columns of blocks that are not full (the last block, or a
            // fold cut at Nyquist) come from sample_column, the host table
            // holding those of full blocks: one check for all folds, so the
            // fold passes below are one region with no branch
            {
                bool part = false;
#pragma unroll
                for (int f = 0; f < kMaxFolds; ++f) {
                    const uint32_t nv = half_u(nvq, f);
                    part = part || (f < nf && inb && (uint32_t)jj < nv && nv != np);
                }
                if (__builtin_expect(__any(part), 0)) {
#pragma unroll
                    for (int f = 0; f < kMaxFolds; ++f) {   // compile-time f: scq stays in registers
                        const uint32_t nv = half_u(nvq, f);
                        if (f < nf && inb && (uint32_t)jj < nv && nv != np)
                            set_half(scq, f,
                                     (uint32_t)sample_column32((uint32_t)jj, nv, (uint32_t)fsx->f[f].nbins,
                                                               (uint32_t)fsx->cap[f]));
                    }
                }
            }
            // fold f of NB bins (NBC: the length as a compile-time constant,
            // or 0 to dispatch on the descriptor's length at run time)
            auto fold_step = [&](const int f, auto nbc, auto kfc, auto bmc) {
                constexpr int NBC = decltype(nbc)::value;
                constexpr int KF = decltype(kfc)::value;        // fold factor, 0: run time
                constexpr uint32_t BM = decltype(bmc)::value;   // bank mask, 0: run time
                const int NB = NBC > 0 ? NBC : fsx->f[f].nbins;
                const bool act = inb && (uint32_t)jj < half_u(nvq, f);
                // this trial's sampled column of the fold, if any: its slot block
                // in region (cl, f), columns round_up(NB, 4) slots apart
                const int sc = half_s(scq, f);
                const ColumnSink sk{
                    srs, (act && sc >= 0)
                             ? 4u * ((uint32_t)(cl * nf + f) * (uint32_t)fu.sld +
                                     (uint32_t)sc * (uint32_t)((NB + 3) & ~3))
                             : kNoStore};
                // fold factor and first DFT tile: compile-time constants in the
                // fold-set kernels (the launcher checked both), else read from
                // the descriptors; the bank mask from lane f of a register
                const int cb = SET > 0 ? set_cbase(SET, f) : mdx->c_base[f];
                const CS_LDS u32x4* th = opaque_lds((CS_LDS u32x4*)(dhi + (size_t)cb * 64));
                const CS_LDS u32x4* tl = opaque_lds((CS_LDS u32x4*)(dlo + (size_t)cb * 64));
                const int k = SET > 0 ? f + 1 : fsx->f[f].k;
                const uint32_t bm = (uint32_t)__builtin_amdgcn_readlane((int)bmv, f);
                const bool narrow = (bm >> 5) == 0;   // bank within widths 1, 2, 3, 4, 6
                float bound = -INFINITY;
                if constexpr (NBC > 0) {
                    bound = run_fold<NBC, WIDE, true, KF, BM>(narrow, lre, lim, th, tl, k, tcol, bm,
                                                              sk);
                } else {
                    switch (NB) {
#define X(SZ)                                                                  \
    case SZ:                                                                   \
        bound = run_fold<SZ, WIDE>(narrow, lre, lim, th, tl, k, tcol, bm, sk); \
        break;
                        CS_MFMA_SIZES(X)
#undef X
                        default:
                            break;
                    }
                }
                // the fold's bound of the trial, stored as soon as it is known
                // (row f of the bounds as its own buffer: offsets stay 32-bit)
                const __amdgpu_buffer_rsrc_t mrs = __builtin_amdgcn_make_buffer_rsrc(
                    (void*)(fu.mraw + (int64_t)f * fu.ldm), (short)0, (int)(fu.ldm * 4), 0x00020000);
                __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(act ? bound : -INFINITY), mrs,
                                                      moff, 0, 0);
            };
            if constexpr (SET == 0) {
#pragma unroll
                for (int f = 0; f < kMaxFolds; ++f)
                    if (f < nf)
                        fold_step(f, std::integral_constant<int, 0>{}, std::integral_constant<int, 0>{},
                                  std::integral_constant<uint32_t, 0u>{});
            } else {
                // the launcher checked that the search's folds are exactly this set
                static_for<fold_set_size(SET)>([&](auto fc) {
                    constexpr int F = decltype(fc)::value;
                    fold_step(F, std::integral_constant<int, kFoldSetNB[SET][F]>{},
                              std::integral_constant<int, set_reads_in_tables(SET) ? F + 1 : 0>{},
                              std::integral_constant<uint32_t, kFoldSetBM[SET][F]>{});
                });
            }
            // the next tile's loads
            if (tile + t_step < t_end) {
                fetch(tile + t_step);
            }
            // every wave is past its folds: the next tile's interpolation may
            // overwrite the stacked amplitudes
            lds_barrier();
        }
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

int mfma_sample_column(int64_t j, int64_t nv, int nbins, int64_t cap) {
    return sample_column(j, nv, nbins, cap);
}

MfmaLds mfma_lds_plan(int nharms, const FoldSet& fs, MfmaDev& md) {
    int tot = 0;
    for (int f = 0; f < fs.nf; ++f)
        tot += md.c_mt[f] * md.c_ks[f] * (geo::c2r_direct(fs.f[f].nbins) ? 1 : 2);
    md.dft_tiles_max = tot;
    md.nbp = 0;
    const MfmaLds L = mfma_lds_layout(nharms, tot);
    md.wpf = L.region > 0 ? 1 : 0;
    return L;
}

hipError_t launch_split_spectrum(hipStream_t st, const SpecDev& sp, uint32_t* spl) {
    const int64_t ld = sp.namps + 2 * kSpecPad;
    hipLaunchKernelGGL(k_split_spectrum, dim3((unsigned)((ld + 255) / 256)), dim3(256), 0, st,
                       sp.amps, sp.namps, ld, spl);
    return hipGetLastError();
}

__global__ void k_iota(uint32_t* __restrict__ out, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (uint32_t)i;
}

// slot[c] = position of block c in the ascending list sel, or -1
__global__ void k_slot_table(const int32_t* __restrict__ sel, int64_t nsel, int64_t nblocks,
                             int32_t* __restrict__ slot) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nblocks) return;
    int64_t lo = 0, hi = nsel;   // first position with sel >= c
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (sel[mid] < c) lo = mid + 1;
        else hi = mid;
    }
    slot[c] = (lo < nsel && sel[lo] == c) ? (int32_t)lo : -1;
}

hipError_t launch_iota(hipStream_t st, uint32_t* out, int64_t n) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_iota, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, out, n);
    return hipGetLastError();
}

hipError_t launch_slot_table(hipStream_t st, const int32_t* sel, int64_t nsel, int64_t nblocks,
                             int32_t* slot) {
    if (nblocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_slot_table, dim3((unsigned)((nblocks + 255) / 256)), dim3(256), 0, st, sel,
                       nsel, nblocks, slot);
    return hipGetLastError();
}

hipError_t launch_rowmask(hipStream_t st, const SpecDev& sp, const TabDev& tb, const GridDev& g,
                          int64_t nblocks, uint64_t* rowmask) {
    if (nblocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_rowmask, dim3((unsigned)((nblocks + 3) / 4)), dim3(256), 0, st, sp, tb, g,
                       nblocks, rowmask);
    return hipGetLastError();
}

// the fold set (kFoldSetNB) whose kernel runs this search's folds, 0 if none:
// the same profile lengths with fold factors 1, 2, ..., the stacked planes of
// the set's nharms and its DFT tile count (the LDS layout its gathers assume)
static int fold_set_of(const FoldSet& fs, const MfmaDev& md, int nharms) {
    if (md.generic) return 0;
    for (int set = 1; set < kFoldSets; ++set) {
        bool same = fs.nf == fold_set_size(set) && nharms == kFoldSetNB[set][0] / 2 &&
                    md.dft_tiles_max == set_tiles(set);
        for (int f = 0; f < fs.nf && same; ++f)
            same = fs.f[f].nbins == kFoldSetNB[set][f] && fs.f[f].k == f + 1 &&
                   md.c_base[f] == set_cbase(set, f) && fs.bankmask[f] == kFoldSetBM[set][f];
        if (same) return set;
    }
    return 0;
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
        const MfmaOps ops{md.wA_hi, md.wA_lo, md.gA, md.hdesc, md.rowmask, md.scol, md.nv16};
        hipLaunchKernelGGL(kern, dim3((unsigned)nwg), dim3(kThreads), shm, st, sp, tb, g, fs_dev, b,
                           fu, md_dev, ops);
        return hipGetLastError();
    };
    if (fu.dump) {
        if (md.ksmax == 2)
            return wide ? go(k_fused_mfma<2, true, true, 0>) : go(k_fused_mfma<2, false, true, 0>);
        if (!wide && fold_set_of(fs, md, tb.nharms) == 1) return go(k_fused_mfma<1, false, true, 1>);
        return wide ? go(k_fused_mfma<1, true, true, 0>) : go(k_fused_mfma<1, false, true, 0>);
    }
    if (md.ksmax == 2)
        return wide ? go(k_fused_mfma<2, true, false, 0>) : go(k_fused_mfma<2, false, false, 0>);
    if (!wide) {
        switch (fold_set_of(fs, md, tb.nharms)) {
            case 1: return go(k_fused_mfma<1, false, false, 1>);
            case 2: return go(k_fused_mfma<1, false, false, 2>);
            case 3: return go(k_fused_mfma<1, false, false, 3>);
            default: break;
        }
    }
    return wide ? go(k_fused_mfma<1, true, false, 0>) : go(k_fused_mfma<1, false, false, 0>);
}

}  // namespace cs
