// Device-side structures and kernel launchers of the coherent-search engine.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "cs_plan.hpp"

namespace cs {

// The spectrum as the kernels see it.
struct SpecDev {
    const float2* amps;
    int64_t namps;
    int64_t N;
    int64_t Nhalf;
};

// Interpolation tables on the device (see HarmDev).
struct TabDev {
    const HarmDev* harm;
    const int32_t* row;
    const uint8_t* carry;
    const float* W32;
    const float2* A32;
    const double* W64;
    const double2* A64;
    double lodr;
    int32_t m;
    int32_t nharms;
    int32_t weights_f64;
    int32_t pad;
};

// Block geometry: block c spans global trials [origin + c*Nprof,
// min(origin + (c+1)*Nprof, total)).
struct GridDev {
    int64_t origin;
    int64_t Nprof;
    int64_t total;
};

// The folds searched and where the current batch keeps their profiles.
struct FoldSet {
    int32_t nf;
    int32_t pad;
    int64_t ldp;                     // leading dimension of every fold's profiles
    FoldDev f[kMaxFolds];
    float* prof[kMaxFolds];          // FP32 profiles, [bin][batch trial]
    const double2* twid[kMaxFolds];  // (cos, sin)(2 pi e / nbins), e < nbins
    int32_t cap[kMaxFolds];          // sigma sample-buffer length per fold
    uint32_t bankmask[kMaxFolds];    // widths as a subset of the default bank
};

// Where a batch sits in the search.  A batch is either a run of consecutive
// blocks (c_first .. c_first + nblocks - 1, trials t_base .. t_base + nt - 1)
// or, on the split-BF16 path, a list of selected blocks: the search numbers
// its selected blocks 0, 1, ... (slot[global block], -1 where unselected;
// sel[position] = global block) and the batch holds positions slot0 ..
// slot0 + nblocks - 1.  The split-BF16 path always passes slot/sel/tiles
// (identity tables when every block is selected).
struct BatchDev {
    int64_t t_base;     // global trial of batch column 0 (runs of blocks)
    int64_t nt;         // trials in the batch (runs of blocks)
    int64_t c_first;    // global block index of the batch's first block (runs)
    int32_t nblocks;    // blocks in the batch
    int32_t slot0;      // selection position of local block 0
    const int32_t* nvalid;   // [global block * nf + f]
    double* sigma;           // [local block * nf + f]
    const int32_t* sel;      // selection position -> global block, or null
    const int32_t* slot;     // global block -> selection position or -1, or null
    const uint32_t* tiles;   // split-BF16 kernel: the 128-trial tiles to run
    int64_t ntile;           //   (tiles touching any block of the batch)
};

// global block of local block cl, and local block of global block cg (-1 if
// the block is not in the batch)
__host__ __device__ inline int64_t batch_block(const BatchDev& b, int cl) {
    return b.sel ? (int64_t)b.sel[b.slot0 + cl] : b.c_first + cl;
}
__host__ __device__ inline int batch_slot(const BatchDev& b, int64_t cg) {
    const int64_t cl = b.slot ? (int64_t)b.slot[cg] - b.slot0 : cg - b.c_first;
    return (cl >= 0 && cl < b.nblocks) ? (int)cl : -1;
}

// Trial of batch column j: a run of blocks maps column j to t_base + j; a
// block list (plain pipeline over selected blocks) maps it to trial j % Nprof
// of block sel[slot0 + j / Nprof], and a column past the band's last trial to -1.
__host__ __device__ inline int64_t batch_trial(const BatchDev& b, const GridDev& g, int64_t j) {
    if (!b.sel) return b.t_base + j;
    const int64_t cl = j / g.Nprof;
    const int64_t t = g.origin + (int64_t)b.sel[b.slot0 + cl] * g.Nprof + (j - cl * g.Nprof);
    return t < g.total ? t : -1;
}

struct OutDev {
    uint64_t* surv;        // packed (t << 4) | fold
    uint32_t* surv_count;
    uint32_t surv_cap;
    uint32_t cand_cap;
    struct cs_cand_dev* cands;
    uint32_t* cand_count;
    double* mall;          // statistics mode: every rescanned metric at
    int64_t mall_ld;       //   mall[fold * mall_ld + t] (else null)
    double threshold;
    double medcut;
    // statistics search: survivors are the trials whose metric in gm
    // ([fold][trial], NaN where invalid) exceeds gmcut, instead of the bound
    const double* gm;
    int64_t gm_ld;
    double gmcut;
};

struct cs_cand_dev {
    int64_t t;
    int32_t k;
    int32_t nharm;
    int64_t block;
    double metric;
};

// ---- launchers (all asynchronous on `st`) ------------------------------------
// columns j < nt: trial t_base + j, or with `sel` trial j % Nprof of block
// sel[slot0 + j / Nprof] (columns past the band's end are zero)
hipError_t launch_interp(hipStream_t st, const SpecDev& sp, const TabDev& tb,
                         const GridDev& g, int64_t t_base, int64_t nt, bool out64,
                         void* out, int64_t ld, const int32_t* sel = nullptr, int32_t slot0 = 0);
hipError_t launch_c2r_f32(hipStream_t st, const float2* ftp, int64_t ldf, int k,
                          int nbins, int64_t nt, float* prof, int64_t ldp,
                          const double2* twid);
hipError_t launch_sigma_f32(hipStream_t st, const FoldSet& fs, const GridDev& g,
                            const BatchDev& b);
// where the search's c2r puts the sampled columns of fold f (a run-of-blocks
// batch starting at block c_first, trial t_base)
struct SampleSink {
    float* samples;
    int64_t sld;
    const int32_t* nvalid;   // [global block * nf + f]
    GridDev g;
    int64_t t_base;
    int64_t c_first;
    int32_t f, nf;
    int32_t cap;
    int32_t slot0;           // with sel: the batch's first selection position
    const int32_t* sel;      // block list of the batch (see batch_trial), or null
};
bool c2r_samples_supported(int nbins);
hipError_t launch_c2r_samples(hipStream_t st, const float2* ftp, int64_t ldf, int k, int nbins,
                              int64_t nt, float* prof, int64_t ldp, const SampleSink& sk);
hipError_t launch_gate_f32(hipStream_t st, const FoldSet& fs, const GridDev& g,
                           const BatchDev& b, const OutDev& o);
hipError_t launch_rescan(hipStream_t st, const SpecDev& sp, const TabDev& tb,
                         const GridDev& g, const FoldSet& fs, const BatchDev& b,
                         const OutDev& o, int nblocks_grid);
hipError_t launch_reset_u32(hipStream_t st, uint32_t* p, int n);
// Zero two u32 ranges with one launch (q may be null)
hipError_t launch_reset2_u32(hipStream_t st, uint32_t* p, int n, uint32_t* q, int m);
hipError_t launch_copy_u32(hipStream_t st, const uint32_t* src, uint32_t* dst);
// Scratch of the FP64 sigma refinement: flags (nblocks * nf + 1 words, the
// last the list count; cleared by the launch itself), the list of flagged
// pairs (nblocks * nf words) and FP64 samples of the first `cap` listed pairs
// (cap * kSigmaSamples doubles).
constexpr int kRefineChunks = 8;      // column chunks (workgroups) per listed pair
constexpr uint32_t kRefineCap = 2048; // listed pairs the sample scratch holds at most
struct RefineBufs {
    uint32_t* flags;
    uint32_t* list;
    double* samples;
    uint32_t cap;
    bool flags_clear;   // the batch's flags and count are already zero (no reset launch)
};
// FP64 sigma of the batch's (block, fold) pairs that hold gate survivors
hipError_t launch_sigma_refine(hipStream_t st, const SpecDev& sp, const TabDev& tb,
                               const GridDev& g, const FoldSet& fs, const BatchDev& b,
                               const OutDev& o, const RefineBufs& rb);

// FP64 bridge path (chunk_metrics / decim / block_metrics).
hipError_t launch_c2r_f64(hipStream_t st, const double2* ftp, int64_t ldf, int k,
                          int nbins, int64_t n, double* prof, int64_t ldp,
                          const double2* twid, double scale);
hipError_t launch_sigma_f64(hipStream_t st, const double* prof, int64_t ldp,
                            int nbins, int64_t n, int64_t cap, double* sigma_out);
// profile j's bin i at prof[i * ldp + j * ldt]
hipError_t launch_exact_f64(hipStream_t st, const double* prof, int64_t ldp, int nbins,
                            int64_t n, const int32_t* widths, int nw,
                            const double* sigma, double* out, int64_t ldt = 1);
hipError_t launch_point_interp_f64(hipStream_t st, const SpecDev& sp, const double* rfund,
                                   int64_t L, int nharms, int m, const uint8_t* rowok,
                                   double2* out, int64_t ld);
// block_metrics(:fft): linear interpolation between fine-grid points
// lob[h-1] + i/nb (i < kgrid[h-1]) of FP64 Eqn. 30
hipError_t launch_point_interp_fft(hipStream_t st, const SpecDev& sp, const double* rfund,
                                   int64_t L, int nharms, int m, int nb, const int64_t* lob,
                                   const int64_t* kgrid, const uint8_t* rowok, double2* out,
                                   int64_t ld);
// candidate profiles: H[c] harmonics (-1: skip), the width bank of candidate c
// at wtab[woff[c] .. woff[c+1]) (widths <= 2 H[c]); profiles to prof_out[c *
// ld ..] when non-null, best widths to best_w
constexpr int kMaxCandH = 512;
hipError_t launch_cand_profiles(hipStream_t st, const SpecDev& sp, const double* r,
                                const int32_t* H, int64_t n, int m, const int32_t* wtab,
                                const int32_t* woff, double* prof_out, int64_t ld,
                                int32_t* best_w);

}  // namespace cs
