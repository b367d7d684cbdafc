// Split-BF16 MFMA fused search: operand tables on the device and the launcher.
#pragma once
#include "cs_fused.hpp"
#include "cs_gpu.hpp"

namespace cs {

// Zero bins on each side of the pre-split spectrum planes: every 16-trial
// interpolation window (at most 64 bins) starting in [-kSpecPad, namps] reads
// inside the allocation.
constexpr int kSpecPad = 64;
// Sample region per (block, fold) of the MFMA mode: at most cap / nbins columns
// of round_up(nbins, 4) slots (whole 16-byte stores); the largest is 273
// columns of 32 slots at 30 bins (8736), 8960 keeps regions 256-byte aligned.
constexpr int kSigmaSampleAlign = 4;
constexpr int kSigmaSamplesPadded = 8960;
// The split-BF16 kernel stores bounds and sigma samples through buffer
// resources with 32-bit byte offsets, an offset of kMfmaNoStore or more being a
// dropped store: a batch's sample regions, and each fold's row of bounds, stay
// below that many bytes.
constexpr uint32_t kMfmaNoStore = 0xFFFFF000u;

struct MfmaDev {
    const uint4* wA_hi;       // interpolation weight tiles, [tile][64 lanes] x 8 bf16
    const uint4* wA_lo;
    const float2* gA;         // [class][parity][16] conj(A(dr)) per trial of the group, Im
                              // negated for parity 0 (the lanes that hold Re)
    const int4* hdesc;        // [(h-1)*U + u]: {first bin of trial 16u (clamped),
                              //  bin advance per occurrence, first weight tile,
                              //  class | K steps << 28}
    const uint4* cA_hi;       // c2r DFT tiles, [tile][64 lanes] x 8 bf16
    const uint4* cA_lo;
    const uint32_t* spl;      // split spectrum: [plane re/im][kSpecPad + bin], (hi | lo << 16)
    const uint64_t* rowmask;  // [global block]: bit h-1 set if harmonic row h is kept
    const uint4* scol;        // [block trial]: sampled column of a full block per fold,
                              //   eight int16 (fold f in half-word f), -1: none
    const uint4* nv16;        // [global block]: valid trials per fold, eight uint16
    int64_t spl_ld;           // plane stride (namps + 2 * kSpecPad)
    int64_t S;                // occurrence stride, lcm(16, q) trials
    int64_t ntiles;           // (ceil(total / (8 S))) * U tiles of 128 trials in the search
    int32_t c_base[kMaxFolds];
    int32_t c_mt[kMaxFolds];
    int32_t c_ks[kMaxFolds];
    int32_t U;                // S / 16
    int32_t q;
    int32_t ksmax;            // largest interpolation K steps (1 or 2)
    int32_t nbp;              // profile row stride in LDS (floats)
    int32_t dft_tiles_max;    // c2r DFT tiles of all folds (hi or lo)
    int32_t tile_order;       // 0: contiguous tile-list ranges per XCD; 1: round-robin
    int32_t generic;          // 1: the run-time-dispatch kernel even for a known fold set
    int32_t wpf;              // 1: the next tile's first conj(A) operands are prefetched into LDS
};

struct MfmaLds {
    size_t ftp, region, dft, total;
};

bool mfma_supported(const FoldSet& fs, int nharms);
// Sampled column of block trial j (whole-column subsample of the sigma estimate), -1 if none.
int mfma_sample_column(int64_t j, int64_t nv, int nbins, int64_t cap);
// Fills md.nbp / md.dft_tiles_max and returns the LDS plan.
MfmaLds mfma_lds_plan(int nharms, const FoldSet& fs, MfmaDev& md);
// Pre-split spectrum planes (2 * (namps + 2 * kSpecPad) words at spl).
hipError_t launch_split_spectrum(hipStream_t st, const SpecDev& sp, uint32_t* spl);
// Per-block harmonic row guards as bit masks (nharms <= 64).
hipError_t launch_rowmask(hipStream_t st, const SpecDev& sp, const TabDev& tb, const GridDev& g,
                          int64_t nblocks, uint64_t* rowmask);
// out[i] = i for i < n (identity block selection and tile lists).
hipError_t launch_iota(hipStream_t st, uint32_t* out, int64_t n);
// slot[c] = position of block c in the ascending block list sel[0, nsel), or -1.
hipError_t launch_slot_table(hipStream_t st, const int32_t* sel, int64_t nsel, int64_t nblocks,
                             int32_t* slot);
// fs_dev and md_dev are device copies of fs and md; nwg persistent workgroups.
hipError_t launch_fused_mfma(hipStream_t st, const SpecDev& sp, const TabDev& tb, const GridDev& g,
                             const FoldSet& fs, const BatchDev& b, const FusedDev& fu,
                             const MfmaDev& md, const FoldSet* fs_dev, const MfmaDev* md_dev,
                             int nwg);

}  // namespace cs
