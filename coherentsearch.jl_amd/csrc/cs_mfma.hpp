// Split-BF16 MFMA fused search: operand tables on the device and the launcher.
#pragma once
#include "cs_fused.hpp"
#include "cs_gpu.hpp"

namespace cs {

struct MfmaDev {
    const uint4* wA_hi;      // interpolation weight tiles, [tile][64 lanes] x 8 bf16
    const uint4* wA_lo;
    const float2* gA;        // [class][16] conj(A(dr)) per trial of the group
    const int32_t* clsmap;   // [(h-1)*q + residue] -> class
    const int32_t* cls_tile; // [class] -> first tile
    const int32_t* ks;       // [h-1] -> K steps
    const int32_t* clsu;     // [(h-1)*U + u] -> {first tile, class}
    const uint4* cA_hi;      // c2r DFT tiles, [tile][64 lanes] x 8 bf16
    const uint4* cA_lo;
    int32_t c_base[kMaxFolds];
    int32_t c_mt[kMaxFolds];
    int32_t c_ks[kMaxFolds];
    int64_t S;               // occurrence stride, lcm(16, q) trials
    int32_t U;               // S / 16
    int32_t q;
    int32_t kwin;            // staged bins per (harmonic, occurrence): 32 * max K steps
    int32_t nbp_max;         // set by the launcher
    int32_t dft_tiles_max;   // set by the launcher
    int32_t pad;
};

bool mfma_supported(const FoldSet& fs, int nharms);
size_t mfma_lds_bytes(int nharms, const FoldSet& fs, int kwin, int* nbp_max, int* dft_max,
                      const MfmaDev& md);
// md.nbp_max / md.dft_tiles_max must be filled by mfma_lds_bytes; fs_dev and
// md_dev are device copies of fs and md.
hipError_t launch_fused_mfma(hipStream_t st, const SpecDev& sp, const TabDev& tb, const GridDev& g,
                             const FoldSet& fs, const BatchDev& b, const FusedDev& fu,
                             const MfmaDev& md, const FoldSet* fs_dev, const MfmaDev* md_dev,
                             int64_t nwg);

}  // namespace cs
