// Fused search pipeline: launchers and the buffers it leaves in HBM.
#pragma once
#include "cs_gpu.hpp"

namespace cs {

struct FusedDev {
    float* mraw;            // [fold][batch column]: zero-baseline bound before sigma
    int64_t ldm;            // leading dimension of mraw
    float* samples;         // [(block_local * nf + fold) * sld + slot]: sigma samples |x|
    int64_t sld;            // samples per (block, fold) region
    int32_t salign;         // sampled columns start at multiples of round_up(nbins, salign)
    int32_t pad;
    const uint8_t* rowok;   // [block_local * nharms + h - 1]: harmonic row kept for the block
    // split-BF16 kernel, diagnostic instantiation only: the stacked amplitudes
    // (hi + lo, as the DFT reads them) of trials dump_t0 .. dump_t0 + dump_n - 1
    // at dump[(h - 1) * dump_n + trial - dump_t0]; null otherwise
    float2* dump;
    int64_t dump_t0;
    int64_t dump_n;
};

bool fused_supported(const FoldSet& fs, int nharms);
hipError_t launch_guard(hipStream_t st, const SpecDev& sp, const TabDev& tb, const GridDev& g,
                        const BatchDev& b, uint8_t* rowok);
hipError_t launch_fused(hipStream_t st, const SpecDev& sp, const TabDev& tb, const GridDev& g,
                        const FoldSet& fs, const BatchDev& b, const FusedDev& fu);
hipError_t launch_select(hipStream_t st, const FoldSet& fs, const GridDev& g, const BatchDev& b,
                         const FusedDev& fu, const OutDev& o);

}  // namespace cs
