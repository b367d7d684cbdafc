// Metric statistics over every trial (the --metricstats diagnostic) and the
// normalised second pass (--normalize).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cs_gpu.hpp"

namespace cs {

constexpr int kStatsMaxBins = 12288;   // histogram bins (one LDS copy per workgroup)

// Per (fold, block): exact moments, extremes and median of the block's valid
// trial metrics, and the trials outside the histogram range.
struct BlockStatDev {
    int64_t n, under, over, pad;
    double vmin, vmax, sum, sumsq, mean, median, stdv;
};

// metrics[f * ld + first + j], j < nvalid[block * nf + f] (the valid prefix);
// block b of fold f adds its histogram to counts[(f * nwin + win[f * nblocks +
// b]) * nb + bin] (bins of width (hi - lo) / nb from lo).
hipError_t launch_block_stats(hipStream_t st, const double* metrics, int64_t ld,
                              const int32_t* nvalid, int nf, int64_t Nprof, int64_t nblocks,
                              const int32_t* win, int nwin, double lo, double hi, int nb,
                              uint32_t* counts, BlockStatDev* out);

// Every valid trial's exact metric of fold fidx from a contiguous batch's FP32
// profiles (prof[bin][batch trial], stride ldp) into mall[fidx * ld + t]; the
// fold's widths are the default-bank entries of bankmask (widths up to 0.3
// nbins).
bool exact_all_supported(int nbins);
hipError_t launch_exact_all(hipStream_t st, const float* prof, int64_t ldp, int nbins,
                            uint32_t bankmask, int fidx, int nf, const GridDev& g,
                            const BatchDev& b, double* mall, int64_t ld);

// sigma of a set of n profiles (bin i of profile j at prof[i * ldp + j * ldt])
// by the search's subsampling rule with cap samples, about zero (center_median
// 0) or about the sample median (1); buf holds min(nbins n, cap) doubles.
hipError_t launch_sigma_any(hipStream_t st, const double* prof, int64_t ldp, int nbins, int64_t n,
                            int64_t cap, int center_median, double* buf, double* sigma_out,
                            int64_t ldt = 1);

// Normalised detection over the resident metrics: for every valid (fold f,
// trial t), r = k * (r_lo + i0 lodr + j lodr) (r_lo + ... alone for k = 1),
// window = last edge <= r / T (clamped), z = (M - loc) / scale, and (t, k,
// Hk, block, z) is appended when z > threshold.
struct NormDev {
    const double* metrics;
    int64_t ld, total, Nprof;
    double r_lo, lodr, T, threshold;
    const double* edges;   // [f][nwin + 1]
    const double* loc;     // [f][nwin]
    const double* scale;   // [f][nwin]
    int nwin, nf;
    int32_t k[kMaxFolds], Hk[kMaxFolds];
};
hipError_t launch_normalize(hipStream_t st, const NormDev& nd, cs_cand_dev* cands,
                            uint32_t* count, uint32_t cap);

}  // namespace cs
