// Block triage (features + MLP scorer) launchers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace cs {

constexpr int kTriageFeatures = 32;
constexpr int kTriageHidden = 64;
// W1 [64][32], b1 [64], w2 [64], b2
constexpr int kTriageParams = kTriageHidden * kTriageFeatures + 2 * kTriageHidden + 1;

// floats of the power pyramid (levels 1..6) for a spectrum of namps bins
int64_t triage_pyramid_floats(int64_t namps);

// features of blocks 0..nblocks-1 of the trial grid r_lo + t*lodr (Nprof
// trials per block, total trials); foldmask bit k = fold k is searched
// kind 0: power-pyramid cell maxima per integer fundamental; kind 1: nearest-bin
// power at h r_t per trial (no pyramid; pyr unused)
hipError_t launch_triage_features(hipStream_t st, int kind, const float2* amps, int64_t namps,
                                  float* pyr, double r_lo, double lodr, int64_t Nprof,
                                  int64_t total, int nh, uint32_t foldmask, int64_t nblocks,
                                  float* out);

// waves (of 32 rows) of an MLP pass over n rows; the partial-gradient buffer
// holds mlp_partial_rows(n) rows of kTriageParams + 1 floats
int64_t mlp_waves(int64_t n);
inline int64_t mlp_partial_rows(int64_t n) { return (mlp_waves(n) + 3) / 4 * 4; }

// logits (if non-null) and, with y / partial / grad, the gradient of the mean
// binary cross-entropy (grad[kTriageParams] = loss)
hipError_t launch_mlp(hipStream_t st, const float* params, const float* norm, const float* X,
                      const float* y, int64_t n, float* logits, float* partial, float* grad);
// Adam step `step` (1-based) of n parameters from grad (loss at grad[n] -> losses[it])
hipError_t launch_adam(hipStream_t st, float* p, float* m1, float* m2, const float* grad, int n,
                       float lr, float b1, float b2, float eps, int step, float* losses, int it);

}  // namespace cs
