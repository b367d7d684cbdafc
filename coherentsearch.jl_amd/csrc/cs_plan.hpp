// Host-side planning for the coherent-search engine: the exact trial grid,
// per-harmonic interpolation tables, boxcar banks and fold bookkeeping.
// Shared structs are laid out for direct upload to the device.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace cs {

constexpr int kSigmaSamples = 8192;   // _BOXCAR_SIGMA_SAMPLES, search.jl:466
constexpr int kLadderWmax = 6;        // _LADDER_WMAX, search.jl:401
constexpr int kMaxWidths = 24;
constexpr int kMaxFolds = 8;
constexpr int kQmax = 1 << 16;        // DIRECT_QMAX, directinterp.jl:111

// Split-BF16 interpolation operands: element e (0..7) of lane group g (lanes
// 16g .. 16g + 15) holds K slot interp_tap(g, e) of a 32-slot step, window
// taps 4g .. 4g + 3 then 16 + 4g .. 16 + 4g + 3, so the first four words of
// the four lane groups are one contiguous 64-byte run of the window (the
// contraction sums over K, so its order is free).
constexpr int interp_tap(int g, int e) { return e < 4 ? 4 * g + e : 12 + 4 * g + e; }
// The last word of a harmonic descriptor: class index in bits 0..23; in bits
// 24..26 the lane groups whose second run (taps 16 + 4g .. 16 + 4g + 3 of
// the first K step) holds a nonzero weight of some trial of the class (0 to
// 4: only those load it); K steps << 28.
constexpr int kDescRun2Shift = 24;
constexpr int32_t kDescClass = (1 << kDescRun2Shift) - 1;

// One harmonic's recipe (DirectPlan, directinterp.jl:179-204).  Phase index p
// (0-based) runs over the P residues the cycle visits; row[row_off + res] = p.
struct HarmDev {
    int64_t rfloor0;   // floor(h * r_lo)
    double dr0;        // h*r_lo - rfloor0 (used only by the per-trial fallback)
    int32_t q;         // trial-grid denominator
    int32_t s;         // (h*pnum) mod q
    int32_t base_adv;  // (h*pnum) div q
    int32_t hp;        // h*pnum
    int32_t P;         // phase-cycle length; 0 => per-trial FP64 fallback
    int32_t row_off;   // into the residue->phase table (q entries)
    int32_t tab_off;   // into carry/A (P entries) and W (P*m entries)
    int32_t pad;
};

struct Plans {
    int nharms = 0, m = 0;
    double hidr = 0.5, r_lo = 0.0, lodr = 0.0;
    bool rational = false;
    int64_t pnum = 0, q = 0;
    std::vector<HarmDev> harm;
    std::vector<int32_t> row;      // sum of q over harmonics
    std::vector<uint8_t> carry;    // sum of P
    std::vector<float> W32;        // sum of P*m
    std::vector<float> A32;        // 2 * sum of P (interleaved)
    std::vector<double> W64;
    std::vector<double> A64;
};

// sin(pi x), cos(pi x) with exact argument reduction.
double sinpi_d(double x);
double cospi_d(double x);

bool trial_grid_rational(double lodr, int64_t* pnum, int64_t* q);
Plans build_plans(int nharms, int m, double hidr, double r_lo);

std::vector<int> boxcar_widths(int nbins, double fsp, double maxfrac);
std::vector<int> ladder_boxcar_widths(int nbins, int k, const std::vector<int>& decims,
                                      double fsp, double maxfrac);

// Operand tables of the split-BF16 MFMA pipeline (16x16x32 bf16 tiles in the
// hardware lane layout: lane l holds row l%16, k = 8*(l/16) .. +8).
//  * interpolation: for each harmonic h and each residue class of 16-trial
//    groups, the 16 x (m + delta) weight block Wx[i][k] = W[k - delta_i, p_i],
//    split into bf16 hi + lo, plus the per-trial conj(A(dr)) factors;
//  * c2r, profile lengths NB > 64: the even/odd halves of the numpy-convention
//    real inverse DFT, rows n = 0..H, columns h = 1..H:
//      Ccos[n][h] = c_h cos(2 pi h n / NB),  Csin[n][h] = -c_h sin(2 pi h n / NB)
//    (c_h = 2, c_H = 1, Csin[.][H] = 0), so x[n] = E + O and x[NB-n] = E - O
//    with E = Ccos . Re X and O = Csin . Im X; split likewise.  Row i = 4g + v
//    of row tile a is entry n = g*4*mt + 4a + v (mt row tiles); tiles
//    [(a*ks + kk)*2 + {cos, sin}].
//  * c2r, NB <= 64 (direct layout, cs_geo.hpp DGeo): one real DFT with
//    K slots (Re h = 1..H | Im h = 1..H-1), row i = 4g + v of row tile a is
//    bin (g*RD + 4a + v) mod NB, RD = ceil(NB/4); tiles [a*ks + kk].  The row
//    tile count depends on the fold's widest boxcar (fold_wm).
struct MfmaTables {
    bool ok = false;
    std::string why;
    int64_t S = 0;                  // occurrence stride lcm(16, q) (trials)
    int32_t U = 0;                  // S / 16
    int32_t q = 0;
    std::vector<uint16_t> wA_hi, wA_lo;   // [tile][64 lanes][8]
    std::vector<float> gA;                // [class][parity][16][2]: (Re A, -Im A | Im A)
    std::vector<int32_t> clsmap;          // [(h-1)*q + residue] -> class, -1 if no group starts there
    std::vector<int32_t> cls_tile;        // [class] -> first tile
    std::vector<uint8_t> cls_run2;        // [class] -> lane groups of the second window run read
    std::vector<int32_t> ks;              // [h-1] -> K steps (1 or 2)
    std::vector<int32_t> clsu;            // [(h-1)*U + u] -> {first tile, class} of the group at 16u
    std::vector<int32_t> hdesc;           // [(h-1)*U + u] x {first bin of trial 16u, bin advance
                                          //  per occurrence, first tile, class | second-run
                                          //  lane groups << 24 | K steps << 28}
    std::vector<uint16_t> cA_hi, cA_lo;   // c2r tiles of every fold, rows permuted
    int32_t c_base[8] = {0}, c_mt[8] = {0}, c_ks[8] = {0};
};

// One decimation fold of the search (k = 1 is the base pass).
struct FoldDev {
    int32_t k, Hk, nbins, nw;
    int32_t widths[kMaxWidths];
};

// fold_wm[f]: widest boxcar the fold's pass evaluates (geo fold_wm);
// fold_k[f]: the fold factor (orders the direct tiles' K slots)
MfmaTables build_mfma_tables(const Plans& pl, const std::vector<int>& fold_nbins,
                             const std::vector<int>& fold_wm, const std::vector<int>& fold_k);

}  // namespace cs
