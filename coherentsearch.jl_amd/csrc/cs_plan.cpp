// Host-side planning: exact rational trial grid and per-harmonic tables.
//
// The tables are the reference's DirectPlan (directinterp.jl:179-204,
// 345-405): for harmonic h the trial grid r_t = r_lo + t*pnum/q visits only
// P = q/gcd(h*pnum, q) fractional offsets dr, so the m weights 1/(dr - j) and
// the scalar conj(sin(pi dr) cispi(dr)/pi) are tabulated per phase.  Every
// entry is computed in FP64 and rounded once to FP32 (the FP64 copy serves the
// Float64-weight bridge).  The device evaluates each trial as one m-term FMA
// chain against its phase's column; the reference's trials-axis group tables
// (directinterp.jl:237-282) are a reordering of exactly those FMAs and need no
// counterpart here.
#include "cs_plan.hpp"
#include "cs_geo.hpp"

#include <algorithm>
#include <cmath>
#include <cfloat>
#include <cstring>
#include <numeric>

namespace cs {

double sinpi_d(double x) {
    double r = std::fmod(x, 2.0);
    if (r < 0) r += 2.0;
    bool neg = false;
    if (r >= 1.0) { r -= 1.0; neg = true; }
    double v;
    if (r <= 0.25) v = std::sin(M_PI * r);
    else if (r <= 0.75) v = std::cos(M_PI * (r - 0.5));
    else v = std::sin(M_PI * (1.0 - r));
    return neg ? -v : v;
}

double cospi_d(double x) {
    double r = std::fabs(std::fmod(x, 2.0));
    if (r > 1.0) r = 2.0 - r;
    if (r <= 0.25) return std::cos(M_PI * r);
    if (r <= 0.75) return -std::sin(M_PI * (r - 0.5));
    return -std::cos(M_PI * (1.0 - r));
}

static int64_t floordiv(int64_t a, int64_t b) {
    int64_t q = a / b;
    if ((a % b) != 0 && ((a < 0) != (b < 0))) --q;
    return q;
}

// Simplest rational within 1e-13 (continued-fraction convergents, as Julia's
// rationalize), accepted only if q <= 2^16 and p/q reproduces lodr to 8 ulp
// (trial_grid_rational, directinterp.jl:217-225).
bool trial_grid_rational(double lodr, int64_t* pnum, int64_t* q) {
    if (!(lodr > 0)) return false;
    const double tol = 1e-13;
    int64_t p0 = 0, q0 = 1, p1 = 1, q1 = 0, P = 0, Q = 1;
    double y = lodr;
    for (int it = 0; it < 64; ++it) {
        double a = std::floor(y);
        if (a > 4e18) break;
        int64_t ai = (int64_t)a;
        int64_t p2 = ai * p1 + p0, q2 = ai * q1 + q0;
        if (q2 <= 0 || q2 > ((int64_t)1 << 40)) break;
        p0 = p1; q0 = q1; p1 = p2; q1 = q2;
        P = p1; Q = q1;
        if (std::fabs((double)P / (double)Q - lodr) <= tol) break;
        double fr = y - a;
        if (fr == 0.0) break;
        y = 1.0 / fr;
    }
    int64_t g = std::gcd(P, Q);
    if (g > 1) { P /= g; Q /= g; }
    if (!(Q <= kQmax && P >= 1)) return false;
    double ulp = std::nextafter(lodr, INFINITY) - lodr;
    if (!(std::fabs((double)P / (double)Q - lodr) <= 8.0 * ulp)) return false;
    *pnum = P;
    *q = Q;
    return true;
}

Plans build_plans(int nharms, int m, double hidr, double r_lo) {
    Plans pl;
    pl.nharms = nharms;
    pl.m = m;
    pl.hidr = hidr;
    pl.r_lo = r_lo;
    pl.lodr = hidr / nharms;
    int64_t pnum = 0, q = 0;
    pl.rational = trial_grid_rational(pl.lodr, &pnum, &q);
    pl.pnum = pnum;
    pl.q = q;
    const int m2 = m / 2;
    pl.harm.resize(nharms);
    for (int h = 1; h <= nharms; ++h) {
        HarmDev& d = pl.harm[h - 1];
        double hr = (double)h * r_lo;
        d.rfloor0 = (int64_t)std::floor(hr);
        d.dr0 = hr - (double)d.rfloor0;
        d.pad = 0;
        if (!pl.rational) {
            d.q = 1; d.s = 0; d.base_adv = 0; d.hp = 0; d.P = 0;
            d.row_off = 0; d.tab_off = 0;
            continue;
        }
        int64_t step = (int64_t)h * pnum;
        d.q = (int32_t)q;
        d.s = (int32_t)(((step % q) + q) % q);
        d.base_adv = (int32_t)floordiv(step, q);
        d.hp = (int32_t)step;
        d.P = (int32_t)(q / std::gcd(step, q));
        d.row_off = (int32_t)pl.row.size();
        d.tab_off = (int32_t)pl.carry.size();
        pl.row.resize(pl.row.size() + q, -1);
        int64_t res = 0;
        for (int32_t p = 0; p < d.P; ++p) {
            pl.row[d.row_off + res] = p;
            double u = d.dr0 + (double)res / (double)q;
            bool carry = u >= 1.0;
            double dr = carry ? u - 1.0 : u;
            pl.carry.push_back(carry ? 1 : 0);
            if (dr == 0.0) {
                // r is exactly a Fourier bin: the kernel is a delta on it.
                pl.A64.push_back(1.0);
                pl.A64.push_back(0.0);
                for (int i = 1; i <= m; ++i) pl.W64.push_back(i == m2 ? 1.0 : 0.0);
            } else {
                double s = sinpi_d(dr), c = cospi_d(dr);
                pl.A64.push_back((s * c) / M_PI);
                pl.A64.push_back(-((s * s) / M_PI));
                for (int i = 1; i <= m; ++i) pl.W64.push_back(1.0 / (dr - (double)(i - m2)));
            }
            res += d.s;
            if (res >= q) res -= q;
        }
    }
    pl.W32.resize(pl.W64.size());
    for (size_t i = 0; i < pl.W64.size(); ++i) pl.W32[i] = (float)pl.W64[i];
    pl.A32.resize(pl.A64.size());
    for (size_t i = 0; i < pl.A64.size(); ++i) pl.A32[i] = (float)pl.A64[i];
    return pl;
}

// boxcar_widths, search.jl:387-397
std::vector<int> boxcar_widths(int nbins, double fsp, double maxfrac) {
    int64_t wmax = (int64_t)std::floor(maxfrac * nbins);
    if (wmax < 1) wmax = 1;
    std::vector<int> ws;
    int64_t w = 1;
    while (w <= wmax && ws.size() < 64) {
        ws.push_back((int)w);
        int64_t nw = (int64_t)std::floor(fsp * (double)w);
        w = std::max(nw, w + 1);
    }
    return ws;
}

// ladder_boxcar_widths, search.jl:443-459
std::vector<int> ladder_boxcar_widths(int nbins, int k, const std::vector<int>& decims,
                                      double fsp, double maxfrac) {
    std::vector<int> ws = boxcar_widths(nbins, fsp, maxfrac);
    std::vector<int> ks = decims;
    std::sort(ks.begin(), ks.end());
    if (ks.size() <= 1) return ws;
    for (size_t i = 0; i + 1 < ks.size(); ++i)
        if (!(ks[i + 1] <= (kLadderWmax / 2) * ks[i])) return ws;
    int kmin = ks.front(), kmax = ks.back();
    std::vector<int> out;
    for (int w : ws)
        if ((w >= 2 || k == kmin) && (w <= kLadderWmax || k == kmax)) out.push_back(w);
    return out.empty() ? ws : out;
}

static uint16_t f2bf(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);      // round to nearest even
    return (uint16_t)(u >> 16);
}
static float bf2f(uint16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}
static void split_bf16(float x, uint16_t& hi, uint16_t& lo) {
    hi = f2bf(x);
    lo = f2bf(x - bf2f(hi));
}

MfmaTables build_mfma_tables(const Plans& pl, const std::vector<int>& fold_nbins,
                             const std::vector<int>& fold_wm, const std::vector<int>& fold_k) {
    MfmaTables t;
    if (!pl.rational) { t.why = "trial step is not a small rational"; return t; }
    const int m = pl.m, nh = pl.nharms;
    const int64_t q = pl.q;
    t.q = (int32_t)q;
    t.S = 16 / std::gcd((int64_t)16, q) * q;
    t.U = (int32_t)(t.S / 16);
    t.clsmap.assign((size_t)nh * q, -1);
    t.ks.resize(nh);
    int ncls = 0, ntile = 0;
    for (int h = 1; h <= nh; ++h) {
        const HarmDev& d = pl.harm[h - 1];
        // residues at which 16-trial groups (anchored at multiples of 16) start
        std::vector<int64_t> starts;
        int64_t g16 = (16 * (int64_t)d.s) % q, r = 0;
        do {
            starts.push_back(r);
            r = (r + g16) % q;
        } while (r != 0 && (int64_t)starts.size() < q);
        // offsets and phases of the 16 trials of each class
        int maxnj = 0;
        std::vector<std::vector<int>> P(starts.size()), D(starts.size());
        for (size_t c = 0; c < starts.size(); ++c) {
            int64_t res = starts[c], B = 0, B0 = 0;
            for (int i = 0; i < 16; ++i) {
                int p = pl.row[d.row_off + res];
                int64_t Bk = B + pl.carry[d.tab_off + p];
                if (i == 0) B0 = Bk;
                P[c].push_back(p);
                D[c].push_back((int)(Bk - B0));
                res += d.s;
                B += d.base_adv;
                if (res >= q) { res -= q; B += 1; }
            }
            maxnj = std::max(maxnj, m + D[c][15]);
        }
        int ks = (maxnj + 31) / 32;
        if (ks > 2) { t.why = "interpolation window wider than 64 bins"; return t; }
        t.ks[h - 1] = ks;
        for (size_t c = 0; c < starts.size(); ++c) {
            t.clsmap[(size_t)(h - 1) * q + starts[c]] = ncls;
            t.cls_tile.push_back(ntile);
            // lane groups whose second run (taps 16 + 4g .. 16 + 4g + 3 of the
            // first K step) reaches a nonzero weight: the trials read taps
            // 0 .. nj - 1
            const int nj = m + D[c][15];
            t.cls_run2.push_back((uint8_t)(ks > 1 ? 4 : std::max(0, std::min(4, (nj - 16 + 3) / 4))));
            // parity 0 (lanes holding Re) carries -Im A, parity 1 Im A
            for (int par = 0; par < 2; ++par)
                for (int i = 0; i < 16; ++i) {
                    const float ai = pl.A32[2 * (size_t)(d.tab_off + P[c][i]) + 1];
                    t.gA.push_back(pl.A32[2 * (size_t)(d.tab_off + P[c][i])]);
                    t.gA.push_back(par ? ai : -ai);
                }
            for (int kk = 0; kk < ks; ++kk) {
                for (int l = 0; l < 64; ++l) {
                    int i = l % 16;
                    for (int e = 0; e < 8; ++e) {
                        int k = kk * 32 + interp_tap(l / 16, e);
                        int tap = k - D[c][i];
                        float w = (tap >= 0 && tap < m)
                                      ? pl.W32[(size_t)(d.tab_off + P[c][i]) * m + tap] : 0.0f;
                        uint16_t hi, lo;
                        split_bf16(w, hi, lo);
                        t.wA_hi.push_back(hi);
                        t.wA_lo.push_back(lo);
                    }
                }
                ++ntile;
            }
            ++ncls;
        }
    }
    if (ncls > kDescClass) { t.why = "too many interpolation classes"; return t; }
    // residue of the group starting at trial 16u (S is a multiple of q, so this
    // is also the residue of every occurrence 16u + o*S)
    t.clsu.assign((size_t)nh * t.U * 2, -1);
    for (int h = 1; h <= nh; ++h) {
        const HarmDev& d = pl.harm[h - 1];
        for (int u = 0; u < t.U; ++u) {
            int64_t res = ((int64_t)16 * u % q) * (d.hp % q) % q;
            int c = t.clsmap[(size_t)(h - 1) * q + res];
            if (c < 0) { t.why = "group residue without a class"; return t; }
            t.clsu[2 * ((size_t)(h - 1) * t.U + u)] = t.cls_tile[c];
            t.clsu[2 * ((size_t)(h - 1) * t.U + u) + 1] = c;
        }
    }
    // per (harmonic, group) descriptors of the persistent kernel: first bin of
    // trial 16u (clamped: rows that far out fail the guard), bin advance per
    // occurrence (S/q periods of h*pnum), first weight tile, class | K steps << 28
    t.hdesc.assign((size_t)nh * t.U * 4, 0);
    for (int h = 1; h <= nh; ++h) {
        const HarmDev& d = pl.harm[h - 1];
        const int64_t hp = d.hp;
        for (int u = 0; u < t.U; ++u) {
            const int64_t T = (int64_t)16 * u * hp;
            const int64_t qint = T / q, res = T % q;
            const int p = pl.row[d.row_off + res];
            int64_t b0 = d.rfloor0 + qint + pl.carry[d.tab_off + p] + 1 - m / 2;
            b0 = std::max<int64_t>(-(1 << 30), std::min<int64_t>(b0, 1 << 30));
            const size_t e = 4 * ((size_t)(h - 1) * t.U + u);
            const int cls = t.clsu[2 * ((size_t)(h - 1) * t.U + u) + 1];
            t.hdesc[e + 0] = (int32_t)b0;
            t.hdesc[e + 1] = (int32_t)(hp * (t.S / q));
            t.hdesc[e + 2] = t.clsu[2 * ((size_t)(h - 1) * t.U + u)];
            t.hdesc[e + 3] = cls | ((int32_t)t.cls_run2[cls] << kDescRun2Shift) | (t.ks[h - 1] << 28);
        }
    }
    int ctile = 0;
    for (size_t f = 0; f < fold_nbins.size() && f < 8; ++f) {
        const int NB = fold_nbins[f], H = NB / 2;
        if (geo::c2r_direct(NB)) {
            // direct tiles: row i = 4g + v of row tile a computes bin
            // (g*RD + 4a + v) mod NB from K slots (Re h | Im h), lane group g's
            // operand e in slot geo::dperm of the fold factor's spacing
            const int sh = geo::dshift(f < fold_k.size() ? fold_k[f] : 1);
            const int wm = f < fold_wm.size() ? fold_wm[f] : 6;
            const int RD = (NB + 3) / 4, SH = 4 * RD - NB, ext = wm - 1;
            const int mt = std::max((RD + 3) / 4, (ext + SH + 3) / 4);
            const int kd = H <= 16 ? 1 : H <= 32 ? 2 : 4, hp = 16 * kd;
            t.c_base[f] = ctile;
            t.c_mt[f] = mt;
            t.c_ks[f] = kd;
            for (int a = 0; a < mt; ++a)
                for (int kk = 0; kk < kd; ++kk) {
                    for (int l = 0; l < 64; ++l) {
                        const int i = l % 16;
                        const int n = ((i / 4) * RD + 4 * a + i % 4) % NB;
                        for (int e = 0; e < 8; ++e) {
                            const geo::KSlot ks =
                                geo::dslot(kk * 32 + geo::dperm(l / 16, e, sh, kd), hp);
                            const int64_t ee = ((int64_t)ks.h * n) % NB;
                            const double ang = 2.0 * (double)ee / (double)NB;
                            double v = 0.0;
                            if (ks.im == 0 && ks.h <= H) v = (ks.h < H ? 2.0 : 1.0) * cospi_d(ang);
                            else if (ks.im == 1 && ks.h < H) v = -2.0 * sinpi_d(ang);
                            uint16_t hi, lo;
                            split_bf16((float)v, hi, lo);
                            t.cA_hi.push_back(hi);
                            t.cA_lo.push_back(lo);
                        }
                    }
                    ++ctile;
                }
            continue;
        }
        // even/odd DFT tiles; row i = 4g + v of row tile a computes profile
        // entry n = g*R + 4a + v (R = 4 * row tiles), so lane group g of a
        // trial column ends up holding the contiguous run n = g*R .. g*R + R - 1
        const int mt = (H + 1 + 15) / 16, ks = (H + 31) / 32, R = 4 * mt;
        t.c_base[f] = ctile;
        t.c_mt[f] = mt;
        t.c_ks[f] = ks;
        for (int a = 0; a < mt; ++a)
            for (int kk = 0; kk < ks; ++kk)
                for (int gsel = 0; gsel < 2; ++gsel) {
                    for (int l = 0; l < 64; ++l) {
                        const int i = l % 16;
                        const int n = (i / 4) * R + 4 * a + i % 4;
                        for (int e = 0; e < 8; ++e) {
                            int h = kk * 32 + (l / 16) * 8 + e + 1;
                            double v = 0.0;
                            if (n <= H && h <= H) {
                                int64_t ee = ((int64_t)h * n) % NB;
                                double ang = 2.0 * (double)ee / (double)NB;
                                double coef = h < H ? 2.0 : 1.0;
                                if (gsel == 0) v = coef * cospi_d(ang);
                                else v = h < H ? -coef * sinpi_d(ang) : 0.0;
                            }
                            uint16_t hi, lo;
                            split_bf16((float)v, hi, lo);
                            t.cA_hi.push_back(hi);
                            t.cA_lo.push_back(lo);
                        }
                    }
                    ++ctile;
                }
    }
    t.ok = true;
    return t;
}

}  // namespace cs
