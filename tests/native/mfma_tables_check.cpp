// Host check of the split-BF16 MFMA operand tables (coherentsearch.jl_amd/csrc/
// cs_plan.cpp build_mfma_tables) against quantities recomputed here from the
// interpolation plans alone:
//  1. window starts: for every (harmonic, 16-trial group) the first bin of any
//     occurrence st*8S + o*S + 16u equals the table's b0(16u) + adv*(8st + o),
//     where the direct value is floor(h r_t) + 1 - m/2 from the exact integer
//     state (t*hp) div/mod q and the carry table;
//  2. weight tiles: row i of a group's tile applied to a window starting at the
//     group's first bin reproduces trial i's own m-tap FP32 weights at its own
//     first bin (up to the bf16 hi+lo split), and conj(A) of the class is the
//     trial's A (Im negated in the copy for the lanes that hold Re);
//  3. c2r tiles: the permuted even/odd tiles (NB > 64) and the direct-layout
//     tiles (NB <= 64, every register of every lane group) reproduce the numpy
//     irfft convention (unnormalised) computed directly in FP64, and every
//     coefficient of a dead K slot is exactly zero (the kernel reads a clamped
//     row for those slots).
// Prints OK on success.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "cs_plan.hpp"

using namespace cs;

static float bf2f(uint16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

static int bad = 0;
#define CHECK(cond, ...)                      \
    do {                                      \
        if (!(cond)) {                        \
            if (bad < 10) printf(__VA_ARGS__); \
            ++bad;                            \
        }                                     \
    } while (0)

// exact first bin and phase of trial t for harmonic plan d
static int64_t first_bin(const Plans& pl, const HarmDev& d, int64_t t, int* phase) {
    const __int128 T = (__int128)t * d.hp;
    const int64_t qint = (int64_t)(T / d.q), res = (int64_t)(T % d.q);
    const int p = pl.row[d.row_off + res];
    *phase = p;
    return d.rfloor0 + qint + pl.carry[d.tab_off + p] + 1 - pl.m / 2;
}

static void check_config(int nharms, int m, double hidr, double r_lo, const std::vector<int>& nbins,
                         unsigned seed, const std::vector<int>& wms = {},
                         const std::vector<int>& fks = {}) {
    Plans pl = build_plans(nharms, m, hidr, r_lo);
    if (!pl.rational) {
        printf("config nharms=%d: not rational\n", nharms);
        ++bad;
        return;
    }
    std::vector<int> wm = wms;
    if (wm.empty()) wm.assign(nbins.size(), 6);
    std::vector<int> fk = fks;   // fold factors 1, 2, ... unless given
    for (size_t f = fk.size(); f < nbins.size(); ++f) fk.push_back((int)f + 1);
    MfmaTables t = build_mfma_tables(pl, nbins, wm, fk);
    if (!t.ok) {
        printf("tables failed: %s\n", t.why.c_str());
        ++bad;
        return;
    }
    srand(seed);
    int nrun[5] = {0, 0, 0, 0, 0};
    for (uint8_t v : t.cls_run2) ++nrun[v];
    printf("nharms %d m %d hidr %g: %zu classes; second-run lane groups 0..4: %d %d %d %d %d\n", nharms, m,
           hidr, t.cls_run2.size(), nrun[0], nrun[1], nrun[2], nrun[3], nrun[4]);
    // 1 + 2: descriptors and weight tiles
    std::vector<double> xr(4096);
    for (int h = 1; h <= nharms; ++h) {
        const HarmDev& d = pl.harm[h - 1];
        for (int u = 0; u < t.U; ++u) {
            const int32_t* e = &t.hdesc[4 * ((size_t)(h - 1) * t.U + u)];
            const int64_t b0u = e[0], adv = e[1];
            const int tile = e[2], cls = e[3] & kDescClass, ks = e[3] >> 28;
            const int run2 = (e[3] >> kDescRun2Shift) & 7;
            CHECK(run2 == t.cls_run2[cls] && run2 <= 4 && (ks == 1 || run2 == 4),
                  "h=%d u=%d: second-run lane groups %d (class %d, ks %d)\n", h, u, run2, cls, ks);
            CHECK(ks == t.ks[h - 1], "h=%d ks %d vs %d\n", h, ks, t.ks[h - 1]);
            for (int trial = 0; trial < 6; ++trial) {
                const int64_t st = rand() % 4000, o = rand() % 8;
                const int64_t tt = st * 8 * t.S + o * t.S + 16 * (int64_t)u;
                int p;
                const int64_t direct = first_bin(pl, d, tt, &p);
                CHECK(direct == b0u + adv * (8 * st + o),
                      "h=%d u=%d st=%lld o=%lld: first bin %lld vs %lld\n", h, u, (long long)st,
                      (long long)o, (long long)direct, (long long)(b0u + adv * (8 * st + o)));
            }
            // tile rows against each trial's own weights
            for (auto& v : xr) v = (double)(rand() % 20001 - 10000) / 10000.0;
            int p0;
            const int64_t g0 = first_bin(pl, d, 16 * (int64_t)u, &p0);
            for (int i = 0; i < 16; ++i) {
                int p;
                const int64_t bi = first_bin(pl, d, 16 * (int64_t)u + i, &p);
                const int64_t off = bi - g0;
                CHECK(off >= 0 && off + m <= 32 * ks, "h=%d u=%d i=%d offset %lld\n", h, u, i,
                      (long long)off);
                // every tap a trial reads lies in a run the kernel loads: the
                // first (taps 0..15) or a second-run lane group below run2
                CHECK(ks > 1 || off + m <= 16 + 4 * run2, "h=%d u=%d i=%d reads tap %lld, run2 %d\n", h,
                      u, i, (long long)(off + m - 1), run2);
                double want = 0.0, scale = 0.0;
                for (int tap = 0; tap < m; ++tap) {
                    const double w = pl.W32[(size_t)(d.tab_off + p) * m + tap];
                    want += w * xr[off + tap];
                    scale += std::fabs(w * xr[off + tap]);
                }
                double got = 0.0;
                for (int kk = 0; kk < ks; ++kk)
                    for (int kg = 0; kg < 4; ++kg)
                        for (int el = 0; el < 8; ++el) {
                            const size_t idx = (((size_t)(tile + kk) * 64) + (kg * 16 + i)) * 8 + el;
                            const double w = (double)bf2f(t.wA_hi[idx]) + (double)bf2f(t.wA_lo[idx]);
                            got += w * xr[kk * 32 + interp_tap(kg, el)];
                        }
                CHECK(std::fabs(got - want) <= 1e-5 * scale + 1e-12,
                      "h=%d u=%d i=%d tile row %g vs %g\n", h, u, i, got, want);
                for (int par = 0; par < 2; ++par) {
                    const float ar = t.gA[2 * (((size_t)cls * 2 + par) * 16 + i)];
                    const float ai = t.gA[2 * (((size_t)cls * 2 + par) * 16 + i) + 1];
                    const float want_ai = pl.A32[2 * (size_t)(d.tab_off + p) + 1];
                    CHECK(ar == pl.A32[2 * (size_t)(d.tab_off + p)] &&
                              ai == (par ? want_ai : -want_ai),
                          "h=%d u=%d i=%d parity %d: conj(A) mismatch\n", h, u, i, par);
                }
            }
        }
    }
    // 3: c2r tiles, with the row permutation of each lane group
    for (size_t f = 0; f < nbins.size(); ++f) {
        const int NB = nbins[f], H = NB / 2, mt = t.c_mt[f], ks = t.c_ks[f], R = 4 * mt;
        std::vector<double> Xr(H + 1), Xi(H + 1), want(NB), got(NB, 0.0);
        for (int h = 0; h <= H; ++h) {
            Xr[h] = (double)(rand() % 2001 - 1000) / 1000.0;
            Xi[h] = (double)(rand() % 2001 - 1000) / 1000.0;
        }
        Xr[0] = Xi[0] = 0.0;   // row 0 (DC) of the stack is zero
        for (int n = 0; n < NB; ++n) {   // numpy irfft * NB: DC/Nyquist imaginary parts ignored
            double s = Xr[0] + ((n % 2) ? -Xr[H] : Xr[H]);
            for (int h = 1; h < H; ++h) {
                const double a = 2.0 * M_PI * (double)((int64_t)h * n % NB) / NB;
                s += 2.0 * (Xr[h] * std::cos(a) - Xi[h] * std::sin(a));
            }
            want[n] = s;
        }
        double scale = 0.0;
        for (int h = 1; h <= H; ++h) scale += std::fabs(Xr[h]) + std::fabs(Xi[h]);
        if (NB <= 64) {
            // direct layout: lane group g's operand e of K chunk kk is the
            // stacked row the kernel's gather reads, min(h, H) of plane Re or Im:
            // one chunk (H <= 16): groups 0-1 Re, 2-3 Im, h = 1 + d(g&1) +
            // (e mod d) + 2d(e div d); two chunks: chunk 0 Re, 1 Im,
            // h = 1 + dg + (e mod d) + 4d(e div d); d = 8 for odd fold factors
            // k, else 8 / (largest power of two dividing k), at least 1.
            // Lane group g, register r = 4a + v holds bin (g*RD + r) mod NB.
            const int RD = (NB + 3) / 4, SH = 4 * RD - NB, ext = wm[f] - 1;
            const int mt = std::max((RD + 3) / 4, (ext + SH + 3) / 4);
            const int kd = H <= 16 ? 1 : H <= 32 ? 2 : 4, hp = 16 * kd;
            CHECK(t.c_mt[f] == mt && t.c_ks[f] == kd, "NB=%d direct tile counts %d %d\n", NB,
                  t.c_mt[f], t.c_ks[f]);
            (void)hp;
            const int k = fk[f];
            int d = 8;
            for (int kk2 = k; d > 1 && kk2 % 2 == 0; kk2 /= 2) d /= 2;
            // the two lane groups of a half-wave (0-1, 2-3) gather rows whose
            // 16-word runs at the 130-word stride fall in disjoint halves of
            // the 32 banks
            for (int el = 0; el < 8; ++el) {
                const int r0 = (1 + (el % d) + (kd == 1 ? 2 : 4) * d * (el / d)) * k - 1;
                const int r1 = r0 + d * k;
                CHECK(((r1 - r0) * 130) % 32 == 16, "NB=%d k=%d: rows %d, %d share banks\n", NB,
                      k, r0, r1);
            }
            std::vector<double> acc(16 * mt, 0.0);
            for (int a = 0; a < mt; ++a)
                for (int kk = 0; kk < kd; ++kk) {
                    const size_t tile = (size_t)t.c_base[f] + a * kd + kk;
                    for (int l = 0; l < 64; ++l)
                        for (int el = 0; el < 8; ++el) {
                            const int i = l % 16, g = l / 16;
                            const bool im = kd == 1 ? g >= 2 : kk == 1;
                            const int h = kd == 1
                                              ? 1 + d * (g & 1) + (el % d) + 2 * d * (el / d)
                                              : 1 + d * g + (el % d) + 4 * d * (el / d);
                            const size_t idx = (tile * 64 + l) * 8 + el;
                            const double c = (double)bf2f(t.cA_hi[idx]) + (double)bf2f(t.cA_lo[idx]);
                            const bool live = im ? h < H : h <= H;
                            if (!live) {
                                CHECK(t.cA_hi[idx] == 0 && t.cA_lo[idx] == 0,
                                      "NB=%d direct tile %zu lane %d el %d: nonzero dead slot\n",
                                      NB, tile, l, el);
                                continue;
                            }
                            const int hr = std::min(h, H);
                            acc[a * 16 + i] += c * (im ? Xi[hr] : Xr[hr]);
                        }
                }
            for (int g = 0; g < 4; ++g)
                for (int r = 0; r < 4 * mt; ++r) {
                    const int a = r / 4, v = r % 4, i = 4 * g + v;
                    const int n = (g * RD + r) % NB;
                    CHECK(std::fabs(acc[a * 16 + i] - want[n]) <= 2e-5 * 2.0 * scale,
                          "NB=%d direct g=%d r=%d bin %d: %g vs %g\n", NB, g, r, n,
                          acc[a * 16 + i], want[n]);
                }
            continue;
        }
        std::vector<double> E(16 * mt, 0.0), O(16 * mt, 0.0);
        for (int a = 0; a < mt; ++a)
            for (int kk = 0; kk < ks; ++kk)
                for (int sel = 0; sel < 2; ++sel) {
                    const size_t tile = (size_t)t.c_base[f] + (a * ks + kk) * 2 + sel;
                    for (int l = 0; l < 64; ++l)
                        for (int el = 0; el < 8; ++el) {
                            const int i = l % 16, hh = kk * 32 + (l / 16) * 8 + el + 1;
                            const size_t idx = (tile * 64 + l) * 8 + el;
                            if (hh > H) {
                                // the kernel reads row H for these slots: the
                                // coefficients must be exactly zero
                                CHECK(t.cA_hi[idx] == 0 && t.cA_lo[idx] == 0,
                                      "NB=%d tile %zu lane %d el %d: nonzero past H\n", NB, tile, l,
                                      el);
                                continue;
                            }
                            const double c = (double)bf2f(t.cA_hi[idx]) + (double)bf2f(t.cA_lo[idx]);
                            (sel ? O : E)[a * 16 + i] += c * (sel ? Xi[hh] : Xr[hh]);
                        }
                }
        // row i = 4g + v of tile a is entry n = g*R + 4a + v
        std::vector<int> seen(NB, 0);
        for (int a = 0; a < mt; ++a)
            for (int i = 0; i < 16; ++i) {
                const int n = (i / 4) * R + 4 * a + i % 4;
                if (n > H) continue;
                const double e = E[a * 16 + i], o = O[a * 16 + i];
                got[n] = e + o;
                seen[n]++;
                if (n > 0 && n < H) {
                    got[NB - n] = e - o;
                    seen[NB - n]++;
                }
            }
        for (int n = 0; n < NB; ++n) {
            CHECK(seen[n] == 1, "NB=%d bin %d produced %d times\n", NB, n, seen[n]);
            CHECK(std::fabs(got[n] - want[n]) <= 2e-5 * 2.0 * scale,
                  "NB=%d bin %d: %g vs %g\n", NB, n, got[n], want[n]);
        }
    }
}

int main() {
    // 2^25 points, dt 64 us: nharms 60, folds 120..20 (r_lo = 0.1 Hz * T)
    const double T2 = 33554432.0 * 64e-6;
    check_config(60, 16, 0.5, 0.1 * T2, {120, 60, 40, 30, 24, 20}, 11);
    // 2^20 points: nharms 8, one fold; and a 2^24-point nharms=32 search
    check_config(8, 16, 0.5, 1.0 * 1048576 * 64e-6, {16}, 12);
    check_config(32, 16, 0.5, 0.1 * 16777216 * 64e-6, {64, 32, 20, 16}, 13);
    // wider windows (two K steps): m = 32
    check_config(32, 32, 0.5, 4000.0, {64, 32, 20, 16}, 14);
    // wide banks (more direct row tiles): nharms 60, folds 120, 60, 40, 30
    check_config(60, 32, 0.5, 0.1 * T2, {120, 60, 40, 30}, 15, {28, 13, 9, 9});
    check_config(32, 16, 0.5, 0.1 * 16777216 * 64e-6, {64, 32, 16}, 16, {19, 9, 6}, {1, 2, 4});
    check_config(12, 16, 0.5, 1000.0, {24}, 17);
    // other trial-grid periods: hidr 0.25 (q = 240, S = 240), hidr 0.3
    // (q = 200, S = 400), hidr 1/3 at nharms 30 (q = 90, S = 720), and the
    // 128-bin even/odd fold (nharms 64)
    check_config(60, 16, 0.25, 523.0, {120, 60, 40, 30, 24, 20}, 18);
    check_config(60, 16, 0.3, 523.0, {120, 60, 40, 30, 24, 20}, 19);
    check_config(30, 16, 1.0 / 3.0, 777.0, {60, 30, 20}, 20);
    check_config(64, 16, 0.5, 911.0, {128, 64}, 21);
    // fold factor 8 (K slots one harmonic apart per lane group)
    check_config(64, 16, 0.5, 911.0, {128, 64, 16}, 22, {}, {1, 2, 8});
    printf(bad ? "FAIL %d\n" : "OK\n", bad);
    return bad ? 1 : 0;
}
