// Host check of the register geometry used by the split-BF16 bound
// (coherentsearch.jl_amd/csrc/cs_geo.hpp): for every supported profile length,
// rebuild each lane's XP/XM sequences (even/odd layout) or its direct-layout
// run and extension from emulated lane registers and compare every boxcar sum
// and per-width maximum with a direct circular boxcar over the profile.
// Integer-valued profiles make the sums exact.
#include <cstdio>
#include <cstdlib>
#include <vector>
#define CS_HD
#include "cs_geo.hpp"

using namespace cs::geo;

static const int kW[9] = {1, 2, 3, 4, 6, 9, 13, 19, 28};

template <int NB, int EXT>
static int check_nb(unsigned seed) {
    using G = Geo<NB, EXT>;
    const int nwid = EXT == 5 ? 5 : EXT == 8 ? 6 : EXT == 12 ? 7 : EXT == 18 ? 8 : 9;
    constexpr int R = G::R, H = G::H, X = G::X;
    static_assert(G::max_rounds() <= G::kRounds, "too many fetch rounds");
    if (seed == 0) printf("-");
    int bad = 0;
    srand(seed);
    for (int trial = 0; trial < 50; ++trial) {
        std::vector<long> x(NB);
        for (int b = 0; b < NB; ++b) x[b] = (rand() % 2001) - 1000;
        // emulated lane registers
        long P[4][R], M[4][R];
        for (int g = 0; g < 4; ++g)
            for (int i = 0; i < R; ++i) {
                int n = g * R + i;
                P[g][i] = n <= H ? x[n] : 7777777 + rand() % 100;
                M[g][i] = (n >= 1 && n <= H - 1) ? x[NB - n] : 5555555 + rand() % 100;
            }
        // direct maxima per width
        long best[9];
        for (int a = 0; a < nwid; ++a) {
            best[a] = -(1L << 60);
            if (kW[a] > NB) continue;
            for (int b = 0; b < NB; ++b) {
                long s = 0;
                for (int j = 0; j < kW[a]; ++j) s += x[(b + j) % NB];
                if (s > best[a]) best[a] = s;
            }
        }
        long got[9];
        for (int a = 0; a < nwid; ++a) got[a] = -(1L << 60);
        std::vector<int> starts(NB, 0);
        for (int g = 0; g < 4; ++g) {
            for (int w = 0; w < 2; ++w) {
                long S[X];
                for (int i = 0; i < X; ++i) {
                    Src s = G::src(w, g, i);
                    S[i] = s.kind ? M[s.lane][s.reg] : P[s.lane][s.reg];
                }
                // every position's window is a real boxcar (a position another
                // lane owns repeats that lane's window); ownership is exact
                for (int i = 0; i < R; ++i) {
                    int bin = (w == 0 ? g * R + i : NB - (g * R + R - 1 - i)) % NB;
                    if (G::start(w, g, i)) starts[bin] += 1;
                    for (int a = 0; a < nwid; ++a) {
                        if (kW[a] > NB) continue;
                        long s = 0, ref = 0;
                        for (int j = 0; j < kW[a]; ++j) {
                            s += S[i + j];
                            ref += x[(bin + j) % NB];
                        }
                        if (s != ref) {
                            if (bad < 5)
                                printf("NB=%d g=%d %s i=%d w=%d: %ld vs %ld\n", NB, g, w ? "M" : "P",
                                       i, kW[a], s, ref);
                            ++bad;
                        }
                        if (s > got[a]) got[a] = s;
                    }
                }
            }
        }
        for (int b = 0; b < NB; ++b)
            if (starts[b] != 1) {
                if (bad < 5) printf("NB=%d bin %d started %d times\n", NB, b, starts[b]);
                ++bad;
            }
        for (int a = 0; a < nwid; ++a)
            if (got[a] != best[a]) {
                if (bad < 5) printf("NB=%d w=%d max %ld vs %ld\n", NB, kW[a], got[a], best[a]);
                ++bad;
            }
    }
    return bad;
}

// Direct layout (DGeo): lane group g's MFMA registers hold bins
// (g*RD + r) mod NB (r < NR); entries r >= NR are fetched the way the kernel
// does it, every lane reading its next group, which offers register r - RD
// (group 0 offers r - RD + SH).  Every window the lanes form must equal the
// circular boxcar, every bin must start exactly one window, and the sample
// writes (registers r < RD of the starts a group owns, at slot g*RD + r) must
// cover every bin exactly once with its own value.
template <int NB, int WM>
static int check_direct(unsigned seed) {
    using D = DGeo<NB, WM>;
    constexpr int RD = D::RD, NR = D::NR, X = D::X, SH = D::SH;
    static_assert(D::EXT + SH <= NR && D::EXT <= NR, "extension source outside the MFMA registers");
    int nwid = 0;
    while (nwid < 9 && kW[nwid] <= WM) ++nwid;
    int bad = 0;
    srand(seed);
    for (int trial = 0; trial < 50; ++trial) {
        std::vector<long> x(NB);
        for (int b = 0; b < NB; ++b) x[b] = (rand() % 2001) - 1000;
        long reg[4][X];
        for (int g = 0; g < 4; ++g)
            for (int r = 0; r < NR; ++r) reg[g][r] = x[D::bin(g, r)];
        for (int g = 0; g < 4; ++g)
            for (int r = NR; r < X; ++r) {
                const int sg = (g + 1) % 4;                         // the lane read
                const int sr = sg == 0 ? r - RD + SH : r - RD;      // what it offers
                if (sr < 0 || sr >= NR) {
                    printf("NB=%d WM=%d: offer register %d out of range\n", NB, WM, sr);
                    return bad + 1;
                }
                if (sr != D::next_reg(g, r)) ++bad;
                reg[g][r] = reg[sg][sr];
            }
        std::vector<int> starts(NB, 0), slots(NB, 0);
        for (int g = 0; g < 4; ++g)
            for (int r = 0; r < RD; ++r) {
                // every position's window (group 3's last SH repeat group 0's
                // first); starts and sample slots only where the group owns it
                const int bin = (g * RD + r) % NB;
                if (D::owns(g, r)) {
                    starts[bin] += 1;
                    slots[bin] += 1;
                    if (reg[g][r] != x[bin]) ++bad;   // sample slot value
                }
                for (int a = 0; a < nwid; ++a) {
                    long s = 0, ref = 0;
                    for (int j = 0; j < kW[a]; ++j) {
                        s += reg[g][r + j];
                        ref += x[(bin + j) % NB];
                    }
                    if (s != ref) {
                        if (bad < 5)
                            printf("direct NB=%d WM=%d g=%d r=%d w=%d: %ld vs %ld\n", NB, WM, g, r,
                                   kW[a], s, ref);
                        ++bad;
                    }
                }
            }
        for (int b = 0; b < NB; ++b)
            if (starts[b] != 1 || slots[b] != 1) {
                if (bad < 5) printf("direct NB=%d bin %d started %d times\n", NB, b, starts[b]);
                ++bad;
            }
    }
    return bad;
}

template <int NB, int EXT>
static void rounds() {
    using G = Geo<NB, EXT>;
    printf("NB=%3d EXT=%2d R=%2d max_rounds=%d total_rounds=%d\n", NB, EXT, G::R, G::max_rounds(),
           G::total_rounds());
}

// every profile length of the fused MFMA mode with the narrow bank (widths <= 6)
// and with its full bank (widest rung <= 0.3 * NB)
#define CS_GEO_CASES(X) X(16, 5) X(20, 5) X(24, 5) X(30, 5) X(32, 5) X(40, 5) X(60, 5) X(64, 5) \
    X(120, 5) X(128, 5) X(30, 8) X(32, 8) X(40, 8) X(60, 12) X(64, 18) X(120, 27) X(128, 27)

int main() {
    int bad = 0;
    unsigned seed = 1;
#define X(NB, EXT) rounds<NB, EXT>();
    CS_GEO_CASES(X)
#undef X
#define X(NB, EXT) bad += check_nb<NB, EXT>(seed++);
    CS_GEO_CASES(X)
#undef X
    // direct layout: every length up to 64, narrow and full banks
#define X(NB, WM) bad += check_direct<NB, WM>(seed++);
    X(16, 6) X(20, 6) X(24, 6) X(30, 6) X(32, 6) X(40, 6) X(60, 6) X(64, 6)
    X(16, 4) X(30, 9) X(32, 9) X(40, 9) X(60, 13) X(64, 19)
#undef X
    printf(bad ? "FAIL %d\n" : "OK\n", bad);
    return bad ? 1 : 0;
}
