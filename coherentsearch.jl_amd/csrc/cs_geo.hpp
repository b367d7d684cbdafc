// Register geometry of one folded profile after the even/odd c2r MFMAs.
//
// A profile of NB = 2H bins is produced as E[n] (cosine half) and O[n] (sine
// half) for n = 0..H; bin n is P[n] = E + O and bin NB - n is M[n] = E - O
// (0 < n < H).  The c2r tile rows are permuted so that the four lanes g = 0..3
// of one trial column hold the contiguous run n = g*R .. g*R + R - 1
// (R = 4 * row tiles), i.e. registers P[i], M[i] with n = g*R + i.
//
// In bin order the profile is the circular chain
//     P[0], P[1], ..., P[H], M[H-1], M[H-2], ..., M[1], (wrap) P[0], ...
// Every boxcar start belongs to exactly one lane: starts at bins n <= H to the
// lane holding P[n] ("P run"), starts at bins NB - n (1 <= n < H) to the lane
// holding M[n] ("M run").  Each lane lays its two runs out as sequences in
// ascending bin order,
//     XP[i] = bin g*R + i                    (i < R + EXT)
//     XM[i] = bin NB - g*R - R + 1 + i       (i < R + EXT)
// so a window of width w <= EXT + 1 starting at XP[i] or XM[i] (i < R) covers
// consecutive sequence entries.  The entries past the lane's own run (and the
// ones that cross the P/M junction at bin H or the wrap at bin NB) come from
// other lanes of the same column; this header gives, at compile time, where
// every entry lives.  Every entry is a real bin, so the window at every
// position i < R is a real boxcar: a position whose start another lane owns
// repeats that lane's window (same values, same sums), which leaves each
// width's maximum unchanged, so the bound evaluates every position unmasked.
//
// Profile lengths up to 64 use the direct layout instead (DGeo below): one
// real c2r with Re and Im stacked along K, bins in ascending order per lane.
#pragma once

#ifndef CS_HD
#if defined(__HIPCC__) || defined(__HIP__)
#define CS_HD __host__ __device__
#else
#define CS_HD
#endif
#endif

namespace cs {

// Boxcar width ladder of the fused kernels: bank entry a has width bank_w(a)
// (1, 2, 3, 4, 6, 9, 13, ...); bank_count(lim) entries have width <= lim.
CS_HD constexpr int bank_w(int a) {
    return a == 0 ? 1 : a == 1 ? 2 : a == 2 ? 3 : a == 3 ? 4 : a == 4 ? 6 : a == 5 ? 9
         : a == 6 ? 13 : a == 7 ? 19 : a == 8 ? 28 : a == 9 ? 42 : a == 10 ? 63 : 94;
}
CS_HD constexpr int bank_count(int lim) {
    int n = 0;
    for (int a = 0; a < 12; ++a) if (bank_w(a) <= lim) n = a + 1;
    return n;
}
// widest bank width the kernels support for a profile of nb bins (0.3 nb)
CS_HD constexpr int gate_wmax(int nb) {
    return bank_w(bank_count((3 * nb) / 10 < 1 ? 1 : (3 * nb) / 10) - 1);
}
// widest width a fold pass evaluates: 6 for banks within widths 1..6, else
// the profile length's widest supported width
CS_HD constexpr int fold_wm(int nb, unsigned bankmask) {
    return (bankmask >> 5) == 0 || gate_wmax(nb) <= 6 ? 6 : gate_wmax(nb);
}

namespace geo {

// ---------------------------------------------------------------------------
// Direct layout (profile lengths NB <= 64).  The c2r tile row i = 4g + v of
// row tile a computes bin (g*RD + 4a + v) mod NB directly, from Re X(1..H)
// in K slots 0..HP-1 and Im X(1..H-1) in K slots HP..2HP-1, so lane group g
// of a trial column holds registers X[r] = bin (g*RD + r) mod NB, r < 4*MT:
// its RD window starts followed by the first bins of the next group's run.
// A window of width w <= WM at start r < RD reads X[r .. r + w - 1]; entries
// r >= 4*MT come from the next lane group, register r - RD (group 3 reads
// group 0's register r - RD + SH, SH = 4*RD - NB, because of the wrap).
// Group 3 owns only RD - SH starts (the sample writes); its windows at the
// other SH positions start at bins 0 .. SH-1 again, repeats of group 0's.
// ---------------------------------------------------------------------------
CS_HD constexpr bool c2r_direct(int nb) { return nb <= 64; }

template <int NB, int WM>
struct DGeo {
    static constexpr int H = NB / 2;
    static constexpr int RD = (NB + 3) / 4;            // starts per lane group
    static constexpr int SH = 4 * RD - NB;             // 0 unless NB % 4 != 0
    static constexpr int EXT = WM - 1;                 // entries past a start
    static constexpr int MT = (RD + 3) / 4 > (EXT + SH + 3) / 4 ? (RD + 3) / 4
                                                               : (EXT + SH + 3) / 4;
    static constexpr int NR = 4 * MT;                  // registers from the MFMAs
    static constexpr int X = RD + EXT;                 // sequence length
    static constexpr int NF = X > NR ? X - NR : 0;     // entries fetched from the next group
    static constexpr int KD = H <= 16 ? 1 : H <= 32 ? 2 : 4;   // K steps
    static constexpr int HP = 16 * KD;                 // K slots per (Re | Im) half
    // bin held by register r of lane group g
    static constexpr int bin(int g, int r) { return (g * RD + r) % NB; }
    // does lane group g own the window start at its register r (r < RD)?
    static constexpr bool owns(int g, int r) { return r < RD && g * RD + r < NB; }
    // register of the next lane group that holds sequence entry r >= NR of group g
    static constexpr int next_reg(int g, int r) { return g == 3 ? r - RD + SH : r - RD; }
};

// c2r coefficient of K slot kidx for bin n of a direct-layout tile (numpy
// irfft convention, unnormalised; DC excluded, Nyquist imaginary part ignored):
// returns {re/im selector 0/1, harmonic h}; the value is
//   re: (h < H ? 2 : 1) cos(2 pi h n / NB) for h <= H, else 0
//   im: -2 sin(2 pi h n / NB) for h < H, else 0
struct KSlot {
    int im;
    int h;
};
CS_HD constexpr KSlot dslot(int kidx, int hp) {
    return kidx < hp ? KSlot{0, kidx + 1} : KSlot{1, kidx - hp + 1};
}

// K-slot order of a direct-layout tile for fold factor k.  Operand e of lane
// group g takes slot dperm(g, e, s, kd) of its 32-slot chunk: runs of
// d = 2^s consecutive harmonics, the lane groups that read one plane
// starting d harmonics apart.  kd = 1: groups 0-1 read Re (slots 0-15),
// groups 2-3 Im (16-31); kd = 2: all four read the chunk's plane.  With
// d*k = 8 (mod 16) the two lane groups of a half-wave gather stacked rows d*k
// apart, 16 (mod 32) banks at a row stride of 2 (mod 32) words, so their
// 16-word runs never share a bank.
CS_HD constexpr int dshift(int k) { return (k & 1) ? 3 : (k & 2) ? 2 : (k & 4) ? 1 : 0; }
CS_HD constexpr int dperm(int g, int e, int s, int kd) {
    return kd == 1 ? 16 * (g >> 1) + ((g & 1) << s) + (e & ((1 << s) - 1)) + ((e >> s) << (s + 1))
                   : (g << s) + (e & ((1 << s) - 1)) + ((e >> s) << (s + 2));
}

struct Src {
    int lane;   // lane group 0..3 of the column that holds the value
    int kind;   // 0 = P register, 1 = M register
    int reg;    // register index 0..R-1
};

// EXT: entries needed past the run, the widest boxcar minus one
template <int NB, int EXT = 5>
struct Geo {
    static constexpr int H = NB / 2;
    static constexpr int MT = (H + 1 + 15) / 16;   // c2r row tiles (n = 0..H)
    static constexpr int R = 4 * MT;               // run length per lane
    static constexpr int KS = (H + 31) / 32;       // c2r K steps (harmonics 1..H)
    static constexpr int X = R + EXT;              // sequence length

    // where bin c (taken modulo NB) lives
    static constexpr Src chain(int c) {
        c %= NB;
        if (c < 0) c += NB;
        if (c <= H) return Src{c / R, 0, c % R};
        const int n = NB - c;
        return Src{n / R, 1, n % R};
    }

    // ---- P run: starts at bins g*R + i <= H ----
    static constexpr bool p_start(int g, int i) { return i < R && g * R + i <= H; }
    static constexpr bool p_need(int g, int i) { return i < X; }
    static constexpr Src p_src(int g, int i) {
        return p_need(g, i) ? chain(g * R + i) : Src{g, 0, i < R ? i : 0};
    }

    // ---- M run: entry i < R is n = g*R + R - 1 - i (bin NB - n), starts 1 <= n <= H - 1 ----
    static constexpr bool m_start(int g, int i) {
        const int n = g * R + R - 1 - i;
        return i < R && n >= 1 && n <= H - 1;
    }
    static constexpr int m_first(int g) {
        for (int i = 0; i < R; ++i)
            if (m_start(g, i)) return i;
        return X;
    }
    static constexpr int m_last(int g) {
        int l = -1;
        for (int i = 0; i < R; ++i)
            if (m_start(g, i)) l = i;
        return l;
    }
    static constexpr bool m_need(int g, int i) { return i < X; }
    static constexpr Src m_src(int g, int i) {
        return m_need(g, i) ? chain(NB - g * R - R + 1 + i) : Src{g, 1, i < R ? i : 0};
    }

    static constexpr bool start(int which, int g, int i) {
        return which == 0 ? p_start(g, i) : m_start(g, i);
    }
    static constexpr Src src(int which, int g, int i) {
        return which == 0 ? p_src(g, i) : m_src(g, i);
    }

    // Remote fetch plan of entry i: groups that need a value held by another
    // group fetch it with one cross-lane read per round; within a round every
    // providing group offers one register (its kind/reg), so requests that ask
    // one provider for different registers go to different rounds.
    static constexpr int kRounds = 8;
    struct Prov {
        static constexpr int kMaxRound = 8;
        int nround;                  // rounds used (0: every group reads its own registers)
        int round[4];                // requester group -> round, -1 if local
        bool has[kRounds][4];        // [round][provider]
        int kind[kRounds][4], reg[kRounds][4];
    };
    static constexpr Prov prov(int which, int i) {
        Prov p{0, {-1, -1, -1, -1}, {}, {}, {}};
        for (int g = 0; g < 4; ++g) {
            const Src s = src(which, g, i);
            if (s.lane == g) continue;
            int r = 0;
            for (; r < kRounds; ++r) {
                if (!p.has[r][s.lane]) break;
                if (p.kind[r][s.lane] == s.kind && p.reg[r][s.lane] == s.reg) break;
            }
            p.round[g] = r;
            p.has[r][s.lane] = true;
            p.kind[r][s.lane] = s.kind;
            p.reg[r][s.lane] = s.reg;
            if (r + 1 > p.nround) p.nround = r + 1;
        }
        return p;
    }
    static constexpr int max_rounds() {
        int m = 0;
        for (int w = 0; w < 2; ++w)
            for (int i = 0; i < X; ++i)
                if (prov(w, i).nround > m) m = prov(w, i).nround;
        return m;
    }
    static constexpr int total_rounds() {
        int m = 0;
        for (int w = 0; w < 2; ++w)
            for (int i = 0; i < X; ++i) m += prov(w, i).nround;
        return m;
    }
};

}  // namespace geo
}  // namespace cs
