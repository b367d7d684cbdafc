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
// every entry lives.  Entries no valid window reads are left as the lane's
// own (finite) registers.
#pragma once

#ifndef CS_HD
#define CS_HD __host__ __device__
#endif

namespace cs {
namespace geo {

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
    static constexpr bool p_need(int g, int i) { return g * R <= H && i <= H - g * R + EXT; }
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
    static constexpr bool m_need(int g, int i) {
        return m_last(g) >= 0 && i >= m_first(g) && i <= m_last(g) + EXT;
    }
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
