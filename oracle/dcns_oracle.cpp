// dcns_oracle.cpp -- CPU restatement of dCNS's alignment hot path.
//
// TEST INFRASTRUCTURE ONLY (the checker).  Only tests/, the smoke entry point
// and bench.py's CPU-baseline leg load this library; the product never does.
// See dcns_oracle.h for what each function restates and how it is pinned.
#include "dcns_oracle.h"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <map>
#include <set>
#include <thread>
#include <vector>

namespace {

const int32_t kOutOfBand = -536870912;  // sequenceAlignment3.cpp:34

// Scorei, model/Score.cpp:105-125: N (code 2) scores 0 against anything.
struct Scoring {
    int32_t match, mismatch;
    int32_t operator()(int8_t a, int8_t b) const {
        if (a == 2 || b == 2) return 0;
        return a == b ? match : mismatch;
    }
};

// --------------------------------------------------------------------------
// K1: local alignment, max score and first (ref-major) position.
// --------------------------------------------------------------------------

// Dispatch condition of SmithWaterman, sequenceAlignment3.cpp:224-225 (int8)
// and :338-339 (int16); both SIMD paths share one recurrence.
bool simd_bound(int64_t refLen, int64_t readLen, int64_t match, int64_t open, int64_t ext, int64_t bound) {
    return ((-refLen * ext - open) < bound || (-readLen * ext - open) < bound) &&
           (refLen * match < bound || readLen * match < bound);
}

// The striped int8/int16 recurrence (sequenceAlignment3.cpp:267-332), as a
// plain row-major DP.  Row i walks the reference, column j the read.
//   Hpre[i][j] = max(H[i-1][j-1] + s, 0, E[i][j])
//   E[i+1][j]  = max(E[i][j] + ext, Hpre[i][j] + open, 0)   (vertical gap, from Hpre)
//   F[i][j]    = max(H[i][j-1] + open, max(F[i][j-1],0) + ext)   (along the read)
//   H[i][j]    = max(Hpre[i][j], F[i][j])
// The first strict maximum in (i, j) order is reported.
void sw_variant(const int8_t *ref, const int8_t *read, int32_t L1, int32_t L2, Scoring sc, int32_t open,
                int32_t ext, int32_t &best, int32_t &bi, int32_t &bj) {
    std::vector<int32_t> H(L2 > 0 ? L2 : 1, 0), E(L2 > 0 ? L2 : 1, 0);
    best = 0; bi = -1; bj = -1;
    for (int32_t i = 0; i < L1; ++i) {
        int32_t diag = 0, left = 0, fc = 0;
        for (int32_t j = 0; j < L2; ++j) {
            int32_t hpre = std::max(std::max(diag + sc(ref[i], read[j]), 0), E[j]);
            diag = H[j];
            E[j] = std::max(std::max(E[j] + ext, hpre + open), 0);
            int32_t f = std::max(left + open, fc + ext);
            int32_t h = std::max(hpre, f);
            fc = std::max(f, 0);
            H[j] = h;
            left = h;
            if (h > best) { best = h; bi = i; bj = j; }
        }
    }
}

// Textbook Gotoh local alignment, SmithWaterman_non_avx (positions=true),
// sequenceAlignment3.cpp:143-216: both gap kinds open from the final H.
void sw_textbook(const int8_t *ref, const int8_t *read, int32_t L1, int32_t L2, Scoring sc, int32_t open,
                 int32_t ext, int32_t &best, int32_t &bi, int32_t &bj) {
    std::vector<int32_t> prev(L2 + 1, 0), cur(L2 + 1, 0), vert(L2 + 1, 0);
    best = 0; bi = -1; bj = -1;
    for (int32_t i = 1; i <= L1; ++i) {
        int32_t horiz = 0;
        for (int32_t j = 1; j <= L2; ++j) {
            horiz = std::max(open + cur[j - 1], ext + horiz);
            vert[j] = std::max(open + prev[j], ext + vert[j]);
            int32_t m = std::max(std::max(sc(ref[i - 1], read[j - 1]) + prev[j - 1], horiz), vert[j]);
            cur[j] = std::max(m, 0);
            vert[j] = std::max(vert[j], 0);
            horiz = std::max(horiz, 0);
            if (m > best) { best = m; bi = i - 1; bj = j - 1; }
        }
        std::swap(prev, cur);
    }
}

void sw_dispatch(const int8_t *ref, const int8_t *read, int32_t L1, int32_t L2, Scoring sc, int32_t open,
                 int32_t ext, int32_t &best, int32_t &bi, int32_t &bj) {
    if (simd_bound(L1, L2, sc.match, open, ext, 32767))
        sw_variant(ref, read, L1, L2, sc, open, ext, best, bi, bj);
    else
        sw_textbook(ref, read, L1, L2, sc, open, ext, best, bi, bj);
}

// --------------------------------------------------------------------------
// K2 / K3: banded x-drop extension anchored at (0,0).
// --------------------------------------------------------------------------

// Traceback store with the semantics of model/Matrix.cpp:24-36 after
// reset(): every cell reads 255 unless written in the current pass.  Kept
// band-sparse (one row segment per DP row) so long inputs fit in memory.
struct BandTrace {
    std::vector<int32_t> lo;
    std::vector<std::vector<uint8_t>> cells;
    void reset(int32_t rows) {
        lo.assign((size_t)rows, 0);
        cells.assign((size_t)rows, std::vector<uint8_t>());
    }
    void open_row(int32_t i, int32_t start) { lo[(size_t)i] = start; cells[(size_t)i].clear(); }
    void set(int32_t i, int32_t j, uint8_t v) {
        std::vector<uint8_t> &r = cells[(size_t)i];
        size_t k = (size_t)(j - lo[(size_t)i]);
        if (r.size() <= k) r.resize(k + 1, 255);
        r[k] = v;
    }
    uint8_t get(int32_t i, int32_t j) const {
        if (i < 0 || (size_t)i >= cells.size()) return 255;
        const std::vector<uint8_t> &r = cells[(size_t)i];
        int64_t k = (int64_t)j - lo[(size_t)i];
        if (k < 0 || (size_t)k >= r.size()) return 255;
        return r[(size_t)k];
    }
};

struct XResult {
    int32_t max, end1, end2;
    int64_t cells;
};

// One pass of SemiGlobal_xextend at band half-width w.  Returns false when
// the reference would restart with 2w (sequenceAlignment3.cpp:1284-1289,
// :1381-1385).  With tb != nullptr the traceback byte of every band cell is
// recorded as in :1358-1379.
bool xextend_pass(const int8_t *s1, const int8_t *s2, int32_t L1, int32_t L2, Scoring sc, int32_t open, int32_t ext,
                  int32_t xdrop, int32_t w, BandTrace *tb, XResult &r) {
    std::vector<int32_t> M1(L2 + 1, kOutOfBand), M2(L2 + 1, kOutOfBand), F(L2 + 1, kOutOfBand);
    M1[0] = 0; M2[0] = 0;
    r.max = 0; r.end1 = 0; r.end2 = 0;
    if (tb) tb->reset(L1 + 1);
    for (int32_t i = 1; i <= L1; ++i) {
        int32_t rowMax = 0, rowJ = -1;
        int32_t start = std::max(i - w, 1);
        int32_t end = std::min(i + w, L2);
        int32_t e = kOutOfBand;
        F[end] = kOutOfBand;
        if (tb) tb->open_row(i, start);
        for (int32_t j = start; j <= end; ++j) {
            ++r.cells;
            int32_t m = sc(s1[i - 1], s2[j - 1]) + M1[j - 1];
            uint8_t d = 0;
            if (!(m > F[j])) { m = F[j]; d = 1; }
            if (!(m > e)) { m = e; d = 2; }
            M2[j] = m;
            if (m > rowMax) { rowMax = m; rowJ = j; }
            int32_t h = m + open;
            int32_t f = F[j] + ext;
            if (f >= h) d |= 8; else f = h;
            F[j] = f;
            e += ext;
            if (e >= h) d |= 16; else e = h;
            if (tb) tb->set(i, j, d);
        }
        if (rowJ != L2 && rowJ != 1 && (rowJ == start || rowJ == end)) return false;
        if (rowMax > r.max) { r.max = rowMax; r.end1 = i; r.end2 = rowJ; }
        if (r.max - rowMax > xdrop) break;
        if (rowJ <= 0) break;
        std::swap(M1, M2);
        M1[0] = kOutOfBand;
        M2[0] = kOutOfBand;
    }
    return true;
}

XResult xextend(const int8_t *s1, const int8_t *s2, int32_t L1, int32_t L2, Scoring sc, int32_t open, int32_t ext,
                int32_t xdrop, int32_t w, BandTrace *tb) {
    XResult r{0, 0, 0, 0};
    int64_t cells = 0;
    while (true) {
        r.cells = 0;
        bool done = xextend_pass(s1, s2, L1, L2, sc, open, ext, xdrop, w, tb, r);
        cells += r.cells;
        if (done) break;
        w *= 2;
    }
    r.cells = cells;
    r.end1 -= 1;
    r.end2 -= 1;
    return r;
}

// Traceback state machine of sequenceAlignment3.cpp:1416-1438 (also used by
// the two-piece DP, :864-887).  Ops: 0 = M, 1 = I (--j), 2 = D (--i).
void push_op(std::vector<uint32_t> &cig, uint32_t op) {
    if (cig.empty() || (cig.back() & 0xf) != op) cig.push_back((1u << 4) | op);
    else cig.back() += 1u << 4;
}

void trace(const BandTrace &T, int32_t &ii, int32_t &jj, std::vector<uint32_t> &cig) {
    int state = 0;
    while (ii > 0 && jj > 0) {
        int t = T.get(ii, jj);
        if (state == 0) state = t & 7;
        else if (!((t >> (state + 2)) & 1)) state = 0;
        if (state == 0) state = t & 7;
        if (state == 0) { push_op(cig, 0); --ii; --jj; }
        else if (state == 1 || state == 3) { push_op(cig, 2); --ii; }
        else { push_op(cig, 1); --jj; }
    }
}

// --------------------------------------------------------------------------
// K4: two-piece affine banded semi-global DP (SemiGlobal_no_avx).
// --------------------------------------------------------------------------
struct SgResult {
    int32_t max, end1, end2;
};

SgResult semiglobal(const int8_t *s1, const int8_t *s2, int32_t L1, int32_t L2, Scoring sc, int32_t o1, int32_t x1,
                    int32_t o2, int32_t x2, int32_t zdrop, int32_t w, BandTrace *tb, std::vector<uint32_t> *cig,
                    int64_t *cells = nullptr) {
    // an empty query reads seq2[-1] in the reference (undefined); both empty
    // cases are taken as "no alignment"
    if (L1 <= 0 || L2 <= 0) return SgResult{0, -2, -2};
    std::vector<int32_t> M1(L2 + 1, kOutOfBand), M2(L2 + 1, kOutOfBand), F1(L2 + 1, kOutOfBand),
        F2(L2 + 1, kOutOfBand);
    M1[0] = 0; M2[0] = 0;
    int32_t best = 0, bi = -1, bj = -1;
    if (tb) tb->reset(L1 + 1);
    for (int32_t i = 1; i <= L1; ++i) {
        int32_t rowMax = 0, rowJ = 0;
        int32_t start = std::max(i - w, 1);
        int32_t end = std::min(i + w, L2);
        int32_t e1 = kOutOfBand, e2 = kOutOfBand;
        F1[end] = kOutOfBand;
        F2[end] = kOutOfBand;
        if (tb) tb->open_row(i, start);
        if (cells) *cells += (int64_t)std::max(0, end - start) + 1;
        for (int32_t j = start; j < end; ++j) {
            int32_t m = sc(s1[i - 1], s2[j - 1]) + M1[j - 1];
            uint8_t d = 0;
            if (!(m > F1[j])) { m = F1[j]; d = 1; }
            if (!(m > e1)) { m = e1; d = 2; }
            if (!(m > F2[j])) { m = F2[j]; d = 3; }
            if (!(m > e2)) { m = e2; d = 4; }
            M2[j] = m;
            if (m > rowMax) { rowMax = m; rowJ = j; }
            int32_t h = m + o1, f = F1[j] + x1;
            if (f >= h) d |= 8; else f = h;
            F1[j] = f;
            e1 += x1;
            if (e1 >= h) d |= 16; else e1 = h;
            int32_t h2 = m + o2, f2 = F2[j] + x2;
            if (f2 >= h2) d |= 32; else f2 = h2;
            F2[j] = f2;
            e2 += x2;
            if (e2 >= h2) d |= 64; else e2 = h2;
            if (tb) tb->set(i, j, d);
        }
        {   // last column of the row: max only, no gap-state update, no trace byte (:766-787)
            int32_t j = end;
            int32_t m = sc(s1[i - 1], s2[j - 1]) + M1[j - 1];
            m = std::max(std::max(std::max(std::max(m, F1[j]), e1), F2[j]), e2);
            M2[j] = m;
            if (m > rowMax) { rowMax = m; rowJ = j; }
        }
        if (rowMax > best) { best = rowMax; bi = i; bj = rowJ; }
        int32_t l = (i - bi) - (rowJ - bj);
        if (l < 0) l = -l;
        if (best - rowMax > zdrop + l * x2) break;
        if (rowJ <= 0) break;
        std::swap(M1, M2);
        M1[0] = kOutOfBand;
        M2[0] = kOutOfBand;
    }
    if (tb && cig) {
        int32_t ii = bi, jj = bj;
        trace(*tb, ii, jj, *cig);
        while (ii > 0) { push_op(*cig, 2); --ii; }
        while (jj > 0) { push_op(*cig, 1); --jj; }
        std::reverse(cig->begin(), cig->end());
    }
    return SgResult{best, bi - 1, bj - 1};
}

// --------------------------------------------------------------------------
// Weighted semi-global DP (ZDP): SemiGlobal(..., weights, Score&, ...),
// sequenceAlignment3.cpp:1483-1706 (two-piece), and
// SemiGlobal_single_gap_penalty, :1710-1890 (one piece).  Row i takes its
// substitution scores, horizontal gap penalties and z-drop from the score
// profile of weights[i-1] and its vertical gap penalties from weights[i];
// weights[-1 .. L1] are read.
// --------------------------------------------------------------------------
struct Profile {
    int32_t match, mismatch, o1, x1, o2, x2, zdrop;
};

struct Profiles {
    std::map<int16_t, Profile> by_cat;
    const Profile &operator[](int16_t c) const { return by_cat.at(c); }
    // Score::getM: 5x5, N row and column 0 (model/Score.cpp:50-64)
    static int32_t sub(const Profile &p, int8_t a, int8_t b) {
        if (a == 2 || b == 2) return 0;
        return a == b ? p.match : p.mismatch;
    }
};

SgResult semiglobal_weighted(const int8_t *s1, const int8_t *s2, int32_t L1, int32_t L2, const int16_t *weights,
                             const Profiles &P, int32_t w, bool two_piece, BandTrace *tb, std::vector<uint32_t> *cig,
                             int64_t *cells = nullptr) {
    if (L1 <= 0 || L2 <= 0) return SgResult{0, -2, -2};
    std::vector<int32_t> M1(L2 + 1, kOutOfBand), M2(L2 + 1, kOutOfBand), F1(L2 + 1, kOutOfBand),
        F2(L2 + 1, kOutOfBand);
    M1[0] = 0; M2[0] = 0;
    int32_t best = 0, bi = -1, bj = -1;
    if (tb) tb->reset(L1 + 1);
    for (int32_t i = 1; i <= L1; ++i) {
        int32_t rowMax = 0, rowJ = two_piece ? kOutOfBand : -1;
        const int32_t start = std::max(i - w, 1), end = std::min(i + w, L2);
        int32_t e1 = kOutOfBand, e2 = kOutOfBand;
        F1[end] = kOutOfBand;
        F2[end] = kOutOfBand;
        const Profile &cur = P[weights[i - 1]], &nxt = P[weights[i]];
        if (tb) tb->open_row(i, start);
        if (cells) *cells += (int64_t)std::max(0, end - start + 1);
        for (int32_t j = start; j <= end; ++j) {
            int32_t m = Profiles::sub(cur, s1[i - 1], s2[j - 1]) + M1[j - 1];
            uint8_t d = 0;
            if (!(m > F1[j])) { m = F1[j]; d = 1; }
            if (!(m > e1)) { m = e1; d = 2; }
            if (two_piece) {
                if (!(m > F2[j])) { m = F2[j]; d = 3; }
                if (!(m > e2)) { m = e2; d = 4; }
            }
            M2[j] = m;
            if (m > rowMax) { rowMax = m; rowJ = j; }
            int32_t h = m + nxt.o1, f = F1[j] + nxt.x1;  // vertical: the next row's profile
            if (f >= h) d |= 8; else f = h;
            F1[j] = f;
            h = m + cur.o1;  // horizontal: this row's profile
            e1 += cur.x1;
            if (e1 >= h) d |= 16; else e1 = h;
            if (two_piece) {
                int32_t h2 = m + nxt.o2, f2 = F2[j] + nxt.x2;
                if (f2 >= h2) d |= 32; else f2 = h2;
                F2[j] = f2;
                h2 = m + cur.o2;
                e2 += cur.x2;
                if (e2 >= h2) d |= 64; else e2 = h2;
            }
            if (tb) tb->set(i, j, d);
        }
        if (rowMax > best) { best = rowMax; bi = i; bj = rowJ; }
        if (two_piece) {
            int32_t l = (i - bi) - (rowJ - bj);
            if (l < 0) l = -l;
            if (best - rowMax > cur.zdrop + l * cur.x2) break;
        } else if (best - rowMax > cur.zdrop) {
            break;
        }
        if (rowJ <= 0) break;
        std::swap(M1, M2);
        M1[0] = kOutOfBand;
        M2[0] = kOutOfBand;
    }
    if (tb && cig) {
        int32_t ii = bi, jj = bj;
        trace(*tb, ii, jj, *cig);
        while (ii > 0) { push_op(*cig, 2); --ii; }
        while (jj > 0) { push_op(*cig, 1); --jj; }
        std::reverse(cig->begin(), cig->end());
    }
    return SgResult{best, bi - 1, bj - 1};
}

// --------------------------------------------------------------------------
// Seeds (model/Seed.h:13-43) and the getSeeds recursion.
// --------------------------------------------------------------------------
struct Seed {
    int32_t start1, end1, start2, end2;
    // Not a strict weak order; std::set behaviour under it is part of the
    // reference semantics (Seed.h:29-42).
    bool operator<(const Seed &c) const {
        if (start1 == c.start1 && start2 == c.start2) return false;
        if (end1 == c.end1 && end2 == c.end2) return false;
        if (start1 < c.start1) return true;
        if (start1 == c.start1) return start2 < c.start2;
        return false;
    }
};

struct SeedCtx {
    const int8_t *seq1, *seq2;
    std::vector<int8_t> rc1, rc2;  // reverse complements (4 - c reversed)
    int32_t L1, L2, W, minScore, minRef;
    Scoring sc;
    int32_t open, ext;
    int64_t cells = 0;
};

// One getSeeds body for the ref sub-range [a, a+len) and the window at sp:
// forward SW, then the reverse SW on reverse complements
// (findSimilarFragmentsForPairedSequence4.cpp:142-166).  Coordinates are
// absolute.  Returns false when the window max is below minScore.
bool window_seed(SeedCtx &c, int32_t a, int32_t len, int32_t sp, Seed &s) {
    int32_t thisW = (sp + c.W > c.L2) ? (c.L2 - sp) : c.W;
    int32_t mx, e1, e2;
    c.cells += (int64_t)len * thisW;
    sw_dispatch(c.seq1 + a, c.seq2 + sp, len, thisW, c.sc, c.open, c.ext, mx, e1, e2);
    if (!(c.minScore <= mx)) return false;
    int32_t r1 = std::min(1 + e1, thisW + c.W);
    int32_t r2 = std::min(1 + e2, thisW + c.W);
    // reverse complement of seq1[a+e1-r1+1 .. a+e1] and seq2[sp+e2-r2+1 .. sp+e2]
    const int8_t *rref = c.rc1.data() + (c.L1 - 1 - (a + e1));
    const int8_t *rread = c.rc2.data() + (c.L2 - 1 - (sp + e2));
    int32_t rmx, q1 = -1, q2 = -1;
    c.cells += (int64_t)r1 * r2;
    sw_dispatch(rref, rread, r1, r2, c.sc, c.open, c.ext, rmx, q1, q2);
    s.end1 = a + e1;
    s.end2 = sp + e2;
    s.start1 = a + e1 - q1;
    s.start2 = sp + e2 - q2;
    return true;
}

// Inner (ungated) getSeeds, findSimilarFragmentsForPairedSequence4.cpp:61-128.
std::set<Seed> inner_seeds(SeedCtx &c, int32_t a, int32_t len, int32_t sp) {
    std::set<Seed> out;
    Seed s;
    if (!window_seed(c, a, len, sp, s)) return out;
    out.insert(s);
    int32_t leftLen = s.start1 - a;
    if (leftLen >= c.minRef)
        for (const Seed &x : inner_seeds(c, a, leftLen, sp)) out.insert(x);
    int32_t rightLen = len - (s.end1 - a) - 1;
    if (rightLen >= c.minRef)
        for (const Seed &x : inner_seeds(c, s.end1 + 1, rightLen, sp)) out.insert(x);
    return out;
}

// Outer getSeeds with the set gate, :131-201.
void outer_seeds(SeedCtx &c, int32_t sp, std::set<Seed> &seeds) {
    Seed s;
    if (!window_seed(c, 0, c.L1, sp, s)) return;
    if (seeds.find(s) != seeds.end()) return;
    seeds.insert(s);
    if (s.start1 >= c.minRef)
        for (const Seed &x : inner_seeds(c, 0, s.start1, sp)) seeds.insert(x);
    if (c.L1 - s.end1 - 1 >= c.minRef)
        for (const Seed &x : inner_seeds(c, s.end1 + 1, c.L1 - s.end1 - 1, sp)) seeds.insert(x);
}

// Window schedule of overload 1, :461-480.
void window_loop(SeedCtx &c, int32_t step, std::set<Seed> &seeds) {
    int32_t sp = 0;
    while (true) {
        outer_seeds(c, sp, seeds);
        sp += step;
        if (sp > c.L2) break;
        if (sp + c.W > c.L2) {
            outer_seeds(c, sp, seeds);
            break;
        }
    }
}

// link_seeds, :288-344.  Sorting is std::sort on start1 (not stable), so the
// result depends on the library's sort; the product uses the same toolchain.
std::vector<Seed> link_seed_list(const std::set<Seed> &seeds) {
    std::vector<Seed> cur(seeds.begin(), seeds.end());
    std::sort(cur.begin(), cur.end(), [](Seed x, Seed y) { return x.start1 < y.start1; });
    bool changed = true;
    while (changed) {
        changed = false;
        std::vector<Seed> next;
        std::vector<char> used(cur.size(), 0);
        for (size_t i = 0; i < cur.size(); ++i) {
            if (used[i]) continue;
            Seed s1 = cur[i];
            bool merged = false;
            for (size_t j = i + 1; j < cur.size(); ++j) {
                if (used[j]) continue;
                Seed s2 = cur[j];
                if (s1.start2 <= s2.start2 && s1.end1 >= s2.end1 && s1.end2 >= s2.end2) {
                    used[j] = 1;
                } else if (s1.start2 <= s2.start2 && s1.end1 <= s2.end1 && s1.end2 <= s2.end2) {
                    int32_t d = s2.start1 - s1.start1;
                    if (d == s2.start2 - s1.start2 && d == s2.end1 - s1.end1 && d == s2.end2 - s1.end2) {
                        next.push_back(Seed{s1.start1, s2.end1, s1.start2, s2.end2});
                        used[i] = 1;
                        used[j] = 1;
                        merged = true;
                        break;
                    }
                } else if (s2.start1 > s1.end1 || s2.start2 > s1.end2) {
                    break;
                }
            }
            if (merged) changed = true;
            else next.push_back(s1);
        }
        cur.swap(next);
    }
    return cur;
}

void link_seeds(std::set<Seed> &seeds) {
    const std::vector<Seed> cur = link_seed_list(seeds);
    seeds.clear();
    for (const Seed &s : cur) seeds.insert(s);
}

struct Fragment {
    uint32_t start1, end1, start2, end2, score;
    std::vector<uint32_t> cigar;
    double p, e;
};

}  // namespace

extern "C" {

// link_seeds on a given list: inserted into the set in the order given,
// linked, written in set order (insertion_order 0) or as the merged list in
// the order link_seeds inserts it back (1); returns the count (or -1 past cap).
int32_t orc_link_seeds(const int32_t *seeds, int32_t n, int32_t insertion_order, int32_t *out, int32_t cap) {
    std::set<Seed> set;
    for (int32_t k = 0; k < n; ++k)
        set.insert(Seed{seeds[4 * k], seeds[4 * k + 1], seeds[4 * k + 2], seeds[4 * k + 3]});
    std::vector<Seed> list = link_seed_list(set);
    if (!insertion_order) {
        set.clear();
        for (const Seed &s : list) set.insert(s);
        list.assign(set.begin(), set.end());
    }
    if ((int64_t)list.size() > cap) return -1;
    for (size_t o = 0; o < list.size(); ++o) {
        out[4 * o] = list[o].start1;
        out[4 * o + 1] = list[o].end1;
        out[4 * o + 2] = list[o].start2;
        out[4 * o + 3] = list[o].end2;
    }
    return (int32_t)list.size();
}

int orc_sw_uses_simd_variant(int32_t refLen, int32_t readLen, int32_t match, int32_t open, int32_t ext) {
    return simd_bound(refLen, readLen, match, open, ext, 32767) ? 1 : 0;
}

int64_t orc_sw_cells(int32_t refLen, int32_t readLen) { return (int64_t)refLen * readLen; }

int orc_sw(const int8_t *ref, const int8_t *read, int32_t refLen, int32_t readLen, int32_t match, int32_t mismatch,
           int32_t open, int32_t ext, int32_t *out3) {
    int32_t b, i, j;
    sw_dispatch(ref, read, refLen, readLen, Scoring{match, mismatch}, open, ext, b, i, j);
    out3[0] = b; out3[1] = i; out3[2] = j;
    return 0;
}

int orc_sw_textbook(const int8_t *ref, const int8_t *read, int32_t refLen, int32_t readLen, int32_t match,
                    int32_t mismatch, int32_t open, int32_t ext, int32_t *out3) {
    int32_t b, i, j;
    sw_textbook(ref, read, refLen, readLen, Scoring{match, mismatch}, open, ext, b, i, j);
    out3[0] = b; out3[1] = i; out3[2] = j;
    return 0;
}

int orc_xextend(const int8_t *s1, const int8_t *s2, int32_t L1, int32_t L2, int32_t match, int32_t mismatch,
                int32_t open, int32_t ext, int32_t xdrop, int32_t w, int32_t *out3) {
    XResult r = xextend(s1, s2, L1, L2, Scoring{match, mismatch}, open, ext, xdrop, w, nullptr);
    out3[0] = r.max; out3[1] = r.end1; out3[2] = r.end2;
    return 0;
}

int orc_xextend_tb(const int8_t *s1, const int8_t *s2, int32_t L1, int32_t L2, int32_t match, int32_t mismatch,
                   int32_t open, int32_t ext, int32_t xdrop, int32_t w, int32_t *out5, uint32_t *cigar,
                   int32_t cap) {
    BandTrace T;
    XResult r = xextend(s1, s2, L1, L2, Scoring{match, mismatch}, open, ext, xdrop, w, &T);
    int32_t ii = r.end1 + 1, jj = r.end2 + 1;
    std::vector<uint32_t> cig;
    trace(T, ii, jj, cig);
    out5[0] = r.max; out5[1] = r.end1; out5[2] = r.end2; out5[3] = ii; out5[4] = jj;
    if ((int32_t)cig.size() > cap) return -1;
    std::copy(cig.begin(), cig.end(), cigar);
    return (int32_t)cig.size();
}

int orc_semiglobal(const int8_t *s1, const int8_t *s2, int32_t L1, int32_t L2, int32_t match, int32_t mismatch,
                   int32_t open1, int32_t ext1, int32_t open2, int32_t ext2, int32_t zdrop, int32_t w,
                   int32_t return_cigar, int32_t *out3, uint32_t *cigar, int32_t cap) {
    BandTrace T;
    std::vector<uint32_t> cig;
    SgResult r = semiglobal(s1, s2, L1, L2, Scoring{match, mismatch}, open1, ext1, open2, ext2, zdrop, w,
                            return_cigar ? &T : nullptr, return_cigar ? &cig : nullptr);
    out3[0] = r.max; out3[1] = r.end1; out3[2] = r.end2;
    if ((int32_t)cig.size() > cap) return -1;
    std::copy(cig.begin(), cig.end(), cigar);
    return (int32_t)cig.size();
}

int orc_seed_set(const int8_t *seq1, int32_t L1, const int8_t *seq2, int32_t L2, int32_t match, int32_t mismatch,
                 int32_t open, int32_t ext, int32_t window, int32_t step, int32_t min_score, int32_t linked,
                 int32_t *out, int32_t cap) {
    SeedCtx c;
    c.seq1 = seq1; c.seq2 = seq2; c.L1 = L1; c.L2 = L2; c.W = window; c.minScore = min_score;
    c.minRef = min_score / match; c.sc = Scoring{match, mismatch}; c.open = open; c.ext = ext;
    c.rc1.resize((size_t)L1 + 1); c.rc2.resize((size_t)L2 + 1);
    for (int32_t i = 0; i < L1; ++i) c.rc1[i] = (int8_t)(4 - seq1[L1 - 1 - i]);
    for (int32_t i = 0; i < L2; ++i) c.rc2[i] = (int8_t)(4 - seq2[L2 - 1 - i]);
    std::set<Seed> seeds;
    window_loop(c, step, seeds);
    if (linked) link_seeds(seeds);
    int32_t n = 0;
    for (const Seed &s : seeds) {
        if (n < cap) {
            out[4 * n] = s.start1; out[4 * n + 1] = s.end1; out[4 * n + 2] = s.start2; out[4 * n + 3] = s.end2;
        }
        ++n;
    }
    return n <= cap ? n : -n;
}

int orc_find_similar(const int8_t *seq1, int32_t L1, const int8_t *seq2, int32_t L2, int32_t match,
                     int32_t mismatch, int32_t open, int32_t ext, int32_t window, int32_t step, int32_t min_score,
                     int32_t w, int32_t xdrop, double pvalues, double lambda, double kvalue, int32_t *ints,
                     double *dbls, int32_t cap, uint32_t *cig, int64_t *cig_off, int64_t cig_cap, int64_t *stats) {
    SeedCtx c;
    c.seq1 = seq1; c.seq2 = seq2; c.L1 = L1; c.L2 = L2; c.W = window; c.minScore = min_score;
    c.minRef = min_score / match; c.sc = Scoring{match, mismatch}; c.open = open; c.ext = ext;
    c.rc1.resize((size_t)L1 + 1); c.rc2.resize((size_t)L2 + 1);
    for (int32_t i = 0; i < L1; ++i) c.rc1[i] = (int8_t)(4 - seq1[L1 - 1 - i]);
    for (int32_t i = 0; i < L2; ++i) c.rc2[i] = (int8_t)(4 - seq2[L2 - 1 - i]);
    std::set<Seed> seeds;
    window_loop(c, step, seeds);
    int64_t nseeds = (int64_t)seeds.size();
    link_seeds(seeds);

    const double L1d = (double)L1, L2d = (double)L2;
    int64_t xcells = 0;
    std::vector<Fragment> frags;
    for (const Seed &sd : seeds) {
        // x_extend_seed with traceback, :245-285, entered at seed start + 1 (:492-495)
        int32_t start1 = sd.start1 + 1, end1 = sd.end1 + 1, start2 = sd.start2 + 1, end2 = sd.end2 + 1;
        int32_t s1 = start1, s2 = start2;
        double p, e = 0;
        std::vector<uint32_t> cigar;
        XResult f = xextend(seq1 + start1, seq2 + start2, L1 - start1, L2 - start2, c.sc, open, ext, xdrop, w,
                            nullptr);
        xcells += f.cells;
        int32_t score = f.max;
        if (f.end1 < 0 || f.end2 < 0) {
            p = 100;
        } else {
            BandTrace T;
            int32_t rl1 = start1 + 1 + f.end1, rl2 = start2 + 1 + f.end2;
            XResult r = xextend(c.rc1.data() + (L1 - 1 - (start1 + f.end1)), c.rc2.data() + (L2 - 1 - (start2 + f.end2)),
                                rl1, rl2, c.sc, open, ext, xdrop, w, &T);
            xcells += r.cells;
            score = r.max;
            int32_t ii = r.end1 + 1, jj = r.end2 + 1;
            trace(T, ii, jj, cigar);
            if (r.end1 < 0 || r.end2 < 0) {
                p = 100;
            } else {
                end1 = s1 + f.end1;
                end2 = s2 + f.end2;
                start1 = s1 + f.end1 - r.end1;
                start2 = s2 + f.end2 - r.end2;
                end1 -= ii;
                end2 -= jj;
                e = kvalue * L1d * L2d * std::exp(-1.0 * lambda * (double)score);
                p = 1.0 - std::exp(-1.0 * e);
            }
        }
        if (!(p < pvalues)) continue;
        // Fragment dedup, :500-521: stored coordinates are 1-based uint32;
        // the comparisons mix 0-based int32 and 1-based uint32 on purpose.
        Fragment nf{(uint32_t)(start1 + 1), (uint32_t)(end1 + 1), (uint32_t)(start2 + 1), (uint32_t)(end2 + 1),
                    (uint32_t)score, cigar, p, e};
        bool fresh = true;
        for (Fragment &g : frags) {
            if (g.start1 == (uint32_t)(start1 + 1) && g.start2 == (uint32_t)(start2 + 1)) {
                fresh = false;
                if ((uint32_t)end1 >= g.end1 && (uint32_t)end2 >= g.end2) g = nf;
                break;
            } else if (g.end1 == (uint32_t)(end1 + 1) && g.end2 == (uint32_t)(end2 + 1)) {
                fresh = false;
                if ((uint32_t)start1 <= g.start1 && (uint32_t)start2 <= g.start2) g = nf;
                break;
            }
        }
        if (fresh) frags.push_back(nf);
    }
    if (stats) { stats[0] = c.cells; stats[1] = xcells; stats[2] = nseeds; stats[3] = (int64_t)seeds.size(); }
    int32_t n = (int32_t)frags.size();
    int64_t total = 0;
    for (auto &g : frags) total += (int64_t)g.cigar.size();
    if (n > cap || total > cig_cap) return -n - 1;
    int64_t o = 0;
    for (int32_t i = 0; i < n; ++i) {
        ints[5 * i] = (int32_t)frags[i].start1; ints[5 * i + 1] = (int32_t)frags[i].end1;
        ints[5 * i + 2] = (int32_t)frags[i].start2; ints[5 * i + 3] = (int32_t)frags[i].end2;
        ints[5 * i + 4] = (int32_t)frags[i].score;
        dbls[2 * i] = frags[i].p; dbls[2 * i + 1] = frags[i].e;
        cig_off[i] = o;
        for (uint32_t x : frags[i].cigar) cig[o++] = x;
    }
    cig_off[n] = o;
    return n;
}

namespace {

// getAllExtendSeed (findSimilarFragmentsForPairedSequence4.cpp:348-429): the
// window loop, link_seeds, the score-only x_extend_seed of every seed, the
// p-value filter and the x-seed merge.
std::vector<Seed> extend_seeds(SeedCtx &c, int32_t open, int32_t ext, int32_t step, int32_t w, int32_t xdrop,
                               double pvalues, double lambda, double kvalue, int64_t &nseeds, int64_t &nlinked,
                               int64_t &xcells) {
    const int8_t *seq1 = c.seq1, *seq2 = c.seq2;
    const int32_t L1 = c.L1, L2 = c.L2;
    // window loop (:377-392): no sp > L2 guard; a window of non-positive
    // width scores 0 (the SIMD SmithWaterman with an empty profile) and adds
    // no seed
    std::set<Seed> seeds;
    {
        int32_t sp = 0;
        while (true) {
            if (std::min(c.W, c.L2 - sp) > 0) outer_seeds(c, sp, seeds);
            sp += step;
            if (sp + c.W > c.L2) {
                if (c.L2 - sp > 0) outer_seeds(c, sp, seeds);
                break;
            }
        }
    }
    nseeds = (int64_t)seeds.size();
    link_seeds(seeds);
    nlinked = (int64_t)seeds.size();
    const double L1d = (double)L1, L2d = (double)L2;
    // score-only x_extend_seed (:205-241), p filter and x-seed merge (:400-427)
    std::vector<Seed> xs;
    for (const Seed &sd : seeds) {
        int32_t s1 = sd.start1 + 1, s2 = sd.start2 + 1;
        XResult f = xextend(seq1 + s1, seq2 + s2, L1 - s1, L2 - s2, c.sc, open, ext, xdrop, w, nullptr);
        xcells += f.cells;
        if (f.end1 < 0 || f.end2 < 0) continue;  // p = 100
        XResult r = xextend(c.rc1.data() + (L1 - 1 - (s1 + f.end1)), c.rc2.data() + (L2 - 1 - (s2 + f.end2)),
                            s1 + 1 + f.end1, s2 + 1 + f.end2, c.sc, open, ext, xdrop, w, nullptr);
        xcells += r.cells;
        if (r.end1 < 0 || r.end2 < 0) continue;
        const int32_t end1 = s1 + f.end1, end2 = s2 + f.end2;
        const int32_t start1 = s1 + f.end1 - r.end1, start2 = s2 + f.end2 - r.end2;
        const double e = kvalue * L1d * L2d * std::exp(-1.0 * lambda * (double)r.max);
        const double p = 1.0 - std::exp(-1.0 * e);
        if (!(p < pvalues)) continue;
        bool never_used = true;
        for (Seed &x : xs) {
            if (x.start1 == start1 && x.start2 == start2) {
                never_used = false;
                if (end1 > x.end1 && end2 > x.end2) { x.end1 = end1; x.end2 = end2; }
            } else if (x.end1 == end1 && x.end2 == end2) {
                if (start1 < x.start1 && start2 < x.start2) { x.start1 = start1; x.start2 = start2; }
                never_used = false;
            }
        }
        if (never_used) xs.push_back(Seed{start1, end1, start2, end2});
    }
    return xs;
}

void init_ctx(SeedCtx &c, const int8_t *seq1, int32_t L1, const int8_t *seq2, int32_t L2, int32_t match,
              int32_t mismatch, int32_t open, int32_t ext, int32_t window, int32_t min_score) {
    c.seq1 = seq1; c.seq2 = seq2; c.L1 = L1; c.L2 = L2; c.W = window; c.minScore = min_score;
    c.minRef = min_score / match; c.sc = Scoring{match, mismatch}; c.open = open; c.ext = ext;
    // one N past each reverse complement: the base before the sequence
    c.rc1.assign((size_t)L1 + 1, 2); c.rc2.assign((size_t)L2 + 1, 2);
    for (int32_t i = 0; i < L1; ++i) c.rc1[i] = (int8_t)(4 - seq1[L1 - 1 - i]);
    for (int32_t i = 0; i < L2; ++i) c.rc2[i] = (int8_t)(4 - seq2[L2 - 1 - i]);
}

// Overload 2's fragment replace / append (:584-611), p-values without a filter.
void add_fragment(std::vector<Fragment> &frags, int32_t start1, int32_t start2, int32_t end1, int32_t end2,
                  int32_t score, const std::vector<uint32_t> &cigar, double L1d, double L2d, double lambda,
                  double kvalue) {
    int32_t to_remove = -1;
    bool found = false;
    for (size_t i = 0; i < frags.size(); ++i) {
        if (frags[i].end1 == (uint32_t)(end1 + 1) && frags[i].end2 == (uint32_t)(end2 + 1)) {
            found = true;
            if ((uint32_t)(start1 + 1) < frags[i].start1 && (uint32_t)(start2 + 1) < frags[i].start2)
                to_remove = (int32_t)i;
        }
    }
    if (found && to_remove < 0) return;
    const double e = kvalue * L1d * L2d * std::exp(-1.0 * lambda * (double)score);
    const double p = 1.0 - std::exp(-1.0 * e);
    Fragment nf{(uint32_t)(start1 + 1), (uint32_t)(end1 + 1), (uint32_t)(start2 + 1), (uint32_t)(end2 + 1),
                (uint32_t)score, cigar, p, e};
    if (to_remove >= 0) frags[(size_t)to_remove] = nf;
    else frags.push_back(nf);
}

int emit(const std::vector<Fragment> &frags, int32_t *ints, double *dbls, int32_t cap, uint32_t *cig,
         int64_t *cig_off, int64_t cig_cap) {
    int32_t n = (int32_t)frags.size();
    int64_t total = 0;
    for (auto &g : frags) total += (int64_t)g.cigar.size();
    if (n > cap || total > cig_cap) return -n - 1;
    int64_t o = 0;
    for (int32_t i = 0; i < n; ++i) {
        ints[5 * i] = (int32_t)frags[i].start1; ints[5 * i + 1] = (int32_t)frags[i].end1;
        ints[5 * i + 2] = (int32_t)frags[i].start2; ints[5 * i + 3] = (int32_t)frags[i].end2;
        ints[5 * i + 4] = (int32_t)frags[i].score;
        dbls[2 * i] = frags[i].p; dbls[2 * i + 1] = frags[i].e;
        cig_off[i] = o;
        for (uint32_t x : frags[i].cigar) cig[o++] = x;
    }
    cig_off[n] = o;
    return n;
}

}  // namespace

int orc_find_similar_2gaps(const int8_t *seq1, int32_t L1, const int8_t *seq2, int32_t L2, int32_t match,
                           int32_t mismatch, int32_t open, int32_t ext, int32_t open2, int32_t ext2,
                           int32_t window, int32_t step, int32_t min_score, int32_t w, int32_t xdrop, int32_t zdrop,
                           int32_t bandwidth, double pvalues, double lambda, double kvalue, int32_t *ints,
                           double *dbls, int32_t cap, uint32_t *cig, int64_t *cig_off, int64_t cig_cap,
                           int64_t *stats) {
    SeedCtx c;
    init_ctx(c, seq1, L1, seq2, L2, match, mismatch, open, ext, window, min_score);
    int64_t nseeds = 0, nlinked = 0, xcells = 0, sgcells = 0;
    std::vector<Seed> xs = extend_seeds(c, open, ext, step, w, xdrop, pvalues, lambda, kvalue, nseeds, nlinked, xcells);
    const double L1d = (double)L1, L2d = (double)L2;
    // overload 2 (:532-616)
    std::vector<Fragment> frags;
    std::set<std::pair<int32_t, int32_t>> starts;
    for (const Seed &x : xs) {
        SgResult rv = semiglobal(c.rc1.data() + (L1 - 1 - x.end1), c.rc2.data() + (L2 - x.end2 - 1), 1 + x.end1,
                                 x.end2 + 1, c.sc, open, ext, open2, ext2, zdrop, bandwidth, nullptr, nullptr,
                                 &sgcells);
        const int32_t start1 = x.end1 - rv.end1, start2 = x.end2 - rv.end2;
        if (!starts.insert({start1, start2}).second) continue;
        std::vector<uint32_t> cigar;
        BandTrace T;
        SgResult fw = semiglobal(seq1 + start1, seq2 + start2, L1 - start1, L2 - start2, c.sc, open, ext, open2, ext2,
                                 zdrop, bandwidth, &T, &cigar, &sgcells);
        if (!(cigar.size() > 0 && fw.end1 > 0 && fw.end2 > 0)) continue;
        add_fragment(frags, start1, start2, fw.end1 + start1, fw.end2 + start2, fw.max, cigar, L1d, L2d, lambda,
                     kvalue);
    }
    if (stats) {
        stats[0] = c.cells; stats[1] = xcells; stats[2] = nseeds; stats[3] = nlinked;
        stats[4] = sgcells;
    }
    return emit(frags, ints, dbls, cap, cig, cig_off, cig_cap);
}

int orc_find_similar_seeded(const int8_t *seq1, int32_t L1, const int8_t *seq2, int32_t L2, const int32_t *seeds,
                            int32_t nseeds, int32_t match, int32_t mismatch, int32_t open, int32_t ext,
                            int32_t open2, int32_t ext2, int32_t zdrop, int32_t bandwidth, int32_t max_len,
                            double lambda, double kvalue, int32_t *ints, double *dbls, int32_t cap, uint32_t *cig,
                            int64_t *cig_off, int64_t cig_cap, int64_t *sgcells) {
    // findSimilarFragmentsForPairedSequence overload 3
    // (findSimilarFragmentsForPairedSequence4.cpp:623-745)
    SeedCtx c;
    init_ctx(c, seq1, L1, seq2, L2, match, mismatch, open, ext, 38, 40);
    const Scoring sc{match, mismatch};
    const double L1d = (double)L1, L2d = (double)L2;
    std::vector<Fragment> frags;
    std::set<std::pair<int32_t, int32_t>> starts;
    int64_t cl = 0;
    for (int32_t k = 0; k < nseeds; ++k) {
        const int32_t s1 = seeds[4 * k], e1s = seeds[4 * k + 1], s2 = seeds[4 * k + 2], e2s = seeds[4 * k + 3];
        bool use = true;
        for (const Fragment &f : frags)  // 1-based uint32 against 0-based seeds
            if (f.start1 <= (uint32_t)s1 && f.end1 >= (uint32_t)e1s && f.start2 <= (uint32_t)s2 &&
                f.end2 >= (uint32_t)e2s) {
                use = false;
                break;
            }
        if (!use) continue;
        int32_t e1 = e1s, e2 = e2s;
        if (e1 < 0 || e1 >= L1 || e2 < 0 || e2 >= L2) continue;  // outside the sequences (undefined in the reference)
        const int32_t r1 = std::min(1 + e1, max_len), r2 = std::min(1 + e2, max_len);
        if (seq1[e1] != seq2[e2]) {  // the reference's patch (:661-665)
            --e1;
            --e2;
        }
        // after the patch the reverse DP keeps r1 x r2 and starts one base
        // earlier: its last row / column is the base before the sequence (N)
        const SgResult rv = semiglobal(c.rc1.data() + (L1 - 1 - e1), c.rc2.data() + (L2 - 1 - e2), r1, r2, sc, open,
                                       ext, open2, ext2, zdrop, bandwidth, nullptr, nullptr, &cl);
        const int32_t start1 = e1 - rv.end1, start2 = e2 - rv.end2;
        if (!starts.insert({start1, start2}).second) continue;
        const int32_t f1 = std::min(L1 - start1, max_len), f2 = std::min(L2 - start2, max_len);
        std::vector<uint32_t> cigar;
        BandTrace T;
        SgResult fw{0, -2, -2};
        if (start1 >= 0 && start2 >= 0 && f1 > 0 && f2 > 0)
            fw = semiglobal(seq1 + start1, seq2 + start2, f1, f2, sc, open, ext, open2, ext2, zdrop, bandwidth, &T,
                            &cigar, &cl);
        if (!(cigar.size() > 0 && fw.end1 > 0 && fw.end2 > 0)) continue;
        add_fragment(frags, start1, start2, fw.end1 + start1, fw.end2 + start2, fw.max, cigar, L1d, L2d, lambda,
                     kvalue);
    }
    if (sgcells) *sgcells = cl;
    return emit(frags, ints, dbls, cap, cig, cig_off, cig_cap);
}

int orc_semiglobal_weighted(const int8_t *s1, const int8_t *s2, int32_t L1, int32_t L2, const int16_t *weights,
                            int32_t nprof, const int16_t *pcat, const int32_t *pval, int32_t w, int32_t two_piece,
                            int32_t return_cigar, int32_t *out3, uint32_t *cigar, int32_t cap, int64_t *cells) {
    Profiles P;
    for (int32_t i = 0; i < nprof; ++i) {
        const int32_t *v = pval + 7 * i;
        P.by_cat[pcat[i]] = Profile{v[0], v[1], v[2], v[3], v[4], v[5], v[6]};
    }
    BandTrace T;
    std::vector<uint32_t> cig;
    int64_t cl = 0;
    SgResult r = semiglobal_weighted(s1, s2, L1, L2, weights, P, w, two_piece != 0, return_cigar ? &T : nullptr,
                                     return_cigar ? &cig : nullptr, &cl);
    out3[0] = r.max; out3[1] = r.end1; out3[2] = r.end2;
    if (cells) *cells = cl;
    if (!return_cigar) return 0;
    if ((int32_t)cig.size() > cap) return -1;
    for (size_t i = 0; i < cig.size(); ++i) cigar[i] = cig[i];
    return (int32_t)cig.size();
}

int orc_find_similar_weighted(const int8_t *seq1, int32_t L1, const int8_t *seq2, int32_t L2, int32_t match,
                              int32_t mismatch, int32_t open, int32_t ext, int32_t window, int32_t step,
                              int32_t min_score, int32_t w, int32_t xdrop, int32_t bandwidth, double pvalues,
                              double lambda, double kvalue, const int16_t *wcat, int32_t nprof, const int16_t *pcat,
                              const int32_t *pval, int32_t two_piece, int32_t *ints, double *dbls, int32_t cap,
                              uint32_t *cig, int64_t *cig_off, int64_t cig_cap, int64_t *stats) {
    Profiles P;
    for (int32_t i = 0; i < nprof; ++i) {
        const int32_t *v = pval + 7 * i;
        P.by_cat[pcat[i]] = Profile{v[0], v[1], v[2], v[3], v[4], v[5], v[6]};
    }
    SeedCtx c;
    init_ctx(c, seq1, L1, seq2, L2, match, mismatch, open, ext, window, min_score);
    int64_t nseeds = 0, nlinked = 0, xcells = 0, sgcells = 0;
    std::vector<Seed> xs = extend_seeds(c, open, ext, step, w, xdrop, pvalues, lambda, kvalue, nseeds, nlinked, xcells);
    // weights of the region and of its reverse complement: W[x] = wcat[x + 1]
    // for x = -1 .. L1; the reversed array holds W[L1 - 1 - k] at k
    const int16_t *Wf = wcat + 1;
    std::vector<int16_t> rev((size_t)L1 + 2);
    for (int32_t k = -1; k <= L1; ++k) rev[(size_t)(k + 1)] = Wf[L1 - 1 - k];
    const int16_t *Wr = rev.data() + 1;
    const double L1d = (double)L1, L2d = (double)L2;
    // findSimilarFragmentsForPairedSequence_wighted (:830-925) and _wighted_1gap (:779-826)
    std::vector<Fragment> frags;
    std::set<std::pair<int32_t, int32_t>> starts;
    for (const Seed &x : xs) {
        SgResult rv = semiglobal_weighted(c.rc1.data() + (L1 - 1 - x.end1), c.rc2.data() + (L2 - x.end2 - 1),
                                          1 + x.end1, x.end2 + 1, Wr + (L1 - 1 - x.end1), P, bandwidth,
                                          two_piece != 0, nullptr, nullptr, &sgcells);
        const int32_t start1 = x.end1 - rv.end1, start2 = x.end2 - rv.end2;
        if (!starts.insert({start1, start2}).second) continue;
        std::vector<uint32_t> cigar;
        BandTrace T;
        SgResult fw = semiglobal_weighted(seq1 + start1, seq2 + start2, L1 - start1, L2 - start2, Wf + start1, P,
                                          bandwidth, two_piece != 0, &T, &cigar, &sgcells);
        if (two_piece && !(cigar.size() > 0 && fw.end1 > 0 && fw.end2 > 0)) continue;
        add_fragment(frags, start1, start2, fw.end1 + start1, fw.end2 + start2, fw.max, cigar, L1d, L2d, lambda,
                     kvalue);
    }
    if (stats) {
        stats[0] = c.cells; stats[1] = xcells; stats[2] = nseeds; stats[3] = nlinked;
        stats[4] = sgcells;
    }
    return emit(frags, ints, dbls, cap, cig, cig_off, cig_cap);
}

int orc_sw_max_batch(const int8_t *seqs, const int64_t *ref_off, const int32_t *ref_len, const int64_t *qry_off,
                     const int32_t *qry_len, int32_t n, int32_t match, int32_t mismatch, int32_t open, int32_t ext,
                     int32_t *out_max, int32_t nthreads) {
    if (nthreads < 1) nthreads = 1;
    std::atomic<int32_t> next(0);
    auto work = [&]() {
        for (;;) {
            int32_t k = next.fetch_add(1);
            if (k >= n) break;
            int32_t b, i, j;
            sw_dispatch(seqs + ref_off[k], seqs + qry_off[k], ref_len[k], qry_len[k], Scoring{match, mismatch}, open,
                        ext, b, i, j);
            out_max[k] = b;
        }
    };
    std::vector<std::thread> th;
    for (int32_t t = 1; t < nthreads; ++t) th.emplace_back(work);
    work();
    for (auto &t : th) t.join();
    return 0;
}

}  // extern "C"
