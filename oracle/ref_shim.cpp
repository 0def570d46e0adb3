// ref_shim.cpp -- C-callable wrappers around the *compiled reference* (dCNS)
// hot-path functions, linked only into oracle/_ref/libdcnsref.so.
//
// Test infrastructure: this file is the checker's bridge to the reference
// build.  It is compiled against the reference's own headers (the include
// path is given by oracle/Makefile) and never ships in the product library.
//
// Each wrapper names the reference function it calls:
//   ref_sw            -> SmithWaterman(..., bool& positions)   impl/sequenceAlignment3.cpp:220
//   ref_sw_textbook   -> SmithWaterman_non_avx                  impl/sequenceAlignment3.cpp:143
//   ref_xextend       -> SemiGlobal_xextend (score only)       impl/sequenceAlignment3.cpp:1234
//   ref_xextend_tb    -> SemiGlobal_xextend (+ traceback)      impl/sequenceAlignment3.cpp:1318
//   ref_semiglobal    -> SemiGlobal (two-piece affine)         impl/sequenceAlignment3.cpp:932
//   ref_seed_set      -> getSeeds window loop + link_seeds     impl/findSimilarFragmentsForPairedSequence4.cpp:131,288,461-483
//   ref_link_seeds    -> link_seeds on a given list             impl/findSimilarFragmentsForPairedSequence4.cpp:288-344
//   ref_find_similar  -> findSimilarFragmentsForPairedSequence impl/findSimilarFragmentsForPairedSequence4.cpp:433
//   ref_find_similar_2gaps -> findSimilarFragmentsForPairedSequence (overload 4) :752
//   ref_semiglobal_weighted -> SemiGlobal(..., weights, Score&) / SemiGlobal_single_gap_penalty
//                                                             impl/sequenceAlignment3.cpp:1483,1710
//   ref_find_similar_weighted -> findSimilarFragmentsForPairedSequence_wighted(_1gap)
//                                                             impl/findSimilarFragmentsForPairedSequence4.cpp:779,830
//   ref_find_similar_seeded -> findSimilarFragmentsForPairedSequence (overload 3) :623
//   ref_score_profile -> Score(folder) getters                model/Score.cpp:9-104
//   ref_gff_categories -> readGffFileWithEveryThing           impl/gffToCategory.cpp:81-238
#include <cstdint>
#include <cstring>
#include <map>
#include <set>
#include <string>
#include <vector>
#include "impl/impl.h"
#include "model/model.h"

// Defined in the reference translation units but not declared in a header.
void SmithWaterman_non_avx(int8_t *ref, int8_t *read, const int32_t &refLen, const int32_t &readLen,
                           const int &_open_gap_penalty, const int &_extend_gap_penalty, int32_t &maxScore,
                           int32_t &endPosition1, int32_t &endPosition2, const Scorei &scorei, bool &positions);
void getSeeds(int8_t *seq1, int8_t *seq2, int8_t *seq1_rev_com, int8_t *seq2_rev_com, const int32_t &length1,
              const int32_t &length2, const int32_t &miniReferenceSize, const int32_t &_open_gap_penalty1,
              const int32_t &_extend_gap_penalty1, const int32_t &matrix_boundary_distance,
              const int32_t &startPosition2, const int32_t &mini_cns_score, const int32_t &windowsSize,
              int32_t &this_windowsSize_return, const Scorei &m, const int32_t &step_size, std::set<Seed> &seeds);
void link_seeds(std::set<Seed> &seeds);

namespace {
std::vector<int8_t> rc_of(const int8_t *s, int32_t n) {
    // one N past the end: the base the reference's patched reverse alignment
    // of overload 3 reads there (outside its array)
    std::vector<int8_t> r((size_t)n + 1, 2);
    for (int32_t i = 0; i < n; ++i) r[i] = (int8_t)(4 - s[n - 1 - i]);
    return r;
}
std::vector<int8_t> copy_of(const int8_t *s, int32_t n) {
    std::vector<int8_t> r((size_t)n + 1);
    if (n > 0) std::memcpy(r.data(), s, (size_t)n);
    return r;
}
}  // namespace

extern "C" {

// out3 = {maxScore, endPosition1, endPosition2}; positions are left at -1 when
// the reference leaves them unwritten (all-zero matrix on the SIMD path).
int ref_sw(const int8_t *ref, const int8_t *read, int32_t refLen, int32_t readLen, int32_t match,
           int32_t mismatch, int32_t open, int32_t ext, int32_t *out3) {
    Scorei m((int8_t)match, (int8_t)mismatch);
    std::vector<int8_t> a = copy_of(ref, refLen), b = copy_of(read, readLen);
    int32_t mx = 0, e1 = -1, e2 = -1;
    bool pos = true;
    SmithWaterman(a.data(), b.data(), refLen, readLen, open, ext, mx, e1, e2, m, pos);
    out3[0] = mx; out3[1] = e1; out3[2] = e2;
    return 0;
}

int ref_sw_textbook(const int8_t *ref, const int8_t *read, int32_t refLen, int32_t readLen, int32_t match,
                    int32_t mismatch, int32_t open, int32_t ext, int32_t *out3) {
    Scorei m((int8_t)match, (int8_t)mismatch);
    std::vector<int8_t> a = copy_of(ref, refLen), b = copy_of(read, readLen);
    int32_t mx = 0, e1 = -1, e2 = -1;
    bool pos = true;
    SmithWaterman_non_avx(a.data(), b.data(), refLen, readLen, open, ext, mx, e1, e2, m, pos);
    out3[0] = mx; out3[1] = e1; out3[2] = e2;
    return 0;
}

int ref_xextend(const int8_t *s1, const int8_t *s2, int32_t L1, int32_t L2, int32_t match, int32_t mismatch,
                int32_t open, int32_t ext, int32_t xdrop, int32_t w, int32_t *out3) {
    Scorei m((int8_t)match, (int8_t)mismatch);
    std::vector<int8_t> a = copy_of(s1, L1), b = copy_of(s2, L2);
    int32_t mx = 0, e1 = 0, e2 = 0;
    SemiGlobal_xextend(a.data(), b.data(), L1, L2, open, ext, mx, e1, e2, m, xdrop, w);
    out3[0] = mx; out3[1] = e1; out3[2] = e2;
    return 0;
}

// out5 = {maxScore, end1, end2, iii, jjj}; returns the cigar length or -1 if cap is too small.
int ref_xextend_tb(const int8_t *s1, const int8_t *s2, int32_t L1, int32_t L2, int32_t match, int32_t mismatch,
                   int32_t open, int32_t ext, int32_t xdrop, int32_t w, int32_t *out5, uint32_t *cigar, int32_t cap) {
    Scorei m((int8_t)match, (int8_t)mismatch);
    std::vector<int8_t> a = copy_of(s1, L1), b = copy_of(s2, L2);
    int32_t mx = 0, e1 = 0, e2 = 0, iii = 0, jjj = 0;
    Matrix T(L1 + 1, L2 + 1);
    std::vector<uint32_t> c = SemiGlobal_xextend(a.data(), b.data(), L1, L2, open, ext, mx, e1, e2, m, xdrop, w, T, iii, jjj);
    out5[0] = mx; out5[1] = e1; out5[2] = e2; out5[3] = iii; out5[4] = jjj;
    if ((int32_t)c.size() > cap) return -1;
    for (size_t i = 0; i < c.size(); ++i) cigar[i] = c[i];
    return (int32_t)c.size();
}

int ref_semiglobal(const int8_t *s1, const int8_t *s2, int32_t L1, int32_t L2, int32_t match, int32_t mismatch,
                   int32_t open1, int32_t ext1, int32_t open2, int32_t ext2, int32_t zdrop, int32_t w,
                   int32_t return_cigar, int32_t *out3, uint32_t *cigar, int32_t cap) {
    Scorei m((int8_t)match, (int8_t)mismatch);
    std::vector<int8_t> a = copy_of(s1, L1), b = copy_of(s2, L2);
    int32_t mx = 0, e1 = 0, e2 = 0;
    Matrix T(return_cigar ? L1 + 1 : 1, return_cigar ? L2 + 1 : 1);
    std::vector<uint32_t> c = SemiGlobal(a.data(), b.data(), L1, L2, open1, ext1, open2, ext2, mx, e1, e2, m,
                                         return_cigar != 0, zdrop, w, T);
    out3[0] = mx; out3[1] = e1; out3[2] = e2;
    if ((int32_t)c.size() > cap) return -1;
    for (size_t i = 0; i < c.size(); ++i) cigar[i] = c[i];
    return (int32_t)c.size();
}

// Runs the window loop of findSimilarFragmentsForPairedSequence (overload 1)
// and returns the seed set in set order, before (linked=0) or after
// (linked=1) link_seeds.  out holds 4 int32 per seed (start1,end1,start2,end2).
int ref_seed_set(const int8_t *seq1, int32_t L1, const int8_t *seq2, int32_t L2, int32_t match, int32_t mismatch,
                 int32_t open, int32_t ext, int32_t window, int32_t step, int32_t min_score, int32_t linked,
                 int32_t *out, int32_t cap) {
    Scorei m((int8_t)match, (int8_t)mismatch);
    std::vector<int8_t> a = copy_of(seq1, L1), b = copy_of(seq2, L2);
    std::vector<int8_t> ar = rc_of(seq1, L1), br = rc_of(seq2, L2);
    int32_t miniReferenceSize = min_score / match;
    int32_t mbd = 0, sp = 0, thisw = 0;
    std::set<Seed> seeds;
    while (true) {
        getSeeds(a.data(), b.data(), ar.data(), br.data(), L1, L2, miniReferenceSize, open, ext, mbd, sp, min_score,
                 window, thisw, m, step, seeds);
        sp += step;
        if (sp > L2) break;
        if (sp + window > L2) {
            getSeeds(a.data(), b.data(), ar.data(), br.data(), L1, L2, miniReferenceSize, open, ext, mbd, sp,
                     min_score, window, thisw, m, step, seeds);
            break;
        }
    }
    if (linked) link_seeds(seeds);
    int32_t n = 0;
    for (const Seed &s : seeds) {
        if (n < cap) {
            out[4 * n + 0] = s.getStart1(); out[4 * n + 1] = s.getEnd1();
            out[4 * n + 2] = s.getStart2(); out[4 * n + 3] = s.getEnd2();
        }
        ++n;
    }
    return n <= cap ? n : -n;
}

// link_seeds on a given list: the seeds inserted into a std::set<Seed> in the
// order given, then the reference's link_seeds; out in set order.
int ref_link_seeds(const int32_t *seeds, int32_t n, int32_t *out, int32_t cap) {
    std::set<Seed> set;
    for (int32_t k = 0; k < n; ++k) {
        int32_t s1 = seeds[4 * k], e1 = seeds[4 * k + 1], s2 = seeds[4 * k + 2], e2 = seeds[4 * k + 3];
        set.insert(Seed(s1, e1, s2, e2));
    }
    link_seeds(set);
    if ((int64_t)set.size() > cap) return -1;
    int32_t o = 0;
    for (const Seed &s : set) {
        out[4 * o + 0] = s.getStart1(); out[4 * o + 1] = s.getEnd1();
        out[4 * o + 2] = s.getStart2(); out[4 * o + 3] = s.getEnd2();
        ++o;
    }
    return o;
}

// findSimilarFragmentsForPairedSequence overload 1.  ints = 5 per fragment
// (start1,end1,start2,end2,score, 1-based as stored), dbls = 2 per fragment
// (pValue, eValue), cigar ops concatenated with cig_off[n+1] offsets.
// Returns the fragment count, or -(count) when a capacity is exceeded.
int ref_find_similar(const int8_t *seq1, int32_t L1, const int8_t *seq2, int32_t L2, int32_t match,
                     int32_t mismatch, int32_t open, int32_t ext, int32_t window, int32_t step, int32_t min_score,
                     int32_t w, int32_t xdrop, double pvalues, double lambda, double kvalue, int32_t *ints,
                     double *dbls, int32_t cap, uint32_t *cig, int64_t *cig_off, int64_t cig_cap) {
    Scorei m((int8_t)match, (int8_t)mismatch);
    std::vector<int8_t> a = copy_of(seq1, L1), b = copy_of(seq2, L2);
    std::vector<int8_t> ar = rc_of(seq1, L1), br = rc_of(seq2, L2);
    std::string s1str, s2str;
    int32_t l1 = L1, l2 = L2, win = window, mbd = 0;
    std::vector<PairedSimilarFragment> f = findSimilarFragmentsForPairedSequence(
        a.data(), ar.data(), b.data(), br.data(), l1, l2, win, min_score, mbd, open, ext, match, mismatch, m, step,
        s1str, s2str, pvalues, lambda, kvalue, w, xdrop);
    int32_t n = (int32_t)f.size();
    int64_t total = 0;
    for (auto &x : f) total += (int64_t)x.getCigar().size();
    if (n > cap || total > cig_cap) return -n - 1;
    int64_t o = 0;
    for (int32_t i = 0; i < n; ++i) {
        ints[5 * i + 0] = (int32_t)f[i].getStart1(); ints[5 * i + 1] = (int32_t)f[i].getEnd1();
        ints[5 * i + 2] = (int32_t)f[i].getStart2(); ints[5 * i + 3] = (int32_t)f[i].getEnd2();
        ints[5 * i + 4] = (int32_t)f[i].getScore();
        dbls[2 * i + 0] = f[i].getPValue(); dbls[2 * i + 1] = f[i].getEValue();
        cig_off[i] = o;
        for (uint32_t c : f[i].getCigar()) cig[o++] = c;
    }
    cig_off[n] = o;
    return n;
}

// findSimilarFragmentsForPairedSequence overload 4 (two-piece path,
// impl/findSimilarFragmentsForPairedSequence4.cpp:752).  Layout as
// ref_find_similar.
int ref_find_similar_2gaps(const int8_t *seq1, int32_t L1, const int8_t *seq2, int32_t L2, int32_t match,
                           int32_t mismatch, int32_t open, int32_t ext, int32_t open2, int32_t ext2, int32_t window,
                           int32_t step, int32_t min_score, int32_t w, int32_t xdrop, int32_t zdrop,
                           int32_t bandwidth, double pvalues, double lambda, double kvalue, int32_t *ints,
                           double *dbls, int32_t cap, uint32_t *cig, int64_t *cig_off, int64_t cig_cap) {
    Scorei m((int8_t)match, (int8_t)mismatch);
    std::vector<int8_t> a = copy_of(seq1, L1), b = copy_of(seq2, L2);
    std::vector<int8_t> ar = rc_of(seq1, L1), br = rc_of(seq2, L2);
    std::string s1str, s2str;
    int32_t l1 = L1, l2 = L2, win = window, mbd = 0;
    std::vector<PairedSimilarFragment> f = findSimilarFragmentsForPairedSequence(
        a.data(), ar.data(), b.data(), br.data(), l1, l2, win, min_score, mbd, open, ext, open2, ext2, match,
        mismatch, m, step, s1str, s2str, pvalues, lambda, kvalue, zdrop, bandwidth, w, xdrop);
    int32_t n = (int32_t)f.size();
    int64_t total = 0;
    for (auto &x : f) total += (int64_t)x.getCigar().size();
    if (n > cap || total > cig_cap) return -n - 1;
    int64_t o = 0;
    for (int32_t i = 0; i < n; ++i) {
        ints[5 * i + 0] = (int32_t)f[i].getStart1(); ints[5 * i + 1] = (int32_t)f[i].getEnd1();
        ints[5 * i + 2] = (int32_t)f[i].getStart2(); ints[5 * i + 3] = (int32_t)f[i].getEnd2();
        ints[5 * i + 4] = (int32_t)f[i].getScore();
        dbls[2 * i + 0] = f[i].getPValue(); dbls[2 * i + 1] = f[i].getEValue();
        cig_off[i] = o;
        for (uint32_t c : f[i].getCigar()) cig[o++] = c;
    }
    cig_off[n] = o;
    return n;
}

// Score(folder) and its getters for one category: vals7 = {match, mismatch,
// open1, ext1, open2, ext2, zdrop}.
int ref_score_profile(const char *folder, int16_t category, int32_t *vals7) {
    Score score{std::string(folder)};
    int32_t **m = score.getM(category);
    if (!m) return -1;
    vals7[0] = m[0][0];
    vals7[1] = m[0][1];
    vals7[2] = score.getOpenGapPenalty1(category);
    vals7[3] = score.getExtendGapPenalty1(category);
    vals7[4] = score.getOpenGapPenalty2(category);
    vals7[5] = score.getExtendGapPenalty2(category);
    vals7[6] = score.getZdrop(category);
    return 0;
}

// Weighted SemiGlobal on one task; weights[-1 .. L1] readable.  The score
// profiles come from the folder.  out3 = {max, end1, end2}; returns the cigar
// length (return_cigar) or 0, -1 when cap is too small.
int ref_semiglobal_weighted(const int8_t *s1, const int8_t *s2, int32_t L1, int32_t L2, const int16_t *weights,
                            const char *folder, int32_t w, int32_t two_piece, int32_t return_cigar, int32_t *out3,
                            uint32_t *cigar, int32_t cap) {
    Score score{std::string(folder)};
    std::vector<int8_t> a = copy_of(s1, L1), b = copy_of(s2, L2);
    int32_t l1 = L1, l2 = L2, mx = 0, e1 = -1, e2 = -1, z = 0, ww = w;
    Matrix T(L1 + 1, L2 + 1);
    std::vector<uint32_t> cig =
        two_piece ? SemiGlobal(a.data(), b.data(), l1, l2, weights, score, mx, e1, e2, return_cigar != 0, z, ww, T)
                  : SemiGlobal_single_gap_penalty(a.data(), b.data(), l1, l2, weights, score, mx, e1, e2,
                                                  return_cigar != 0, z, ww, T);
    out3[0] = mx; out3[1] = e1; out3[2] = e2;
    if (!return_cigar) return 0;
    if ((int32_t)cig.size() > cap) return -1;
    for (size_t i = 0; i < cig.size(); ++i) cigar[i] = cig[i];
    return (int32_t)cig.size();
}

// findSimilarFragmentsForPairedSequence_wighted (two_piece = 1) or _1gap.
// wcat: categories of reference positions -1 .. L1 (L1 + 2 values), i.e. the
// reference's weight pointer at [-1 .. L1]; the reversed array is built as
// readGffFileWithEveryThing builds categories_rev.  Layout as ref_find_similar.
int ref_find_similar_weighted(const int8_t *seq1, int32_t L1, const int8_t *seq2, int32_t L2, int32_t match,
                              int32_t mismatch, int32_t open, int32_t ext, int32_t window, int32_t step,
                              int32_t min_score, int32_t w, int32_t xdrop, int32_t bandwidth, double pvalues,
                              double lambda, double kvalue, const int16_t *wcat, const char *folder,
                              int32_t two_piece, int32_t *ints, double *dbls, int32_t cap, uint32_t *cig,
                              int64_t *cig_off, int64_t cig_cap) {
    Scorei m((int8_t)match, (int8_t)mismatch);
    Score score{std::string(folder)};
    std::vector<int8_t> a = copy_of(seq1, L1), b = copy_of(seq2, L2);
    std::vector<int8_t> ar = rc_of(seq1, L1), br = rc_of(seq2, L2);
    std::vector<int16_t> fwd(wcat, wcat + L1 + 2), rev((size_t)L1 + 2);
    for (int32_t k = -1; k <= L1; ++k) rev[(size_t)(k + 1)] = wcat[(L1 - 1 - k) + 1];
    std::string s1str, s2str;
    int32_t l1 = L1, l2 = L2, win = window, mbd = 0, ww = w, zd = 0;
    std::vector<PairedSimilarFragment> f =
        two_piece ? findSimilarFragmentsForPairedSequence_wighted(
                        a.data(), ar.data(), b.data(), br.data(), l1, l2, win, min_score, mbd, open, ext, match,
                        mismatch, m, step, s1str, s2str, pvalues, lambda, kvalue, zd, bandwidth, ww, xdrop, score,
                        fwd.data() + 1, rev.data() + 1)
                  : findSimilarFragmentsForPairedSequence_wighted_1gap(
                        a.data(), ar.data(), b.data(), br.data(), l1, l2, win, min_score, mbd, open, ext, match,
                        mismatch, m, step, s1str, s2str, pvalues, lambda, kvalue, zd, bandwidth, ww, xdrop, score,
                        fwd.data() + 1, rev.data() + 1);
    int32_t n = (int32_t)f.size();
    int64_t total = 0;
    for (auto &x : f) total += (int64_t)x.getCigar().size();
    if (n > cap || total > cig_cap) return -n - 1;
    int64_t o = 0;
    for (int32_t i = 0; i < n; ++i) {
        ints[5 * i + 0] = (int32_t)f[i].getStart1(); ints[5 * i + 1] = (int32_t)f[i].getEnd1();
        ints[5 * i + 2] = (int32_t)f[i].getStart2(); ints[5 * i + 3] = (int32_t)f[i].getEnd2();
        ints[5 * i + 4] = (int32_t)f[i].getScore();
        dbls[2 * i + 0] = f[i].getPValue(); dbls[2 * i + 1] = f[i].getEValue();
        cig_off[i] = o;
        for (uint32_t c : f[i].getCigar()) cig[o++] = c;
    }
    cig_off[n] = o;
    return n;
}

// readGffFileWithEveryThing over chromosomes (names, sizes); copies the
// categories of chromosome `which` at [from, from + count) into out.
int ref_gff_categories(const char *gff, const char *const *names, const int32_t *sizes, int32_t nchr, int32_t which,
                       int64_t from, int64_t count, int16_t *out) {
    std::map<std::string, int32_t> chrSize;
    for (int32_t i = 0; i < nchr; ++i) chrSize[names[i]] = sizes[i];
    std::map<std::string, int16_t *> cats;
    readGffFileWithEveryThing(std::string(gff), chrSize, cats);
    const int16_t *c = cats[names[which]];
    for (int64_t k = 0; k < count; ++k) out[k] = c[from + k];
    for (auto &kv : cats) delete[] kv.second;
    return 0;
}

// findSimilarFragmentsForPairedSequence overload 3 (the cut2Gaps2 path,
// impl/findSimilarFragmentsForPairedSequence4.cpp:623): seeds as 4 ints
// (start1, end1, start2, end2); layout as ref_find_similar.
int ref_find_similar_seeded(const int8_t *seq1, int32_t L1, const int8_t *seq2, int32_t L2, const int32_t *seeds,
                            int32_t nseeds, int32_t match, int32_t mismatch, int32_t open, int32_t ext,
                            int32_t open2, int32_t ext2, int32_t zdrop, int32_t bandwidth, int32_t max_len,
                            double lambda, double kvalue, int32_t *ints, double *dbls, int32_t cap, uint32_t *cig,
                            int64_t *cig_off, int64_t cig_cap) {
    Scorei m((int8_t)match, (int8_t)mismatch);
    std::vector<int8_t> a = copy_of(seq1, L1), b = copy_of(seq2, L2);
    std::vector<int8_t> ar = rc_of(seq1, L1), br = rc_of(seq2, L2);
    std::string s1str, s2str;
    std::vector<Seed> xs;
    for (int32_t k = 0; k < nseeds; ++k) {
        int32_t s1 = seeds[4 * k], e1 = seeds[4 * k + 1], s2 = seeds[4 * k + 2], e2 = seeds[4 * k + 3];
        xs.push_back(Seed(s1, e1, s2, e2));
    }
    int32_t l1 = L1, l2 = L2, win = 38, ml = max_len;
    const int32_t tdim = std::min(max_len, std::max(L1, L2) + 2) + 1;
    Matrix T(tdim, tdim);
    std::vector<PairedSimilarFragment> f = findSimilarFragmentsForPairedSequence(
        a.data(), ar.data(), b.data(), br.data(), l1, l2, win, open, ext, open2, ext2, match, mismatch, m, s1str,
        s2str, 0.1, lambda, kvalue, zdrop, bandwidth, xs, T, ml);
    int32_t n = (int32_t)f.size();
    int64_t total = 0;
    for (auto &x : f) total += (int64_t)x.getCigar().size();
    if (n > cap || total > cig_cap) return -n - 1;
    int64_t o = 0;
    for (int32_t i = 0; i < n; ++i) {
        ints[5 * i + 0] = (int32_t)f[i].getStart1(); ints[5 * i + 1] = (int32_t)f[i].getEnd1();
        ints[5 * i + 2] = (int32_t)f[i].getStart2(); ints[5 * i + 3] = (int32_t)f[i].getEnd2();
        ints[5 * i + 4] = (int32_t)f[i].getScore();
        dbls[2 * i + 0] = f[i].getPValue(); dbls[2 * i + 1] = f[i].getEValue();
        cig_off[i] = o;
        for (uint32_t c : f[i].getCigar()) cig[o++] = c;
    }
    cig_off[n] = o;
    return n;
}

}  // extern "C"
