/* compat.h -- the reference's declared signatures of its dead lambda and
 * Smith-Waterman helpers, as thin C++ aliases over the C-ABI, for callers
 * that include them:
 *   calculateLambda  impl/calculateLambda.h:11   -> dcns_calculate_lambda
 *   calculateK       impl/calculateK.h:12         (the reference's body is empty:
 *                    it returns no value; lambda and K come from
 *                    dcns_fit_lambda_k or the surrogate instead, so no alias)
 *   SmithWaterman    impl/SmithWaterman_bas.h:15-17 -> dcns_sw_bas_batch: the
 *                    body of impl/SmithWaterman_bas.cpp:59-201 (a file the
 *                    reference does not compile; its body has no returnCigar
 *                    argument and walks the traceback whenever
 *                    reverseAlignment is set).  It prints "maxScore <n>" as
 *                    that body does, and the walk's "line 110" message; where
 *                    that body's walk would never end, the alias throws
 *                    std::runtime_error instead of hanging.
 * The Smith-Waterman alias runs on a process-wide context created on first
 * use (device 0). */
#ifndef DCNS_COMPAT_H
#define DCNS_COMPAT_H

#include <cstdint>
#include <iostream>
#include <stdexcept>
#include <vector>

#include "../dcns.h"

inline double calculateLambda(const int8_t *seq1, const int8_t *seq2, const uint32_t &length1,
                              const uint32_t &length2, const int &matchingScore, const int &mismatchingPenalty,
                              const double &epsilon, const double &precision) {
    return dcns_calculate_lambda(seq1, seq2, length1, length2, matchingScore, mismatchingPenalty, epsilon, precision);
}

inline dcns_ctx *dcns_compat_ctx() {
    static dcns_ctx *ctx = [] {
        dcns_ctx *c = nullptr;
        if (dcns_ctx_create(0, 0, &c) != DCNS_OK) throw std::runtime_error(dcns_last_error(nullptr));
        return c;
    }();
    return ctx;
}

inline std::vector<uint32_t> SmithWaterman(int8_t *seq1, int8_t *seq2, const uint32_t &length1, const uint32_t &length2,
                                           const int &_open_gap_penalty, const int &_extend_gap_penalty,
                                           uint32_t &maxScore, uint32_t &endPosition1, uint32_t &endPosition2,
                                           bool reverseAlignment, bool returnCigar) {
    (void)returnCigar;  // no body of the reference takes it
    std::vector<int8_t> buf(seq1, seq1 + length1);
    buf.insert(buf.end(), seq2, seq2 + length2);
    const int64_t off1 = 0, off2 = length1;
    const int32_t l1 = (int32_t)length1, l2 = (int32_t)length2;
    dcns_batch b{buf.data(), &off1, &l1, &off2, &l2, 1, 0};
    uint32_t out[3];
    std::vector<uint32_t> cigar(reverseAlignment ? 2 * ((size_t)length1 + length2) + 4 : 0);
    int64_t cig_off[2] = {0, 0};
    int32_t status = 0;
    if (dcns_sw_bas_batch(dcns_compat_ctx(), &b, _open_gap_penalty, _extend_gap_penalty, reverseAlignment ? 1 : 0, out,
                          cigar.data(), cig_off, (int64_t)cigar.size(), &status) != DCNS_OK)
        throw std::runtime_error(dcns_last_error(dcns_compat_ctx()));
    maxScore = out[0];
    endPosition1 = out[1];
    endPosition2 = out[2];
    std::cout << "maxScore " << maxScore << std::endl;
    if (!reverseAlignment) return std::vector<uint32_t>();
    if (status == 2)
        throw std::runtime_error("SmithWaterman: the traceback of impl/SmithWaterman_bas.cpp never ends on this input");
    if (status == 1) std::cerr << "there is something wrong with smith-waterman algorithm in line 110" << std::endl;
    cigar.resize((size_t)cig_off[1]);
    return cigar;
}

#endif /* DCNS_COMPAT_H */
