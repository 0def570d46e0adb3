/* compat.h -- the reference's declared signatures of its dead lambda and
 * Smith-Waterman helpers, as thin C++ aliases over the C-ABI, for callers
 * that include them:
 *   calculateLambda  impl/calculateLambda.h:11   -> dcns_calculate_lambda
 *   calculateK       impl/calculateK.h:12         (the reference's body is empty:
 *                    it returns no value; lambda and K come from
 *                    dcns_fit_lambda_k or the surrogate instead, so no alias)
 *   SmithWaterman    impl/SmithWaterman_bas.h:15-17 (a file the reference does
 *                    not compile) -> dcns_sw_batch with its fixed +2 / -1
 *                    scoring; maximum and first end position only, no cigar.
 * The Smith-Waterman alias runs on a process-wide context created on first
 * use (device 0). */
#ifndef DCNS_COMPAT_H
#define DCNS_COMPAT_H

#include <cstdint>
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
    (void)reverseAlignment;
    (void)returnCigar;
    std::vector<int8_t> buf(seq1, seq1 + length1);
    buf.insert(buf.end(), seq2, seq2 + length2);
    const int64_t off1 = 0, off2 = length1;
    const int32_t l1 = (int32_t)length1, l2 = (int32_t)length2;
    dcns_batch b{buf.data(), &off1, &l1, &off2, &l2, 1, 0};
    int32_t out[3];
    if (dcns_sw_batch(dcns_compat_ctx(), &b, 2, -1, _open_gap_penalty, _extend_gap_penalty, out) != DCNS_OK)
        throw std::runtime_error(dcns_last_error(dcns_compat_ctx()));
    maxScore = (uint32_t)out[0];
    endPosition1 = (uint32_t)out[1];
    endPosition2 = (uint32_t)out[2];
    return std::vector<uint32_t>();
}

#endif /* DCNS_COMPAT_H */
