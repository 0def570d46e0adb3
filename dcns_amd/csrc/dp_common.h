// dp_common.h -- device helpers shared by the gfx950 DP kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

#include "dcns_types.h"

namespace dcns {

typedef short s16x2 __attribute__((ext_vector_type(2)));
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ s16x2 as_s2(uint32_t v) { return __builtin_bit_cast(s16x2, v); }
__device__ __forceinline__ u16x2 as_u2(uint32_t v) { return __builtin_bit_cast(u16x2, v); }
__device__ __forceinline__ uint32_t bits(s16x2 v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ uint32_t bits(u16x2 v) { return __builtin_bit_cast(uint32_t, v); }

// v_pk_add_u16 (wrapping; operands stay far from the int16 limits)
__device__ __forceinline__ uint32_t pk_add(uint32_t a, uint32_t b) { return bits(as_s2(a) + as_s2(b)); }
// v_pk_max_i16
__device__ __forceinline__ uint32_t pk_max_s(uint32_t a, uint32_t b) {
    return bits(__builtin_elementwise_max(as_s2(a), as_s2(b)));
}
// v_pk_max_u16
__device__ __forceinline__ uint32_t pk_max_u(uint32_t a, uint32_t b) {
    return bits(__builtin_elementwise_max(as_u2(a), as_u2(b)));
}
// v_pk_sub_u16 clamp: max(a - b, 0) per half for non-negative a
__device__ __forceinline__ uint32_t pk_subsat(uint32_t a, uint32_t b) {
    return bits(__builtin_elementwise_sub_sat(as_u2(a), as_u2(b)));
}

// Substitution scores through v_perm_b32.  For a reference base c the
// 64-bit table {hi, lo} holds s(c, A), s(c, C), s(c, G), s(c, T) in bytes
// 1, 3, 5, 7 (bytes 0-3 = lo, 4-7 = hi).  A per-column selector picks the
// byte of the read base and fills the upper bytes with its sign
// (selector values 8..11 copy bit 7 of bytes 1, 3, 5, 7), so one v_perm_b32
// yields a signed score.  Selector 12 yields 0: the score of N, and of
// padding columns past the end of a window.
struct ScoreTab {
    uint32_t lo[8], hi[8];  // indexed by base code; codes outside A,C,G,T give zero scores
};

__host__ __device__ inline int8_t score_of(int32_t match, int32_t mismatch, int a, int b) {
    // Scorei, model/Score.cpp:105-125
    if (a == 2 || b == 2 || a < 0 || b < 0 || a > 4 || b > 4) return 0;
    return (int8_t)(a == b ? match : mismatch);
}

inline ScoreTab make_score_tab(int32_t match, int32_t mismatch) {
    ScoreTab t{};
    const int cls[4] = {0, 1, 3, 4};
    for (int c = 0; c < 8; ++c) {
        uint32_t lo = 0, hi = 0;
        for (int k = 0; k < 4; ++k) {
            uint32_t s = (uint8_t)score_of(match, mismatch, c, cls[k]);
            if (k < 2) lo |= s << (8 * (2 * k + 1));
            else hi |= s << (8 * (2 * (k - 2) + 1));
        }
        t.lo[c] = lo;
        t.hi[c] = hi;
    }
    return t;
}

// Selector of a read base (branch-free): A,C,G,T -> class k = 0..3 picks
// score byte 2k+1 and sign code 8+k; N and anything else -> 12 (zero).
__host__ __device__ __forceinline__ uint32_t sel16(int q) {
    const int k = q - (q > 2);
    const bool base = (q >= 0) & (q <= 4) & (q != 2);
    return base ? (uint32_t)((2 * k + 1) | ((8 + k) << 8)) : 0x0C0Cu;
}
// 32-bit selector: sign-extended int32 score.
__host__ __device__ __forceinline__ uint32_t sel32(int q) {
    const int k = q - (q > 2);
    const bool base = (q >= 0) & (q <= 4) & (q != 2);
    return base ? (uint32_t)((2 * k + 1) | ((8 + k) * 0x010101u << 8)) : 0x0C0C0C0Cu;
}

// Packed-f16 variant for the window scan: every DP value there is an integer
// of magnitude <= 2048 (window * match <= 2048, penalties >= -128), which f16
// represents exactly, and gfx950 has a three-input packed max
// (v_pk_maximum3_f16).  Table bytes (2k, 2k+1) hold the f16 score of class k;
// selector 0x0C0C still yields +0.
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ f16x2 as_h2(uint32_t v) { return __builtin_bit_cast(f16x2, v); }
__device__ __forceinline__ uint32_t bits(f16x2 v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ f16x2 hmax3(f16x2 a, f16x2 b, f16x2 c) {
    return __builtin_elementwise_maximum(__builtin_elementwise_maximum(a, b), c);
}

inline uint16_t f16_bits_of_int(int v) {  // exact for |v| <= 2048
    if (v == 0) return 0;
    const uint16_t sign = v < 0 ? 0x8000u : 0u;
    uint32_t a = (uint32_t)(v < 0 ? -v : v);
    const int e = 31 - __builtin_clz(a);  // a in [2^e, 2^(e+1)), e <= 11
    const uint32_t mant = (e <= 10 ? (a << (10 - e)) : (a >> (e - 10))) & 0x3ffu;
    return (uint16_t)(sign | ((uint32_t)(e + 15) << 10) | mant);
}

inline ScoreTab make_score_tab_f16(int32_t match, int32_t mismatch) {
    ScoreTab t{};
    const int cls[4] = {0, 1, 3, 4};
    for (int c = 0; c < 8; ++c) {
        uint32_t lo = 0, hi = 0;
        for (int k = 0; k < 4; ++k) {
            const uint32_t h = f16_bits_of_int(score_of(match, mismatch, c, cls[k]));
            if (k < 2) lo |= h << (16 * k);
            else hi |= h << (16 * (k - 2));
        }
        t.lo[c] = lo;
        t.hi[c] = hi;
    }
    return t;
}

__host__ __device__ __forceinline__ uint32_t sel16_f16(int q) {
    const int k = q - (q > 2);
    const bool base = (q >= 0) & (q <= 4) & (q != 2);
    return base ? (uint32_t)((2 * k) | ((2 * k + 1) << 8)) : 0x0C0Cu;
}

__device__ __forceinline__ uint32_t perm(uint32_t hi, uint32_t lo, uint32_t sel) {
    return __builtin_amdgcn_perm(hi, lo, sel);
}

__device__ __forceinline__ int32_t max3i(int32_t a, int32_t b, int32_t c) {
    return __builtin_elementwise_max(__builtin_elementwise_max(a, b), c);
}

// Which recurrence SmithWaterman applies (sequenceAlignment3.cpp:224-225,
// :338-339): the striped int8/int16 variant, or the scalar textbook one.
__host__ __device__ inline bool sw_simd_variant(int64_t refLen, int64_t readLen, int64_t match, int64_t open,
                                                int64_t ext) {
    const int64_t b = 32767;
    return ((-refLen * ext - open) < b || (-readLen * ext - open) < b) && (refLen * match < b || readLen * match < b);
}

}  // namespace dcns
