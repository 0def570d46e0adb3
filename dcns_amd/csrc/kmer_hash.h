// kmer_hash.h -- minimizer buckets of the maskGenome k-mer table, shared by
// the host that builds the table and the scan kernel that reads it.
//
// A k-mer (2-bit codes, base j in bits 2j..2j+1) belongs to the bucket of its
// minimizer: of its w = k - l + 1 l-mers, the one with the smallest
// lmer_hash.  Consecutive windows of a genome mostly share their minimizer,
// so a run of windows reads one bucket.  A bucket is one 64-byte slot: word
// 0 holds the key count (low 32 bits) and the start of the keys past the
// seventh in an overflow array (high 32 bits); words 1..7 hold the first
// keys, the unused ones all ones (no code of k <= 31 bases).
// Which bucket a k-mer lands in depends on the k-mer alone, so the lookup is
// exact whatever the hash's quality.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace dcns {

constexpr int kMinimizerL = 12;  // l-mer length for k >= 12
constexpr int kMinimizerW = 9;   // l-mers per k-mer at most
constexpr int kSlotWords = 8;   // 64-bit words per bucket slot
constexpr int kSlotKeys = kSlotWords - 1;

// w for a given k: 1 (the k-mer is its own minimizer) below l + 1
__host__ __device__ inline int minimizer_w(int k) {
    const int w = k - kMinimizerL + 1;
    return w < 1 ? 1 : (w > kMinimizerW ? kMinimizerW : w);
}

__host__ __device__ inline uint32_t fmix32(uint32_t h) {
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}

__host__ __device__ inline uint32_t lmer_hash(uint64_t code) {
    return fmix32((uint32_t)code ^ ((uint32_t)(code >> 32) * 0x9E3779B1u));
}

__host__ __device__ inline uint32_t minimizer_bucket(uint32_t min_hash, uint32_t nb_mask) {
    return fmix32(min_hash ^ 0x5bd1e995u) & nb_mask;
}

// the bucket of a k-mer code (host side; the kernel computes the same from
// the l-mers it already holds)
inline uint32_t kmer_bucket(uint64_t code, int k, uint32_t nb_mask) {
    const int w = minimizer_w(k), l = k - w + 1;
    const uint64_t lmask = l >= 32 ? ~0ull : ((1ull << (2 * l)) - 1);
    uint32_t mn = 0xffffffffu;
    for (int p = 0; p < w; ++p) {
        const uint32_t h = lmer_hash((code >> (2 * p)) & lmask);
        mn = h < mn ? h : mn;
    }
    return minimizer_bucket(mn, nb_mask);
}

}  // namespace dcns
