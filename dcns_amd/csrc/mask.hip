// mask.hip -- maskGenome on the device (impl/maskGenome.cpp:50-121): the
// k-mer scan that marks every position covered by a window [i, i + k) whose
// text is one of a set of k-mers, the CDS ranges painted over that mask, and
// the FASTA bodies written from sequence and mask.
//
// The genome is one byte buffer of concatenated sequences with offsets
// seg_off[0..nseg] (seg_off[nseg] = N).  Windows never cross a sequence end.
//
// K-mer set, two parts:
//   * members made of upper-case A/C/G/T only, k <= 31: 2-bit codes
//     ((b >> 1) & 3 per byte: A 0, C 1, T 2, G 3) in a bucketed open-addressing
//     table (4 keys = 32 bytes per bucket, linear probing over buckets, load
//     <= 1/2, empty = all ones, which no code of k <= 31 bases equals);
//   * every other member (other bytes, or k >= 32): a table of member
//     indices keyed by a 64-bit polynomial hash of the bytes, each hit
//     verified byte by byte.
// A window of upper-case A/C/G/T can only equal a member of the first part;
// any other window only one of the second.  Lower-case windows never match,
// as the reference compares the text as read against upper-cased members.
//
// k_kmer_scan: one lane per CHUNK positions.  It evaluates the windows from
// k - 1 (rounded up to the group size) before its chunk, keeps the start of
// the last window that hit, and writes mask[j] = 1 when a hit window starts
// in [j - k + 1, j].  Windows go in groups of G: the G bucket loads are issued
// together so each lane has G probes in flight.  HBM traffic per position: 1
// byte read, 1 byte written, plus the probes into the table.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace dcns {
namespace {

constexpr uint64_t kEmpty = ~0ull;
constexpr int CHUNK = 1024;  // positions per lane in the scan
constexpr int G = 8;         // windows per probe group
constexpr int BUCKET = 4;    // keys per bucket
constexpr int EMIT = 16;     // output bytes per lane in the FASTA writer

__device__ __forceinline__ uint64_t mix(uint64_t x) {
    // splitmix64 finaliser (a bijection on 64 bits)
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebull;
    x ^= x >> 31;
    return x;
}

// (b >> 1) & 3 for upper-case A, C, G, T; -1 for every other byte
__device__ __forceinline__ int base2(uint32_t b) {
    const uint32_t d = b - 64u;
    return (d < 32u && ((0x10008Au >> d) & 1u)) ? (int)((b >> 1) & 3u) : -1;
}

__global__ void k_kmer_insert(const uint64_t *__restrict__ codes, int64_t n, uint64_t *__restrict__ table,
                              uint64_t bmask) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t key = codes[i];
    uint64_t b = mix(key) & bmask;
    for (uint64_t t = 0; t <= bmask; ++t) {  // at most half the slots are used: an empty slot is always found
        unsigned long long *bk = (unsigned long long *)(table + b * BUCKET);
        for (int s = 0; s < BUCKET; ++s) {
            const unsigned long long prev = atomicCAS(bk + s, (unsigned long long)kEmpty, (unsigned long long)key);
            if (prev == kEmpty || prev == key) return;
        }
        b = (b + 1) & bmask;
    }
}

__global__ void k_kmer_insert_general(const uint64_t *__restrict__ hashes, int64_t n, uint32_t *__restrict__ gtab,
                                      uint64_t gmask) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t s = mix(hashes[i]) & gmask;
    for (uint64_t t = 0; t <= gmask; ++t) {
        if (atomicCAS(gtab + s, 0u, (uint32_t)(i + 1)) == 0u) return;
        s = (s + 1) & gmask;
    }
}

// Continues a lookup whose home bucket was full and did not hold the key.
__device__ __noinline__ bool probe_on(const uint64_t *__restrict__ table, uint64_t bmask, uint64_t b, uint64_t key) {
    for (uint64_t t = 0; t < bmask; ++t) {
        b = (b + 1) & bmask;
        const uint64_t *bk = table + b * BUCKET;
        for (int s = 0; s < BUCKET; ++s) {
            const uint64_t v = bk[s];
            if (v == key) return true;
            if (v == kEmpty) return false;
        }
    }
    return false;
}

__device__ __noinline__ bool general_lookup(const KmerGeneral &g, const uint8_t *__restrict__ win, uint64_t h,
                                            int32_t k) {
    uint64_t s = mix(h) & g.mask;
    for (uint64_t t = 0; t <= g.mask; ++t) {
        const uint32_t v = g.table[s];
        if (v == 0) return false;
        const int64_t idx = (int64_t)v - 1;
        if (g.hashes[idx] == h) {
            const uint8_t *m = g.bytes + idx * (int64_t)k;
            int32_t j = 0;
            while (j < k && m[j] == win[j]) ++j;
            if (j == k) return true;
        }
        s = (s + 1) & g.mask;
    }
    return false;
}

template <bool CODES, bool GEN>
__global__ __launch_bounds__(256) void k_kmer_scan(const uint8_t *__restrict__ seq, int64_t N,
                                                   const int64_t *__restrict__ seg_off, int32_t nseg, int32_t k,
                                                   const uint64_t *__restrict__ table, uint64_t bmask,
                                                   KmerGeneral g, uint8_t *__restrict__ mask) {
    const int64_t c0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * CHUNK;
    if (c0 >= N) return;
    const int64_t c1 = c0 + CHUNK < N ? c0 + CHUNK : N;
    const int32_t back = (k - 1 + G - 1) / G * G;
    const int64_t ib = c0 - back > 0 ? c0 - back : 0;  // first window evaluated (a multiple of G)

    // the sequence holding byte ib: the last s with seg_off[s] <= ib
    int32_t lo = 0, hi = nseg - 1;
    while (lo < hi) {
        const int32_t mid = (lo + hi + 1) >> 1;
        if (seg_off[mid] <= ib) lo = mid;
        else hi = mid - 1;
    }
    int32_t s = lo;
    int64_t seg_from = seg_off[s], seg_next = seg_off[s + 1];

    const uint64_t cmask = (k >= 32) ? ~0ull : ((1ull << (2 * k)) - 1ull);
    uint64_t code = 0, h = 0;
    int64_t bad_from = ib;  // windows starting before this hold a byte other than upper-case A/C/G/T

    auto feed = [&](int64_t p) {
        const uint32_t b = seq[p];
        const int v = base2(b);
        if (v < 0) bad_from = p + 1;
        code = ((code << 2) | (uint64_t)(v & 3)) & cmask;
        while (p >= seg_next) {
            ++s;
            seg_from = seg_off[s];
            seg_next = seg_off[s + 1];
        }
        if (GEN) {
            const uint64_t out = (p - k >= ib) ? (uint64_t)seq[p - k] : 0ull;
            h = (h - out * g.pk1) * kKmerHashBase + (uint64_t)b;
        }
    };

    for (int64_t p = ib; p < ib + k - 1 && p < N; ++p) feed(p);

    int64_t last_hit = INT64_MIN / 2;
    for (int64_t i0 = ib; i0 < c1; i0 += G) {
        uint64_t key[G];
        bool va[G], vg[G];
        uint64_t hh[G];
#pragma unroll
        for (int t = 0; t < G; ++t) {
            const int64_t i = i0 + t, p = i + k - 1;
            bool in_seq = false;
            if (p < N) {
                feed(p);
                in_seq = i >= seg_from;
            }
            va[t] = CODES && in_seq && i >= bad_from;
            vg[t] = GEN && in_seq && (!CODES || i < bad_from);
            key[t] = code;
            hh[t] = h;
        }
        bool hit[G];
        if (CODES) {
            ulonglong2 x0[G], x1[G];
            uint64_t home[G];
#pragma unroll
            for (int t = 0; t < G; ++t) {
                home[t] = va[t] ? (mix(key[t]) & bmask) : 0;
                const ulonglong2 *bk = (const ulonglong2 *)(table + home[t] * BUCKET);
                x0[t] = bk[0];
                x1[t] = bk[1];
            }
#pragma unroll
            for (int t = 0; t < G; ++t) {
                const uint64_t kk = key[t];
                const bool found = x0[t].x == kk || x0[t].y == kk || x1[t].x == kk || x1[t].y == kk;
                const bool open = x0[t].x == kEmpty || x0[t].y == kEmpty || x1[t].x == kEmpty || x1[t].y == kEmpty;
                hit[t] = va[t] && (found || (!open && probe_on(table, bmask, home[t], kk)));
            }
        } else {
#pragma unroll
            for (int t = 0; t < G; ++t) hit[t] = false;
        }
        if (GEN) {
#pragma unroll
            for (int t = 0; t < G; ++t)
                if (vg[t]) hit[t] = general_lookup(g, seq + i0 + t, hh[t], k);
        }
        uint32_t lo4 = 0, hi4 = 0;
#pragma unroll
        for (int t = 0; t < G; ++t) {
            const int64_t i = i0 + t;
            if (hit[t]) last_hit = i;
            const uint32_t m = (last_hit > i - k) ? 1u : 0u;
            if (t < 4) lo4 |= m << (8 * t);
            else hi4 |= m << (8 * (t - 4));
        }
        if (i0 >= c0) *(uint2 *)(mask + i0) = make_uint2(lo4, hi4);  // the buffer is padded to CHUNK
    }
}

__global__ __launch_bounds__(256) void k_paint_ranges(const int64_t *__restrict__ ranges, int64_t n,
                                                      uint8_t *__restrict__ mask) {
    for (int64_t r = blockIdx.x; r < n; r += gridDim.x) {
        const int64_t a = ranges[2 * r], b = ranges[2 * r + 1];
        for (int64_t x = a + threadIdx.x; x < b; x += blockDim.x) mask[x] = 1;
    }
}

__global__ __launch_bounds__(256) void k_mask_emit(const uint8_t *__restrict__ seq, const uint8_t *__restrict__ mask,
                                                   const int64_t *__restrict__ seg_off,
                                                   const int64_t *__restrict__ body_off, int32_t nseg, int64_t total,
                                                   int32_t width, uint32_t alt, int32_t soft,
                                                   uint8_t *__restrict__ out) {
    const int64_t q0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * EMIT;
    if (q0 >= total) return;
    int32_t lo = 0, hi = nseg - 1;
    while (lo < hi) {
        const int32_t mid = (lo + hi + 1) >> 1;
        if (body_off[mid] <= q0) lo = mid;
        else hi = mid - 1;
    }
    int32_t s = lo;
    int64_t r = q0 - body_off[s], blen = body_off[s + 1] - body_off[s], base = seg_off[s];
    int64_t col = 0, i = r;
    if (width > 0) {
        const int64_t line = r / (width + 1);
        col = r - line * (width + 1);
        i = line * width + col;
    }
    uint32_t w[EMIT / 4] = {0, 0, 0, 0};
#pragma unroll
    for (int t = 0; t < EMIT; ++t) {
        uint32_t v = 0;
        if (q0 + t < total) {
            if (r == blen - 1 || (width > 0 && col == width)) {
                v = '\n';
            } else {
                v = seq[base + i];
                if (v - 'a' < 26u) v -= 32u;  // upper-cased for output
                if (mask[base + i]) v = soft ? (v - 'A' < 26u ? v + 32u : v) : alt;
                ++i;
            }
            if (width > 0) col = (col == width) ? 0 : col + 1;
            if (++r == blen && q0 + t + 1 < total) {
                ++s;
                r = 0;
                col = 0;
                i = 0;
                blen = body_off[s + 1] - body_off[s];
                base = seg_off[s];
            }
        }
        w[t >> 2] |= (v & 0xffu) << (8 * (t & 3));
    }
    *(uint4 *)(out + q0) = make_uint4(w[0], w[1], w[2], w[3]);  // the buffer is padded to EMIT
}

}  // namespace

int64_t kmer_table_buckets(int64_t nkeys) {
    int64_t slots = 1024;
    while (slots < 2 * nkeys) slots *= 2;
    return slots / BUCKET;
}

int64_t kmer_general_slots(int64_t n) {
    int64_t slots = 1024;
    while (slots < 2 * n) slots *= 2;
    return slots;
}

int64_t kmer_mask_padded(int64_t n) { return (n + CHUNK - 1) / CHUNK * CHUNK + CHUNK; }
int64_t kmer_body_padded(int64_t n) { return (n + EMIT - 1) / EMIT * EMIT + EMIT; }

hipError_t launch_kmer_table(hipStream_t st, const uint64_t *codes, int64_t n, uint64_t *table, int64_t buckets) {
    hipError_t e = hipMemsetAsync(table, 0xff, (size_t)buckets * BUCKET * sizeof(uint64_t), st);
    if (e != hipSuccess || n == 0) return e;
    hipLaunchKernelGGL(k_kmer_insert, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, codes, n, table,
                       (uint64_t)(buckets - 1));
    return hipGetLastError();
}

hipError_t launch_kmer_general(hipStream_t st, const uint64_t *hashes, int64_t n, uint32_t *gtab, int64_t slots) {
    hipError_t e = hipMemsetAsync(gtab, 0, (size_t)slots * sizeof(uint32_t), st);
    if (e != hipSuccess || n == 0) return e;
    hipLaunchKernelGGL(k_kmer_insert_general, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, hashes, n, gtab,
                       (uint64_t)(slots - 1));
    return hipGetLastError();
}

hipError_t launch_kmer_scan(hipStream_t st, const uint8_t *seq, int64_t N, const int64_t *seg_off, int32_t nseg,
                            int32_t k, bool codes, const uint64_t *table, int64_t buckets, const KmerGeneral &g,
                            uint8_t *mask) {
    if (N <= 0) return hipSuccess;
    const int64_t lanes = (N + CHUNK - 1) / CHUNK;
    const dim3 grid((unsigned)((lanes + 255) / 256)), block(256);
    const uint64_t bmask = buckets > 0 ? (uint64_t)(buckets - 1) : 0;
    const bool gen = g.n > 0;
    if (codes && gen)
        hipLaunchKernelGGL((k_kmer_scan<true, true>), grid, block, 0, st, seq, N, seg_off, nseg, k, table, bmask, g,
                           mask);
    else if (codes)
        hipLaunchKernelGGL((k_kmer_scan<true, false>), grid, block, 0, st, seq, N, seg_off, nseg, k, table, bmask, g,
                           mask);
    else if (gen)
        hipLaunchKernelGGL((k_kmer_scan<false, true>), grid, block, 0, st, seq, N, seg_off, nseg, k, table, bmask, g,
                           mask);
    else
        return hipMemsetAsync(mask, 0, (size_t)N, st);
    return hipGetLastError();
}

hipError_t launch_paint_ranges(hipStream_t st, const int64_t *ranges, int64_t n, uint8_t *mask) {
    if (n <= 0) return hipSuccess;
    const unsigned blocks = (unsigned)(n < 65536 ? n : 65536);
    hipLaunchKernelGGL(k_paint_ranges, dim3(blocks), dim3(256), 0, st, ranges, n, mask);
    return hipGetLastError();
}

hipError_t launch_mask_emit(hipStream_t st, const uint8_t *seq, const uint8_t *mask, const int64_t *seg_off,
                            const int64_t *body_off, int32_t nseg, int64_t total, int32_t width, uint8_t alt,
                            bool soft, uint8_t *out) {
    if (total <= 0) return hipSuccess;
    const int64_t lanes = (total + EMIT - 1) / EMIT;
    hipLaunchKernelGGL(k_mask_emit, dim3((unsigned)((lanes + 255) / 256)), dim3(256), 0, st, seq, mask, seg_off,
                       body_off, nseg, total, width, (uint32_t)alt, soft ? 1 : 0, out);
    return hipGetLastError();
}

}  // namespace dcns
