// mask.hip -- maskGenome on the device (impl/maskGenome.cpp:50-121): the
// k-mer scan that marks every position covered by a window [i, i + k) whose
// text is one of a set of k-mers, the CDS ranges painted over that mask, and
// the FASTA bodies written from sequence and mask.
//
// The genome is one byte buffer of concatenated sequences with offsets
// seg_off[0..nseg] (seg_off[nseg] = N).  Windows never cross a sequence end.
//
// K-mer set, two parts:
//   * members made of upper-case A/C/G/T only, k <= 31: 2-bit codes (base j
//     in bits 2j..2j+1 as (b >> 1) & 3: A 0, C 1, T 2, G 3) in minimizer
//     buckets of one 64-byte slot each (kmer_hash.h), keys past the seventh
//     of a bucket in an overflow array;
//   * every other member (other bytes, or k >= 32): a table of member
//     indices keyed by a 64-bit polynomial hash of the bytes, each hit
//     verified byte by byte.
// A window of upper-case A/C/G/T can only equal a member of the first part;
// any other window only one of the second.  Lower-case windows never match,
// as the reference compares the text as read against upper-cased members.
//
// k <= 31, k_kmer_scan16: a thread owns 16 window starts / 16 positions.  It
// reads its 48 bytes with three aligned 16-byte loads (consecutive threads,
// consecutive blocks), packs them into 96 bits of 2-bit codes and a 48-bit
// mask of other bytes with SWAR, and takes each window's code and l-mers by
// constant shifts.  A window whose bucket differs from the previous window's
// loads the bucket's slot (four 16-byte loads, one round trip), at 8 waves
// per SIMD to keep many slot loads in flight;
// the windows of a genome share their minimizer over ~(w + 1) / 2
// consecutive starts, so one line serves several windows.  The 16 hit bits go through LDS to the next
// two threads; a position is masked when one of the k windows ending on it
// hit, which is a shift-OR of the 48 hit bits of the thread and its two
// predecessors.  The first two threads of a workgroup recompute the previous
// workgroup's last 32 windows.  HBM traffic per position: 1 byte read, 1
// byte written, plus the bucket reads.
//
// k >= 32, k_kmer_scan_long: one lane per CHUNK positions keeps a rolling
// hash of the last k bytes and the start of the last hit window.
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "kmer_hash.h"

namespace dcns {
namespace {

constexpr uint64_t kEmpty = ~0ull;
constexpr int WG = 256;                       // threads per workgroup of the k <= 31 scan
constexpr int HALO = 2;                       // threads recomputing the previous workgroup's windows
constexpr int SPAN = 16;                      // windows and positions per thread
constexpr int WG_POS = (WG - HALO) * SPAN;    // positions a workgroup writes
constexpr int CHUNK = 1024;                   // positions per lane of the k >= 32 scan
constexpr int G = 8;                          // windows per probe group
constexpr int EMIT = 16;                      // output bytes per lane in the FASTA writer

__device__ __forceinline__ uint64_t mix(uint64_t x) {
    // splitmix64 finaliser (a bijection on 64 bits)
    x ^= x >> 30;
    x *= 0xbf58476d1ce4e5b9ull;
    x ^= x >> 27;
    x *= 0x94d049bb133111ebull;
    x ^= x >> 31;
    return x;
}

// 0x80 in every byte lane of w that equals the byte of c4 (c4 = byte x 0x01010101), else 0
__device__ __forceinline__ uint32_t eq_lanes(uint32_t w, uint32_t c4) {
    const uint32_t t = w ^ c4;
    return ~(((t & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | t | 0x7F7F7F7Fu);
}

// bit b set for each byte b of w other than upper-case A, C, G, T
__device__ __forceinline__ uint32_t other_bytes4(uint32_t w) {
    const uint32_t ok = eq_lanes(w, 0x41414141u) | eq_lanes(w, 0x43434343u) | eq_lanes(w, 0x47474747u) |
                        eq_lanes(w, 0x54545454u);
    const uint32_t bad = ~ok & 0x80808080u;
    return ((bad >> 7) & 1u) | ((bad >> 14) & 2u) | ((bad >> 21) & 4u) | ((bad >> 28) & 8u);
}

// (b >> 1) & 3 of the 4 bytes of w, byte j in bits 2j..2j+1
__device__ __forceinline__ uint32_t codes4(uint32_t w) {
    uint32_t v = (w >> 1) & 0x03030303u;
    v = (v | (v >> 6)) & 0x000F000Fu;
    return (v | (v >> 12)) & 0xFFu;
}

// the sequence holding position pos: the last s in [lo, hi] with seg_off[s] <= pos
__device__ __forceinline__ int32_t seg_of(const int64_t *__restrict__ seg_off, int32_t lo, int32_t hi, int64_t pos) {
    while (lo < hi) {
        const int32_t mid = (lo + hi + 1) >> 1;
        if (seg_off[mid] <= pos) lo = mid;
        else hi = mid - 1;
    }
    return lo;
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

__device__ __noinline__ bool general_window(const KmerGeneral &g, const uint8_t *__restrict__ win, int32_t k) {
    uint64_t h = 0;
    for (int32_t j = 0; j < k; ++j) h = h * kKmerHashBase + win[j];
    return general_lookup(g, win, h, k);
}

// keys of a bucket past its slot (sorted)
__device__ __noinline__ bool overflow_lookup(const uint64_t *__restrict__ keys, uint32_t n, uint64_t key) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        const uint64_t v = keys[mid];
        if (v == key) return true;
        if (v < key) lo = mid + 1;
        else hi = mid;
    }
    return false;
}

template <bool GEN>
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(8))) void k_kmer_scan16(const uint8_t *__restrict__ seq, int64_t N,
                                                    const int64_t *__restrict__ seg_off, int32_t nseg, int32_t k,
                                                    const uint64_t *__restrict__ slots,
                                                    const uint64_t *__restrict__ overflow, uint32_t nb_mask,
                                                    KmerGeneral g, uint8_t *__restrict__ mask) {
    __shared__ uint32_t lds_hits[WG];
    __shared__ int32_t lds_seg[2];
    const int tid = threadIdx.x;
    const int64_t w0 = (int64_t)blockIdx.x * WG_POS - HALO * SPAN;  // first window of thread 0
    const int64_t i0 = w0 + (int64_t)tid * SPAN;                     // this thread's windows [i0, i0 + 16)
    if (tid == 0) lds_seg[0] = seg_of(seg_off, 0, nseg - 1, w0 > 0 ? w0 : 0);
    if (tid == WG - 1) {
        const int64_t last = w0 + WG * SPAN + 48;
        lds_seg[1] = seg_of(seg_off, 0, nseg - 1, last < N ? last : N - 1);
    }
    __syncthreads();

    uint32_t hits = 0;
    if (i0 >= 0 && i0 < N) {
        const uint4 *src = (const uint4 *)(seq + i0);  // 16-byte aligned; the buffer holds N + 128 bytes
        const uint4 b0 = src[0], b1 = src[1], b2 = src[2];
        const uint32_t wd[12] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w, b2.x, b2.y, b2.z, b2.w};
        uint64_t lo = 0, bad = 0;
        uint32_t hi = 0;
#pragma unroll
        for (int q = 0; q < 12; ++q) {
            const uint32_t c = codes4(wd[q]);
            if (q < 8) lo |= (uint64_t)c << (8 * q);
            else hi |= c << (8 * (q - 8));
            bad |= (uint64_t)other_bytes4(wd[q]) << (4 * q);
        }
        if (N - i0 < 48) bad |= ~0ull << (N - i0);  // past the genome
        // sequence starts inside (i0, i0 + 48)
        uint64_t bnd = 0;
        int32_t s = seg_of(seg_off, lds_seg[0], lds_seg[1], i0);
        while (s + 1 < nseg && seg_off[s + 1] < i0 + 48) {
            ++s;
            bnd |= 1ull << (seg_off[s] - i0);
        }
        const uint64_t kbits = (1ull << k) - 1, kin = (1ull << (k - 1)) - 1;
        const uint64_t cmask = (1ull << (2 * k)) - 1;
        const int w = minimizer_w(k), l = k - w + 1;
        const uint64_t lmask = (1ull << (2 * l)) - 1;
        // l-mer hashes at offsets 0 .. 15 + w - 1, then each window's bucket
        uint32_t lh[SPAN + kMinimizerW - 1];
#pragma unroll
        for (int p = 0; p < SPAN + kMinimizerW - 1; ++p)
            lh[p] = lmer_hash((p == 0 ? lo : ((lo >> (2 * p)) | ((uint64_t)hi << (64 - 2 * p)))) & lmask);
        // a window's code, taken again where it is compared (keeping all 16 would cost 32 VGPRs)
        auto key_of = [&](int t) {
            return (t == 0 ? lo : ((lo >> (2 * t)) | ((uint64_t)hi << (64 - 2 * t)))) & cmask;
        };
        uint32_t bk[SPAN];
        uint32_t valid = 0;
#pragma unroll
        for (int t = 0; t < SPAN; ++t) {
            uint32_t mn = lh[t];
#pragma unroll
            for (int q = 1; q < kMinimizerW; ++q)
                if (q < w) mn = min(mn, lh[t + q]);
            bk[t] = minimizer_bucket(mn, nb_mask);
            if (((bad >> t) & kbits) == 0 && ((bnd >> (t + 1)) & kin) == 0) valid |= 1u << t;
        }
        // a window whose bucket differs from the one held loads its slot
        ulonglong2 sv[kSlotWords / 2];
        uint32_t held = 0;
        bool have = false;
#pragma unroll
        for (int t = 0; t < SPAN; ++t) {
            if ((valid >> t) & 1u) {
                if (!have || bk[t] != held) {
                    const ulonglong2 *sl = (const ulonglong2 *)(slots + (uint64_t)bk[t] * kSlotWords);
#pragma unroll
                    for (int q = 0; q < kSlotWords / 2; ++q) sv[q] = sl[q];
                    held = bk[t];
                    have = true;
                }
                const uint64_t kt = key_of(t);
                bool m = sv[0].y == kt;
#pragma unroll
                for (int q = 1; q < kSlotWords / 2; ++q) m |= sv[q].x == kt || sv[q].y == kt;
                const uint32_t cnt = (uint32_t)sv[0].x;
                if (!m && cnt > (uint32_t)kSlotKeys)
                    m = overflow_lookup(overflow + (sv[0].x >> 32), cnt - kSlotKeys, kt);
                hits |= (m ? 1u : 0u) << t;
            }
        }
        if (GEN) {
#pragma unroll
            for (int t = 0; t < SPAN; ++t)
                if (!((valid >> t) & 1u) && ((bnd >> (t + 1)) & kin) == 0 && i0 + t + k <= N &&
                    general_window(g, seq + i0 + t, k))
                    hits |= 1u << t;
        }
    }
    lds_hits[tid] = hits;
    __syncthreads();
    if (tid < HALO || i0 >= N) return;
    // windows i0 - 32 .. i0 + 15 in bits 0..47; position i0 + u is masked when a
    // window starting in [i0 + u - k + 1, i0 + u] hit
    uint64_t r = (uint64_t)lds_hits[tid - 2] | ((uint64_t)lds_hits[tid - 1] << 16) | ((uint64_t)hits << 32);
    for (int32_t c = 1; c < k;) {
        const int32_t d = c < k - c ? c : k - c;
        r |= r << d;
        c += d;
    }
    const uint32_t m = (uint32_t)(r >> 32) & 0xFFFFu;
    uint32_t out[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const uint32_t b = m >> (4 * q);
        out[q] = (b & 1u) | ((b & 2u) << 7) | ((b & 4u) << 14) | ((b & 8u) << 21);
    }
    *(uint4 *)(mask + i0) = make_uint4(out[0], out[1], out[2], out[3]);  // the buffer is padded past N
}

// k >= 32: every member is in the general table
__global__ __launch_bounds__(256) void k_kmer_scan_long(const uint8_t *__restrict__ seq, int64_t N,
                                                        const int64_t *__restrict__ seg_off, int32_t nseg, int32_t k,
                                                        KmerGeneral g, uint8_t *__restrict__ mask) {
    const int64_t c0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * CHUNK;
    if (c0 >= N) return;
    const int64_t c1 = c0 + CHUNK < N ? c0 + CHUNK : N;
    const int32_t back = (k - 1 + G - 1) / G * G;
    const int64_t ib = c0 - back > 0 ? c0 - back : 0;  // first window evaluated (a multiple of G)
    int32_t s = seg_of(seg_off, 0, nseg - 1, ib);
    int64_t seg_from = seg_off[s], seg_next = seg_off[s + 1];
    uint64_t h = 0;
    auto feed = [&](int64_t p) {
        while (p >= seg_next) {
            ++s;
            seg_from = seg_off[s];
            seg_next = seg_off[s + 1];
        }
        const uint64_t out = (p - k >= ib) ? (uint64_t)seq[p - k] : 0ull;
        h = (h - out * g.pk1) * kKmerHashBase + (uint64_t)seq[p];
    };
    for (int64_t p = ib; p < ib + k - 1 && p < N; ++p) feed(p);
    int64_t last_hit = INT64_MIN / 2;
    for (int64_t i0 = ib; i0 < c1; i0 += G) {
        uint32_t lo4 = 0, hi4 = 0;
#pragma unroll
        for (int t = 0; t < G; ++t) {
            const int64_t i = i0 + t, p = i + k - 1;
            if (p < N) {
                feed(p);
                if (i >= seg_from && general_lookup(g, seq + i, h, k)) last_hit = i;
            }
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

// FASTA bodies, one workgroup per EMIT_WG * EMIT input bases: every output
// byte that is not a base is a newline, so the workgroup's output range is
// filled with newlines in LDS, each thread scatters its 16 bases (read with
// one 16-byte load of sequence and of mask) to their output offsets, and the
// range is written with aligned 16-byte stores.  Needs no empty sequence.
constexpr int EMIT_WG = 256;
constexpr int EMIT_TILE = 3 * EMIT_WG * EMIT + 64;  // width 1: a base, its line end, a sequence end

__global__ __launch_bounds__(EMIT_WG) void k_mask_emit(const uint8_t *__restrict__ seq,
                                                       const uint8_t *__restrict__ mask,
                                                       const int64_t *__restrict__ seg_off,
                                                       const int64_t *__restrict__ body_off, int32_t nseg, int64_t N,
                                                       int64_t total, int32_t width, uint32_t alt, int32_t soft,
                                                       uint8_t *__restrict__ out) {
    __shared__ uint4 tile4[EMIT_TILE / 16];
    __shared__ int64_t lds_range[2];
    __shared__ int32_t lds_seg[2];
    uint8_t *tile = (uint8_t *)tile4;
    const int tid = threadIdx.x;
    const int64_t in0 = (int64_t)blockIdx.x * EMIT_WG * EMIT;
    const int64_t in1 = in0 + EMIT_WG * EMIT < N ? in0 + EMIT_WG * EMIT : N;
    auto out_pos = [&](int32_t s, int64_t i) {  // output offset of base i of sequence s
        const int64_t r = i - seg_off[s];
        return body_off[s] + r + (width > 0 ? r / width : 0);
    };
    if (tid == 0) {
        const int32_t s0 = seg_of(seg_off, 0, nseg - 1, in0);
        lds_seg[0] = s0;
        lds_range[0] = out_pos(s0, in0);
    } else if (tid == EMIT_WG - 1) {
        const int32_t s1 = seg_of(seg_off, 0, nseg - 1, in1 - 1);
        lds_seg[1] = s1;
        lds_range[1] = in1 < N ? out_pos(seg_of(seg_off, 0, nseg - 1, in1), in1) : total;
    }
    for (int q = tid; q < EMIT_TILE / 16; q += EMIT_WG) tile4[q] = make_uint4(0x0A0A0A0Au, 0x0A0A0A0Au, 0x0A0A0A0Au,
                                                                              0x0A0A0A0Au);
    __syncthreads();
    const int64_t o0 = lds_range[0], o1 = lds_range[1];
    const int64_t a = in0 + (int64_t)tid * EMIT;
    if (a < in1) {
        const uint4 sv = *(const uint4 *)(seq + a), mv = *(const uint4 *)(mask + a);  // buffers padded to 16
        const uint32_t sw[4] = {sv.x, sv.y, sv.z, sv.w}, mw[4] = {mv.x, mv.y, mv.z, mv.w};
        int32_t s = seg_of(seg_off, lds_seg[0], lds_seg[1], a);
        int64_t next = seg_off[s + 1];
        int64_t r = a - seg_off[s];
        int64_t o = body_off[s] + r;
        int64_t col = 0;
        if (width > 0) {
            o += r / width;
            col = r % width;
        }
#pragma unroll
        for (int t = 0; t < EMIT; ++t) {
            const int64_t i = a + t;
            if (i < in1) {
                if (i == next) {  // the first base of the next sequence (none is empty)
                    ++s;
                    next = seg_off[s + 1];
                    o = body_off[s];
                    col = 0;
                }
                uint32_t v = (sw[t >> 2] >> (8 * (t & 3))) & 0xFFu;
                if (v - 'a' < 26u) v -= 32u;  // upper-cased for output
                if ((mw[t >> 2] >> (8 * (t & 3))) & 0xFFu) v = soft ? (v - 'A' < 26u ? v + 32u : v) : alt;
                tile[o - o0] = (uint8_t)v;
                ++o;
                if (width > 0 && ++col == width) {
                    col = 0;
                    ++o;  // the line end
                }
            }
        }
    }
    __syncthreads();
    // [o0, o1): aligned 16-byte blocks inside, single bytes at the two ends
    const int64_t a16 = (o0 + 15) & ~(int64_t)15, b16 = o1 & ~(int64_t)15;
    if (a16 < b16) {
        for (int64_t q = a16 + (int64_t)tid * 16; q < b16; q += EMIT_WG * 16) {
            const int64_t l = q - o0;
            uint32_t wv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u)
                wv[u] = (uint32_t)tile[l + 4 * u] | ((uint32_t)tile[l + 4 * u + 1] << 8) |
                        ((uint32_t)tile[l + 4 * u + 2] << 16) | ((uint32_t)tile[l + 4 * u + 3] << 24);
            *(uint4 *)(out + q) = make_uint4(wv[0], wv[1], wv[2], wv[3]);
        }
        if (tid < a16 - o0) out[o0 + tid] = tile[tid];
        if (tid < o1 - b16) out[b16 + tid] = tile[b16 - o0 + tid];
    } else {
        for (int64_t q = o0 + tid; q < o1; q += EMIT_WG) out[q] = tile[q - o0];
    }
}

__global__ __launch_bounds__(256) void k_mask_emit_bytes(const uint8_t *__restrict__ seq, const uint8_t *__restrict__ mask,
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

int64_t kmer_general_slots(int64_t n) {
    int64_t slots = 1024;
    while (slots < 2 * n) slots *= 2;
    return slots;
}

int64_t kmer_mask_padded(int64_t n) { return (n + CHUNK - 1) / CHUNK * CHUNK + CHUNK; }
int64_t kmer_body_padded(int64_t n) { return (n + EMIT - 1) / EMIT * EMIT + EMIT; }

hipError_t launch_kmer_general(hipStream_t st, const uint64_t *hashes, int64_t n, uint32_t *gtab, int64_t slots) {
    hipError_t e = hipMemsetAsync(gtab, 0, (size_t)slots * sizeof(uint32_t), st);
    if (e != hipSuccess || n == 0) return e;
    hipLaunchKernelGGL(k_kmer_insert_general, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, hashes, n, gtab,
                       (uint64_t)(slots - 1));
    return hipGetLastError();
}

hipError_t launch_kmer_scan(hipStream_t st, const uint8_t *seq, int64_t N, const int64_t *seg_off, int32_t nseg,
                            int32_t k, const uint64_t *slots, const uint64_t *overflow, int64_t buckets,
                            const KmerGeneral &g, uint8_t *mask) {
    if (N <= 0) return hipSuccess;
    const bool gen = g.n > 0;
    if (k >= 1 && k <= 31 && buckets > 0) {  // buckets: the bucket table exists (it may hold no key)
        const dim3 grid((unsigned)((N + WG_POS - 1) / WG_POS)), block(WG);
        const uint32_t nb_mask = (uint32_t)(buckets - 1);
        if (gen)
            hipLaunchKernelGGL(k_kmer_scan16<true>, grid, block, 0, st, seq, N, seg_off, nseg, k, slots, overflow,
                               nb_mask, g, mask);
        else
            hipLaunchKernelGGL(k_kmer_scan16<false>, grid, block, 0, st, seq, N, seg_off, nseg, k, slots, overflow,
                               nb_mask, g, mask);
    } else if (k >= 32 && gen) {
        const int64_t lanes = (N + CHUNK - 1) / CHUNK;
        hipLaunchKernelGGL(k_kmer_scan_long, dim3((unsigned)((lanes + 255) / 256)), dim3(256), 0, st, seq, N,
                           seg_off, nseg, k, g, mask);
    } else {
        return hipMemsetAsync(mask, 0, (size_t)N, st);
    }
    return hipGetLastError();
}

hipError_t launch_paint_ranges(hipStream_t st, const int64_t *ranges, int64_t n, uint8_t *mask) {
    if (n <= 0) return hipSuccess;
    const unsigned blocks = (unsigned)(n < 65536 ? n : 65536);
    hipLaunchKernelGGL(k_paint_ranges, dim3(blocks), dim3(256), 0, st, ranges, n, mask);
    return hipGetLastError();
}

hipError_t launch_mask_emit(hipStream_t st, const uint8_t *seq, const uint8_t *mask, const int64_t *seg_off,
                            const int64_t *body_off, int32_t nseg, int64_t N, int64_t total, int32_t width,
                            bool any_empty, uint8_t alt, bool soft, uint8_t *out) {
    if (total <= 0) return hipSuccess;
    if (!any_empty && N > 0) {
        const int64_t wgs = (N + EMIT_WG * EMIT - 1) / (EMIT_WG * EMIT);
        hipLaunchKernelGGL(k_mask_emit, dim3((unsigned)wgs), dim3(EMIT_WG), 0, st, seq, mask, seg_off, body_off, nseg,
                           N, total, width, (uint32_t)alt, soft ? 1 : 0, out);
    } else {  // empty sequences (k = 0 only): one lane per 16 output bytes
        const int64_t lanes = (total + EMIT - 1) / EMIT;
        hipLaunchKernelGGL(k_mask_emit_bytes, dim3((unsigned)((lanes + 255) / 256)), dim3(256), 0, st, seq, mask,
                           seg_off, body_off, nseg, total, width, (uint32_t)alt, soft ? 1 : 0, out);
    }
    return hipGetLastError();
}

}  // namespace dcns
