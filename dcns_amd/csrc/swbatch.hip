// swbatch.hip -- SmithWaterman max score and first position for many
// independent pairs of arbitrary length (the inner call of ranSco,
// permutationLsqLambdaK.cpp:97, and any direct SmithWaterman caller).
//
// Semantics follow SmithWaterman(...bool& positions)
// (sequenceAlignment3.cpp:220-469): maximum local score and its first strict
// occurrence in (ref row, read column) order.  The reference switches between
// a striped recurrence and the textbook one by a bound test; both give the
// same H matrix, so the textbook form (one shared H + open term) is used.
//
// One lane per pair, int32.  The read is cut into 32-column stripes held in
// VGPRs; each stripe sweeps all reference rows and hands its last column
// (H and the clamped horizontal gap) to the next stripe through a per-pair
// boundary array (8 bytes per reference row).  Rows are taken 16 at a time:
// one 16-byte load brings their reference codes and eight 16-byte loads /
// stores move their boundary values, so a lane's memory traffic is a few wide
// accesses per 512 cells.  Pairs are processed in the host's cost order so
// that lanes of one wave run similar loop counts.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "dp_common.h"
#include "kernels.h"

namespace dcns {
namespace {

constexpr int S = 32;   // query columns per stripe (VGPRs)
constexpr int RB = 16;  // reference rows per block: one 16-byte load of codes and
                        // 8 x 16-byte loads / stores of the stripe boundary

// Sequences start at 16-byte aligned offsets with at least 16 readable bytes
// past their end; the boundary array holds one int2 (H, clamped horizontal
// gap) per reference row, 16-byte aligned, rows rounded up to RB.
// Codes from memory: reference rows 16 at a time, query stripes 32 at a time.
// Rows at or past the lane's own length are not read (the wave runs its
// widest lane's rectangle; those rows and columns are N).
struct MemCodes {
    const int8_t *ref, *qry;
    int32_t L1, L2;
    __device__ __forceinline__ void rows16(int32_t ib, uint32_t w[4]) const {
        if (ib < L1) {
            const uint4 v = *(const uint4 *)(ref + ib);
            w[0] = v.x; w[1] = v.y; w[2] = v.z; w[3] = v.w;
        } else {
            w[0] = w[1] = w[2] = w[3] = 0x02020202u;
        }
    }
    __device__ __forceinline__ void cols32(int32_t j0, uint32_t w[8]) const {
        if (j0 < L2) {
            const uint4 a = *(const uint4 *)(qry + j0), b = *(const uint4 *)(qry + j0 + 16);
            w[0] = a.x; w[1] = a.y; w[2] = a.z; w[3] = a.w; w[4] = b.x; w[5] = b.y; w[6] = b.z; w[7] = b.w;
        } else {
#pragma unroll
            for (int q = 0; q < 8; ++q) w[q] = 0x02020202u;
        }
    }
};

// i.i.d. bases of composition gc (P(C) = P(G) = gc/2, P(A) = P(T) = (1-gc)/2),
// base k of a sequence = threshold of 32 bits of mix64(key + k/2).
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
struct RandCodes {
    uint64_t kref, kqry;
    uint32_t tC, tG, tA;  // cumulative thresholds of C, G, A (T above)
    __device__ __forceinline__ uint32_t code(uint32_t u) const {
        return u < tC ? 1u : u < tG ? 3u : u < tA ? 0u : 4u;
    }
    __device__ __forceinline__ void fill(uint64_t key, int32_t k0, uint32_t *w, int nw) const {
        for (int q = 0; q < nw; ++q) {
            uint32_t v = 0;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const uint64_t x = mix64(key + (uint64_t)((k0 >> 1) + 2 * q + h));
                v |= code((uint32_t)x) << (16 * h);
                v |= code((uint32_t)(x >> 32)) << (16 * h + 8);
            }
            w[q] = v;
        }
    }
    __device__ __forceinline__ void rows16(int32_t ib, uint32_t w[4]) const { fill(kref, ib, w, 4); }
    __device__ __forceinline__ void cols32(int32_t j0, uint32_t w[8]) const { fill(kqry, j0, w, 8); }
};

// Largest value of v over the wave's active lanes, wave-uniform.
__device__ __forceinline__ int32_t wave_max(int32_t v) {
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) v = max(v, __shfl_xor(v, d, 64));
    return __builtin_amdgcn_readfirstlane(v);
}

// The lanes of a wave run the same loops: the widest lane's stripes and rows
// (scalar loop control); a lane's rows and columns past its own lengths are
// N, which scores 0 against everything, so no cell there exceeds the best of
// the cells before it and the maximum and its first position are unchanged.
// The boundary slot of every lane must hold the wave's rows; a lane without
// a pair (bnd null, lengths 0) takes part in the loops but touches no memory.
// NN: an N row against an N column scores -1 instead of 0 (the score matrix
// of SmithWaterman_bas, impl/SmithWaterman_bas.cpp:13-20); padding rows and
// columns past the sequences keep 0.
template <class Codes, bool NN = false>
__device__ void sw_pair(const Codes &src, int32_t L1, int32_t L2, const Params &P, const uint32_t *shi,
                        const uint32_t *slo, int2 *__restrict__ bnd, int32_t &best, int32_t &bi, int32_t &bj) {
    best = 0;
    bi = -1;
    bj = -1;
    const int32_t O = P.open, X = P.ext;
    const int32_t L1w = wave_max(L1);
    const int nstripes = wave_max((L2 + S - 1) / S);
    for (int s = 0; s < nstripes; ++s) {
        const int j0 = s * S;
        uint32_t sel[S];
        int32_t H[S], E[S];
        uint32_t qn = 0;  // NN: bit j = column j0 + j is a real N
        {
            uint32_t qw[8];
            src.cols32(j0, qw);
#pragma unroll
            for (int j = 0; j < S; ++j) {
                const int code = (int)(int8_t)((qw[j >> 2] >> (8 * (j & 3))) & 0xffu);
                sel[j] = sel32(j0 + j < L2 ? code : 2);
                if (NN && j0 + j < L2 && code == 2) qn |= 1u << j;
                H[j] = 0;
                E[j] = 0;
            }
        }
        const bool has_left = s > 0, has_right = s + 1 < nstripes;
        int32_t sbest = 0, si = -1, sj = -1;
        int32_t hb_prev = 0;  // boundary H of the previous row (diagonal of column j0)
        for (int32_t ib = 0; ib < L1w; ib += RB) {
            uint32_t rw[4];
            src.rows16(ib, rw);
            int2 b[RB];
            if (has_left && bnd) {
#pragma unroll
                for (int q = 0; q < RB / 2; ++q) {
                    const int4 v = *(const int4 *)(bnd + ib + 2 * q);
                    b[2 * q] = make_int2(v.x, v.y);
                    b[2 * q + 1] = make_int2(v.z, v.w);
                }
            } else {
#pragma unroll
                for (int r = 0; r < RB; ++r) b[r] = make_int2(0, 0);
            }
            const int rows = min(RB, L1w - ib);  // wave-uniform
#pragma unroll
            for (int r = 0; r < RB; ++r) {
                if (r < rows) {
                    const int c = ib + r < L1 ? (int)(int8_t)((rw[r >> 2] >> (8 * (r & 3))) & 0xffu) : 2;
                    const int ci = (unsigned)c < 8u ? c : 2;
                    const uint32_t thi = shi[ci], tlo = slo[ci];
                    const uint32_t rn = (NN && ib + r < L1 && c == 2) ? qn : 0u;  // N row: its N columns score -1
                    int32_t left = b[r].x, fc = b[r].y;
                    int32_t diag = hb_prev;
                    hb_prev = left;
                    int32_t rmax = 0;
#pragma unroll
                    for (int j = 0; j < S; ++j) {
                        int32_t sc = (int32_t)perm(thi, tlo, sel[j]);
                        if (NN) sc -= (int32_t)((rn >> j) & 1u);
                        const int32_t t = diag + sc;
                        diag = H[j];
                        // fc here is the clamped horizontal gap entering column j
                        const int32_t h = max3i(t, E[j], fc);
                        const int32_t ho = h + O;
                        E[j] = max3i(E[j] + X, ho, 0);
                        fc = max3i(ho, fc + X, 0);
                        H[j] = h;
                        left = h;
                        if (j & 1) rmax = max3i(rmax, H[j - 1], h);
                    }
                    b[r] = make_int2(left, fc);
                    if (rmax > sbest) {
                        int jj = 0;
#pragma unroll
                        for (int j = S - 1; j >= 0; --j)
                            if (H[j] == rmax) jj = j;
                        sbest = rmax;
                        si = ib + r;
                        sj = j0 + jj;
                    }
                }
            }
            if (has_right && bnd) {
#pragma unroll
                for (int q = 0; q < RB / 2; ++q)
                    *(int4 *)(bnd + ib + 2 * q) = make_int4(b[2 * q].x, b[2 * q].y, b[2 * q + 1].x, b[2 * q + 1].y);
            }
        }
        if (sbest > best || (sbest == best && sbest > 0 && (si < bi || (si == bi && sj < bj)))) {
            best = sbest;
            bi = si;
            bj = sj;
        }
    }
}

// Pairs order[gid] (or first + gid without an order) for gid < n.  With
// max_rows > 0, a pair of more reference rows than its boundary slot holds
// is skipped and flagged in *err (bit 0), its outputs left as they are.
// A lane's boundary rows: with per-wave offsets (wave_off[w] .. wave_off[w+1]
// int32 of wave w = blockIdx.x, relative to the launch, split evenly among its
// 64 lanes), else gid * stride.
__device__ __forceinline__ int2 *boundary_slot(int32_t *bound, int64_t stride, const int64_t *wave_off, int64_t gid) {
    if (!wave_off) return (int2 *)(bound + gid * stride);
    const int64_t a = wave_off[blockIdx.x], per_lane = (wave_off[blockIdx.x + 1] - a) / 64;
    return (int2 *)(bound + a + (int64_t)threadIdx.x * per_lane);
}

template <bool NN>
__global__ __launch_bounds__(64) void k_sw_batch(const int8_t *__restrict__ seq, const int64_t *__restrict__ off1,
                                                 const int32_t *__restrict__ len1, const int64_t *__restrict__ off2,
                                                 const int32_t *__restrict__ len2, const int32_t *__restrict__ order,
                                                 int32_t first, int32_t n, Params P, ScoreTab tab,
                                                 int32_t *__restrict__ out, int32_t *__restrict__ bound,
                                                 int64_t bound_stride, const int64_t *__restrict__ wave_off,
                                                 int32_t max_rows, int32_t *__restrict__ err) {
    __shared__ uint32_t shi[8], slo[8];
    if (threadIdx.x < 8) {
        shi[threadIdx.x] = tab.hi[threadIdx.x];
        slo[threadIdx.x] = tab.lo[threadIdx.x];
    }
    __syncthreads();
    // every lane of the wave stays for the wave-uniform loops of sw_pair
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    bool ok = gid < n;
    int p = 0;
    int32_t L1 = 0, L2 = 0;
    if (ok) {
        p = order ? order[gid] : first + gid;
        L1 = len1[p];
        L2 = len2[p];
        if (max_rows > 0 && (L1 > max_rows || L1 < 0 || L2 < 0)) {
            atomicOr(err, 1);
            ok = false;
            L1 = L2 = 0;
        }
    }
    int2 *bnd = ok ? boundary_slot(bound, bound_stride, wave_off, gid) : nullptr;
    int32_t best, bi, bj;
    sw_pair<MemCodes, NN>(MemCodes{ok ? seq + off1[p] : seq, ok ? seq + off2[p] : seq, L1, L2}, L1, L2, P, shi, slo,
                          bnd, best, bi, bj);
    if (!ok) return;
    out[3 * p + 0] = best;
    out[3 * p + 1] = bi;
    out[3 * p + 2] = bj;
}

// Max scores of random fragment pairs generated on the device: pair r of
// point k has lengths len1[k] x len2[k] and base thresholds thr[3k..3k+2]
// (C, G, A cumulative, out of 2^32); its keys are mix64(seed[k] + 2r)
// (reference) and mix64(seed[k] + 2r + 1) (query).
// The lane slots take the pairs in the host's plan: order[slot] is the pair
// (slot_order) or order[slot / per_point] the point (pairs of a point
// adjacent), so that the lanes of a wave share their loop shape and the
// launch's waves come longest first.
__global__ __launch_bounds__(64) void k_sw_random(const uint32_t *__restrict__ thr, const int32_t *__restrict__ len1,
                                                  const int32_t *__restrict__ len2, const uint64_t *__restrict__ seed,
                                                  const int32_t *__restrict__ order, int32_t slot_order,
                                                  int32_t per_point, int64_t first, int64_t n, Params P, ScoreTab tab,
                                                  int32_t *__restrict__ out, int32_t *__restrict__ bound,
                                                  int64_t bound_stride, const int64_t *__restrict__ wave_off) {
    __shared__ uint32_t shi[8], slo[8];
    if (threadIdx.x < 8) {
        shi[threadIdx.x] = tab.hi[threadIdx.x];
        slo[threadIdx.x] = tab.lo[threadIdx.x];
    }
    __syncthreads();
    // every lane of the wave stays for the wave-uniform loops of sw_pair
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool ok = gid < n;
    const int64_t slot = first + (ok ? gid : 0);
    int64_t k, r;
    if (order && slot_order) {
        const int64_t q = order[slot];
        k = q / per_point;
        r = q % per_point;
    } else {
        k = order ? (int64_t)order[slot / per_point] : slot / per_point;
        r = slot % per_point;
    }
    const int64_t pair = k * per_point + r;
    RandCodes src;
    src.kref = mix64(seed[k] + 2 * (uint64_t)r);
    src.kqry = mix64(seed[k] + 2 * (uint64_t)r + 1);
    src.tC = thr[3 * k];
    src.tG = thr[3 * k + 1];
    src.tA = thr[3 * k + 2];
    int2 *bnd = ok ? boundary_slot(bound, bound_stride, wave_off, gid) : nullptr;
    int32_t best, bi, bj;
    sw_pair(src, ok ? len1[k] : 0, ok ? len2[k] : 0, P, shi, slo, bnd, best, bi, bj);
    if (ok) out[pair] = best;
}

// One random pair per wave for the longest pairs, which would otherwise hold
// a launch for one lane's walk over every row of every stripe: stripe j of
// the query (columns 32j .. 32j + 31, at most 64 stripes) in lane j, rows on
// an anti-diagonal.  At step t lane j evaluates row t - j; the last column's
// H and clamped horizontal gap of that row in stripe j - 1 arrive from lane
// j - 1 by DPP wave_shr:1 (lane 0 takes the left boundary, zeros), and so does
// the row's reference base (lane 0 takes it from the 16 rows generated every
// 16 steps).  int32, one pair per wave: the pairs list[0, n) (point * per_point
// + r); only the maximum is kept.
__device__ void sw_random_wave_pair(const uint32_t *__restrict__ thr, const int32_t *__restrict__ len1,
                                    const int32_t *__restrict__ len2, const uint64_t *__restrict__ seed, int64_t q,
                                    int32_t per_point, const Params &P, const uint32_t *shi, const uint32_t *slo,
                                    int32_t *__restrict__ out) {
    const int lane = threadIdx.x;
    const int64_t k = q / per_point, r = q % per_point;
    RandCodes src;
    src.kref = mix64(seed[k] + 2 * (uint64_t)r);
    src.kqry = mix64(seed[k] + 2 * (uint64_t)r + 1);
    src.tC = thr[3 * k];
    src.tG = thr[3 * k + 1];
    src.tA = thr[3 * k + 2];
    const int32_t L1 = len1[k], L2 = len2[k];
    const int nstripes = (L2 + S - 1) / S;  // <= 64 (the host's routing)
    const int j0 = lane * S;
    uint32_t sel[S];
    int32_t H[S], E[S];
    {
        uint32_t qw[8];
        src.cols32(j0, qw);
#pragma unroll
        for (int j = 0; j < S; ++j) {
            const int code = (int)(int8_t)((qw[j >> 2] >> (8 * (j & 3))) & 0xffu);
            sel[j] = sel32(j0 + j < L2 ? code : 2);
            H[j] = 0;
            E[j] = 0;
        }
    }
    const int32_t O = P.open, X = P.ext;
    const bool stripe = lane < nstripes;
    int32_t h_out = 0, f_out = 0, c_out = 2, hb_prev = 0, best = 0;
    uint32_t rbuf[4] = {0x02020202u, 0x02020202u, 0x02020202u, 0x02020202u};
    const int32_t steps = L1 + nstripes - 1;
    for (int32_t t = 0; t < steps; ++t) {
        if ((t & 15) == 0) src.rows16(t, rbuf);  // wave-uniform: the pair's next 16 rows
        const int c0 = t < L1 ? (int)(int8_t)((rbuf[(t & 15) >> 2] >> (8 * (t & 3))) & 0xffu) : 2;
        const int c = __builtin_amdgcn_update_dpp(c0, c_out, 0x138, 0xf, 0xf, false);        // row t - lane
        const int32_t hin = __builtin_amdgcn_update_dpp(0, h_out, 0x138, 0xf, 0xf, false);  // H(row, j0 - 1)
        const int32_t fin = __builtin_amdgcn_update_dpp(0, f_out, 0x138, 0xf, 0xf, false);  // gap entering j0
        const int32_t row = t - lane;
        if (stripe && row >= 0 && row < L1) {
            const int ci = (unsigned)c < 8u ? c : 2;
            const uint32_t thi = shi[ci], tlo = slo[ci];
            int32_t fc = fin, rmax = 0, h = 0;
            int32_t tt = hb_prev + (int32_t)perm(thi, tlo, sel[0]);
#pragma unroll
            for (int j = 0; j < S; ++j) {
                h = max3i(tt, E[j], fc);  // E, fc >= 0: includes the 0 floor
                if (j + 1 < S) tt = H[j] + (int32_t)perm(thi, tlo, sel[j + 1]);
                const int32_t ho = h + O;
                E[j] = max3i(E[j] + X, ho, 0);
                fc = max3i(ho, fc + X, 0);
                H[j] = h;
                if (j & 1) rmax = max3i(rmax, H[j - 1], h);
            }
            h_out = h;
            f_out = fc;
            best = max(best, rmax);
        } else {
            h_out = 0;
            f_out = 0;
        }
        hb_prev = hin;
        c_out = c;
    }
    best = wave_max(best);
    if (lane == 0) out[q] = best;
}

__global__ __launch_bounds__(64) void k_sw_random_wave(const uint32_t *__restrict__ thr,
                                                       const int32_t *__restrict__ len1,
                                                       const int32_t *__restrict__ len2,
                                                       const uint64_t *__restrict__ seed,
                                                       const int32_t *__restrict__ list, int32_t per_point,
                                                       int64_t n, Params P, ScoreTab tab,
                                                       int32_t *__restrict__ out) {
    __shared__ uint32_t shi[8], slo[8];
    if (threadIdx.x < 8) {
        shi[threadIdx.x] = tab.hi[threadIdx.x];
        slo[threadIdx.x] = tab.lo[threadIdx.x];
    }
    __syncthreads();
    if ((int64_t)blockIdx.x >= n) return;
    sw_random_wave_pair(thr, len1, len2, seed, list[blockIdx.x], per_point, P, shi, slo, out);
}

// Maxima of two random pairs per lane, pair A in the low f16 half of every
// register and pair B in the high half (exact integers while every value
// stays below 2048).  The stripes, 16-row blocks and boundary rows are those
// of sw_pair; a boundary entry holds both halves.  A row's reference bases of
// A and B give two words of score bytes (the high bytes of the f16 scores
// against A, C, G, T: mismatch, with match at the base's own class), and one
// v_perm_b32 per column picks both halves' scores with a selector built from
// the column's query bases.  Only the maximum is kept (the random pairs need
// no position): one packed max per row.
__device__ void sw_pair_rand_h(const RandCodes &A, const RandCodes &B, int32_t L1a, int32_t L2a, int32_t L1b,
                               int32_t L2b, uint32_t open_h, uint32_t ext_h, uint32_t mis4, uint32_t diff,
                               int2 *__restrict__ bnd, uint32_t &best_bits) {
    const f16x2 zero = as_h2(0u), O = as_h2(open_h * 0x10001u), X = as_h2(ext_h * 0x10001u);
    const int32_t L1w = wave_max(max(L1a, L1b));
    const int nstripes = wave_max((max(L2a, L2b) + S - 1) / S);
    f16x2 best = zero;
    // score-byte selector of a query base: its class (A0 C1 G2 T3), or a zero byte for N
    auto cls = [](int q) -> uint32_t { return (q >= 0) & (q <= 4) & (q != 2) ? (uint32_t)(q - (q > 2)) : 0x0Cu; };
    auto row_word = [&](int c) -> uint32_t {
        const bool base = (c >= 0) & (c <= 4) & (c != 2);
        return base ? mis4 ^ (diff << (8 * (c - (c > 2)))) : 0u;
    };
    for (int s = 0; s < nstripes; ++s) {
        const int j0 = s * S;
        uint32_t sel[S];
        f16x2 H[S], E[S];
        {
            uint32_t qa[8], qb[8];
            A.cols32(j0, qa);
            B.cols32(j0, qb);
#pragma unroll
            for (int j = 0; j < S; ++j) {
                const int ca = j0 + j < L2a ? (int)(int8_t)((qa[j >> 2] >> (8 * (j & 3))) & 0xffu) : 2;
                const int cb = j0 + j < L2b ? (int)(int8_t)((qb[j >> 2] >> (8 * (j & 3))) & 0xffu) : 2;
                const uint32_t ka = cls(ca), kb = cls(cb);
                sel[j] = 0x000C000Cu | (ka << 8) | ((kb == 0x0Cu ? 0x0Cu : kb + 4u) << 24);
                H[j] = zero;
                E[j] = zero;
            }
        }
        const bool has_left = s > 0, has_right = s + 1 < nstripes;
        uint32_t hb_prev = 0;  // boundary H of the previous row (diagonal of column j0)
        for (int32_t ib = 0; ib < L1w; ib += RB) {
            uint32_t ra[4], rb[4];
            A.rows16(ib, ra);
            B.rows16(ib, rb);
            int2 b[RB];
            if (has_left && bnd) {
#pragma unroll
                for (int q = 0; q < RB / 2; ++q) {
                    const int4 v = *(const int4 *)(bnd + ib + 2 * q);
                    b[2 * q] = make_int2(v.x, v.y);
                    b[2 * q + 1] = make_int2(v.z, v.w);
                }
            } else {
#pragma unroll
                for (int r = 0; r < RB; ++r) b[r] = make_int2(0, 0);
            }
            const int rows = min(RB, L1w - ib);  // wave-uniform
#pragma unroll
            for (int r = 0; r < RB; ++r) {
                if (r < rows) {
                    const int ca = ib + r < L1a ? (int)(int8_t)((ra[r >> 2] >> (8 * (r & 3))) & 0xffu) : 2;
                    const int cb = ib + r < L1b ? (int)(int8_t)((rb[r >> 2] >> (8 * (r & 3))) & 0xffu) : 2;
                    const uint32_t wa = row_word(ca), wb = row_word(cb);
                    f16x2 fc = as_h2((uint32_t)b[r].y);
                    // the diagonal term of column j: H(i-1, j-1) plus the score,
                    // formed before H[j - 1] is overwritten
                    f16x2 t = as_h2(hb_prev) + as_h2(perm(wb, wa, sel[0]));
                    hb_prev = (uint32_t)b[r].x;
                    f16x2 rmax = zero, h = zero;
#pragma unroll
                    for (int j = 0; j < S; ++j) {
                        h = hmax3(t, E[j], fc);  // E, fc >= 0: includes the 0 floor
                        if (j + 1 < S) t = H[j] + as_h2(perm(wb, wa, sel[j + 1]));
                        const f16x2 ho = h + O;
                        E[j] = hmax3(E[j] + X, ho, zero);
                        fc = hmax3(ho, fc + X, zero);
                        H[j] = h;
                        if (j & 1) rmax = hmax3(rmax, H[j - 1], h);
                    }
                    b[r] = make_int2((int32_t)bits(h), (int32_t)bits(fc));
                    best = __builtin_elementwise_maximum(best, rmax);
                }
            }
            if (has_right && bnd) {
#pragma unroll
                for (int q = 0; q < RB / 2; ++q)
                    *(int4 *)(bnd + ib + 2 * q) = make_int4(b[2 * q].x, b[2 * q].y, b[2 * q + 1].x, b[2 * q + 1].y);
            }
        }
    }
    best_bits = bits(best);
}

// k_sw_random with two pairs per lane in packed f16 (scorings whose f16
// match and mismatch have a zero low byte): lane l of lane-wave w takes slots
// first + 128 w + l (low half) and + 64 (high half) of the plan's order.  A
// pair whose maximum reaches 2048 (where f16 integers stop being exact) gets
// -1 and is listed in ovf[0, *novf) for the int32 kernels.  The launch's
// first nwave blocks score the pairs wlist[0, nwave) one per wave instead
// (sw_random_wave_pair, int32 with the table tab), so that the longest
// chains start first and the lane waves fill the chip around them.
__global__ __launch_bounds__(64) void k_sw_random_h(const uint32_t *__restrict__ thr, const int32_t *__restrict__ len1,
                                                    const int32_t *__restrict__ len2, const uint64_t *__restrict__ seed,
                                                    const int32_t *__restrict__ order, int32_t slot_order,
                                                    int32_t per_point, int64_t first, int64_t n, uint32_t open_h,
                                                    uint32_t ext_h, uint32_t mis4, uint32_t diff,
                                                    int32_t *__restrict__ out, int32_t *__restrict__ bound,
                                                    const int64_t *__restrict__ wave_off, int32_t *__restrict__ ovf,
                                                    int32_t *__restrict__ novf, const int32_t *__restrict__ wlist,
                                                    int32_t nwave, Params P, ScoreTab tab) {
    if ((int32_t)blockIdx.x < nwave) {
        __shared__ uint32_t shi[8], slo[8];
        if (threadIdx.x < 8) {
            shi[threadIdx.x] = tab.hi[threadIdx.x];
            slo[threadIdx.x] = tab.lo[threadIdx.x];
        }
        __syncthreads();
        sw_random_wave_pair(thr, len1, len2, seed, wlist[blockIdx.x], per_point, P, shi, slo, out);
        return;
    }
    const int32_t lw = (int32_t)blockIdx.x - nwave;  // the lane wave
    // every lane of the wave stays for the wave-uniform loops
    const int64_t g0 = (int64_t)lw * 128 + threadIdx.x;
    int64_t pair[2];
    RandCodes src[2];
    int32_t L1[2], L2[2];
    bool ok[2];
#pragma unroll
    for (int hh = 0; hh < 2; ++hh) {
        const int64_t gid = g0 + 64 * hh;
        ok[hh] = gid < n;
        const int64_t slot = first + (ok[hh] ? gid : 0);
        int64_t k, r;
        if (order && slot_order) {
            const int64_t q = order[slot];
            k = q / per_point;
            r = q % per_point;
        } else {
            k = order ? (int64_t)order[slot / per_point] : slot / per_point;
            r = slot % per_point;
        }
        pair[hh] = k * per_point + r;
        src[hh].kref = mix64(seed[k] + 2 * (uint64_t)r);
        src[hh].kqry = mix64(seed[k] + 2 * (uint64_t)r + 1);
        src[hh].tC = thr[3 * k];
        src[hh].tG = thr[3 * k + 1];
        src[hh].tA = thr[3 * k + 2];
        L1[hh] = ok[hh] ? len1[k] : 0;
        L2[hh] = ok[hh] ? len2[k] : 0;
    }
    int2 *bnd = nullptr;
    if (ok[0]) {
        const int64_t a = wave_off[lw], per_lane = (wave_off[lw + 1] - a) / 64;
        bnd = (int2 *)(bound + a + (int64_t)threadIdx.x * per_lane);
    }
    uint32_t bb;
    sw_pair_rand_h(src[0], src[1], L1[0], L2[0], L1[1], L2[1], open_h, ext_h, mis4, diff, bnd, bb);
#pragma unroll
    for (int hh = 0; hh < 2; ++hh) {
        if (!ok[hh]) continue;
        const uint32_t v = hh ? bb >> 16 : bb & 0xffffu;  // non-negative f16: bits order like values
        if (v >= 0x6800u) {  // 2048
            out[pair[hh]] = -1;
            ovf[atomicAdd(novf, 1)] = (int32_t)pair[hh];
        } else {
            out[pair[hh]] = (int32_t)(float)__builtin_bit_cast(_Float16, (uint16_t)v);
        }
    }
}

// SmithWaterman_bas (impl/SmithWaterman_bas.cpp:59-201) as written, for one
// task per thread: the full M, E, F matrices (in ws, 3 x (L1+1) x (L2+1)
// int32 from ws_off[t]) under the reference's initialisation (reverse:
// E = open + ext on row and column 0, F(0, j) = open + ext + ext * j), its
// three-way recurrences and comparisons, the first strict maximum, then its
// traceback walk.  The walk as written keeps going without moving as soon as
// its state variable (assigned the int8 of a score) is neither 0 nor, after a
// move, taken back to 0; status 2 reports that the reference's loop would
// not end, status 1 the walk's stop at a zero cell (the reference prints
// "... line 110" to stderr), status 0 the walk reaching row or column 0.
// Cigar entries as the walk appends them (len << 4 | op).
__global__ __launch_bounds__(64) void k_sw_bas_full(const int8_t *__restrict__ seq, const int64_t *__restrict__ off1,
                                                    const int32_t *__restrict__ len1,
                                                    const int64_t *__restrict__ off2,
                                                    const int32_t *__restrict__ len2, int32_t n, int32_t open,
                                                    int32_t ext, int32_t reverse, int32_t *__restrict__ ws,
                                                    const int64_t *__restrict__ ws_off, uint32_t *__restrict__ out3,
                                                    uint32_t *__restrict__ cig, const int64_t *__restrict__ cig_off,
                                                    int32_t *__restrict__ ncig, int32_t *__restrict__ status) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int8_t *s1 = seq + off1[t], *s2 = seq + off2[t];
    const int32_t L1 = len1[t], L2 = len2[t], W = L2 + 1;
    int32_t *M = ws + ws_off[t];
    int32_t *E = M + (int64_t)(L1 + 1) * W;
    int32_t *F = E + (int64_t)(L1 + 1) * W;
    const int32_t q = open + ext, e = ext;
    auto sc = [&](int a, int b) -> int32_t {  // the file's score matrix
        if (a == 2 || b == 2) return (a == 2 && b == 2) ? -1 : 0;
        return a == b ? 2 : -1;
    };
    for (int32_t i = 0; i <= L1; ++i)
        for (int32_t j = 0; j <= L2; ++j) {
            const int64_t x = (int64_t)i * W + j;
            M[x] = 0;
            E[x] = 0;
            F[x] = 0;
        }
    if (reverse) {
        for (int32_t j = 0; j <= L2; ++j) {
            E[j] = q;
            F[j] = q + e * j;
        }
        for (int32_t i = 0; i <= L1; ++i) {
            E[(int64_t)i * W] = q;
            F[(int64_t)i * W] = q;
        }
    }
    uint32_t maxScore = 0, end1 = 0, end2 = 0;
    for (int32_t i = 1; i <= L1; ++i) {
        const int a = s1[i - 1];
        for (int32_t j = 1; j <= L2; ++j) {
            const int64_t x = (int64_t)i * W + j, xl = x - 1, xu = x - W, xd = x - W - 1;
            int32_t ev = (q + M[xl]) >= (e + E[xl]) ? q + M[xl] : e + E[xl];
            if (ev < q + F[xl]) ev = q + F[xl];
            int32_t fv = (q + M[xu]) >= (q + E[xu]) ? q + M[xu] : q + E[xu];
            if (fv < e + F[xu]) fv = e + F[xu];
            const int32_t dg = sc(a, s2[j - 1]) + M[xd];
            int32_t mv = dg >= ev ? dg : ev;
            if (mv < fv) mv = fv;
            if (mv < 0) mv = 0;
            E[x] = ev;
            F[x] = fv;
            M[x] = mv;
            if ((uint32_t)mv > maxScore) {
                maxScore = (uint32_t)mv;
                end1 = (uint32_t)i;
                end2 = (uint32_t)j;
            }
        }
    }
    out3[3 * t] = maxScore;
    out3[3 * t + 1] = end1 - 1u;  // the file's final decrement (0-based; all ones when nothing scored)
    out3[3 * t + 2] = end2 - 1u;
    int32_t nc = 0, st = 0;
    if (reverse) {
        uint32_t *cg = cig + cig_off[t];
        const int64_t cap = cig_off[t + 1] - cig_off[t];
        int32_t ii = (int32_t)end1, jj = (int32_t)end2;
        int8_t which = 0;
        uint32_t op = 0;
        while (ii > 0 && jj > 0) {
            bool moved = false;
            if (which == 0) {
                op = 0;
                const int64_t x = (int64_t)ii * W + jj, xd = x - W - 1;
                which = (int8_t)M[xd];
                if (M[x] == M[xd] + sc(s1[ii - 1], s2[jj - 1])) {
                    --ii;
                    --jj;
                    moved = true;
                } else if (M[x] == E[x]) {
                    --jj;
                    op = 1;
                    moved = true;
                } else if (M[x] == F[x]) {
                    --ii;
                    op = 2;
                    moved = true;
                } else if (M[x] == 0) {
                    st = 1;
                    break;
                }
            } else if (which == 1) {
                op = 1;
                which = (int8_t)M[(int64_t)(ii - 1) * W + (jj - 1)];
            }
            if (nc == 0 || op != (cg[nc - 1] & 0xfu)) {
                if (nc >= cap) {
                    st = 2;
                    break;
                }
                cg[nc++] = (1u << 4) | op;
            } else {
                cg[nc - 1] += 1u << 4;
            }
            // without a move the state can only change from 1 to the diagonal
            // score; once it is neither 0 nor 1 after such a step, or the
            // step repeats, nothing changes again: the reference never ends
            if (!moved && !(which == 0 && op == 1)) {
                st = 2;
                break;
            }
        }
    }
    ncig[t] = nc;
    status[t] = st;
}

}  // namespace

hipError_t launch_sw_batch(hipStream_t st, const int8_t *seq, const int64_t *off1, const int32_t *len1,
                           const int64_t *off2, const int32_t *len2, const int32_t *order, int32_t n, const Params &P,
                           const ScoreTab &tab, int32_t *out, int32_t *bound, const int64_t *wave_off, bool nn) {
    if (n <= 0) return hipSuccess;
    auto k = nn ? k_sw_batch<true> : k_sw_batch<false>;
    hipLaunchKernelGGL(k, dim3((n + 63) / 64), dim3(64), 0, st, seq, off1, len1, off2, len2, order, 0, n, P, tab, out,
                       bound, (int64_t)0, wave_off, 0, nullptr);
    return hipGetLastError();
}

hipError_t launch_sw_bas_full(hipStream_t st, const int8_t *seq, const int64_t *off1, const int32_t *len1,
                              const int64_t *off2, const int32_t *len2, int32_t n, int32_t open, int32_t ext,
                              bool reverse, int32_t *ws, const int64_t *ws_off, uint32_t *out3, uint32_t *cig,
                              const int64_t *cig_off, int32_t *ncig, int32_t *status) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sw_bas_full, dim3((n + 63) / 64), dim3(64), 0, st, seq, off1, len1, off2, len2, n, open, ext,
                       (int32_t)reverse, ws, ws_off, out3, cig, cig_off, ncig, status);
    return hipGetLastError();
}

hipError_t launch_sw_random(hipStream_t st, const uint32_t *thr, const int32_t *len1, const int32_t *len2,
                            const uint64_t *seed, const int32_t *order, bool slot_order, int32_t per_point,
                            int64_t first, int64_t n, const Params &P, const ScoreTab &tab, int32_t *out,
                            int32_t *bound, const int64_t *wave_off) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sw_random, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, st, thr, len1, len2, seed, order,
                       (int32_t)slot_order, per_point, first, n, P, tab, out, bound, (int64_t)0, wave_off);
    return hipGetLastError();
}

hipError_t launch_sw_random_wave(hipStream_t st, const uint32_t *thr, const int32_t *len1, const int32_t *len2,
                                 const uint64_t *seed, const int32_t *list, int32_t per_point, int64_t n,
                                 const Params &P, const ScoreTab &tab, int32_t *out) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sw_random_wave, dim3((unsigned)n), dim3(64), 0, st, thr, len1, len2, seed, list, per_point, n,
                       P, tab, out);
    return hipGetLastError();
}

bool sw_random_packed(const Params &P, uint32_t &mis4, uint32_t &diff) {
    const uint32_t m = f16_bits_of_int(P.match), x = f16_bits_of_int(P.mismatch);
    if ((m & 0xffu) || (x & 0xffu)) return false;
    if (P.open < -2048 || P.ext < -2048) return false;
    mis4 = (x >> 8) * 0x01010101u;
    diff = (m ^ x) >> 8;
    return true;
}

hipError_t launch_sw_random_h(hipStream_t st, const uint32_t *thr, const int32_t *len1, const int32_t *len2,
                              const uint64_t *seed, const int32_t *order, bool slot_order, int32_t per_point,
                              int64_t first, int64_t n, const Params &P, const ScoreTab &tab, int32_t *out,
                              int32_t *bound, const int64_t *wave_off, int32_t *ovf, int32_t *novf,
                              const int32_t *wlist, int32_t nwave) {
    if (n <= 0 && nwave <= 0) return hipSuccess;
    uint32_t mis4, diff;
    if (!sw_random_packed(P, mis4, diff)) return hipErrorInvalidValue;
    const unsigned lane_waves = n > 0 ? (unsigned)((n + 127) / 128) : 0u;
    hipLaunchKernelGGL(k_sw_random_h, dim3(lane_waves + (unsigned)nwave), dim3(64), 0, st, thr, len1, len2, seed,
                       order, (int32_t)slot_order, per_point, first, n, f16_bits_of_int(P.open),
                       f16_bits_of_int(P.ext), mis4, diff, out, bound, wave_off, ovf, novf, wlist, nwave, P, tab);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Lane plans of the K1 launches.
// ---------------------------------------------------------------------------
void sw_plan(const int32_t *len1, const int32_t *len2, int64_t npoints, int32_t per_point, SwPlan &plan,
             int32_t wslots, const std::vector<uint8_t> *skip) {
    plan.order.clear();
    plan.waves.clear();
    plan.per_point = per_point;
    int64_t kept = npoints;
    if (skip)
        for (int64_t k = 0; k < npoints; ++k) kept -= (*skip)[(size_t)k] ? 1 : 0;
    plan.total = kept * per_point;
    plan.wslots = wslots;
    const int64_t WS = wslots;
    if (plan.total == 0) return;
    auto blk = [&](int64_t k) { return (std::max(len1[k], 1) + 15) / 16; };
    auto str = [&](int64_t k) { return (std::max(len2[k], 1) + 31) / 32; };
    plan.slot_order = per_point < WS;
    if (plan.slot_order) {
        // pairs by loop shape (stripes, then 16-row blocks, both descending):
        // two stable counting sorts, then the pairs of each point in turn
        int32_t max_blk = 1, max_str = 1;
        for (int64_t k = 0; k < npoints; ++k) {
            max_blk = std::max(max_blk, blk(k) + 1);
            max_str = std::max(max_str, str(k) + 1);
        }
        auto counting = [&](const std::vector<int32_t> &in, std::vector<int32_t> &out, auto key, int32_t range) {
            std::vector<int64_t> start((size_t)range + 1, 0);
            for (int32_t k : in) ++start[(size_t)(range - 1 - key(k)) + 1];
            for (size_t b = 1; b < start.size(); ++b) start[b] += start[b - 1];
            out.resize(in.size());
            for (int32_t k : in) out[(size_t)start[(size_t)(range - 1 - key(k))]++] = k;
        };
        std::vector<int32_t> ident((size_t)npoints), by_blk, by_shape;
        for (int64_t k = 0; k < npoints; ++k) ident[(size_t)k] = (int32_t)k;
        counting(ident, by_blk, blk, max_blk);
        counting(by_blk, by_shape, str, max_str);
        std::vector<int32_t> slots;
        slots.reserve((size_t)plan.total);
        for (int32_t k : by_shape) {
            if (skip && (*skip)[(size_t)k]) continue;
            for (int32_t r = 0; r < per_point; ++r) slots.push_back(k * per_point + r);
        }
        // waves of WS slots, their loops the widest slot's; the waves longest first
        const int64_t nw = (plan.total + WS - 1) / WS;
        std::vector<int64_t> cost((size_t)nw);
        std::vector<int32_t> rows((size_t)nw), widx((size_t)nw);
        for (int64_t w = 0; w < nw; ++w) {
            int32_t b = 1, t = 1;
            for (int64_t s = WS * w; s < std::min(plan.total, WS * w + WS); ++s) {
                b = std::max(b, blk(slots[(size_t)s] / per_point));
                t = std::max(t, str(slots[(size_t)s] / per_point));
            }
            cost[(size_t)w] = (int64_t)b * t;
            rows[(size_t)w] = 16 * b;
            widx[(size_t)w] = (int32_t)w;
        }
        // a partial last wave stays last, so that every wave before it is full
        if (plan.total % WS) cost[(size_t)(nw - 1)] = -1;
        std::stable_sort(widx.begin(), widx.end(), [&](int32_t a, int32_t b) { return cost[(size_t)a] > cost[(size_t)b]; });
        plan.order.reserve((size_t)plan.total);
        for (int32_t w : widx) {
            const int64_t s0 = WS * (int64_t)w, s1 = std::min(plan.total, s0 + WS);
            for (int64_t s = s0; s < s1; ++s) plan.order.push_back(slots[(size_t)s]);
            plan.waves.push_back(rows[(size_t)w]);
        }
    } else {
        // points longest first; a wave holds at most two points' pairs
        for (int64_t k = 0; k < npoints; ++k)
            if (!skip || !(*skip)[(size_t)k]) plan.order.push_back((int32_t)k);
        std::stable_sort(plan.order.begin(), plan.order.end(), [&](int32_t a, int32_t b) {
            return (int64_t)blk(a) * str(a) > (int64_t)blk(b) * str(b);
        });
        const int64_t nw = (plan.total + WS - 1) / WS;
        plan.waves.resize((size_t)nw);
        for (int64_t w = 0; w < nw; ++w) {
            const int64_t a = plan.order[(size_t)(WS * w / per_point)];
            const int64_t b = plan.order[(size_t)(std::min(plan.total - 1, WS * w + WS - 1) / per_point)];
            plan.waves[(size_t)w] = 16 * std::max(blk(a), blk(b));
        }
    }
}

// Issue slots (wave-instructions) of one row of one 32-column stripe, in the
// lane kernels (shared by 64 int32 or 128 packed pairs) and per step of the
// wave kernel (one pair), and the cycles such a step takes at two waves per
// SIMD.
constexpr double kSwStepIssue = 300.0, kSwStepLatency = 2400.0;

std::vector<uint8_t> sw_wave_points(const int32_t *len1, const int32_t *len2, int64_t npoints, int32_t per_point,
                                    bool packed, int32_t cus) {
    // The model runs over quarter-octave buckets of a point's lane chain
    // (rows x stripes), longest bucket first: routing only sets the time, so
    // the cut need not fall between two particular points.
    std::vector<uint8_t> wave((size_t)npoints, 0);
    const double lanes = packed ? 128.0 : 64.0, pp = per_point;
    auto st = [&](int64_t k) { return (int64_t)((std::max(len2[k], 1) + 31) / 32); };
    auto rows = [&](int64_t k) { return (int64_t)std::max(len1[k], 1); };
    constexpr int NB = 4 * 64;
    auto bucket = [](int64_t key) {
        const int lg = 63 - __builtin_clzll((uint64_t)key | 1);
        const int q = lg >= 2 ? (int)(((uint64_t)key >> (lg - 2)) & 3) : 0;
        return std::min(NB - 1, 4 * lg + q);
    };
    std::vector<double> b_lane(NB, 0), b_wave(NB, 0), b_wchain(NB, 0), b_chain(NB, 0);
    double lane_issue = 0, fixed_chain = 0;
    std::vector<uint16_t> bk((size_t)npoints);
    for (int64_t k = 0; k < npoints; ++k) {
        const int64_t key = rows(k) * st(k);
        const double li = pp * (double)key * kSwStepIssue / lanes;
        lane_issue += li;
        if (st(k) > 64) {  // past 64 stripes: lanes only
            fixed_chain = std::max(fixed_chain, (double)key * kSwStepLatency);
            bk[(size_t)k] = 0xffff;
            continue;
        }
        const int b = bucket(key);
        bk[(size_t)k] = (uint16_t)b;
        b_lane[(size_t)b] += li;
        b_wave[(size_t)b] += pp * (double)(rows(k) + st(k) - 1) * kSwStepIssue;
        b_wchain[(size_t)b] = std::max(b_wchain[(size_t)b], (double)(rows(k) + st(k) - 1) * kSwStepLatency);
        b_chain[(size_t)b] = std::max(b_chain[(size_t)b], (double)key * kSwStepLatency);
    }
    const double cap = (double)std::max(cus, 1);  // wave-instructions per cycle over the chip: cus x 4 SIMDs / 4
    auto chain_below = [&](int b) {  // the longest lane chain of the buckets below b
        double c = fixed_chain;
        for (int x = b - 1; x >= 0; --x)
            if (b_chain[(size_t)x] > 0) return std::max(c, b_chain[(size_t)x]);
        return c;
    };
    double best = std::max(chain_below(NB), lane_issue / cap), wave_issue = 0, wave_chain = 0;
    int cut = NB;  // buckets >= cut go to the wave kernel
    for (int b = NB - 1; b >= 0; --b) {
        if (b_chain[(size_t)b] == 0) continue;
        lane_issue -= b_lane[(size_t)b];
        wave_issue += b_wave[(size_t)b];
        wave_chain = std::max(wave_chain, b_wchain[(size_t)b]);
        const double t = std::max(std::max(chain_below(b), wave_chain), (lane_issue + wave_issue) / cap);
        if (t < best) {
            best = t;
            cut = b;
        }
    }
    for (int64_t k = 0; k < npoints; ++k) wave[(size_t)k] = bk[(size_t)k] != 0xffff && bk[(size_t)k] >= cut;
    return wave;
}

// Lane plan over listed pairs (pair = point * per_point + r), one per slot:
// waves of 64 slots ordered by loop shape (stripes, then 16-row blocks, both
// descending), so the longest waves come first.
void sw_plan_pairs(const int32_t *len1, const int32_t *len2, int32_t per_point, const std::vector<int32_t> &pairs,
                   SwPlan &plan) {
    plan = SwPlan{};
    plan.per_point = per_point;
    plan.wslots = 64;
    plan.slot_order = true;
    plan.total = (int64_t)pairs.size();
    auto blk = [&](int32_t q) { return (std::max(len1[q / per_point], 1) + 15) / 16; };
    auto str = [&](int32_t q) { return (std::max(len2[q / per_point], 1) + 31) / 32; };
    plan.order = pairs;
    std::stable_sort(plan.order.begin(), plan.order.end(), [&](int32_t a, int32_t b) {
        return str(a) != str(b) ? str(a) > str(b) : blk(a) > blk(b);
    });
    for (int64_t s0 = 0; s0 < plan.total; s0 += 64) {
        int32_t b = 1;
        for (int64_t s = s0; s < std::min(plan.total, s0 + 64); ++s) b = std::max(b, blk(plan.order[(size_t)s]));
        plan.waves.push_back(16 * b);
    }
}

// Launches [w0, w1) of the plan's waves fit `budget` int32 of boundary rows
// (at least one wave).
int64_t sw_plan_chunk(const SwPlan &plan, int64_t w0, int64_t budget, std::vector<int64_t> &wave_off) {
    wave_off.assign(1, 0);
    int64_t w = w0;
    const int64_t nw = (int64_t)plan.waves.size();
    while (w < nw) {
        const int64_t ints = (int64_t)plan.waves[(size_t)w] * 2 * 64;
        if (w > w0 && wave_off.back() + ints > budget) break;
        wave_off.push_back(wave_off.back() + ints);
        ++w;
    }
    return w;
}

// ---------------------------------------------------------------------------
// dcns_sw_batch_noalloc: dcns_sw_batch on device-resident pairs with the
// stripe boundaries in a library-owned device array, for callers that cannot
// allocate device memory.  The array is one per device (its address taken
// once per device at context creation); every call on a device shares it.
// ---------------------------------------------------------------------------
namespace {
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;  // 64 MiB per device
constexpr size_t kNoAllocHeader = 256;                     // error word, then the boundary rows
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

__global__ void k_clear_word(int32_t *w) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *w = 0;
}
}  // namespace

hipError_t sw_noalloc_init(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return hipSuccess;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));  // the device is current
    if (e == hipSuccess) g_scratch_ptr[device] = ptr;
    return e;
}

size_t sw_noalloc_scratch_bytes() { return kNoAllocScratchBytes; }

// Rows of the longest reference a call can take: one pair's boundary slot
// must fit the array.
int64_t sw_noalloc_max_rows() { return (int64_t)((kNoAllocScratchBytes - kNoAllocHeader) / 8) - 16; }

hipError_t sw_noalloc_run(int device, hipStream_t st, const int8_t *seq, const int64_t *off1, const int32_t *len1,
                          const int64_t *off2, const int32_t *len2, int32_t n, int32_t max_len1, const Params &P,
                          const ScoreTab &tab, int32_t *out3, int32_t *err_host) {
    void *base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (device >= 0 && device < kMaxDevices) base = g_scratch_ptr[device];
    }
    if (!base) return hipErrorNotInitialized;
    int32_t *err = (int32_t *)base;
    int32_t *bound = (int32_t *)((unsigned char *)base + kNoAllocHeader);
    const int64_t rows = ((int64_t)std::max(max_len1, 1) + 15) / 16 * 16;  // rows rounded to the 16-row blocks
    const int64_t stride = rows * 2;                                        // int2 per row, in int32 units
    const int64_t per_launch = (int64_t)((kNoAllocScratchBytes - kNoAllocHeader) / 4) / stride;
    hipLaunchKernelGGL(k_clear_word, dim3(1), dim3(64), 0, st, err);
    hipError_t e = hipGetLastError();
    // launches in turn reuse the boundary array; stream order keeps them apart
    for (int32_t s = 0; e == hipSuccess && s < n; s += (int32_t)std::min<int64_t>(per_launch, n - s)) {
        const int32_t m = (int32_t)std::min<int64_t>(per_launch, n - s);
        hipLaunchKernelGGL(k_sw_batch<false>, dim3((m + 63) / 64), dim3(64), 0, st, seq, off1, len1, off2, len2, nullptr, s,
                           m, P, tab, out3, bound, stride, (const int64_t *)nullptr, max_len1, err);
        e = hipGetLastError();
    }
    if (e != hipSuccess) return e;
    e = hipMemcpyAsync(err_host, err, sizeof(int32_t), hipMemcpyDeviceToHost, st);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(st);
}

}  // namespace dcns
