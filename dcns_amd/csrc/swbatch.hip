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

// ---------------------------------------------------------------------------
// Lane plans of the K1 launches.
// ---------------------------------------------------------------------------
void sw_plan(const int32_t *len1, const int32_t *len2, int64_t npoints, int32_t per_point, SwPlan &plan) {
    plan.order.clear();
    plan.waves.clear();
    plan.per_point = per_point;
    plan.total = npoints * per_point;
    auto blk = [&](int64_t k) { return (std::max(len1[k], 1) + 15) / 16; };
    auto str = [&](int64_t k) { return (std::max(len2[k], 1) + 31) / 32; };
    plan.slot_order = per_point < 64;
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
        for (int32_t k : by_shape)
            for (int32_t r = 0; r < per_point; ++r) slots.push_back(k * per_point + r);
        // waves of 64 slots, their loops the widest lane's; the waves longest first
        const int64_t nw = (plan.total + 63) / 64;
        std::vector<int64_t> cost((size_t)nw);
        std::vector<int32_t> rows((size_t)nw), widx((size_t)nw);
        for (int64_t w = 0; w < nw; ++w) {
            int32_t b = 1, t = 1;
            for (int64_t s = 64 * w; s < std::min(plan.total, 64 * w + 64); ++s) {
                b = std::max(b, blk(slots[(size_t)s] / per_point));
                t = std::max(t, str(slots[(size_t)s] / per_point));
            }
            cost[(size_t)w] = (int64_t)b * t;
            rows[(size_t)w] = 16 * b;
            widx[(size_t)w] = (int32_t)w;
        }
        // a partial last wave stays last, so that every wave before it is full
        if (plan.total % 64) cost[(size_t)(nw - 1)] = -1;
        std::stable_sort(widx.begin(), widx.end(), [&](int32_t a, int32_t b) { return cost[(size_t)a] > cost[(size_t)b]; });
        plan.order.reserve((size_t)plan.total);
        for (int32_t w : widx) {
            const int64_t s0 = 64 * (int64_t)w, s1 = std::min(plan.total, s0 + 64);
            for (int64_t s = s0; s < s1; ++s) plan.order.push_back(slots[(size_t)s]);
            plan.waves.push_back(rows[(size_t)w]);
        }
    } else {
        // points longest first; a wave holds at most two points' pairs
        plan.order.resize((size_t)npoints);
        for (int64_t k = 0; k < npoints; ++k) plan.order[(size_t)k] = (int32_t)k;
        std::stable_sort(plan.order.begin(), plan.order.end(), [&](int32_t a, int32_t b) {
            return (int64_t)blk(a) * str(a) > (int64_t)blk(b) * str(b);
        });
        const int64_t nw = (plan.total + 63) / 64;
        plan.waves.resize((size_t)nw);
        for (int64_t w = 0; w < nw; ++w) {
            const int64_t a = plan.order[(size_t)(64 * w / per_point)];
            const int64_t b = plan.order[(size_t)(std::min(plan.total - 1, 64 * w + 63) / per_point)];
            plan.waves[(size_t)w] = 16 * std::max(blk(a), blk(b));
        }
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
