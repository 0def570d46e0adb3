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
struct MemCodes {
    const int8_t *ref, *qry;
    __device__ __forceinline__ void rows16(int32_t ib, uint32_t w[4]) const {
        const uint4 v = *(const uint4 *)(ref + ib);
        w[0] = v.x; w[1] = v.y; w[2] = v.z; w[3] = v.w;
    }
    __device__ __forceinline__ void cols32(int32_t j0, uint32_t w[8]) const {
        const uint4 a = *(const uint4 *)(qry + j0), b = *(const uint4 *)(qry + j0 + 16);
        w[0] = a.x; w[1] = a.y; w[2] = a.z; w[3] = a.w; w[4] = b.x; w[5] = b.y; w[6] = b.z; w[7] = b.w;
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

template <class Codes>
__device__ void sw_pair(const Codes &src, int32_t L1, int32_t L2, const Params &P, const uint32_t *shi,
                        const uint32_t *slo, int2 *__restrict__ bnd, int32_t &best, int32_t &bi, int32_t &bj) {
    best = 0;
    bi = -1;
    bj = -1;
    const int32_t O = P.open, X = P.ext;
    const int nstripes = (L2 + S - 1) / S;
    for (int s = 0; s < nstripes; ++s) {
        const int j0 = s * S;
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
        const bool has_left = s > 0, has_right = j0 + S < L2;
        int32_t sbest = 0, si = -1, sj = -1;
        int32_t hb_prev = 0;  // boundary H of the previous row (diagonal of column j0)
        for (int32_t ib = 0; ib < L1; ib += RB) {
            uint32_t rw[4];
            src.rows16(ib, rw);
            int2 b[RB];
            if (has_left) {
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
            const int rows = min(RB, L1 - ib);
#pragma unroll
            for (int r = 0; r < RB; ++r) {
                if (r < rows) {
                    const int c = (int)(int8_t)((rw[r >> 2] >> (8 * (r & 3))) & 0xffu);
                    const int ci = (unsigned)c < 8u ? c : 2;
                    const uint32_t thi = shi[ci], tlo = slo[ci];
                    int32_t left = b[r].x, fc = b[r].y;
                    int32_t diag = hb_prev;
                    hb_prev = left;
                    int32_t rmax = 0;
#pragma unroll
                    for (int j = 0; j < S; ++j) {
                        const int32_t sc = (int32_t)perm(thi, tlo, sel[j]);
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
            if (has_right) {
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

__global__ __launch_bounds__(64) void k_sw_batch(const int8_t *__restrict__ seq, const int64_t *__restrict__ off1,
                                                 const int32_t *__restrict__ len1, const int64_t *__restrict__ off2,
                                                 const int32_t *__restrict__ len2, const int32_t *__restrict__ order,
                                                 int32_t n, Params P, ScoreTab tab, int32_t *__restrict__ out,
                                                 int32_t *__restrict__ bound, int64_t bound_stride) {
    __shared__ uint32_t shi[8], slo[8];
    if (threadIdx.x < 8) {
        shi[threadIdx.x] = tab.hi[threadIdx.x];
        slo[threadIdx.x] = tab.lo[threadIdx.x];
    }
    __syncthreads();
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= n) return;
    const int p = order ? order[gid] : gid;
    const int32_t L1 = len1[p], L2 = len2[p];
    int2 *bnd = (int2 *)(bound + (int64_t)gid * bound_stride);
    int32_t best, bi, bj;
    sw_pair(MemCodes{seq + off1[p], seq + off2[p]}, L1, L2, P, shi, slo, bnd, best, bi, bj);
    out[3 * p + 0] = best;
    out[3 * p + 1] = bi;
    out[3 * p + 2] = bj;
}

// Max scores of random fragment pairs generated on the device: pair r of
// point k has lengths len1[k] x len2[k] and base thresholds thr[3k..3k+2]
// (C, G, A cumulative, out of 2^32); its keys are mix64(seed[k] + 2r)
// (reference) and mix64(seed[k] + 2r + 1) (query).
// Lanes of a wave share a point, so their loops are uniform.
__global__ __launch_bounds__(64) void k_sw_random(const uint32_t *__restrict__ thr, const int32_t *__restrict__ len1,
                                                  const int32_t *__restrict__ len2, const uint64_t *__restrict__ seed,
                                                  int32_t per_point, int64_t first, int64_t n, Params P, ScoreTab tab,
                                                  int32_t *__restrict__ out, int32_t *__restrict__ bound,
                                                  int64_t bound_stride) {
    __shared__ uint32_t shi[8], slo[8];
    if (threadIdx.x < 8) {
        shi[threadIdx.x] = tab.hi[threadIdx.x];
        slo[threadIdx.x] = tab.lo[threadIdx.x];
    }
    __syncthreads();
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= n) return;
    const int64_t pair = first + gid;
    const int64_t k = pair / per_point, r = pair % per_point;
    RandCodes src;
    src.kref = mix64(seed[k] + 2 * (uint64_t)r);
    src.kqry = mix64(seed[k] + 2 * (uint64_t)r + 1);
    src.tC = thr[3 * k];
    src.tG = thr[3 * k + 1];
    src.tA = thr[3 * k + 2];
    int2 *bnd = (int2 *)(bound + gid * bound_stride);
    int32_t best, bi, bj;
    sw_pair(src, len1[k], len2[k], P, shi, slo, bnd, best, bi, bj);
    out[pair] = best;
}

}  // namespace

hipError_t launch_sw_batch(hipStream_t st, const int8_t *seq, const int64_t *off1, const int32_t *len1,
                           const int64_t *off2, const int32_t *len2, const int32_t *order, int32_t n, const Params &P,
                           const ScoreTab &tab, int32_t *out, int32_t *bound, int64_t bound_stride) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sw_batch, dim3((n + 63) / 64), dim3(64), 0, st, seq, off1, len1, off2, len2, order, n, P, tab,
                       out, bound, bound_stride);
    return hipGetLastError();
}

hipError_t launch_sw_random(hipStream_t st, const uint32_t *thr, const int32_t *len1, const int32_t *len2,
                            const uint64_t *seed, int32_t per_point, int64_t first, int64_t n, const Params &P,
                            const ScoreTab &tab, int32_t *out, int32_t *bound, int64_t bound_stride) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sw_random, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, st, thr, len1, len2, seed, per_point,
                       first, n, P, tab, out, bound, bound_stride);
    return hipGetLastError();
}

}  // namespace dcns
