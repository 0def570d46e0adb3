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
__device__ void sw_pair(const int8_t *__restrict__ ref, int32_t L1, const int8_t *__restrict__ qry, int32_t L2,
                        const Params &P, const uint32_t *shi, const uint32_t *slo, int2 *__restrict__ bnd,
                        int32_t &best, int32_t &bi, int32_t &bj) {
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
            const uint4 q0 = *(const uint4 *)(qry + j0), q1 = *(const uint4 *)(qry + j0 + 16);
            const uint32_t qw[8] = {q0.x, q0.y, q0.z, q0.w, q1.x, q1.y, q1.z, q1.w};
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
            const uint4 rc4 = *(const uint4 *)(ref + ib);
            const uint32_t rw[4] = {rc4.x, rc4.y, rc4.z, rc4.w};
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
    sw_pair(seq + off1[p], L1, seq + off2[p], L2, P, shi, slo, bnd, best, bi, bj);
    out[3 * p + 0] = best;
    out[3 * p + 1] = bi;
    out[3 * p + 2] = bj;
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

}  // namespace dcns
