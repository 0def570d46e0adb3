// semiglobal.hip -- the two-piece affine semi-global DP with traceback
// (SemiGlobal -> SemiGlobal_no_avx, impl/sequenceAlignment3.cpp:681-941) for
// many independent tasks.
//
// Recurrence (row i = seq1 base, column j = seq2 base, band [max(i-w,1),
// min(i+w,C)]), as the reference evaluates it:
//   m = s(i,j) + M(i-1,j-1); then F1, e1, F2, e2 in that order, each taking
//   over when not strictly below the running value (d = 0..4, later wins);
//   F1 = max(F1 + x1, m + o1) (bit 3 when the extension is kept), e1 likewise
//   (bit 4), F2 / e2 with the second piece (bits 5, 6).  The last column of a
//   row only takes the maximum: no gap update, no trace byte.  Row maxima are
//   the first strict maximum; the caller's termination test (z-drop, empty
//   row) runs over the row results afterwards.
//
// Device layout: one wave per task.  Columns are cut into strips of 64 lanes
// x 4 columns; inside a strip lane l works on row r = t - l at step t (an
// anti-diagonal skew), so the values a row carries from left to right (M, the
// two horizontal gap states, the running row maximum, the reference base) move
// one lane per step by a DPP wave shift.  The strip's last lane hands each
// row's carried values to the next strip through a per-task boundary array;
// the first lane reads them 64 rows at a time.  Every row in the band is
// computed (the reference's termination test usually never fires with the
// default long-gap piece, whose column state keeps every later row within
// open2 of the best score); the finish kernel cuts the rows at the reference's
// termination row.  Trace bytes are stored step-major (64 lanes x 4 bytes per
// step, one coalesced 256-byte store), 255 outside the band as after
// Matrix::reset.
#include <hip/hip_runtime.h>

#include "dp_common.h"
#include "kernels.h"

namespace dcns {
namespace {

constexpr int K = 4;                 // columns per lane
constexpr int SW = 64 * K;           // strip width
constexpr int32_t NEG = -536870912;  // SCORE_OUT_BANDED_ALIGNMENT_REGION
constexpr int BND = 5;               // ints per boundary record: M, e1, e2, row max, row max column

__device__ __forceinline__ int32_t shr1(int32_t v) {
    // wave_shr:1 -- lane l receives lane l-1's value; lane 0 keeps its own
    return __builtin_amdgcn_update_dpp(v, v, 0x138, 0xf, 0xf, false);
}

__device__ __forceinline__ int32_t code_at(const int8_t *base, int32_t a, int32_t k, int32_t rev) {
    // k-th base (0-based) of the forward suffix starting at a, or of the
    // reverse complement of the prefix ending at a
    return rev ? 4 - (int32_t)base[a - k] : (int32_t)base[a + k];
}

// Scores of one column for reference codes 0..3 (bytes of lo) and 4 (hi).
__device__ __forceinline__ void column_table(int32_t qc, int32_t match, int32_t mismatch, uint32_t &lo, uint32_t &hi) {
    lo = 0;
    for (int rc = 0; rc < 4; ++rc) {
        const int32_t s = (rc == 2 || qc == 2) ? 0 : (rc == qc ? match : mismatch);
        lo |= ((uint32_t)s & 0xffu) << (8 * rc);
    }
    hi = (uint32_t)((qc == 2) ? 0 : (qc == 4 ? match : mismatch)) & 0xffu;
}

struct StripGeom {
    int32_t jb, je, rlo, rhi;
};
__device__ __forceinline__ StripGeom strip_geom(int32_t s, int32_t R, int32_t C, int32_t w) {
    StripGeom g;
    g.jb = s * SW + 1;
    g.je = min(g.jb + SW - 1, C);
    g.rlo = max(1, g.jb - w);
    g.rhi = min(min(R, g.je + w), C + w);  // rows past C + w are handled by the finish kernel
    return g;
}

template <bool TB>
__global__ __launch_bounds__(64) void k_sg_dp(const int8_t *__restrict__ seq, const PairDesc *__restrict__ pairs,
                                              const SgTask *__restrict__ tasks, Params P, int32_t *__restrict__ scr,
                                              uint8_t *__restrict__ tb, SgOut *__restrict__ out) {
    const SgTask T = tasks[blockIdx.x];
    const PairDesc pd = pairs[T.pair];
    const int8_t *s1 = seq + pd.off1, *s2 = seq + pd.off2;
    const int lane = threadIdx.x;
    const int32_t R = T.R, C = T.C, w = P.bandwidth;
    const int32_t o1 = P.open, x1 = P.ext, o2 = P.open2, x2 = P.ext2;
    int32_t *bnd[2] = {scr + T.scr, scr + T.scr + (int64_t)(R + 2) * BND};
    int32_t *rowres = scr + T.scr + (int64_t)(R + 2) * BND * 2;
    const int32_t nstrips = (C + SW - 1) / SW;
    // the lane and column that hold column C-1 (for the stale diagonals)
    const int32_t stale_lane = (C >= 2) ? ((C - 2) % SW) / K : -1, stale_c = (C >= 2) ? (C - 2) % K : -1;
    const int32_t stale_strip = (C >= 2) ? (C - 2) / SW : -1;
    int64_t tb_base = T.tb;
    int32_t plo = 1, phi = 0;  // row range of the previous strip
    for (int32_t s = 0; s < nstrips; ++s) {
        const StripGeom g = strip_geom(s, R, C, w);
        if (g.rlo > g.rhi) break;  // band does not reach this strip (nor any later one)
        const int32_t *bin = bnd[s & 1];
        int32_t *bout = bnd[(s + 1) & 1];
        const int32_t jfirst = g.jb + lane * K;
        uint32_t tlo[K], thi[K];
#pragma unroll
        for (int c = 0; c < K; ++c) {
            const int32_t j = jfirst + c;
            column_table(j <= C ? code_at(s2, T.a2, j - 1, T.rev) : 2, P.match, P.mismatch, tlo[c], thi[c]);
        }
        int32_t F1[K], F2[K], Mp[K];
#pragma unroll
        for (int c = 0; c < K; ++c) F1[c] = F2[c] = Mp[c] = NEG;

        // lane-0 input of row r: the previous strip's boundary record, or the
        // left edge of the matrix (M(0,0) = 0, M(r,0) = NEG)
        auto load_in = [&](int32_t r, int32_t &m, int32_t &e1, int32_t &e2, int32_t &rm, int32_t &rj, int32_t &rc) {
            m = NEG;
            e1 = e2 = NEG;
            rm = rj = 0;
            rc = 2;
            if (s == 0) {
                if (r == 0) m = 0;
            } else if (r >= plo && r <= phi) {
                const int32_t *b = bin + (int64_t)r * BND;
                m = __builtin_nontemporal_load(b + 0);
                e1 = __builtin_nontemporal_load(b + 1);
                e2 = __builtin_nontemporal_load(b + 2);
                rm = __builtin_nontemporal_load(b + 3);
                rj = __builtin_nontemporal_load(b + 4);
            }
            if (r >= 1 && r <= R) rc = code_at(s1, T.a1, r - 1, T.rev);
        };
        int32_t cm, ce1, ce2, crm, crj, crc;  // rows rlo + 64q + lane (current chunk)
        int32_t nm, ne1, ne2, nrm, nrj, nrc;  // next chunk
        load_in(g.rlo + lane, cm, ce1, ce2, crm, crj, crc);
        load_in(g.rlo + 64 + lane, nm, ne1, ne2, nrm, nrj, nrc);
        int32_t prev_im;
        {
            int32_t a, b, c2, d, e, f;
            load_in(g.rlo - 1, prev_im, a, b, c2, d, f);
            (void)e;
            if (lane != 0) prev_im = NEG;
        }
        int32_t om = NEG, oe1 = NEG, oe2 = NEG, orm = 0, orj = 0, orc = 2;
        const int32_t nrows = g.rhi - g.rlo + 1, nsteps = nrows + 63;
        for (int32_t t = 0; t < nsteps; ++t) {
            const int32_t q = t & 63;
            if (q == 0 && t > 0) {
                cm = nm; ce1 = ne1; ce2 = ne2; crm = nrm; crj = nrj; crc = nrc;
                load_in(g.rlo + t + 64 + lane, nm, ne1, ne2, nrm, nrj, nrc);
            }
            int32_t im = shr1(om), ie1 = shr1(oe1), ie2 = shr1(oe2), irm = shr1(orm), irj = shr1(orj),
                    irc = shr1(orc);
            {
                const int32_t a = __builtin_amdgcn_readlane(cm, q), b = __builtin_amdgcn_readlane(ce1, q),
                              c2 = __builtin_amdgcn_readlane(ce2, q), d = __builtin_amdgcn_readlane(crm, q),
                              e = __builtin_amdgcn_readlane(crj, q), f = __builtin_amdgcn_readlane(crc, q);
                if (lane == 0) {
                    im = a; ie1 = b; ie2 = c2; irm = d; irj = e; irc = f;
                }
            }
            const int32_t r = g.rlo + t - lane;
            const bool valid = lane <= t && r <= g.rhi;
            const int32_t diag = prev_im;
            prev_im = im;
            uint32_t tword = 0xffffffffu;
            if (valid) {
                const int32_t st = max(r - w, 1), en = min(r + w, C);
                const int32_t lo = st - jfirst, hi = en - jfirst;  // in-band columns: lo <= c < hi; c == hi: last
                const bool hsel = irc >= 4;
                const uint32_t sh = (uint32_t)(irc & 3) * 8u;
                int32_t e1 = ie1, e2 = ie2, rm = irm, rj = irj, dg = diag;
#pragma unroll
                for (int c = 0; c < K; ++c) {
                    const int32_t sc = __builtin_amdgcn_sbfe((int32_t)(hsel ? thi[c] : tlo[c]), sh, 8);
                    const int32_t m0 = sc + dg;
                    dg = Mp[c];
                    const int32_t m = max(max(max(m0, F1[c]), max(e1, F2[c])), e2);
                    const bool inb = c >= lo && c < hi;
                    if (c >= lo && c <= hi && m > rm) {
                        rm = m;
                        rj = jfirst + c;
                    }
                    const int32_t h = m + o1, f = F1[c] + x1, ne1 = e1 + x1;
                    const int32_t h2 = m + o2, f2 = F2[c] + x2, ne2 = e2 + x2;
                    if (TB) {
                        // source: the last of diag, F1, e1, F2, e2 that reaches the maximum
                        int32_t d = (e2 == m) ? 4 : (F2[c] == m) ? 3 : (e1 == m) ? 2 : (F1[c] == m) ? 1 : 0;
                        d |= (f >= h ? 8 : 0) | (ne1 >= h ? 16 : 0) | (f2 >= h2 ? 32 : 0) | (ne2 >= h2 ? 64 : 0);
                        if (inb) tword = (tword & ~(0xffu << (8 * c))) | ((uint32_t)d << (8 * c));
                    }
                    if (inb) {
                        F1[c] = max(f, h);
                        F2[c] = max(f2, h2);
                        e1 = max(ne1, h);
                        e2 = max(ne2, h2);
                    } else {
                        if (c == hi) F1[c] = F2[c] = NEG;  // the band's last column keeps no vertical state
                        e1 = e2 = NEG;
                    }
                    Mp[c] = m;
                }
                om = Mp[K - 1];
                oe1 = e1;
                oe2 = e2;
                orm = rm;
                orj = rj;
                if (s == stale_strip && lane == stale_lane && (r == C + w - 2 || r == C + w - 1))
                    out[blockIdx.x].stale[r - (C + w - 2)] = Mp[stale_c];
            } else {
                om = NEG;
                oe1 = oe2 = NEG;
                orm = orj = 0;
            }
            orc = irc;
            if (TB) *(uint32_t *)(tb + tb_base + (int64_t)t * SW + lane * K) = tword;
            if (lane == 63 && valid) {
                int32_t *b = bout + (int64_t)r * BND;
                b[0] = om;
                b[1] = oe1;
                b[2] = oe2;
                b[3] = orm;
                b[4] = orj;
                rowres[2 * (int64_t)r] = orm;
                rowres[2 * (int64_t)r + 1] = orj;
            }
        }
        if (TB) tb_base += (int64_t)nsteps * SW;
        plo = g.rlo;
        phi = g.rhi;
        __threadfence();  // the boundary records are read back by the next strip
    }
}

// One thread per task: the reference's row loop termination over the row
// results (z-drop, empty row), the maximum, and the traceback (trace tasks).
__device__ __forceinline__ uint32_t trace_byte(const uint8_t *tb, int64_t base, int32_t R, int32_t C, int32_t w,
                                               int32_t ii, int32_t jj, int32_t &cs, int64_t &cbase) {
    const int32_t s = (jj - 1) / SW;
    while (cs > s) {  // walk the strip base back (jj only decreases)
        --cs;
        const StripGeom g = strip_geom(cs, R, C, w);
        cbase -= (int64_t)(g.rhi - g.rlo + 1 + 63) * SW;
    }
    const StripGeom g = strip_geom(s, R, C, w);
    if (ii < g.rlo || ii > g.rhi) return 255u;
    const int32_t within = (jj - 1) % SW, l = within / K, c = within % K;
    return tb[base + cbase + (int64_t)(ii - g.rlo + l) * SW + l * K + c];
}

__global__ void k_sg_finish(const int8_t *__restrict__ seq, const PairDesc *__restrict__ pairs,
                            const SgTask *__restrict__ tasks, int32_t n, Params P, const int32_t *__restrict__ scr,
                            const uint8_t *__restrict__ tb, SgOut *__restrict__ out, uint32_t *__restrict__ cigar) {
    const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const SgTask T = tasks[k];
    const PairDesc pd = pairs[T.pair];
    SgOut &o = out[k];
    const int32_t R = T.R, C = T.C, w = P.bandwidth;
    const int32_t *rowres = scr + T.scr + (int64_t)(R + 2) * BND * 2;
    int32_t best = 0, e1 = -1, e2 = -1, brk = R;
    int64_t cells = 0;
    for (int32_t i = 1; i <= R; ++i) {
        const int32_t st = max(i - w, 1), en = min(i + w, C);
        cells += (int64_t)max(0, en - st) + 1;
        int32_t rm, rj;
        if (i <= C + w) {
            rm = rowres[2 * (int64_t)i];
            rj = rowres[2 * (int64_t)i + 1];
        } else {
            // the band has left the matrix: only column C, whose diagonal is
            // the value the row buffer of this parity last held at C-1
            const int32_t rc = code_at(seq + pd.off1, T.a1, i - 1, T.rev);
            const int32_t qc = code_at(seq + pd.off2, T.a2, C - 1, T.rev);
            const int32_t sc = (rc == 2 || qc == 2) ? 0 : (rc == qc ? P.match : P.mismatch);
            const int32_t dg = (C >= 2) ? o.stale[(i - (C + w + 1)) & 1] : NEG;
            const int32_t m = sc + dg;
            rm = m > 0 ? m : 0;
            rj = m > 0 ? C : 0;
        }
        if (rm > best) {
            best = rm;
            e1 = i;
            e2 = rj;
        }
        int32_t l = (i - e1) - (rj - e2);
        l = l > 0 ? l : -l;
        if (best - rm > P.zdrop + l * P.ext2 || rj <= 0) {
            brk = i;
            break;
        }
    }
    o.max = best;
    o.e1 = e1;
    o.e2 = e2;
    o.brk = brk;
    o.cells = cells;
    o.ncig = 0;
    if (T.tb < 0) return;
    // traceback (impl/sequenceAlignment3.cpp:864-915); runs are written in
    // traceback order and reversed by the host
    uint32_t *cg = cigar + o.cig_off;
    int32_t nc = 0;
    auto push = [&](uint32_t op) {
        if (nc == 0 || op != (cg[nc - 1] & 0xfu))
            cg[nc++] = (1u << 4) | op;
        else
            cg[nc - 1] += 1u << 4;
    };
    const int32_t nstrips = (C + SW - 1) / SW;
    int32_t cs = nstrips;
    int64_t cbase = 0;
    for (int32_t s = 0; s < nstrips; ++s) {
        const StripGeom g = strip_geom(s, R, C, w);
        if (g.rlo > g.rhi) {
            cs = s;
            break;
        }
        cbase += (int64_t)(g.rhi - g.rlo + 1 + 63) * SW;
    }
    int32_t ii = e1, jj = e2, state = 0;
    while (ii > 0 && jj > 0) {
        const int32_t tmp = (int32_t)trace_byte(tb, T.tb, R, C, w, ii, jj, cs, cbase);
        if (state == 0)
            state = tmp & 7;
        else if (!((tmp >> (state + 2)) & 1))
            state = 0;
        if (state == 0) state = tmp & 7;
        if (state == 0) {
            push(0);
            --ii;
            --jj;
        } else if (state == 1 || state == 3) {
            push(2);
            --ii;
        } else {
            push(1);
            --jj;
        }
    }
    while (ii > 0) {
        push(2);
        --ii;
    }
    while (jj > 0) {
        push(1);
        --jj;
    }
    o.ncig = nc;
}

}  // namespace

int64_t sg_trace_bytes(int32_t R, int32_t C, int32_t w) {
    int64_t total = 0;
    for (int32_t s = 0; s * SW < C; ++s) {
        const int32_t jb = s * SW + 1, je = std::min(jb + SW - 1, C);
        const int32_t rlo = std::max(1, jb - w), rhi = std::min(std::min(R, je + w), C + w);
        if (rlo > rhi) break;
        total += (int64_t)(rhi - rlo + 1 + 63) * SW;
    }
    return total;
}

int64_t sg_scratch_ints(int32_t R) { return (int64_t)(R + 2) * (2 * BND + 2); }

hipError_t launch_sg_dp(hipStream_t st, bool trace, const int8_t *seq, const PairDesc *pairs, const SgTask *tasks,
                        int32_t n, const Params &P, int32_t *scr, uint8_t *tb, SgOut *out) {
    if (n <= 0) return hipSuccess;
    if (trace)
        hipLaunchKernelGGL(k_sg_dp<true>, dim3(n), dim3(64), 0, st, seq, pairs, tasks, P, scr, tb, out);
    else
        hipLaunchKernelGGL(k_sg_dp<false>, dim3(n), dim3(64), 0, st, seq, pairs, tasks, P, scr, tb, out);
    return hipGetLastError();
}

hipError_t launch_sg_finish(hipStream_t st, const int8_t *seq, const PairDesc *pairs, const SgTask *tasks, int32_t n,
                            const Params &P, const int32_t *scr, const uint8_t *tb, SgOut *out, uint32_t *cigar) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sg_finish, dim3((n + 63) / 64), dim3(64), 0, st, seq, pairs, tasks, n, P, scr, tb, out,
                       cigar);
    return hipGetLastError();
}

}  // namespace dcns
