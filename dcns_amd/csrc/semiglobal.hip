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
// Device layout: one workgroup of up to 16 waves per task.  Columns are cut
// into strips of 64 lanes x 4 columns, strip s on wave s % nw; inside a strip
// lane l works on row r = t - l at step t (an anti-diagonal skew), so the
// values a row carries from left to right (M, the two horizontal gap states,
// the running row maximum, the reference base) move one lane per step by a
// DPP wave shift.  The strip's last lane hands each row's carried values to
// the next strip through a per-task boundary array; the next strip's first
// lane reads them 64 rows at a time, a barrier-separated phase later.  Every row in the band is
// computed (the reference's termination test usually never fires with the
// default long-gap piece, whose column state keeps every later row within
// open2 of the best score); the finish kernel cuts the rows at the reference's
// termination row.  Trace bytes are stored step-major (64 lanes x 4 bytes per
// step, one coalesced 256-byte store), 255 outside the band as after
// Matrix::reset.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

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

// Lane-0 input of row r of a strip: the previous strip's boundary record,
// or the left edge of the matrix (M(0,0) = 0, M(r,0) = NEG); and the
// reference code of row r.
struct StripIn {
    const int32_t *bin;
    int32_t first, plo, phi, R, a1, rev;
    const int8_t *s1;
    __device__ __forceinline__ void load(int32_t r, int32_t &m, int32_t &e1, int32_t &e2, int32_t &rm, int32_t &rj,
                                         int32_t &rc) const {
        m = NEG;
        e1 = e2 = NEG;
        rm = rj = 0;
        rc = 2;
        if (first) {
            if (r == 0) m = 0;
        } else if (r >= plo && r <= phi) {
            const int32_t *b = bin + (int64_t)r * BND;
            m = b[0];
            e1 = b[1];
            e2 = b[2];
            rm = b[3];
            rj = b[4];
        }
        if (r >= 1 && r <= R) rc = code_at(s1, a1, r - 1, rev);
    }
};

// One workgroup of nw waves per task.  Strip s runs on wave s % nw, starting
// at the phase the host's schedule gives it; a phase is 64 steps of every
// active wave followed by a workgroup barrier, which publishes the strip
// boundaries written in the phase.  The schedule starts each strip late
// enough that every boundary row its first lane loads at the start of a phase
// was written in an earlier phase, and a wave takes its next strip only after
// its current one, so nw + 1 boundary arrays suffice (strip s reads array
// s % (nw+1) and writes (s+1) % (nw+1)).
template <bool TB>
__global__ __launch_bounds__(1024) void k_sg_dp(const int8_t *__restrict__ seq, const PairDesc *__restrict__ pairs,
                                                const SgTask *__restrict__ tasks, const SgStrip *__restrict__ strips,
                                                Params P, int32_t nw, int32_t *__restrict__ scr,
                                                uint8_t *__restrict__ tb, SgOut *__restrict__ out) {
    const SgTask T = tasks[blockIdx.x];
    const PairDesc pd = pairs[T.pair];
    const int8_t *s1 = seq + pd.off1, *s2 = seq + pd.off2;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int32_t R = T.R, C = T.C, w = P.bandwidth;
    const int32_t o1 = P.open, x1 = P.ext, o2 = P.open2, x2 = P.ext2;
    const int64_t bsz = (int64_t)(R + 2) * BND;
    int32_t *rowres = scr + T.scr + bsz * (nw + 1);
    const SgStrip *sched = strips + T.sched;
    // the lane and column that hold column C-1 (for the stale diagonals)
    const int32_t stale_lane = (C >= 2) ? ((C - 2) % SW) / K : -1, stale_c = (C >= 2) ? (C - 2) % K : -1;
    const int32_t stale_strip = (C >= 2) ? (C - 2) / SW : -1;

    int32_t cs = wv;  // this wave's current strip
    uint32_t tlo[K], thi[K];
    int32_t F1[K], F2[K], Mp[K];
    int32_t prev_im = NEG, om = NEG, oe1 = NEG, oe2 = NEG, orm = 0, orj = 0, orc = 2;
    for (int32_t ph = 0; ph < T.phases; ++ph) {
        if (cs < T.nstrips && ph >= sched[cs].start) {
            const SgStrip S = sched[cs];
            const int32_t t0 = (ph - S.start) * 64, nsteps = S.rhi - S.rlo + 1 + 63;
            const int32_t jb = cs * SW + 1, jfirst = jb + lane * K;
            StripIn in;
            in.bin = scr + T.scr + bsz * (cs % (nw + 1));
            in.first = cs == 0;
            in.plo = cs > 0 ? sched[cs - 1].rlo : 1;
            in.phi = cs > 0 ? sched[cs - 1].rhi : 0;
            in.R = R;
            in.a1 = T.a1;
            in.rev = T.rev;
            in.s1 = s1;
            int32_t *bout = scr + T.scr + bsz * ((cs + 1) % (nw + 1));
            if (t0 == 0) {  // a new strip: column tables, empty state above row rlo
#pragma unroll
                for (int c = 0; c < K; ++c) {
                    const int32_t j = jfirst + c;
                    column_table(j <= C ? code_at(s2, T.a2, j - 1, T.rev) : 2, P.match, P.mismatch, tlo[c], thi[c]);
                    F1[c] = F2[c] = Mp[c] = NEG;
                }
                int32_t a, b, c2, d, f;
                in.load(S.rlo - 1, prev_im, a, b, c2, d, f);
                if (lane != 0) prev_im = NEG;
                om = oe1 = oe2 = NEG;
                orm = orj = 0;
                orc = 2;
            }
            int32_t cm, ce1, ce2, crm, crj, crc;  // rows rlo + t0 + lane
            in.load(S.rlo + t0 + lane, cm, ce1, ce2, crm, crj, crc);
            const int32_t t1 = min(t0 + 64, nsteps);
            for (int32_t t = t0; t < t1; ++t) {
                const int32_t q = t - t0;
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
                const int32_t r = S.rlo + t - lane;
                const bool valid = lane <= t && r <= S.rhi;
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
                    if (cs == stale_strip && lane == stale_lane && (r == C + w - 2 || r == C + w - 1))
                        out[blockIdx.x].stale[r - (C + w - 2)] = Mp[stale_c];
                } else {
                    om = NEG;
                    oe1 = oe2 = NEG;
                    orm = orj = 0;
                }
                orc = irc;
                if (TB) *(uint32_t *)(tb + T.tb + S.tb + (int64_t)t * SW + lane * K) = tword;
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
            if (t1 >= nsteps) cs += nw;
        }
        __syncthreads();
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
                            const SgTask *__restrict__ tasks, int32_t n, Params P, int32_t nw,
                            const int32_t *__restrict__ scr, const uint8_t *__restrict__ tb, SgOut *__restrict__ out,
                            uint32_t *__restrict__ cigar) {
    const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const SgTask T = tasks[k];
    const PairDesc pd = pairs[T.pair];
    SgOut &o = out[k];
    const int32_t R = T.R, C = T.C, w = P.bandwidth;
    const int32_t *rowres = scr + T.scr + (int64_t)(R + 2) * BND * (nw + 1);
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

int64_t sg_scratch_ints(int32_t R, int32_t nw) { return (int64_t)(R + 2) * ((nw + 1) * BND + 2); }

// The strip schedule of one task for nw waves (see k_sg_dp).  A strip may
// start once (a) its predecessor has run floor((126 + d) / 64) + 1 phases,
// d = the rise of the first row between the two strips, so that the rows its
// first lane loads at the start of any phase (64 rows ahead, first-lane rows
// lagging the producer's last lane by 63 steps) were all written in earlier
// phases, and (b) its wave has finished the strip nw places before it.
int32_t sg_schedule(int32_t R, int32_t C, int32_t w, int32_t nw, std::vector<SgStrip> &out) {
    out.clear();
    int32_t phases = 0;
    std::vector<int32_t> len;
    for (int32_t s = 0; (int64_t)s * SW < C; ++s) {
        const int32_t jb = s * SW + 1, je = std::min(jb + SW - 1, C);
        const int32_t rlo = std::max(1, jb - w), rhi = std::min(std::min(R, je + w), C + w);
        if (rlo > rhi) break;
        SgStrip g{};
        g.rlo = rlo;
        g.rhi = rhi;
        g.tb = s == 0 ? 0 : out.back().tb + (int64_t)(out.back().rhi - out.back().rlo + 1 + 63) * SW;
        const int32_t p = (rhi - rlo + 1 + 63 + 63) / 64;
        int32_t start = 0;
        if (s > 0) {
            const int32_t d = std::max(0, rlo - out[(size_t)s - 1].rlo);
            start = out[(size_t)s - 1].start + (126 + d) / 64 + 1;
        }
        if (s >= nw) start = std::max(start, out[(size_t)(s - nw)].start + len[(size_t)(s - nw)]);
        g.start = start;
        out.push_back(g);
        len.push_back(p);
        phases = std::max(phases, start + p);
    }
    return phases;
}

hipError_t launch_sg_dp(hipStream_t st, bool trace, const int8_t *seq, const PairDesc *pairs, const SgTask *tasks,
                        const SgStrip *strips, int32_t n, int32_t nw, const Params &P, int32_t *scr, uint8_t *tb,
                        SgOut *out) {
    if (n <= 0) return hipSuccess;
    if (nw < 1 || nw > 16) return hipErrorInvalidValue;
    if (trace)
        hipLaunchKernelGGL(k_sg_dp<true>, dim3(n), dim3(64 * nw), 0, st, seq, pairs, tasks, strips, P, nw, scr, tb,
                           out);
    else
        hipLaunchKernelGGL(k_sg_dp<false>, dim3(n), dim3(64 * nw), 0, st, seq, pairs, tasks, strips, P, nw, scr, tb,
                           out);
    return hipGetLastError();
}

hipError_t launch_sg_finish(hipStream_t st, const int8_t *seq, const PairDesc *pairs, const SgTask *tasks, int32_t n,
                            int32_t nw, const Params &P, const int32_t *scr, const uint8_t *tb, SgOut *out,
                            uint32_t *cigar) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sg_finish, dim3((n + 63) / 64), dim3(64), 0, st, seq, pairs, tasks, n, P, nw, scr, tb, out,
                       cigar);
    return hipGetLastError();
}

}  // namespace dcns
