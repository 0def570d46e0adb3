// xdrop.hip -- banded x-drop extension of dCNS seeds on gfx950.
//
// Reference: x_extend_seed (findSimilarFragmentsForPairedSequence4.cpp:245-285)
// runs SemiGlobal_xextend forward, score only (sequenceAlignment3.cpp:1234-1314),
// from the seed start + 1, then backwards on reverse complements with
// traceback (:1318-1451).  A row's band is j in [max(1,i-w), min(i+w,L2)];
// a pass restarts with 2w when the row maximum sits on a band edge.
//
// One lane per seed while the band is at most 21 cells, then (the tasks
// whose band doubled) one lane per seed at 41 cells.  The band is held
// band-relative in VGPRs: in band coordinates k = j - (i - w) the diagonal
// predecessor of (i, k) is (i-1, k) and the vertical one is (i-1, k+1), so
// rows advance without moving data.  Query bases ride in a byte window packed
// four per VGPR that shifts by one byte per row.  Two passes:
//   pass 1: forward + reverse, score only; reports the final band of each
//           direction and the rows the final reverse pass evaluated;
//   pass 2: the final reverse pass again, storing one traceback byte per
//           band cell (rows x band, default 255 like Matrix::reset), then the
//           traceback state machine writes the cigar.
// A task whose band outgrows the lane kernel's (a row maximum on the band
// edge, so the pass restarts at twice the width) is listed for the next
// kernel: the 41-cell lane kernel, then the wave kernels: one
// wave per task, the band's cells across the lanes, 192 cells at most, which
// keeps a long extension at a wide band from holding a launch for the
// length of one lane's walk across every cell.  Past 192 cells a kernel keeps
// the band in global scratch.
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include "dp_common.h"
#include "kernels.h"

namespace dcns {
namespace {

struct Dir {
    const int8_t *s1;  // forward: s1[x] = seq1[base1 + x]; reverse: 4 - seq1[base1 - x]
    const int8_t *s2;
    int32_t L1, L2;
    bool rev;
    __device__ __forceinline__ int c1(int x) const { return rev ? 4 - s1[-x] : s1[x]; }
    __device__ __forceinline__ int c2(int x) const { return rev ? 4 - s2[-x] : s2[x]; }
};

__device__ __forceinline__ int32_t score(int a, int b, int32_t match, int32_t mism) {
    return (a == 2 || b == 2) ? 0 : (a == b ? match : mism);
}

struct PassOut {
    int32_t max, e1, e2;  // 1-based end (0 when none)
    int32_t rows;         // rows evaluated
    int64_t cells;
    bool restart;
    bool stopped;         // the row budget ran out before the pass ended
};

// One pass at half-width w (2w+1 <= BMAX).  With TB, row i's traceback bytes
// are written to tb[i * BP + k] (k in band coordinates, 255 outside the band).
// A pass that would evaluate a row past max_rows stops there (stopped).
template <int BMAX, bool TB>
__device__ PassOut xpass(const Dir &D, int32_t w, const Params &P, uint8_t *__restrict__ tb, int32_t BP,
                         int32_t max_rows = INT32_MAX) {
    constexpr int NQ = BMAX / 4 + 2;
    int32_t Mp[BMAX], Fp[BMAX];
    uint32_t qw[NQ];
#pragma unroll
    for (int k = 0; k < BMAX; ++k) {
        Mp[k] = kOutOfBand;
        Fp[k] = kOutOfBand;
    }
    // query window for row 1: columns j = 1 - w + k, code of column j is c2(j-1)
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        uint32_t v = 0;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int j = 1 - w + 4 * q + b;
            const int code = (j >= 1 && j <= D.L2) ? D.c2(j - 1) : 2;
            v |= (uint32_t)(code & 0xff) << (8 * b);
        }
        qw[q] = v;
    }
    PassOut r{0, 0, 0, 0, 0, false, false};
    const int32_t O = P.open, X = P.ext;
    for (int32_t i = 1; i <= D.L1; ++i) {
        if (i > max_rows) {
            r.stopped = true;
            return r;
        }
        const int32_t st = max(i - w, 1);
        const int32_t en = min(i + w, D.L2);
        const int32_t lo = i - w;  // column of band slot 0
        const int c = D.c1(i - 1);
        int32_t e = kOutOfBand;
        int32_t rowMax = 0, rowJ = -1;
        uint32_t tw[(BMAX + 3) / 4];
#pragma unroll
        for (int q = 0; q < (BMAX + 3) / 4; ++q) tw[q] = 0xffffffffu;
        if (st <= en) r.cells += en - st + 1;
#pragma unroll
        for (int k = 0; k < BMAX; ++k) {
            const int32_t j = lo + k;
            const bool valid = (j >= st) && (j <= en);
            const int qk = (qw[k >> 2] >> (8 * (k & 3))) & 0xff;
            const int32_t s = score(c, qk, P.match, P.mismatch);
            const int32_t diag = (j == 1) ? (i == 1 ? 0 : kOutOfBand) : Mp[k];
            const int32_t fv = (k + 1 < BMAX && j != en) ? Fp[k + 1 < BMAX ? k + 1 : k] : kOutOfBand;
            int32_t m = s + diag;
            uint32_t d = 0;
            if (!(m > fv)) { m = fv; d = 1; }
            if (!(m > e)) { m = e; d = 2; }
            const int32_t h = m + O;
            int32_t f = fv + X;
            if (f >= h) d |= 8; else f = h;
            int32_t en2 = e + X;
            if (en2 >= h) d |= 16; else en2 = h;
            if (valid) {
                if (m > rowMax) { rowMax = m; rowJ = j; }
                e = en2;
                Mp[k] = m;
                Fp[k] = f;
                if (TB) tw[k >> 2] = (tw[k >> 2] & ~(0xffu << (8 * (k & 3)))) | (d << (8 * (k & 3)));
            }
        }
        if (TB) {
            uint32_t *row = (uint32_t *)(tb + (int64_t)i * BP);
#pragma unroll
            for (int q = 0; q < (BMAX + 3) / 4; ++q)
                if (4 * q < BP) row[q] = tw[q];
        }
        r.rows = i;
        if (rowJ != D.L2 && rowJ != 1 && (rowJ == st || rowJ == en)) {
            r.restart = true;
            return r;
        }
        if (rowMax > r.max) { r.max = rowMax; r.e1 = i; r.e2 = rowJ; }
        if (r.max - rowMax > P.xdrop) break;
        if (rowJ <= 0) break;
        // shift the query window one column right for row i+1
        const int jn = lo + 4 * NQ;  // column entering slot 4*NQ-1 of row i+1
        const int code = (jn >= 1 && jn <= D.L2) ? D.c2(jn - 1) : 2;
#pragma unroll
        for (int q = 0; q < NQ - 1; ++q) qw[q] = (qw[q] >> 8) | (qw[q + 1] << 24);
        qw[NQ - 1] = (qw[NQ - 1] >> 8) | ((uint32_t)(code & 0xff) << 24);
        // band slot k of row i+1 is column j+1: diagonal = Mp[k], vertical = Fp[k+1]
    }
    return r;
}

// The same pass for any band width, with the band-relative rows in
// per-task global scratch (Mp, Fp of 2w+1 entries each) instead of VGPRs.
// Used only for the tasks whose band outgrows the widest register kernel.
template <bool TB>
__device__ PassOut xpass_mem(const Dir &D, int32_t w, const Params &P, uint8_t *__restrict__ tb, int32_t BP,
                             int32_t *__restrict__ Mp, int32_t *__restrict__ Fp) {
    const int32_t B = 2 * w + 1;
    for (int32_t k = 0; k < B; ++k) {
        Mp[k] = kOutOfBand;
        Fp[k] = kOutOfBand;
    }
    PassOut r{0, 0, 0, 0, 0, false, false};
    const int32_t O = P.open, X = P.ext;
    for (int32_t i = 1; i <= D.L1; ++i) {
        const int32_t st = max(i - w, 1);
        const int32_t en = min(i + w, D.L2);
        const int32_t lo = i - w;
        const int c = D.c1(i - 1);
        int32_t e = kOutOfBand;
        int32_t rowMax = 0, rowJ = -1;
        uint8_t *row = TB ? tb + (int64_t)i * BP : nullptr;
        if (st <= en) r.cells += en - st + 1;
        for (int32_t k = 0; k < B; ++k) {
            const int32_t j = lo + k;
            const bool valid = (j >= st) && (j <= en);
            const int qk = (j >= 1 && j <= D.L2) ? D.c2(j - 1) : 2;
            const int32_t s = score(c, qk, P.match, P.mismatch);
            const int32_t diag = (j == 1) ? (i == 1 ? 0 : kOutOfBand) : Mp[k];
            const int32_t fv = (k + 1 < B && j != en) ? Fp[k + 1] : kOutOfBand;
            int32_t m = s + diag;
            uint32_t d = 0;
            if (!(m > fv)) { m = fv; d = 1; }
            if (!(m > e)) { m = e; d = 2; }
            const int32_t h = m + O;
            int32_t f = fv + X;
            if (f >= h) d |= 8; else f = h;
            int32_t en2 = e + X;
            if (en2 >= h) d |= 16; else en2 = h;
            if (valid) {
                if (m > rowMax) { rowMax = m; rowJ = j; }
                e = en2;
                Mp[k] = m;
                Fp[k] = f;
            }
            if (TB) row[k] = valid ? (uint8_t)d : (uint8_t)255;
        }
        if (TB)
            for (int32_t k = B; k < BP; ++k) row[k] = 255;
        r.rows = i;
        if (rowJ != D.L2 && rowJ != 1 && (rowJ == st || rowJ == en)) {
            r.restart = true;
            return r;
        }
        if (rowMax > r.max) { r.max = rowMax; r.e1 = i; r.e2 = rowJ; }
        if (r.max - rowMax > P.xdrop) break;
        if (rowJ <= 0) break;
    }
    return r;
}

// Passes at w, 2w, ... until one ends without a restart.  too_wide: the next
// band is wider than BMAX, or (stopped) a pass ran out of its max_rows
// budget; the pending pass is then at band w, and a stopped pass's cells are
// not counted (the kernel that takes the task over runs it again).
template <int BMAX, bool TB>
__device__ PassOut xrun(const Dir &D, int32_t &w, const Params &P, uint8_t *tb, int32_t BP, int64_t &cells,
                        bool &too_wide, bool &stopped, int32_t max_rows) {
    too_wide = false;
    stopped = false;
    while (true) {
        if (2 * w + 1 > BMAX) {
            too_wide = true;
            return PassOut{0, 0, 0, 0, 0, false, false};
        }
        PassOut r = xpass<BMAX, TB>(D, w, P, tb, BP, max_rows);
        if (r.stopped) {
            too_wide = stopped = true;
            return r;
        }
        cells += r.cells;
        if (!r.restart) return r;
        w *= 2;
    }
}

// ---------------------------------------------------------------------------
// Wave passes: one task per wave, for the bands past the lane kernel's 21
// cells.  A row's band slots lie across the wave, S consecutive slots per
// lane (slot k in lane k / S).  The diagonal predecessor of slot k is slot k
// of the previous row (same register); the vertical one, slot k + 1, comes
// from the next lane by DPP wave_shl:1 for a lane's last slot.  The
// horizontal gap is a max-plus recurrence along the row,
//   e(k + 1) = max(e(k) + D, m'(k) + open),  D = max(ext, open),
// where m'(k) is the cell's value before the horizontal term (a cell's
// horizontal output max(e + ext, max(m', e) + open) equals that, as
// max(e + ext, e + open) = e + D), so the wave evaluates it as a prefix scan:
// each lane folds its slots, DPP row_shr 1/2/4/8 and row_bcast 15/31 combine
// the lanes, and each lane re-runs its slots from its incoming value.  The
// row's band starts with e = kOutOfBand at its first valid slot, entered as
// a term at the slot before it.  Values, traceback bits, the row maximum
// (first column on ties) and every stop rule are those of xpass.
// ---------------------------------------------------------------------------
constexpr int kWaveSlots = 3;  // slots per lane of the widest wave band
static_assert(64 * kWaveSlots == kXdropWaveBand, "wave band");
constexpr int32_t kScanNeg = -(1 << 30);  // below every value of a row, decays included

template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ int32_t dpp(int32_t old, int32_t src) {
    return __builtin_amdgcn_update_dpp(old, src, CTRL, ROWS, 0xf, false);
}

// y(l) = max over l' <= l of x(l') + (l - l') * step, across the wave
__device__ __forceinline__ int32_t wave_maxplus_scan(int32_t x, int32_t step, int lane) {
    int32_t y = x;
    y = max(y, dpp<0x111>(kScanNeg, y) + step);      // row_shr:1
    y = max(y, dpp<0x112>(kScanNeg, y) + 2 * step);  // row_shr:2
    y = max(y, dpp<0x114>(kScanNeg, y) + 4 * step);  // row_shr:4
    y = max(y, dpp<0x118>(kScanNeg, y) + 8 * step);  // row_shr:8
    y = max(y, dpp<0x142, 0xa>(kScanNeg, y) + ((lane & 15) + 1) * step);  // row_bcast:15 into rows 1, 3
    y = max(y, dpp<0x143, 0xc>(kScanNeg, y) + (lane - 31) * step);        // row_bcast:31 into rows 2, 3
    return y;
}

__device__ __forceinline__ int32_t wave_max(int32_t v) {
    v = max(v, dpp<0x111>(INT32_MIN, v));
    v = max(v, dpp<0x112>(INT32_MIN, v));
    v = max(v, dpp<0x114>(INT32_MIN, v));
    v = max(v, dpp<0x118>(INT32_MIN, v));
    v = max(v, dpp<0x142, 0xa>(INT32_MIN, v));
    v = max(v, dpp<0x143, 0xc>(INT32_MIN, v));
    return __builtin_amdgcn_readlane(v, 63);
}

template <int S, bool TB>
__device__ PassOut wave_xpass(const Dir &D, int32_t w, const Params &P, uint8_t *__restrict__ tb, int32_t BP) {
    const int lane = threadIdx.x;
    const int32_t O = P.open, X = P.ext, DK = max(P.ext, P.open);
    int32_t Mp[S], Fp[S], qk[S];
#pragma unroll
    for (int u = 0; u < S; ++u) {
        Mp[u] = kOutOfBand;
        Fp[u] = kOutOfBand;
        const int32_t j = 1 - w + lane * S + u;  // column of slot k in row 1
        qk[u] = (j >= 1 && j <= D.L2) ? D.c2(j - 1) : 2;
    }
    // 64 rows at a time, one per lane: the row's reference base, and the
    // query base entering the band's last slot for the next row (column
    // i - w + 64 S); the next 64 in flight while these are used
    auto chunk = [&](int32_t i0, int32_t &cb, int32_t &qb) {
        const int32_t ri = i0 + lane;
        cb = ri <= D.L1 ? D.c1(ri - 1) : 2;
        const int32_t jn = ri - w + 64 * S;
        qb = (jn >= 1 && jn <= D.L2) ? D.c2(jn - 1) : 2;
    };
    int32_t ccur, qcur, cnext, qnext;
    chunk(1, ccur, qcur);
    chunk(65, cnext, qnext);
    PassOut r{0, 0, 0, 0, 0, false, false};
    for (int32_t i = 1; i <= D.L1; ++i) {
        const int t = (i - 1) & 63;
        if (t == 0 && i > 1) {
            ccur = cnext;
            qcur = qnext;
            chunk(i + 64, cnext, qnext);
        }
        const int c = __builtin_amdgcn_readlane(ccur, t);
        const int32_t st = max(i - w, 1), en = min(i + w, D.L2), lo = i - w;
        const int32_t kst = st - lo;  // first valid slot
        if (st <= en) r.cells += en - st + 1;
        const int32_t fnext = dpp<0x130>(kOutOfBand, Fp[0]);  // wave_shl:1: the next lane's first slot
        int32_t A[S], d0[S], fv[S];
        int32_t L = kScanNeg;  // the lane's slots folded from an empty input
#pragma unroll
        for (int u = 0; u < S; ++u) {
            const int32_t k = lane * S + u, j = lo + k;
            const bool valid = (j >= st) && (j <= en);
            const int32_t sc = score(c, qk[u], P.match, P.mismatch);
            const int32_t diag = (j == 1) ? (i == 1 ? 0 : kOutOfBand) : Mp[u];
            const int32_t fr = (u + 1 < S) ? Fp[u + 1 < S ? u + 1 : u] : fnext;
            fv[u] = (j != en) ? fr : kOutOfBand;
            int32_t m = sc + diag;
            int32_t d = 0;
            if (!(m > fv[u])) {
                m = fv[u];
                d = 1;
            }
            A[u] = m;
            d0[u] = d;
            L = max(L + DK, valid ? m + O : (k == kst - 1 ? kOutOfBand : kScanNeg));
        }
        const int32_t y = wave_maxplus_scan(L, S * DK, lane);
        int32_t e = max(dpp<0x138>(kScanNeg, y), (kst == 0 ? kOutOfBand : kScanNeg) + lane * S * DK);  // wave_shr:1
        int32_t lm = 0, lj = -1;
#pragma unroll
        for (int u = 0; u < S; ++u) {
            const int32_t k = lane * S + u, j = lo + k;
            const bool valid = (j >= st) && (j <= en);
            int32_t m = A[u];
            uint32_t d = (uint32_t)d0[u];
            if (!(m > e)) {
                m = e;
                d = 2;
            }
            const int32_t h = m + O;
            int32_t f = fv[u] + X;
            if (f >= h) d |= 8;
            else f = h;
            if (e + X >= h) d |= 16;
            e = max(e + DK, valid ? A[u] + O : (k == kst - 1 ? kOutOfBand : kScanNeg));
            if (valid) {
                if (m > lm) {
                    lm = m;
                    lj = j;
                }
                Mp[u] = m;
                Fp[u] = f;
            }
            if (TB && k < BP) tb[(int64_t)i * BP + k] = valid ? (uint8_t)d : (uint8_t)255;
        }
        const int32_t rowMax = wave_max(lm);
        int32_t rowJ = -1;
        if (rowMax > 0) {
            const uint64_t bal = __ballot(lm == rowMax);
            rowJ = __builtin_amdgcn_readlane(lj, (int)__builtin_ctzll(bal));
        }
        r.rows = i;
        if (rowJ != D.L2 && rowJ != 1 && (rowJ == st || rowJ == en)) {
            r.restart = true;
            return r;
        }
        if (rowMax > r.max) {
            r.max = rowMax;
            r.e1 = i;
            r.e2 = rowJ;
        }
        if (r.max - rowMax > P.xdrop) break;
        if (rowJ <= 0) break;
        // row i + 1: slot k holds column j + 1, the column of slot k + 1 now
        const int32_t q0 = dpp<0x130>(__builtin_amdgcn_readlane(qcur, t), qk[0]);  // wave_shl:1
#pragma unroll
        for (int u = 0; u + 1 < S; ++u) qk[u] = qk[u + 1];
        qk[S - 1] = q0;
    }
    return r;
}

// The narrowest slot count that holds the band (fewer slots: a shorter row)
template <bool TB>
__device__ PassOut wave_xpass_fit(const Dir &D, int32_t w, const Params &P, uint8_t *__restrict__ tb, int32_t BP) {
    const int32_t B = 2 * w + 1;
    if (B <= 64) return wave_xpass<1, TB>(D, w, P, tb, BP);
    if (B <= 128) return wave_xpass<2, TB>(D, w, P, tb, BP);
    return wave_xpass<kWaveSlots, TB>(D, w, P, tb, BP);
}

__device__ PassOut wave_xrun(const Dir &D, int32_t &w, const Params &P, int64_t &cells, bool &too_wide) {
    too_wide = false;
    while (true) {
        if (2 * w + 1 > kXdropWaveBand) {
            too_wide = true;
            return PassOut{0, 0, 0, 0, 0, false, false};
        }
        PassOut r = wave_xpass_fit<false>(D, w, P, nullptr, 0);
        cells += r.cells;
        if (!r.restart) return r;
        w *= 2;
    }
}

__device__ Dir forward_dir(const int8_t *seq, const PairDesc &pd, const XTask &t) {
    Dir D;
    D.s1 = seq + pd.off1 + t.start1;
    D.s2 = seq + pd.off2 + t.start2;
    D.L1 = pd.len1 - t.start1;
    D.L2 = pd.len2 - t.start2;
    D.rev = false;
    return D;
}

// Reverse complements ending at (start1 + fe1, start2 + fe2), lengths
// start+1+fe (:268-271).
__device__ Dir reverse_dir(const int8_t *seq, const PairDesc &pd, const XTask &t, int32_t fe1, int32_t fe2) {
    Dir D;
    D.s1 = seq + pd.off1 + t.start1 + fe1;
    D.s2 = seq + pd.off2 + t.start2 + fe2;
    D.L1 = t.start1 + 1 + fe1;
    D.L2 = t.start2 + 1 + fe2;
    D.rev = true;
    return D;
}

// Pass 1, one lane per task, for the tasks in_list[0, n) (all tasks when
// in_list is null); tasks whose band outgrows BMAX, or whose pass runs past
// max_rows rows (a lane walks a row per step, so one long extension would
// hold the whole wave), are listed in list for the next kernel.  A task a
// narrower lane kernel stopped for its length is listed on at once.
template <int BMAX>
__global__ __launch_bounds__(64) void k_xdrop_pass1(const int8_t *__restrict__ seq, const PairDesc *__restrict__ pairs,
                                                    const XTask *__restrict__ tasks,
                                                    const int32_t *__restrict__ in_list, int32_t n, Params P,
                                                    XOut *__restrict__ out, int32_t *__restrict__ pending,
                                                    int32_t *__restrict__ list, int32_t max_rows) {
    const int lane_gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (lane_gid >= n) return;
    const int gid = in_list ? in_list[lane_gid] : lane_gid;
    const XTask t = tasks[gid];
    XOut o = out[gid];
    if (o.status == 0 && o.fw != 0) return;  // already finished by a narrower kernel
    if (o.lane_stop) {
        list[atomicAdd(pending, 1)] = gid;
        return;
    }
    const PairDesc pd = pairs[t.pair];
    bool wide, stopped;
    if (o.status != 2) {  // forward pass still to do
        int32_t w = o.fw > 0 ? o.fw : P.w;
        const Dir F = forward_dir(seq, pd, t);
        PassOut f = xrun<BMAX, false>(F, w, P, nullptr, 0, o.cells, wide, stopped, max_rows);
        if (wide) {
            o.status = 1;
            o.fw = w;
            o.lane_stop = stopped;
            out[gid] = o;
            list[atomicAdd(pending, 1)] = gid;
            return;
        }
        o.fmax = f.max;
        o.fe1 = f.e1 - 1;
        o.fe2 = f.e2 - 1;
        o.fw = w;
        o.rw = P.w;
        if (o.fe1 < 0 || o.fe2 < 0) {
            o.status = 0;
            o.rmax = 0;
            o.re1 = o.re2 = -1;
            o.rrows = 0;
            out[gid] = o;
            return;
        }
    }
    int32_t w = o.rw > 0 ? o.rw : P.w;
    const Dir R = reverse_dir(seq, pd, t, o.fe1, o.fe2);
    PassOut r = xrun<BMAX, false>(R, w, P, nullptr, 0, o.cells, wide, stopped, max_rows);
    if (wide) {
        o.status = 2;
        o.rw = w;
        o.lane_stop = stopped;
        out[gid] = o;
        list[atomicAdd(pending, 1)] = gid;
        return;
    }
    o.status = 0;
    o.rmax = r.max;
    o.re1 = r.e1 - 1;
    o.re2 = r.e2 - 1;
    o.rw = w;
    o.rrows = r.rows;
    out[gid] = o;
}

// Pass 1, one wave per task, for the tasks of list[0, n) (all tasks when
// list is null): those the lane kernel handed back (status 1: forward pass
// pending at band fw, 2: reverse pending at rw), or every task when the
// initial band is wider than the lane kernel's.  Tasks whose band outgrows
// the wave's are counted in `pending` for the scratch-band kernel.
__global__ __launch_bounds__(64) void k_xdrop_wave_pass1(const int8_t *__restrict__ seq,
                                                         const PairDesc *__restrict__ pairs,
                                                         const XTask *__restrict__ tasks,
                                                         const int32_t *__restrict__ list, Params P,
                                                         XOut *__restrict__ out, int32_t *__restrict__ pending) {
    const int gid = list ? list[blockIdx.x] : (int)blockIdx.x;
    const bool lead = threadIdx.x == 0;
    const XTask t = tasks[gid];
    XOut o = out[gid];
    if (o.status == 0 && o.fw != 0) return;  // already finished
    const PairDesc pd = pairs[t.pair];
    bool wide;
    if (o.status != 2) {
        int32_t w = o.fw > 0 ? o.fw : P.w;
        const Dir F = forward_dir(seq, pd, t);
        PassOut f = wave_xrun(F, w, P, o.cells, wide);
        if (wide) {
            o.status = 1;
            o.fw = w;
            if (lead) {
                out[gid] = o;
                atomicAdd(pending, 1);
            }
            return;
        }
        o.fmax = f.max;
        o.fe1 = f.e1 - 1;
        o.fe2 = f.e2 - 1;
        o.fw = w;
        o.rw = P.w;
        if (o.fe1 < 0 || o.fe2 < 0) {
            o.status = 0;
            o.rmax = 0;
            o.re1 = o.re2 = -1;
            o.rrows = 0;
            if (lead) out[gid] = o;
            return;
        }
    }
    int32_t w = o.rw > 0 ? o.rw : P.w;
    const Dir R = reverse_dir(seq, pd, t, o.fe1, o.fe2);
    PassOut r = wave_xrun(R, w, P, o.cells, wide);
    if (wide) {
        o.status = 2;
        o.rw = w;
        if (lead) {
            out[gid] = o;
            atomicAdd(pending, 1);
        }
        return;
    }
    o.status = 0;
    o.rmax = r.max;
    o.re1 = r.e1 - 1;
    o.re2 = r.e2 - 1;
    o.rw = w;
    o.rrows = r.rows;
    if (lead) out[gid] = o;
}

// Pass 1 for the tasks the register kernels handed back (status 1: forward
// pass pending, 2: reverse pass pending).  scr_off[t] is the int32 offset of
// the task's 2 x (2 wmax + 1) scratch; wmax[t] is the first doubling of w
// that reaches the pair's longer length, where a pass can no longer restart
// (the band then spans whole rows).
__global__ void k_xdrop_wide_pass1(const int8_t *__restrict__ seq, const PairDesc *__restrict__ pairs,
                                   const XTask *__restrict__ tasks, int32_t n, Params P, XOut *__restrict__ out,
                                   int32_t *__restrict__ scr, const int64_t *__restrict__ scr_off,
                                   const int32_t *__restrict__ wmax) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= n) return;
    if (scr_off[gid] < 0) return;
    const XTask t = tasks[gid];
    XOut o = out[gid];
    const PairDesc pd = pairs[t.pair];
    int32_t *Mp = scr + scr_off[gid];
    int32_t *Fp = Mp + 2 * wmax[gid] + 1;
    if (o.status == 1) {
        int32_t w = o.fw;
        const Dir F = forward_dir(seq, pd, t);
        PassOut f;
        while (true) {
            f = xpass_mem<false>(F, w, P, nullptr, 0, Mp, Fp);
            o.cells += f.cells;
            if (!f.restart || 2 * w > wmax[gid]) break;
            w *= 2;
        }
        o.fmax = f.max;
        o.fe1 = f.e1 - 1;
        o.fe2 = f.e2 - 1;
        o.fw = w;
        o.rw = P.w;
        if (o.fe1 < 0 || o.fe2 < 0) {
            o.status = 0;
            o.rmax = 0;
            o.re1 = o.re2 = -1;
            o.rrows = 0;
            out[gid] = o;
            return;
        }
    }
    int32_t w = o.rw;
    const Dir R = reverse_dir(seq, pd, t, o.fe1, o.fe2);
    PassOut r;
    while (true) {
        r = xpass_mem<false>(R, w, P, nullptr, 0, Mp, Fp);
        o.cells += r.cells;
        if (!r.restart || 2 * w > wmax[gid]) break;
        w *= 2;
    }
    o.status = 0;
    o.rmax = r.max;
    o.re1 = r.e1 - 1;
    o.re2 = r.e2 - 1;
    o.rw = w;
    o.rrows = r.rows;
    out[gid] = o;
}

__device__ __forceinline__ void push_op(uint32_t *cig, int32_t &n, uint32_t op) {
    if (n == 0 || (cig[n - 1] & 0xfu) != op) cig[n++] = (1u << 4) | op;
    else cig[n - 1] += 1u << 4;
}

// Traceback state machine (:1416-1438) over the band-compressed store of the
// final reverse pass; cells outside the stored band read 255.
__device__ void trace_cigar(XOut &o, int32_t w, const uint8_t *__restrict__ T, int32_t BP, uint32_t *__restrict__ cigar) {
    const int32_t B = 2 * w + 1;
    int32_t ii = o.re1 + 1, jj = o.re2 + 1;
    uint32_t *cig = cigar + o.cig_off;
    int32_t nc = 0;
    int state = 0;
    while (ii > 0 && jj > 0) {
        const int32_t k = jj - (ii - w);
        int tv = 255;
        if (k >= 0 && k < B && ii <= o.rrows) tv = T[(int64_t)ii * BP + k];
        if (state == 0) state = tv & 7;
        else if (!((tv >> (state + 2)) & 1)) state = 0;
        if (state == 0) state = tv & 7;
        if (state == 0) { push_op(cig, nc, 0); --ii; --jj; }
        else if (state == 1 || state == 3) { push_op(cig, nc, 2); --ii; }
        else { push_op(cig, nc, 1); --jj; }
    }
    o.iii = ii;
    o.jjj = jj;
    o.ncig = nc;
}

// Pass 2: reverse pass with traceback at the final band from pass 1.
template <int BMAX>
__global__ __launch_bounds__(64) void k_xdrop_pass2(const int8_t *__restrict__ seq, const PairDesc *__restrict__ pairs,
                                                    const XTask *__restrict__ tasks, const int32_t *__restrict__ order,
                                                    int32_t n, Params P, XOut *__restrict__ out,
                                                    const int64_t *__restrict__ tb_off, uint8_t *__restrict__ tb,
                                                    uint32_t *__restrict__ cigar) {
    const int lane_gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (lane_gid >= n) return;
    const int gid = order ? order[lane_gid] : lane_gid;  // the host orders tasks by band class and traced cells
    const XTask t = tasks[gid];
    XOut o = out[gid];
    if (o.fe1 < 0 || o.fe2 < 0) return;
    const int32_t w = o.rw;
    if (2 * w + 1 > BMAX) return;  // a wider kernel's task (the order keeps the narrower ones out)
    const PairDesc pd = pairs[t.pair];
    const Dir R = reverse_dir(seq, pd, t, o.fe1, o.fe2);
    const int32_t B = 2 * w + 1;
    const int32_t BP = (B + 3) & ~3;
    uint8_t *T = tb + tb_off[gid];
    PassOut r = xpass<BMAX, true>(R, w, P, T, BP);
    (void)r;
    trace_cigar(o, w, T, BP, cigar);
    out[gid] = o;
}

// Pass 2, one wave per task, for the tasks order[0, n): those whose final
// reverse band is wider than the lane kernels' and fits the wave, and the
// long ones of the lane bands; the first lane walks the traceback once the
// wave's bytes are stored.
__global__ __launch_bounds__(64) void k_xdrop_wave_pass2(const int8_t *__restrict__ seq,
                                                         const PairDesc *__restrict__ pairs,
                                                         const XTask *__restrict__ tasks,
                                                         const int32_t *__restrict__ order, Params P,
                                                         XOut *__restrict__ out, const int64_t *__restrict__ tb_off,
                                                         uint8_t *__restrict__ tb, uint32_t *__restrict__ cigar) {
    const int gid = order[blockIdx.x];
    const XTask t = tasks[gid];
    XOut o = out[gid];
    if (o.fe1 < 0 || o.fe2 < 0) return;
    const int32_t w = o.rw;
    if (2 * w + 1 > kXdropWaveBand) return;  // the scratch-band kernel's
    const PairDesc pd = pairs[t.pair];
    const Dir R = reverse_dir(seq, pd, t, o.fe1, o.fe2);
    const int32_t B = 2 * w + 1;
    const int32_t BP = (B + 3) & ~3;
    uint8_t *T = tb + tb_off[gid];
    wave_xpass_fit<true>(R, w, P, T, BP);
    __syncthreads();  // the wave's traceback bytes are stored before the walk reads them
    if (threadIdx.x == 0) {
        trace_cigar(o, w, T, BP, cigar);
        out[gid] = o;
    }
}

// Pass 2 for the tasks whose final reverse band is wider than the wave's.
__global__ void k_xdrop_wide_pass2(const int8_t *__restrict__ seq, const PairDesc *__restrict__ pairs,
                                   const XTask *__restrict__ tasks, int32_t n, Params P, XOut *__restrict__ out,
                                   const int64_t *__restrict__ tb_off, uint8_t *__restrict__ tb,
                                   uint32_t *__restrict__ cigar, int32_t *__restrict__ scr,
                                   const int64_t *__restrict__ scr_off) {
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= n) return;
    const XTask t = tasks[gid];
    XOut o = out[gid];
    if (o.fe1 < 0 || o.fe2 < 0) return;
    const int32_t w = o.rw;
    if (2 * w + 1 <= kXdropWaveBand || scr_off[gid] < 0) return;
    const PairDesc pd = pairs[t.pair];
    const Dir R = reverse_dir(seq, pd, t, o.fe1, o.fe2);
    const int32_t B = 2 * w + 1;
    const int32_t BP = (B + 3) & ~3;
    uint8_t *T = tb + tb_off[gid];
    int32_t *Mp = scr + scr_off[gid];
    PassOut r = xpass_mem<true>(R, w, P, T, BP, Mp, Mp + B);
    (void)r;
    trace_cigar(o, w, T, BP, cigar);
    out[gid] = o;
}

// Cigar runs of each traced task (0 for the others), then the runs moved to
// their packed offsets; cig_off is rewritten to the packed offset.
__global__ void k_cig_counts(const XOut *__restrict__ out, int32_t n, int64_t *__restrict__ cnt) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const XOut &o = out[i];
    cnt[i] = (o.fe1 >= 0 && o.fe2 >= 0) ? o.ncig : 0;
}

__global__ void k_cig_pack(XOut *__restrict__ out, int32_t n, const int64_t *__restrict__ off,
                           const uint32_t *__restrict__ cig, uint32_t *__restrict__ packed) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    XOut &o = out[i];
    if (o.fe1 < 0 || o.fe2 < 0) return;
    const uint32_t *src = cig + o.cig_off;
    uint32_t *dst = packed + off[i];
    for (int32_t q = 0; q < o.ncig; ++q) dst[q] = src[q];
    o.cig_off = off[i];
}

}  // namespace

size_t pack_cigars_scratch_bytes(int32_t n) {
    size_t bytes = 0;
    hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, (const int64_t *)nullptr, (int64_t *)nullptr, std::max(n, 1));
    return bytes + 2 * (size_t)std::max(n, 1) * sizeof(int64_t) + 256;
}

hipError_t launch_pack_cigars(hipStream_t st, XOut *out, int32_t n, void *scratch, size_t scratch_bytes,
                              const uint32_t *cig, uint32_t *packed) {
    if (n <= 0) return hipSuccess;
    int64_t *cnt = (int64_t *)scratch;
    int64_t *off = cnt + n;
    void *tmp = off + n;
    size_t tmp_bytes = scratch_bytes - 2 * (size_t)n * sizeof(int64_t);
    const dim3 g((n + 255) / 256), b(256);
    hipLaunchKernelGGL(k_cig_counts, g, b, 0, st, out, n, cnt);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, cnt, off, n, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_cig_pack, g, b, 0, st, out, n, off, cig, packed);
    return hipGetLastError();
}

hipError_t launch_xdrop_pass1(hipStream_t st, int band, const int8_t *seq, const PairDesc *pairs, const XTask *tasks,
                              const int32_t *in_list, int32_t n, const Params &P, XOut *out, int32_t *pending,
                              int32_t *list, int32_t max_rows) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(band <= kXdropLaneBand ? k_xdrop_pass1<kXdropLaneBand> : k_xdrop_pass1<kXdropLane2Band>,
                       dim3((n + 63) / 64), dim3(64), 0, st, seq, pairs, tasks, in_list, n, P, out, pending, list,
                       max_rows);
    return hipGetLastError();
}

hipError_t launch_xdrop_wave_pass1(hipStream_t st, const int8_t *seq, const PairDesc *pairs, const XTask *tasks,
                                   const int32_t *list, int32_t n, const Params &P, XOut *out, int32_t *pending) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_xdrop_wave_pass1, dim3(n), dim3(64), 0, st, seq, pairs, tasks, list, P, out, pending);
    return hipGetLastError();
}

hipError_t launch_xdrop_wide_pass1(hipStream_t st, const int8_t *seq, const PairDesc *pairs, const XTask *tasks,
                                   int32_t n, const Params &P, XOut *out, int32_t *scr, const int64_t *scr_off,
                                   const int32_t *wmax) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_xdrop_wide_pass1, dim3((n + 63) / 64), dim3(64), 0, st, seq, pairs, tasks, n, P, out, scr,
                       scr_off, wmax);
    return hipGetLastError();
}

hipError_t launch_xdrop_wide_pass2(hipStream_t st, const int8_t *seq, const PairDesc *pairs, const XTask *tasks,
                                   int32_t n, const Params &P, XOut *out, const int64_t *tb_off, uint8_t *tb,
                                   uint32_t *cigar, int32_t *scr, const int64_t *scr_off) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_xdrop_wide_pass2, dim3((n + 63) / 64), dim3(64), 0, st, seq, pairs, tasks, n, P, out, tb_off,
                       tb, cigar, scr, scr_off);
    return hipGetLastError();
}

hipError_t launch_xdrop_pass2(hipStream_t st, int band, const int8_t *seq, const PairDesc *pairs, const XTask *tasks,
                              const int32_t *order, int32_t n, const Params &P, XOut *out, const int64_t *tb_off,
                              uint8_t *tb, uint32_t *cigar) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(band <= kXdropLaneBand ? k_xdrop_pass2<kXdropLaneBand> : k_xdrop_pass2<kXdropLane2Band>,
                       dim3((n + 63) / 64), dim3(64), 0, st, seq, pairs, tasks, order, n, P, out, tb_off, tb, cigar);
    return hipGetLastError();
}

hipError_t launch_xdrop_wave_pass2(hipStream_t st, const int8_t *seq, const PairDesc *pairs, const XTask *tasks,
                                   const int32_t *order, int32_t n, const Params &P, XOut *out,
                                   const int64_t *tb_off, uint8_t *tb, uint32_t *cigar) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_xdrop_wave_pass2, dim3(n), dim3(64), 0, st, seq, pairs, tasks, order, P, out, tb_off, tb,
                       cigar);
    return hipGetLastError();
}

}  // namespace dcns
