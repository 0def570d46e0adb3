// seedscan.hip -- windowed seed scan of dCNS on gfx950.
//
// Reference: getSeeds (findSimilarFragmentsForPairedSequence4.cpp:61-201)
// calls the striped SmithWaterman (sequenceAlignment3.cpp:220-469) once per
// 38-bp query window over the whole reference, then recurses on the
// reference left and right of every hit.  Here:
//
//  * k_scan_uniform  -- level 0.  One wave owns 128 consecutive windows of one
//    pair: lane l holds windows base+l and base+64+l as the two int16 halves
//    of each VGPR.  The reference base of the current row is the same for the
//    whole wave (scalar load); the substitution score of a column is one
//    v_perm_b32 of a uniform 64-bit score table by a per-lane selector.
//    H, E and the selectors of all window columns stay in VGPRs for the whole
//    scan, so the only memory traffic is the reference stream.
//  * k_scan_general_h -- later levels (right sub-ranges of the recursion).
//    One lane per task, per-lane reference rows, the window split in two
//    column strips held as packed-f16 halves (the second strip one row
//    behind); tasks of more than 4096 rows one per wave instead (columns
//    across the lanes on an anti-diagonal), so that a long rescan does not
//    hold a whole level.  k_scan_general is the int32 version for scorings
//    the packed one cannot represent.
//  * k_resolve       -- walks the left spine of the recursion for every
//    scanned node: the left child's scan is a prefix of its parent's, so its
//    maximum is read from the parent's trajectory; the reverse SW (<= 76 x 38)
//    runs inline; right children are queued as new scans.
//
// Every scan records its "trajectory": each (row, col, score) where the
// running maximum rose strictly above both every earlier cell and
// min_score - 1.  The first maximum of any row prefix [a, b') of the scan is
// then the last entry with row < b'.
//
// Level 0 also records, per window, the last row whose maximum reaches
// min_score (lastrow).  A rescan is a fresh SmithWaterman over rows [a, b)
// of the same window: every value is >= 0 and every step of the recurrence
// is monotone, so starting from zeros at row a instead of from the level-0
// values there never raises a cell.  A rescan's trajectory entry (a cell of
// at least min_score) can therefore only lie on a row where level 0 reached
// min_score: k_resolve stops every rescan after the window's lastrow, and
// resolves a right child that starts past it as empty without scanning it.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "dp_common.h"
#include "kernels.h"

namespace dcns {
namespace {

__device__ __forceinline__ void uniform_table(const ScoreTab &t, int c, uint32_t &hi, uint32_t &lo) {
    switch (c) {
        case 0: hi = t.hi[0]; lo = t.lo[0]; break;
        case 1: hi = t.hi[1]; lo = t.lo[1]; break;
        case 3: hi = t.hi[3]; lo = t.lo[3]; break;
        case 4: hi = t.hi[4]; lo = t.lo[4]; break;
        default: hi = 0; lo = 0; break;
    }
}

__device__ __forceinline__ void record(TrajEntry *__restrict__ traj, int64_t slot, int32_t cap, int32_t &cnt,
                                       int32_t row, int32_t score, int32_t col) {
    if (cnt < cap) {
        TrajEntry e;
        e.row = row;
        e.score_col = (score << 8) | col;
        traj[slot * cap + cnt] = e;
    }
    ++cnt;
}

// lastrow of a 128-window task's windows (base + lane: the low half, base +
// 64 + lane: the high half) whose halves of hit are nonzero, set to row.
// The lane index comes from mbcnt here rather than from a register kept
// through the row loop, which has none to spare.
__device__ __forceinline__ void mark_lastrow(int32_t *__restrict__ lastrow, uint32_t base, uint32_t hit, int32_t row) {
    const uint32_t ln = __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
    if (hit & 0xffffu) lastrow[base + ln] = row;
    if (hit >> 16) lastrow[base + 64u + ln] = row;
}

// mark_lastrow for segments that run at once: the largest row wins.
__device__ __forceinline__ void mark_lastrow_max(int32_t *__restrict__ lastrow, uint32_t base, uint32_t hit,
                                                 int32_t row) {
    const uint32_t ln = __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
    if (hit & 0xffffu) atomicMax(&lastrow[base + ln], row);
    if (hit >> 16) atomicMax(&lastrow[base + 64u + ln], row);
}

// Appends one strong row (window w, entry rv) to the calling lane's chunk of
// the sink.  sk: the lane's two LDS words, its chunk (-1 before the first)
// and fill.  Lanes append independently (no cross-lane operation, so the
// row loops around the calls keep their unrolling).
__device__ __forceinline__ void strong_push(const StrongSink &S, volatile int32_t *sk, uint32_t w, uint64_t rv) {
    int32_t base = sk[0], fill = sk[1];
    if (base < 0 || fill == kStrongChunk) {
        base = atomicAdd(S.nchunks, 1);
        fill = 0;
        sk[0] = base;
    }
    if (base < S.cap_chunks) {
        const int64_t at = (int64_t)base * kStrongChunk + fill;
        S.win[at] = w;
        S.rv[at] = rv;
        S.chunk_fill[base] = fill + 1;
    }
    sk[1] = fill + 1;
}

// The lane's two words of a kernel's strong-row state (2 per lane).
__device__ __forceinline__ volatile int32_t *lane_sk(int32_t *sk) {
    return sk + 2 * (int32_t)__builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

__device__ __forceinline__ uint64_t strong_entry(int32_t row, int32_t v, int32_t col) {
    return ((uint64_t)(uint32_t)row << 32) | (uint64_t)(uint32_t)((v << 8) | col);
}

// One long ScanTask across a wave: lane j owns window column j and works on
// row a + s - j at step s (an anti-diagonal), so a cell's left and diagonal
// neighbours were computed by lane j - 1 one and two steps earlier and reach
// lane j by a DPP shift (wave_shr:1, lane 0 taking the row's inputs).  The
// reference base travels the same way as its score selector: with symmetric
// scoring the lane's table is the one of its own query base.  The row
// maximum travels as a partial (maximum, first column) and is complete at the
// window's last column, whose lane keeps the running best and records the
// trajectory.  The N columns past the window that the lane kernel also
// evaluates never hold a new maximum (each is bounded by the previous row's
// last real cell), so the records are the same.  R rows take R + width - 1
// steps instead of R x width lane steps.
// L0: a level-0 window (rows [0, L1) of its pair): a window with a record
// joins the hit list, and its cells are counted on the host instead.
template <int WM, bool L0 = false, bool REC = false>
__device__ void scan_task_wave(const int8_t *__restrict__ seq, const PairDesc *__restrict__ pairs, const ScanTask &t,
                               const Params &P, const ScoreTab &tab, TrajEntry *__restrict__ traj,
                               int32_t *__restrict__ tcount, int32_t cap, Counters *__restrict__ ctr,
                               int32_t *__restrict__ hitlist = nullptr, int32_t *__restrict__ lastrow = nullptr,
                               const StrongSink *sink = nullptr, int32_t *sk = nullptr) {
    const int lane = threadIdx.x;
    const PairDesc pd = pairs[t.pair];
    const int8_t *__restrict__ ref = seq + pd.off1;
    const int8_t *__restrict__ qry = seq + pd.off2;
    const int sp = t.win * P.step;
    const int tw = min(P.window, pd.len2 - sp);
    const int qc = lane < tw ? qry[sp + lane] : 2;
    const int qi = (unsigned)qc < 8u ? qc : 2;
    const uint32_t thi = tab.hi[qi], tlo = tab.lo[qi];
    const int32_t O = P.open, X = P.ext;
    const int rows = t.b - t.a, steps = rows + tw - 1, last = tw - 1;
    int32_t E = 0, diag = 0, h_out = 0, f_out = 0, pm_out = -1, pj_out = 0;
    uint32_t sel_out = 0x0C0C0C0Cu;
    int32_t best = P.min_score - 1, cnt = 0;
    // reference bases 64 rows at a time, one per lane (2 past the task's
    // rows), the next 64 in flight while these are used
    int32_t cur = lane < rows ? ref[t.a + lane] : 2;
    int32_t nxt = 64 + lane < rows ? ref[t.a + 64 + lane] : 2;
    for (int st = 0; st < steps; ++st) {
        const int c0 = __builtin_amdgcn_readlane(cur, st & 63);  // row a + st
        if ((st & 63) == 63) {
            cur = nxt;
            const int r = st + 65 + lane;
            nxt = r < rows ? ref[t.a + r] : 2;
        }
        const uint32_t sel = (uint32_t)__builtin_amdgcn_update_dpp((int32_t)sel32(c0), (int32_t)sel_out, 0x138, 0xf,
                                                                   0xf, false);
        const int32_t fin = __builtin_amdgcn_update_dpp(0, f_out, 0x138, 0xf, 0xf, false);
        const int32_t hin = __builtin_amdgcn_update_dpp(0, h_out, 0x138, 0xf, 0xf, false);
        const int32_t pmin = __builtin_amdgcn_update_dpp(-1, pm_out, 0x138, 0xf, 0xf, false);
        const int32_t pjin = __builtin_amdgcn_update_dpp(0, pj_out, 0x138, 0xf, 0xf, false);
        const int row = st - lane;
        const bool valid = row >= 0 && row < rows && lane < tw;
        int32_t h = 0, fo = 0, pm = pmin, pj = pjin;
        if (valid) {
            const int32_t tt = diag + (int32_t)perm(thi, tlo, sel);
            h = max3i(tt, E, fin);  // E, fin >= 0: includes the 0 floor
            const int32_t ho = h + O;
            E = max3i(E + X, ho, 0);
            fo = max3i(ho, fin + X, 0);
            if (h > pmin) {
                pm = h;
                pj = lane;
            }
            if (lane == last && pm >= P.min_score) {
                if (L0) lastrow[t.scan] = t.a + row;  // rows rise: the last store is the last such row
                if (L0 && REC) strong_push(*sink, lane_sk(sk), (uint32_t)t.scan, strong_entry(t.a + row, pm, pj));
                if (pm > best) {
                    record(traj, t.scan, cap, cnt, t.a + row, pm, pj);
                    best = pm;
                }
            }
        }
        diag = hin;  // H(row, lane - 1): the diagonal of this lane's next row
        h_out = h;
        f_out = fo;
        pm_out = pm;
        pj_out = pj;
        sel_out = sel;
    }
    if (lane == last) {
        tcount[t.scan] = cnt;
        if (cnt > 0) {
            if (L0) hitlist[atomicAdd(&ctr->hits, 1)] = t.scan;
            atomicAdd(&ctr->entries, min(cnt, cap));
        }
        if (!L0) atomicAdd(&ctr->gen_cells, (unsigned long long)rows * (unsigned long long)max(tw, 0));
    }
}

// A level-0 window of one of the longest references, one wave on the
// anti-diagonal of scan_task_wave: a reference of R rows takes R + width - 1
// steps instead of R rows of the 128-window lane kernel, at ~10x the VALU per
// cell.  The level-0 kernels run these windows in their first blocks (wins,
// nwins; tab32: the int32 score table); the host picks the windows whose
// lane-kernel rows would otherwise set the launch's duration.
template <int WM, bool REC = false>
__device__ __forceinline__ void l0_wave_window(const int2 w, const int8_t *__restrict__ seq,
                                               const PairDesc *__restrict__ pairs, const Params &P,
                                               const ScoreTab &tab32, TrajEntry *__restrict__ traj,
                                               int32_t *__restrict__ tcount, int32_t cap,
                                               int32_t *__restrict__ hitlist, Counters *__restrict__ ctr,
                                               int32_t *__restrict__ lastrow, const StrongSink &sink,
                                               int32_t *sk) {
    const PairDesc pd = pairs[w.x];
    ScanTask t;
    t.pair = w.x;
    t.win = w.y;
    t.a = 0;
    t.b = pd.len1;
    t.node = -1;
    t.scan = pd.win_base + w.y;
    scan_task_wave<WM, true, REC>(seq, pairs, t, P, tab32, traj, tcount, cap, ctr, hitlist, lastrow, &sink, sk);
}

// The next level-0 task of a persistent wave: lane 0 pops the queue head,
// the wave reads the index from it (uniform).
__device__ __forceinline__ int32_t pop_task(int32_t *queue) {
    int32_t t = 0;
    if (threadIdx.x == 0) t = atomicAdd(queue, 1);
    return __builtin_amdgcn_readfirstlane(t);
}

template <int WM>
__device__ __forceinline__ void scan_uniform_task(const int2 task, const int8_t *__restrict__ seq,
                                                  const PairDesc *__restrict__ pairs, const Params &P,
                                                  const ScoreTab &tab, TrajEntry *__restrict__ traj,
                                                  int32_t *__restrict__ tcount, int32_t cap,
                                                  int32_t *__restrict__ hitlist, Counters *__restrict__ ctr,
                                                  int32_t *__restrict__ lastrow) {
    const int lane = threadIdx.x;
    const PairDesc pd = pairs[task.x];
    const int8_t *__restrict__ ref = seq + pd.off1;
    const int8_t *__restrict__ qry = seq + pd.off2;
    const int L1 = pd.len1, L2 = pd.len2;
    const int wl = task.y + lane, wh = task.y + 64 + lane;
    const bool vl = wl < pd.nwin, vh = wh < pd.nwin;
    const int spl = wl * P.step, sph = wh * P.step;
    const int twl = vl ? min(P.window, L2 - spl) : 0;
    const int twh = vh ? min(P.window, L2 - sph) : 0;

    uint32_t sel[WM], H[WM], E[WM];
#pragma unroll
    for (int j = 0; j < WM; ++j) {
        int ql = j < twl ? qry[spl + j] : 2;
        int qh = j < twh ? qry[sph + j] : 2;
        sel[j] = sel16(ql) | (sel16(qh) << 16);
        H[j] = 0;
        E[j] = 0;
    }
    const uint32_t OPEN = (uint32_t)(-P.open) * 0x10001u;
    const uint32_t EXT = (uint32_t)(-P.ext) * 0x10001u;
    const uint32_t thr = (uint32_t)(P.min_score - 1);
    const uint32_t thr2 = thr | (thr << 16);
    uint32_t best = thr2;
    int32_t cntl = 0, cnth = 0;

    uint32_t word = L1 > 0 ? *(const uint32_t *)ref : 0u;
    uint32_t next = 0;
    for (int i = 0; i < L1; ++i) {
        const int sub = i & 3;
        if (sub == 0 && i + 4 < L1) next = *(const uint32_t *)(ref + i + 4);
        const int c = (int)(int8_t)(word >> (8 * sub));
        if (sub == 3) word = next;
        uint32_t thi, tlo;
        uniform_table(tab, c, thi, tlo);
        uint32_t diag = 0, fc = 0, rmax = 0;
#pragma unroll
        for (int j = 0; j < WM; ++j) {
            const uint32_t s = perm(thi, tlo, sel[j]);
            const uint32_t t = pk_add(diag, s);
            diag = H[j];
            const uint32_t h = pk_max_s(pk_max_s(t, E[j]), fc);  // E, fc >= 0: includes the 0 floor
            const uint32_t ho = pk_subsat(h, OPEN);                // both gap kinds open from the final H
            E[j] = pk_max_u(pk_subsat(E[j], EXT), ho);
            fc = pk_max_u(ho, pk_subsat(fc, EXT));
            H[j] = h;
            rmax = pk_max_s(rmax, h);
        }
        const uint32_t hit = pk_subsat(rmax, thr2);  // halves whose row reaches min_score
        if (__builtin_expect(hit != 0, 0)) {
            mark_lastrow(lastrow, (uint32_t)(pd.win_base + task.y), hit, i);
            const uint32_t imp = pk_subsat(rmax, best);
            if (imp & 0xffffu) {
                const uint32_t v = rmax & 0xffffu;
                int jj = 0;
#pragma unroll
                for (int j = WM - 1; j >= 0; --j)
                    if ((H[j] & 0xffffu) == v) jj = j;
                record(traj, pd.win_base + wl, cap, cntl, i, (int32_t)v, jj);
                best = (best & 0xffff0000u) | v;
            }
            if (imp >> 16) {
                const uint32_t v = rmax >> 16;
                int jj = 0;
#pragma unroll
                for (int j = WM - 1; j >= 0; --j)
                    if ((H[j] >> 16) == v) jj = j;
                record(traj, pd.win_base + wh, cap, cnth, i, (int32_t)v, jj);
                best = (best & 0xffffu) | (v << 16);
            }
        }
    }
    if (vl) {
        tcount[pd.win_base + wl] = cntl;
        if (cntl > 0) {
            hitlist[atomicAdd(&ctr->hits, 1)] = pd.win_base + wl;
            atomicAdd(&ctr->entries, min(cntl, cap));
        }
    }
    if (vh) {
        tcount[pd.win_base + wh] = cnth;
        if (cnth > 0) {
            hitlist[atomicAdd(&ctr->hits, 1)] = pd.win_base + wh;
            atomicAdd(&ctr->entries, min(cnth, cap));
        }
    }
}

// Level-0 window scan, one wave per 128-window task: task blockIdx.x of the
// length-sorted list or, with a queue (persistent waves, a grid that fills
// the chip once), the tasks popped from its head in list order.
template <int WM, bool PERSIST, bool WAVES>
__global__ __launch_bounds__(64) void k_scan_uniform(const int8_t *__restrict__ seq, const PairDesc *__restrict__ pairs,
                                                     const int2 *__restrict__ wtasks, int32_t nwt,
                                                     int32_t *__restrict__ queue, Params P, ScoreTab tab,
                                                     TrajEntry *__restrict__ traj, int32_t *__restrict__ tcount,
                                                     int32_t cap, int32_t *__restrict__ hitlist,
                                                     Counters *__restrict__ ctr, const int2 *__restrict__ wins,
                                                     int32_t nwins, int32_t *__restrict__ lastrow) {
    const int32_t lane_blocks = (int32_t)gridDim.x - nwins;
    if (WAVES && (int32_t)blockIdx.x >= lane_blocks) {
        // this kernel's scorings never record strong rows
        l0_wave_window<WM>(wins[blockIdx.x - lane_blocks], seq, pairs, P, tab, traj, tcount, cap, hitlist, ctr,
                           lastrow, StrongSink{nullptr, nullptr, nullptr, nullptr, 0}, nullptr);
        return;
    }
    if (!PERSIST) {
        scan_uniform_task<WM>(wtasks[blockIdx.x], seq, pairs, P, tab, traj, tcount, cap, hitlist, ctr, lastrow);
        return;
    }
    for (int32_t t = pop_task(queue); t < nwt; t = pop_task(queue))
        scan_uniform_task<WM>(wtasks[t], seq, pairs, P, tab, traj, tcount, cap, hitlist, ctr, lastrow);
}

// Packed-f16 version of k_scan_uniform (same wave layout).  Valid when
// window * match <= 2048: every H, E, F and score is then an exact f16
// integer, and v_pk_maximum3_f16 merges the three-way maxima.  Non-negative
// f16 values order like their bit patterns, so the improvement test and the
// first-column search compare bits.
template <int WM, bool REC, bool SEG = false>
__device__ __forceinline__ void scan_uniform_task_h(const int2 task, const int8_t *__restrict__ seq,
                                                    const PairDesc *__restrict__ pairs, const Params &P,
                                                    const ScoreTab &tab, TrajEntry *__restrict__ traj,
                                                    int32_t *__restrict__ tcount, int32_t cap,
                                                    int32_t *__restrict__ hitlist, Counters *__restrict__ ctr,
                                                    uint32_t open_h, uint32_t ext_h, uint32_t thr_h,
                                                    int32_t *__restrict__ lastrow, const StrongSink &sink,
                                                    int32_t *sk, const SegL0 &sg = SegL0{}, int32_t r0 = 0,
                                                    int32_t r1 = 0, int32_t tid = 0) {
    const int lane = threadIdx.x;
    const PairDesc pd = pairs[task.x];
    const int8_t *__restrict__ ref = seq + pd.off1;
    const int8_t *__restrict__ qry = seq + pd.off2;
    const int L1 = pd.len1, L2 = pd.len2;
    const int wl = task.y + lane, wh = task.y + 64 + lane;
    const bool vl = wl < pd.nwin, vh = wh < pd.nwin;
    const int spl = wl * P.step, sph = wh * P.step;
    const int twl = vl ? min(P.window, L2 - spl) : 0;
    const int twh = vh ? min(P.window, L2 - sph) : 0;

    uint32_t sel[WM];
    f16x2 H[WM], E[WM];
    const f16x2 zero = as_h2(0u);
#pragma unroll
    for (int j = 0; j < WM; ++j) {
        const int ql = j < twl ? qry[spl + j] : 2;
        const int qh = j < twh ? qry[sph + j] : 2;
        sel[j] = sel16_f16(ql) | (sel16_f16(qh) << 16);
        H[j] = zero;
        E[j] = zero;
    }
    const f16x2 O = as_h2(open_h * 0x10001u), X = as_h2(ext_h * 0x10001u);
    const uint32_t thr2 = thr_h * 0x10001u;
    uint32_t best = thr2;
    int32_t cntl = 0, cnth = 0;

    // SEG: rows [r0, r1) after at least sg.warm rows from zeros (the level-0
    // tail bound makes rows from r0 on level 0's values), the first row a
    // multiple of 4 like level 0's own; records only from r0
    const int i0 = SEG ? max(0, r0 - sg.warm) & ~3 : 0, i1 = SEG ? r1 : L1;
    uint32_t word = i0 < L1 ? *(const uint32_t *)(ref + i0) : 0u;
    uint32_t next = 0;
    // two rows per iteration: the register allocator then keeps H, E in
    // place across rows instead of rotating them with moves at the loop end
#pragma unroll 2
    for (int i = i0; i < i1; ++i) {
        const int sub = i & 3;
        if (sub == 0 && i + 4 < L1) next = *(const uint32_t *)(ref + i + 4);
        const int c = (int)(int8_t)(word >> (8 * sub));
        if (sub == 3) word = next;
        uint32_t thi, tlo;
        uniform_table(tab, c, thi, tlo);
        // t: the diagonal term of column j, the previous row's H[j - 1] plus
        // the score (0 before the first column), formed before H[j - 1] is
        // overwritten
        f16x2 t = as_h2(perm(thi, tlo, sel[0])), fc = zero, rmax = zero;
#pragma unroll
        for (int j = 0; j < WM; ++j) {
            const f16x2 h = hmax3(t, E[j], fc);  // E, fc >= 0: includes the 0 floor
            if (j + 1 < WM) t = H[j] + as_h2(perm(thi, tlo, sel[j + 1]));
            const f16x2 ho = h + O;              // both gap kinds open from the final H
            E[j] = hmax3(E[j] + X, ho, zero);
            fc = hmax3(ho, fc + X, zero);
            H[j] = h;
            if (j & 1) rmax = hmax3(rmax, H[j - 1], h);
            else if (j == WM - 1) rmax = __builtin_elementwise_maximum(rmax, h);
        }
        const uint32_t hit = pk_subsat(bits(rmax), thr2);  // halves whose row reaches min_score
        if (__builtin_expect(hit != 0 && (!SEG || i >= r0), 0)) {
            if (SEG) mark_lastrow_max(lastrow, (uint32_t)(pd.win_base + task.y), hit, i);
            else mark_lastrow(lastrow, (uint32_t)(pd.win_base + task.y), hit, i);
            const uint32_t rb = bits(rmax);
            const uint32_t imp = pk_subsat(rb, best);
            // first column of each half's row maximum, where a record needs
            // it (a strong-row entry leaves it to the tail)
            const uint32_t need = imp;
            int jl = 0, jh = 0;
            if (need & 0xffffu) {
                const uint32_t v = rb & 0xffffu;
#pragma unroll
                for (int j = WM - 1; j >= 0; --j)
                    if ((bits(H[j]) & 0xffffu) == v) jl = j;
            }
            if (need >> 16) {
                const uint32_t v = rb >> 16;
#pragma unroll
                for (int j = WM - 1; j >= 0; --j)
                    if ((bits(H[j]) >> 16) == v) jh = j;
            }
            const int32_t vl = (int32_t)(float)rmax[0], vh = (int32_t)(float)rmax[1];
            if (imp & 0xffffu) {
                if (SEG) record(sg.traj, (int64_t)tid * 128 + lane, cap, cntl, i, vl, jl);
                else record(traj, pd.win_base + wl, cap, cntl, i, vl, jl);
                best = (best & 0xffff0000u) | (rb & 0xffffu);
            }
            if (imp >> 16) {
                if (SEG) record(sg.traj, (int64_t)tid * 128 + 64 + lane, cap, cnth, i, vh, jh);
                else record(traj, pd.win_base + wh, cap, cnth, i, vh, jh);
                best = (best & 0xffffu) | (rb & 0xffff0000u);
            }
            if (REC) {
                volatile int32_t *my = lane_sk(sk);
                if (hit & 0xffffu)
                    strong_push(sink, my, (uint32_t)(pd.win_base + wl), strong_entry(i, vl, kStrongNoCol));
                if (hit >> 16)
                    strong_push(sink, my, (uint32_t)(pd.win_base + wh), strong_entry(i, vh, kStrongNoCol));
            }
        }
    }
    if (SEG) {
        // the segments' trajectories merge in k_l0_merge
        sg.tcount[(int64_t)tid * 128 + lane] = vl ? cntl : 0;
        sg.tcount[(int64_t)tid * 128 + 64 + lane] = vh ? cnth : 0;
        return;
    }
    if (vl) {
        tcount[pd.win_base + wl] = cntl;
        if (cntl > 0) {
            hitlist[atomicAdd(&ctr->hits, 1)] = pd.win_base + wl;
            atomicAdd(&ctr->entries, min(cntl, cap));
        }
    }
    if (vh) {
        tcount[pd.win_base + wh] = cnth;
        if (cnth > 0) {
            hitlist[atomicAdd(&ctr->hits, 1)] = pd.win_base + wh;
            atomicAdd(&ctr->entries, min(cnth, cap));
        }
    }
}

// Level 0 in row segments: one wave per (128-window group, segment) task,
// sg.tasks[blockIdx.x] = (pair, first window, r0, r1).
template <int WM, bool REC>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void k_scan_uniform_seg(
    const int8_t *__restrict__ seq, const PairDesc *__restrict__ pairs, SegL0 sg, Params P, ScoreTab tab,
    int32_t cap, Counters *__restrict__ ctr, uint32_t open_h, uint32_t ext_h, uint32_t thr_h,
    int32_t *__restrict__ lastrow, StrongSink sink) {
    __shared__ int32_t sk[REC ? 128 : 1];  // each lane's strong-row chunk and fill
    if (REC) {
        sk[2 * threadIdx.x] = -1;
        sk[2 * threadIdx.x + 1] = 0;
    }
    const int4 t = sg.tasks[blockIdx.x];
    scan_uniform_task_h<WM, REC, true>(make_int2(t.x, t.y), seq, pairs, P, tab, nullptr, nullptr, cap, nullptr, ctr,
                                       open_h, ext_h, thr_h, lastrow, sink, sk, sg, t.z, t.w, (int32_t)blockIdx.x);
}

// A segmented group's windows: the segments' trajectories in row order,
// each entry kept when it beats every earlier one (a segment's own running
// maximum starts at min_score - 1, so its records are a superset of the
// window's past its first row), into the window's trajectory slot; windows
// with an entry join the hit list.  sg.groups[g] = (pair, first window,
// first task, segments).
__global__ __launch_bounds__(128) void k_l0_merge(const PairDesc *__restrict__ pairs, SegL0 sg, int32_t ngroups,
                                                  TrajEntry *__restrict__ traj, int32_t *__restrict__ tcount,
                                                  int32_t cap, int32_t min_score, int32_t *__restrict__ hitlist,
                                                  Counters *__restrict__ ctr) {
    const int32_t g = blockIdx.x, k = threadIdx.x;
    if (g >= ngroups) return;
    const int4 G = sg.groups[g];
    const PairDesc pd = pairs[G.x];
    const int32_t w = G.y + k;
    if (w >= pd.nwin) return;
    const int64_t gw = pd.win_base + w;
    int32_t best = min_score - 1, cnt = 0;
    for (int32_t s = 0; s < G.w; ++s) {
        const int64_t slot = (int64_t)(G.z + s) * 128 + k;
        const int32_t n = min(sg.tcount[slot], cap);
        const TrajEntry *e = sg.traj + slot * cap;
        for (int32_t q = 0; q < n; ++q) {
            const TrajEntry x = e[q];
            if ((x.score_col >> 8) > best) {
                if (cnt < cap) traj[gw * cap + cnt] = x;
                ++cnt;
                best = x.score_col >> 8;
            }
        }
    }
    tcount[gw] = cnt;
    if (cnt > 0) {
        hitlist[atomicAdd(&ctr->hits, 1)] = (int32_t)gw;
        atomicAdd(&ctr->entries, min(cnt, cap));
    }
}

// One wave per task: task blockIdx.x of the length-sorted list (3 waves
// per SIMD fit its 168 registers).
template <int WM, bool WAVES, bool REC>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void k_scan_uniform_h(
    const int8_t *__restrict__ seq, const PairDesc *__restrict__ pairs, const int2 *__restrict__ wtasks, int32_t nwt,
    int32_t *__restrict__ queue, Params P, ScoreTab tab, TrajEntry *__restrict__ traj, int32_t *__restrict__ tcount,
    int32_t cap, int32_t *__restrict__ hitlist, Counters *__restrict__ ctr, uint32_t open_h, uint32_t ext_h,
    uint32_t thr_h, ScoreTab tab32, const int2 *__restrict__ wins, int32_t nwins, int32_t *__restrict__ lastrow, StrongSink sink) {
    __shared__ int32_t sk[REC ? 128 : 1];  // each lane's strong-row chunk and fill
    if (REC) {
        sk[2 * threadIdx.x] = -1;
        sk[2 * threadIdx.x + 1] = 0;
    }
    const int32_t lane_blocks = (int32_t)gridDim.x - nwins;
    if (WAVES && (int32_t)blockIdx.x >= lane_blocks) {
        l0_wave_window<WM, REC>(wins[blockIdx.x - lane_blocks], seq, pairs, P, tab32, traj, tcount, cap, hitlist,
                                ctr, lastrow, sink, sk);
        return;
    }
    scan_uniform_task_h<WM, REC>(wtasks[blockIdx.x], seq, pairs, P, tab, traj, tcount, cap, hitlist, ctr, open_h,
                                 ext_h, thr_h, lastrow, sink, sk);
}

// Persistent waves: each pops tasks from the queue head in list order until
// the list is exhausted.  The loop's own state does not fit beside the task's
// registers at 3 waves per SIMD without spilling, so this one runs 2.
template <int WM, bool WAVES, bool REC>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void k_scan_uniform_hp(
    const int8_t *__restrict__ seq, const PairDesc *__restrict__ pairs, const int2 *__restrict__ wtasks, int32_t nwt,
    int32_t *__restrict__ queue, Params P, ScoreTab tab, TrajEntry *__restrict__ traj, int32_t *__restrict__ tcount,
    int32_t cap, int32_t *__restrict__ hitlist, Counters *__restrict__ ctr, uint32_t open_h, uint32_t ext_h,
    uint32_t thr_h, ScoreTab tab32, const int2 *__restrict__ wins, int32_t nwins, int32_t *__restrict__ lastrow, StrongSink sink) {
    __shared__ int32_t sk[REC ? 128 : 1];  // each lane's strong-row chunk and fill
    if (REC) {
        sk[2 * threadIdx.x] = -1;
        sk[2 * threadIdx.x + 1] = 0;
    }
    const int32_t lane_blocks = (int32_t)gridDim.x - nwins;
    if (WAVES && (int32_t)blockIdx.x >= lane_blocks) {
        l0_wave_window<WM, REC>(wins[blockIdx.x - lane_blocks], seq, pairs, P, tab32, traj, tcount, cap, hitlist,
                                ctr, lastrow, sink, sk);
        return;
    }
    for (int32_t t = pop_task(queue); t < nwt; t = pop_task(queue))
        scan_uniform_task_h<WM, REC>(wtasks[t], seq, pairs, P, tab, traj, tcount, cap, hitlist, ctr, open_h, ext_h,
                                     thr_h, lastrow, sink, sk);
}

// One lane per ScanTask: rows [a, b) of seq1 against one window, int32.
// Like every kernel here it applies the textbook recurrence (both gap kinds
// open from the final H): the striped variant of the reference, where the
// vertical gap opens from H before the horizontal fix-up, yields the same H
// in every cell (a horizontal run followed by a vertical one can always be
// swapped without changing score or end cell), and the shared H + open term
// saves two operations per cell.

// Tasks of more than wave_rows rows among the first wave_blocks of the
// order (longest first) run one per wave (the first wave_blocks blocks);
// every other task runs one per lane (the blocks after them, which stride
// over the order when they are fewer than its tasks).  The task count is
// read from the device, so the launch can be enqueued before it is known.
template <int WM>
__global__ __launch_bounds__(64) void k_scan_general(const int8_t *__restrict__ seq, const PairDesc *__restrict__ pairs,
                                                     const ScanTask *__restrict__ tasks, const int32_t *__restrict__ order,
                                                     const int32_t *__restrict__ ntasks_p, int32_t wave_blocks,
                                                     int32_t wave_rows, Params P, ScoreTab tab,
                                                     TrajEntry *__restrict__ traj, int32_t *__restrict__ tcount,
                                                     int32_t cap, Counters *__restrict__ ctr) {
    const int ntasks = *ntasks_p;
    if ((int)blockIdx.x < wave_blocks) {
        if ((int)blockIdx.x >= ntasks) return;
        const ScanTask t = tasks[order ? order[blockIdx.x] : blockIdx.x];
        if (t.b - t.a > wave_rows) scan_task_wave<WM>(seq, pairs, t, P, tab, traj, tcount, cap, ctr);
        return;
    }
    __shared__ uint32_t shi[8], slo[8];
    if (threadIdx.x < 8) {
        shi[threadIdx.x] = tab.hi[threadIdx.x];
        slo[threadIdx.x] = tab.lo[threadIdx.x];
    }
    __syncthreads();
    const int stride = ((int)gridDim.x - wave_blocks) * (int)blockDim.x;
    for (int gid = ((int)blockIdx.x - wave_blocks) * blockDim.x + threadIdx.x; gid < ntasks; gid += stride) {
    const ScanTask t = tasks[order ? order[gid] : gid];
    if (gid < wave_blocks && t.b - t.a > wave_rows) continue;  // a wave's task
    const PairDesc pd = pairs[t.pair];
    const int8_t *__restrict__ ref = seq + pd.off1;
    const int8_t *__restrict__ qry = seq + pd.off2;
    const int sp = t.win * P.step;
    const int tw = min(P.window, pd.len2 - sp);
    uint32_t sel[WM];
    int32_t H[WM], E[WM];
#pragma unroll
    for (int j = 0; j < WM; ++j) {
        sel[j] = sel32(j < tw ? qry[sp + j] : 2);
        H[j] = 0;
        E[j] = 0;
    }
    const int32_t O = P.open, X = P.ext;
    int32_t best = P.min_score - 1;
    int32_t cnt = 0;
    int c_next = t.a < t.b ? ref[t.a] : 2;
    for (int i = t.a; i < t.b; ++i) {
        const int c = c_next;
        if (i + 1 < t.b) c_next = ref[i + 1];
        const int ci = (unsigned)c < 8u ? c : 2;
        const uint32_t thi = shi[ci], tlo = slo[ci];
        int32_t diag = 0, fc = 0, rmax = 0;
#pragma unroll
        for (int j = 0; j < WM; ++j) {
            const int32_t s = (int32_t)perm(thi, tlo, sel[j]);
            const int32_t t = diag + s;
            diag = H[j];
            const int32_t h = max3i(t, E[j], fc);  // E, fc >= 0: includes the 0 floor
            const int32_t ho = h + O;
            E[j] = max3i(E[j] + X, ho, 0);
            fc = max3i(ho, fc + X, 0);
            H[j] = h;
            if (j & 1) rmax = max3i(rmax, H[j - 1], h);
            else if (j == WM - 1) rmax = max(rmax, h);
        }
        if (rmax > best) {
            int jj = 0;
#pragma unroll
            for (int j = WM - 1; j >= 0; --j)
                if (H[j] == rmax) jj = j;
            record(traj, t.scan, cap, cnt, i, rmax, jj);
            best = rmax;
        }
    }
    tcount[t.scan] = cnt;
    if (cnt > 0) atomicAdd(&ctr->entries, min(cnt, cap));
    atomicAdd(&ctr->gen_cells, (unsigned long long)(t.b - t.a) * (unsigned long long)max(tw, 0));
    }
}

// Packed-f16 k_scan_general: one task per lane, its window split in two
// column strips, columns [0, C) in the low f16 half and [C, 2C) in the high
// half, the high half one row behind: at step s the low half computes row s
// and the high half row s - 1, whose left neighbours (F out of column C - 1,
// and H(s - 2, C - 1) for the diagonal) the low half left at the previous
// step.  A row thus costs C packed columns instead of 2C int32 ones, with as
// many tasks in flight.  The two halves score different reference rows:
// besides the f16 bound of the window scan this needs scores whose f16
// patterns have a zero low byte (the small integers of the usual scorings,
// e.g. 2 = 0x4000, -3 = 0xC200), so that the high bytes of s(c, A..T) of a
// base c fit one word and one v_perm_b32 over the two rows' words yields
// both scores.  A row's maximum is complete when its high half is done; the
// first column of a low-half maximum is taken the step before, when that
// half alone exceeds the running best.  Tasks the int32 wave path takes (the
// first wave_blocks of the order, more than wave_rows rows) run there as in
// k_scan_general.
template <int WM>
__global__ __launch_bounds__(64) void k_scan_general_h(
    const int8_t *__restrict__ seq, const PairDesc *__restrict__ pairs, const ScanTask *__restrict__ tasks,
    const int32_t *__restrict__ order, const int32_t *__restrict__ ntasks_p, int32_t wave_blocks, int32_t wave_rows,
    Params P, ScoreTab tab, TrajEntry *__restrict__ traj, int32_t *__restrict__ tcount, int32_t cap,
    Counters *__restrict__ ctr, uint32_t open_h, uint32_t ext_h, uint32_t thr_h, uint32_t mis4, uint32_t diff) {
    constexpr int C = WM / 2;
    const int ntasks = *ntasks_p;
    if ((int)blockIdx.x < wave_blocks) {
        if ((int)blockIdx.x >= ntasks) return;
        const ScanTask t = tasks[order ? order[blockIdx.x] : blockIdx.x];
        if (t.b - t.a > wave_rows) scan_task_wave<WM>(seq, pairs, t, P, tab, traj, tcount, cap, ctr);
        return;
    }
    const int stride = ((int)gridDim.x - wave_blocks) * 64;
    for (int gid = ((int)blockIdx.x - wave_blocks) * 64 + (int)threadIdx.x; gid < ntasks; gid += stride) {
    const ScanTask t = tasks[order ? order[gid] : gid];
    if (gid < wave_blocks && t.b - t.a > wave_rows) continue;  // a wave's task
    const PairDesc pd = pairs[t.pair];
    const int8_t *__restrict__ ref = seq + pd.off1 + t.a;
    const int8_t *__restrict__ qry = seq + pd.off2;
    const int sp = t.win * P.step;
    const int tw = min(P.window, pd.len2 - sp);
    const int rows = t.b - t.a;
    // selector bytes: 1 = the low half's base class in the row word of
    // bytes 0-3, 3 = the high half's in bytes 4-7; bytes 0, 2 (and N) zero
    uint32_t sel[C];
#pragma unroll
    for (int j = 0; j < C; ++j) {
        const int ql = j < tw ? qry[sp + j] : 2;
        const int qh = C + j < tw ? qry[sp + C + j] : 2;
        const bool bl = (ql >= 0) & (ql <= 4) & (ql != 2), bh = (qh >= 0) & (qh <= 4) & (qh != 2);
        const uint32_t kl = bl ? (uint32_t)(ql - (ql > 2)) : 0x0Cu;
        const uint32_t kh = bh ? (uint32_t)(qh - (qh > 2)) + 4u : 0x0Cu;
        sel[j] = 0x000C000Cu | (kl << 8) | (kh << 24);
    }
    f16x2 H[C], E[C];
    const f16x2 zero = as_h2(0u);
#pragma unroll
    for (int j = 0; j < C; ++j) {
        H[j] = zero;
        E[j] = zero;
    }
    const f16x2 O = as_h2(open_h * 0x10001u), X = as_h2(ext_h * 0x10001u);
    uint32_t best = thr_h, lo_prev = 0, w_prev = 0, d0 = 0, f0 = 0;
    int32_t cnt = 0, jj_prev = 0;
    int c1 = rows > 0 ? ref[0] : 2, c2 = rows > 1 ? ref[1] : 2;  // two rows ahead
    for (int s = 0; s <= rows; ++s) {
        const int c = c1;
        c1 = c2;
        c2 = s + 2 < rows ? ref[s + 2] : 2;
        const bool base = (c >= 0) & (c <= 4) & (c != 2);
        // high bytes of s(c, A), s(c, C), s(c, G), s(c, T): mismatch but at c's own class
        const uint32_t w = base ? mis4 ^ (diff << (8 * (c - (c > 2)))) : 0u;
        f16x2 diag = as_h2(d0), fc = as_h2(f0), rmax = zero;
#pragma unroll
        for (int j = 0; j < C; ++j) {
            const f16x2 sc = as_h2(perm(w_prev, w, sel[j]));
            const f16x2 tt = diag + sc;
            diag = H[j];
            const f16x2 hv = hmax3(tt, E[j], fc);  // E, fc >= 0: includes the 0 floor
            const f16x2 ho = hv + O;
            E[j] = hmax3(E[j] + X, ho, zero);
            fc = hmax3(ho, fc + X, zero);
            H[j] = hv;
            if (j & 1) rmax = hmax3(rmax, H[j - 1], hv);
            else if (j == C - 1) rmax = __builtin_elementwise_maximum(rmax, hv);
        }
        const uint32_t rb = bits(rmax), lo = rb & 0xffffu, hi = rb >> 16;
        // row s - 1 is complete: its low-half maximum lo_prev, high-half hi
        if (s > 0 && max(lo_prev, hi) > best) {
            int jj = jj_prev;
            if (hi > lo_prev) {
                jj = 0;
#pragma unroll
                for (int j = C - 1; j >= 0; --j)
                    if ((bits(H[j]) >> 16) == hi) jj = C + j;
            }
            best = max(lo_prev, hi);
            record(traj, t.scan, cap, cnt, t.a + s - 1, (int32_t)(float)__builtin_bit_cast(_Float16, (uint16_t)best), jj);
        }
        if (lo > best && s < rows) {  // first column of row s's low-half maximum
            jj_prev = 0;
#pragma unroll
            for (int j = C - 1; j >= 0; --j)
                if ((bits(H[j]) & 0xffffu) == lo) jj_prev = j;
        }
        lo_prev = lo;
        d0 = bits(diag) << 16;  // H(s - 1, C - 1) of the low half
        f0 = bits(fc) << 16;    // F out of column C - 1, row s
        w_prev = w;
    }
    tcount[t.scan] = cnt;
    if (cnt > 0) atomicAdd(&ctr->entries, min(cnt, cap));
    atomicAdd(&ctr->gen_cells, (unsigned long long)rows * (unsigned long long)max(tw, 0));
    }
}

// Reverse SW of getSeeds (:152-159): reverse complements of seq1[e1-r1+1..e1]
// and seq2[q0-r2+1..q0], max and first position, int32 in registers.
template <int WM>
__device__ void reverse_sw(const int8_t *__restrict__ s1, int32_t e1, int32_t r1, const int8_t *__restrict__ s2,
                           int32_t q0, int32_t r2, const Params &P, const uint32_t *shi, const uint32_t *slo,
                           int32_t &best, int32_t &bi, int32_t &bj) {
    uint32_t sel[WM];
    int32_t H[WM], E[WM];
#pragma unroll
    for (int j = 0; j < WM; ++j) {
        sel[j] = sel32(j < r2 ? 4 - s2[q0 - j] : 2);
        H[j] = 0;
        E[j] = 0;
    }
    const int32_t O = P.open, X = P.ext;
    best = 0;
    bi = -1;
    bj = -1;
    for (int i = 0; i < r1; ++i) {
        const int c = 4 - s1[e1 - i];
        const int ci = (unsigned)c < 8u ? c : 2;
        const uint32_t thi = shi[ci], tlo = slo[ci];
        int32_t diag = 0, fc = 0, rmax = 0;
#pragma unroll
        for (int j = 0; j < WM; ++j) {
            const int32_t s = (int32_t)perm(thi, tlo, sel[j]);
            const int32_t t = diag + s;
            diag = H[j];
            const int32_t h = max3i(t, E[j], fc);  // E, fc >= 0: includes the 0 floor
            const int32_t ho = h + O;
            E[j] = max3i(E[j] + X, ho, 0);
            fc = max3i(ho, fc + X, 0);
            H[j] = h;
            if (j & 1) rmax = max3i(rmax, H[j - 1], h);
            else if (j == WM - 1) rmax = max(rmax, h);
        }
        if (rmax > best) {
            int jj = 0;
#pragma unroll
            for (int j = WM - 1; j >= 0; --j)
                if (H[j] == rmax) jj = j;
            best = rmax;
            bi = i;
            bj = jj;
        }
    }
}

// Resolves nodes whose scan trajectory is complete.  level0: the inputs are
// hit windows (node id = global window id); otherwise ScanTasks.
template <int WM>
__global__ __launch_bounds__(64) void k_resolve(const int8_t *__restrict__ seq, const PairDesc *__restrict__ pairs,
                                                const int32_t *__restrict__ pair_of_window, Params P, ScoreTab tab,
                                                int level0, const int32_t *__restrict__ hitlist,
                                                const ScanTask *__restrict__ tasks, const int32_t *__restrict__ n_p,
                                                Node *__restrict__ nodes, int32_t node_cap,
                                                const TrajEntry *__restrict__ traj, const int32_t *__restrict__ tcount,
                                                int32_t cap, int32_t scan_cap, ScanTask *__restrict__ next,
                                                int32_t next_cap, Counters *__restrict__ ctr,
                                                const int32_t *__restrict__ lastrow, int32_t tail_rows) {
    __shared__ uint32_t shi[8], slo[8];
    if (threadIdx.x < 8) {
        shi[threadIdx.x] = tab.hi[threadIdx.x];
        slo[threadIdx.x] = tab.lo[threadIdx.x];
    }
    __syncthreads();
    const int n_in = *n_p;
    for (int gid = blockIdx.x * blockDim.x + threadIdx.x; gid < n_in; gid += gridDim.x * blockDim.x) {
    int32_t node, pair, win;
    if (level0) {
        node = hitlist[gid];
        pair = pair_of_window[node];
        win = node - pairs[pair].win_base;
        nodes[node].a = 0;
        nodes[node].b = pairs[pair].len1;
        nodes[node].scan = node;
    } else {
        const ScanTask t = tasks[gid];
        node = t.node;
        pair = t.pair;
        win = t.win;
    }
    const PairDesc pd = pairs[pair];
    const int8_t *__restrict__ s1 = seq + pd.off1;
    const int8_t *__restrict__ s2 = seq + pd.off2;
    const int sp = win * P.step;
    const int tw = min(P.window, pd.len2 - sp);
    const int32_t scan_end = lastrow[pd.win_base + win] + 1;  // no rescan entry at or past this row
    while (true) {
        Node nd = nodes[node];
        const int cnt = min(tcount[nd.scan], cap);
        const TrajEntry *tr = traj + (int64_t)nd.scan * cap;
        int k = cnt - 1;
        while (k >= 0 && tr[k].row >= nd.b) --k;
        if (k < 0) {
            nd.e1 = -1;
            nd.s1 = nd.s2 = nd.e2 = -1;
            nd.left = nd.right = -1;
            nodes[node] = nd;
            break;
        }
        const int32_t e1 = tr[k].row;
        const int32_t col = tr[k].score_col & 0xff;
        const int32_t r1 = min(1 + (e1 - nd.a), tw + P.window);
        const int32_t r2 = min(1 + col, tw + P.window);
        int32_t rmax, q1, q2;
        reverse_sw<WM>(s1, e1, r1, s2, sp + col, r2, P, shi, slo, rmax, q1, q2);
        nd.e1 = e1;
        nd.e2 = sp + col;
        nd.s1 = e1 - q1;
        nd.s2 = sp + col - q2;
        nd.left = nd.right = -1;
        if (nd.b - e1 - 1 >= P.min_ref && e1 + 1 >= scan_end) {
            // the right child's rows hold no cell of min_score: resolved empty, not scanned
            const int32_t r = atomicAdd(&ctr->nodes, 1);
            if (r < node_cap) {
                Node rn;
                rn.a = e1 + 1;
                rn.b = nd.b;
                rn.scan = -1;
                rn.s1 = rn.e1 = rn.s2 = rn.e2 = -1;
                rn.left = rn.right = -1;
                nodes[r] = rn;
                nd.right = r;
            } else {
                atomicOr(&ctr->overflow, 1);
            }
        } else if (nd.b - e1 - 1 >= P.min_ref) {
            const int32_t r = atomicAdd(&ctr->nodes, 1);
            const int32_t sc = atomicAdd(&ctr->scans, 1);
            const int32_t q = atomicAdd(&ctr->next_tasks, 1);
            if (r < node_cap && sc < scan_cap && q < next_cap) {
                Node rn;
                rn.a = e1 + 1;
                rn.b = nd.b;
                rn.scan = sc;
                rn.s1 = rn.e1 = rn.s2 = rn.e2 = -1;
                rn.left = rn.right = -1;
                nodes[r] = rn;
                ScanTask st;
                st.pair = pair;
                st.win = win;
                st.a = e1 + 1;
                st.b = min(nd.b, scan_end);
                // with level-0 tails the scan runs the rows whose values can
                // differ from level 0's; k_rescan_tail adds the rest
                if (tail_rows > 0) st.b = min(st.b, st.a + tail_rows);
                st.node = r;
                st.scan = sc;
                next[q] = st;
                nd.right = r;
            } else {
                atomicOr(&ctr->overflow, 1);
            }
        }
        int32_t lchild = -1;
        if (nd.s1 - nd.a >= P.min_ref) {
            const int32_t l = atomicAdd(&ctr->nodes, 1);
            if (l < node_cap) {
                Node ln;
                ln.a = nd.a;
                ln.b = nd.s1;
                ln.scan = nd.scan;
                ln.s1 = ln.e1 = ln.s2 = ln.e2 = -1;
                ln.left = ln.right = -1;
                nodes[l] = ln;
                lchild = l;
            } else {
                atomicOr(&ctr->overflow, 1);
            }
        }
        nd.left = lchild;
        nodes[node] = nd;
        if (lchild < 0) break;
        node = lchild;
    }
    }
}

// Counters at the start of a seed phase: the window nodes and trajectory
// slots allocated, everything else zero.
__global__ void k_seed_init(Counters *__restrict__ ctr, int32_t nw) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        Counters c{};
        c.nodes = nw;
        c.scans = nw;
        *ctr = c;
    }
}

// Start of a recursion level: the tasks the previous resolve queued become
// the level's (none once a pool has overflowed, so that every later launch
// of an enqueued sequence is empty), and the order histogram is cleared.
__global__ void k_level_begin(Counters *__restrict__ ctr, int32_t *__restrict__ hist) {
    for (int b = threadIdx.x; b < kOrderBuckets; b += blockDim.x) hist[b] = 0;
    if (threadIdx.x == 0) {
        const int32_t n = ctr->overflow ? 0 : ctr->next_tasks;
        ctr->level_tasks = n;
        ctr->next_tasks = 0;
        if (n > 0) ctr->levels += 1;
    }
}

// Bucket of a scan task in the level order: 16 buckets per doubling of the
// row count, longest first (bucket 0 holds the longest).
__device__ __forceinline__ int order_bucket(int32_t rows) {
    if (rows <= 0) return kOrderBuckets - 1;
    const int lg = 31 - __clz(rows);
    const int frac = lg >= 4 ? (rows >> (lg - 4)) & 15 : (rows << (4 - lg)) & 15;
    return kOrderBuckets - 1 - min(lg * 16 + frac, kOrderBuckets - 1);
}

// The level order: a counting sort of the tasks by bucket (histogram,
// exclusive scan, scatter).  Within a bucket the order is that of the
// scatter's atomics; only the load balance depends on it.
__global__ __launch_bounds__(256) void k_order_count(const ScanTask *__restrict__ tasks, const int32_t *__restrict__ n_p,
                                                     int32_t *__restrict__ hist) {
    __shared__ int32_t h[kOrderBuckets];
    for (int b = threadIdx.x; b < kOrderBuckets; b += blockDim.x) h[b] = 0;
    __syncthreads();
    const int n = *n_p;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        atomicAdd(&h[order_bucket(tasks[i].b - tasks[i].a)], 1);
    __syncthreads();
    for (int b = threadIdx.x; b < kOrderBuckets; b += blockDim.x)
        if (h[b]) atomicAdd(&hist[b], h[b]);
}

__global__ __launch_bounds__(kOrderBuckets) void k_order_scan(int32_t *__restrict__ hist) {
    __shared__ int32_t v[kOrderBuckets];
    const int b = threadIdx.x;
    v[b] = hist[b];
    __syncthreads();
    for (int d = 1; d < kOrderBuckets; d <<= 1) {
        const int32_t x = b >= d ? v[b - d] : 0;
        __syncthreads();
        v[b] += x;
        __syncthreads();
    }
    hist[kOrderBuckets + b] = v[b] - hist[b];  // exclusive start of bucket b
}

__global__ __launch_bounds__(256) void k_order_scatter(const ScanTask *__restrict__ tasks,
                                                       const int32_t *__restrict__ n_p, int32_t *__restrict__ hist,
                                                       int32_t *__restrict__ order) {
    __shared__ int32_t h[kOrderBuckets], base[kOrderBuckets];
    const int n = *n_p;
    for (int i0 = blockIdx.x * blockDim.x; i0 < n; i0 += gridDim.x * blockDim.x) {
        for (int b = threadIdx.x; b < kOrderBuckets; b += blockDim.x) h[b] = 0;
        __syncthreads();
        const int i = i0 + threadIdx.x;
        int bk = 0, rank = 0;
        if (i < n) {
            bk = order_bucket(tasks[i].b - tasks[i].a);
            rank = atomicAdd(&h[bk], 1);
        }
        __syncthreads();
        for (int b = threadIdx.x; b < kOrderBuckets; b += blockDim.x)
            if (h[b]) base[b] = atomicAdd(&hist[kOrderBuckets + b], h[b]);
        __syncthreads();
        if (i < n) order[base[bk] + rank] = i;
        __syncthreads();
    }
}

__global__ void k_widen_i32(const int32_t *__restrict__ a, int32_t n, int64_t *__restrict__ b) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

__global__ void k_fill_i32(int32_t *p, int64_t n, int32_t v) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_gather_roots(const Node *__restrict__ nodes, const int32_t *__restrict__ hitlist,
                               const int32_t *__restrict__ n_p, Node *__restrict__ out) {
    const int n = *n_p;
    for (int gid = blockIdx.x * blockDim.x + threadIdx.x; gid < n; gid += gridDim.x * blockDim.x)
        out[gid] = nodes[hitlist[gid]];
}

// Strong-row index.  Chunk c's entries move to chunk_off[c] onward (one
// workgroup per chunk).
// The sort key is window << 32 | row (segments of one window run at once, so
// a window's entries reach the sink out of row order).
__global__ __launch_bounds__(64) void k_strong_compact(const uint32_t *__restrict__ win,
                                                        const uint64_t *__restrict__ rv,
                                                        const int32_t *__restrict__ fill,
                                                        const int64_t *__restrict__ off, uint64_t *__restrict__ key2,
                                                        uint64_t *__restrict__ rv2) {
    const int64_t c = blockIdx.x;
    const int32_t n = fill[c];
    for (int32_t i = threadIdx.x; i < n; i += blockDim.x) {
        const uint64_t e = rv[c * kStrongChunk + i];
        key2[off[c] + i] = ((uint64_t)win[c * kStrongChunk + i] << 32) | (e >> 32);
        rv2[off[c] + i] = e;
    }
}

// Total entries: the last chunk's offset plus its fill.
__global__ void k_strong_total(const int32_t *__restrict__ fill, const int64_t *__restrict__ off, int32_t nchunks,
                               int64_t *__restrict__ total) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *total = nchunks > 0 ? off[nchunks - 1] + fill[nchunks - 1] : 0;
}

// seg[w] = [first, last + 1) of window w's entries in the sorted list (seg
// cleared to empty beforehand).
__global__ void k_strong_seg(const uint64_t *__restrict__ key, int64_t n, int2 *__restrict__ seg) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t w = (uint32_t)(key[i] >> 32);
    if (i == 0 || (uint32_t)(key[i - 1] >> 32) != w) seg[w].x = (int32_t)i;
    if (i == n - 1 || (uint32_t)(key[i + 1] >> 32) != w) seg[w].y = (int32_t)(i + 1);
}

__device__ __forceinline__ int32_t strong_val(uint64_t e) { return (int32_t)((uint32_t)e >> 8); }
__device__ __forceinline__ int32_t strong_row(uint64_t e) { return (int32_t)(e >> 32); }

// Next strictly larger maximum of each entry within its window, one lane per
// window from the back: the candidates after i are i + 1 and then the
// entries reachable through ng, which skip only values that are not larger.
__global__ void k_strong_ng(const int2 *__restrict__ seg, int32_t nw, const uint64_t *__restrict__ rv,
                            int32_t *__restrict__ ng) {
    const int32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= nw) return;
    const int2 g = seg[w];
    for (int32_t i = g.y - 1; i >= g.x; --i) {
        const int32_t v = strong_val(rv[i]);
        int32_t j = i + 1;
        while (j < g.y && strong_val(rv[j]) <= v) j = ng[j];
        ng[i] = j;
    }
}

// First column of level 0's maximum v on row r of window (pd, win): the
// window's DP restarted from zeros at row max(0, r - delta) has level 0's
// values on row r (the tail bound), so r - r0 + 1 rows of it give the row.
// The same recurrence as every scan kernel, int32, the window's tw real
// columns (a tail entry's maximum is never only in the N padding: a padded
// cell is bounded by an earlier real cell of the same rescan, which the
// running maximum already holds).
template <int WM>
__device__ int32_t l0_row_col(const int8_t *__restrict__ s1, const int8_t *__restrict__ win, int32_t tw, int32_t r,
                              int32_t delta, const Params &P, int32_t v) {
    int32_t H[WM], E[WM], q[WM];
#pragma unroll
    for (int j = 0; j < WM; ++j) {
        const int32_t c = j < tw ? (int32_t)win[j] : 2;
        q[j] = (c >= 0 && c <= 4 && c != 2) ? c : -1;
        H[j] = 0;
        E[j] = 0;
    }
    for (int32_t i = max(0, r - delta); i <= r; ++i) {
        const int32_t c0 = s1[i];
        const int32_t c = (c0 >= 0 && c0 <= 4 && c0 != 2) ? c0 : -2;
        int32_t diag = 0, f = 0;
#pragma unroll
        for (int j = 0; j < WM; ++j) {
            const int32_t sc = (q[j] < 0 || c < 0) ? 0 : (q[j] == c ? P.match : P.mismatch);
            const int32_t h = max(max(diag + sc, E[j]), max(f, 0));
            const int32_t ho = h + P.open;
            E[j] = max(max(E[j] + P.ext, ho), 0);
            f = max(max(ho, f + P.ext), 0);
            diag = H[j];
            H[j] = h;
        }
    }
    int32_t col = kStrongNoCol;
#pragma unroll
    for (int j = WM - 1; j >= 0; --j)
        if (j < tw && H[j] == v) col = j;
    return col;
}

// A rescan task's trajectory continued with level 0's strong rows: rows at or
// past a + delta, before the rescan's end (its node's b, cut at the window's
// lastrow + 1 as the scan was), that exceed the running maximum; the first
// such entry is found by following ng from the first entry on or past
// a + delta, and so is each next one.  Every step raises the maximum, so a
// task takes at most about 2 * cap steps.  An entry without its column gets
// it from l0_row_col.
template <int WM>
__global__ __launch_bounds__(64) void k_rescan_tail(const int8_t *__restrict__ seq, const PairDesc *__restrict__ pairs,
                                                    Params P, const ScanTask *__restrict__ tasks,
                                                    const int32_t *__restrict__ n_p, const Node *__restrict__ nodes,
                                                    const int32_t *__restrict__ lastrow, StrongIndex I,
                                                    TrajEntry *__restrict__ traj, int32_t *__restrict__ tcount,
                                                    int32_t cap, Counters *__restrict__ ctr) {
    const int n = *n_p;
    for (int gid = blockIdx.x * blockDim.x + threadIdx.x; gid < n; gid += gridDim.x * blockDim.x) {
        const ScanTask t = tasks[gid];
        const PairDesc pd = pairs[t.pair];
        const int32_t gw = pd.win_base + t.win;
        const int32_t end = min(nodes[t.node].b, lastrow[gw] + 1);
        const int32_t from = t.a + I.delta;
        if (end <= from) continue;
        const int2 g = I.seg[gw];
        // first entry with row >= from
        int32_t lo = g.x, hi = g.y;
        while (lo < hi) {
            const int32_t mid = (lo + hi) >> 1;
            if (strong_row(I.rv[mid]) < from) lo = mid + 1;
            else hi = mid;
        }
        int32_t cnt = tcount[t.scan];
        TrajEntry *tr = traj + (int64_t)t.scan * cap;
        int32_t best = cnt > 0 ? (tr[min(cnt, cap) - 1].score_col >> 8) : P.min_score - 1;
        const int32_t c0 = cnt;
        const int sp = t.win * P.step;
        const int tw = min(P.window, pd.len2 - sp);
        int32_t p = lo;
        while (p < g.y) {
            const uint64_t e = I.rv[p];
            const int32_t row = strong_row(e);
            if (row >= end) break;
            const int32_t v = strong_val(e);
            if (v > best) {
                if (cnt < cap) {
                    int32_t col = (int32_t)(e & 0xffu);
                    if (col == kStrongNoCol)
                        col = l0_row_col<WM>(seq + pd.off1, seq + pd.off2 + sp, tw, row, I.delta, P, v);
                    TrajEntry x;
                    x.row = row;
                    x.score_col = (v << 8) | col;
                    tr[cnt] = x;
                }
                ++cnt;
                best = v;
            }
            p = I.ng[p];
        }
        if (cnt != c0) {
            tcount[t.scan] = cnt;
            atomicAdd(&ctr->entries, min(cnt, cap) - min(c0, cap));
        }
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Launchers (host side).
// ---------------------------------------------------------------------------
hipError_t launch_scan_uniform(hipStream_t st, int wm, const int8_t *seq, const PairDesc *pairs, const int2 *wtasks,
                               int32_t nwt, int32_t persistent_waves, const Params &P, const ScoreTab &tab,
                               TrajEntry *traj, int32_t *tcount, int32_t cap, int32_t *hitlist, Counters *ctr,
                               const int2 *wins, int32_t nwins, int32_t *lastrow, const StrongSink &sink) {
    if (nwt <= 0 && nwins <= 0) return hipSuccess;
    // lane tasks first (persistent waves popping the task list from the
    // counter block's queue head, zeroed by k_seed_init, or one workgroup per
    // task), then the one-per-wave windows in the last nwins workgroups; the
    // kernels without the wave path (WAVES false) keep the lane kernels' own
    // register allocation
    const bool persist = persistent_waves > 0;
    int32_t *queue = &ctr->queue;
    const dim3 g((unsigned)nwins + (persist ? (unsigned)std::min(persistent_waves, nwt) : (unsigned)std::max(nwt, 0))),
        b(64);
    const bool waves = nwins > 0;
    if (scan_uses_f16(P)) {
        const ScoreTab th = make_score_tab_f16(P.match, P.mismatch);
        const uint32_t oh = f16_bits_of_int(P.open), xh = f16_bits_of_int(P.ext), thr = f16_bits_of_int(P.min_score - 1);
        // REC: the strong rows go to the sink (level-0 tails on); its own
        // instantiation, so that the others carry no trace of it
        const bool rec = sink.cap_chunks > 0;
        decltype(&k_scan_uniform_h<38, false, false>) k;
#define DCNS_L0_PICK(W, R)                                                                                          \
    (persist ? (waves ? k_scan_uniform_hp<W, true, R> : k_scan_uniform_hp<W, false, R>)                             \
             : (waves ? k_scan_uniform_h<W, true, R> : k_scan_uniform_h<W, false, R>))
        if (wm <= 38)
            k = rec ? DCNS_L0_PICK(38, true) : DCNS_L0_PICK(38, false);
        else
            k = rec ? DCNS_L0_PICK(64, true) : DCNS_L0_PICK(64, false);
#undef DCNS_L0_PICK
        hipLaunchKernelGGL(k, g, b, 0, st, seq, pairs, wtasks, nwt, queue, P, th, traj, tcount, cap, hitlist, ctr, oh,
                           xh, thr, tab, wins, nwins, lastrow, sink);
    } else {
        decltype(&k_scan_uniform<38, false, false>) k;
        if (wm <= 38)
            k = persist ? (waves ? k_scan_uniform<38, true, true> : k_scan_uniform<38, true, false>)
                        : (waves ? k_scan_uniform<38, false, true> : k_scan_uniform<38, false, false>);
        else
            k = persist ? (waves ? k_scan_uniform<64, true, true> : k_scan_uniform<64, true, false>)
                        : (waves ? k_scan_uniform<64, false, true> : k_scan_uniform<64, false, false>);
        hipLaunchKernelGGL(k, g, b, 0, st, seq, pairs, wtasks, nwt, queue, P, tab, traj, tcount, cap, hitlist, ctr,
                           wins, nwins, lastrow);
    }
    return hipGetLastError();
}

hipError_t launch_scan_general(hipStream_t st, int wm, const int8_t *seq, const PairDesc *pairs,
                               const ScanTask *tasks, const int32_t *order, const int32_t *ntasks_dev,
                               int32_t ntasks_max, int32_t wave_tasks, int32_t wave_rows, int32_t lane_blocks_max,
                               const Params &P, const ScoreTab &tab, TrajEntry *traj, int32_t *tcount, int32_t cap,
                               Counters *ctr) {
    if (ntasks_max <= 0) return hipSuccess;
    // the long tasks sort first: up to wave_tasks of them get a wave each (a
    // wave step costs ~11x the SIMD time per cell of the lane path)
    const int32_t wave_blocks = std::min<int32_t>(ntasks_max, wave_tasks);
    int32_t lane_blocks = (ntasks_max + 63) / 64;
    if (lane_blocks_max > 0) lane_blocks = std::min(lane_blocks, lane_blocks_max);
    const dim3 g(wave_blocks + lane_blocks), b(64);
    uint32_t mis4, diff;
    if (scan_general_packed(P, mis4, diff)) {
        const uint32_t oh = f16_bits_of_int(P.open), xh = f16_bits_of_int(P.ext), thr = f16_bits_of_int(P.min_score - 1);
        if (wm <= 38)
            hipLaunchKernelGGL(k_scan_general_h<38>, g, b, 0, st, seq, pairs, tasks, order, ntasks_dev, wave_blocks,
                               wave_rows, P, tab, traj, tcount, cap, ctr, oh, xh, thr, mis4, diff);
        else
            hipLaunchKernelGGL(k_scan_general_h<64>, g, b, 0, st, seq, pairs, tasks, order, ntasks_dev, wave_blocks,
                               wave_rows, P, tab, traj, tcount, cap, ctr, oh, xh, thr, mis4, diff);
        return hipGetLastError();
    }
    if (wm <= 38)
        hipLaunchKernelGGL(k_scan_general<38>, g, b, 0, st, seq, pairs, tasks, order, ntasks_dev, wave_blocks, wave_rows,
                           P, tab, traj, tcount, cap, ctr);
    else
        hipLaunchKernelGGL(k_scan_general<64>, g, b, 0, st, seq, pairs, tasks, order, ntasks_dev, wave_blocks, wave_rows,
                           P, tab, traj, tcount, cap, ctr);
    return hipGetLastError();
}

hipError_t launch_resolve(hipStream_t st, int wm, const int8_t *seq, const PairDesc *pairs,
                          const int32_t *pair_of_window, const Params &P, const ScoreTab &tab, int level0,
                          const int32_t *hitlist, const ScanTask *tasks, const int32_t *n_dev, int32_t n_max,
                          int32_t blocks_max, Node *nodes, int32_t node_cap, const TrajEntry *traj,
                          const int32_t *tcount, int32_t cap, int32_t scan_cap, ScanTask *next, int32_t next_cap,
                          Counters *ctr, const int32_t *lastrow, int32_t tail_rows) {
    if (n_max <= 0) return hipSuccess;
    int32_t blocks = (n_max + 63) / 64;
    if (blocks_max > 0) blocks = std::min(blocks, blocks_max);
    dim3 g(blocks), b(64);
    if (wm <= 38)
        hipLaunchKernelGGL(k_resolve<38>, g, b, 0, st, seq, pairs, pair_of_window, P, tab, level0, hitlist, tasks, n_dev,
                           nodes, node_cap, traj, tcount, cap, scan_cap, next, next_cap, ctr, lastrow, tail_rows);
    else
        hipLaunchKernelGGL(k_resolve<64>, g, b, 0, st, seq, pairs, pair_of_window, P, tab, level0, hitlist, tasks, n_dev,
                           nodes, node_cap, traj, tcount, cap, scan_cap, next, next_cap, ctr, lastrow, tail_rows);
    return hipGetLastError();
}

hipError_t launch_seed_init(hipStream_t st, Counters *ctr, int32_t nw) {
    hipLaunchKernelGGL(k_seed_init, dim3(1), dim3(64), 0, st, ctr, nw);
    return hipGetLastError();
}

hipError_t launch_level_begin(hipStream_t st, Counters *ctr, int32_t *hist) {
    hipLaunchKernelGGL(k_level_begin, dim3(1), dim3(256), 0, st, ctr, hist);
    return hipGetLastError();
}

hipError_t launch_order_tasks(hipStream_t st, const ScanTask *tasks, const int32_t *n_dev, int32_t n_max,
                              int32_t *hist, int32_t *order) {
    if (n_max <= 0) return hipSuccess;
    const unsigned blocks = (unsigned)std::min<int32_t>((n_max + 255) / 256, 1024);
    hipLaunchKernelGGL(k_order_count, dim3(blocks), dim3(256), 0, st, tasks, n_dev, hist);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_order_scan, dim3(1), dim3(kOrderBuckets), 0, st, hist);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_order_scatter, dim3(blocks), dim3(256), 0, st, tasks, n_dev, hist, order);
    return hipGetLastError();
}

hipError_t launch_fill_i32(hipStream_t st, int32_t *p, int64_t n, int32_t v) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

hipError_t launch_scan_uniform_seg(hipStream_t st, int wm, const int8_t *seq, const PairDesc *pairs, const SegL0 &sg,
                                   int32_t ntasks, const Params &P, int32_t cap, Counters *ctr, int32_t *lastrow,
                                   const StrongSink &sink) {
    if (ntasks <= 0) return hipSuccess;
    if (!scan_uses_f16(P)) return hipErrorInvalidValue;
    const ScoreTab th = make_score_tab_f16(P.match, P.mismatch);
    const uint32_t oh = f16_bits_of_int(P.open), xh = f16_bits_of_int(P.ext), thr = f16_bits_of_int(P.min_score - 1);
    const bool rec = sink.cap_chunks > 0;
    decltype(&k_scan_uniform_seg<38, false>) k =
        wm <= 38 ? (rec ? k_scan_uniform_seg<38, true> : k_scan_uniform_seg<38, false>)
                 : (rec ? k_scan_uniform_seg<64, true> : k_scan_uniform_seg<64, false>);
    hipLaunchKernelGGL(k, dim3((unsigned)ntasks), dim3(64), 0, st, seq, pairs, sg, P, th, cap, ctr, oh, xh, thr, lastrow,
                       sink);
    return hipGetLastError();
}

hipError_t launch_l0_merge(hipStream_t st, const PairDesc *pairs, const SegL0 &sg, int32_t ngroups, TrajEntry *traj,
                           int32_t *tcount, int32_t cap, int32_t min_score, int32_t *hitlist, Counters *ctr) {
    if (ngroups <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_l0_merge, dim3((unsigned)ngroups), dim3(128), 0, st, pairs, sg, ngroups, traj, tcount, cap,
                       min_score, hitlist, ctr);
    return hipGetLastError();
}

size_t strong_index_scratch(int64_t nchunks, int64_t n) {
    size_t a = 0, b = 0;
    hipcub::DeviceScan::ExclusiveSum(nullptr, a, (const int64_t *)nullptr, (int64_t *)nullptr,
                                     (int)std::max<int64_t>(nchunks, 1));
    hipcub::DeviceRadixSort::SortPairs(nullptr, b, (const uint64_t *)nullptr, (uint64_t *)nullptr,
                                       (const uint64_t *)nullptr, (uint64_t *)nullptr, (int)std::max<int64_t>(n, 1));
    return std::max(a, b) + 256;
}

hipError_t launch_strong_count(hipStream_t st, const StrongSink &S, int32_t nchunks, void *scratch,
                               size_t scratch_bytes, int64_t *fill64, int64_t *chunk_off, int64_t *total_dev) {
    if (nchunks <= 0) {
        hipLaunchKernelGGL(k_strong_total, dim3(1), dim3(64), 0, st, S.chunk_fill, chunk_off, 0, total_dev);
        return hipGetLastError();
    }
    // chunk fills widen to int64, then their exclusive scan into chunk_off
    hipLaunchKernelGGL(k_widen_i32, dim3((unsigned)((nchunks + 255) / 256)), dim3(256), 0, st, S.chunk_fill, nchunks,
                       fill64);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    size_t bytes = scratch_bytes;
    e = hipcub::DeviceScan::ExclusiveSum(scratch, bytes, fill64, chunk_off, nchunks, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_strong_total, dim3(1), dim3(64), 0, st, S.chunk_fill, chunk_off, nchunks, total_dev);
    return hipGetLastError();
}

hipError_t launch_strong_index(hipStream_t st, const StrongSink &S, int32_t nchunks, int64_t n, int32_t nw,
                               int32_t win_bits, void *scratch, size_t scratch_bytes, uint64_t *key2, uint64_t *rv2,
                               uint64_t *key, int64_t *chunk_off, int2 *seg, int32_t *ng) {
    hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)((2 * (int64_t)nw + 255) / 256)), dim3(256), 0, st, (int32_t *)seg,
                       2 * (int64_t)nw, 0);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || n <= 0) return e;
    // the chunks' entries, compacted into (key2, rv2), sort by (window, row)
    // into (key, the sink's rv array), which then holds the index's entries
    hipLaunchKernelGGL(k_strong_compact, dim3((unsigned)nchunks), dim3(64), 0, st, S.win, S.rv, S.chunk_fill,
                       chunk_off, key2, rv2);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    size_t bytes = scratch_bytes;
    e = hipcub::DeviceRadixSort::SortPairs(scratch, bytes, key2, key, rv2, S.rv, (int)n, 0, 32 + win_bits, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_strong_seg, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, key, n, seg);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_strong_ng, dim3((unsigned)((nw + 63) / 64)), dim3(64), 0, st, seg, nw, S.rv, ng);
    return hipGetLastError();
}

hipError_t launch_rescan_tail(hipStream_t st, int wm, const int8_t *seq, const PairDesc *pairs, const Params &P,
                              const ScanTask *tasks, const int32_t *n_dev, int32_t n_max, int32_t blocks_max,
                              const Node *nodes, const int32_t *lastrow, const StrongIndex &I, TrajEntry *traj,
                              int32_t *tcount, int32_t cap, Counters *ctr) {
    if (n_max <= 0) return hipSuccess;
    int32_t blocks = (n_max + 63) / 64;
    if (blocks_max > 0) blocks = std::min(blocks, blocks_max);
    if (wm <= 38)
        hipLaunchKernelGGL(k_rescan_tail<38>, dim3(blocks), dim3(64), 0, st, seq, pairs, P, tasks, n_dev, nodes, lastrow,
                           I, traj, tcount, cap, ctr);
    else
        hipLaunchKernelGGL(k_rescan_tail<64>, dim3(blocks), dim3(64), 0, st, seq, pairs, P, tasks, n_dev, nodes, lastrow,
                           I, traj, tcount, cap, ctr);
    return hipGetLastError();
}

hipError_t launch_gather_roots(hipStream_t st, const Node *nodes, const int32_t *hitlist, const int32_t *n_dev,
                               int32_t n_max, Node *out) {
    if (n_max <= 0) return hipSuccess;
    const unsigned blocks = (unsigned)std::min<int32_t>((n_max + 255) / 256, 4096);
    hipLaunchKernelGGL(k_gather_roots, dim3(blocks), dim3(256), 0, st, nodes, hitlist, n_dev, out);
    return hipGetLastError();
}

}  // namespace dcns
