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
// Device layout: columns are cut into strips of 64 lanes x 4 columns and
// every strip of every task is one wave, so one long task spreads over as
// many CUs as it has strips.  Inside a strip lane l works on row r = t - l at
// step t (an anti-diagonal skew), so the values a row carries from left to
// right (M, the two horizontal gap states, the running row maximum, the
// reference base) move one lane per step by a DPP wave shift.  The strip's
// last lane puts each row's carried values in LDS, and every 64 steps the
// wave stores that block of records to the strip's boundary array (one store
// per field) and publishes how many boundary rows are complete (a progress
// word); the next strip waits for the rows its first lane needs in its next
// 64 steps.  Every row in
// the band is computed (the reference's termination test usually never
// fires with the default long-gap piece, whose column state keeps every later
// row within open2 of the best score); the finish kernel cuts the rows at the
// reference's termination row.  Trace bytes are stored step-major per strip
// (64 lanes x 4 bytes per step, one coalesced 256-byte store), 255 outside the
// band as after Matrix::reset.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <type_traits>
#include <vector>

#include "dp_common.h"
#include "kernels.h"

namespace dcns {
namespace {

constexpr int K = 4;                 // columns per lane
constexpr int SW = 64 * K;           // strip width
constexpr int32_t NEG = -536870912;  // SCORE_OUT_BANDED_ALIGNMENT_REGION
constexpr int BND = 5;               // ints per boundary record: M, e1, e2, row max, row max column
// Longest wait for a predecessor strip, in wall-clock ticks (100 MHz): far
// beyond any legitimate wait, which is the time the predecessor needs for 64
// rows plus its own start.
constexpr uint64_t kWaitTicks = 500000000ull;

__device__ __forceinline__ int32_t code_at(const int8_t *base, int32_t a, int32_t k, int32_t rev) {
    // k-th base (0-based) of the forward suffix starting at a, or of the
    // reverse complement of the prefix ending at a
    return rev ? 4 - (int32_t)base[a - k] : (int32_t)base[a + k];
}

// Boundary records and row results cross CUs (and XCD L2s): agent-scope
// atomic accesses.
__device__ __forceinline__ int32_t ald(const int32_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void ast(int32_t *p, int32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Strip-to-strip hand-over.  Every value one strip reads from another (the
// boundary records, the progress words, the row results) is written and read
// with agent-scope atomic accesses, which the gfx9.5 memory model makes
// coherent across CUs and XCD L2s by themselves.  An agent-scope release or
// acquire would add an L2 write-back before each progress store and an L2
// invalidate after each progress load; both exist for plain (non-atomic)
// data, of which strips exchange none, and the write-back of every strip's
// trace bytes every 64 steps costs more than the DP.  So the hand-over is the
// remaining part of release / acquire: the producer waits for all its vector
// memory operations to complete (its boundary stores are then visible at
// agent scope) before it stores the progress word, and the consumer issues
// its boundary loads only after it has observed the progress word.
__device__ __forceinline__ void publish_rows(int32_t *prog, int32_t rows) {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0): earlier loads and stores complete
    __hip_atomic_store(prog, rows, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// Waits until *prog >= need.  False when another strip has failed or the
// wait exceeds kWaitTicks; the caller then abandons its strip.
__device__ bool wait_rows(const int32_t *prog, int32_t need, const int32_t *err) {
    bool ok = ald(prog) >= need;
    if (!ok) {
        const uint64_t t0 = wall_clock64();
        for (;;) {
            __builtin_amdgcn_s_sleep(2);
            if (ald(prog) >= need) {
                ok = true;
                break;
            }
            if (ald(err) != 0 || wall_clock64() - t0 > kWaitTicks) break;
        }
    }
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    return ok;
}

// Scores of one column as v_perm_b32 sources: the score against reference
// code 0 and 1 in bytes 1 and 3 of lo, against 3 and 4 in bytes 1 and 3 of hi
// (N scores 0 and has no byte).  row_selector(rc) then takes the row's byte
// into byte 0 and replicates its sign into bytes 1-3 (selectors 8-11 are the
// signs of bytes 1, 3, 5, 7), or yields 0 for N (selector 12): one v_perm
// gives the sign-extended score.
__device__ __forceinline__ void column_table(int32_t qc, int32_t match, int32_t mismatch, uint32_t &lo, uint32_t &hi) {
    auto s = [&](int32_t rc) { return (uint32_t)((qc == 2) ? 0 : (rc == qc ? match : mismatch)) & 0xffu; };
    lo = (s(0) << 8) | (s(1) << 24);
    hi = (s(3) << 8) | (s(4) << 24);
}
__device__ __forceinline__ uint32_t row_selector(int32_t rc) {
    return rc == 0 ? 0x08080801u : rc == 1 ? 0x09090903u : rc == 3 ? 0x0A0A0A05u : rc == 4 ? 0x0B0B0B07u : 0x0C0C0C0Cu;
}
// The same lookup one byte up: byte 0 is zero (selector 12), so the v_perm
// yields the score times 256, the scale of the tagged values below.
__device__ __forceinline__ uint32_t row_selector_tag(int32_t rc) {
    return rc == 0 ? 0x0808010Cu : rc == 1 ? 0x0909030Cu : rc == 3 ? 0x0A0A050Cu : rc == 4 ? 0x0B0B070Cu : 0x0C0C0C0Cu;
}

// Tagged values (the traceback kernel where the keyed row maximum applies).
// A DP value v is held as v * 256 plus a low byte: bits 4-6 name the state
// that holds it (0 the diagonal term, 1 F1, 2 e1, 3 F2, 4 e2, the reference's
// order of the five-way maximum) and bits 0-3 are the extension flags of the
// gap states (bit 0 F1, 1 e1, 2 F2, 3 e2), each set only in its own state.
// So the maximum of the five tagged values is the reference's maximum with
// its tie order (the later source wins on equal scores), and its bits 4-6 are
// the source the trace byte records.  A new gap state is the maximum of its
// extension, which carries its flag, and its opening, which does not: on
// equal scores the extension wins, as the reference's test "extension >=
// opening" does, and the state's flag is that test's outcome.  The trace
// byte is then source << 4 | the four new states' flags, with no compare
// and select per bit.  Out-of-band values are NEGK (-2^22 * 256) with their
// state's tag: in-band scores stay below 2^21 in magnitude wherever the keyed
// row maximum is used, every in-band cell has an in-band predecessor, and
// the out-of-band values that in-band cells read are NEGK plus one step's
// penalties, so no maximum or test that an in-band cell takes differs from
// the reference's with its NEG.
constexpr int32_t NEGK = -(1 << 30);
constexpr int32_t kTagF1 = 0x10, kTagE1 = 0x20, kTagF2 = 0x30, kTagE2 = 0x40;

// Keyed row maximum: (maximum, first column) as one int, maximum << 9 | a
// tie rank, so that one v_max keeps the first strict maximum.  Inside a strip
// a local column lc (0..255) ranks 510 - lc; the value carried in from the
// strips to the left ranks 511 and keeps its column beside the key.  The host
// uses it only where every in-band score fits 22 bits.
constexpr int kKeyShift = 9;
constexpr int32_t kKeyIn = 511;
__device__ __forceinline__ int32_t key_max(int32_t key) { return key >> kKeyShift; }
__device__ __forceinline__ int32_t key_col(int32_t key, int32_t col_in, int32_t jb) {
    const int32_t rank = key & kKeyIn;
    return rank == kKeyIn ? col_in : jb + (kKeyIn - 1) - rank;
}

// Lane-0 input of row r of a strip: the previous strip's boundary record,
// or the left edge of the matrix (M(0,0) = 0, M(r,0) = NEG); and the
// reference code of row r.
struct StripIn {
    const int32_t *bin;  // previous strip's records, row plo first
    int32_t first, plo, phi, R, a1, rev;
    bool tag;            // tagged values (see NEGK)
    const int8_t *s1;
    const uint8_t *wc;   // weighted: profile indices of reference positions -1 .. len1 (wc[x + 1])
    __device__ __forceinline__ void load(int32_t r, int32_t &m, int32_t &e1, int32_t &e2, int32_t &rm, int32_t &rj,
                                         int32_t &rc) const {
        m = tag ? NEGK : NEG;
        e1 = tag ? NEGK | kTagE1 : NEG;
        e2 = tag ? NEGK | kTagE2 : NEG;
        rm = rj = 0;
        rc = 2;
        if (first) {
            if (r == 0) m = 0;
        } else if (r >= plo && r <= phi) {
            const int32_t *b = bin + (int64_t)(r - plo) * BND;
            m = ald(b);
            e1 = ald(b + 1);
            e2 = ald(b + 2);
            rm = ald(b + 3);
            rj = ald(b + 4);
        }
        if (r >= 1 && r <= R) {
            rc = code_at(s1, a1, r - 1, rev);
            if (wc) {
                // row r's profile (its reference base) in bits 4-11, the next
                // row's (the vertical gap penalties) in bits 12-19
                const int32_t pos = rev ? a1 - (r - 1) : a1 + r - 1, pn = rev ? pos - 1 : pos + 1;
                rc |= ((int32_t)wc[pos + 1] << 4) | ((int32_t)wc[pn + 1] << 12);
            }
        }
    }
};

// One step of a lane's K columns on row r: scores from the column tables,
// the five-way maximum, the row maximum, the trace byte and the gap states.
// In-band columns are lo <= c < hi, and c == hi is the band's last column.
// Columns left of the band start never re-enter the band, so their state may
// hold anything; columns right of the band end keep the NEG vertical state
// they start with until they become the band's last column, which takes the
// maximum only.  EDGE = false: every column is inside the band.
// KEYED: rm is the keyed row maximum (rank base kc for column 0 of the lane)
// and rj is unused.
template <bool TB, bool EDGE, bool KEYED>
__device__ __forceinline__ void columns(const uint32_t (&tlo)[K], const uint32_t (&thi)[K], int32_t (&F1)[K],
                                        int32_t (&F2)[K], int32_t (&Mp)[K], uint32_t sel, int32_t lo, int32_t hi,
                                        int32_t jfirst, int32_t kc, int32_t o1, int32_t x1, int32_t o2, int32_t x2,
                                        int32_t dg, int32_t &e1, int32_t &e2, int32_t &rm, int32_t &rj,
                                        uint32_t &tword) {
#pragma unroll
    for (int c = 0; c < K; ++c) {
        const int32_t sc = (int32_t)perm(thi[c], tlo[c], sel);
        const int32_t m0 = sc + dg;
        dg = Mp[c];
        const int32_t f1 = F1[c], f2o = F2[c];
        const int32_t m = max(max(max(max(m0, f1), e1), f2o), e2);
        const bool left = EDGE && c < lo, inner = !EDGE || c < hi;
        const bool inb = EDGE ? (!left & (c <= hi)) : true;
        if (KEYED) {
            const int32_t key = (int32_t)(((uint32_t)m << kKeyShift) | (uint32_t)(kc - c));
            rm = inb ? max(rm, key) : rm;
        } else {
            const bool up = inb & (m > rm);
            rm = up ? m : rm;
            rj = up ? jfirst + c : rj;
        }
        const int32_t h = m + o1, f = f1 + x1, ne1 = e1 + x1;
        const int32_t h2 = m + o2, f2 = f2o + x2, ne2 = e2 + x2;
        if (TB) {
            // source: the last of diag, F1, e1, F2, e2 that reaches the maximum
            int32_t src = (f1 == m) ? 1 : 0;
            src = (e1 == m) ? 2 : src;
            src = (f2o == m) ? 3 : src;
            src = (e2 == m) ? 4 : src;
            const int32_t bits =
                ((f >= h) ? 8 : 0) | ((ne1 >= h) ? 16 : 0) | ((f2 >= h2) ? 32 : 0) | ((ne2 >= h2) ? 64 : 0);
            const uint32_t byte = (!left & inner) ? (uint32_t)(src | bits) : 0xffu;
            tword = c == 0 ? byte : (tword | (byte << (8 * c)));
        }
        F1[c] = inner ? max(f, h) : f1;
        F2[c] = inner ? max(f2, h2) : f2o;
        e1 = left ? NEG : max(ne1, h);
        e2 = left ? NEG : max(ne2, h2);
        Mp[c] = m;
    }
}

// columns() with traceback on tagged values (see NEGK): the same cells and
// band rules; o1f/o1e/o2f/o2e are the openings times 256 with the tag of the
// state they open (F1, e1, F2, e2), x1s/x2s the extensions times 256, and kc
// the keyed row maximum's rank of the lane's column 0.  Stored diagonals (Mp)
// and the carried M are untagged (low byte 0).
template <bool EDGE, bool X2Z>
__device__ __forceinline__ void columns_tag(const uint32_t (&tlo)[K], const uint32_t (&thi)[K], int32_t (&F1)[K],
                                            int32_t (&F2)[K], int32_t (&Mp)[K], uint32_t sel, int32_t lo, int32_t hi,
                                            int32_t kc, int32_t o1f, int32_t o1e, int32_t x1s, int32_t o2f,
                                            int32_t o2e, int32_t x2s, int32_t dg, int32_t &e1, int32_t &e2,
                                            int32_t &rm, uint32_t &tword) {
    uint32_t bc[K];
#pragma unroll
    for (int c = 0; c < K; ++c) {
        const int32_t m0 = (int32_t)perm(thi[c], tlo[c], sel) + dg;
        dg = Mp[c];
        const int32_t f1 = F1[c], f2o = F2[c];
        const int32_t mk = max(max(max(max(m0, f1), e1), f2o), e2);
        const int32_t mc = mk & ~0xff;
        const bool left = EDGE && c < lo, inner = !EDGE || c < hi;
        const bool inb = EDGE ? (!left & (c <= hi)) : true;
        // keyed row maximum: score << 9 | rank = mc * 2 + rank
        const int32_t key = (int32_t)(((uint32_t)mc << 1) + (uint32_t)(kc - c));
        rm = inb ? max(rm, key) : rm;
        const int32_t f = (f1 | 1) + x1s, ne1 = (e1 | 2) + x1s;
        const int32_t f2 = X2Z ? (f2o | 4) : (f2o | 4) + x2s, ne2 = X2Z ? (e2 | 8) : (e2 | 8) + x2s;
        const int32_t nf1 = max(f, mc + o1f), ne1n = max(ne1, mc + o1e);
        const int32_t nf2 = max(f2, mc + o2f), ne2n = max(ne2, mc + o2e);
        const uint32_t b = ((uint32_t)(nf1 | ne1n | nf2 | ne2n) & 0x0fu) | ((uint32_t)mk & ~0x0fu);
        bc[c] = (!left & inner) ? b : 0xffu;
        F1[c] = inner ? nf1 : f1;
        F2[c] = inner ? nf2 : f2o;
        e1 = left ? (NEGK | kTagE1) : ne1n;
        e2 = left ? (NEGK | kTagE2) : ne2n;
        Mp[c] = mc;
    }
    // byte 0 of each column's value, columns 0..3 in bytes 0..3
    tword = perm(bc[1], bc[0], 0x0C0C0400u) | perm(bc[3], bc[2], 0x04000C0Cu);
}

// The weighted counterpart of columns() (SemiGlobal with a Score profile per
// reference base, sequenceAlignment3.cpp:1483-1706; single gap piece when
// TWO is false, :1710-1890).  Substitution scores, horizontal gap penalties
// and the row's profile come from this row's base, the vertical gap
// penalties (vo*, vx*) from the next row's.  Every band column, the last one
// included, is a full cell; the last column's vertical states start from NEG
// on every row (F[end] is reset before the row).
template <bool TB, bool EDGE, bool TWO>
__device__ __forceinline__ void columns_w(const int32_t (&qc)[K], int32_t (&F1)[K], int32_t (&F2)[K],
                                          int32_t (&Mp)[K], int32_t rc, int32_t match, int32_t mis, int32_t lo,
                                          int32_t hi, int32_t jfirst, int32_t o1, int32_t x1, int32_t o2, int32_t x2,
                                          int32_t vo1, int32_t vx1, int32_t vo2, int32_t vx2, int32_t dg, int32_t &e1,
                                          int32_t &e2, int32_t &rm, int32_t &rj, uint32_t &tword) {
#pragma unroll
    for (int c = 0; c < K; ++c) {
        const int32_t sc = (qc[c] == 2) ? 0 : (qc[c] == rc ? match : mis);
        const int32_t m0 = sc + dg;
        dg = Mp[c];
        const bool last = EDGE && c == hi;
        const int32_t f1 = last ? NEG : F1[c];
        const int32_t f2o = TWO ? (last ? NEG : F2[c]) : NEG;
        const int32_t m = TWO ? max(max(max(m0, f1), max(e1, f2o)), e2) : max(max(m0, f1), e1);
        const bool left = EDGE && c < lo;
        const bool inb = EDGE ? (!left & (c <= hi)) : true;
        const bool up = inb & (m > rm);
        rm = up ? m : rm;
        rj = up ? jfirst + c : rj;
        const int32_t hv = m + vo1, f = f1 + vx1, hh = m + o1, ne1 = e1 + x1;
        const int32_t hv2 = m + vo2, f2 = f2o + vx2, hh2 = m + o2, ne2 = e2 + x2;
        if (TB) {
            // source: the last of diag, F1, e1(, F2, e2) that reaches the maximum
            int32_t src = (f1 == m) ? 1 : 0;
            src = (e1 == m) ? 2 : src;
            int32_t bits = ((f >= hv) ? 8 : 0) | ((ne1 >= hh) ? 16 : 0);
            if (TWO) {
                src = (f2o == m) ? 3 : src;
                src = (e2 == m) ? 4 : src;
                bits |= ((f2 >= hv2) ? 32 : 0) | ((ne2 >= hh2) ? 64 : 0);
            }
            const uint32_t byte = inb ? (uint32_t)(src | bits) : 0xffu;
            tword = c == 0 ? byte : (tword | (byte << (8 * c)));
        }
        // columns right of the band keep the NEG vertical states they start with
        const bool upd = EDGE ? (c <= hi) : true;
        F1[c] = upd ? max(f, hv) : F1[c];
        e1 = left ? NEG : max(ne1, hh);
        if (TWO) {
            F2[c] = upd ? max(f2, hv2) : F2[c];
            e2 = left ? NEG : max(ne2, hh2);
        }
        Mp[c] = m;
    }
}

// One wave per strip.  The strip list holds every task's strips in order
// (strip k's predecessor is entry k - 1 when it is not the task's first).
// Strips start in launch order: each wave takes the next position from a
// launch-wide counter in the order the waves start running and runs strip
// order[position] (bits 0-29; bit 30 raises the wave's issue priority).  The
// host puts every strip after its predecessor in that order, so a strip's
// predecessor has always started: every wait ends, whatever number of waves
// the device holds at once.  sync[0] = the counter,
// sync[1] = the error word, sync[2 + k] = rows of strip k whose boundary
// records are complete.  MODE 0: SemiGlobal_no_avx; 1: weighted two-piece;
// 2: weighted single gap (profiles in W).
template <bool TB, int MODE, bool X2Z, bool KEYED>
__global__ __launch_bounds__(64) void k_sg_strip(const int8_t *__restrict__ seq, const PairDesc *__restrict__ pairs,
                                                 const SgTask *__restrict__ tasks, const SgStrip *__restrict__ strips,
                                                 const int32_t *__restrict__ order, int32_t total, Params P,
                                                 int32_t *__restrict__ scr,
                                                 uint8_t *__restrict__ tb, SgOut *__restrict__ out, SgWeights W,
                                                 int32_t *__restrict__ sync) {
    __shared__ int32_t wtab[kMaxProfiles * 8];
    __shared__ int32_t bstage[64 * BND];  // lane 63's boundary records of one block of 64 steps
    __shared__ __align__(16) int32_t cstage[64 * 8];  // lane 0's inputs of one block of 64 steps
    const int lane = threadIdx.x;
    if (MODE != 0) {
        for (int i = lane; i < W.ncat * 8; i += 64) wtab[i] = W.tab[i];
        __syncthreads();
    }
    int32_t k = 0;
    if (lane == 0) k = atomicAdd(sync, 1);
    k = __builtin_amdgcn_readfirstlane(k);
    if (k >= total) return;
    {
        // bit 30: a strip on or near the launch's longest chain of strips
        // issues ahead of the other waves of its SIMD
        const int32_t ok = order[k];
        k = ok & 0x3fffffff;
        if (ok & 0x40000000) __builtin_amdgcn_s_setprio(3);
    }
    int32_t *err = sync + 1, *prog = sync + 2;
    const SgStrip S = strips[k];
    const SgTask T = tasks[S.task];
    const PairDesc pd = pairs[T.pair];
    const int8_t *s1 = seq + pd.off1, *s2 = seq + pd.off2;
    const int32_t R = T.R, C = T.C, w = P.bandwidth;
    const int32_t o1 = P.open, x1 = P.ext, o2 = P.open2, x2 = X2Z ? 0 : P.ext2;
    // traceback with the keyed row maximum: tagged values (see NEGK)
    constexpr bool TAG = TB && KEYED && MODE == 0;
    const int32_t o1f = o1 * 256 + kTagF1, o1e = o1 * 256 + kTagE1, o2f = o2 * 256 + kTagF2, o2e = o2 * 256 + kTagE2,
                  x1s = x1 * 256, x2s = x2 * 256;
    constexpr int32_t NM = TAG ? NEGK : NEG, NE1 = TAG ? NEGK | kTagE1 : NEG, NE2 = TAG ? NEGK | kTagE2 : NEG;
    int32_t *rowres = scr + T.scr;
    int32_t *bout = scr + S.bnd;
    // the lane and column that hold column C-1 (for the stale diagonals)
    const int32_t stale_lane = (C >= 2) ? ((C - 2) % SW) / K : -1, stale_c = (C >= 2) ? (C - 2) % K : -1;
    const int32_t stale_strip = (C >= 2) ? (C - 2) / SW : -1;

    StripIn in;
    in.first = S.s == 0;
    in.plo = 1;
    in.phi = 0;
    in.bin = nullptr;
    if (!in.first) {
        const SgStrip Q = strips[k - 1];
        in.plo = Q.rlo;
        in.phi = Q.rhi;
        in.bin = scr + Q.bnd;
    }
    in.R = R;
    in.a1 = T.a1;
    in.rev = T.rev;
    in.s1 = s1;
    in.wc = MODE != 0 ? W.cat + W.cat_off[T.pair] : nullptr;
    in.tag = TAG;

    const int32_t nsteps = S.rhi - S.rlo + 1 + 63;
    const int32_t jfirst = S.s * SW + 1 + lane * K;
    const int32_t hi_c = C - jfirst;  // the lane's band end in right-edge blocks
    const int32_t jb = S.s * SW + 1;  // the strip's first column
    const int32_t kc = (kKeyIn - 1) - lane * K;  // keyed row maximum: rank of the lane's first column
    uint32_t tlo[K], thi[K];
    int32_t qc[K], F1[K], F2[K], Mp[K];
#pragma unroll
    for (int c = 0; c < K; ++c) {
        const int32_t j = jfirst + c;
        qc[c] = j <= C ? code_at(s2, T.a2, j - 1, T.rev) : 2;
        if (MODE == 0) column_table(qc[c], P.match, P.mismatch, tlo[c], thi[c]);
        F1[c] = TAG ? NEGK | kTagF1 : NEG;
        F2[c] = TAG ? NEGK | kTagF2 : NEG;
        Mp[c] = NM;
    }
    // carried row maximum: (max, column), or keyed (key, column carried in)
    constexpr int32_t kRm0 = KEYED ? kKeyIn : 0;
    int32_t prev_im = NM, om = NM, oe1 = NE1, oe2 = NE2, orm = kRm0, orj = 0,
            orc = MODE == 0 ? (int32_t)row_selector(2) : 2;
    bool ok = true;
    for (int32_t t0 = 0; t0 < nsteps; t0 += 64) {
        if (!in.first) {
            const int32_t need = min(S.rlo + t0 + 63, in.phi);
            if (need >= in.plo && !wait_rows(prog + k - 1, need, err)) {
                ok = false;
                break;
            }
        }
        if (t0 == 0) {  // the diagonal of the strip's first row
            int32_t a, b, c2, d, f;
            in.load(S.rlo - 1, prev_im, a, b, c2, d, f);
            if (lane != 0) prev_im = NM;
        }
        {
            // lane 0's inputs of the block's steps (rows rlo + t0 + q), one
            // 8-int record per step in LDS, read back by every lane at once
            int32_t cm, ce1, ce2, crm, crj, crc;
            in.load(S.rlo + t0 + lane, cm, ce1, ce2, crm, crj, crc);
            // the row's score selector travels instead of its code
            if (MODE == 0) crc = (int32_t)(TAG ? row_selector_tag(crc) : row_selector(crc));
            if (KEYED) {
                // the previous strip's record holds its keyed maximum and the
                // column carried into it: decode, then rank it above every
                // column of this strip
                if (!in.first && crm != 0) crj = key_col(crm, crj, jb - SW);
                crm = ((in.first ? 0 : key_max(crm)) << kKeyShift) | kKeyIn;
                if (in.first) crj = 0;
            }
            int32_t *cs = cstage + lane * 8;
            cs[0] = cm;
            cs[1] = ce1;
            cs[2] = ce2;
            cs[3] = crm;
            cs[4] = crj;
            cs[5] = crc;
        }
        __syncthreads();
        auto load_in = [&](int32_t q, int4 &a, int2 &b) {
            const int32_t *p = cstage + min(q, 63) * 8;
            a = *(const int4 *)p;
            b = *(const int2 *)(p + 4);
        };
        const int32_t t1 = min(t0 + 64, nsteps);
        // One step of the strip.  KIND 0: any block.  KIND 1 (full): every
        // lane holds a valid row whose band covers the lane's K columns and
        // the next one (no edge masks), the row's result lies in a later strip
        // and no stale diagonal is recorded, so the step needs no row, band or
        // lane tests.  KIND 2 (right edge): as full, except that the band of
        // every row ends at column C inside this strip, so each lane's band
        // end hi = C - jfirst is the same on every row and every row's result
        // is written here.
        auto step = [&](auto kind_c, int32_t t, const int4 &a, const int2 &b) {
            constexpr int KIND = decltype(kind_c)::value;
            constexpr bool FULL = KIND != 0;
            const int32_t q = t - t0;
            // lanes 1..63 take the carried values of lane l-1; lane 0 takes row
            // rlo + t's input (a, b: the step's LDS record, which the DPP shift
            // leaves in lane 0 as its old operand)
            const int32_t im = __builtin_amdgcn_update_dpp(a.x, om, 0x138, 0xf, 0xf, false);
            const int32_t ie1 = __builtin_amdgcn_update_dpp(a.y, oe1, 0x138, 0xf, 0xf, false);
            const int32_t ie2 = __builtin_amdgcn_update_dpp(a.z, oe2, 0x138, 0xf, 0xf, false);
            const int32_t irm = __builtin_amdgcn_update_dpp(a.w, orm, 0x138, 0xf, 0xf, false);
            const int32_t irj = __builtin_amdgcn_update_dpp(b.x, orj, 0x138, 0xf, 0xf, false);
            const int32_t irc = __builtin_amdgcn_update_dpp(b.y, orc, 0x138, 0xf, 0xf, false);
            const int32_t r = S.rlo + t - lane;
            const bool valid = FULL || (lane <= t && r <= S.rhi);
            const int32_t diag = prev_im;
            prev_im = im;
            uint32_t tword = 0xffffffffu;
            if (valid) {
                const int32_t st = FULL ? 0 : max(r - w, 1), en = FULL ? 0 : min(r + w, C);
                // in-band columns: lo <= c < hi; c == hi: last
                const int32_t lo = KIND == 2 ? 0 : st - jfirst, hi = KIND == 2 ? hi_c : en - jfirst;
                int32_t e1 = ie1, e2 = ie2, rm = irm, rj = irj;
                // every lane's columns inside the band (the usual case): no edge masks
                const bool interior = KIND == 1 || (KIND == 0 && __ballot(lo > 0 || hi < K) == 0);
                if (TAG) {
                    const uint32_t sel = (uint32_t)irc;
                    if (interior)
                        columns_tag<false, X2Z>(tlo, thi, F1, F2, Mp, sel, lo, hi, kc, o1f, o1e, x1s, o2f, o2e, x2s,
                                                diag, e1, e2, rm, tword);
                    else
                        columns_tag<true, X2Z>(tlo, thi, F1, F2, Mp, sel, lo, hi, kc, o1f, o1e, x1s, o2f, o2e, x2s,
                                               diag, e1, e2, rm, tword);
                } else if (MODE == 0) {
                    const uint32_t sel = (uint32_t)irc;
                    if (interior)
                        columns<TB, false, KEYED>(tlo, thi, F1, F2, Mp, sel, lo, hi, jfirst, kc, o1, x1, o2, x2, diag,
                                                  e1, e2, rm, rj, tword);
                    else
                        columns<TB, true, KEYED>(tlo, thi, F1, F2, Mp, sel, lo, hi, jfirst, kc, o1, x1, o2, x2, diag,
                                                 e1, e2, rm, rj, tword);
                } else {
                    // this row's profile and the next row's (vertical gaps)
                    const int32_t rc = irc & 15;
                    const int32_t *pc = wtab + ((irc >> 4) & 255) * 8, *pv = wtab + (irc >> 12) * 8;
                    const int32_t match = rc == 2 ? 0 : pc[0], mis = rc == 2 ? 0 : pc[1];
                    if (interior)
                        columns_w<TB, false, MODE == 1>(qc, F1, F2, Mp, rc, match, mis, lo, hi, jfirst, pc[2], pc[3],
                                                        pc[4], pc[5], pv[2], pv[3], pv[4], pv[5], diag, e1, e2, rm,
                                                        rj, tword);
                    else
                        columns_w<TB, true, MODE == 1>(qc, F1, F2, Mp, rc, match, mis, lo, hi, jfirst, pc[2], pc[3],
                                                       pc[4], pc[5], pv[2], pv[3], pv[4], pv[5], diag, e1, e2, rm,
                                                       rj, tword);
                }
                om = Mp[K - 1];
                oe1 = e1;
                oe2 = e2;
                orm = rm;
                orj = rj;
                if (!FULL && MODE == 0 && S.s == stale_strip && lane == stale_lane &&
                    (r == C + w - 2 || r == C + w - 1))
                    out[S.task].stale[r - (C + w - 2)] = TAG ? Mp[stale_c] >> 8 : Mp[stale_c];
            } else {
                om = NM;
                oe1 = NE1;
                oe2 = NE2;
                orm = kRm0;
                orj = 0;
            }
            orc = irc;
            if (TB) *(uint32_t *)(tb + T.tb + S.tb + (int64_t)t * SW + lane * K) = tword;
            if (lane == 63 && valid) {
                // the row's boundary record, staged for the block's store
                int32_t *b = bstage + q * BND;
                b[0] = om;
                b[1] = oe1;
                b[2] = oe2;
                b[3] = orm;
                b[4] = orj;
                if (KIND == 2 || (KIND == 0 && (min(r + w, C) - 1) / SW == S.s)) {  // the row's last strip holds its result
                    ast(rowres + 2 * (int64_t)r, KEYED ? key_max(orm) : orm);
                    ast(rowres + 2 * (int64_t)r + 1, KEYED ? key_col(orm, orj, jb) : orj);
                }
            }
        };
        // The block's rows run from rlo + t0 - 63 (lane 63 at t0) to rlo + t0 + 63
        // (lane 0 at t0 + 63).  It is FULL when all its 64 steps hold a row in
        // every lane, the band of its last row starts at or before the strip's
        // first column, the band of its first row reaches past the strip's last
        // column (so lane 63's last column is not the band's last), and no row
        // of the block is one whose stale diagonal is recorded.  Then for every
        // row of the block max(r - w, 1) <= jfirst and min(r + w, C) >=
        // jfirst + K in every lane, which is the interior case, and the row's
        // last strip lies to the right, so no row result is written here.
        // A right-edge block has the same rows and left edge, but column C lies
        // in this strip and the band of the block's first row already reaches
        // it, so every row's band ends at C.
        const int32_t rmin_b = S.rlo + t0 - 63, rmax_b = S.rlo + t0 + 63;
        const bool rows_left = t1 - t0 == 64 && t0 >= 63 && rmax_b <= S.rhi && rmax_b - w <= jb &&
                               !(MODE == 0 && S.s == stale_strip && rmax_b >= C + w - 2);
        // two steps per iteration, each step's LDS record read one step ahead
        auto run = [&](auto kind_c) {
            int4 a;
            int2 b;
            load_in(0, a, b);
            for (int32_t t = t0; t < t1; t += 2) {
                int4 a2;
                int2 b2;
                load_in(t + 1 - t0, a2, b2);
                step(kind_c, t, a, b);
                if (t + 1 < t1) {
                    load_in(t + 2 - t0, a, b);
                    step(kind_c, t + 1, a2, b2);
                }
            }
        };
        if (rows_left && min(rmin_b + w, C) >= jb + SW)
            run(std::integral_constant<int, 1>{});
        else if (rows_left && C < jb + SW && rmin_b + w >= C)
            run(std::integral_constant<int, 2>{});
        else
            run(std::integral_constant<int, 0>{});
        // The block's boundary records leave LDS in one store per field: lane
        // l stores the record lane 63 wrote at step t0 + l (row rlo + t0 + l -
        // 63), where that step held a row.  A single-wave workgroup, so the
        // barriers only order the LDS accesses around the copy.
        __syncthreads();
        {
            const int32_t rl = S.rlo + t0 + lane - 63;
            if (lane < t1 - t0 && t0 + lane >= 63 && rl <= S.rhi) {
                int32_t *b = bout + (int64_t)(rl - S.rlo) * BND;
                const int32_t *src = bstage + lane * BND;
#pragma unroll
                for (int f = 0; f < BND; ++f) ast(b + f, src[f]);
            }
        }
        __syncthreads();
        // boundary rows up to rlo + t1 - 64 are written (the last lane trails by 63 steps)
        publish_rows(prog + k, min(S.rlo + t1 - 64, S.rhi));
    }
    if (!ok) {
        if (lane == 0) ast(err, 1);
        publish_rows(prog + k, INT_MAX);
    }
}

// Ordering key of a row result for the first strict maximum: the larger
// score, then the earlier row.
__device__ __forceinline__ uint64_t best_key(int32_t rm, int32_t row) {
    return ((uint64_t)((uint32_t)rm ^ 0x80000000u) << 32) | (uint64_t)(0xffffffffu - (uint32_t)row);
}

__device__ __forceinline__ uint64_t shfl_up64(uint64_t v, int d) {
    const uint32_t lo = (uint32_t)__shfl_up((int32_t)(uint32_t)v, d), hi = (uint32_t)__shfl_up((int32_t)(v >> 32), d);
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ uint64_t shfl64(uint64_t v, int src) {
    const uint32_t lo = (uint32_t)__shfl((int32_t)(uint32_t)v, src), hi = (uint32_t)__shfl((int32_t)(v >> 32), src);
    return ((uint64_t)hi << 32) | lo;
}

__device__ __forceinline__ int64_t wave_sum64(int64_t v) {
    for (int d = 32; d >= 1; d >>= 1) {
        const uint32_t lo = (uint32_t)__shfl_xor((int32_t)(uint32_t)v, d);
        const uint32_t hi = (uint32_t)__shfl_xor((int32_t)((uint64_t)v >> 32), d);
        v += (int64_t)(((uint64_t)hi << 32) | lo);
    }
    return v;
}

// One wave per task: the reference's row loop termination over the row
// results (z-drop, empty row), 64 rows at a time, the maximum, and the
// traceback (trace tasks) through 64 x 64 tiles of trace bytes loaded by the
// whole wave at once.  tagfmt: the trace bytes are source << 4 | flags (the
// tagged kernel); they are turned into source | flags << 3 as they load.
__global__ __launch_bounds__(64) void k_sg_finish(const int8_t *__restrict__ seq, const PairDesc *__restrict__ pairs,
                                                  const SgTask *__restrict__ tasks,
                                                  const SgStrip *__restrict__ strips, int32_t n, Params P,
                                                  const int32_t *__restrict__ scr, const uint8_t *__restrict__ tb,
                                                  SgOut *__restrict__ out, uint32_t *__restrict__ cigar,
                                                  SgWeights W, int32_t tagfmt) {
    __shared__ uint32_t tile[64 * 16];
    const int32_t k = blockIdx.x;
    if (k >= n) return;
    const int lane = threadIdx.x;
    const SgTask T = tasks[k];
    const PairDesc pd = pairs[T.pair];
    SgOut &o = out[k];
    const int32_t R = T.R, C = T.C, w = P.bandwidth;
    const int32_t *rowres = scr + T.scr;
    const int32_t stale0 = o.stale[0], stale1 = o.stale[1];
    // the best so far: key (score, row; row 0 = none) and its column
    uint64_t ckey = best_key(0, 0);
    int32_t ce2 = -1, brk = R;
    int64_t cells = 0;
    for (int32_t i0 = 1; i0 <= R; i0 += 64) {
        const int32_t i = i0 + lane;
        const bool live = i <= R;
        int32_t rm = 0, rj = 0;
        int64_t cl = 0;
        int32_t zd = P.zdrop, x2 = P.ext2;
        if (live && W.mode != 0) {
            // weighted: every band column is a cell; rows past C + w have an
            // empty band, no positive cell and end the row loop
            const int32_t st = max(i - w, 1), en = min(i + w, C);
            cl = (int64_t)max(0, en - st + 1);
            rj = -1;
            if (i <= C + w) {
                rm = rowres[2 * (int64_t)i];
                rj = rowres[2 * (int64_t)i + 1];
            }
            const int32_t pos = T.rev ? T.a1 - (i - 1) : T.a1 + i - 1;
            const int32_t *pc = W.tab + 8 * (int32_t)W.cat[W.cat_off[T.pair] + pos + 1];
            zd = pc[6];
            x2 = W.mode == 1 ? pc[5] : 0;  // the single-gap z-drop has no l term
        } else if (live) {
            const int32_t st = max(i - w, 1), en = min(i + w, C);
            cl = (int64_t)max(0, en - st) + 1;
            if (i <= C + w) {
                rm = rowres[2 * (int64_t)i];
                rj = rowres[2 * (int64_t)i + 1];
            } else {
                // the band has left the matrix: only column C, whose diagonal is
                // the value the row buffer of this parity last held at C-1
                const int32_t rc = code_at(seq + pd.off1, T.a1, i - 1, T.rev);
                const int32_t qc = code_at(seq + pd.off2, T.a2, C - 1, T.rev);
                const int32_t sc = (rc == 2 || qc == 2) ? 0 : (rc == qc ? P.match : P.mismatch);
                const int32_t dg = (C >= 2) ? (((i - (C + w + 1)) & 1) ? stale1 : stale0) : NEG;
                const int32_t m = sc + dg;
                rm = m > 0 ? m : 0;
                rj = m > 0 ? C : 0;
            }
        }
        // best over rows 1..i (inclusive prefix maximum across the wave)
        uint64_t key = live ? best_key(rm, i) : 0;
        for (int d = 1; d < 64; d <<= 1) {
            const uint64_t v = shfl_up64(key, d);
            if (lane >= d && v > key) key = v;
        }
        if (ckey > key) key = ckey;
        const int32_t best = (int32_t)((uint32_t)(key >> 32) ^ 0x80000000u);
        const int32_t row = (int32_t)(0xffffffffu - (uint32_t)key);
        const int32_t rj_best = __shfl(rj, row >= i0 ? row - i0 : 0);
        const int32_t e1 = row > 0 ? row : -1;
        const int32_t e2 = row >= i0 ? rj_best : (row > 0 ? ce2 : -1);
        int32_t l = (i - e1) - (rj - e2);
        l = l > 0 ? l : -l;
        // zdrop + l * ext2 in 32-bit two's complement, as the reference's int32
        const int32_t lim = (int32_t)((uint32_t)zd + (uint32_t)l * (uint32_t)x2);
        const bool stop = live && (best - rm > lim || rj <= 0);
        const uint64_t hits = __ballot(stop);
        const int last = hits ? __ffsll((unsigned long long)hits) - 1 : 63;
        cells += wave_sum64(lane <= last ? cl : 0);
        ckey = shfl64(key, last);
        ce2 = __shfl(e2, last);
        if (hits) {
            brk = i0 + last;
            break;
        }
    }
    const int32_t brow = (int32_t)(0xffffffffu - (uint32_t)ckey);
    const int32_t e1 = brow > 0 ? brow : -1, e2 = brow > 0 ? ce2 : -1;
    if (lane == 0) {
        o.max = (int32_t)((uint32_t)(ckey >> 32) ^ 0x80000000u);
        o.e1 = e1;
        o.e2 = e2;
        o.brk = brk;
        o.cells = cells;
        o.ncig = 0;
    }
    if (T.tb < 0) return;
    // traceback (impl/sequenceAlignment3.cpp:864-915); runs are written in
    // traceback order and reversed by the host.  Every lane walks the same
    // path; lane 0 writes the runs.
    uint32_t *cg = cigar + o.cig_off;
    int32_t nc = 0, run_op = -1, run_len = 0;
    auto push = [&](int32_t op, int32_t len) {
        if (op == run_op) {
            run_len += len;
            return;
        }
        if (run_op >= 0) {
            if (lane == 0) cg[nc] = ((uint32_t)run_len << 4) | (uint32_t)run_op;
            ++nc;
        }
        run_op = op;
        run_len = len;
    };
    int32_t ii = e1, jj = e2, state = 0;
    while (ii > 0 && jj > 0) {
        // the tile of rows ii-63..ii and the 16 four-column groups ending at jj's
        const int32_t r0 = ii - 63, c0 = ((jj - 1) & ~3) + 1 - 60;
        {
            // 64 columns span at most two strips: those of max(c0, 1) and of jj
            const int32_t sb = (jj - 1) / SW, sa = (max(c0, 1) - 1) / SW;
            const SgStrip A = strips[T.sched + sa], B = strips[T.sched + sb];
            const int32_t r = r0 + lane;
            uint32_t wv[16];
#pragma unroll
            for (int g = 0; g < 16; ++g) {  // 16 independent loads, one round trip
                const int32_t j = c0 + 4 * g;
                const bool inb = j >= 1 && (j - 1) / SW == sb;
                const int32_t rlo = inb ? B.rlo : A.rlo, rhi = inb ? B.rhi : A.rhi;
                const int64_t sto = inb ? B.tb : A.tb;
                const int32_t l = ((j - 1) & (SW - 1)) / K;
                const bool ok = j >= 1 && r >= 1 && r >= rlo && r <= rhi;
                const int64_t off = ok ? T.tb + sto + (int64_t)(r - rlo + l) * SW + l * K : T.tb;
                wv[g] = *(const uint32_t *)(tb + off);
                if (tagfmt) wv[g] = ((wv[g] >> 4) & 0x07070707u) | ((wv[g] & 0x0F0F0F0Fu) << 3);
                wv[g] = ok ? wv[g] : 0xffffffffu;
            }
#pragma unroll
            for (int g = 0; g < 16; ++g) tile[lane * 16 + g] = wv[g];
        }
        __syncthreads();
        // The walk one run at a time.  A cell's step is: in state 0 take the
        // byte's source (tmp & 7); in a gap state keep it while the byte's
        // extension bit (state + 2) is set, else fall to 0 and take the
        // source; then move diagonally (0), up (1, 3) or left.  So from
        // state 0 the walk goes diagonally over the bytes whose source is 0,
        // and from a gap state straight on over the bytes with its extension
        // bit; the first other byte sets the state to its source and takes
        // one move.  Lane k reads the k-th cell of the run's line at once;
        // a ballot finds where the run stops (that byte, the tile's edge or
        // the matrix's).
        while (ii > 0 && jj > 0 && ii >= r0 && jj >= c0) {
            const int32_t st = state;
            const bool vert = st == 1 || st == 3;
            const int32_t dy = (st == 0 || vert) ? 1 : 0, dx = vert ? 0 : 1;
            const int32_t ci = ii - lane * dy, cj = jj - lane * dx;
            const bool in = ci > 0 && cj > 0 && ci >= r0 && cj >= c0;
            int32_t tmp = 0;
            if (in) {
                const uint32_t wd = tile[(ci - r0) * 16 + ((cj - c0) >> 2)];
                tmp = (int32_t)((wd >> (8 * ((cj - c0) & 3))) & 0xffu);
            }
            const bool cont = in && (st == 0 ? (tmp & 7) == 0 : ((tmp >> (st + 2)) & 1) != 0);
            const uint64_t stop = __ballot(!cont);
            const int32_t n = stop ? __ffsll((unsigned long long)stop) - 1 : 64;
            if (n > 0) {
                push(st == 0 ? 0 : (vert ? 2 : 1), n);
                ii -= n * dy;
                jj -= n * dx;
            }
            if (n == 64 || !__builtin_amdgcn_readlane((int32_t)in, n)) continue;  // the tile's or matrix's edge
            state = __builtin_amdgcn_readlane(tmp, n) & 7;
            if (state == 0) {
                push(0, 1);
                --ii;
                --jj;
            } else if (state == 1 || state == 3) {
                push(2, 1);
                --ii;
            } else {
                push(1, 1);
                --jj;
            }
        }
        __syncthreads();
    }
    if (ii > 0) push(2, ii);
    if (jj > 0) push(1, jj);
    push(-2, 0);  // flush the last run
    if (lane == 0) o.ncig = nc;
}

// Strip geometry of one task: the strips the band reaches, with their trace
// store offsets (boundary offsets are set by the caller).
int32_t strip_list(int32_t R, int32_t C, int32_t w, std::vector<SgStrip> *out) {
    int32_t n = 0;
    int64_t tbo = 0;
    for (int32_t s = 0; (int64_t)s * SW < C; ++s) {
        const int32_t jb = s * SW + 1, je = std::min(jb + SW - 1, C);
        const int32_t rlo = std::max(1, jb - w), rhi = std::min(std::min(R, je + w), C + w);
        if (rlo > rhi) break;
        if (out) {
            SgStrip g{};
            g.s = s;
            g.rlo = rlo;
            g.rhi = rhi;
            g.tb = tbo;
            out->push_back(g);
        }
        tbo += (int64_t)(rhi - rlo + 1 + 63) * SW;
        ++n;
    }
    return n;
}

}  // namespace

int64_t sg_trace_bytes(int32_t R, int32_t C, int32_t w) {
    std::vector<SgStrip> v;
    strip_list(R, C, w, &v);
    return v.empty() ? 0 : v.back().tb + (int64_t)(v.back().rhi - v.back().rlo + 1 + 63) * SW;
}

int64_t sg_scratch_ints(int32_t R, int32_t C, int32_t w) {
    std::vector<SgStrip> v;
    strip_list(R, C, w, &v);
    int64_t n = 2 * ((int64_t)R + 2);
    for (const SgStrip &g : v) n += (int64_t)(g.rhi - g.rlo + 1) * BND;
    return n;
}

int32_t sg_task_strips(int32_t R, int32_t C, int32_t w, int32_t task, int64_t bnd, std::vector<SgStrip> &out) {
    const size_t first = out.size();
    const int32_t n = strip_list(R, C, w, &out);
    for (size_t i = first; i < out.size(); ++i) {
        out[i].task = task;
        out[i].bnd = bnd;
        bnd += (int64_t)(out[i].rhi - out[i].rlo + 1) * BND;
    }
    return n;
}

bool sg_keyed_ok(const Params &P, int64_t max_rows, int64_t max_cols) {
    // Every in-band M(i, j) lies between the score of a diagonal of
    // mismatches to (min, min) followed by one gap run (which stays in the
    // band), and match * min(rows, cols); the key holds scores below 2^22 in
    // magnitude (here 2^21, with margin), the tagged values of the traceback
    // kernel scores and penalties times 256.
    const int64_t L = std::max<int64_t>(max_cols, 1);
    const int64_t gap = std::max((int64_t)P.open + (L - 1) * P.ext, (int64_t)P.open2 + (L - 1) * P.ext2);
    const int64_t low = std::min<int64_t>(P.mismatch, 0) * std::max<int64_t>(max_rows, 0) + std::min<int64_t>(gap, 0);
    const int64_t high = (int64_t)std::max(P.match, 0) * std::min(max_rows, max_cols);
    const int64_t lim = (int64_t)1 << 21;
    // the tagged traceback values hold every penalty times 256
    const int64_t pen = (int64_t)1 << 20;
    const bool pen_ok = -(int64_t)P.open < pen && -(int64_t)P.ext < pen && -(int64_t)P.open2 < pen &&
                        -(int64_t)P.ext2 < pen && P.open <= 0 && P.ext <= 0 && P.open2 <= 0 && P.ext2 <= 0;
    return -low < lim && high < lim && pen_ok;
}

hipError_t launch_sg_dp(hipStream_t st, bool trace, bool keyed, const int8_t *seq, const PairDesc *pairs,
                        const SgTask *tasks, const SgStrip *strips, const int32_t *order, int32_t nstrips,
                        const Params &P, int32_t *scr, uint8_t *tb,
                        SgOut *out, const SgWeights &W, int32_t *sync) {
    if (nstrips <= 0) return hipSuccess;
    if (W.mode != 0 && (W.ncat < 1 || W.ncat > kMaxProfiles)) return hipErrorInvalidValue;
#define DCNS_SG_LAUNCH(TBV, MV, X2V, KV)                                                                             \
    hipLaunchKernelGGL((k_sg_strip<TBV, MV, X2V, KV>), dim3(nstrips), dim3(64), 0, st, seq, pairs, tasks, strips,     \
                       order,                                                                                       \
                       nstrips, P,                                                                                  \
                       scr, tb, out, W, sync)
    switch (W.mode * 2 + (trace ? 1 : 0)) {
        // the unweighted kernel with a flat second gap piece (ext2 0, the
        // cut2Gaps default) drops the piece's two extension adds per cell
        // keyed: the row maximum and its column as one int (see key_max),
        // where every in-band score fits 22 bits
        case 0:
            if (P.ext2 == 0) {
                if (keyed)
                    DCNS_SG_LAUNCH(false, 0, true, true);
                else
                    DCNS_SG_LAUNCH(false, 0, true, false);
            } else {
                if (keyed)
                    DCNS_SG_LAUNCH(false, 0, false, true);
                else
                    DCNS_SG_LAUNCH(false, 0, false, false);
            }
            break;
        case 1:
            if (P.ext2 == 0) {
                if (keyed)
                    DCNS_SG_LAUNCH(true, 0, true, true);
                else
                    DCNS_SG_LAUNCH(true, 0, true, false);
            } else {
                if (keyed)
                    DCNS_SG_LAUNCH(true, 0, false, true);
                else
                    DCNS_SG_LAUNCH(true, 0, false, false);
            }
            break;
        case 2: DCNS_SG_LAUNCH(false, 1, false, false); break;
        case 3: DCNS_SG_LAUNCH(true, 1, false, false); break;
        case 4: DCNS_SG_LAUNCH(false, 2, false, false); break;
        case 5: DCNS_SG_LAUNCH(true, 2, false, false); break;
        default: return hipErrorInvalidValue;
    }
#undef DCNS_SG_LAUNCH
    return hipGetLastError();
}

hipError_t launch_sg_finish(hipStream_t st, const int8_t *seq, const PairDesc *pairs, const SgTask *tasks,
                            const SgStrip *strips, int32_t n, const Params &P, const int32_t *scr, const uint8_t *tb,
                            SgOut *out, uint32_t *cigar, const SgWeights &W, bool tagfmt) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_sg_finish, dim3(n), dim3(64), 0, st, seq, pairs, tasks, strips, n, P, scr, tb, out, cigar,
                       W, (int32_t)tagfmt);
    return hipGetLastError();
}

}  // namespace dcns
