// kernels.h -- host-side launchers of the gfx950 kernels (internal API).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "dcns_types.h"
#include "dp_common.h"

namespace dcns {

// seedscan.hip
// The level-0 scan runs in packed f16 when every DP value is an exact f16 integer.
inline bool scan_uses_f16(const Params &P) { return (int64_t)P.window * P.match <= 2048; }
// The later levels' scan runs two tasks per lane in packed f16 when, besides,
// the f16 patterns of match and mismatch have a zero low byte: a reference
// base's scores against A, C, G, T are then the four high bytes of one word,
// mis4 (mismatch in every byte) with diff (match ^ mismatch) xored into the
// base's own byte.
inline bool scan_general_packed(const Params &P, uint32_t &mis4, uint32_t &diff) {
    if (!scan_uses_f16(P)) return false;
    const uint32_t m = f16_bits_of_int(P.match), x = f16_bits_of_int(P.mismatch);
    if ((m & 0xffu) || (x & 0xffu)) return false;
    mis4 = (x >> 8) * 0x01010101u;
    diff = (m ^ x) >> 8;
    return true;
}
// Level-0 tails of the rescans.  A rescan of rows [a, b) of a window starts
// from zeros at row a.  The scan kernels evaluate wm columns (the window,
// padded with N columns, which score 0).  A path of cells that starts above
// row a and ends on row r enters at most wm of its rows diagonally, gains at
// most match on each of at most `window` real columns, and pays at least
// g = -max(open, ext) for each row it enters by a vertical gap step, so its
// score is at most window * match - g * (r - a + 1 - wm).  From row
// a + level0_tail_rows(P, wm) on that is <= 0, so no such path reaches a
// cell (every cell and gap state is floored at 0): the rescan's values
// there equal level 0's.  Such a rescan runs only its first
// level0_tail_rows rows and takes the rest of its trajectory from level 0's
// strong rows (rows whose window maximum reaches min_score).  0: the bound
// does not hold (g <= 0).
inline int32_t level0_tail_rows(const Params &P, int32_t wm) {
    const int32_t g = -std::max(P.open, P.ext);
    if (g <= 0 || P.match <= 0) return 0;
    return wm + (P.window * P.match + g - 1) / g;
}
// The column of a strong-row entry is not searched in level 0's lane
// kernel: kStrongNoCol there, found again by the tail from a short rescan.
constexpr int32_t kStrongNoCol = 0xff;
// Level 0 appends every strong row of every window to the sink: chunks of
// kStrongChunk entries handed out one at a time from *nchunks to each lane;
// an entry is the global window id and row << 32 | maximum << 8 | first
// column of the maximum.  cap_chunks 0: no recording.  When more chunks are
// asked for than the buffers hold, the later entries are dropped and
// *nchunks tells the host so.
constexpr int32_t kStrongChunk = 64;
struct StrongSink {
    uint32_t *win;
    uint64_t *rv;
    int32_t *chunk_fill;  // entries used per chunk
    int32_t *nchunks;
    int32_t cap_chunks;
};
// Level 0 in row segments (a batch whose longest references would hold the
// launch): task (pair, first window, r0, r1) scans rows [r0, r1) of its 128
// windows after `warm` rows from zeros (level0_tail_rows: rows from r0 on are
// then level 0's values), into 128 trajectory slots of its own; a group
// (pair, first window, first task, segments) lists its consecutive tasks in
// row order, and k_l0_merge turns them into the windows' trajectories.
struct SegL0 {
    const int4 *tasks = nullptr;
    const int4 *groups = nullptr;
    TrajEntry *traj = nullptr;
    int32_t *tcount = nullptr;
    int32_t warm = 0;
};
hipError_t launch_scan_uniform_seg(hipStream_t st, int wm, const int8_t *seq, const PairDesc *pairs, const SegL0 &sg,
                                   int32_t ntasks, const Params &P, int32_t cap, Counters *ctr, int32_t *lastrow,
                                   const StrongSink &sink);
hipError_t launch_l0_merge(hipStream_t st, const PairDesc *pairs, const SegL0 &sg, int32_t ngroups, TrajEntry *traj,
                           int32_t *tcount, int32_t cap, int32_t min_score, int32_t *hitlist, Counters *ctr);
// The strong rows sorted by window (row order within a window): seg[w] =
// [begin, end) of global window w, ng[i] = the next index of the same window
// whose maximum is strictly larger (end of the window's range when none).
struct StrongIndex {
    const int2 *seg;
    const uint64_t *rv;
    const int32_t *ng;
    int32_t delta;  // level0_tail_rows
};
// Builds the index from the sink's first nchunks chunks: scratch_bytes of
// scratch (sized by strong_index_scratch) at scratch; key2, rv2 and key hold
// n entries (the total the host read back), seg nw windows.
size_t strong_index_scratch(int64_t nchunks, int64_t n);
hipError_t launch_strong_index(hipStream_t st, const StrongSink &S, int32_t nchunks, int64_t n, int32_t nw,
                               int32_t win_bits, void *scratch, size_t scratch_bytes, uint64_t *key2, uint64_t *rv2,
                               uint64_t *key, int64_t *chunk_off, int2 *seg, int32_t *ng);
// Entries of the chunks: the exclusive scan of chunk_fill (widened into
// fill64) into chunk_off and the total in *total_dev (one int64).
hipError_t launch_strong_count(hipStream_t st, const StrongSink &S, int32_t nchunks, void *scratch,
                               size_t scratch_bytes, int64_t *fill64, int64_t *chunk_off, int64_t *total_dev);
// After a level's rescans (of at most index.delta rows each): every task
// whose rows reach past a + delta appends level 0's strong rows of
// [a + delta, min(node b, lastrow + 1)) that raise its running maximum.
hipError_t launch_rescan_tail(hipStream_t st, int wm, const int8_t *seq, const PairDesc *pairs, const Params &P,
                              const ScanTask *tasks, const int32_t *n_dev, int32_t n_max, int32_t blocks_max,
                              const Node *nodes, const int32_t *lastrow, const StrongIndex &I, TrajEntry *traj,
                              int32_t *tcount, int32_t cap, Counters *ctr);
// Level-0 window scan: one workgroup per 128-window task, or, with
// persistent_waves > 0, that many waves popping the task list from ctr->queue.
// wins[0, nwins): level-0 windows (pair, window) scanned one per wave on the
// anti-diagonal path, in the launch's first workgroups.
hipError_t launch_scan_uniform(hipStream_t st, int wm, const int8_t *seq, const PairDesc *pairs, const int2 *wtasks,
                               int32_t nwt, int32_t persistent_waves, const Params &P, const ScoreTab &tab,
                               TrajEntry *traj, int32_t *tcount, int32_t cap, int32_t *hitlist, Counters *ctr,
                               const int2 *wins, int32_t nwins, int32_t *lastrow, const StrongSink &sink);
// Later-level rescans: *ntasks_dev tasks (at most ntasks_max) in order[];
// up to wave_tasks of the first ones with more than wave_rows rows run one
// per wave, the rest one per lane on at most lane_blocks_max waves (0: one
// lane per task).
hipError_t launch_scan_general(hipStream_t st, int wm, const int8_t *seq, const PairDesc *pairs,
                               const ScanTask *tasks, const int32_t *order, const int32_t *ntasks_dev,
                               int32_t ntasks_max, int32_t wave_tasks, int32_t wave_rows, int32_t lane_blocks_max,
                               const Params &P, const ScoreTab &tab, TrajEntry *traj, int32_t *tcount, int32_t cap,
                               Counters *ctr);
hipError_t launch_resolve(hipStream_t st, int wm, const int8_t *seq, const PairDesc *pairs,
                          const int32_t *pair_of_window, const Params &P, const ScoreTab &tab, int level0,
                          const int32_t *hitlist, const ScanTask *tasks, const int32_t *n_dev, int32_t n_max,
                          int32_t blocks_max, Node *nodes, int32_t node_cap, const TrajEntry *traj,
                          const int32_t *tcount, int32_t cap, int32_t scan_cap, ScanTask *next, int32_t next_cap,
                          Counters *ctr, const int32_t *lastrow, int32_t tail_rows);
// Counter block at the start of a seed phase (nodes = scans = nw, the rest 0).
hipError_t launch_seed_init(hipStream_t st, Counters *ctr, int32_t nw);
// Start of a recursion level: level_tasks = next_tasks (0 after an
// overflow), next_tasks = 0, the order histogram (2 * kOrderBuckets ints) cleared.
constexpr int kOrderBuckets = 512;
hipError_t launch_level_begin(hipStream_t st, Counters *ctr, int32_t *hist);
// order[0, *n_dev): the level's tasks by row count, longest first (16 buckets
// per doubling); hist: 2 * kOrderBuckets ints cleared by launch_level_begin.
hipError_t launch_order_tasks(hipStream_t st, const ScanTask *tasks, const int32_t *n_dev, int32_t n_max,
                              int32_t *hist, int32_t *order);
hipError_t launch_fill_i32(hipStream_t st, int32_t *p, int64_t n, int32_t v);
hipError_t launch_gather_roots(hipStream_t st, const Node *nodes, const int32_t *hitlist, const int32_t *n_dev,
                               int32_t n_max, Node *out);

// xdrop.hip
// Bands of up to kXdropLaneBand cells run one task per lane, up to
// kXdropLane2Band one task per lane of the wider lane kernel, up to
// kXdropWaveBand one task per wave, wider ones with the band in scratch.
constexpr int kXdropLaneBand = 21;
constexpr int kXdropLane2Band = 41;
constexpr int kXdropWaveBand = 192;
// Rows a lane kernel's pass may evaluate before its task moves to the wave
// kernel (dcns_tuning.xdrop_lane_rows overrides it).
constexpr int kXdropLaneRows = 4096;
// Pass 1: forward score-only extension + reverse score-only extension, one
// lane per task of in_list[0, n) (all tasks when null) at bands of up to
// `band` cells (kXdropLaneBand or kXdropLane2Band); tasks whose band
// outgrows it, or whose pass runs past max_rows rows, are counted in
// *pending and listed in list[0, *pending).
hipError_t launch_xdrop_pass1(hipStream_t st, int band, const int8_t *seq, const PairDesc *pairs, const XTask *tasks,
                              const int32_t *in_list, int32_t n, const Params &P, XOut *out, int32_t *pending,
                              int32_t *list, int32_t max_rows);
// Pass 1 one wave per task for list[0, n) (every task when list is null);
// tasks too wide for the wave are counted in *pending.
hipError_t launch_xdrop_wave_pass1(hipStream_t st, const int8_t *seq, const PairDesc *pairs, const XTask *tasks,
                                   const int32_t *list, int32_t n, const Params &P, XOut *out, int32_t *pending);
// Pass 2: reverse extension with traceback into a band-compressed store, then
// the cigar; the lane kernel of `band` cells takes order[0, n)'s tasks, the
// wave kernel one task of the wave band per order entry.
hipError_t launch_xdrop_pass2(hipStream_t st, int band, const int8_t *seq, const PairDesc *pairs, const XTask *tasks,
                              const int32_t *order, int32_t n, const Params &P, XOut *out, const int64_t *tb_off,
                              uint8_t *tb, uint32_t *cigar);
hipError_t launch_xdrop_wave_pass2(hipStream_t st, const int8_t *seq, const PairDesc *pairs, const XTask *tasks,
                                   const int32_t *order, int32_t n, const Params &P, XOut *out,
                                   const int64_t *tb_off, uint8_t *tb, uint32_t *cigar);

// Packs the cigar runs of out[0, n) (each task's arena holds o.ncig of them
// at o.cig_off) into packed[], task after task, and points cig_off there.
size_t pack_cigars_scratch_bytes(int32_t n);
hipError_t launch_pack_cigars(hipStream_t st, XOut *out, int32_t n, void *scratch, size_t scratch_bytes,
                              const uint32_t *cig, uint32_t *packed);

// Wide bands (2w+1 > kXdropWaveBand): band rows in per-task global scratch (scr_off[t] < 0: task not handled).
hipError_t launch_xdrop_wide_pass1(hipStream_t st, const int8_t *seq, const PairDesc *pairs, const XTask *tasks,
                                   int32_t n, const Params &P, XOut *out, int32_t *scr, const int64_t *scr_off,
                                   const int32_t *wmax);
hipError_t launch_xdrop_wide_pass2(hipStream_t st, const int8_t *seq, const PairDesc *pairs, const XTask *tasks,
                                   int32_t n, const Params &P, XOut *out, const int64_t *tb_off, uint8_t *tb,
                                   uint32_t *cigar, int32_t *scr, const int64_t *scr_off);

// swbatch.hip: SmithWaterman max score and first position for arbitrary pairs.
// Lane plan of a K1 launch sequence over npoints points of per_point pairs
// (pair = point * per_point + r): the waves of wslots slots (64, or 128 for
// the kernel with two pairs per lane) share their loop shape and run longest
// first.  slot_order: order[slot] is the pair; else order lists the points
// (pairs of a point adjacent; per_point >= wslots, so a wave spans at most
// two points).  waves[w]: boundary rows per lane of wave w.
struct SwPlan {
    std::vector<int32_t> order;
    std::vector<int32_t> waves;
    bool slot_order = false;
    int32_t per_point = 1;
    int32_t wslots = 64;
    int64_t total = 0;
};
void sw_plan(const int32_t *len1, const int32_t *len2, int64_t npoints, int32_t per_point, SwPlan &plan,
             int32_t wslots = 64, const std::vector<uint8_t> *skip = nullptr);
// Plan over listed pairs, one per slot, waves of 64 slots of one loop shape.
void sw_plan_pairs(const int32_t *len1, const int32_t *len2, int32_t per_point, const std::vector<int32_t> &pairs,
                   SwPlan &plan);
// The waves [w0, return) that fit `budget` int32 of boundary rows (at least
// one); wave_off gets their offsets relative to the first (one more entry).
int64_t sw_plan_chunk(const SwPlan &plan, int64_t w0, int64_t budget, std::vector<int64_t> &wave_off);
// order[0, n) are the pairs of the launch's slots; wave_off as sw_plan_chunk;
// nn: N against N scores -1 (SmithWaterman_bas's matrix) instead of 0.
hipError_t launch_sw_batch(hipStream_t st, const int8_t *seq, const int64_t *off1, const int32_t *len1,
                           const int64_t *off2, const int32_t *len2, const int32_t *order, int32_t n,
                           const Params &P, const ScoreTab &tab, int32_t *out, int32_t *bound,
                           const int64_t *wave_off, bool nn = false);
// SmithWaterman_bas as written, one pair per thread with its full matrices in
// ws (3 x (L1+1) x (L2+1) int32 at ws_off[t]), and with reverse the
// reference's traceback walk (cigar at cig_off[t] .. cig_off[t+1], ncig,
// status 0 / 1 / 2 = ended / stopped at a zero cell / never ends).
hipError_t launch_sw_bas_full(hipStream_t st, const int8_t *seq, const int64_t *off1, const int32_t *len1,
                              const int64_t *off2, const int32_t *len2, int32_t n, int32_t open, int32_t ext,
                              bool reverse, int32_t *ws, const int64_t *ws_off, uint32_t *out3, uint32_t *cig,
                              const int64_t *cig_off, int32_t *ncig, int32_t *status);
// Max scores of random fragment pairs generated on the device: slots
// [first, first + n) of the plan (order, slot_order) with the launch's wave_off.
hipError_t launch_sw_random(hipStream_t st, const uint32_t *thr, const int32_t *len1, const int32_t *len2,
                            const uint64_t *seed, const int32_t *order, bool slot_order, int32_t per_point,
                            int64_t first, int64_t n, const Params &P, const ScoreTab &tab, int32_t *out,
                            int32_t *bound, const int64_t *wave_off);
// The same with two pairs per lane in packed f16 (the plan made with wslots
// 128), for scorings sw_random_packed accepts: match and mismatch whose f16
// patterns have a zero low byte, gaps within f16's exact integers.  A pair
// whose maximum reaches 2048 gets -1 in out and is appended to ovf (count in
// *novf, zeroed by the caller) for the int32 kernel.
bool sw_random_packed(const Params &P, uint32_t &mis4, uint32_t &diff);
// One pair per wave, stripe j of the query in lane j (at most 64 stripes,
// len2 <= 2048): the pairs list[0, n), int32, maxima only.
hipError_t launch_sw_random_wave(hipStream_t st, const uint32_t *thr, const int32_t *len1, const int32_t *len2,
                                 const uint64_t *seed, const int32_t *list, int32_t per_point, int64_t n,
                                 const Params &P, const ScoreTab &tab, int32_t *out);
// The points of dcns_sw_random_max whose pairs take the wave kernel: the
// longest (rows x stripes) for as long as that shortens the launch in an
// issue-slot model of the two kernels.
std::vector<uint8_t> sw_wave_points(const int32_t *len1, const int32_t *len2, int64_t npoints, int32_t per_point,
                                    bool packed, int32_t cus);
// wlist[0, nwave): pairs scored one per wave (int32, as launch_sw_random_wave)
// in the same launch's first blocks.
hipError_t launch_sw_random_h(hipStream_t st, const uint32_t *thr, const int32_t *len1, const int32_t *len2,
                              const uint64_t *seed, const int32_t *order, bool slot_order, int32_t per_point,
                              int64_t first, int64_t n, const Params &P, const ScoreTab &tab, int32_t *out,
                              int32_t *bound, const int64_t *wave_off, int32_t *ovf, int32_t *novf,
                              const int32_t *wlist = nullptr, int32_t nwave = 0);
// Allocation-free K1 (dcns_sw_batch_noalloc): boundary rows in a static device
// array.  sw_noalloc_init takes its address on the current device (once per
// device); sw_noalloc_run launches over pairs [0, n) in order, in as many
// launches as the array needs, and returns after the stream has finished,
// with *err_host bit 0 set when a pair had more than max_len1 rows.
hipError_t sw_noalloc_init(int device);
size_t sw_noalloc_scratch_bytes();
int64_t sw_noalloc_max_rows();
hipError_t sw_noalloc_run(int device, hipStream_t st, const int8_t *seq, const int64_t *off1, const int32_t *len1,
                          const int64_t *off2, const int32_t *len2, int32_t n, int32_t max_len1, const Params &P,
                          const ScoreTab &tab, int32_t *out3, int32_t *err_host);
// Max scores of random fragment pairs generated on the device (pairs [first, first + n)).


// semiglobal.hip: two-piece semi-global DP, one wave per 256-column strip.
int64_t sg_trace_bytes(int32_t R, int32_t C, int32_t w);   // traceback store of one task
int64_t sg_scratch_ints(int32_t R, int32_t C, int32_t w);  // row results + strip boundaries of one task
// Appends the strips of one task (launch-local index `task`, boundary records
// from int32 offset bnd on); returns their number.
int32_t sg_task_strips(int32_t R, int32_t C, int32_t w, int32_t task, int64_t bnd, std::vector<SgStrip> &out);
// sync: 2 + nstrips zeroed ints (start counter, error word, per-strip progress).
// W.mode 0: SemiGlobal_no_avx with P's scores; 1 / 2: weighted two-piece /
// single gap with W's profiles (P's scores unused).
// keyed: every in-band score of the launch's tasks fits 22 bits and every gap
// penalty 20 bits (sg_keyed_ok), so the unweighted kernel may carry the row
// maximum and its column as one int, and its traceback form uses tagged
// values, whose trace bytes the finish kernel reads with tagfmt set.
hipError_t launch_sg_dp(hipStream_t st, bool trace, bool keyed, const int8_t *seq, const PairDesc *pairs,
                        const SgTask *tasks, const SgStrip *strips, const int32_t *order, int32_t nstrips,
                        const Params &P, int32_t *scr, uint8_t *tb, SgOut *out, const SgWeights &W, int32_t *sync);
bool sg_keyed_ok(const Params &P, int64_t max_rows, int64_t max_cols);
hipError_t launch_sg_finish(hipStream_t st, const int8_t *seq, const PairDesc *pairs, const SgTask *tasks,
                            const SgStrip *strips, int32_t n, const Params &P, const int32_t *scr, const uint8_t *tb,
                            SgOut *out, uint32_t *cigar, const SgWeights &W, bool tagfmt);

// mask.hip: maskGenome's k-mer scan, CDS ranges and FASTA bodies.
int64_t kmer_general_slots(int64_t n);      // slots of the general members' table
int64_t kmer_mask_padded(int64_t n);        // bytes to allocate for a mask of n positions
int64_t kmer_body_padded(int64_t n);        // bytes to allocate for n bytes of FASTA bodies
hipError_t launch_kmer_general(hipStream_t st, const uint64_t *hashes, int64_t n, uint32_t *gtab, int64_t slots);
// mask[j] = 1 when a window [i, i + k) with i <= j < i + k, inside one sequence, is in the set.
// slots / overflow: the minimizer buckets of the code members (kmer_hash.h), buckets a power of two.
hipError_t launch_kmer_scan(hipStream_t st, const uint8_t *seq, int64_t N, const int64_t *seg_off, int32_t nseg,
                            int32_t k, const uint64_t *slots, const uint64_t *overflow, int64_t buckets,
                            const KmerGeneral &g, uint8_t *mask);
// mask[x] = 1 for x in [ranges[2r], ranges[2r + 1]).
hipError_t launch_paint_ranges(hipStream_t st, const int64_t *ranges, int64_t n, uint8_t *mask);
// FASTA bodies: sequence s occupies out[body_off[s], body_off[s + 1]).
hipError_t launch_mask_emit(hipStream_t st, const uint8_t *seq, const uint8_t *mask, const int64_t *seg_off,
                            const int64_t *body_off, int32_t nseg, int64_t N, int64_t total, int32_t width,
                            bool any_empty, uint8_t alt, bool soft, uint8_t *out);

}  // namespace dcns
