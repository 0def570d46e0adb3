// dcns_types.h -- plain data shared by host code and gfx950 kernels.
//
// Sequences live in one device buffer of int8 codes (A0 C1 N2 G3 T4, the
// alphabet of impl/SequenceCharToUInt8.cpp:10-30); every pair addresses its
// reference (seq1) and query (seq2) by offset.  Reverse complements are never
// materialised: a kernel reads s[L-1-k] and maps c -> 4-c
// (SequenceCharToUInt8.cpp:62-68).
#pragma once
#include <cstdint>

namespace dcns {

constexpr int32_t kOutOfBand = -536870912;  // sequenceAlignment3.cpp:34
constexpr int kMaxWindow = 64;             // widest seed window the scan kernels hold in registers

// Scoring and seeding parameters of cut1Gap (controlLayer4.cpp:28-52).
struct Params {
    int32_t match, mismatch, open, ext;   // Scorei(match, mismatch); gap open/extend (<= 0)
    int32_t open2, ext2, zdrop, bandwidth; // two-piece piece 2, z-drop, band (SemiGlobal)
    int32_t window, step, min_score;       // seed window (38), step (8), minimum seed score (40)
    int32_t w, xdrop;                      // x-drop band half-width (10) and drop (20)
    int32_t min_ref;                       // miniReferenceSize = min_score / match
};

struct PairDesc {
    int64_t off1, off2;   // offsets of seq1 / seq2 in the sequence buffer
    int32_t len1, len2;
    int32_t nwin;         // windows in the schedule of overload 1
    int32_t win_base;     // global index of the pair's first window
};

// One forward local-alignment scan: rows [a, b) of seq1 against the window
// that starts at query position sp = win * step.
struct ScanTask {
    int32_t pair, win, a, b;
    int32_t node;   // node resolved once the scan's trajectory is known
    int32_t scan;   // trajectory slot
};

// A node of one window's getSeeds recursion tree
// (findSimilarFragmentsForPairedSequence4.cpp:61-128).  Seed coordinates
// are absolute and 0-based; e1 < 0 marks "no seed in [a, b)".
struct Node {
    int32_t a, b, scan;
    int32_t s1, e1, s2, e2;
    int32_t left, right;
};

// Trajectory entry: the scan's running maximum rose strictly above every
// earlier cell (and above min_score - 1) at (row, col) with this score.
struct TrajEntry {
    int32_t row;
    int32_t score_col;  // score << 8 | col
};

// Device counters used by the level loop and the x-drop passes.
struct Counters {
    int32_t nodes;        // allocated nodes
    int32_t scans;        // allocated trajectory slots
    int32_t next_tasks;   // scan tasks queued for the next level
    int32_t hits;         // level-0 windows with at least one trajectory entry
    int32_t overflow;     // set when a pool capacity was exceeded
    int32_t entries;      // trajectory entries recorded by the last scan launches
    int32_t level_tasks;  // scan tasks of the current level (read by the level's kernels)
    int32_t pending;      // x-drop tasks that need a wider band kernel
    unsigned long long gen_cells;  // DP cells evaluated by the later-level scans
    int32_t queue;        // next level-0 task of the persistent scan's queue
    int32_t levels;       // recursion levels after level 0 that had tasks
    int32_t strong_chunks;  // level-0 strong-row chunks handed out
};

// x-drop extension task for one seed (x_extend_seed,
// findSimilarFragmentsForPairedSequence4.cpp:245-285); start1/start2 are the
// seed start + 1 as passed at :492-495.
struct XTask {
    int32_t pair;
    int32_t start1, start2;
    int32_t hint;  // the seed's length (informational)
};

struct XOut {
    int32_t fmax, fe1, fe2;       // forward pass (score only)
    int32_t rmax, re1, re2;       // reverse pass
    int32_t fw, rw;               // band half-width of the final pass of each direction
    int32_t rrows;                // rows evaluated by the final reverse pass
    int32_t iii, jjj;             // traceback leftovers
    int32_t ncig;                 // cigar length
    int32_t status;               // 0 done, 1 forward / 2 reverse pass pending for a wider band kernel
    int32_t lane_stop;            // a lane kernel stopped the pending pass at its row budget: the wave kernel takes it
    int64_t cig_off;              // offset into the cigar arena
    int64_t cells;                // DP cells evaluated (all passes)
};

// One two-piece semi-global DP (SemiGlobal): rows = seq1 from a1, columns =
// seq2 from a2, forward, or on the reverse complements of the prefixes that
// end at a1 / a2 (rev = 1).
struct SgTask {
    int32_t pair, rev, a1, a2;
    int32_t R, C;     // rows (seq1 length) and columns (seq2 length), both >= 1
    int64_t tb;       // byte offset of the traceback store, -1: score only
    int64_t scr;      // int32 offset of the row results (2 ints per row)
    int64_t sched;    // first SgStrip of the task in the launch's strip list
    int32_t nstrips;  // strips the band reaches
    int32_t pad;
};

// One 256-column strip of a SemiGlobal task.  A launch lists the strips of
// all its tasks, each task's strips consecutive and in column order.
struct SgStrip {
    int32_t task;      // task index within the launch
    int32_t s;         // strip index within the task (columns s*256+1 ..)
    int32_t rlo, rhi;  // rows the band reaches in this strip
    int64_t tb;        // byte offset of the strip's trace store inside the task's
    int64_t bnd;       // int32 offset of the strip's right-boundary records (rows rlo..rhi)
};

// Per-row score profiles of the weighted (ZDP) semi-global DP
// (sequenceAlignment3.cpp:1483-1890).  cat[cat_off[pair] + x + 1] is the
// profile index of reference position x (x = -1 .. len1); tab holds 8 ints per
// profile: match, mismatch, open1, ext1, open2, ext2, zdrop, 0.
struct SgWeights {
    const uint8_t *cat;
    const int64_t *cat_off;
    const int32_t *tab;
    int32_t ncat;
    int32_t mode;  // 0: unweighted SemiGlobal_no_avx, 1: weighted two-piece, 2: weighted single gap
};
constexpr int kMaxProfiles = 64;

struct SgOut {
    int32_t max, e1, e2;   // maximum and its 1-based row/column (-1: none)
    int32_t brk;           // last row evaluated (the termination row, or R)
    int64_t cells;         // cells of rows 1..brk, as the reference evaluates them
    int32_t ncig, stale_set;
    int64_t cig_off;       // cigar arena offset (runs in traceback order)
    int32_t stale[2];      // M(C+w-2, C-1), M(C+w-1, C-1): the diagonals rows past C+w read
};

}  // namespace dcns

namespace dcns {

// Members of a maskGenome k-mer set that are not 2-bit codes: any member
// holding a byte other than upper-case A/C/G/T, or every member when k >= 32.
// Looked up by the polynomial hash sum(b[j] * base^(k-1-j)) mod 2^64 of the
// window's bytes, each candidate verified byte by byte.
struct KmerGeneral {
    const uint32_t *table;   // member index + 1 per slot, 0 = empty
    const uint64_t *hashes;  // hash of each member
    const uint8_t *bytes;    // the members, k bytes each
    int64_t n;               // members (0: no such member)
    uint64_t mask;           // slots - 1
    uint64_t pk1;            // base^(k-1) mod 2^64
};
constexpr uint64_t kKmerHashBase = 0x100000001b3ull;

}  // namespace dcns
