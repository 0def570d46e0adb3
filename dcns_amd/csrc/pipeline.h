// pipeline.h -- host orchestration of the device kernels (internal API).
#pragma once
#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <memory>
#include <new>
#include <type_traits>
#include <set>
#include <vector>

#include "../../include/dcns.h"
#include "../../include/dcns/model.h"
#include "ctx.h"
#include "dcns_types.h"

namespace dcns {

void validate_params(const dcns_params &p);
Params to_params(const dcns_params &p);

// Sequences of a batch staged on the device.
struct Staged {
    std::vector<PairDesc> pd;
    int64_t nwin = 0;  // windows over all pairs
    int32_t window = 0, step = 0;  // schedule the window layout was made for
    std::vector<int2> wt;                 // level-0 wave tasks (pair, first window)
    std::vector<int2> ww;                 // level-0 windows (pair, window) scanned one per wave
    std::vector<int32_t> pair_of_window;  // global window -> pair
    std::vector<int64_t> level0_cells;    // per pair: L1 * sum of window widths
};
Staged stage_batch(dcns_ctx *c, const dcns_batch &b, const Params &P);
// Recomputes the window layout of staged pairs for another window/step.
void relayout(dcns_ctx *c, Staged &S, const Params &P);
void upload_layout(dcns_ctx *c, Staged &S);

// Seed sets of every pair (set order), before and after link_seeds.
struct SeedResult {
    std::vector<std::vector<Seed>> linked;  // per pair, set order after link_seeds
    std::vector<int64_t> before_link;        // seeds.size() before link_seeds (the "seeds size" line)
    std::vector<std::vector<Seed>> unlinked; // filled only when asked
    int64_t scan_cells = 0, gpu_cells = 0;
    int32_t levels = 0;
};
void seed_sets(dcns_ctx *c, const Staged &S, const Params &P, bool keep_unlinked, SeedResult &out);

// Tree nodes of one seed set, carved from blocks that grow geometrically
// (64 KiB up to 8 MiB) and are freed together with the set: a set of a
// million seeds allocates ~20 blocks instead of a million nodes.  Only the
// allocator differs from std::set<Seed>; the tree, its balancing and every
// comparison are the same, so the contents and order are too.
template <class T>
struct NodeArena {
    using value_type = T;
    // a set takes its blocks along when swapped or moved
    using propagate_on_container_swap = std::true_type;
    using propagate_on_container_move_assignment = std::true_type;
    struct Pool {
        std::vector<char *> blocks;
        size_t used = 0, cap = 0;
        ~Pool() {
            for (char *b : blocks) std::free(b);
        }
    };
    std::shared_ptr<Pool> pool;
    NodeArena() : pool(std::make_shared<Pool>()) {}
    template <class U>
    NodeArena(const NodeArena<U> &o) : pool(std::reinterpret_pointer_cast<Pool>(o.pool)) {}
    T *allocate(size_t n) {
        const size_t bytes = (n * sizeof(T) + alignof(std::max_align_t) - 1) & ~(alignof(std::max_align_t) - 1);
        if (pool->blocks.empty() || pool->used + bytes > pool->cap) {
            const size_t next = std::min<size_t>(std::max<size_t>(pool->cap * 2, 64 << 10), 8 << 20);
            const size_t cap = std::max(next, bytes);
            char *b = (char *)std::malloc(cap);
            if (!b) throw std::bad_alloc();
            pool->blocks.push_back(b);
            pool->used = 0;
            pool->cap = cap;
        }
        T *p = (T *)(pool->blocks.back() + pool->used);
        pool->used += bytes;
        return p;
    }
    void deallocate(T *, size_t) {}
    template <class U>
    bool operator==(const NodeArena<U> &o) const {
        return (const void *)pool.get() == (const void *)o.pool.get();
    }
    template <class U>
    bool operator!=(const NodeArena<U> &o) const {
        return !(*this == o);
    }
};
using SeedSet = std::set<Seed, std::less<Seed>, NodeArena<Seed>>;

// link_seeds, findSimilarFragmentsForPairedSequence4.cpp:288-344.
void link_seeds(SeedSet &seeds);
// link_seeds before its final insertion: the list in insertion order, from
// the set's contents in set order.
std::vector<Seed> link_seed_list(std::vector<Seed> cur);
// Inserts a list into a set one seed at a time, in list order, with the
// same result as plain inserts; fast when the list is in start1 order.
template <class Set>
void insert_in_start1_order(Set &seeds, const std::vector<Seed> &list);

// x_extend_seed with traceback for a list of tasks (seed start + 1 already applied).
struct XResult {
    int32_t score, start1, end1, start2, end2;
    bool ok;  // false: the reference's p = 100 case
    // the cigar, a view into the context's host cigar buffer (valid until the
    // context's next x-drop call)
    const uint32_t *cig;
    int32_t ncig;
};
// score_only: the variant without traceback (:205-241): untrimmed coordinates, no cigar.
void xdrop_extend(dcns_ctx *c, const Staged &S, const Params &P, const std::vector<XTask> &tasks,
                  std::vector<XResult> &out, int64_t &cells, bool score_only = false);

// SemiGlobal (impl/sequenceAlignment3.cpp:932 -> :681) for many tasks on staged
// pairs.  Task fields used: pair, rev, a1, a2, R, C (R or C <= 0: no DP, the
// empty result).  end1/end2 are 0-based (-2 when no cell is positive), as the
// reference returns them; cigar only when trace is set.
struct SgResult {
    int32_t max, end1, end2;
    std::vector<uint32_t> cigar;
    int64_t cells = 0;  // cells the reference evaluates for this task
};
// W (may be null): the weighted (ZDP) DP with per-row score profiles
// (sequenceAlignment3.cpp:1483-1890) instead of SemiGlobal_no_avx.
void semiglobal_batch(dcns_ctx *c, const Params &P, const std::vector<SgTask> &tasks, bool trace,
                      std::vector<SgResult> &out, int64_t &cells, const SgWeights *W = nullptr);

// One score profile of the weighted paths: a file of the score folder
// (model/Score.cpp:9-71).
struct Profile {
    int16_t category;
    int32_t match, mismatch, open1, ext1, open2, ext2, zdrop;
};
// Categories of the pairs' reference positions -1 .. len1: len1 + 2 values
// from cat + off[pair] (the reference's weight pointer at [-1 .. len1]).
struct PairCategories {
    const int16_t *cat;
    const int64_t *off;
};



struct FragSet {
    std::vector<PairedSimilarFragment> frags;
    std::vector<int64_t> first, count, seeds_size;
};
void find_similar(dcns_ctx *c, const dcns_batch &b, const dcns_params &p, FragSet &out);
// Same on sequences already resident on the device (dcns_stage).
void find_similar_staged(dcns_ctx *c, Staged &S, const dcns_params &p, FragSet &out);
// findSimilarFragmentsForPairedSequence overload 4 (:752-774): getAllExtendSeed
// then overload 2 (:532-616).
void find_similar_2gaps_staged(dcns_ctx *c, Staged &S, const dcns_params &p, FragSet &out);
void find_similar_2gaps(dcns_ctx *c, const dcns_batch &b, const dcns_params &p, FragSet &out);
// findSimilarFragmentsForPairedSequence overload 3 (:623-745, the cut2Gaps2
// path): SemiGlobal from given seeds (seeds[pair], 0-based) with both DPs
// capped at max_len rows and columns.
void find_similar_seeded(dcns_ctx *c, const dcns_batch &b, const dcns_params &p,
                         const std::vector<std::vector<Seed>> &seeds, int32_t max_len, FragSet &out);
// Uploads the profiles and every pair's categories (as profile indices) for
// the weighted DP; mode 1 (two-piece) or 2 (single gap).
SgWeights stage_weights(dcns_ctx *c, const dcns_batch &b, const std::vector<Profile> &prof, const PairCategories &pc,
                        int mode);
// findSimilarFragmentsForPairedSequence_wighted (two_piece, :830-925) and
// _wighted_1gap (:779-826): getAllExtendSeed with p's scoring, then the
// weighted reverse / forward DP per x-seed with the profiles and categories.
void find_similar_weighted(dcns_ctx *c, const dcns_batch &b, const dcns_params &p,
                           const std::vector<Profile> &prof, const PairCategories &pc, bool two_piece, FragSet &out);

}  // namespace dcns

struct dcns_frag_set {
    dcns::FragSet fs;
};
