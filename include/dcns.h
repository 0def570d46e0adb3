/* dcns.h -- C-ABI of the MI355X (gfx950) implementation of dCNS's alignment
 * hot path.  Plain pointers and sizes only; no C++ or torch types cross this
 * boundary.  Every entry point returns 0 on success and a negative code on
 * failure; dcns_last_error() then describes the failure.  Nothing is
 * silently computed on the CPU: when no usable gfx950 device is present,
 * dcns_ctx_create fails.
 *
 * Sequences are int8 codes A0 C1 N2 G3 T4 (the encoding of the reference's
 * SequenceCharToUInt8::seq_to_int8, impl/SequenceCharToUInt8.cpp:46-52),
 * passed in host memory; the library stages them to device memory.
 *
 * Reference interfaces each entry point replaces (paths in the reference
 * tree):
 *   dcns_sw_bas_batch        SmithWaterman(...) of impl/SmithWaterman_bas.h:15-17
 *                            (body impl/SmithWaterman_bas.cpp:59-201)
 *   dcns_sw_batch            SmithWaterman(..., bool& positions)
 *                            impl/sequenceAlignment3.cpp:220 (decl.
 *                            impl/sequenceAlignment.h:26-28); caller
 *                            impl/permutationLsqLambdaK.cpp:97
 *   dcns_xdrop_extend        x_extend_seed (with traceback)
 *                            impl/findSimilarFragmentsForPairedSequence4.cpp:245
 *                            over SemiGlobal_xextend
 *                            impl/sequenceAlignment3.cpp:1234 and :1318
 *   dcns_seed_sets           the getSeeds window loop and link_seeds of
 *                            impl/findSimilarFragmentsForPairedSequence4.cpp:61-201,
 *                            288-344, 461-483
 *   dcns_find_similar_batch  findSimilarFragmentsForPairedSequence (overload 1)
 *                            impl/findSimilarFragmentsForPairedSequence4.cpp:433
 *                            (decl. impl/findSimilarFragmentsForPairedSequence.h:26-32);
 *                            callers controlLayer4.cpp:212, 1411
 */
#ifndef DCNS_H
#define DCNS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DCNS_OK 0
#define DCNS_ERR_ARG -1      /* invalid argument or unsupported parameter */
#define DCNS_ERR_DEVICE -2   /* HIP runtime / device failure */
#define DCNS_ERR_CAPACITY -3 /* caller-provided output buffer too small */
#define DCNS_ERR_INTERNAL -4

typedef struct dcns_ctx dcns_ctx;
typedef struct dcns_frag_set dcns_frag_set;

/* Scoring, seeding and significance parameters; defaults are the cut1Gap
 * defaults of controlLayer4.cpp:28-52. */
typedef struct dcns_params {
    int32_t match, mismatch;        /* Scorei(match, mismatch)           (2, -3) */
    int32_t open, ext;              /* gap open / extend, <= 0           (-4, -2) */
    int32_t open2, ext2;            /* second gap piece (SemiGlobal)     (-45, 0) */
    int32_t zdrop, bandwidth;       /* SemiGlobal z-drop and band        (50, 20000) */
    int32_t window, step, min_score;/* seed window / step / min score    (38, 8, 40) */
    int32_t w, xdrop;               /* x-drop band half-width and drop   (10, 20) */
    int32_t pad;
    double pvalue, lambda, k;       /* report p < pvalue; E = K*L1*L2*exp(-lambda*S) (0.1, 0.382291, 0.006662) */
} dcns_params;

/* A batch of sequence pairs in one host buffer of codes. */
typedef struct dcns_batch {
    const int8_t *seqs;
    const int64_t *off1;  /* reference (seq1) offset per pair */
    const int32_t *len1;
    const int64_t *off2;  /* query (seq2) offset per pair */
    const int32_t *len2;
    int32_t n;
    int32_t pad;
} dcns_batch;

/* Counters of the last pipeline call (reference-equivalent work). */
typedef struct dcns_stats {
    int64_t scan_cells;     /* DP cells of every SmithWaterman call the reference makes */
    int64_t xdrop_cells;    /* band cells of every SemiGlobal_xextend pass the reference makes */
    int64_t seeds;          /* seeds before link_seeds, summed over pairs */
    int64_t linked_seeds;   /* seeds after link_seeds */
    int64_t fragments;      /* reported fragments */
    int64_t gpu_scan_cells; /* cells the GPU evaluated in the seed scan (incl. speculative work) */
    int32_t levels;         /* recursion levels run on the device */
    int32_t pad;
    double ms_seed, ms_replay, ms_xdrop, ms_total; /* wall-clock phases */
    int64_t semiglobal_cells; /* cells of every SemiGlobal call the reference makes (two-piece path) */
    double ms_semiglobal;     /* wall-clock of the two-piece DP phase */
} dcns_stats;

const char *dcns_version(void);
void dcns_default_params(dcns_params *p);

/* Context: one HIP device and stream, device buffers reused across calls.
 * flags: bit 0 = record per-kernel HIP event timings. */
int dcns_ctx_create(int device, uint32_t flags, dcns_ctx **out);
int dcns_ctx_destroy(dcns_ctx *ctx);
const char *dcns_last_error(const dcns_ctx *ctx);
int dcns_device_info(dcns_ctx *ctx, int32_t *cu_count, int32_t *clock_mhz, char *name, int32_t name_cap);
int dcns_last_stats(const dcns_ctx *ctx, dcns_stats *out);

/* Launch shapes of the seed phase (the batching of the reference's
 * per-window, per-process calls; no counterpart in the reference, which runs
 * one pair per process).  Fields at 0 take the defaults; results never
 * depend on them, only the time does.
 *   level0_mode   1 (default): the level-0 window scan runs one workgroup
 *                 per 128-window task of the length-sorted task list; 2:
 *                 persistent waves (level0_waves per CU, default 8) pop the
 *                 list from a queue head in device memory.
 *   wave_tasks    later levels: at most this many of the longest rescans run
 *                 one per wave (default 1024) ...
 *   wave_rows     ... those of more than wave_rows rows (default 4096).
 *   graph_levels  level 0's resolve and this many recursion levels are
 *                 enqueued with no host read between them and replayed as
 *                 one captured HIP graph (default 4; -1: no graph, the host
 *                 reads the level counters after every launch).  Levels past
 *                 them run host-driven; when a pool overflows inside the
 *                 graph, the recursion runs again host-driven with grown
 *                 pools (kept for later calls).
 *   pool_min      > 0: the recursion's node, trajectory and task pools start
 *                 this many entries above the window count instead of a
 *                 quarter of it (makes the pools grow; for tests).
 *   xdrop_lane_rows  x-drop extension: a pass on the one-task-per-lane
 *                 kernels that runs past this many rows moves its task to
 *                 the one-task-per-wave kernel, and so does a traceback pass
 *                 of more rows (default 4096; -1: no limit).
 *   k1_lanes      dcns_sw_random_max: 0 (default): the longest points' pairs
 *                 one per wave (query stripes across the lanes, int32) where
 *                 an issue-slot model shows that shortens the launch, the
 *                 others two per lane in packed f16 when the scoring's f16
 *                 match and mismatch have a zero low byte and the gap costs
 *                 are within f16's exact integers (each pair whose maximum
 *                 reaches 2048 scored again in int32), else one per lane in
 *                 int32; 1: every pair one per lane, int32; 2: every pair on
 *                 the lane kernels as in 0; 3: every pair whose query fits
 *                 64 stripes (2048 bases) one per wave.
 *   level0_tail   cut1Gap seed recursion: 0 (default): on when a pair's
 *                 reference reaches 65,536 bases; 1: on; -1: off.  When on,
 *                 level 0 also records every row where a window's maximum
 *                 reaches min_score, and a rescan of rows [a, b) runs only
 *                 rows [a, a + D) (D = window + ceil(window * match / g),
 *                 g = -max(open, ext)); from row a + D on its values are
 *                 level 0's, whose recorded rows complete its trajectory.
 *                 Off when g <= 0 or the level-0 scan is not packed f16, and
 *                 for a batch whose recorded rows overflow their buffers.
 *   level0_split  cut1Gap level-0 scan: 0 (default): when the batch's
 *                 longest references would hold the launch, every 128-window
 *                 group runs its rows in segments (each started D rows early
 *                 from zeros, D as above, so its rows equal level 0's), as
 *                 many waves as segments, and the segments' trajectories
 *                 merge; 1: always, in segments of 4 D rows; -1: never.
 *                 Needs the conditions of level0_tail's D. */
typedef struct dcns_tuning {
    int32_t level0_mode, level0_waves;
    int32_t wave_tasks, wave_rows;
    int32_t graph_levels;
    int32_t pool_min;
    int32_t xdrop_lane_rows;
    int32_t k1_lanes;
    int32_t level0_tail;
    int32_t level0_split;
} dcns_tuning;
int dcns_get_tuning(const dcns_ctx *ctx, dcns_tuning *out);
int dcns_set_tuning(dcns_ctx *ctx, const dcns_tuning *in);

/* Per-kernel timing (flags bit 0): total milliseconds and launch count of the
 * kernel family `name` ("scan_uniform", "seed_levels", "scan_general",
 * "resolve", "xdrop_pass1", "xdrop_pass1_41", "xdrop_pass1_wave",
 * "xdrop_pass2", "xdrop_pass2_41", "xdrop_pass2_wave", "sw_batch") since the
 * last reset. */
int dcns_kernel_time(const dcns_ctx *ctx, const char *name, double *ms, int64_t *launches);
int dcns_kernel_time_reset(dcns_ctx *ctx);

/* SmithWaterman max score and first (ref-major) position for n pairs:
 * out3[3*i] = max, out3[3*i+1] = end1, out3[3*i+2] = end2 (0-based, -1 when
 * the max is 0). */
int dcns_sw_batch(dcns_ctx *ctx, const dcns_batch *b, int32_t match, int32_t mismatch, int32_t open, int32_t ext,
                  int32_t *out3);

/* SmithWaterman_bas for n pairs: the body in impl/SmithWaterman_bas.cpp:59-201
 * (declared at impl/SmithWaterman_bas.h:15-17 with a returnCigar argument no
 * body takes; the reference does not compile the file).  Fixed scores: match
 * 2, mismatch -1, N against a base 0, N against N -1; a gap of k bases costs
 * open + k * ext (open, ext <= 0, open + ext >= -128).  out3 (uint32, 3 per
 * pair): maxScore and the end positions as the function leaves them, 0-based
 * after its final decrement (0xFFFFFFFF when no cell scores above 0).
 * reverse != 0: the reverse initialisation and the traceback walk, on the
 * full matrices (12 bytes per cell of device memory): pair i's cigar entries
 * (len << 4 | op, in the walk's order) at cigar[cig_off[i] .. cig_off[i+1]),
 * status[i] 0 when the walk reached row or column 0, 1 when it stopped at a
 * zero cell (the reference then prints "there is something wrong with
 * smith-waterman algorithm in line 110" to stderr), 2 when the reference's
 * walk never ends (its state stops changing; the function does not return).
 * With reverse == 0, cigar may be NULL; cig_off and status, when given, get
 * zeros.  The reference also prints "maxScore <n>" to stdout on every call;
 * this entry point prints nothing. */
int dcns_sw_bas_batch(dcns_ctx *ctx, const dcns_batch *b, int32_t open, int32_t ext, int32_t reverse, uint32_t *out3,
                      uint32_t *cigar, int64_t *cig_off, int64_t cig_cap, int32_t *status);

/* dcns_sw_batch for pairs already in device memory, allocating nothing.  No
 * counterpart in the reference: it is for callers that cannot allocate device
 * memory (tight serving loops, harnesses that forbid hipMalloc).  The stripe
 * boundaries live in a 64 MiB device array the library holds on each device
 * from its first context there to the end of the process; every call on a
 * device shares it, so calls that overlap on one device race.  All pointers
 * are device pointers on the context's device: pair i's reference is
 * d_seqs[d_off1[i], + d_len1[i]) and its query d_seqs[d_off2[i], + d_len2[i])
 * (codes as for dcns_sw_batch), offsets multiples of 16, d_seqs 16-byte
 * aligned, and round_up(len, 16) bytes of a reference and round_up(len, 32)
 * of a query readable (the bytes past the length are ignored).  d_out3
 * receives 3 int32 per pair as dcns_sw_batch's out3.  max_len1 bounds every
 * d_len1[i]; it may not exceed (64 MiB - 256) / 8 - 16 rows (DCNS_ERR_ARG).
 * A pair longer than max_len1 is not computed and the call returns
 * DCNS_ERR_ARG.  Runs on the context's stream and returns when it is done. */
int dcns_sw_batch_noalloc(dcns_ctx *ctx, const int8_t *d_seqs, const int64_t *d_off1, const int32_t *d_len1,
                          const int64_t *d_off2, const int32_t *d_len2, int32_t n, int32_t max_len1, int32_t match,
                          int32_t mismatch, int32_t open, int32_t ext, int32_t *d_out3);

/* Max local scores of random fragment pairs generated on the device: the
 * sampling loop of permutationLsqLambdaK (impl/permutationLsqLambdaK.cpp:32-102),
 * which scores random fragment pairs with SmithWaterman, with i.i.d. fragments
 * of a given composition in place of fragments cut from genome files.  Point k
 * has pairs_per_point pairs of lengths len1[k] x len2[k]; bases are C or G
 * with probability gc[k]/2 each and A or T with (1 - gc[k])/2 each, drawn from
 * a counter-based hash of seed[k] (base b of pair r: bits of
 * splitmix64(seed[k] + 2r (+1 for the query)) + b/2, low half for even b).
 * out[k * pairs_per_point + r] = the pair's SmithWaterman maximum. */
int dcns_sw_random_max(dcns_ctx *ctx, int32_t match, int32_t mismatch, int32_t open, int32_t ext, const double *gc,
                       const int32_t *len1, const int32_t *len2, const uint64_t *seed, int32_t npoints,
                       int32_t pairs_per_point, int32_t *out);

/* x_extend_seed (with traceback) for n seeds.  seeds[2*i] = start1,
 * seeds[2*i+1] = start2 (the seed start + 1, as the reference passes them);
 * pair_of_seed[i] indexes b.  out_ints (6 per seed): maxScore, start1, end1,
 * start2, end2 (0-based, after the traceback trim), ok (1 = both passes
 * found a positive cell, else the reference's p = 100 case).  Cigars are
 * written to cigar (capacity cig_cap ops) with cig_off[n+1] offsets. */
int dcns_xdrop_extend(dcns_ctx *ctx, const dcns_batch *b, const dcns_params *p, const int32_t *pair_of_seed,
                      const int32_t *seeds, int32_t n, int32_t *out_ints, uint32_t *cigar, int64_t *cig_off,
                      int64_t cig_cap);

/* The seed set of every pair (window loop + getSeeds recursion + std::set
 * semantics), optionally after link_seeds.  counts[i] = seeds of pair i;
 * seeds receives 4 int32 per seed (start1, end1, start2, end2, 0-based) in
 * set order, pair after pair.  Returns DCNS_ERR_CAPACITY (with counts
 * filled) when cap seeds do not fit. */
int dcns_seed_sets(dcns_ctx *ctx, const dcns_batch *b, const dcns_params *p, int32_t linked, int32_t *counts,
                   int32_t *seeds, int64_t cap);

/* link_seeds (impl/findSimilarFragmentsForPairedSequence4.cpp:287-344) on the
 * host: the n seeds (4 int32 each: start1, end1, start2, end2) are inserted
 * into a std::set<Seed> in the order given, as the reference's window loop
 * inserts them, then merged.  insertion_order = 0: out receives the linked
 * set in set order; 1: the merged list in the order link_seeds inserts it
 * back into its set (inserting it so reproduces that set: Seed's comparator
 * is not a strict weak order, so the order matters).  *n_out = the count;
 * DCNS_ERR_CAPACITY when it exceeds cap (never above n).  Needs no device or
 * context. */
int dcns_link_seeds(const int32_t *seeds, int64_t n, int32_t insertion_order, int32_t *out, int64_t cap,
                    int64_t *n_out);

/* findSimilarFragmentsForPairedSequence (overload 1) for every pair. */
int dcns_find_similar_batch(dcns_ctx *ctx, const dcns_batch *b, const dcns_params *p, dcns_frag_set **out);
/* findSimilarFragmentsForPairedSequence overload 4 -- the two-piece affine
 * path of cut2Gaps / pairCns2Gaps
 * (impl/findSimilarFragmentsForPairedSequence4.cpp:752-774: getAllExtendSeed
 * :348-429, then overload 2 :532-616; callers controlLayer4.cpp:355, 1584).
 * Uses open2/ext2 (second gap piece), zdrop and bandwidth of p. */
int dcns_find_similar_2gaps_batch(dcns_ctx *ctx, const dcns_batch *b, const dcns_params *p, dcns_frag_set **out);

/* SemiGlobal (two-piece affine, z-drop, banded; impl/sequenceAlignment3.cpp:932
 * -> SemiGlobal_no_avx :681) for n tasks.  Task i aligns pair
 * pair_of_task[i] from anchors[2i], anchors[2i+1] (0-based): forward on
 * seq1[a1..), seq2[a2..) or, when reverse[i] != 0, on the reverse complements
 * of seq1[..a1] and seq2[..a2] (the reverse alignment of overload 2, :564).
 * out3 = max, end1, end2 (0-based; -2 when no cell is positive).  With
 * with_cigar, cigars (len << 4 | op, op 0 M, 1 I, 2 D) go to cigar with
 * cig_off[n+1] offsets. */
int dcns_semiglobal_batch(dcns_ctx *ctx, const dcns_batch *b, const dcns_params *p, const int32_t *pair_of_task,
                          const int32_t *anchors, const int32_t *reverse, int32_t n, int32_t with_cigar,
                          int32_t *out3, uint32_t *cigar, int64_t *cig_off, int64_t cig_cap);

/* findSimilarFragmentsForPairedSequence overload 3 -- the cut2Gaps2 path
 * (impl/findSimilarFragmentsForPairedSequence4.cpp:623-745, called from
 * controlLayer4.cpp:1928 and :1967).  Pair i's seeds are seeds[4k .. 4k+3] =
 * start1, end1, start2, end2 (0-based) for k in [seed_first[i],
 * seed_first[i+1]), in the reference's order; both SemiGlobal DPs of a seed
 * are capped at max_align_length rows and columns.  Uses open2/ext2, zdrop,
 * bandwidth, lambda and k of p.  Seeds whose end lies outside the sequences
 * are skipped (the reference reads outside its arrays for them); where the
 * reference's end correction (:661-665) makes the reverse DP reach one base
 * before a sequence, that base is N. */
int dcns_find_similar_seeded_batch(dcns_ctx *ctx, const dcns_batch *b, const dcns_params *p, const int32_t *seeds,
                                   const int64_t *seed_first, int32_t max_align_length, dcns_frag_set **out);

/* ---- Weighted (ZDP) paths: weighted1Gap / weighted2Gaps ------------------ */

/* One score profile: the seven lines of a score-folder file
 * (scoreMatrix/<category>, model/Score.cpp:9-71). */
typedef struct dcns_profile {
    int32_t category;
    int32_t match, mismatch, open1, ext1, open2, ext2, zdrop;
} dcns_profile;

/* Score::Score(folder) (model/Score.cpp:9-71): the profiles listed in
 * folder/matrixList, in list order (a category listed again takes the later
 * file's values).  n receives the count; DCNS_ERR_CAPACITY when it exceeds
 * cap.  No device needed. */
int dcns_read_score_folder(const char *folder, dcns_profile *out, int32_t cap, int32_t *n);

/* readGffFileWithEveryThing (impl/gffToCategory.cpp:81-238): the category of
 * each position [from, from + count) (0-based) of chromosome chr (chr_size
 * bases) under the GFF3 annotation: 10 intergenic, 70 gene-like records, 85
 * merged exons (without each chromosome's last merged interval, as
 * removeDuplications :11-33 leaves it), 100 CDS, 1 transposable elements.
 * Positions outside the chromosome, which the reference reads outside its
 * array, get 10.  Records of other chromosomes are ignored.  No device
 * needed. */
int dcns_gff_categories(const char *gff_path, const char *chr, int64_t chr_size, int64_t from, int64_t count,
                        int16_t *out);

/* findSimilarFragmentsForPairedSequence_wighted (two_piece != 0,
 * impl/findSimilarFragmentsForPairedSequence4.cpp:830-925) and
 * findSimilarFragmentsForPairedSequence_wighted_1gap (two_piece == 0,
 * :779-826), called from controlLayer4.cpp:701 (weighted2Gaps) and :525
 * (weighted1Gap), for every pair.  p's scoring drives getAllExtendSeed; the
 * weighted DP takes its scores from the profiles.  cat + cat_off[i] holds the
 * categories of pair i's reference positions -1 .. len1[i] (len1[i] + 2
 * values: the reference's weight pointer at [-1 .. len1]); every category
 * needs a profile.  p->bandwidth is the DP band. */
int dcns_find_similar_weighted_batch(dcns_ctx *ctx, const dcns_batch *b, const dcns_params *p,
                                     const dcns_profile *prof, int32_t nprof, const int16_t *cat,
                                     const int64_t *cat_off, int32_t two_piece, dcns_frag_set **out);

/* The weighted SemiGlobal (impl/sequenceAlignment3.cpp:1483, two-piece) or
 * SemiGlobal_single_gap_penalty (:1710) for n tasks, laid out as
 * dcns_semiglobal_batch; categories and profiles as
 * dcns_find_similar_weighted_batch. */
int dcns_semiglobal_weighted_batch(dcns_ctx *ctx, const dcns_batch *b, const dcns_params *p,
                                   const dcns_profile *prof, int32_t nprof, const int16_t *cat,
                                   const int64_t *cat_off, int32_t two_piece, const int32_t *pair_of_task,
                                   const int32_t *anchors, const int32_t *reverse, int32_t n, int32_t with_cigar,
                                   int32_t *out3, uint32_t *cigar, int64_t *cig_off, int64_t cig_cap);

/* Keeps a batch resident in device memory (replacing any earlier one; the
 * other entry points that take a batch drop it) and runs
 * dcns_find_similar_batch on it without re-uploading.  No reference
 * counterpart: the reference reads each region per process. */
int dcns_stage(dcns_ctx *ctx, const dcns_batch *b, const dcns_params *p);
int dcns_find_similar_staged(dcns_ctx *ctx, const dcns_params *p, dcns_frag_set **out);
int64_t dcns_frag_set_total(const dcns_frag_set *fs);
/* Fragments of one pair occupy [first, first + count) in the set. */
int dcns_frag_set_pair(const dcns_frag_set *fs, int32_t pair, int64_t *first, int64_t *count, int64_t *seeds_size);
/* Fragment idx: ints5 = start1, end1, start2, end2 (1-based as stored by
 * PairedSimilarFragment), score; dbls2 = pValue, eValue. *ncig receives the
 * cigar length; up to cig_cap ops (len << 4 | op, op 0 M, 1 I, 2 D) are
 * copied to cigar. */
int dcns_frag_set_get(const dcns_frag_set *fs, int64_t idx, int32_t *ints5, double *dbls2, uint32_t *cigar,
                      int32_t cig_cap, int32_t *ncig);
void dcns_frag_set_free(dcns_frag_set *fs);

/* Least-squares fit of lambda and K to maximum local scores of random pairs
 * (the FP64 step of the permutation method; host code).  Replaces the
 * reference's R fit, scripts/scriptToestimateKandLambda.r:4-15 (nlsLM of
 * y ~ 1 - exp(-K*N*exp(-lambda*x)) over the empirical survival y(x) of the
 * sorted unique scores x, fed by ranSco, permutationLsqLambdaK.cpp:97-100).
 * N is fixed_space for every pair when log_space is NULL (the reference's
 * N = 1e6), else pair i contributes ln(m_i*n_i) = log_space[i] and the model
 * averages the survival over the pairs' search spaces.  out4 = K, lambda,
 * residual sum of squares, iterations. */
int dcns_fit_lambda_k(const int32_t *scores, int64_t n, const double *log_space, double fixed_space, double k0,
                      double lambda0, int32_t maxiter, double *out4);
/* calculateLambda (impl/calculateLambda.h:11, calculateLambda.cpp:64): the
 * ungapped Karlin-Altschul lambda for the A/C/G/T composition of seq1 (seq2
 * is ignored, as in the reference), Newton's method with a forward-difference
 * derivative of step epsilon until a step is <= precision.  The reference
 * reads uninitialised frequencies and truncates the step size to an integer;
 * this is the well-defined computation it describes.  NaN on bad input. */
double dcns_calculate_lambda(const int8_t *seq1, const int8_t *seq2, uint32_t len1, uint32_t len2, int32_t match,
                             int32_t mismatch, double epsilon, double precision);

/* ngroups independent fits, fixed search space each: group g = scores[g *
 * group_size, (g + 1) * group_size), N = space[g]; out4 as above per group. */
int dcns_fit_lambda_k_batch(const int32_t *scores, int64_t ngroups, int32_t group_size, const double *space,
                            double k0, double lambda0, int32_t maxiter, double *out4);

/* Commands with the reference's flags, files and output text (argv[0] is the
 * command name):
 *   dcns_cut1gap_main  int cut1Gap(int argc, char** argv), controlLayer4.cpp:1441
 *                      (decl. controlLayer.h:109); all query regions of the
 *                      input are aligned in one batched device call
 *   dcns_ransco_main   int smitherWaterManScoreOfRandomFragments(int argc, char** argv),
 *                      controlLayer4.cpp:2088 (decl. controlLayer.h:113) */
int dcns_cut1gap_main(int argc, char **argv);
int dcns_ransco_main(int argc, char **argv);
/*   dcns_cut2gaps_main      void/int cut2Gaps(...), controlLayer4.cpp:1534 / :1608
 *                           (decl. controlLayer.h); all query regions in one batch
 *   dcns_paircns2gaps_main  int pairCns2Gaps(int argc, char** argv), controlLayer4.cpp:378
 *                           (the queries are the input's sequences; the reference
 *                           clears them first, :346, and aborts)
 *   dcns_paircnsxextend_main int pairCnsXExtend(int argc, char** argv), controlLayer4.cpp:238
 *   dcns_weighted1gap_main   int weighted1Gap(int argc, char** argv), controlLayer4.cpp:554
 *   dcns_weighted2gaps_main  int weighted2Gaps(int argc, char** argv), controlLayer4.cpp:729
 *                           (all query regions in one batch; the GFF3 categories
 *                           of the reference region only)
 *   dcns_cut2gaps2_main      int cut2Gaps2(int argc, char** argv), controlLayer4.cpp:1815
 *                           (every chromosome pair of the SAM file in one batch) */
int dcns_cut2gaps_main(int argc, char **argv);
int dcns_paircns2gaps_main(int argc, char **argv);
int dcns_paircnsxextend_main(int argc, char **argv);
int dcns_weighted1gap_main(int argc, char **argv);
int dcns_weighted2gaps_main(int argc, char **argv);
int dcns_cut2gaps2_main(int argc, char **argv);

/* ---- maskGenome (controlLayer4.cpp:1999-2084) ------------------------------
 * The reference masks a genome in one function,
 *   void maskGenome(std::map<std::string, std::string>& genome, char& alterChar,
 *                   std::string& kmerProfileFile, int32_t& maskKmerFrequency,
 *                   int32_t& outputLineWidth, std::map<std::string, std::string>& ifCds,
 *                   std::string& outputFile, std::vector<std::string>& seqNames,
 *                   bool& ifSoftmask)                       impl/maskGenome.cpp:8
 * after gffToMask / gffToMaskGene / samToMask have filled ifCds
 * (impl/getAllTheCdsPositions.cpp:9, :59, :105).  Here the same steps are
 * separate calls: the k-mer profile (host), the CDS masks (host), and the
 * scan plus FASTA writer on the device over a staged genome. */
typedef struct dcns_kmer_set dcns_kmer_set;
typedef struct dcns_cds_mask dcns_cds_mask;

/* The k-mer profile, impl/maskGenome.cpp:15-48: the records of a jellyfish
 * dump ("><count>" line, then sequence lines) whose count is above freq,
 * upper-cased; k is the length of the dump's last record.  freq <= 0 reads
 * nothing and gives the empty set with k = 10, as the reference does. */
int dcns_kmer_set_read(const char *path, int32_t freq, dcns_kmer_set **out);
/* members: members of length k, the only ones a window can equal (a k-mer
 * listed twice counts twice); records: all records above the threshold. */
int dcns_kmer_set_info(const dcns_kmer_set *s, int32_t *k, int64_t *members, int64_t *records);
/* The members of length k, k bytes each, back to back (cap >= members * k). */
int dcns_kmer_set_get(const dcns_kmer_set *s, char *out, int64_t cap);
void dcns_kmer_set_free(dcns_kmer_set *s);

/* The reference's ifCds over a genome of n uniquely named sequences. */
int dcns_cds_mask_create(const char *const *names, const int64_t *sizes, int32_t n, dcns_cds_mask **out);
/* gffToMask (gene = 0: CDS, start_codon, stop_codon), getAllTheCdsPositions.cpp:9-56, or
 * gffToMaskGene (gene = 1: gene records with "biotype=protein_coding"), :59-101.  Starts from
 * empty masks.  A record starting past the end of its sequence's mask string is an error
 * (the reference throws std::out_of_range). */
int dcns_cds_mask_gff(dcns_cds_mask *m, const char *gff_path, int32_t gene);
/* samToMask, getAllTheCdsPositions.cpp:105-173: M/D runs of the SAM records whose cs-tag
 * identical bases reach similarity x the CDS length (CDS sequences from cds_fasta, read as
 * impl/fasta.cpp:7-44 reads them).  *rebuilt = 1 when the masks were first rebuilt empty
 * (no GFF masks, or one ran past a sequence end): the reference then prints "GFF was not
 * used for genome masking". */
int dcns_cds_mask_sam(dcns_cds_mask *m, const char *sam_path, const char *cds_fasta, double similarity,
                      int32_t *rebuilt);
/* 0/1 per position of sequence s (cap >= its size). */
int dcns_cds_mask_get(const dcns_cds_mask *m, int32_t s, uint8_t *out, int64_t cap);
void dcns_cds_mask_free(dcns_cds_mask *m);

/* Stages a genome on the context's device for maskGenome.cpp:50-61: sequences
 * seq[offsets[s], offsets[s + 1]) (offsets[0] = 0, n + 1 entries) and the k-mer
 * set (NULL: no dump, k = 10).  Every sequence must hold at least k bases when
 * k >= 1: the reference's scan throws std::out_of_range on a shorter one. */
int dcns_mask_stage(dcns_ctx *ctx, const dcns_kmer_set *kmers, const char *seq, const int64_t *offsets,
                    int32_t nseq);
/* Bytes of one FASTA body written by dcns_mask_run: the bases, a newline after every width-th
 * base (none for width < 0), and the closing newline (maskGenome.cpp:66-121). */
int64_t dcns_mask_body_size(int64_t len, int32_t width);
/* maskGenome.cpp:50-121 on the staged genome: a position is masked when a window of k bases
 * inside its sequence covering it equals a member, or when cds (NULL: none; same sequences,
 * order and sizes as staged) marks it.  mask_out (N bytes, 0/1) and body_out (the FASTA
 * bodies back to back: bases upper-cased, masked ones written as alt, or lower-cased when
 * soft != 0) are filled when non-NULL.  width 0 is an error (the reference divides by it). */
int dcns_mask_run(dcns_ctx *ctx, const dcns_cds_mask *cds, char alt, int32_t soft, int32_t width,
                  uint8_t *mask_out, char *body_out);
/* int maskGenome(int argc, char** argv), controlLayer4.cpp:1999: the command with its flags,
 * stdout lines and output file. */
int dcns_mask_genome_main(int argc, char **argv);
/* No counterpart in the reference, whose regions fan out as one process per
 * region under GNU parallel (README.md:117-122).  Runs such a command file
 * (one "dCNS <command> <flags>" line per job) in one process: consecutive
 * cut1Gap lines that differ only in -i / -o are aligned in chunked batches
 * (about 4096 pairs, "-n" to change), the next chunk's regions read while the
 * device aligns the current one; every other line runs through its command.
 * argv: "batch" <file> [-n pairs].  Files and stdout are those of running the
 * lines in order; the status is 1 when any line failed. */
int dcns_batch_main(int argc, char **argv);

/* lambda/K surrogate network.  No counterpart in the reference, which obtains
 * lambda and K only by permutation (ranSco, permutationLsqLambdaK.cpp:32-102,
 * then scripts/scriptToestimateKandLambda.r).  Inputs per row (8 floats):
 * match, mismatch, gap open, gap extend, long-gap open, GC fraction,
 * ln(reference length), ln(query length); outputs (2 floats): ln(lambda),
 * ln(K).  8 -> 4 x 256 (GELU) -> 2, BF16 operands on the matrix cores with
 * FP32 accumulation, FP32 master weights trained by AdamW on the device.
 *
 * Parameters cross the boundary as FP32 in the logical order: for each of the
 * five layers, W [out x in] row-major, then b [out] (200,194 values).  Inputs
 * and outputs are standardised inside the network by the scaling set with
 * dcns_mlp_set_scaling (identity by default); dcns_mlp_predict returns
 * unstandardised outputs; dcns_mlp_loss_grad and dcns_mlp_train use the mean
 * squared error of the standardised outputs. */
typedef struct dcns_mlp dcns_mlp;
/* A network lives on its context's device and stream: destroy it before the
 * context. */
int dcns_mlp_create(dcns_ctx *ctx, uint64_t seed, dcns_mlp **out);
void dcns_mlp_destroy(dcns_mlp *m);
int64_t dcns_mlp_num_params(const dcns_mlp *m);
int dcns_mlp_set_scaling(dcns_mlp *m, const float *in_mean8, const float *in_std8, const float *out_mean2,
                         const float *out_std2);
int dcns_mlp_get_params(const dcns_mlp *m, float *out, int64_t cap);
int dcns_mlp_set_params(dcns_mlp *m, const float *in, int64_t n);
/* X [n x 8] -> Y [n x 2] (ln lambda, ln K), batches of 65,536 rows. */
int dcns_mlp_predict(dcns_mlp *m, const float *X, int64_t n, float *Y);
/* One forward and backward pass over all n rows, no update: loss and the
 * gradient in the logical parameter order (grad may be NULL). */
int dcns_mlp_loss_grad(dcns_mlp *m, const float *X, const float *Y, int64_t n, float *loss, float *grad,
                       int64_t grad_cap);
/* AdamW (beta 0.9 / 0.999, eps 1e-8, decoupled decay on weights only) over
 * `steps` minibatches of `batch` rows drawn with replacement (seeded), or,
 * when batch >= n, `steps` full-batch steps over the rows in order;
 * losses[steps] (optional) receives each step's loss. */
int dcns_mlp_train(dcns_mlp *m, const float *X, const float *Y, int64_t n, int32_t steps, int32_t batch, float lr,
                   float weight_decay, uint64_t seed, float *losses);

#ifdef __cplusplus
}
#endif

#endif /* DCNS_H */
