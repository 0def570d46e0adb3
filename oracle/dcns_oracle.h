// dcns_oracle.h -- CPU restatement of dCNS's alignment hot path.
//
// TEST INFRASTRUCTURE ONLY.  This library is the checker: tests/, the smoke
// entry point and the CPU-baseline leg of bench.py may load it; the product
// library (dcns_amd) never links, loads or calls it.
//
// Every function restates one reference routine (file:line in the
// reference tree) in plain scalar C++, with full-matrix or full-row state so
// that it is easy to read against the reference.  Sequences are the
// reference's int8 codes: A0 C1 N2 G3 T4 (SequenceCharToUInt8.cpp:10-30).
//
// Pinning (how each function is checked, see tests/test_oracle_*.py):
//   orc_sw, orc_sw_textbook   golden vectors from the compiled reference,
//                             live comparison with oracle/_ref, brute-force
//                             path enumeration on tiny inputs, closed forms
//                             (identical pair -> 2L), textbook==variant under
//                             default scoring.
//   orc_xextend(_tb)          golden vectors + live _ref comparison.
//   orc_semiglobal            golden vectors + live _ref comparison.
//   orc_seed_set              golden vectors + live _ref comparison.
//   orc_find_similar          golden vectors + live _ref comparison, closed
//                             form of the p-value.
//   orc_semiglobal_weighted,  golden vectors + live _ref comparison; with one
//   orc_find_similar_weighted profile for every category the two-piece form
//                             equals SemiGlobal's recurrence except at the
//                             band's last column (tested).
//   orc_mask_genome,          golden files of the compiled reference's
//   orc_kmer_mask             maskGenome command + live comparison with
//                             oracle/_ref/dCNS_ref on random genomes, dumps,
//                             GFF3 and SAM files; brute force (a window
//                             compared with every member) for the scan.
#pragma once
#include <cstdint>

#ifdef __cplusplus
extern "C" {
#endif

// SmithWaterman(...bool& positions) dispatch, sequenceAlignment3.cpp:220-469.
// out3 = {max, end1, end2} (0-based; -1/-1 when no positive cell).
int orc_sw(const int8_t *ref, const int8_t *read, int32_t refLen, int32_t readLen, int32_t match,
           int32_t mismatch, int32_t open, int32_t ext, int32_t *out3);
// SmithWaterman_non_avx (positions=true), sequenceAlignment3.cpp:143-216.
int orc_sw_textbook(const int8_t *ref, const int8_t *read, int32_t refLen, int32_t readLen, int32_t match,
                    int32_t mismatch, int32_t open, int32_t ext, int32_t *out3);
// Which recurrence SmithWaterman uses: 1 = striped int8/int16 variant, 0 = textbook.
int orc_sw_uses_simd_variant(int32_t refLen, int32_t readLen, int32_t match, int32_t open, int32_t ext);
// Number of DP cells the reference evaluates for one SmithWaterman call.
int64_t orc_sw_cells(int32_t refLen, int32_t readLen);

// SemiGlobal_xextend score only, sequenceAlignment3.cpp:1234-1314.  out3 = {max, end1, end2}.
int orc_xextend(const int8_t *s1, const int8_t *s2, int32_t L1, int32_t L2, int32_t match, int32_t mismatch,
                int32_t open, int32_t ext, int32_t xdrop, int32_t w, int32_t *out3);
// SemiGlobal_xextend + traceback, sequenceAlignment3.cpp:1318-1451.
// out5 = {max, end1, end2, iii, jjj}; returns cigar length or -1 when cap is too small.
int orc_xextend_tb(const int8_t *s1, const int8_t *s2, int32_t L1, int32_t L2, int32_t match, int32_t mismatch,
                   int32_t open, int32_t ext, int32_t xdrop, int32_t w, int32_t *out5, uint32_t *cigar,
                   int32_t cap);
// SemiGlobal / SemiGlobal_no_avx (two-piece affine), sequenceAlignment3.cpp:681-941.
int orc_semiglobal(const int8_t *s1, const int8_t *s2, int32_t L1, int32_t L2, int32_t match, int32_t mismatch,
                   int32_t open1, int32_t ext1, int32_t open2, int32_t ext2, int32_t zdrop, int32_t w,
                   int32_t return_cigar, int32_t *out3, uint32_t *cigar, int32_t cap);

// Window loop + getSeeds recursion (+ link_seeds when linked != 0),
// findSimilarFragmentsForPairedSequence4.cpp:61-201, 288-344, 461-483.
// out = 4 int32 per seed in set order.  Returns the count (negative when cap too small).
int orc_seed_set(const int8_t *seq1, int32_t L1, const int8_t *seq2, int32_t L2, int32_t match, int32_t mismatch,
                 int32_t open, int32_t ext, int32_t window, int32_t step, int32_t min_score, int32_t linked,
                 int32_t *out, int32_t cap);

// findSimilarFragmentsForPairedSequence overload 1,
// findSimilarFragmentsForPairedSequence4.cpp:433-524.  Same output layout as
// ref_find_similar in ref_shim.cpp.  stats (may be null) receives
// {seed scan cells, x-drop cells, seeds before link, seeds after link}.
int orc_find_similar(const int8_t *seq1, int32_t L1, const int8_t *seq2, int32_t L2, int32_t match,
                     int32_t mismatch, int32_t open, int32_t ext, int32_t window, int32_t step, int32_t min_score,
                     int32_t w, int32_t xdrop, double pvalues, double lambda, double kvalue, int32_t *ints,
                     double *dbls, int32_t cap, uint32_t *cig, int64_t *cig_off, int64_t cig_cap, int64_t *stats);

// findSimilarFragmentsForPairedSequence overload 4 (the two-piece path of
// cut2Gaps / pairCns2Gaps), findSimilarFragmentsForPairedSequence4.cpp:752-774
// = getAllExtendSeed :348-429 + overload 2 :532-616.  Output layout as
// orc_find_similar; stats (may be null) receives {seed scan cells, x-drop
// cells, seeds before link, seeds after link, SemiGlobal cells}.
int orc_find_similar_2gaps(const int8_t *seq1, int32_t L1, const int8_t *seq2, int32_t L2, int32_t match,
                           int32_t mismatch, int32_t open, int32_t ext, int32_t open2, int32_t ext2,
                           int32_t window, int32_t step, int32_t min_score, int32_t w, int32_t xdrop, int32_t zdrop,
                           int32_t bandwidth, double pvalues, double lambda, double kvalue, int32_t *ints,
                           double *dbls, int32_t cap, uint32_t *cig, int64_t *cig_off, int64_t cig_cap,
                           int64_t *stats);

// Weighted semi-global DP: SemiGlobal(..., const int16_t *weights, Score &, ...),
// sequenceAlignment3.cpp:1483-1706 (two_piece = 1), or
// SemiGlobal_single_gap_penalty, :1710-1890 (two_piece = 0).  weights[-1 ..
// L1] must be readable (row i uses the score profile of weights[i-1] and the
// vertical gap penalties of weights[i]).  Profiles: nprof categories pcat[k]
// with pval[7k..7k+6] = {match, mismatch, open1, ext1, open2, ext2, zdrop}
// (the score-folder files, model/Score.cpp:9-71).  out3 = {max, end1, end2}
// (0-based); returns the cigar length (traceback) or 0; cells (may be null)
// receives the evaluated band cells.
int orc_semiglobal_weighted(const int8_t *s1, const int8_t *s2, int32_t L1, int32_t L2, const int16_t *weights,
                            int32_t nprof, const int16_t *pcat, const int32_t *pval, int32_t w, int32_t two_piece,
                            int32_t return_cigar, int32_t *out3, uint32_t *cigar, int32_t cap, int64_t *cells);

// findSimilarFragmentsForPairedSequence_wighted (two_piece = 1,
// findSimilarFragmentsForPairedSequence4.cpp:830-925) and
// findSimilarFragmentsForPairedSequence_wighted_1gap (two_piece = 0, :779-826):
// getAllExtendSeed with the fixed scoring, then per x-seed the weighted
// reverse and forward DP.  wcat holds the categories of reference positions
// -1 .. L1 (L1 + 2 values; the reference reads one base past each end of the
// region).  Output layout and stats as orc_find_similar_2gaps.
int orc_find_similar_weighted(const int8_t *seq1, int32_t L1, const int8_t *seq2, int32_t L2, int32_t match,
                              int32_t mismatch, int32_t open, int32_t ext, int32_t window, int32_t step,
                              int32_t min_score, int32_t w, int32_t xdrop, int32_t bandwidth, double pvalues,
                              double lambda, double kvalue, const int16_t *wcat, int32_t nprof, const int16_t *pcat,
                              const int32_t *pval, int32_t two_piece, int32_t *ints, double *dbls, int32_t cap,
                              uint32_t *cig, int64_t *cig_off, int64_t cig_cap, int64_t *stats);

// findSimilarFragmentsForPairedSequence overload 3 (the cut2Gaps2 path,
// findSimilarFragmentsForPairedSequence4.cpp:623-745): SemiGlobal from every
// seed (4 ints each: start1, end1, start2, end2, 0-based) in order, skipping
// seeds inside an earlier fragment, with both DPs capped at max_len rows and
// columns.  Seeds whose end lies outside the sequences are skipped.  Output
// layout as orc_find_similar; sgcells (may be null) receives the SemiGlobal
// cells.
int orc_find_similar_seeded(const int8_t *seq1, int32_t L1, const int8_t *seq2, int32_t L2, const int32_t *seeds,
                            int32_t nseeds, int32_t match, int32_t mismatch, int32_t open, int32_t ext,
                            int32_t open2, int32_t ext2, int32_t zdrop, int32_t bandwidth, int32_t max_len,
                            double lambda, double kvalue, int32_t *ints, double *dbls, int32_t cap, uint32_t *cig,
                            int64_t *cig_off, int64_t cig_cap, int64_t *sgcells);

// Max local score of many pairs (the ranSco inner call,
// permutationLsqLambdaK.cpp:97).  seqs is one buffer; offsets/lengths per pair.
int orc_sw_max_batch(const int8_t *seqs, const int64_t *ref_off, const int32_t *ref_len, const int64_t *qry_off,
                     const int32_t *qry_len, int32_t n, int32_t match, int32_t mismatch, int32_t open, int32_t ext,
                     int32_t *out_max, int32_t nthreads);

// maskGenome, controlLayer4.cpp:1999-2084 with impl/maskGenome.cpp:8-123 and
// impl/getAllTheCdsPositions.cpp:9-173 (mask_oracle.cpp): reads fasta, the
// jellyfish dump (null: none), the GFF3 (null: none; gene != 0 is -z) and the
// SAM + CDS fasta (null: none) and writes the masked fasta to out.  log
// receives the command's stdout lines, or the failure on error (returns -1).
int orc_mask_genome(const char *fasta, const char *out, const char *dump, int32_t freq, const char *gff,
                    int32_t gene, const char *sam, const char *cds_fasta, double similarity, char alt, int32_t soft,
                    int32_t width, char *log, int32_t logcap);
// The k-mer scan alone (maskGenome.cpp:15-61) over sequences
// seq[offsets[s], offsets[s + 1]): mask = 0/1 per position.  Returns k, or -1
// when a sequence is shorter than k or the dump cannot be read.
int orc_kmer_mask(const char *dump, int32_t freq, const char *seq, const int64_t *offsets, int32_t nseq,
                  uint8_t *mask);
// The distinct members of length k of the dump's k-mer profile, k bytes each
// in sorted order; returns their number (-2: cap too small, -1: error).
int orc_kmer_members(const char *dump, int32_t freq, char *out, int64_t cap, int32_t *k);

#ifdef __cplusplus
}
#endif
