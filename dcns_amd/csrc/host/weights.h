// weights.h -- inputs of the weighted (ZDP) paths: score profiles and the
// per-base categories of a GFF3 annotation.
#pragma once
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "pipeline.h"

namespace dcns {

// split, model/utils.cpp:9-17: tab-separated fields, empty fields dropped.
std::vector<std::string> split_fields(const std::string &line);

// Score::Score(folder), model/Score.cpp:9-71: one profile per line of
// folder/matrixList, in list order; a category listed again takes the later
// file's values (the reference's maps keep a value a shorter later file does
// not set).
std::vector<Profile> read_score_folder(const std::string &folder);

// Feature intervals of one chromosome (0-based, inclusive, start <= end),
// by the four classes of readGffFileWithEveryThing
// (impl/gffToCategory.cpp:98-150).
struct GffChrom {
    std::vector<std::pair<int64_t, int64_t>> introns, exons, cds, transposons;
};
struct GffAnnotation {
    std::map<std::string, GffChrom> chrom;
    std::vector<std::string> unknown_lines;  // lines of unknown feature type, in file order
};
// The reading pass of readGffFileWithEveryThing (:151-181).
GffAnnotation read_gff(const std::string &path);

// The categories readGffFileWithEveryThing assigns (:86-92, 198-221) to
// positions [from, from + count) of a chromosome of chr_size bases: 10
// (intergenic), then 70 over gene-like records, 85 over merged exons (the
// merge drops each chromosome's last merged interval, :11-33), 100 over CDS,
// 1 over transposable elements, each class painted over the previous one.
// Positions outside [0, chr_size) -- where the reference reads outside its
// array -- and records past the chromosome end are given / clipped to 10.
void gff_categories(const GffAnnotation &g, const std::string &chr, int64_t chr_size, int64_t from, int64_t count,
                    int16_t *out);

}  // namespace dcns
