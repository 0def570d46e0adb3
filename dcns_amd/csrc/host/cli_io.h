// cli_io.h -- command-line and file I/O of the dCNS commands this library
// serves (cut1Gap, ranSco): argument lookup, FASTA with region metadata,
// samtools-indexed FASTA, sub-sequences, reverse complements, and the
// SAM / "o" writer.  Behaviour follows the reference byte for byte:
//   InputParser            InputParser.cpp:43-88
//   str2bool               InputParser.cpp:27-37
//   readFastaFile (meta)   impl/fasta.cpp:46-111
//   readFastaFile (plain)  impl/fasta.cpp:7-44
//   FastaWithIndex         impl/FastaWithIndex.cpp:11-71, impl/FastaIndexEntryImpl.cpp:8-28
//   getSubSequence / getReverseComplementary  impl/fasta.cpp:152-217
//   outputMasking          controlLayer4.cpp:835-1358
//   outputNoMasking        controlLayer4.cpp:56-117
#pragma once
#include <cstdint>
#include <fstream>
#include <map>
#include <ostream>
#include <string>
#include <vector>

#include "../../../include/dcns/model.h"

namespace dcns {
namespace cli {

class Args {
  public:
    Args(int argc, char **argv);
    bool has(const std::string &opt) const;
    std::string get(const std::string &opt) const;                 // "" when absent or last token
    std::vector<std::string> many(const std::string &opt) const;   // values up to the next '-'/'>' token
  private:
    std::vector<std::string> tokens_;
};

bool str2bool(std::string p, bool dflt);

struct Meta {
    std::string species, chr;
    int32_t start = 0, end = 0;
    int strand = 0;
};

void read_meta_fasta(const std::string &path, std::map<std::string, std::string> &seqs,
                     std::map<std::string, Meta> &meta, std::vector<std::string> &names);
void read_plain_fasta(const std::string &path, std::map<std::string, std::string> &seqs,
                      std::vector<std::string> &names);

class IndexedFasta {
  public:
    explicit IndexedFasta(const std::string &fasta);
    std::string fetch(const std::string &chr, int32_t start1, int32_t end1, int8_t strand);
    // clears the stream's error state (a fetch past the end of the file sets it)
    void clear() { in_.clear(); }
    struct Entry {
        int32_t length = 0;
        int64_t offset = 0;
        int32_t line_blen = 0, line_len = 0;
    };
    const std::map<std::string, Entry> &entries() const { return entries_; }

  private:
    std::map<std::string, Entry> entries_;
    std::ifstream in_;
};

// samFileToSeeds (impl/samFileToSeeds.cpp:19-141): the records of a cut1Gap
// SAM file as seeds in stripped (no 'n') coordinates, keyed by reference then
// query chromosome, plus-strand records in pos and minus-strand ones in neg
// (query coordinates on the reverse complement).  Prints the reference's
// progress line every 1000 lines.  Unknown chromosomes read as empty
// sequences, as the reference's map lookups make them.
using SeedMap = std::map<std::string, std::map<std::string, std::vector<Seed>>>;
void sam_to_seeds(const std::string &sam, std::map<std::string, std::string> &ref_genome,
                  std::map<std::string, std::string> &query_genome, SeedMap &pos, SeedMap &neg);

std::string reverse_complement(const std::string &s);
std::string sub_sequence(std::string s, const int &start, const int &end);
std::string sub_sequence(std::string &s, const int &start, const int &end, const int &strand);
std::vector<int8_t> encode(const std::string &s);

void write_sam_and_o(const std::string &prefix, const std::vector<PairedSimilarFragment> &frags, int32_t refStrand,
                     int32_t queStrand, int32_t refStart, int32_t queStart, const std::string &refChr,
                     const std::string &queChr, const std::string &seq1_0, const std::string &seq2_0);
// The same into streams the caller has opened (in append mode, .sam first).
void write_sam_and_o(std::ostream &sam, std::ostream &o, const std::vector<PairedSimilarFragment> &frags,
                     int32_t refStrand, int32_t queStrand, int32_t refStart, int32_t queStart,
                     const std::string &refChr, const std::string &queChr, const std::string &seq1_0,
                     const std::string &seq2_0);

// outputNoMasking, controlLayer4.cpp:56-117: coordinates without masked
// runs, SAM cigar ops "MID", header line first and "#done" last in the o file.
void write_sam_and_o_unmasked(const std::string &prefix, const std::vector<PairedSimilarFragment> &frags,
                              int32_t refStrand, int32_t queStrand, int32_t refStart, int32_t queStart,
                              const std::string &refChr, const std::string &queChr, const std::string &seq1,
                              const std::string &seq2);

}  // namespace cli
}  // namespace dcns
