// mask_io.h -- host side of maskGenome (controlLayer4.cpp:1999-2084): the
// k-mer profile of a jellyfish dump (impl/maskGenome.cpp:15-48) and the CDS
// masks of gffToMask, gffToMaskGene and samToMask
// (impl/getAllTheCdsPositions.cpp:9-173).
#pragma once
#include <cstdint>
#include <iosfwd>
#include <map>
#include <string>
#include <utility>
#include <vector>

namespace dcns {

// The k-mers of a dump whose count is above a threshold, upper-cased.  k is
// the length of the dump's last record (10 when no dump is read, as the
// reference's kmerSize); only members of that length can equal a window.
struct KmerSet {
    int32_t k = 10;
    std::vector<uint64_t> codes;   // members of upper-case A/C/G/T (k <= 31): base j in bits 2j, 2j+1 as (b >> 1) & 3
    std::vector<uint8_t> general;  // every other member of length k, k bytes each
    int64_t records = 0;           // records above the threshold, any length
    int64_t n_general() const { return k > 0 ? (int64_t)general.size() / k : 0; }
};

// freq <= 0 reads nothing (the reference only reads the dump for a positive
// threshold) and gives the empty set with k = 10.
KmerSet read_kmer_dump(const std::string &path, int32_t freq);

// 2-bit code of a byte as the scan kernel reads it; -1 for other bytes.
inline int kmer_base(uint8_t b) {
    const uint32_t d = (uint32_t)b - 64u;
    return (d < 32u && ((0x10008Au >> d) & 1u)) ? (int)((b >> 1) & 3u) : -1;
}
uint64_t kmer_hash(const uint8_t *p, int32_t k);  // the general members' polynomial hash
uint64_t kmer_hash_pk1(int32_t k);                // base^(k-1)

// The reference's ifCds strings, kept as ranges: per genome sequence the
// positions [a, b) set to '1' (clipped to the sequence), and the string's
// length, which replace() grows when a range runs past it.  A range starting
// past the string's end makes the reference throw std::out_of_range; here it
// is an error.
class CdsMask {
  public:
    CdsMask(const std::vector<std::string> &names, const std::vector<int64_t> &sizes);
    // gffToMask (CDS, start_codon, stop_codon) or, with gene, gffToMaskGene
    // (gene records whose attributes hold "biotype=protein_coding").
    void gff(const std::string &path, bool gene);
    // samToMask with the CDS lengths of cds_sizes (sequence name -> length);
    // returns whether the masks were rebuilt empty first (no GFF masks, or a
    // GFF range ran past a sequence end), which is reported on log (when
    // given) as "GFF was not used for genome masking".
    bool sam(const std::string &path, const std::map<std::string, int64_t> &cds_sizes, double similarity,
             std::ostream *log = nullptr);
    int32_t size() const { return (int32_t)names_.size(); }
    int64_t seq_size(int32_t s) const { return size_[(size_t)s]; }
    const std::vector<std::pair<int64_t, int64_t>> &ranges(int32_t s) const { return ranges_[(size_t)s]; }
    int64_t length(int32_t s) const { return len_[(size_t)s]; }

  private:
    void reset();
    void mark(const std::string &chr, int64_t pos, int64_t len, const std::string &what);
    std::vector<std::string> names_;
    std::map<std::string, int32_t> index_;
    std::vector<int64_t> size_, len_;
    std::map<std::string, int64_t> other_len_;  // sequences the genome does not hold
    std::vector<std::vector<std::pair<int64_t, int64_t>>> ranges_;
    bool any_ = false;  // the reference's ifCds map holds an entry
};

}  // namespace dcns
