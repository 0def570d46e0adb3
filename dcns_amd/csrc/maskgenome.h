// maskgenome.h -- maskGenome's device pipeline (impl/maskGenome.cpp:50-121):
// a genome and a k-mer set staged on the device, then the k-mer scan, the CDS
// ranges and the FASTA bodies.
#pragma once
#include <cstdint>
#include <vector>

#include "ctx.h"
#include "host/mask_io.h"

namespace dcns {

struct MaskStaged {
    int32_t nseq = 0;
    int64_t N = 0;             // bases of all sequences
    std::vector<int64_t> off;  // nseq + 1 sequence offsets into the concatenation
    int32_t k = 10;
    int64_t buckets = 0;       // minimizer buckets of the code members (0: no table)
    KmerGeneral general{};     // device pointers set at run time
    int64_t gslots = 0;
};

// Bytes of one sequence's FASTA body: the bases, a newline after every
// width-th base (width > 0; none for width < 0) and the closing newline.
int64_t mask_body_size(int64_t len, int32_t width);

// Uploads the sequences seq[off[s], off[s + 1]) and builds the k-mer tables.
// Every sequence must hold at least k bases when k >= 1 (the reference's scan
// throws std::out_of_range on a shorter one).
void mask_stage(dcns_ctx *c, const KmerSet &ks, const uint8_t *seq, const int64_t *off, int32_t nseq);

// Masks the staged genome: k-mer windows, then the CDS ranges of cds (null:
// none; else one entry per staged sequence, same order and sizes), then the
// bodies; mask_out (N bytes, 0/1) and body_out (the bodies back to back) are
// filled when non-null.
void mask_run(dcns_ctx *c, const CdsMask *cds, uint8_t alt, bool soft, int32_t width, uint8_t *mask_out,
              uint8_t *body_out);

}  // namespace dcns
