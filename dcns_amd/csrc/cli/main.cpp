// main.cpp -- the dCNS command-line drop-in for the commands this library
// implements on the GPU (dispatch as in the reference's main.cpp:8-50).
#include <cstring>
#include <iostream>

#include "../../../include/dcns.h"

static void usage() {
    std::cout << "Program: dCNS (MI355X build, " << dcns_version() << ")\n"
              << "Usage: dCNS <command> [options]\n"
              << "Commands:\n"
              << "    cut1Gap             masking genome, sliding window seed, x-drop extension\n"
              << "    cut2Gaps            masking genome, sliding window seed, x-drop and two-piece gap extension\n"
              << "    cut2Gaps2           two-piece gap extension of the alignments of a cut1Gap sam file\n"
              << "    pairCnsXExtend      sliding window seed and x-drop extension of the sequences of one fasta file\n"
              << "    pairCns2Gaps        the same with two-piece gap extension\n"
              << "    weighted1Gap        x-drop seeds extended by the weighted (per-base GFF class) alignment\n"
              << "    weighted2Gaps       the same with the weighted two-piece gap alignment\n"
              << "    maskGenome          mask repetitive k-mers, CDS (GFF3) or aligned CDS (SAM) in a genome\n"
              << "    ranSco              smith-waterman scores of random sequence fragments\n"
              << "                        for lambda and k fitting\n"
              << "    batch               run a file of dCNS command lines in one process (cut1Gap lines batched)\n";
}

int main(int argc, char **argv) {
    if (argc <= 1) {
        usage();
        return 1;
    }
    const char *p = argv[1];
    if (!std::strcmp(p, "-h") || !std::strcmp(p, "--help")) {
        usage();
        return 1;
    }
    if (!std::strcmp(p, "cut1Gap")) return dcns_cut1gap_main(argc - 1, argv + 1);
    if (!std::strcmp(p, "ranSco")) return dcns_ransco_main(argc - 1, argv + 1);
    if (!std::strcmp(p, "cut2Gaps")) return dcns_cut2gaps_main(argc - 1, argv + 1);
    if (!std::strcmp(p, "cut2Gaps2")) return dcns_cut2gaps2_main(argc - 1, argv + 1);
    if (!std::strcmp(p, "pairCns2Gaps")) return dcns_paircns2gaps_main(argc - 1, argv + 1);
    if (!std::strcmp(p, "pairCnsXExtend")) return dcns_paircnsxextend_main(argc - 1, argv + 1);
    if (!std::strcmp(p, "weighted1Gap")) return dcns_weighted1gap_main(argc - 1, argv + 1);
    if (!std::strcmp(p, "weighted2Gaps")) return dcns_weighted2gaps_main(argc - 1, argv + 1);
    if (!std::strcmp(p, "maskGenome")) return dcns_mask_genome_main(argc - 1, argv + 1);
    if (!std::strcmp(p, "batch")) return dcns_batch_main(argc - 1, argv + 1);
    std::cerr << "dCNS: command " << p << " is not part of this build" << std::endl;
    usage();
    return 1;
}
