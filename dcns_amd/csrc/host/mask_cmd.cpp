// mask_cmd.cpp -- the maskGenome command (controlLayer4.cpp:1999-2084): flags,
// defaults, stdout lines and the masked FASTA file.  The k-mer scan and the
// FASTA bodies are computed on the GPU.
#include <cstdio>
#include <iostream>
#include <map>
#include <string>
#include <vector>

#include "../../../include/dcns.h"
#include "../ctx.h"
#include "../maskgenome.h"
#include "cli_io.h"
#include "mask_io.h"

namespace dcns {
namespace cli {
namespace {

void mask_usage(int32_t freq, int32_t width, char alt) {
    std::cerr << "Usage: dCNS maskGenome -i input.fa -o output.fa [options]\n"
              << "Options\n"
              << " -i FILE    genome to mask, fasta\n"
              << " -o FILE    masked genome, fasta\n"
              << " -g FILE    GFF3 annotation: mask CDS (start/stop codons included)\n"
              << " -z         with -g: mask protein-coding genes instead of CDS\n"
              << " -s FILE    SAM file of CDS sequences aligned to the genome (needs -c)\n"
              << " -c FILE    the CDS sequences of -s, fasta\n"
              << " -p FLOAT   with -s: fraction of a CDS that must align identically (default: 0.8)\n"
              << " -k FILE    k-mer counts, fasta dump of jellyfish\n"
              << " -l CHAR    character written for masked positions (default: " << alt << ")\n"
              << " -d         soft mask (lower case) instead of -l\n"
              << " -f INT     mask k-mers counted more than INT times (default: " << freq << ")\n"
              << " -w INT     output line width (default: " << width << ")\n";
}

int32_t int_flag(const Args &a, const char *opt) {
    try {
        return std::stoi(a.get(opt));
    } catch (const std::exception &) {
        throw Error(DCNS_ERR_ARG, std::string("option ") + opt + ": not an integer");
    }
}

int mask_genome(int argc, char **argv) {
    int32_t freq = 50, width = 60;
    double similarity = 0.8;
    char alt = 'n';
    bool soft = false;
    Args a(argc, argv);
    if (a.has("-f")) freq = int_flag(a, "-f");
    if (a.has("-w")) width = int_flag(a, "-w");
    if (a.has("-l")) alt = a.get("-l")[0];  // '\0' when the flag has no value, as in the reference
    if (a.has("-d")) soft = true;
    if (a.has("-h") || a.has("--help") || !(a.has("-i") && a.has("-o"))) {
        mask_usage(freq, width, alt);
        return 1;
    }
    const std::string input = a.get("-i"), output = a.get("-o");
    std::string dump;
    if (a.has("-k")) dump = a.get("-k");
    else freq = -1;

    std::map<std::string, std::string> genome;
    std::vector<std::string> order;
    read_plain_fasta(input, genome, order);
    std::vector<std::string> names;
    std::vector<int64_t> sizes;
    std::map<std::string, int32_t> index;
    for (const auto &kv : genome) {
        index[kv.first] = (int32_t)names.size();
        names.push_back(kv.first);
        sizes.push_back((int64_t)kv.second.size());
    }
    CdsMask cds(names, sizes);
    bool use_cds = false;
    if (a.has("-g")) {
        std::cout << "masking genome using GFF file" << std::endl;
        cds.gff(a.get("-g"), a.has("-z"));
        use_cds = true;
    }
    if (a.has("-s") && a.has("-c")) {
        std::cout << "masking genome using sam file" << std::endl;
        const std::string sam = a.get("-s");
        if (a.has("-p")) {
            try {
                similarity = std::stod(a.get("-p"));
            } catch (const std::exception &) {
                throw Error(DCNS_ERR_ARG, "option -p: not a number");
            }
        }
        std::map<std::string, std::string> cds_seqs;
        std::vector<std::string> cds_names;
        read_plain_fasta(a.get("-c"), cds_seqs, cds_names);
        std::map<std::string, int64_t> cds_sizes;
        for (const auto &kv : cds_seqs) cds_sizes[kv.first] = (int64_t)kv.second.size();
        cds.sam(sam, cds_sizes, similarity, &std::cout);
        use_cds = true;
    }
    std::cout << "masking genome using kmer" << std::endl;
    const KmerSet ks = read_kmer_dump(dump, freq);
    if (width == 0) throw Error(DCNS_ERR_ARG, "line width 0 (the reference divides by it)");

    // the genome back to back, in name order
    std::vector<int64_t> off(1, 0);
    for (int64_t s : sizes) off.push_back(off.back() + s);
    std::string all;
    all.reserve((size_t)off.back());
    for (const auto &kv : genome) all += kv.second;
    genome.clear();
    std::vector<int64_t> boff(1, 0);
    for (int64_t s : sizes) boff.push_back(boff.back() + mask_body_size(s, width));
    std::vector<char> bodies((size_t)boff.back());

    dcns_ctx *c = nullptr;
    if (dcns_ctx_create(0, 0, &c) != DCNS_OK) throw Error(DCNS_ERR_DEVICE, dcns_last_error(nullptr));
    std::unique_ptr<dcns_ctx, int (*)(dcns_ctx *)> ctx(c, &dcns_ctx_destroy);
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    mask_stage(c, ks, (const uint8_t *)all.data(), off.data(), (int32_t)names.size());
    mask_run(c, use_cds ? &cds : nullptr, (uint8_t)alt, soft, width, nullptr, (uint8_t *)bodies.data());

    std::FILE *f = std::fopen(output.c_str(), "wb");
    if (!f) throw Error(DCNS_ERR_ARG, "error in opening output file " + output);
    for (const std::string &chr : order) {
        const int32_t s = index.at(chr);
        std::fputc('>', f);
        std::fwrite(chr.data(), 1, chr.size(), f);
        std::fputc('\n', f);
        std::fwrite(bodies.data() + boff[(size_t)s], 1, (size_t)(boff[(size_t)s + 1] - boff[(size_t)s]), f);
    }
    const bool bad = std::ferror(f) != 0;
    if (std::fclose(f) != 0 || bad) throw Error(DCNS_ERR_ARG, "error in writing " + output);
    return 0;
}

}  // namespace

int mask_genome_main(int argc, char **argv) {
    try {
        return mask_genome(argc, argv);
    } catch (const std::exception &e) {
        std::cout.flush();
        std::cerr << "dCNS maskGenome: " << e.what() << std::endl;
        return 1;
    }
}

}  // namespace cli
}  // namespace dcns

extern "C" int dcns_mask_genome_main(int argc, char **argv) { return dcns::cli::mask_genome_main(argc, argv); }
