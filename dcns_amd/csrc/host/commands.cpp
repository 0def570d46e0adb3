// commands.cpp -- the dCNS commands served by this library, with the
// reference's flags, defaults, files and output text:
//   cut1Gap  controlLayer4.cpp:1361-1529
//   ranSco   controlLayer4.cpp:2088-2149, impl/permutationLsqLambdaK.cpp:7-102
// The alignment work runs on the GPU; the host keeps the reference's order
// of file operations and stdout lines.  All query regions of one cut1Gap
// input are aligned in one batched call.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <memory>
#include <sstream>

#include "../../../include/dcns.h"
#include "../ctx.h"
#include "../pipeline.h"
#include "cli_io.h"

namespace dcns {
namespace cli {
namespace {

// Defaults of controlLayer4.cpp:28-52.
struct Defaults {
    int32_t window = 38, min_score = 40, step = 8, boundary = 0;
    int32_t match = 2, mismatch = -3, open1 = -4, ext1 = -2, open2 = -45, ext2 = 0;
    int32_t zdrop = 50, bandwidth = 20000;
    double lambda = 0.382291, k = 0.006662, pvalues = 0.1;
    bool syntenic = false;
    int32_t w = 10, xdrop = 20;
};

dcns_ctx *make_ctx() {
    dcns_ctx *c = nullptr;
    if (dcns_ctx_create(0, 0, &c) != DCNS_OK) {
        std::cerr << "dCNS: " << dcns_last_error(nullptr) << std::endl;
        std::exit(1);
    }
    return c;
}

void cut1gap_usage(const Defaults &d) {
    std::cerr << "Usage: dCNS cut1Gap -i input -r reference -ra refGenome.fa -qa queryGenome.fa -o output\n"
              << "Options\n"
              << " -i FILE    input file in fasta format (regions with species/chr/strand/start/end)\n"
              << " -r STRING  reference sequence entry name in input fasta file\n"
              << " -ra STRING masked reference genome sequence in fasta format (.fai index required)\n"
              << " -qa STRING masked query genome sequence in fasta format (.fai index required)\n"
              << " -o FILE    output prefix; <o><species>.sam and <o><species>o are appended to\n"
              << " -A INT     matching score (default: " << d.match << ")\n"
              << " -B INT     mismatching penalty (default: " << d.mismatch << ")\n"
              << " -O1 INT    open gap penalty (default: " << d.open1 << ")\n"
              << " -E1 INT    extend gap penalty (default: " << d.ext1 << ")\n"
              << " -u INT     x-drop extension band width (default: " << d.w << ")\n"
              << " -w INT     seed window size (default: " << d.window << ")\n"
              << " -c INT     minimum seed score (default: " << d.min_score << ")\n"
              << " -s INT     seed window step (default: " << d.step << ")\n"
              << " -l DOUBLE  lambda for p-values (default: " << d.lambda << ")\n"
              << " -k DOUBLE  K for p-values (default: " << d.k << ")\n"
              << " -p DOUBLE  p-value threshold (default: " << d.pvalues << ")\n"
              << std::endl;
}

}  // namespace

int cut1gap_main(int argc, char **argv) {
    Defaults d;
    Args a(argc, argv);
    if (a.has("-A")) d.match = std::stoi(a.get("-A"));
    if (a.has("-B")) d.mismatch = std::stoi(a.get("-B"));
    if (a.has("-O1")) d.open1 = std::stoi(a.get("-O1"));
    if (a.has("-E1")) d.ext1 = std::stoi(a.get("-E1"));
    if (a.has("-u")) d.w = std::stoi(a.get("-u"));
    if (a.has("-y")) d.syntenic = str2bool(a.get("-y"), d.syntenic);
    if (a.has("-w")) d.window = std::stoi(a.get("-w"));
    if (a.has("-c")) d.min_score = std::stoi(a.get("-c"));
    if (a.has("-s")) d.step = std::stoi(a.get("-s"));
    if (a.has("-b")) d.boundary = std::stoi(a.get("-b"));
    if (a.has("-p")) d.pvalues = std::stod(a.get("-p"));
    if (a.has("-l")) d.lambda = std::stod(a.get("-l"));
    if (a.has("-k")) d.k = std::stod(a.get("-k"));
    if (a.has("-h") || a.has("--help")) {
        cut1gap_usage(d);
        return 1;
    }
    if (!(a.has("-i") && a.has("-r") && a.has("-o") && a.has("-ra") && a.has("-qa"))) {
        cut1gap_usage(d);
        return 1;
    }
    if (d.syntenic) {
        std::cerr << "dCNS: syntenic filtering (-y) is outside this build's scope" << std::endl;
        return 1;
    }
    const std::string input = a.get("-i"), reference = a.get("-r"), output = a.get("-o");
    IndexedFasta refFa(a.get("-ra")), queFa(a.get("-qa"));
    std::map<std::string, Meta> meta;
    std::map<std::string, std::string> seqs;
    std::vector<std::string> names;
    read_meta_fasta(input, seqs, meta, names);

    const Meta refMeta = meta[reference];
    const std::string seq1_0 = refFa.fetch(refMeta.chr, refMeta.start, refMeta.end, (int8_t)refMeta.strand);
    std::string seq1 = seq1_0;
    seq1.erase(std::remove(seq1.begin(), seq1.end(), 'n'), seq1.end());
    const std::vector<int8_t> c1 = encode(seq1);

    // every query region first (the alignment runs as one batch), then the
    // file operations in the reference's order
    struct Q {
        std::string name, seq2_0;
        std::vector<int8_t> codes;
        bool run;
    };
    std::vector<Q> qs;
    for (const std::string &q : names) {
        if (q.compare(reference) == 0) continue;
        Q x;
        x.name = q;
        x.run = seqs[q].size() > 0;
        if (x.run) {
            const Meta &m = meta[q];
            x.seq2_0 = queFa.fetch(m.chr, m.start, m.end, (int8_t)m.strand);
            std::string s2 = x.seq2_0;
            s2.erase(std::remove(s2.begin(), s2.end(), 'n'), s2.end());
            x.codes = encode(s2);
        }
        qs.push_back(std::move(x));
    }
    std::vector<int8_t> buf(c1.begin(), c1.end());
    std::vector<int64_t> off1, off2;
    std::vector<int32_t> len1, len2;
    std::vector<size_t> which;
    for (size_t i = 0; i < qs.size(); ++i) {
        if (!qs[i].run) continue;
        off1.push_back(0);
        len1.push_back((int32_t)c1.size());
        off2.push_back((int64_t)buf.size());
        len2.push_back((int32_t)qs[i].codes.size());
        buf.insert(buf.end(), qs[i].codes.begin(), qs[i].codes.end());
        which.push_back(i);
    }
    FragSet fs;
    if (!which.empty()) {
        dcns_params p;
        dcns_default_params(&p);
        p.match = d.match;
        p.mismatch = d.mismatch;
        p.open = d.open1;
        p.ext = d.ext1;
        p.window = d.window;
        p.step = d.step;
        p.min_score = d.min_score;
        p.w = d.w;
        p.xdrop = d.xdrop;
        p.pvalue = d.pvalues;
        p.lambda = d.lambda;
        p.k = d.k;
        dcns_batch b{buf.data(), off1.data(), len1.data(), off2.data(), len2.data(), (int32_t)which.size(), 0};
        dcns_ctx *ctx = make_ctx();
        try {
            find_similar(ctx, b, p, fs);
        } catch (const std::exception &e) {
            std::cerr << "dCNS: " << e.what() << std::endl;
            dcns_ctx_destroy(ctx);
            return 1;
        }
        dcns_ctx_destroy(ctx);
    }
    size_t k = 0;
    for (size_t i = 0; i < qs.size(); ++i) {
        const Meta &m = meta[qs[i].name];
        const std::string prefix = output + m.species;
        { std::ofstream s(prefix + ".sam", std::ios_base::app); std::ofstream o(prefix + "o", std::ios_base::app); }
        if (qs[i].run) {
            std::cout << "seeds size" << fs.seeds_size[k] << std::endl;
            std::vector<PairedSimilarFragment> fr(fs.frags.begin() + fs.first[k],
                                                  fs.frags.begin() + fs.first[k] + fs.count[k]);
            write_sam_and_o(prefix, fr, refMeta.strand, m.strand, refMeta.start, m.start, refMeta.chr, m.chr, seq1_0,
                            qs[i].seq2_0);
            ++k;
        }
        std::ofstream o(prefix + "o", std::ios_base::app);
        o << "#done" << std::endl;
    }
    return 0;
}

// ranSco: sample random reference/query fragments with the C library's
// rand() stream exactly as permutationLsqLambdaK does, score every pair's
// maximum local alignment on the GPU in one batch, print the maxima in
// sampling order.
int ransco_main(int argc, char **argv) {
    Defaults d;
    int32_t length = 1000, times = 10000, seed = 1;
    bool removen = true;
    Args a(argc, argv);
    if (a.has("-n")) removen = str2bool(a.get("-n"), d.syntenic);
    if (a.has("-l")) length = std::stoi(a.get("-l"));
    if (a.has("-p")) times = std::stoi(a.get("-p"));
    if (a.has("-s")) seed = std::stoi(a.get("-s"));
    if (a.has("-A")) d.match = std::stoi(a.get("-A"));
    if (a.has("-B")) d.mismatch = std::stoi(a.get("-B"));
    if (a.has("-O")) d.open1 = std::stoi(a.get("-O"));
    if (a.has("-E")) d.ext1 = std::stoi(a.get("-E"));
    if (a.has("-h") || a.has("--help") || !(a.has("-r") && a.has("-i"))) {
        std::cerr << "Usage: dCNS ranSco -r reference.fa -i query1.fa [query2.fa ...]\n"
                  << " -l INT length (" << length << ")  -p INT samples (" << times << ")  -s INT seed (" << seed
                  << ")\n -n bool remove n (true)  -A/-B/-O/-E scores (" << d.match << "," << d.mismatch << ","
                  << d.open1 << "," << d.ext1 << ")" << std::endl;
        return 1;
    }
    const std::vector<std::string> queryFastas = a.many("-i");
    const std::string refFasta = a.get("-r");
    std::srand(seed);
    std::map<std::string, std::string> refSeqs;
    std::vector<std::string> refNames;
    read_plain_fasta(refFasta, refSeqs, refNames);
    std::map<std::string, std::map<std::string, std::string>> qSeqs;
    std::map<std::string, std::vector<std::string>> qNames;
    for (const std::string &qf : queryFastas) {
        std::map<std::string, std::string> s;
        std::vector<std::string> n;
        read_plain_fasta(qf, s, n);
        qSeqs[qf] = s;
        qNames[qf] = n;
    }
    if (removen) {
        for (auto &kv : refSeqs) kv.second.erase(std::remove(kv.second.begin(), kv.second.end(), 'n'), kv.second.end());
        for (const std::string &qf : queryFastas)
            for (auto &kv : qSeqs[qf]) kv.second.erase(std::remove(kv.second.begin(), kv.second.end(), 'n'), kv.second.end());
    }
    auto rnd = [](size_t range) -> size_t { return (size_t)std::rand() % (range + 1); };
    auto fragment = [&](std::map<std::string, std::string> &s, std::vector<std::string> &names) {
        std::string name = names[rnd(names.size() - 1)];
        std::string out;
        if (s[name].size() > (size_t)length) {
            const size_t start = rnd(s[name].size() - (size_t)length) + 1;
            const size_t end = start + (size_t)length - 1;
            out = sub_sequence(s[name], (int)start, (int)end);
        }
        return out;
    };
    auto bad = [&](const std::string &x) {
        return x.length() < (size_t)length || x.find('n') != std::string::npos || x.find('N') != std::string::npos ||
               x.find('-') != std::string::npos;
    };
    std::vector<int8_t> buf;
    std::vector<int64_t> off1, off2;
    std::vector<int32_t> len1, len2;
    for (int i = 0; i < times; ++i) {
        std::string r = fragment(refSeqs, refNames);
        while (bad(r)) r = fragment(refSeqs, refNames);
        auto qfrag = [&]() {
            const std::string &sp = queryFastas[rnd(queryFastas.size() - 1)];
            return fragment(qSeqs[sp], qNames[sp]);
        };
        std::string q = qfrag();
        while (bad(q)) q = qfrag();
        const std::vector<int8_t> cr = encode(r), cq = encode(q);
        off1.push_back((int64_t)buf.size());
        len1.push_back((int32_t)cr.size());
        buf.insert(buf.end(), cr.begin(), cr.end());
        off2.push_back((int64_t)buf.size());
        len2.push_back((int32_t)cq.size());
        buf.insert(buf.end(), cq.begin(), cq.end());
    }
    if (times <= 0) return 0;
    std::vector<int32_t> out((size_t)times * 3);
    dcns_batch b{buf.data(), off1.data(), len1.data(), off2.data(), len2.data(), times, 0};
    dcns_ctx *ctx = make_ctx();
    const int rc = dcns_sw_batch(ctx, &b, d.match, d.mismatch, d.open1, d.ext1, out.data());
    if (rc != DCNS_OK) {
        std::cerr << "dCNS: " << dcns_last_error(ctx) << std::endl;
        dcns_ctx_destroy(ctx);
        return 1;
    }
    dcns_ctx_destroy(ctx);
    for (int i = 0; i < times; ++i) std::cout << out[(size_t)i * 3] << std::endl;
    return 0;
}

}  // namespace cli
}  // namespace dcns

extern "C" int dcns_cut1gap_main(int argc, char **argv) { return dcns::cli::cut1gap_main(argc, argv); }
extern "C" int dcns_ransco_main(int argc, char **argv) { return dcns::cli::ransco_main(argc, argv); }
