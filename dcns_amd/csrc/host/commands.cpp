// commands.cpp -- the dCNS commands served by this library, with the
// reference's flags, defaults, files and output text:
//   cut1Gap  controlLayer4.cpp:1361-1529
//   ranSco   controlLayer4.cpp:2088-2149, impl/permutationLsqLambdaK.cpp:7-102
//   cut2Gaps controlLayer4.cpp:1534-1713
//   pairCns2Gaps controlLayer4.cpp:318-462
//   pairCnsXExtend controlLayer4.cpp:181-317
//   weighted1Gap controlLayer4.cpp:481-652, weighted2Gaps :655-829
//   cut2Gaps2 controlLayer4.cpp:1815-1994
// The alignment work runs on the GPU; the host keeps the reference's order
// of file operations and stdout lines.  All query regions of one cut1Gap
// input are aligned in one batched call.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <future>
#include <iostream>
#include <memory>
#include <sstream>

#include "../../../include/dcns.h"
#include "../ctx.h"
#include "../pipeline.h"
#include "cli_io.h"
#include "weights.h"

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

void two_gaps_usage(const char *cmd, const Defaults &d, bool genomes) {
    std::cerr << "Usage: dCNS " << cmd << " -i input -r reference -o output"
              << (genomes ? " -ra refGenome.fa -qa queryGenome.fa" : "") << "\n"
              << "Options\n"
              << " -i FILE    input file in fasta format\n"
              << " -r STRING  reference sequence entry name in input fasta file\n";
    if (genomes)
        std::cerr << " -ra STRING masked reference genome sequence in fasta format\n"
                  << " -qa STRING masked query genome sequence in fasta format\n";
    std::cerr << " -o FILE    output prefix\n"
              << " -A INT     matching score (default: " << d.match << ")\n"
              << " -B INT     mismatching penalty (default: " << d.mismatch << ")\n"
              << " -O1 INT    open gap penalty (default: " << d.open1 << ")\n"
              << " -E1 INT    extend gap penalty (default: " << d.ext1 << ")\n"
              << " -u INT     x-drop extension band width (default: " << d.w << ")\n"
              << " -O2 INT    second-piece open gap penalty (default: " << d.open2 << ")\n"
              << " -E2 INT    second-piece extend gap penalty (default: " << d.ext2 << ")\n"
              << " -u2 INT    z-drop alignment band width (default: " << d.bandwidth << ")\n"
              << " -w INT     seed window size (default: " << d.window << ")\n"
              << " -c INT     minimum seed score (default: " << d.min_score << ")\n"
              << " -s INT     seed window step (default: " << d.step << ")\n"
              << " -l DOUBLE  lambda for p-values (default: " << d.lambda << ")\n"
              << " -k DOUBLE  K for p-values (default: " << d.k << ")\n"
              << " -p DOUBLE  p-value threshold (default: " << d.pvalues << ")\n"
              << std::endl;
}

dcns_params to_dcns_params(const Defaults &d) {
    dcns_params p;
    dcns_default_params(&p);
    p.match = d.match;
    p.mismatch = d.mismatch;
    p.open = d.open1;
    p.ext = d.ext1;
    p.open2 = d.open2;
    p.ext2 = d.ext2;
    p.zdrop = d.zdrop;
    p.bandwidth = d.bandwidth;
    p.window = d.window;
    p.step = d.step;
    p.min_score = d.min_score;
    p.w = d.w;
    p.xdrop = d.xdrop;
    p.pvalue = d.pvalues;
    p.lambda = d.lambda;
    p.k = d.k;
    return p;
}

}  // namespace

namespace {

// The flags of cut1Gap (controlLayer4.cpp:1361-1400) into d; false when the
// usage text is to be printed instead.
bool cut1gap_flags(const Args &a, Defaults &d) {
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
    if (a.has("-h") || a.has("--help")) return false;
    return a.has("-i") && a.has("-r") && a.has("-o") && a.has("-ra") && a.has("-qa");
}

// One cut1Gap input file: the reference region and the query regions, read
// from the genomes and stripped of their masked ('n') bases
// (controlLayer4.cpp:1401-1410).
struct Cut1Job {
    std::string output, error;
    Meta ref;
    std::string seq1_0;
    std::vector<int8_t> c1;
    struct Q {
        std::string name, seq2_0;
        Meta meta;
        std::vector<int8_t> codes;
        bool run = false;
    };
    std::vector<Q> qs;
    int64_t pairs() const {
        int64_t n = 0;
        for (const Q &q : qs) n += q.run;
        return n;
    }
};

Cut1Job cut1gap_job(const std::string &input, const std::string &reference, const std::string &output,
                    IndexedFasta &refFa, IndexedFasta &queFa) {
    Cut1Job j;
    j.output = output;
    // each line starts with clean genome streams: a fetch that ran past the
    // end of a genome file (a truncated FASTA, a stale index) affects only
    // the rest of its own line, as it does in the reference's one process per line
    refFa.clear();
    queFa.clear();
    if (!std::ifstream(input).good()) {
        j.error = "error in opening fasta file " + input;
        return j;
    }
    std::map<std::string, Meta> meta;
    std::map<std::string, std::string> seqs;
    std::vector<std::string> names;
    read_meta_fasta(input, seqs, meta, names);
    j.ref = meta[reference];
    j.seq1_0 = refFa.fetch(j.ref.chr, j.ref.start, j.ref.end, (int8_t)j.ref.strand);
    std::string seq1 = j.seq1_0;
    seq1.erase(std::remove(seq1.begin(), seq1.end(), 'n'), seq1.end());
    j.c1 = encode(seq1);
    for (const std::string &q : names) {
        if (q.compare(reference) == 0) continue;
        Cut1Job::Q x;
        x.name = q;
        x.meta = meta[q];
        x.run = seqs[q].size() > 0;
        if (x.run) {
            x.seq2_0 = queFa.fetch(x.meta.chr, x.meta.start, x.meta.end, (int8_t)x.meta.strand);
            std::string s2 = x.seq2_0;
            s2.erase(std::remove(s2.begin(), s2.end(), 'n'), s2.end());
            x.codes = encode(s2);
        }
        j.qs.push_back(std::move(x));
    }
    return j;
}

// All pairs of the jobs in one batch through overload 1.
FragSet cut1gap_align(dcns_ctx *ctx, const dcns_params &p, const std::vector<Cut1Job> &jobs) {
    std::vector<int8_t> buf;
    std::vector<int64_t> off1, off2;
    std::vector<int32_t> len1, len2;
    for (const Cut1Job &j : jobs) {
        if (!j.error.empty() || j.pairs() == 0) continue;
        const int64_t o1 = (int64_t)buf.size();
        buf.insert(buf.end(), j.c1.begin(), j.c1.end());
        for (const Cut1Job::Q &q : j.qs) {
            if (!q.run) continue;
            off1.push_back(o1);
            len1.push_back((int32_t)j.c1.size());
            off2.push_back((int64_t)buf.size());
            len2.push_back((int32_t)q.codes.size());
            buf.insert(buf.end(), q.codes.begin(), q.codes.end());
        }
    }
    FragSet fs;
    if (!off1.empty()) {
        dcns_batch b{buf.data(), off1.data(), len1.data(), off2.data(), len2.data(), (int32_t)off1.size(), 0};
        find_similar(ctx, b, p, fs);
    }
    return fs;
}

// Per job the file operations and stdout lines of controlLayer4.cpp:1411-1527,
// in job order.
void cut1gap_write(const std::vector<Cut1Job> &jobs, const FragSet &fs) {
    size_t k = 0;
    for (const Cut1Job &j : jobs) {
        if (!j.error.empty()) {
            std::cerr << j.error << std::endl;
            continue;
        }
        for (const Cut1Job::Q &q : j.qs) {
            // the reference opens the files three times per query (created,
            // appended, "#done" appended); one opening each gives the same bytes
            const std::string prefix = j.output + q.meta.species;
            std::ofstream sam(prefix + ".sam", std::ios_base::app);
            std::ofstream o(prefix + "o", std::ios_base::app);
            if (q.run) {
                std::cout << "seeds size" << fs.seeds_size[k] << std::endl;
                std::vector<PairedSimilarFragment> fr(fs.frags.begin() + fs.first[k],
                                                      fs.frags.begin() + fs.first[k] + fs.count[k]);
                write_sam_and_o(sam, o, fr, j.ref.strand, q.meta.strand, j.ref.start, q.meta.start, j.ref.chr,
                                q.meta.chr, j.seq1_0, q.seq2_0);
                ++k;
            }
            o << "#done\n";
        }
    }
}

}  // namespace

int cut1gap_main(int argc, char **argv) {
    Defaults d;
    Args a(argc, argv);
    if (!cut1gap_flags(a, d)) {
        cut1gap_usage(d);
        return 1;
    }
    if (d.syntenic) {
        std::cerr << "dCNS: syntenic filtering (-y) is outside this build's scope" << std::endl;
        return 1;
    }
    IndexedFasta refFa(a.get("-ra")), queFa(a.get("-qa"));
    std::vector<Cut1Job> jobs;
    jobs.push_back(cut1gap_job(a.get("-i"), a.get("-r"), a.get("-o"), refFa, queFa));
    if (!jobs[0].error.empty()) {
        std::cerr << jobs[0].error << std::endl;
        return 1;
    }
    dcns_ctx *ctx = make_ctx();
    try {
        cut1gap_write(jobs, cut1gap_align(ctx, to_dcns_params(d), jobs));
    } catch (const std::exception &e) {
        std::cerr << "dCNS: " << e.what() << std::endl;
        dcns_ctx_destroy(ctx);
        return 1;
    }
    dcns_ctx_destroy(ctx);
    return 0;
}

int dispatch(int argc, char **argv);

// batch: the command file of a GNU-parallel fan-out (README.md:117-122: one
// "dCNS cut1Gap ... -i <region> -o <prefix>" line per region) run in one
// process.  Consecutive cut1Gap lines that differ only in -i / -o form a
// group: its regions are read on a second thread, chunk by chunk (about
// max_pairs pairs), while the device aligns the previous chunk and a third
// thread writes the chunk before; every chunk is one batched call.  Files and stdout are those of running the lines
// one after another.  Lines of other commands run through their own entry
// points, in order.
int batch_main(int argc, char **argv) {
    if (argc < 2) {
        std::cerr << "Usage: dCNS batch commands.txt [-n pairs_per_chunk]\n"
                  << "  commands.txt: one dCNS command line per line (e.g. dCNS cut1Gap ... -i region.fa -o out)\n";
        return 1;
    }
    int64_t max_pairs = 2048;  // measured best of 1024-8192 on 2-12 kb regions
    for (int i = 2; i + 1 < argc; ++i)
        if (!std::strcmp(argv[i], "-n")) max_pairs = std::max<int64_t>(1, std::atoll(argv[i + 1]));
    std::ifstream in(argv[1]);
    if (!in.good()) {
        std::cerr << "error in opening command file " << argv[1] << std::endl;
        return 1;
    }
    std::vector<std::vector<std::string>> lines;
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ss(line);
        std::vector<std::string> t;
        std::string w;
        while (ss >> w) t.push_back(w);
        if (t.empty() || t[0][0] == '#') continue;
        if (t[0] == "dCNS" || (t[0].size() >= 5 && t[0].compare(t[0].size() - 5, 5, "/dCNS") == 0))
            t.erase(t.begin());  // the program name
        if (!t.empty()) lines.push_back(t);
    }
    // a cut1Gap line's group: every token but the -i / -o pairs
    auto group_key = [](const std::vector<std::string> &t) {
        std::string key;
        for (size_t i = 0; i < t.size(); ++i) {
            if ((t[i] == "-i" || t[i] == "-o") && i + 1 < t.size()) {
                ++i;
                continue;
            }
            key += t[i] + '\x1f';
        }
        return key;
    };
    int status = 0;
    dcns_ctx *ctx = nullptr;
    for (size_t li = 0; li < lines.size();) {
        const std::vector<std::string> &t = lines[li];
        std::vector<char *> av;
        for (const std::string &w : t) av.push_back(const_cast<char *>(w.c_str()));
        if (t[0] != "cut1Gap") {
            std::cout.flush();
            const int rc = dcns::cli::dispatch((int)av.size(), av.data());
            status = status ? status : rc;
            ++li;
            continue;
        }
        size_t le = li + 1;
        const std::string key = group_key(t);
        while (le < lines.size() && lines[le][0] == "cut1Gap" && group_key(lines[le]) == key) ++le;
        Defaults d;
        Args a((int)av.size(), av.data());
        bool ok = false;
        try {
            ok = cut1gap_flags(a, d) && !d.syntenic;
        } catch (const std::exception &) {
            std::cerr << "dCNS: invalid number in: " << key << std::endl;
            status = 1;
            li = le;
            continue;
        }
        if (!ok) {
            cut1gap_main((int)av.size(), av.data());  // its usage / refusal text, once per line
            status = 1;
            li = le;
            continue;
        }
        // an unreadable genome index fails this group's lines only (the index
        // reader ends the process, as the reference's does for its one line)
        bool genomes_ok = true;
        for (const std::string g : {a.get("-ra"), a.get("-qa")})
            if (!std::ifstream(g + ".fai").good() || !std::ifstream(g).good()) {
                std::cerr << "error in opening fasta index file " << g << ".fai." << std::endl
                          << " Please index your fasta file using samtools samtools faidx command." << std::endl;
                genomes_ok = false;
            }
        if (!genomes_ok) {
            status = 1;
            li = le;
            continue;
        }
        if (!ctx) ctx = make_ctx();
        const dcns_params p = to_dcns_params(d);
        IndexedFasta refFa(a.get("-ra")), queFa(a.get("-qa"));
        const std::string reference = a.get("-r");
        size_t cursor = li;
        auto next_chunk = [&]() {
            std::vector<Cut1Job> jobs;
            int64_t pairs = 0;
            while (cursor < le && (jobs.empty() || pairs < max_pairs)) {
                const std::vector<std::string> &u = lines[cursor++];
                std::vector<char *> uv;
                for (const std::string &w : u) uv.push_back(const_cast<char *>(w.c_str()));
                Args ua((int)uv.size(), uv.data());
                // a line whose input cannot be read (a malformed region header)
                // fails alone, as its own process would
                try {
                    jobs.push_back(cut1gap_job(ua.get("-i"), reference, ua.get("-o"), refFa, queFa));
                } catch (const std::exception &e) {
                    Cut1Job bad;
                    bad.output = ua.get("-o");
                    bad.error = "dCNS: " + ua.get("-i") + ": " + e.what();
                    jobs.push_back(std::move(bad));
                }
                pairs += jobs.back().pairs();
            }
            return jobs;
        };
        // three stages at once: the next chunk is read while this one aligns
        // and the previous one's files are written; writes stay in chunk order
        std::future<std::vector<Cut1Job>> ahead = std::async(std::launch::async, next_chunk);
        std::future<void> writing;
        auto finish_writing = [&]() {
            if (!writing.valid()) return;
            try {
                writing.get();
            } catch (const std::exception &e) {
                std::cerr << "dCNS: " << e.what() << std::endl;
                status = 1;
            }
        };
        for (;;) {
            std::vector<Cut1Job> jobs = ahead.get();
            if (jobs.empty()) break;
            ahead = std::async(std::launch::async, next_chunk);
            for (const Cut1Job &j : jobs) status = (status || !j.error.empty()) ? 1 : 0;
            FragSet fs;
            bool aligned = false;
            try {
                fs = cut1gap_align(ctx, p, jobs);
                aligned = true;
            } catch (const std::exception &e) {
                finish_writing();
                std::cerr << "dCNS: " << e.what() << std::endl;
                status = 1;
            }
            finish_writing();
            if (aligned)
                writing = std::async(std::launch::async, [jobs = std::move(jobs), fs = std::move(fs)]() {
                    cut1gap_write(jobs, fs);
                });
        }
        finish_writing();
        li = le;
    }
    if (ctx) dcns_ctx_destroy(ctx);
    return status;
}

// cut2Gaps (controlLayer4.cpp:1534-1713): whole genomes read without an
// index, every query region aligned by the two-piece path (overload 4), SAM /
// o output without "#done" lines.  All query regions run as one batch.
int cut2gaps_main(int argc, char **argv) {
    Defaults d;
    Args a(argc, argv);
    if (a.has("-A")) d.match = std::stoi(a.get("-A"));
    if (a.has("-B")) d.mismatch = std::stoi(a.get("-B"));
    if (a.has("-O1")) d.open1 = std::stoi(a.get("-O1"));
    if (a.has("-E1")) d.ext1 = std::stoi(a.get("-E1"));
    if (a.has("-O2")) d.open2 = std::stoi(a.get("-O2"));
    if (a.has("-E2")) d.ext2 = std::stoi(a.get("-E2"));
    if (a.has("-u")) d.w = std::stoi(a.get("-u"));
    if (a.has("-u2")) d.bandwidth = std::stoi(a.get("-u2"));
    if (a.has("-y")) d.syntenic = str2bool(a.get("-y"), d.syntenic);
    if (a.has("-w")) d.window = std::stoi(a.get("-w"));
    if (a.has("-c")) d.min_score = std::stoi(a.get("-c"));
    if (a.has("-s")) d.step = std::stoi(a.get("-s"));
    if (a.has("-b")) d.boundary = std::stoi(a.get("-b"));
    if (a.has("-l")) d.lambda = std::stod(a.get("-l"));
    if (a.has("-k")) d.k = std::stod(a.get("-k"));
    if (a.has("-p")) d.pvalues = std::stod(a.get("-p"));
    if (a.has("-h") || a.has("--help") ||
        !(a.has("-i") && a.has("-r") && a.has("-o") && a.has("-ra") && a.has("-qa"))) {
        two_gaps_usage("cut2Gaps", d, true);
        return 1;
    }
    if (d.syntenic) {
        std::cerr << "dCNS: syntenic filtering (-y) is outside this build's scope" << std::endl;
        return 1;
    }
    const std::string input = a.get("-i"), reference = a.get("-r"), output = a.get("-o");
    std::map<std::string, std::string> refSeqs, queSeqs, seqs;
    std::vector<std::string> refNames, queNames, names;
    read_plain_fasta(a.get("-ra"), refSeqs, refNames);
    read_plain_fasta(a.get("-qa"), queSeqs, queNames);
    std::map<std::string, Meta> meta;
    read_meta_fasta(input, seqs, meta, names);
    const Meta refMeta = meta[reference];
    std::string seq1_0 = sub_sequence(refSeqs[refMeta.chr], refMeta.start, refMeta.end, refMeta.strand);
    std::string seq1 = seq1_0;
    seq1.erase(std::remove(seq1.begin(), seq1.end(), 'n'), seq1.end());
    const std::vector<int8_t> c1 = encode(seq1);
    refSeqs.clear();
    struct Q {
        std::string name, seq2_0;
        std::vector<int8_t> codes;
    };
    std::vector<Q> qs;
    std::vector<int8_t> buf(c1.begin(), c1.end());
    std::vector<int64_t> off1, off2;
    std::vector<int32_t> len1, len2;
    for (const std::string &q : names) {
        if (q.compare(reference) == 0) continue;
        Q x;
        x.name = q;
        const Meta &m = meta[q];
        x.seq2_0 = sub_sequence(queSeqs[m.chr], m.start, m.end, m.strand);
        std::string s2 = x.seq2_0;
        s2.erase(std::remove(s2.begin(), s2.end(), 'n'), s2.end());
        x.codes = encode(s2);
        off1.push_back(0);
        len1.push_back((int32_t)c1.size());
        off2.push_back((int64_t)buf.size());
        len2.push_back((int32_t)x.codes.size());
        buf.insert(buf.end(), x.codes.begin(), x.codes.end());
        qs.push_back(std::move(x));
    }
    FragSet fs;
    if (!qs.empty()) {
        dcns_params p = to_dcns_params(d);
        dcns_batch b{buf.data(), off1.data(), len1.data(), off2.data(), len2.data(), (int32_t)qs.size(), 0};
        dcns_ctx *ctx = make_ctx();
        try {
            find_similar_2gaps(ctx, b, p, fs);
        } catch (const std::exception &e) {
            std::cerr << "dCNS: " << e.what() << std::endl;
            dcns_ctx_destroy(ctx);
            return 1;
        }
        dcns_ctx_destroy(ctx);
    }
    for (size_t i = 0; i < qs.size(); ++i) {
        const Meta &m = meta[qs[i].name];
        std::vector<PairedSimilarFragment> fr(fs.frags.begin() + fs.first[i], fs.frags.begin() + fs.first[i] + fs.count[i]);
        write_sam_and_o(output + m.species, fr, refMeta.strand, m.strand, refMeta.start, m.start, refMeta.chr, m.chr,
                        seq1_0, qs[i].seq2_0);
    }
    return 0;
}

// The pairCns commands: reference and query sequences inline in one FASTA
// file with region metadata, outputNoMasking.  pairCnsXExtend
// (controlLayer4.cpp:181-317) runs overload 1 and prints each sequence;
// pairCns2Gaps (controlLayer4.cpp:318-462) runs the two-piece overload 4.
//
// The reference's pairCns2Gaps clears its sequence map after taking the
// reference (controlLayer4.cpp:346), so every query it aligns is empty, and
// getAllExtendSeed then scores a window of negative width (sp = 8 > L2 = 0),
// whose SmithWaterman allocation throws std::bad_alloc: the reference command
// aborts on every input that has a query.  Here the queries are the
// sequences of the input file, as in pairCnsXExtend, whose map is not
// cleared (controlLayer4.cpp:201).
namespace {
int paircns_common(int argc, char **argv, bool two_piece) {
    Defaults d;
    Args a(argc, argv);
    if (a.has("-A")) d.match = std::stoi(a.get("-A"));
    if (a.has("-B")) d.mismatch = std::stoi(a.get("-B"));
    if (two_piece) {
        if (a.has("-O1")) d.open1 = std::stoi(a.get("-O1"));
        if (a.has("-E1")) d.ext1 = std::stoi(a.get("-E1"));
        if (a.has("-O2")) d.open2 = std::stoi(a.get("-O2"));
        if (a.has("-E2")) d.ext2 = std::stoi(a.get("-E2"));
    } else {
        if (a.has("-O")) d.open1 = std::stoi(a.get("-O"));
        if (a.has("-E")) d.ext1 = std::stoi(a.get("-E"));
    }
    if (a.has("-u")) d.w = std::stoi(a.get("-u"));
    if (two_piece && a.has("-u2")) d.bandwidth = std::stoi(a.get("-u2"));
    if (a.has("-y")) d.syntenic = str2bool(a.get("-y"), d.syntenic);
    if (a.has("-w")) d.window = std::stoi(a.get("-w"));
    if (a.has("-c")) d.min_score = std::stoi(a.get("-c"));
    if (a.has("-s")) d.step = std::stoi(a.get("-s"));
    if (a.has("-b")) d.boundary = std::stoi(a.get("-b"));
    if (a.has("-l")) d.lambda = std::stod(a.get("-l"));
    if (a.has("-k")) d.k = std::stod(a.get("-k"));
    if (a.has("-p")) d.pvalues = std::stod(a.get("-p"));
    if (a.has("-h") || a.has("--help") || !(a.has("-i") && a.has("-r") && a.has("-o"))) {
        two_gaps_usage(two_piece ? "pairCns2Gaps" : "pairCnsXExtend", d, false);
        return 1;
    }
    if (d.syntenic) {
        std::cerr << "dCNS: syntenic filtering (-y) is outside this build's scope" << std::endl;
        return 1;
    }
    const std::string input = a.get("-i"), reference = a.get("-r"), output = a.get("-o");
    std::map<std::string, std::string> seqs;
    std::map<std::string, Meta> meta;
    std::vector<std::string> names;
    read_meta_fasta(input, seqs, meta, names);
    const std::string seq1 = seqs[reference];
    const std::vector<int8_t> c1 = encode(seq1);
    std::vector<std::string> qn;
    std::vector<int8_t> buf(c1.begin(), c1.end());
    std::vector<int64_t> off1, off2;
    std::vector<int32_t> len1, len2;
    for (const std::string &q : names) {
        if (q.compare(reference) == 0) continue;
        const std::vector<int8_t> c2 = encode(seqs[q]);
        qn.push_back(q);
        off1.push_back(0);
        len1.push_back((int32_t)c1.size());
        off2.push_back((int64_t)buf.size());
        len2.push_back((int32_t)c2.size());
        buf.insert(buf.end(), c2.begin(), c2.end());
    }
    FragSet fs;
    if (!qn.empty()) {
        dcns_params p = to_dcns_params(d);
        dcns_batch b{buf.data(), off1.data(), len1.data(), off2.data(), len2.data(), (int32_t)qn.size(), 0};
        dcns_ctx *ctx = make_ctx();
        try {
            if (two_piece)
                find_similar_2gaps(ctx, b, p, fs);
            else
                find_similar(ctx, b, p, fs);
        } catch (const std::exception &e) {
            std::cerr << "dCNS: " << e.what() << std::endl;
            dcns_ctx_destroy(ctx);
            return 1;
        }
        dcns_ctx_destroy(ctx);
    }
    if (!two_piece) std::cout << "reference:" << seq1 << std::endl;
    const Meta refMeta = meta[reference];
    for (size_t i = 0; i < qn.size(); ++i) {
        const Meta &m = meta[qn[i]];
        if (!two_piece) {
            std::cout << "_query:" << qn[i] << ": " << seqs[qn[i]] << std::endl;
            std::cout << "seeds size" << fs.seeds_size[i] << std::endl;
        }
        std::vector<PairedSimilarFragment> fr(fs.frags.begin() + fs.first[i], fs.frags.begin() + fs.first[i] + fs.count[i]);
        write_sam_and_o_unmasked(output + m.species, fr, refMeta.strand, m.strand, refMeta.start, m.start, refMeta.chr,
                                 m.chr, seq1, seqs[qn[i]]);
    }
    return 0;
}
}  // namespace

namespace {

void weighted_usage(const char *cmd, const Defaults &d, bool two_piece) {
    std::cerr << "Usage: dCNS " << cmd << " -i input -r reference -q query -o output\n"
              << "Options\n"
              << " -i FILE    input file in fasta format\n"
              << " -r STRING  reference sequence entry name in input fasta file\n"
              << " -f STRING  gff3 format reference genome annotation file\n"
              << " -ra STRING masked reference genome sequence in fasta format\n"
              << " -sp STRING the folder where score parameters located\n"
              << " -o FILE    output file, exists file will be overwrote\n"
              << " -A INT     Matching score (default: " << d.match << ")\n"
              << " -B INT     Mismatching penalty (default: " << d.mismatch << ")\n"
              << " -O1 INT    open gap penalty (default: " << d.open1 << ")\n"
              << " -E1 INT    extend gap penalty (default: " << d.ext1 << ")\n"
              << " -u INT     xextend alignment band width (default: " << d.w << ")\n";
    if (two_piece) std::cerr << " -u2 INT    z drop alignment band width (default: " << d.bandwidth << ")\n";
    std::cerr << " -w INT     seed window size (default: " << d.window << ")\n"
              << " -c INT     minimum seed score (default: " << d.min_score << ")\n"
              << " -s INT     seed window step (default: " << d.step << ")\n"
              << " -l DOUBLE  lambda for p-values (default: " << d.lambda << ")\n"
              << " -k DOUBLE  K for p-values (default: " << d.k << ")\n"
              << " -p DOUBLE  p-value threshold (default: " << d.pvalues << ")\n"
              << std::endl;
}

void stamp(const char *what) {
    const time_t t = time(nullptr);
    std::cout << ctime(&t) << what << std::endl;
}

// weighted1Gap / weighted2Gaps: the reference region (-r) against every other
// region of -i, with per-base score profiles (-sp) chosen by the GFF3 class of
// the reference base (-f over the chromosomes of -ra).
int weighted_common(int argc, char **argv, bool two_piece) {
    Defaults d;
    Args a(argc, argv);
    if (a.has("-A")) d.match = std::stoi(a.get("-A"));
    if (a.has("-B")) d.mismatch = std::stoi(a.get("-B"));
    if (a.has("-O1")) d.open1 = std::stoi(a.get("-O1"));
    if (a.has("-E1")) d.ext1 = std::stoi(a.get("-E1"));
    if (a.has("-u")) d.w = std::stoi(a.get("-u"));
    if (two_piece && a.has("-u2")) d.bandwidth = std::stoi(a.get("-u2"));
    if (a.has("-y")) d.syntenic = str2bool(a.get("-y"), d.syntenic);
    if (a.has("-w")) d.window = std::stoi(a.get("-w"));
    if (a.has("-c")) d.min_score = std::stoi(a.get("-c"));
    if (a.has("-s")) d.step = std::stoi(a.get("-s"));
    if (a.has("-b")) d.boundary = std::stoi(a.get("-b"));
    if (a.has("-l")) d.lambda = std::stod(a.get("-l"));
    if (a.has("-k")) d.k = std::stod(a.get("-k"));
    if (a.has("-p")) d.pvalues = std::stod(a.get("-p"));
    const char *cmd = two_piece ? "weighted2Gaps" : "weighted1Gap";
    if (a.has("-h") || a.has("--help") ||
        !(a.has("-i") && a.has("-r") && a.has("-o") && a.has("-ra") && a.has("-sp") && a.has("-f"))) {
        weighted_usage(cmd, d, two_piece);
        return 1;
    }
    if (d.syntenic) {
        std::cerr << "dCNS: syntenic filtering (-y) is outside this build's scope" << std::endl;
        return 1;
    }
    const std::string input = a.get("-i"), reference = a.get("-r"), output = a.get("-o");
    std::map<std::string, std::string> seqs;
    std::map<std::string, Meta> meta;
    std::vector<std::string> names;
    read_meta_fasta(input, seqs, meta, names);
    std::vector<Profile> prof;
    std::map<std::string, std::string> genome;
    std::vector<std::string> chroms;
    GffAnnotation gff;
    try {
        prof = read_score_folder(a.get("-sp"));
        read_plain_fasta(a.get("-ra"), genome, chroms);
        gff = read_gff(a.get("-f"));
    } catch (const std::exception &e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }
    // the messages of readGffFileWithEveryThing (impl/gffToCategory.cpp:176-237)
    for (const std::string &l : gff.unknown_lines)
        std::cout << "we could not analysis the line in the gff/gtf file: " << l << std::endl;
    stamp("gff file reading done");
    stamp("gff file intron remove duplication done");
    stamp("gff file exons remove duplication done");
    stamp("gff file cds remove duplication done");
    stamp("gff file transposons remove duplication done");
    stamp("update weigth introns done");
    stamp("update weigth exon done");
    stamp("update weigth cds done");
    stamp("update weigth transposons done");
    const Meta refMeta = meta[reference];
    const std::string seq1 = seqs[reference];
    auto gchr = genome.find(refMeta.chr);
    if (gchr == genome.end()) {
        std::cerr << "dCNS: chromosome " << refMeta.chr << " of the reference region is not in " << a.get("-ra")
                  << std::endl;
        return 1;
    }
    // categories of the region's bases -1 .. len (weight[chr] + start - 1, :523)
    const int64_t L1 = (int64_t)seq1.size();
    std::vector<int16_t> wcat((size_t)L1 + 2);
    gff_categories(gff, refMeta.chr, (int64_t)gchr->second.size(), (int64_t)refMeta.start - 2, L1 + 2, wcat.data());
    genome.clear();
    const std::vector<int8_t> c1 = encode(seq1);
    std::vector<std::string> qn;
    std::vector<int8_t> buf(c1.begin(), c1.end());
    std::vector<int64_t> off1, off2, cat_off;
    std::vector<int32_t> len1, len2;
    for (const std::string &q : names) {
        if (q.compare(reference) == 0) continue;
        const std::vector<int8_t> c2 = encode(seqs[q]);
        qn.push_back(q);
        off1.push_back(0);
        len1.push_back((int32_t)c1.size());
        off2.push_back((int64_t)buf.size());
        len2.push_back((int32_t)c2.size());
        cat_off.push_back(0);  // every query pairs with the same reference region
        buf.insert(buf.end(), c2.begin(), c2.end());
    }
    FragSet fs;
    if (!qn.empty()) {
        dcns_params p = to_dcns_params(d);
        dcns_batch b{buf.data(), off1.data(), len1.data(), off2.data(), len2.data(), (int32_t)qn.size(), 0};
        dcns_ctx *ctx = make_ctx();
        try {
            find_similar_weighted(ctx, b, p, prof, PairCategories{wcat.data(), cat_off.data()}, two_piece, fs);
        } catch (const std::exception &e) {
            std::cerr << "dCNS: " << e.what() << std::endl;
            dcns_ctx_destroy(ctx);
            return 1;
        }
        dcns_ctx_destroy(ctx);
    }
    for (size_t i = 0; i < qn.size(); ++i) {
        const Meta &m = meta[qn[i]];
        std::vector<PairedSimilarFragment> fr(fs.frags.begin() + fs.first[i],
                                              fs.frags.begin() + fs.first[i] + fs.count[i]);
        write_sam_and_o_unmasked(output + m.species, fr, refMeta.strand, m.strand, refMeta.start, m.start, refMeta.chr,
                                 m.chr, seq1, seqs[qn[i]]);
    }
    return 0;
}

}  // namespace

namespace {

void cut2gaps2_usage(const Defaults &d, int32_t max_len) {
    std::cerr << "Usage: dCNS cut2Gaps2 -i input -r reference -q query -o output\n"
              << "          the default parameters use about 20GB RAM\n"
              << "Options\n"
              << " -ra STRING masked reference genome sequence in fasta format\n"
              << " -qa STRING masked query genome sequence in fasta format\n"
              << " -sa STRING sam file from the cut1Gap \n"
              << " -o  FILE   output file, exists file will be overwrote\n"
              << " -A  INT    Matching score (default: " << d.match << ")\n"
              << " -B  INT    Mismatching penalty (default: " << d.mismatch << ")\n"
              << " -O1 INT    open gap penalty (default: " << d.open1 << ")\n"
              << " -E1 INT    extend gap penalty (default: " << d.ext1 << ")\n"
              << " -O2 INT    open gap penalty (default: " << d.open2 << ")\n"
              << " -E2 INT    extend gap penalty (default: " << d.ext2 << ")\n"
              << " -u2 INT    z drop alignment band width (default: " << d.bandwidth << ")\n"
              << " -m  INT    maximum aligned length (default: " << max_len << ")\n"
              << " -w  INT    seed window size (default: " << d.window << ")\n"
              << std::endl;
}

}  // namespace

// cut2Gaps2: the alignments of a cut1Gap SAM file re-extended with the
// two-piece SemiGlobal over whole chromosomes (overload 3), every chromosome
// pair of the file in one batched call.
int cut2gaps2_main(int argc, char **argv) {
    Defaults d;
    int32_t max_len = 140000;
    Args a(argc, argv);
    if (a.has("-A")) d.match = std::stoi(a.get("-A"));
    if (a.has("-B")) d.mismatch = std::stoi(a.get("-B"));
    if (a.has("-O1")) d.open1 = std::stoi(a.get("-O1"));
    if (a.has("-E1")) d.ext1 = std::stoi(a.get("-E1"));
    if (a.has("-O2")) d.open2 = std::stoi(a.get("-O2"));
    if (a.has("-E2")) d.ext2 = std::stoi(a.get("-E2"));
    if (a.has("-m")) max_len = std::stoi(a.get("-m"));
    if (a.has("-u2")) d.bandwidth = std::stoi(a.get("-u2"));
    if (a.has("-y")) d.syntenic = str2bool(a.get("-y"), d.syntenic);
    if (a.has("-w")) d.window = std::stoi(a.get("-w"));
    if (a.has("-h") || a.has("--help") || !(a.has("-sa") && a.has("-o") && a.has("-ra") && a.has("-qa"))) {
        cut2gaps2_usage(d, max_len);
        return 1;
    }
    if (d.syntenic) {
        std::cerr << "dCNS: syntenic filtering (-y) is outside this build's scope" << std::endl;
        return 1;
    }
    const std::string output = a.get("-o");
    std::map<std::string, std::string> query_genome, ref_genome;
    std::vector<std::string> qnames, rnames;
    read_plain_fasta(a.get("-qa"), query_genome, qnames);
    std::cout << "query genome reading done" << std::endl;
    read_plain_fasta(a.get("-ra"), ref_genome, rnames);
    std::cout << "reference genome reading done" << std::endl;
    SeedMap pos, neg;
    sam_to_seeds(a.get("-sa"), ref_genome, query_genome, pos, neg);
    std::cout << "sam file reading done" << std::endl;
    { std::ofstream(output + ".sam", std::ios_base::app); std::ofstream(output + "o", std::ios_base::app); }
    // every (reference, query) chromosome pair: plus strand first, then minus
    // (the query's reverse complement), in map order; 'n' stripped
    struct Job {
        std::string ref_chr, query_chr, s1_0, s2_0;
        const std::vector<Seed> *seeds;
    };
    std::vector<Job> jobs;
    for (int strand = 0; strand < 2; ++strand)
        for (auto &rc : (strand == 0 ? pos : neg))
            for (auto &qc : rc.second) {
                Job j;
                j.ref_chr = rc.first;
                j.query_chr = qc.first;
                j.s1_0 = ref_genome[rc.first];
                j.s2_0 = strand == 0 ? query_genome[qc.first] : reverse_complement(query_genome[qc.first]);
                j.seeds = &qc.second;
                jobs.push_back(std::move(j));
            }
    auto strip = [](std::string s) {
        s.erase(std::remove(s.begin(), s.end(), 'n'), s.end());
        return s;
    };
    std::vector<int8_t> buf;
    std::vector<int64_t> off1, off2;
    std::vector<int32_t> len1, len2;
    std::vector<std::vector<Seed>> seeds;
    std::map<std::string, std::pair<int64_t, int32_t>> ref_at;  // one copy of each reference chromosome
    for (const Job &j : jobs) {
        auto it = ref_at.find(j.ref_chr);
        if (it == ref_at.end()) {
            const std::vector<int8_t> c1 = encode(strip(j.s1_0));
            it = ref_at.emplace(j.ref_chr, std::make_pair((int64_t)buf.size(), (int32_t)c1.size())).first;
            buf.insert(buf.end(), c1.begin(), c1.end());
        }
        const std::vector<int8_t> c2 = encode(strip(j.s2_0));
        off1.push_back(it->second.first);
        len1.push_back(it->second.second);
        off2.push_back((int64_t)buf.size());
        len2.push_back((int32_t)c2.size());
        buf.insert(buf.end(), c2.begin(), c2.end());
        seeds.push_back(*j.seeds);
    }
    ref_genome.clear();
    query_genome.clear();
    FragSet fs;
    if (!jobs.empty()) {
        dcns_params p = to_dcns_params(d);
        dcns_batch b{buf.data(), off1.data(), len1.data(), off2.data(), len2.data(), (int32_t)jobs.size(), 0};
        dcns_ctx *ctx = make_ctx();
        try {
            find_similar_seeded(ctx, b, p, seeds, max_len, fs);
        } catch (const std::exception &e) {
            std::cerr << "dCNS: " << e.what() << std::endl;
            dcns_ctx_destroy(ctx);
            return 1;
        }
        dcns_ctx_destroy(ctx);
    }
    for (size_t i = 0; i < jobs.size(); ++i) {
        std::vector<PairedSimilarFragment> fr(fs.frags.begin() + fs.first[i], fs.frags.begin() + fs.first[i] + fs.count[i]);
        write_sam_and_o(output, fr, 1, 1, 1, 1, jobs[i].ref_chr, jobs[i].query_chr, jobs[i].s1_0, jobs[i].s2_0);
    }
    std::ofstream(output + "o", std::ios_base::app) << "#done" << std::endl;
    return 0;
}

int weighted1gap_main(int argc, char **argv) { return weighted_common(argc, argv, false); }
int weighted2gaps_main(int argc, char **argv) { return weighted_common(argc, argv, true); }
int paircns2gaps_main(int argc, char **argv) { return paircns_common(argc, argv, true); }
int paircnsxextend_main(int argc, char **argv) { return paircns_common(argc, argv, false); }

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

// One command line (argv[0] = the command) through its entry point.
int dispatch(int argc, char **argv) {
    const std::string c = argv[0];
    if (c == "cut1Gap") return dcns_cut1gap_main(argc, argv);
    if (c == "ranSco") return dcns_ransco_main(argc, argv);
    if (c == "cut2Gaps") return dcns_cut2gaps_main(argc, argv);
    if (c == "cut2Gaps2") return dcns_cut2gaps2_main(argc, argv);
    if (c == "pairCns2Gaps") return dcns_paircns2gaps_main(argc, argv);
    if (c == "pairCnsXExtend") return dcns_paircnsxextend_main(argc, argv);
    if (c == "weighted1Gap") return dcns_weighted1gap_main(argc, argv);
    if (c == "weighted2Gaps") return dcns_weighted2gaps_main(argc, argv);
    if (c == "maskGenome") return dcns_mask_genome_main(argc, argv);
    std::cerr << "dCNS: command " << c << " is not part of this build" << std::endl;
    return 1;
}

}  // namespace cli
}  // namespace dcns

namespace {
// No exception crosses the C boundary: a command that throws (a malformed
// number in a header or a flag, a failed allocation) reports it and fails.
template <class F>
int cli_guard(F &&f, int argc, char **argv) {
    try {
        return f(argc, argv);
    } catch (const std::exception &e) {
        std::cout.flush();
        std::cerr << "dCNS: " << (argc > 0 ? argv[0] : "") << ": " << e.what() << std::endl;
    } catch (...) {
        std::cout.flush();
        std::cerr << "dCNS: " << (argc > 0 ? argv[0] : "") << ": unknown error" << std::endl;
    }
    return 1;
}
}  // namespace

extern "C" int dcns_cut1gap_main(int argc, char **argv) { return cli_guard(dcns::cli::cut1gap_main, argc, argv); }
extern "C" int dcns_ransco_main(int argc, char **argv) { return cli_guard(dcns::cli::ransco_main, argc, argv); }
extern "C" int dcns_cut2gaps_main(int argc, char **argv) { return cli_guard(dcns::cli::cut2gaps_main, argc, argv); }
extern "C" int dcns_paircns2gaps_main(int argc, char **argv) {
    return cli_guard(dcns::cli::paircns2gaps_main, argc, argv);
}
extern "C" int dcns_paircnsxextend_main(int argc, char **argv) {
    return cli_guard(dcns::cli::paircnsxextend_main, argc, argv);
}
extern "C" int dcns_weighted1gap_main(int argc, char **argv) {
    return cli_guard(dcns::cli::weighted1gap_main, argc, argv);
}
extern "C" int dcns_weighted2gaps_main(int argc, char **argv) {
    return cli_guard(dcns::cli::weighted2gaps_main, argc, argv);
}
extern "C" int dcns_cut2gaps2_main(int argc, char **argv) { return cli_guard(dcns::cli::cut2gaps2_main, argc, argv); }
extern "C" int dcns_batch_main(int argc, char **argv) { return cli_guard(dcns::cli::batch_main, argc, argv); }
