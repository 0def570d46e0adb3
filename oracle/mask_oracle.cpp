// mask_oracle.cpp -- CPU restatement of dCNS's maskGenome command.
//
// TEST INFRASTRUCTURE ONLY (the checker), like the rest of liboracle.so:
// tests/, the smoke entry point and bench.py's CPU-baseline leg load it; the
// product never does.
//
// Restated, plainly and per position:
//   readFastaFile (names, sequences)   impl/fasta.cpp:7-44
//   the k-mer profile                  impl/maskGenome.cpp:15-48
//   the k-mer window scan              impl/maskGenome.cpp:50-61
//   gffToMask / gffToMaskGene          impl/getAllTheCdsPositions.cpp:9-101
//   samToMask                          impl/getAllTheCdsPositions.cpp:105-173
//   the masked FASTA writer            impl/maskGenome.cpp:63-122
//   the command's order of steps       controlLayer4.cpp:1999-2084
// Inputs that make the reference throw (a window scan over a sequence shorter
// than k, a mask range starting past its string, an unparsable number) or
// divide by zero (line width 0) return an error code here instead.
#include <cctype>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <map>
#include <regex>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "dcns_oracle.h"

namespace {

struct Failure {
    std::string what;
};

int32_t parse_int(const std::string &s) {
    try {
        return std::stoi(s);
    } catch (const std::exception &) {
        throw Failure{"not an integer: " + s};
    }
}

std::string upper(std::string s) {
    for (char &c : s)
        if (c >= 'a' && c <= 'z') c = (char)(c - 'a' + 'A');
    return s;
}

// fasta.cpp:7-44.  Header: text after '>' up to the first blank; then, when
// that text holds a tab, it is cut one character before the tab.  Only the
// last record is upper-cased (the others' toupper has no effect there) and
// only the last record is dropped when empty.
void read_fasta(const std::string &path, std::map<std::string, std::string> &seqs, std::vector<std::string> &order) {
    std::ifstream in(path);
    if (!in.good()) throw Failure{"cannot open " + path};
    std::string name, body, line;
    bool open = false;
    while (std::getline(in, line)) {
        if (!line.empty() && line[0] == '>') {
            if (open && !name.empty()) {
                seqs[name] = body;
                order.push_back(name);
            }
            const size_t sp = line.find(' ');
            const std::string head = sp == std::string::npos ? line.substr(1) : line.substr(1, sp - 1);
            const size_t tab = head.find('\t');
            if (!head.empty() && head[0] == '>')  // ">>name": the second '>' dropped, cut at the tab
                name = tab == std::string::npos ? head.substr(1) : head.substr(1, tab - 1);
            else  // cut one character before the tab; a leading tab keeps the whole text
                name = (tab == std::string::npos || tab == 0) ? head : head.substr(0, tab - 1);
            open = true;
            body.clear();
        } else {
            body += line;
        }
    }
    if (open && !name.empty()) {
        body = upper(body);
        if (!body.empty()) {
            seqs[name] = body;
            order.push_back(name);
        }
    }
}

// maskGenome.cpp:15-48 for freq > 0; returns the members and sets k.
std::set<std::string> kmer_profile(const std::string &path, int32_t freq, int32_t &k) {
    std::set<std::string> members;
    k = 10;
    if (freq <= 0) return members;
    std::ifstream in(path);
    if (!in.good()) throw Failure{"cannot open " + path};
    std::string line, body;
    int32_t count = 0;
    while (std::getline(in, line)) {
        if (!line.empty() && line[0] == '>') {
            if (count > freq) members.insert(upper(body));
            const size_t sp = line.find(' ');
            count = parse_int(sp == std::string::npos ? line.substr(1) : line.substr(1, sp - 1));
            body.clear();
        } else {
            body += line;
        }
    }
    const std::string last = upper(body);
    k = (int32_t)last.size();
    if (count > freq) members.insert(last);
    return members;
}

// maskGenome.cpp:50-61: every window of k bases equal to a member marks its k positions.
std::string kmer_marks(const std::string &seq, const std::set<std::string> &members, int32_t k) {
    std::string marks(seq.size(), '0');
    if (k == 0) return marks;
    if (seq.size() < (size_t)k) throw Failure{"sequence shorter than k"};
    for (size_t i = 0; i + k <= seq.size(); ++i)
        if (members.count(seq.substr(i, (size_t)k)))
            for (size_t j = i; j < i + (size_t)k; ++j) marks[j] = '1';
    return marks;
}

// std::string::replace(pos, n, n, '1') as the CDS functions use it.
void set_run(std::string &s, int64_t pos, int64_t n) {
    if (pos < 0 || (uint64_t)pos > s.size()) throw Failure{"mask range past the string end"};
    if ((uint64_t)(pos + n) > s.size()) s.resize((size_t)(pos + n), '0');
    for (int64_t j = pos; j < pos + n; ++j) s[(size_t)j] = '1';
}

using Masks = std::map<std::string, std::string>;

void fresh(Masks &cds, const std::map<std::string, std::string> &genome) {
    cds.clear();
    for (const auto &kv : genome) cds[kv.first] = std::string(kv.second.size(), '0');
}

std::vector<std::string> tab_fields(const std::string &line) {
    std::vector<std::string> f;
    std::string cur;
    for (char c : line) {
        if (c == '\t') {
            if (!cur.empty()) f.push_back(cur);
            cur.clear();
        } else {
            cur += c;
        }
    }
    if (!cur.empty()) f.push_back(cur);
    return f;
}

// getAllTheCdsPositions.cpp:9-56 (gene = false) and :59-101 (gene = true)
void gff_masks(const std::string &path, bool gene, const std::map<std::string, std::string> &genome, Masks &cds) {
    fresh(cds, genome);
    std::ifstream in(path);
    if (!in.good()) throw Failure{"cannot open " + path};
    std::string line;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] == '#' || line.size() < 9) continue;
        const std::vector<std::string> f = tab_fields(line);
        if (f.size() != 9) continue;
        int64_t a = parse_int(f[3]), b = parse_int(f[4]);
        if (a > b) std::swap(a, b);
        a -= 1;
        bool take;
        if (gene) take = f[2] == "gene" && f[8].find("biotype=protein_coding") != std::string::npos;
        else take = f[2] == "CDS" || f[2] == "start_codon" || f[2] == "stop_codon";
        if (take) set_run(cds[f[0]], a, b - a);
    }
}

// getAllTheCdsPositions.cpp:105-173
bool sam_masks(const std::string &path, const std::map<std::string, std::string> &genome, Masks &cds,
               double similarity, const std::map<std::string, std::string> &cds_seqs, std::string &log) {
    bool rebuild = cds.empty();
    for (const auto &kv : genome)  // a sequence without a mask gets an empty one first
        if (cds[kv.first].size() != kv.second.size()) rebuild = true;
    if (rebuild) {
        fresh(cds, genome);
        log += "GFF was not used for genome masking\n";
    }
    std::ifstream in(path);
    if (!in.good()) throw Failure{"cannot open " + path};
    const std::regex rec("^(.*?)\\s+(\\d+)\\s+(\\S+)\\s+(\\d+)\\s+(\\d+)\\s+(\\w+)\\t.*NM:i:(\\d+).*cs:(\\S+)");
    const std::regex op("(\\d+)([MIDNSH])");
    const std::regex run(":(\\d+)");
    std::string line;
    while (std::getline(in, line)) {
        if (line.size() < 9 || line[0] == '@') continue;
        std::smatch m;
        if (!std::regex_search(line, m, rec)) continue;
        double same = 0;
        const std::string cs = m[8];
        for (std::sregex_iterator r(cs.begin(), cs.end(), run), e; r != e; ++r) same += parse_int((*r)[1]);
        auto c = cds_seqs.find(m[1].str());
        const double len = c == cds_seqs.end() ? 0.0 : (double)c->second.size();
        if (!(same >= similarity * len)) continue;
        std::string &s = cds[m[3].str()];
        int64_t pos = parse_int(m[4]);
        const std::string cig = m[6];
        for (std::sregex_iterator r(cig.begin(), cig.end(), op), e; r != e; ++r) {
            const int64_t n = parse_int((*r)[1]);
            const char t = (*r)[2].str()[0];
            if (t == 'M' || t == 'D') {
                set_run(s, pos - 1, n);
                pos += n;
            } else if (t == 'N') {
                pos += n;
            }
        }
    }
    return rebuild;
}

// maskGenome.cpp:63-122
std::string write_fasta(const std::vector<std::string> &order, const std::map<std::string, std::string> &genome,
                        const std::map<std::string, std::string> &marks, const Masks &cds, char alt, bool soft,
                        int32_t width) {
    std::string out;
    for (const std::string &chr : order) {
        out += '>';
        out += chr;
        out += '\n';
        const std::string seq = upper(genome.at(chr));
        const std::string &km = marks.at(chr);
        auto c = cds.find(chr);
        for (size_t i = 0; i < seq.size(); ++i) {
            const bool masked = km[i] == '1' || (c != cds.end() && c->second[i] == '1');
            char ch = seq[i];
            if (masked) ch = soft ? (char)std::tolower((unsigned char)ch) : alt;
            out += ch;
            if (width > 0 && (int64_t)(i % (size_t)width) == width - 1) out += '\n';
        }
        out += '\n';
    }
    return out;
}

int fail(const Failure &f, char *log, int32_t logcap) {
    if (log && logcap > 0) {
        std::strncpy(log, f.what.c_str(), (size_t)logcap - 1);
        log[logcap - 1] = 0;
    }
    return -1;
}

}  // namespace

extern "C" {

int orc_mask_genome(const char *fasta, const char *out, const char *dump, int32_t freq, const char *gff,
                    int32_t gene, const char *sam, const char *cds_fasta, double similarity, char alt, int32_t soft,
                    int32_t width, char *log, int32_t logcap) {
    std::string said;
    try {
        std::map<std::string, std::string> genome;
        std::vector<std::string> order;
        read_fasta(fasta, genome, order);
        Masks cds;
        if (gff) {
            said += "masking genome using GFF file\n";
            gff_masks(gff, gene != 0, genome, cds);
        }
        if (sam && cds_fasta) {
            said += "masking genome using sam file\n";
            std::map<std::string, std::string> cs;
            std::vector<std::string> cn;
            read_fasta(cds_fasta, cs, cn);
            sam_masks(sam, genome, cds, similarity, cs, said);
        }
        said += "masking genome using kmer\n";
        int32_t k = 10;
        const std::set<std::string> members = kmer_profile(dump ? dump : "", dump ? freq : -1, k);
        std::map<std::string, std::string> marks;
        for (const auto &kv : genome) marks[kv.first] = kmer_marks(kv.second, members, k);
        if (width == 0) throw Failure{"line width 0"};
        std::ofstream o(out, std::ios::binary);
        o << write_fasta(order, genome, marks, cds, alt, soft != 0, width);
        if (!o.good()) throw Failure{"cannot write " + std::string(out)};
    } catch (const Failure &f) {
        return fail(f, log, logcap);
    }
    if (log && logcap > 0) {
        std::strncpy(log, said.c_str(), (size_t)logcap - 1);
        log[logcap - 1] = 0;
    }
    return 0;
}

int orc_kmer_mask(const char *dump, int32_t freq, const char *seq, const int64_t *offsets, int32_t nseq,
                  uint8_t *mask) {
    try {
        int32_t k = 10;
        const std::set<std::string> members = kmer_profile(dump ? dump : "", dump ? freq : -1, k);
        for (int32_t s = 0; s < nseq; ++s) {
            const std::string one(seq + offsets[s], (size_t)(offsets[s + 1] - offsets[s]));
            const std::string m = kmer_marks(one, members, k);
            for (size_t j = 0; j < m.size(); ++j) mask[offsets[s] + (int64_t)j] = m[j] == '1';
        }
        return k;
    } catch (const Failure &) {
        return -1;
    }
}

}  // extern "C"

extern "C" int orc_kmer_members(const char *dump, int32_t freq, char *out, int64_t cap, int32_t *k_out) {
    try {
        int32_t k = 10;
        const std::set<std::string> members = kmer_profile(dump ? dump : "", dump ? freq : -1, k);
        int64_t n = 0;
        for (const std::string &m : members) {
            if ((int32_t)m.size() != k) continue;
            if ((n + 1) * k <= cap) std::memcpy(out + n * k, m.data(), (size_t)k);
            ++n;
        }
        if (k_out) *k_out = k;
        return n * k <= cap ? (int)n : -2;
    } catch (const Failure &) {
        return -1;
    }
}
