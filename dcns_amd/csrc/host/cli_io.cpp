// cli_io.cpp -- see cli_io.h.  The text formats and their quirks are part of
// the reference's observable output and are kept as they are.
#include "cli_io.h"

#include <algorithm>
#include <memory>
#include <regex>
#include <cctype>
#include <cstdlib>
#include <iostream>
#include <regex>
#include <sstream>

namespace dcns {
namespace cli {

Args::Args(int argc, char **argv) {
    for (int i = 1; i < argc; ++i) tokens_.emplace_back(argv[i]);
}

bool Args::has(const std::string &opt) const { return std::find(tokens_.begin(), tokens_.end(), opt) != tokens_.end(); }

std::string Args::get(const std::string &opt) const {
    auto it = std::find(tokens_.begin(), tokens_.end(), opt);
    if (it != tokens_.end() && ++it != tokens_.end()) return *it;
    return "";
}

std::vector<std::string> Args::many(const std::string &opt) const {
    std::vector<std::string> out;
    auto it = std::find(tokens_.begin(), tokens_.end(), opt);
    while (it != tokens_.end() && ++it != tokens_.end() && (*it)[0] != '-' && (*it)[0] != '>') out.push_back(*it);
    return out;
}

// The reference tests std::string::compare(...) == 1, so only a handful of
// spellings flip the flag; both branches set true.
bool str2bool(std::string p, bool dflt) {
    std::transform(p.begin(), p.end(), p.begin(), ::toupper);
    for (const char *t : {"YES", "TRUE", "Y", "T", "1"})
        if (p.compare(t) == 1) return true;
    for (const char *t : {"NO", "FALSE", "F", "N", "0"})
        if (p.compare(t) == 1) return true;
    return dflt;
}

namespace {
void store_meta(const std::string &meta_line, const std::string &name, std::map<std::string, Meta> &meta) {
    static const std::regex metaRegex("species:(\\S+)\\s+chr:(\\S+)\\s+strand:(\\S+)\\s+start:(\\S+)\\s+end:(\\S+)");
    std::smatch m;
    std::regex_search(meta_line, m, metaRegex);
    if (!m.empty()) {
        Meta x;
        x.strand = std::stoi(m[3]);
        x.start = std::stoi(m[4]);
        x.end = std::stoi(m[5]);
        x.species = m[1];
        x.chr = m[2];
        meta[name] = x;
    }
}
}  // namespace

// Only the last record is upper-cased (the other records' toupper is a
// no-op expression in the reference).
void read_meta_fasta(const std::string &path, std::map<std::string, std::string> &seqs,
                     std::map<std::string, Meta> &meta, std::vector<std::string> &names) {
    std::ifstream in(path);
    if (!in.good()) {
        std::cerr << "error in opening fasta file " << path << std::endl;
        std::exit(1);
    }
    static const std::regex headRegex(">(\\S+)\\s");
    std::smatch m;
    std::string name, meta_line, line;
    std::stringstream body;
    while (std::getline(in, line)) {
        if (line[0] == '>') {
            if (!name.empty()) {
                seqs[name] = body.str();
                names.push_back(name);
                store_meta(meta_line, name, meta);
            }
            meta_line = line;
            name.clear();
            std::regex_search(line, m, headRegex);
            if (!m.empty()) name = m[1];
            body.str(std::string());
        } else {
            body << line;
        }
    }
    if (!name.empty()) {
        std::string s = body.str();
        std::transform(s.begin(), s.end(), s.begin(), ::toupper);
        seqs[name] = s;
        names.push_back(name);
        store_meta(meta_line, name, meta);
    }
}

// Header name: text after '>' up to the first space, then cut before a tab
// dropping the character just before it (impl/fasta.cpp:24-30).
void read_plain_fasta(const std::string &path, std::map<std::string, std::string> &seqs,
                      std::vector<std::string> &names) {
    std::ifstream in(path);
    if (!in.good()) {
        std::cerr << "error in opening fasta file " << path << std::endl;
        std::exit(1);
    }
    std::string name, line;
    std::stringstream body;
    while (std::getline(in, line)) {
        if (line[0] == '>') {
            if (!name.empty()) {
                seqs[name] = body.str();
                names.push_back(name);
            }
            name = line.substr(1, line.find(" ", 0) - 1);
            if (name[0] == '>') name = name.substr(1, name.find("\t", 0) - 1);
            else name = name.substr(0, name.find("\t", 0) - 1);
            body.str(std::string());
        } else {
            body << line;
        }
    }
    if (!name.empty()) {
        std::string s = body.str();
        std::transform(s.begin(), s.end(), s.begin(), ::toupper);
        if (!s.empty()) {
            seqs[name] = s;
            names.push_back(name);
        }
    }
}

IndexedFasta::IndexedFasta(const std::string &fasta) {
    const std::string fai = fasta + ".fai";
    std::ifstream idx(fai);
    if (!idx.good()) {
        std::cerr << "error in opening fasta index file " << fai << "." << std::endl
                  << " Please index your fasta file using samtools samtools faidx command." << std::endl;
        std::exit(1);
    }
    static const std::regex faiRegex("^(\\S+)\\s+(\\S+)\\s+(\\S+)\\s+(\\S+)\\s+(\\S+)");
    std::smatch m;
    std::string line;
    while (std::getline(idx, line)) {
        std::regex_search(line, m, faiRegex);
        if (!m.empty()) {
            Entry e;
            e.length = std::stoi(m[2]);
            e.offset = std::stol(m[3]);
            e.line_blen = std::stoi(m[4]);
            e.line_len = std::stoi(m[5]);
            entries_[m[1]] = e;
        }
    }
    in_.open(fasta, std::ios::in);
}

// Region fetch through the .fai (newline arithmetic on bases-per-line, one
// stream kept open across calls), then reverse complement for strand 0.
std::string IndexedFasta::fetch(const std::string &chr, int32_t start1, int32_t end1, int8_t strand) {
    int start, end;
    if (start1 < end1) {
        start = start1;
        end = end1;
    } else {
        start = end1;
        end = start1;
    }
    std::string sub;
    auto it = entries_.find(chr);
    if (it != entries_.end()) {
        const Entry e = it->second;
        start--;
        if (start < 0) start = 0;
        if (e.length < end) end = e.length;
        if (e.length < start) start = e.length;
        if (end < 0) end = 0;
        const int length = end - start;
        const int nl_before = start > 0 ? (start - 1) / e.line_blen : 0;
        const int nl_end = (start + length - 1) / e.line_blen;
        const int seqlen = length + (nl_end - nl_before);
        in_.seekg(e.offset + nl_before + start, std::ios::beg);
        char c = 'N';  // a get() that fails keeps the previous character (none before the first)
        for (int32_t i = 0; i < seqlen && !in_.eof(); ++i) {
            in_.get(c);
            sub += c;
        }
        sub.erase(std::remove_if(sub.begin(), sub.end(), ::isspace), sub.end());
    }
    if (strand == 0) return reverse_complement(sub);
    return sub;
}

// Upper-case IUPAC complement; every other character (lower case included)
// is kept as it is.
namespace {
// numberOfNInAInterval (impl/samFileToSeeds.cpp:8-16), inside the string
int32_t n_count(const std::string &s, int64_t start, int64_t end) {
    int32_t n = 0;
    const int64_t lo = std::max<int64_t>(start, 0), hi = std::min<int64_t>(end, (int64_t)s.size() - 1);
    for (int64_t i = lo; i <= hi; ++i)
        if (s[(size_t)i] == 'n') ++n;
    return n;
}
}  // namespace

void sam_to_seeds(const std::string &sam, std::map<std::string, std::string> &ref_genome,
                  std::map<std::string, std::string> &query_genome, SeedMap &pos, SeedMap &neg) {
    std::ifstream in(sam);
    if (!in.good()) {
        std::cerr << "error in opening sam file " << sam << std::endl;
        std::exit(1);
    }
    // the reference's patterns (a SAM line, a cigar operation, the leading clip)
    static const std::regex line_re("^(.*?)\\s+(\\d+)\\s+(\\S+)\\s+(\\d+)\\s+(\\d+)\\s+(\\w+)");
    static const std::regex op_re("(\\d+)([MIDNSH])");
    static const std::regex clip_re("^(\\d+)([SH])");
    std::string line;
    int32_t lineno = 0;
    while (std::getline(in, line)) {
        if (lineno % 1000 == 0) std::cout << "reading samfile file line " << lineno << std::endl;
        ++lineno;
        if (line.empty() || line[0] == '@' || line.size() < 9) continue;
        std::smatch m;
        std::regex_search(line, m, line_re);
        if (m.empty()) continue;
        std::string cigar = m[6];
        std::smatch clip;
        std::regex_search(cigar, clip, clip_re);
        if (clip.empty()) continue;
        const std::string ref_chr = m[3], query_chr = m[1];
        int32_t ref_start = std::stoi(m[4].str()), query_start = std::stoi(clip[1].str());
        int32_t ref_end = ref_start, query_end = query_start;
        std::smatch op;
        while (std::regex_search(cigar, op, op_re)) {
            const char c = op[2].str()[0];
            const int32_t len = std::stoi(op[1].str());
            if (c == 'M') {
                ref_end += len;
                query_end += len;
            } else if (c == 'D' || c == 'N') {
                ref_end += len;
            } else if (c == 'I') {
                query_end += len;
            }
            cigar = op.suffix().str();
        }
        const int32_t flag = std::stoi(m[2].str());
        ref_start -= 1;
        ref_end -= 2;
        query_start -= 1;
        query_end -= 2;
        const std::string &rg = ref_genome[ref_chr];
        int32_t n1 = n_count(rg, 0, ref_start - 1), n2 = n_count(rg, ref_start, ref_end);
        ref_start -= n1;
        ref_end -= n1;
        ref_end -= n2;
        const std::string &qg = query_genome[query_chr];
        if (flag == 0) {
            n1 = n_count(qg, 0, query_start - 1);
            n2 = n_count(qg, query_start, query_end);
            query_start -= n1;
            query_end -= n1;
            query_end -= n2;
            pos[ref_chr][query_chr].push_back(Seed(ref_start, ref_end, query_start, query_end));
        } else {
            n1 = n_count(qg, query_start, query_end);
            n2 = n_count(qg, query_end + 1, (int64_t)qg.size() - 1);
            const int32_t qlen = (int32_t)qg.size();
            query_end = qlen - 1 - query_start;
            query_start = qlen - 1 - query_end;
            query_start -= n2;
            query_end -= n2;
            query_end -= n1;
            neg[ref_chr][query_chr].push_back(Seed(ref_start, ref_end, query_start, query_end));
        }
    }
}

std::string reverse_complement(const std::string &s) {
    std::string r;
    r.reserve(s.size());
    for (int i = (int)s.size() - 1; i >= 0; --i) {
        char c = s[(size_t)i];
        switch (c) {
            case 'A': c = 'T'; break;
            case 'T': c = 'A'; break;
            case 'U': c = 'A'; break;
            case 'C': c = 'G'; break;
            case 'G': c = 'C'; break;
            case 'R': c = 'Y'; break;
            case 'Y': c = 'R'; break;
            case 'K': c = 'M'; break;
            case 'M': c = 'K'; break;
            case 'B': c = 'V'; break;
            case 'V': c = 'B'; break;
            case 'D': c = 'H'; break;
            case 'H': c = 'D'; break;
            default: break;
        }
        r += c;
    }
    return r;
}

std::string sub_sequence(std::string s, const int &_start, const int &_end) {
    size_t start = (size_t)_start, end = (size_t)_end;
    if (start > end) std::swap(start, end);
    if (start < 1) start = 1;
    if (start > s.size()) {
        start = s.size();
        end = start;
    } else if (end > s.size()) {
        end = s.size();
    }
    return s.substr(start - 1, end - start + 1);
}

std::string sub_sequence(std::string &s, const int &start, const int &end, const int &strand) {
    if (strand == 1) return sub_sequence(s, start, end);
    return reverse_complement(sub_sequence(s, start, end));
}

std::vector<int8_t> encode(const std::string &s) {
    std::vector<int8_t> out(s.size());
    for (size_t i = 0; i < s.size(); ++i) {
        switch (s[i]) {
            case 'A': case 'a': out[i] = 0; break;
            case 'C': case 'c': out[i] = 1; break;
            case 'G': case 'g': out[i] = 3; break;
            case 'T': case 't': case 'U': case 'u': out[i] = 4; break;
            default: out[i] = 2; break;
        }
    }
    return out;
}

namespace {
void push(std::vector<uint32_t> &cig, uint32_t op) {
    if (cig.empty() || op != (cig.back() & 0xfu)) cig.push_back((1u << 4) | op);
    else cig.back() += 1u << 4;
}

// Re-inserts the masked ('n') stretches of the unstripped sequences into the
// cigar: op 3 for reference n's, op 6 for query n's, never merged with the
// alignment's own D/I (controlLayer4.cpp:894-960).
std::vector<uint32_t> unmask_cigar(const std::vector<uint32_t> &cigar, const std::string &ref0, const std::string &que0,
                                   int32_t ref_i, int32_t que_i) {
    std::vector<uint32_t> out;
    for (uint32_t c : cigar) {
        const uint32_t n = c >> 4, type = c & 0xfu;
        for (uint32_t h = 0; h < n; ++h) {
            if (type == 0) {
                while (ref0[(size_t)ref_i] == 'n') { push(out, 3); ++ref_i; }
                ++ref_i;
                while (que0[(size_t)que_i] == 'n') { push(out, 6); ++que_i; }
                ++que_i;
            } else if (type == 1) {
                while (que0[(size_t)que_i] == 'n') { push(out, 6); ++que_i; }
                ++que_i;
            } else {
                while (ref0[(size_t)ref_i] == 'n') { push(out, 3); ++ref_i; }
                ++ref_i;
            }
            push(out, type);
        }
    }
    return out;
}
}  // namespace

// Positions of the non-'n' characters of a masked sequence: the c-th
// (1-based) sits at pos[c - 1] and has pos[c - 1] + 1 - c n's before it, which
// is what outputMasking's scans (controlLayer4.cpp:857-888) find by walking the
// sequence for every fragment.
struct NonN {
    std::vector<int32_t> pos;
    explicit NonN(const std::string &s) {
        pos.reserve(s.size());
        for (size_t j = 0; j < s.size(); ++j)
            if (s[j] != 'n') pos.push_back((int32_t)j);
    }
    bool has(uint32_t c) const { return c >= 1 && c <= pos.size(); }
    int32_t at(uint32_t c) const { return pos[c - 1]; }                  // position j where the count reaches c
    int32_t ns(uint32_t c) const { return pos[c - 1] + 1 - (int32_t)c; }  // n's up to and including j
};

namespace {
// sub_sequence(s, start, end[, 0]) without copying s
std::string sub_text(const std::string &s, int start, int end, bool revcomp) {
    size_t a = (size_t)start, b = (size_t)end;
    if (a > b) std::swap(a, b);
    if (a < 1) a = 1;
    if (a > s.size()) {
        a = s.size();
        b = a;
    } else if (b > s.size()) {
        b = s.size();
    }
    const std::string t = s.substr(a - 1, b - a + 1);
    return revcomp ? reverse_complement(t) : t;
}
}  // namespace

// outputMasking, controlLayer4.cpp:835-1358.  Stripped (n-free) 1-based
// fragment coordinates are mapped back onto the masked region by counting
// n's; the four strand combinations differ in which end is anchored, where
// the n count is applied, the SAM flag, the cigar direction and whether the
// query text is reverse-complemented.  The counts come from NonN indexes
// built once per call; a fragment whose coordinates fall outside the usual
// 1 <= start <= end <= length takes the reference's own walk.
void write_sam_and_o(std::ostream &sam, std::ostream &o, const std::vector<PairedSimilarFragment> &frags,
                     int32_t refStrand, int32_t queStrand, int32_t refStart, int32_t queStart,
                     const std::string &refChr, const std::string &queChr, const std::string &seq1_0,
                     const std::string &seq2_0) {
    o << "#" << refChr << ":" << refStrand << ":" << refStart << "-" << queChr << ":" << queStrand << ":" << queStart
      << '\n';
    const bool strands_ok = (refStrand == 1 || refStrand == 0) && (queStrand == 1 || queStrand == 0);
    std::unique_ptr<NonN> i1, i2;
    if (strands_ok && !frags.empty()) {
        i1.reset(new NonN(seq1_0));
        i2.reset(new NonN(seq2_0));
    }
    for (const PairedSimilarFragment &f : frags) {
        const bool qplus = queStrand == 1, rplus = refStrand == 1;
        if (strands_ok) {
            int32_t quePos = qplus ? (int32_t)(queStart + f.getStart2() - 1)
                                   : (int32_t)(queStart + seq2_0.size() - 1 - f.getEnd2() + 1);
            int32_t refPos = rplus ? (int32_t)(refStart + f.getStart1() - 1)
                                   : (int32_t)(refStart + seq1_0.size() - 1 - f.getEnd1() + 1);
            int32_t start = 1, end = 1;
            if (i2->has(f.getStart2()) && i2->has(f.getEnd2()) && f.getStart2() <= f.getEnd2()) {
                start += i2->at(f.getStart2());
                if (qplus) quePos += i2->ns(f.getStart2());  // plus strand: anchored at the start
                end += i2->at(f.getEnd2());
                if (!qplus) quePos -= i2->ns(f.getEnd2());   // minus strand: anchored at the end
            } else {
                int32_t nchr = 0, nn = 0, started = 0;
                for (int32_t j = 0; j < (int32_t)seq2_0.size(); ++j) {
                    if (seq2_0[(size_t)j] == 'n') ++nn;
                    else ++nchr;
                    if ((uint32_t)nchr == f.getStart2() && started == 0) {
                        start += j;
                        if (qplus) quePos += nn;
                        started = 1;
                    }
                    if ((uint32_t)nchr == f.getEnd2()) {
                        end += j;
                        if (!qplus) quePos -= nn;
                        break;
                    }
                }
            }
            int32_t ref_start = 1;
            if (rplus && i1->has(f.getStart1())) {
                ref_start += i1->at(f.getStart1());
                refPos += i1->ns(f.getStart1());
            } else if (!rplus && i1->has(f.getStart1()) && i1->has(f.getEnd1()) && f.getStart1() <= f.getEnd1()) {
                ref_start += i1->at(f.getStart1());
                refPos -= i1->ns(f.getEnd1());
            } else {
                int32_t nchr = 0, nn = 0, started = 0;
                for (int32_t j = 0; j < (int32_t)seq1_0.size(); ++j) {
                    if (seq1_0[(size_t)j] != 'n') ++nchr;
                    else ++nn;
                    if (rplus) {
                        if ((uint32_t)nchr == f.getStart1() && (started == 0 || !qplus)) {
                            ref_start += j;
                            refPos += nn;
                            break;
                        }
                    } else {
                        if ((uint32_t)nchr == f.getStart1() && started == 0) {
                            ref_start += j;
                            started = 1;
                        }
                        if ((uint32_t)nchr == f.getEnd1()) {
                            refPos -= nn;
                            break;
                        }
                    }
                }
            }
            const int flag = (refStrand == queStrand) ? 0 : 16;
            sam << queChr << "\t" << flag << "\t" << refChr << "\t" << refPos << "\t" << f.getScore() << "\t" << quePos
                << "H";
            std::vector<uint32_t> nc = unmask_cigar(f.getCigar(), seq1_0, seq2_0, ref_start - 1, start - 1);
            if (rplus) {
                for (uint32_t c : nc) sam << (c >> 4) << "MIDDSHI=XB"[c & 0xfu];
            } else {
                for (int32_t j = (int32_t)nc.size() - 1; j >= 0; --j)
                    sam << (nc[(size_t)j] >> 4) << "MIDDSHI=XB"[nc[(size_t)j] & 0xfu];
            }
            sam << "\t*\t0\t0\t" << sub_text(seq2_0, start, end, !rplus) << "\t*" << '\n';
        }
        o << f.getStart1() << " " << f.getEnd1() << " " << f.getStart2() << " " << f.getEnd2() << " " << f.getScore()
          << " " << f.getPValue() << '\n';
    }
}

void write_sam_and_o(const std::string &prefix, const std::vector<PairedSimilarFragment> &frags, int32_t refStrand,
                     int32_t queStrand, int32_t refStart, int32_t queStart, const std::string &refChr,
                     const std::string &queChr, const std::string &seq1_0, const std::string &seq2_0) {
    std::ofstream sam(prefix + ".sam", std::ios_base::app);
    std::ofstream o(prefix + "o", std::ios_base::app);
    write_sam_and_o(sam, o, frags, refStrand, queStrand, refStart, queStart, refChr, queChr, seq1_0, seq2_0);
}

void write_sam_and_o_unmasked(const std::string &prefix, const std::vector<PairedSimilarFragment> &frags,
                              int32_t refStrand, int32_t queStrand, int32_t refStart, int32_t queStart,
                              const std::string &refChr, const std::string &queChr, const std::string &seq1,
                              const std::string &seq2) {
    std::ofstream sam(prefix + ".sam", std::ios_base::app);
    std::ofstream o(prefix + "o", std::ios_base::app);
    o << "#" << refChr << ":" << refStrand << ":" << refStart << "-" << queChr << ":" << queStrand << ":" << queStart
      << std::endl;
    for (const PairedSimilarFragment &f : frags) {
        if ((refStrand == 1 || refStrand == 0) && (queStrand == 1 || queStrand == 0)) {
            const bool rplus = refStrand == 1, qplus = queStrand == 1;
            const int32_t quePos = qplus ? (int32_t)(queStart + f.getStart2() - 1)
                                         : (int32_t)(queStart + seq2.size() - 1 - f.getEnd2() + 1);
            const int32_t refPos = rplus ? (int32_t)(refStart + f.getStart1() - 1)
                                         : (int32_t)(refStart + seq1.size() - 1 - f.getEnd1() + 1);
            const int flag = (refStrand == queStrand) ? 0 : 16;
            sam << queChr << "\t" << flag << "\t" << refChr << "\t" << refPos << "\t" << f.getScore() << "\t"
                << quePos << "H";
            const std::vector<uint32_t> &cg = f.getCigar();
            if (rplus) {
                for (uint32_t c : cg) sam << (c >> 4) << "MID"[c & 0xfu];
            } else {
                for (int32_t j = (int32_t)cg.size() - 1; j >= 0; --j) sam << (cg[(size_t)j] >> 4) << "MID"[cg[(size_t)j] & 0xfu];
            }
            std::string s2 = seq2;
            const std::string text = rplus ? sub_sequence(s2, (int)f.getStart2(), (int)f.getEnd2())
                                           : sub_sequence(s2, (int)f.getStart2(), (int)f.getEnd2(), 0);
            sam << "\t*\t0\t0\t" << text << "\t*" << std::endl;
        }
        o << f.getStart1() << " " << f.getEnd1() << " " << f.getStart2() << " " << f.getEnd2() << " " << f.getScore()
          << " " << f.getPValue() << std::endl;
    }
    sam.close();
    o << "#done" << std::endl;
}

}  // namespace cli
}  // namespace dcns
