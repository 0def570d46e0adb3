// mask_io.cpp -- the jellyfish dump reader and the CDS masks of maskGenome.
#include "host/mask_io.h"

#include <algorithm>
#include <cstdio>
#include <fstream>
#include <memory>
#include <ostream>
#include <regex>

#include "ctx.h"
#include "host/weights.h"

namespace dcns {
namespace {

int64_t to_int32(const std::string &s, const std::string &what) {
    try {
        return std::stoi(s);  // the reference's parse: leading blanks, sign, digits, rest ignored
    } catch (const std::exception &) {
        throw Error(DCNS_ERR_ARG, what + ": not a 32-bit integer: '" + s + "'");
    }
}

// count of a dump header line ">count ...": the text between '>' and the
// first space (impl/maskGenome.cpp:34)
int32_t header_count(const char *p, size_t n, const std::string &path) {
    size_t e = 1;
    while (e < n && p[e] != ' ') ++e;
    const size_t len = e - 1;
    if (len >= 1 && len <= 9) {
        int32_t v = 0;
        size_t j = 1;
        for (; j < e && p[j] >= '0' && p[j] <= '9'; ++j) v = v * 10 + (p[j] - '0');
        if (j == e) return v;
    }
    return (int32_t)to_int32(std::string(p + 1, len), "k-mer count in " + path);
}

// the first base in the lowest two bits, as the scan kernel packs windows
bool pack(const std::string &s, uint64_t &code) {
    code = 0;
    for (size_t j = 0; j < s.size(); ++j) {
        const int b = kmer_base((unsigned char)s[j]);
        if (b < 0) return false;
        code |= (uint64_t)b << (2 * j);
    }
    return true;
}

}  // namespace

uint64_t kmer_hash(const uint8_t *p, int32_t k) {
    uint64_t h = 0;
    for (int32_t j = 0; j < k; ++j) h = h * kKmerHashBase + p[j];
    return h;
}

uint64_t kmer_hash_pk1(int32_t k) {
    uint64_t r = 1;
    for (int32_t j = 1; j < k; ++j) r *= kKmerHashBase;
    return r;
}

KmerSet read_kmer_dump(const std::string &path, int32_t freq) {
    KmerSet ks;
    if (freq <= 0) return ks;
    std::unique_ptr<FILE, int (*)(FILE *)> f(std::fopen(path.c_str(), "rb"), &std::fclose);
    if (!f) throw Error(DCNS_ERR_ARG, "error in opening kmer Profile file " + path);

    std::string cur;  // the current record's bytes, upper-cased
    int32_t count = 0;
    int64_t first_len = -1;  // records of this length and of A/C/G/T only go straight to codes
    std::vector<uint8_t> other;
    std::vector<int64_t> other_off(1, 0);
    auto keep = [&]() {
        ++ks.records;
        if (first_len < 0) first_len = (int64_t)cur.size();
        uint64_t code;
        if ((int64_t)cur.size() == first_len && cur.size() <= 31 && pack(cur, code)) {
            ks.codes.push_back(code);
        } else {
            other.insert(other.end(), cur.begin(), cur.end());
            other_off.push_back((int64_t)other.size());
        }
    };
    auto line = [&](const char *p, size_t n) {
        if (n > 0 && p[0] == '>') {
            if (count > freq) keep();
            count = header_count(p, n, path);
            cur.clear();
        } else {
            const size_t o = cur.size();
            cur.append(p, n);
            for (size_t j = o; j < cur.size(); ++j)
                if (cur[j] >= 'a' && cur[j] <= 'z') cur[j] = (char)(cur[j] - 32);
        }
    };

    std::vector<char> buf((size_t)1 << 24);
    std::string carry;
    size_t got;
    while ((got = std::fread(buf.data(), 1, buf.size(), f.get())) > 0) {
        const char *p = buf.data(), *end = p + got;
        while (p < end) {
            const char *nl = (const char *)std::memchr(p, '\n', (size_t)(end - p));
            if (!nl) {
                carry.append(p, (size_t)(end - p));
                break;
            }
            if (!carry.empty()) {
                carry.append(p, (size_t)(nl - p));
                line(carry.data(), carry.size());
                carry.clear();
            } else {
                line(p, (size_t)(nl - p));
            }
            p = nl + 1;
        }
    }
    if (std::ferror(f.get())) throw Error(DCNS_ERR_ARG, "error in reading kmer Profile file " + path);
    if (!carry.empty()) line(carry.data(), carry.size());

    ks.k = (int32_t)cur.size();  // kmerSize: the length of the last record
    if (count > freq) keep();
    const int32_t k = ks.k;
    if (k == 0 || first_len != k) ks.codes.clear();
    if (k == 0) return ks;
    for (size_t r = 0; r + 1 < other_off.size(); ++r) {
        if (other_off[r + 1] - other_off[r] != k) continue;
        const std::string s(other.begin() + other_off[r], other.begin() + other_off[r + 1]);
        uint64_t code;
        if (k <= 31 && pack(s, code)) ks.codes.push_back(code);
        else ks.general.insert(ks.general.end(), s.begin(), s.end());
    }
    return ks;
}

CdsMask::CdsMask(const std::vector<std::string> &names, const std::vector<int64_t> &sizes)
    : names_(names), size_(sizes), len_(sizes), ranges_(names.size()) {
    if (names.size() != sizes.size()) throw Error(DCNS_ERR_ARG, "names and sizes differ in number");
    for (size_t i = 0; i < names.size(); ++i) {
        if (sizes[i] < 0) throw Error(DCNS_ERR_ARG, "negative sequence size");
        if (!index_.emplace(names[i], (int32_t)i).second) throw Error(DCNS_ERR_ARG, "duplicate sequence " + names[i]);
    }
}

// ifCds.clear(), then a string of '0' per genome sequence
void CdsMask::reset() {
    len_ = size_;
    for (auto &r : ranges_) r.clear();
    other_len_.clear();
    any_ = !names_.empty();
}

// ifCds[chr].replace(pos, len, len, '1')
void CdsMask::mark(const std::string &chr, int64_t pos, int64_t len, const std::string &what) {
    any_ = true;
    auto it = index_.find(chr);
    int64_t &cur = it != index_.end() ? len_[(size_t)it->second] : other_len_[chr];
    if (pos < 0 || pos > cur)
        throw Error(DCNS_ERR_ARG, what + ": range at " + std::to_string(pos + 1) + " of " + chr +
                                      " starts past the end of that sequence's mask (" + std::to_string(cur) + ")");
    cur = std::max(cur, pos + len);
    if (it != index_.end()) {
        const int64_t b = std::min(pos + len, size_[(size_t)it->second]);
        if (pos < b) ranges_[(size_t)it->second].emplace_back(pos, b);
    }
}

void CdsMask::gff(const std::string &path, bool gene) {
    reset();
    std::ifstream in(path);
    if (!in.good()) throw Error(DCNS_ERR_ARG, "error in opening GFF/GTF file " + path);
    std::string line;
    while (std::getline(in, line)) {
        if (line.size() < 9 || line[0] == '#') continue;
        const std::vector<std::string> e = split_fields(line);
        if (e.size() != 9) continue;
        int64_t start = to_int32(e[3], path), end = to_int32(e[4], path);
        if (start > end) std::swap(start, end);
        --start;  // 0-based
        const bool take = gene ? (e[2] == "gene" && e[8].find("biotype=protein_coding") != std::string::npos)
                               : (e[2] == "CDS" || e[2] == "start_codon" || e[2] == "stop_codon");
        if (take) mark(e[0], start, end - start, path);
    }
}

bool CdsMask::sam(const std::string &path, const std::map<std::string, int64_t> &cds_sizes, double similarity,
                  std::ostream *log) {
    bool recreate = !any_;
    for (size_t i = 0; i < names_.size(); ++i)
        if (len_[i] != size_[i]) recreate = true;
    if (recreate) {
        reset();
        if (log) *log << "GFF was not used for genome masking" << std::endl;
    }
    std::ifstream in(path);
    if (!in.good()) throw Error(DCNS_ERR_ARG, "error in opening file " + path);
    // the record and tag patterns of impl/getAllTheCdsPositions.cpp:129-135
    static const std::regex rec("^(.*?)\\s+(\\d+)\\s+(\\S+)\\s+(\\d+)\\s+(\\d+)\\s+(\\w+)\\t.*NM:i:(\\d+).*cs:(\\S+)");
    static const std::regex op("(\\d+)([MIDNSH])");
    static const std::regex run(":(\\d+)");
    std::smatch m, om, rm;
    std::string line;
    while (std::getline(in, line)) {
        if (line.size() < 9 || line[0] == '@') continue;
        if (!std::regex_search(line, m, rec)) continue;
        double identical = 0;
        std::string cs = m[8];
        while (std::regex_search(cs, rm, run)) {
            identical += (double)to_int32(rm[1], path);
            cs = rm.suffix().str();
        }
        auto it = cds_sizes.find(m[1].str());
        const double cds_len = it == cds_sizes.end() ? 0.0 : (double)it->second;
        if (!(identical >= similarity * cds_len)) continue;
        const std::string chr = m[3];
        int64_t start = to_int32(m[4], path);
        std::string cigar = m[6];
        while (std::regex_search(cigar, om, op)) {
            const char c = om[2].str()[0];
            const int64_t len = to_int32(om[1], path);
            if (c == 'M' || c == 'D') {
                mark(chr, start - 1, len, path);
                start += len;
            } else if (c == 'N') {
                start += len;
            }
            cigar = om.suffix().str();
        }
    }
    return recreate;
}

}  // namespace dcns
