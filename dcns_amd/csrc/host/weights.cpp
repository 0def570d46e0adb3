// weights.cpp -- score folders and GFF3 categories for the weighted paths.
#include "host/weights.h"

#include <algorithm>
#include <fstream>
#include <set>
#include <sstream>

namespace dcns {
namespace {

int32_t to_int(const std::string &s, const std::string &what) {
    try {
        return std::stoi(s);
    } catch (const std::exception &) {
        throw Error(DCNS_ERR_ARG, "not an integer in " + what + ": '" + s + "'");
    }
}

const std::set<std::string> &gene_like() {
    static const std::set<std::string> s = {
        "transcript", "pseudogenic_transcript", "mRNA", "miRNA", "tRNA", "ncRNA", "mRNA_TE_gene", "rRNA",
        "snoRNA", "snRNA", "lincRNA", "gene", "pseudogene", "transposable_element_gene", "lincRNA_gene",
        "tRNA_gene", "protein", "miRNA_gene", "ncRNA_gene", "lnc_RNA", "pre_miRNA", "SRP_RNA"};
    return s;
}
const std::set<std::string> &exon_like() {
    static const std::set<std::string> s = {"exon", "five_prime_utr", "five_prime_UTR", "three_prime_UTR",
                                            "three_prime_utr"};
    return s;
}
const std::set<std::string> &cds_like() {
    static const std::set<std::string> s = {"CDS", "start_codon", "stop_codon"};
    return s;
}
const std::set<std::string> &te_like() {
    static const std::set<std::string> s = {"transposable_element", "transposon_fragment", "repeat_region",
                                            "LINE_element", "LTR_retrotransposon", "SINE_element", "helitron",
                                            "solo_LTR", "terminal_inverted_repeat_element"};
    return s;
}

// removeDuplications (impl/gffToCategory.cpp:11-33): sort by start, merge
// overlapping intervals; the last merged interval is never appended.
std::vector<std::pair<int64_t, int64_t>> merged_without_last(std::vector<std::pair<int64_t, int64_t>> v) {
    std::vector<std::pair<int64_t, int64_t>> out;
    if (v.empty()) return out;
    std::stable_sort(v.begin(), v.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
    std::pair<int64_t, int64_t> e = v[0];
    for (size_t i = 1; i < v.size(); ++i) {
        if (e.second >= v[i].second) {
        } else if (v[i].first <= e.second) {
            e.second = std::max(e.second, v[i].second);
        } else {
            out.push_back(e);
            e = v[i];
        }
    }
    return out;
}

void paint(const std::vector<std::pair<int64_t, int64_t>> &iv, int16_t value, int64_t chr_size, int64_t from,
           int64_t count, int16_t *out) {
    const int64_t lo = std::max<int64_t>(from, 0), hi = std::min<int64_t>(from + count, chr_size);  // [lo, hi)
    for (const auto &p : iv) {
        const int64_t a = std::max(p.first, lo), b = std::min(p.second + 1, hi);
        for (int64_t x = a; x < b; ++x) out[x - from] = value;
    }
}

}  // namespace

// split of model/utils.cpp:9-17: tab-separated, empty fields dropped
std::vector<std::string> split_fields(const std::string &line) {
    std::vector<std::string> out;
    std::stringstream ss(line);
    std::string item;
    while (std::getline(ss, item, '\t'))
        if (!item.empty()) out.push_back(item);
    return out;
}

std::vector<Profile> read_score_folder(const std::string &folder) {
    std::ifstream list(folder + "/matrixList");
    if (!list.good()) throw Error(DCNS_ERR_ARG, "error in opening score matrix list file " + folder + "/matrixList");
    std::vector<Profile> out;
    std::string line;
    while (std::getline(list, line)) {
        if (line.empty()) continue;
        const int16_t category = (int16_t)to_int(line, folder + "/matrixList");
        std::ifstream f(folder + "/" + line);
        if (!f.good()) throw Error(DCNS_ERR_ARG, "error in opening score matrix list file " + folder + "/" + line);
        auto it = std::find_if(out.begin(), out.end(), [&](const Profile &p) { return p.category == category; });
        if (it == out.end()) {
            out.push_back(Profile{category, 0, 0, 0, 0, 0, 0, 0});
            it = out.end() - 1;
        }
        // match and mismatch are rebuilt from this file (0 when absent); the
        // penalties and z-drop keep earlier values this file does not set
        it->match = 0;
        it->mismatch = 0;
        int32_t *fields7[7] = {&it->match, &it->mismatch, &it->open1, &it->ext1, &it->open2, &it->ext2, &it->zdrop};
        std::string l2;
        int k = 0;
        while (std::getline(f, l2)) {
            if (k < 7) *fields7[k] = to_int(l2, folder + "/" + line);
            ++k;
        }
    }
    return out;
}

GffAnnotation read_gff(const std::string &path) {
    std::ifstream in(path);
    if (!in.good()) throw Error(DCNS_ERR_ARG, "error in opening GFF/GTF file " + path);
    GffAnnotation g;
    std::string line;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] == '#' || line.size() < 9) continue;
        const std::vector<std::string> e = split_fields(line);
        if (e.size() != 9) continue;
        int64_t start = to_int(e[3], path), end = to_int(e[4], path);
        if (start > end) std::swap(start, end);
        --start;  // 0-based
        --end;
        const std::string &type = e[2];
        if (gene_like().count(type))
            g.chrom[e[0]].introns.emplace_back(start, end);
        else if (exon_like().count(type))
            g.chrom[e[0]].exons.emplace_back(start, end);
        else if (cds_like().count(type))
            g.chrom[e[0]].cds.emplace_back(start, end);
        else if (te_like().count(type))
            g.chrom[e[0]].transposons.emplace_back(start, end);
        else if (type != "chromosome" && type != "contig")
            g.unknown_lines.push_back(line);
    }
    return g;
}

void gff_categories(const GffAnnotation &g, const std::string &chr, int64_t chr_size, int64_t from, int64_t count,
                    int16_t *out) {
    std::fill(out, out + count, (int16_t)10);
    auto it = g.chrom.find(chr);
    if (it == g.chrom.end()) return;
    const GffChrom &c = it->second;
    paint(c.introns, 70, chr_size, from, count, out);
    paint(merged_without_last(c.exons), 85, chr_size, from, count, out);
    paint(c.cds, 100, chr_size, from, count, out);
    paint(c.transposons, 1, chr_size, from, count, out);
}

}  // namespace dcns
