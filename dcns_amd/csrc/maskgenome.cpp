// maskgenome.cpp -- staging and launches of maskGenome's device pipeline.
#include "maskgenome.h"

#include <algorithm>
#include <memory>
#include <string>

#include "kernels.h"
#include "kmer_hash.h"

namespace dcns {

int64_t mask_body_size(int64_t len, int32_t width) { return len + (width > 0 ? len / width : 0) + 1; }

// The code members grouped by minimizer bucket into 64-byte slots
// (kmer_hash.h), distinct keys, about two per bucket on average (the k-mers
// sharing a minimizer stay together); keys past the seventh of a bucket go to
// the overflow array.
void build_buckets(const std::vector<uint64_t> &codes, int32_t k, std::vector<uint64_t> &slots,
                   std::vector<uint64_t> &overflow, uint32_t &nb_mask) {
    const int64_t n = (int64_t)codes.size();
    if (n >= (int64_t)1 << 32) throw Error(DCNS_ERR_CAPACITY, "more than 2^32 k-mers");
    int64_t nb = 256;
    while (nb < n / 2) nb *= 2;
    nb_mask = (uint32_t)(nb - 1);
    std::vector<uint32_t> b((size_t)n);
    parallel_for(n, [&](int64_t i) { b[(size_t)i] = kmer_bucket(codes[(size_t)i], k, nb_mask); });
    std::vector<int64_t> start((size_t)nb + 1, 0);
    for (int64_t i = 0; i < n; ++i) ++start[(size_t)b[(size_t)i] + 1];
    for (int64_t j = 0; j < nb; ++j) start[(size_t)j + 1] += start[(size_t)j];
    std::vector<uint64_t> grouped((size_t)n);
    {
        std::vector<int64_t> at(start.begin(), start.end() - 1);
        for (int64_t i = 0; i < n; ++i) grouped[(size_t)at[b[(size_t)i]]++] = codes[(size_t)i];
    }
    std::vector<uint32_t> cnt((size_t)nb);
    const int64_t chunks = 256, per = (nb + chunks - 1) / chunks;
    parallel_for(chunks, [&](int64_t ch) {
        for (int64_t j = ch * per; j < std::min(nb, (ch + 1) * per); ++j) {
            auto lo = grouped.begin() + start[(size_t)j], hi = grouped.begin() + start[(size_t)j + 1];
            std::sort(lo, hi);
            cnt[(size_t)j] = (uint32_t)(std::unique(lo, hi) - lo);
        }
    });
    slots.assign((size_t)nb * kSlotWords, ~0ull);
    overflow.clear();
    for (int64_t j = 0; j < nb; ++j) {
        const uint64_t *key = grouped.data() + start[(size_t)j];
        const uint32_t c = cnt[(size_t)j];
        uint64_t *slot = slots.data() + j * kSlotWords;
        slot[0] = (uint64_t)c | ((uint64_t)overflow.size() << 32);
        for (uint32_t q = 0; q < c; ++q) {
            if (q < (uint32_t)kSlotKeys) slot[1 + q] = key[q];
            else overflow.push_back(key[q]);
        }
    }
    if (overflow.size() >= ((size_t)1 << 32)) throw Error(DCNS_ERR_CAPACITY, "k-mer overflow keys past 2^32");
}

void mask_stage(dcns_ctx *c, const KmerSet &ks, const uint8_t *seq, const int64_t *off, int32_t nseq) {
    if (nseq < 0 || (nseq > 0 && (!off || !seq))) throw Error(DCNS_ERR_ARG, "invalid genome arrays");
    if (ks.k < 0) throw Error(DCNS_ERR_ARG, "negative k-mer size");
    auto ms = std::make_shared<MaskStaged>();
    ms->nseq = nseq;
    ms->off.assign(off, off + nseq + 1);
    if (nseq == 0) ms->off.assign(1, 0);
    if (ms->off[0] != 0) throw Error(DCNS_ERR_ARG, "the first sequence offset must be 0");
    for (int32_t s = 0; s < nseq; ++s) {
        const int64_t len = ms->off[(size_t)s + 1] - ms->off[(size_t)s];
        if (len < 0) throw Error(DCNS_ERR_ARG, "sequence offsets must not decrease");
        if (ks.k >= 1 && len < ks.k)
            throw Error(DCNS_ERR_ARG, "sequence " + std::to_string(s) + " holds " + std::to_string(len) +
                                          " bases, fewer than the k-mer size " + std::to_string(ks.k));
    }
    ms->N = ms->off.back();
    ms->k = ks.k;
    hipStream_t st = c->stream;
    c->mk_seq.ensure((size_t)ms->N + 128);  // the scan reads 48 bytes from each 16-byte block start
    if (ms->N > 0) hip_check(hipMemcpyAsync(c->mk_seq.p, seq, (size_t)ms->N, hipMemcpyHostToDevice, st), "upload genome");
    c->mk_off.ensure(ms->off.size() * sizeof(int64_t));
    hip_check(hipMemcpyAsync(c->mk_off.p, ms->off.data(), ms->off.size() * sizeof(int64_t), hipMemcpyHostToDevice, st),
              "upload offsets");
    // k <= 31: the scan looks every ACGT window up in the minimizer buckets,
    // so they are built (possibly empty) whenever the set has a member of length k
    if (ks.k >= 1 && ks.k <= 31 && (!ks.codes.empty() || ks.n_general() > 0)) {
        std::vector<uint64_t> slots, overflow;
        uint32_t nb_mask = 0;
        build_buckets(ks.codes, ks.k, slots, overflow, nb_mask);
        ms->buckets = (int64_t)nb_mask + 1;
        c->mk_desc.ensure(slots.size() * sizeof(uint64_t));
        c->mk_keys.ensure(overflow.size() * sizeof(uint64_t) + 8);
        hip_check(hipMemcpyAsync(c->mk_desc.p, slots.data(), slots.size() * sizeof(uint64_t), hipMemcpyHostToDevice,
                                 st),
                  "upload k-mer buckets");
        if (!overflow.empty())
            hip_check(hipMemcpyAsync(c->mk_keys.p, overflow.data(), overflow.size() * sizeof(uint64_t),
                                     hipMemcpyHostToDevice, st),
                      "upload k-mer overflow keys");
        c->sync();  // slots and overflow are local
    }
    const int64_t gn = ks.k >= 1 ? ks.n_general() : 0;
    if (gn > 0) {
        std::vector<uint64_t> h((size_t)gn);
        parallel_for(gn, [&](int64_t i) { h[(size_t)i] = kmer_hash(ks.general.data() + i * ks.k, ks.k); });
        ms->gslots = kmer_general_slots(gn);
        c->mk_gbytes.ensure(ks.general.size());
        c->mk_ghash.ensure((size_t)gn * sizeof(uint64_t));
        c->mk_gtab.ensure((size_t)ms->gslots * sizeof(uint32_t));
        hip_check(hipMemcpyAsync(c->mk_gbytes.p, ks.general.data(), ks.general.size(), hipMemcpyHostToDevice, st),
                  "upload k-mers");
        hip_check(hipMemcpyAsync(c->mk_ghash.p, h.data(), h.size() * sizeof(uint64_t), hipMemcpyHostToDevice, st),
                  "upload k-mer hashes");
        c->launch("k_kmer_insert_general", [&] {
            return launch_kmer_general(st, c->mk_ghash.as<uint64_t>(), gn, c->mk_gtab.as<uint32_t>(), ms->gslots);
        });
        ms->general.n = gn;
        ms->general.mask = (uint64_t)(ms->gslots - 1);
        ms->general.pk1 = kmer_hash_pk1(ks.k);
    }
    c->sync();  // the host vectors above die with this call
    c->mstaged = ms;
}

void mask_run(dcns_ctx *c, const CdsMask *cds, uint8_t alt, bool soft, int32_t width, uint8_t *mask_out,
              uint8_t *body_out) {
    if (!c->mstaged) throw Error(DCNS_ERR_ARG, "no staged genome: call dcns_mask_stage first");
    if (width == 0) throw Error(DCNS_ERR_ARG, "line width 0 (the reference divides by it)");
    const MaskStaged &ms = *c->mstaged;
    hipStream_t st = c->stream;
    const int64_t N = ms.N;
    c->mk_mask.ensure((size_t)kmer_mask_padded(N));
    KmerGeneral g = ms.general;
    g.table = c->mk_gtab.as<uint32_t>();
    g.hashes = c->mk_ghash.as<uint64_t>();
    g.bytes = c->mk_gbytes.as<uint8_t>();
    c->launch("k_kmer_scan", [&] {
        return launch_kmer_scan(st, c->mk_seq.as<uint8_t>(), N, c->mk_off.as<int64_t>(), ms.nseq, ms.k,
                                c->mk_desc.as<uint64_t>(), c->mk_keys.as<uint64_t>(), ms.buckets, g,
                                c->mk_mask.as<uint8_t>());
    });

    std::vector<int64_t> ranges;
    if (cds) {
        if (cds->size() != ms.nseq) throw Error(DCNS_ERR_ARG, "CDS masks and staged genome differ in sequences");
        for (int32_t s = 0; s < ms.nseq; ++s) {
            if (cds->seq_size(s) != ms.off[(size_t)s + 1] - ms.off[(size_t)s])
                throw Error(DCNS_ERR_ARG, "CDS mask size differs from staged sequence " + std::to_string(s));
            for (const auto &r : cds->ranges(s)) {
                ranges.push_back(ms.off[(size_t)s] + r.first);
                ranges.push_back(ms.off[(size_t)s] + r.second);
            }
        }
    }
    const int64_t nr = (int64_t)ranges.size() / 2;
    if (nr > 0) {
        c->mk_ranges.ensure(ranges.size() * sizeof(int64_t));
        hip_check(hipMemcpyAsync(c->mk_ranges.p, ranges.data(), ranges.size() * sizeof(int64_t),
                                 hipMemcpyHostToDevice, st),
                  "upload CDS ranges");
        c->launch("k_paint_ranges", [&] {
            return launch_paint_ranges(st, c->mk_ranges.as<int64_t>(), nr, c->mk_mask.as<uint8_t>());
        });
    }

    std::vector<int64_t> boff((size_t)ms.nseq + 1, 0);
    for (int32_t s = 0; s < ms.nseq; ++s)
        boff[(size_t)s + 1] = boff[(size_t)s] + mask_body_size(ms.off[(size_t)s + 1] - ms.off[(size_t)s], width);
    const int64_t total = boff.back();
    c->mk_boff.ensure(boff.size() * sizeof(int64_t));
    hip_check(hipMemcpyAsync(c->mk_boff.p, boff.data(), boff.size() * sizeof(int64_t), hipMemcpyHostToDevice, st),
              "upload body offsets");
    c->mk_body.ensure((size_t)kmer_body_padded(total));
    bool any_empty = false;
    for (int32_t s = 0; s < ms.nseq; ++s) any_empty |= ms.off[(size_t)s + 1] == ms.off[(size_t)s];
    c->launch("k_mask_emit", [&] {
        return launch_mask_emit(st, c->mk_seq.as<uint8_t>(), c->mk_mask.as<uint8_t>(), c->mk_off.as<int64_t>(),
                                c->mk_boff.as<int64_t>(), ms.nseq, N, total, width, any_empty, alt, soft,
                                c->mk_body.as<uint8_t>());
    });
    if (mask_out && N > 0)
        hip_check(hipMemcpyAsync(mask_out, c->mk_mask.p, (size_t)N, hipMemcpyDeviceToHost, st), "download mask");
    if (body_out && total > 0)
        hip_check(hipMemcpyAsync(body_out, c->mk_body.p, (size_t)total, hipMemcpyDeviceToHost, st), "download bodies");
    c->sync();  // ranges / boff are host vectors local to this call
}

}  // namespace dcns
