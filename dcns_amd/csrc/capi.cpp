// capi.cpp -- the extern "C" boundary (include/dcns.h).  Exceptions never
// cross it: every entry point converts them into a status code and keeps
// the message for dcns_last_error().
#include <algorithm>
#include <cstring>
#include <fstream>
#include <map>

#include "../../include/dcns.h"
#include "ctx.h"
#include "host/cli_io.h"
#include "host/mask_io.h"
#include "host/weights.h"
#include "kernels.h"
#include "maskgenome.h"
#include "pipeline.h"

using namespace dcns;

struct dcns_kmer_set {
    KmerSet ks;
};
struct dcns_cds_mask {
    CdsMask m;
};

namespace {
thread_local std::string g_last_error;

template <class F>
int guarded(dcns_ctx *c, F &&f) {
    try {
        f();
        return DCNS_OK;
    } catch (const Error &e) {
        if (c) c->err = e.what();
        g_last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        if (c) c->err = "host allocation failed";
        g_last_error = "host allocation failed";
        return DCNS_ERR_INTERNAL;
    } catch (const std::exception &e) {
        if (c) c->err = e.what();
        g_last_error = e.what();
        return DCNS_ERR_INTERNAL;
    }
}

void check_batch(const dcns_batch *b) {
    if (!b || b->n < 0) throw Error(DCNS_ERR_ARG, "invalid batch");
    if (b->n > 0 && (!b->seqs || !b->off1 || !b->len1 || !b->off2 || !b->len2))
        throw Error(DCNS_ERR_ARG, "batch arrays must be non-null");
}
}  // namespace

extern "C" {

const char *dcns_version(void) { return "dcns-mi355x 0.1 (gfx950)"; }

void dcns_default_params(dcns_params *p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->match = 2;
    p->mismatch = -3;
    p->open = -4;
    p->ext = -2;
    p->open2 = -45;
    p->ext2 = 0;
    p->zdrop = 50;
    p->bandwidth = 20000;
    p->window = 38;
    p->step = 8;
    p->min_score = 40;
    p->w = 10;
    p->xdrop = 20;
    p->pvalue = 0.1;
    p->lambda = 0.382291;
    p->k = 0.006662;
}

int dcns_ctx_create(int device, uint32_t flags, dcns_ctx **out) {
    if (!out) return DCNS_ERR_ARG;
    *out = nullptr;
    dcns_ctx *c = nullptr;
    int rc = guarded(nullptr, [&] {
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) throw Error(DCNS_ERR_DEVICE, "no HIP device available");
        if (device < 0 || device >= n) throw Error(DCNS_ERR_ARG, "device index out of range");
        hipDeviceProp_t prop;
        hip_check(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
        if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
            throw Error(DCNS_ERR_DEVICE, std::string("kernels are built for gfx950, device is ") + prop.gcnArchName);
        hip_check(hipSetDevice(device), "hipSetDevice");
        c = new dcns_ctx();
        c->device = device;
        c->timing = (flags & 1u) != 0;
        hip_check(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate");
        hip_check(sw_noalloc_init(device), "allocation-free scratch address");
    });
    if (rc != DCNS_OK) {
        delete c;
        return rc;
    }
    *out = c;
    return DCNS_OK;
}

int dcns_ctx_destroy(dcns_ctx *c) {
    if (!c) return DCNS_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto &pe : c->pending) {
        (void)hipEventDestroy(pe.second.first);
        (void)hipEventDestroy(pe.second.second);
    }
    for (hipEvent_t e : c->event_pool) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return DCNS_OK;
}

const char *dcns_last_error(const dcns_ctx *c) { return c ? c->err.c_str() : g_last_error.c_str(); }

int dcns_device_info(dcns_ctx *c, int32_t *cu_count, int32_t *clock_mhz, char *name, int32_t name_cap) {
    if (!c) return DCNS_ERR_ARG;
    return guarded(c, [&] {
        hipDeviceProp_t prop;
        hip_check(hipGetDeviceProperties(&prop, c->device), "hipGetDeviceProperties");
        if (cu_count) *cu_count = prop.multiProcessorCount;
        if (clock_mhz) *clock_mhz = prop.clockRate / 1000;
        if (name && name_cap > 0) {
            std::strncpy(name, prop.gcnArchName, (size_t)name_cap - 1);
            name[name_cap - 1] = 0;
        }
    });
}

int dcns_last_stats(const dcns_ctx *c, dcns_stats *out) {
    if (!c || !out) return DCNS_ERR_ARG;
    *out = c->stats;
    return DCNS_OK;
}

int dcns_get_tuning(const dcns_ctx *c, dcns_tuning *out) {
    if (!c || !out) return DCNS_ERR_ARG;
    *out = c->tune;
    return DCNS_OK;
}

int dcns_set_tuning(dcns_ctx *c, const dcns_tuning *in) {
    if (!c || !in) return DCNS_ERR_ARG;
    return guarded(c, [&] {
        const dcns_tuning &t = *in;
        if (t.level0_mode < 0 || t.level0_mode > 2) throw Error(DCNS_ERR_ARG, "level0_mode must be 0, 1 or 2");
        if (t.level0_waves < 0 || t.level0_waves > 32) throw Error(DCNS_ERR_ARG, "level0_waves must be in [0, 32]");
        if (t.wave_tasks < 0 || t.wave_rows < 0 || t.pool_min < 0) throw Error(DCNS_ERR_ARG, "negative tuning value");
        if (t.graph_levels < -1 || t.graph_levels > 64) throw Error(DCNS_ERR_ARG, "graph_levels must be in [-1, 64]");
        if (t.xdrop_lane_rows < -1) throw Error(DCNS_ERR_ARG, "xdrop_lane_rows must be >= -1");
        if (t.k1_lanes < 0 || t.k1_lanes > 3) throw Error(DCNS_ERR_ARG, "k1_lanes must be in [0, 3]");
        if (t.level0_tail < -1 || t.level0_tail > 1) throw Error(DCNS_ERR_ARG, "level0_tail must be in [-1, 1]");
        if (t.level0_split < -1 || t.level0_split > 1) throw Error(DCNS_ERR_ARG, "level0_split must be in [-1, 1]");
        c->tune = t;
    });
}

int dcns_kernel_time(const dcns_ctx *c, const char *name, double *ms, int64_t *launches) {
    if (!c || !name) return DCNS_ERR_ARG;
    auto it = c->timers.find(name);
    if (ms) *ms = it == c->timers.end() ? 0.0 : it->second.ms;
    if (launches) *launches = it == c->timers.end() ? 0 : it->second.launches;
    return DCNS_OK;
}

int dcns_kernel_time_reset(dcns_ctx *c) {
    if (!c) return DCNS_ERR_ARG;
    return guarded(c, [&] {
        c->sync();
        c->timers.clear();
    });
}

namespace {
// SmithWaterman max and first end for the pairs of b (dcns_sw_batch); nn:
// N against N scores -1.  Scores already validated.
void sw_batch_run(dcns_ctx *c, const dcns_batch *b, int32_t match, int32_t mismatch, int32_t open, int32_t ext,
                  bool nn, int32_t *out3) {
    hip_check(hipSetDevice(c->device), "hipSetDevice");
    const int32_t n = b->n;
    if (n == 0) return;
    dcns_params dp;
    dcns_default_params(&dp);
    dp.match = match;
    dp.mismatch = mismatch;
    dp.open = open;
    dp.ext = ext;
    Params P = to_params(dp);
    ScoreTab tab = make_score_tab(match, mismatch);
    // stage sequences
    std::vector<int64_t> o1((size_t)n), o2((size_t)n);
    std::vector<int32_t> l1(b->len1, b->len1 + n), l2(b->len2, b->len2 + n);
    int64_t total = 0;
    for (int32_t i = 0; i < n; ++i) {
        if (l1[(size_t)i] < 0 || l2[(size_t)i] < 0) throw Error(DCNS_ERR_ARG, "negative length");
        total += (l1[(size_t)i] + 31) / 16 * 16 + (l2[(size_t)i] + 47) / 16 * 16;
    }
    std::vector<int8_t> host((size_t)total + 64, 2);
    int64_t o = 0;
    for (int32_t i = 0; i < n; ++i) {
        o1[(size_t)i] = o;
        std::memcpy(&host[(size_t)o], b->seqs + b->off1[i], (size_t)l1[(size_t)i]);
        o += (l1[(size_t)i] + 31) / 16 * 16;  // 16-byte aligned, >= 16 bytes of padding (the kernel reads 16 at a time)
        o2[(size_t)i] = o;
        std::memcpy(&host[(size_t)o], b->seqs + b->off2[i], (size_t)l2[(size_t)i]);
        o += (l2[(size_t)i] + 47) / 16 * 16;  // query stripes read 32 bytes from a multiple of 32
    }
    // lane plan: waves of one loop shape, longest first
    SwPlan plan;
    sw_plan(l1.data(), l2.data(), n, 1, plan);
    auto up = [&](DBuf &d, const void *src, size_t bytes) {
        d.ensure(std::max<size_t>(bytes, 64));
        hip_check(hipMemcpyAsync(d.p, src, bytes, hipMemcpyHostToDevice, c->stream), "upload");
    };
    c->staged.reset();
    up(c->seq, host.data(), host.size());
    up(c->sw_off1, o1.data(), o1.size() * 8);
    up(c->sw_len1, l1.data(), l1.size() * 4);
    up(c->sw_off2, o2.data(), o2.size() * 8);
    up(c->sw_len2, l2.data(), l2.size() * 4);
    up(c->sw_order, plan.order.data(), plan.order.size() * 4);
    c->sw_out.ensure((size_t)n * 12 + 64);
    // boundary rows: launches of whole waves within 1 GiB
    const int64_t budget = ((int64_t)1 << 30) / 4;
    std::vector<int64_t> woff;
    for (int64_t w0 = 0; w0 < (int64_t)plan.waves.size();) {
        const int64_t w1 = sw_plan_chunk(plan, w0, budget, woff);
        c->sw_bound.ensure((size_t)woff.back() * 4 + 64);
        up(c->sw_woff, woff.data(), woff.size() * 8);
        const int32_t s = (int32_t)(64 * w0), m = (int32_t)std::min<int64_t>(64 * w1, n) - s;
        c->launch("sw_batch", [&] {
            return launch_sw_batch(c->stream, c->seq.as<int8_t>(), c->sw_off1.as<int64_t>(), c->sw_len1.as<int32_t>(),
                                   c->sw_off2.as<int64_t>(), c->sw_len2.as<int32_t>(),
                                   c->sw_order.as<int32_t>() + s, m, P, tab, c->sw_out.as<int32_t>(),
                                   c->sw_bound.as<int32_t>(), c->sw_woff.as<int64_t>(), nn);
        });
        w0 = w1;
    }
    hip_check(hipMemcpyAsync(out3, c->sw_out.p, (size_t)n * 12, hipMemcpyDeviceToHost, c->stream), "download");
    c->sync();
}

// Sequences of a batch staged for one-thread-per-pair kernels: 16-byte
// aligned offsets with slack past every sequence.
void stage_pairs(dcns_ctx *c, const dcns_batch *b, std::vector<int64_t> &o1, std::vector<int64_t> &o2) {
    const int32_t n = b->n;
    o1.resize((size_t)n);
    o2.resize((size_t)n);
    int64_t total = 0;
    for (int32_t i = 0; i < n; ++i) total += (b->len1[i] + 31) / 16 * 16 + (b->len2[i] + 47) / 16 * 16;
    std::vector<int8_t> host((size_t)total + 64, 2);
    int64_t o = 0;
    for (int32_t i = 0; i < n; ++i) {
        o1[(size_t)i] = o;
        std::memcpy(&host[(size_t)o], b->seqs + b->off1[i], (size_t)b->len1[i]);
        o += (b->len1[i] + 31) / 16 * 16;
        o2[(size_t)i] = o;
        std::memcpy(&host[(size_t)o], b->seqs + b->off2[i], (size_t)b->len2[i]);
        o += (b->len2[i] + 47) / 16 * 16;
    }
    c->staged.reset();
    c->seq.ensure(std::max<size_t>(host.size(), 64));
    hip_check(hipMemcpyAsync(c->seq.p, host.data(), host.size(), hipMemcpyHostToDevice, c->stream), "upload");
}
}  // namespace

int dcns_sw_batch(dcns_ctx *c, const dcns_batch *b, int32_t match, int32_t mismatch, int32_t open, int32_t ext,
                  int32_t *out3) {
    if (!c) return DCNS_ERR_ARG;
    return guarded(c, [&] {
        check_batch(b);
        if (b->n > 0 && !out3) throw Error(DCNS_ERR_ARG, "out3 must be non-null");
        if (match < 1 || match > 127 || mismatch > 0 || mismatch < -128 || open > 0 || open < -128 || ext > 0 ||
            ext < -128)
            throw Error(DCNS_ERR_ARG, "scores must fit int8 with match > 0 and penalties <= 0");
        sw_batch_run(c, b, match, mismatch, open, ext, false, out3);
    });
}

int dcns_sw_bas_batch(dcns_ctx *c, const dcns_batch *b, int32_t open, int32_t ext, int32_t reverse, uint32_t *out3,
                      uint32_t *cigar, int64_t *cig_off, int64_t cig_cap, int32_t *status) {
    if (!c) return DCNS_ERR_ARG;
    return guarded(c, [&] {
        check_batch(b);
        const int32_t n = b->n;
        if (n > 0 && !out3) throw Error(DCNS_ERR_ARG, "out3 must be non-null");
        if (open > 0 || ext > 0 || open + ext < -128)
            throw Error(DCNS_ERR_ARG, "gap penalties must be <= 0 with open + ext >= -128");
        for (int32_t i = 0; i < n; ++i)
            if (b->len1[i] < 0 || b->len2[i] < 0) throw Error(DCNS_ERR_ARG, "negative length");
        if (!reverse) {
            // max and ends: the batched K1 kernel with the file's scores (match
            // 2, mismatch -1, N against N -1) and gap open = open + ext
            std::vector<int32_t> o3((size_t)std::max(n, 1) * 3);
            if (n > 0) sw_batch_run(c, b, 2, -1, open + ext, ext, true, o3.data());
            for (int64_t k = 0; k < 3 * (int64_t)n; ++k) out3[k] = (uint32_t)o3[(size_t)k];
            if (status)
                for (int32_t i = 0; i < n; ++i) status[i] = 0;
            if (cig_off)
                for (int32_t i = 0; i <= n; ++i) cig_off[i] = 0;
            return;
        }
        if (n > 0 && (!cig_off || !status || (cig_cap > 0 && !cigar)))
            throw Error(DCNS_ERR_ARG, "cigar offsets and status must be non-null with reverse");
        if (n == 0) {
            if (cig_off) cig_off[0] = 0;
            return;
        }
        hip_check(hipSetDevice(c->device), "hipSetDevice");
        // the full matrices (12 bytes per cell) and a walk of at most 2 (L1 + L2) + 4 entries per pair
        std::vector<int64_t> wso((size_t)n), co((size_t)n + 1, 0);
        int64_t ws = 0;
        for (int32_t i = 0; i < n; ++i) {
            wso[(size_t)i] = ws;
            ws += 3 * ((int64_t)b->len1[i] + 1) * ((int64_t)b->len2[i] + 1);
            co[(size_t)i + 1] = co[(size_t)i] + 2 * ((int64_t)b->len1[i] + b->len2[i]) + 4;
        }
        size_t free_b = 0, total_b = 0;
        hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
        if (ws * 4 > (int64_t)(free_b / 2))
            throw Error(DCNS_ERR_CAPACITY, "SmithWaterman_bas full matrices exceed half the free device memory");
        std::vector<int64_t> o1, o2;
        stage_pairs(c, b, o1, o2);
        auto up = [&](DBuf &d, const void *src, size_t bytes) {
            d.ensure(std::max<size_t>(bytes, 64));
            hip_check(hipMemcpyAsync(d.p, src, bytes, hipMemcpyHostToDevice, c->stream), "upload");
        };
        up(c->sw_off1, o1.data(), o1.size() * 8);
        up(c->sw_len1, b->len1, (size_t)n * 4);
        up(c->sw_off2, o2.data(), o2.size() * 8);
        up(c->sw_len2, b->len2, (size_t)n * 4);
        up(c->sw_order, wso.data(), wso.size() * 8);  // workspace offsets
        up(c->sw_woff, co.data(), co.size() * 8);     // walk capacity offsets
        c->sw_bound.ensure((size_t)std::max<int64_t>(ws, 16) * 4);
        c->sw_out.ensure((size_t)n * 12 + 64);
        c->bas_cig.ensure((size_t)std::max<int64_t>(co[(size_t)n], 16) * 4);
        c->bas_st.ensure((size_t)n * 8 + 64);
        c->launch("sw_bas", [&] {
            return launch_sw_bas_full(c->stream, c->seq.as<int8_t>(), c->sw_off1.as<int64_t>(),
                                      c->sw_len1.as<int32_t>(), c->sw_off2.as<int64_t>(), c->sw_len2.as<int32_t>(), n,
                                      open, ext, true, c->sw_bound.as<int32_t>(), c->sw_order.as<int64_t>(),
                                      c->sw_out.as<uint32_t>(), c->bas_cig.as<uint32_t>(), c->sw_woff.as<int64_t>(),
                                      c->bas_st.as<int32_t>(), c->bas_st.as<int32_t>() + n);
        });
        std::vector<uint32_t> cg((size_t)co[(size_t)n]);
        std::vector<int32_t> nst((size_t)2 * n);
        hip_check(hipMemcpyAsync(out3, c->sw_out.p, (size_t)n * 12, hipMemcpyDeviceToHost, c->stream), "download");
        hip_check(hipMemcpyAsync(nst.data(), c->bas_st.p, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream), "status");
        hip_check(hipMemcpyAsync(cg.data(), c->bas_cig.p, cg.size() * 4, hipMemcpyDeviceToHost, c->stream), "cigar");
        c->sync();
        int64_t used = 0;
        for (int32_t i = 0; i < n; ++i) {
            cig_off[i] = used;
            status[i] = nst[(size_t)n + i];
            const int32_t k = nst[(size_t)i];
            if (used + k > cig_cap) throw Error(DCNS_ERR_CAPACITY, "cigar buffer too small");
            std::memcpy(cigar + used, cg.data() + co[(size_t)i], (size_t)k * 4);
            used += k;
        }
        cig_off[n] = used;
    });
}

int dcns_sw_batch_noalloc(dcns_ctx *c, const int8_t *d_seqs, const int64_t *d_off1, const int32_t *d_len1,
                          const int64_t *d_off2, const int32_t *d_len2, int32_t n, int32_t max_len1, int32_t match,
                          int32_t mismatch, int32_t open, int32_t ext, int32_t *d_out3) {
    if (!c) return DCNS_ERR_ARG;
    return guarded(c, [&] {
        if (n < 0) throw Error(DCNS_ERR_ARG, "n must be >= 0");
        if (n == 0) return;
        if (!d_seqs || !d_off1 || !d_len1 || !d_off2 || !d_len2 || !d_out3)
            throw Error(DCNS_ERR_ARG, "device pointers must be non-null");
        if (((uintptr_t)d_seqs & 15u) != 0) throw Error(DCNS_ERR_ARG, "d_seqs must be 16-byte aligned");
        if (max_len1 < 0 || (int64_t)max_len1 > sw_noalloc_max_rows())
            throw Error(DCNS_ERR_ARG, "max_len1 must be in [0, " + std::to_string(sw_noalloc_max_rows()) +
                                          "] (the library's boundary array holds one reference of that many rows)");
        if (match < 1 || match > 127 || mismatch > 0 || mismatch < -128 || open > 0 || open < -128 || ext > 0 ||
            ext < -128)
            throw Error(DCNS_ERR_ARG, "scores must fit int8 with match > 0 and penalties <= 0");
        hip_check(hipSetDevice(c->device), "hipSetDevice");
        dcns_params dp;
        dcns_default_params(&dp);
        dp.match = match;
        dp.mismatch = mismatch;
        dp.open = open;
        dp.ext = ext;
        const Params P = to_params(dp);
        const ScoreTab tab = make_score_tab(match, mismatch);
        int32_t err = 0;
        hip_check(sw_noalloc_run(c->device, c->stream, d_seqs, d_off1, d_len1, d_off2, d_len2, n, max_len1, P, tab,
                                 d_out3, &err),
                  "dcns_sw_batch_noalloc");
        if (err) throw Error(DCNS_ERR_ARG, "a pair's reference is longer than max_len1 (or a length is negative)");
    });
}

namespace dcns {
// The launch plan of dcns_sw_random_max for one set of point shapes: the
// routing to the wave kernel, the lane plan (uploaded to rnd_order) and the
// wave pairs (uploaded to rnd_wlist).  A call with the same lengths, pairs
// per point and launch shape (the permutation method's settings in turn)
// reuses it.
struct RandPlan {
    uint64_t key = 0;
    SwPlan plan;
    std::vector<int32_t> wids;
};
}  // namespace dcns

int dcns_sw_random_max(dcns_ctx *c, int32_t match, int32_t mismatch, int32_t open, int32_t ext, const double *gc,
                       const int32_t *len1, const int32_t *len2, const uint64_t *seed, int32_t npoints,
                       int32_t pairs_per_point, int32_t *out) {
    if (!c) return DCNS_ERR_ARG;
    return guarded(c, [&] {
        if (npoints < 0 || pairs_per_point < 1 || (npoints > 0 && (!gc || !len1 || !len2 || !seed || !out)))
            throw Error(DCNS_ERR_ARG, "invalid point arrays");
        if (match < 1 || match > 127 || mismatch > 0 || mismatch < -128 || open > 0 || open < -128 || ext > 0 ||
            ext < -128)
            throw Error(DCNS_ERR_ARG, "scores must fit int8 with match > 0 and penalties <= 0");
        if (npoints == 0) return;
        hip_check(hipSetDevice(c->device), "hipSetDevice");
        std::vector<uint32_t> thr((size_t)npoints * 3);
        int32_t maxl1 = 1;
        for (int32_t k = 0; k < npoints; ++k) {
            if (!(gc[k] >= 0.0 && gc[k] <= 1.0)) throw Error(DCNS_ERR_ARG, "gc must be in [0, 1]");
            if (len1[k] < 1 || len2[k] < 1 || len1[k] > (1 << 24) || len2[k] > (1 << 24))
                throw Error(DCNS_ERR_ARG, "lengths must be in [1, 2^24]");
            const double g = gc[k], two32 = 4294967296.0;
            auto t = [&](double f) { return (uint32_t)std::min(4294967295.0, std::floor(f * two32)); };
            thr[3 * (size_t)k] = t(g / 2);
            thr[3 * (size_t)k + 1] = t(g);
            thr[3 * (size_t)k + 2] = t(g + (1 - g) / 2);
            maxl1 = std::max(maxl1, len1[k]);
        }
        dcns_params dp;
        dcns_default_params(&dp);
        dp.match = match;
        dp.mismatch = mismatch;
        dp.open = open;
        dp.ext = ext;
        const Params P = to_params(dp);
        const ScoreTab tab = make_score_tab(match, mismatch);
        auto up = [&](DBuf &d, const void *src, size_t bytes) {
            d.ensure(std::max<size_t>(bytes, 64));
            hip_check(hipMemcpyAsync(d.p, src, bytes, hipMemcpyHostToDevice, c->stream), "upload");
        };
        up(c->rnd_thr, thr.data(), thr.size() * 4);
        up(c->rnd_len1, len1, (size_t)npoints * 4);
        up(c->rnd_len2, len2, (size_t)npoints * 4);
        up(c->rnd_seed, seed, (size_t)npoints * 8);
        // lane plan: waves of one loop shape, longest first; boundary rows
        // in launches of whole waves within 4 GiB.  Two pairs per lane in
        // packed f16 where the scoring allows (a pair whose maximum reaches
        // 2048 runs again on the int32 kernel), else one pair per lane.
        // The longest points' pairs go one per wave (stripes across the
        // lanes) when that shortens the launch.
        uint32_t mis4 = 0, diff = 0;
        const int32_t mode = c->tune.k1_lanes;
        const bool packed = mode != 1 && sw_random_packed(P, mis4, diff);
        const int32_t WS = packed ? 128 : 64;
        uint64_t key = 0xcbf29ce484222325ull;  // FNV-1a over the shapes and the launch shape
        auto mixin = [&](uint64_t v) {
            key ^= v;
            key *= 0x100000001b3ull;
        };
        mixin((uint64_t)npoints);
        mixin((uint64_t)pairs_per_point);
        mixin((uint64_t)(mode * 2 + (packed ? 1 : 0)));
        mixin((uint64_t)c->cu_count());
        for (int32_t k = 0; k < npoints; ++k) mixin(((uint64_t)(uint32_t)len1[k] << 32) | (uint32_t)len2[k]);
        if (!c->rnd_plan || c->rnd_plan->key != key) {
            auto rp = std::make_shared<RandPlan>();
            std::vector<uint8_t> wave_pts((size_t)npoints, 0);
            if (mode == 0) wave_pts = sw_wave_points(len1, len2, npoints, pairs_per_point, packed, c->cu_count());
            if (mode == 3)
                for (int32_t k = 0; k < npoints; ++k) wave_pts[(size_t)k] = len2[k] <= 64 * 32;
            sw_plan(len1, len2, npoints, pairs_per_point, rp->plan, WS, &wave_pts);
            for (int32_t k = 0; k < npoints; ++k)  // the pairs scored one per wave
                if (wave_pts[(size_t)k])
                    for (int32_t r = 0; r < pairs_per_point; ++r) rp->wids.push_back(k * pairs_per_point + r);
            c->rnd_plan.reset();
            up(c->rnd_order, rp->plan.order.data(), rp->plan.order.size() * 4);
            if (packed && !rp->wids.empty()) up(c->rnd_wlist, rp->wids.data(), rp->wids.size() * 4);
            rp->key = key;
            c->rnd_plan = rp;
        }
        const SwPlan &plan = c->rnd_plan->plan;
        const std::vector<int32_t> &wids = c->rnd_plan->wids;
        const int64_t total = (int64_t)npoints * pairs_per_point;
        c->rnd_out.ensure((size_t)total * 4 + 64);
        auto run_waves = [&](const std::vector<int32_t> &ids) {
            if (ids.empty()) return;
            up(c->rnd_wtmp, ids.data(), ids.size() * 4);
            c->launch("sw_random_wave", [&] {
                return launch_sw_random_wave(c->stream, c->rnd_thr.as<uint32_t>(), c->rnd_len1.as<int32_t>(),
                                             c->rnd_len2.as<int32_t>(), c->rnd_seed.as<uint64_t>(),
                                             c->rnd_wtmp.as<int32_t>(), pairs_per_point, (int64_t)ids.size(), P,
                                             tab, c->rnd_out.as<int32_t>());
            });
        };
        if (packed) {
            c->rnd_ovf.ensure((size_t)total * 4 + 64);
            c->ctr.ensure(sizeof(Counters) + 64);
            hip_check(hipMemsetAsync(&c->ctr.as<Counters>()->pending, 0, sizeof(int32_t), c->stream), "memset");
        } else {
            run_waves(wids);
        }
        const int64_t budget = ((int64_t)4 << 30) / 4;
        std::vector<int64_t> woff;
        // with h (packed), the first launch also takes the nwave pairs of
        // rnd_wlist in its first blocks
        auto run_plan = [&](const SwPlan &pl, const int32_t *order, bool h, int32_t nwave) {
            if (h && pl.waves.empty() && nwave > 0) {
                c->launch("sw_random", [&] {
                    return launch_sw_random_h(c->stream, c->rnd_thr.as<uint32_t>(), c->rnd_len1.as<int32_t>(),
                                              c->rnd_len2.as<int32_t>(), c->rnd_seed.as<uint64_t>(), order,
                                              pl.slot_order, pairs_per_point, 0, 0, P, tab, c->rnd_out.as<int32_t>(),
                                              nullptr, nullptr, c->rnd_ovf.as<int32_t>(),
                                              &c->ctr.as<Counters>()->pending, c->rnd_wlist.as<int32_t>(), nwave);
                });
                return;
            }
            for (int64_t w0 = 0; w0 < (int64_t)pl.waves.size();) {
                const int64_t w1 = sw_plan_chunk(pl, w0, budget, woff);
                c->sw_bound.ensure((size_t)woff.back() * 4 + 64);
                up(c->sw_woff, woff.data(), woff.size() * 8);
                const int64_t first = pl.wslots * w0, n = std::min<int64_t>(pl.wslots * w1, pl.total) - first;
                const int32_t nw = w0 == 0 ? nwave : 0;
                if (h)
                    c->launch("sw_random", [&] {
                        return launch_sw_random_h(c->stream, c->rnd_thr.as<uint32_t>(), c->rnd_len1.as<int32_t>(),
                                                  c->rnd_len2.as<int32_t>(), c->rnd_seed.as<uint64_t>(), order,
                                                  pl.slot_order, pairs_per_point, first, n, P, tab,
                                                  c->rnd_out.as<int32_t>(), c->sw_bound.as<int32_t>(),
                                                  c->sw_woff.as<int64_t>(), c->rnd_ovf.as<int32_t>(),
                                                  &c->ctr.as<Counters>()->pending, c->rnd_wlist.as<int32_t>(), nw);
                    });
                else
                    c->launch("sw_random", [&] {
                        return launch_sw_random(c->stream, c->rnd_thr.as<uint32_t>(), c->rnd_len1.as<int32_t>(),
                                                c->rnd_len2.as<int32_t>(), c->rnd_seed.as<uint64_t>(), order,
                                                pl.slot_order, pairs_per_point, first, n, P, tab,
                                                c->rnd_out.as<int32_t>(), c->sw_bound.as<int32_t>(),
                                                c->sw_woff.as<int64_t>());
                    });
                w0 = w1;
            }
        };
        run_plan(plan, c->rnd_order.as<int32_t>(), packed, packed ? (int32_t)wids.size() : 0);
        if (packed) {
            // pairs whose f16 maximum reached 2048: the int32 kernel, one pair per lane
            int32_t novf = 0;
            hip_check(hipMemcpyAsync(&novf, &c->ctr.as<Counters>()->pending, sizeof(int32_t), hipMemcpyDeviceToHost,
                                     c->stream), "overflow count");
            c->sync();
            if (novf > 0) {
                std::vector<int32_t> ids((size_t)novf), wide, narrow;
                hip_check(hipMemcpyAsync(ids.data(), c->rnd_ovf.p, (size_t)novf * 4, hipMemcpyDeviceToHost, c->stream),
                          "overflow list");
                c->sync();
                // the wave kernel where the query fits its lanes, else int32 lanes
                for (int32_t q : ids) (len2[q / pairs_per_point] <= 64 * 32 ? wide : narrow).push_back(q);
                run_waves(wide);
                if (!narrow.empty()) {
                    SwPlan re;
                    sw_plan_pairs(len1, len2, pairs_per_point, narrow, re);
                    up(c->rnd_ovf, re.order.data(), re.order.size() * 4);
                    run_plan(re, c->rnd_ovf.as<int32_t>(), false, 0);
                }
            }
        }
        hip_check(hipMemcpyAsync(out, c->rnd_out.p, (size_t)total * 4, hipMemcpyDeviceToHost, c->stream), "download");
        c->sync();
    });
}

int dcns_xdrop_extend(dcns_ctx *c, const dcns_batch *b, const dcns_params *p, const int32_t *pair_of_seed,
                      const int32_t *seeds, int32_t n, int32_t *out_ints, uint32_t *cigar, int64_t *cig_off,
                      int64_t cig_cap) {
    if (!c) return DCNS_ERR_ARG;
    return guarded(c, [&] {
        check_batch(b);
        if (!p) throw Error(DCNS_ERR_ARG, "params must be non-null");
        validate_params(*p);
        if (n < 0 || (n > 0 && (!pair_of_seed || !seeds || !out_ints || !cig_off)))
            throw Error(DCNS_ERR_ARG, "invalid seed arrays");
        hip_check(hipSetDevice(c->device), "hipSetDevice");
        Params P = to_params(*p);
        c->staged.reset();
        Staged S = stage_batch(c, *b, P);
        std::vector<XTask> tasks((size_t)n);
        for (int32_t i = 0; i < n; ++i) {
            const int32_t pr = pair_of_seed[i];
            if (pr < 0 || pr >= b->n) throw Error(DCNS_ERR_ARG, "pair index out of range");
            const int32_t s1 = seeds[2 * i], s2 = seeds[2 * i + 1];
            if (s1 < 0 || s1 > b->len1[pr] || s2 < 0 || s2 > b->len2[pr])
                throw Error(DCNS_ERR_ARG, "seed start outside its sequences");
            tasks[(size_t)i] = XTask{pr, s1, s2, 0};
        }
        std::vector<XResult> xr;
        int64_t cells = 0;
        xdrop_extend(c, S, P, tasks, xr, cells);
        int64_t o = 0;
        for (int32_t i = 0; i < n; ++i) {
            const XResult &r = xr[(size_t)i];
            int32_t *q = out_ints + 6 * i;
            q[0] = r.score;
            q[1] = r.start1;
            q[2] = r.end1;
            q[3] = r.start2;
            q[4] = r.end2;
            q[5] = r.ok ? 1 : 0;
            cig_off[i] = o;
            if (o + r.ncig > cig_cap) throw Error(DCNS_ERR_CAPACITY, "cigar capacity");
            if (r.ncig > 0) std::memcpy(cigar + o, r.cig, (size_t)r.ncig * 4);
            o += r.ncig;
        }
        cig_off[n] = o;
        c->stats = dcns_stats{};
        c->stats.xdrop_cells = cells;
    });
}

int dcns_seed_sets(dcns_ctx *c, const dcns_batch *b, const dcns_params *p, int32_t linked, int32_t *counts,
                   int32_t *seeds, int64_t cap) {
    if (!c) return DCNS_ERR_ARG;
    return guarded(c, [&] {
        check_batch(b);
        if (!p) throw Error(DCNS_ERR_ARG, "params must be non-null");
        validate_params(*p);
        hip_check(hipSetDevice(c->device), "hipSetDevice");
        Params P = to_params(*p);
        c->staged.reset();
        Staged S = stage_batch(c, *b, P);
        SeedResult sr;
        seed_sets(c, S, P, !linked, sr);
        const auto &sets = linked ? sr.linked : sr.unlinked;
        int64_t o = 0;
        bool fits = true;
        for (int32_t i = 0; i < b->n; ++i) {
            if (counts) counts[i] = (int32_t)sets[(size_t)i].size();
            for (const Seed &s : sets[(size_t)i]) {
                if (o < cap && seeds) {
                    seeds[4 * o] = s.getStart1();
                    seeds[4 * o + 1] = s.getEnd1();
                    seeds[4 * o + 2] = s.getStart2();
                    seeds[4 * o + 3] = s.getEnd2();
                } else {
                    fits = false;
                }
                ++o;
            }
        }
        c->stats = dcns_stats{};
        c->stats.scan_cells = sr.scan_cells;
        c->stats.gpu_scan_cells = sr.gpu_cells;
        c->stats.levels = sr.levels;
        if (!fits) throw Error(DCNS_ERR_CAPACITY, "seed capacity");
    });
}

int dcns_find_similar_batch(dcns_ctx *c, const dcns_batch *b, const dcns_params *p, dcns_frag_set **out) {
    if (!c || !out) return DCNS_ERR_ARG;
    *out = nullptr;
    dcns_frag_set *fs = new (std::nothrow) dcns_frag_set();
    if (!fs) return DCNS_ERR_INTERNAL;
    int rc = guarded(c, [&] {
        check_batch(b);
        if (!p) throw Error(DCNS_ERR_ARG, "params must be non-null");
        hip_check(hipSetDevice(c->device), "hipSetDevice");
        c->staged.reset();
        find_similar(c, *b, *p, fs->fs);
    });
    if (rc != DCNS_OK) {
        delete fs;
        return rc;
    }
    *out = fs;
    return DCNS_OK;
}

int dcns_find_similar_2gaps_batch(dcns_ctx *c, const dcns_batch *b, const dcns_params *p, dcns_frag_set **out) {
    if (!c || !out) return DCNS_ERR_ARG;
    *out = nullptr;
    dcns_frag_set *fs = new (std::nothrow) dcns_frag_set();
    if (!fs) return DCNS_ERR_INTERNAL;
    int rc = guarded(c, [&] {
        check_batch(b);
        if (!p) throw Error(DCNS_ERR_ARG, "params must be non-null");
        hip_check(hipSetDevice(c->device), "hipSetDevice");
        c->staged.reset();
        find_similar_2gaps(c, *b, *p, fs->fs);
    });
    if (rc != DCNS_OK) {
        delete fs;
        return rc;
    }
    *out = fs;
    return DCNS_OK;
}

}  // extern "C"

namespace {

std::vector<Profile> profiles_of(const dcns_profile *prof, int32_t nprof) {
    if (!prof || nprof < 1) throw Error(DCNS_ERR_ARG, "score profiles are required");
    std::vector<Profile> v((size_t)nprof);
    for (int32_t i = 0; i < nprof; ++i) {
        const dcns_profile &q = prof[i];
        if (q.category < INT16_MIN || q.category > INT16_MAX) throw Error(DCNS_ERR_ARG, "category outside int16");
        v[(size_t)i] = Profile{(int16_t)q.category, q.match, q.mismatch, q.open1, q.ext1, q.open2, q.ext2, q.zdrop};
    }
    return v;
}

// Task-level SemiGlobal (unweighted when W is null) on a staged batch.
void semiglobal_tasks(dcns_ctx *c, const dcns_batch *b, const Params &P, const SgWeights *W,
                      const int32_t *pair_of_task, const int32_t *anchors, const int32_t *reverse, int32_t n,
                      int32_t with_cigar, int32_t *out3, uint32_t *cigar, int64_t *cig_off, int64_t cig_cap) {
    if (n < 0 || (n > 0 && (!pair_of_task || !anchors || !out3)) || (with_cigar && n > 0 && !cig_off))
        throw Error(DCNS_ERR_ARG, "invalid task arrays");
    std::vector<SgTask> tasks((size_t)n);
    for (int32_t i = 0; i < n; ++i) {
        const int32_t pr = pair_of_task[i];
        if (pr < 0 || pr >= b->n) throw Error(DCNS_ERR_ARG, "pair index out of range");
        const int32_t a1 = anchors[2 * i], a2 = anchors[2 * i + 1], L1 = b->len1[pr], L2 = b->len2[pr];
        const bool rev = reverse && reverse[i];
        SgTask t{};
        t.pair = pr;
        t.rev = rev ? 1 : 0;
        t.a1 = a1;
        t.a2 = a2;
        if (rev) {
            if (a1 < 0 || a1 >= L1 || a2 < 0 || a2 >= L2) throw Error(DCNS_ERR_ARG, "reverse anchor outside its sequences");
            t.R = a1 + 1;
            t.C = a2 + 1;
        } else {
            if (a1 < 0 || a1 > L1 || a2 < 0 || a2 > L2) throw Error(DCNS_ERR_ARG, "anchor outside its sequences");
            t.R = L1 - a1;
            t.C = L2 - a2;
        }
        tasks[(size_t)i] = t;
    }
    std::vector<SgResult> res;
    int64_t cells = 0;
    semiglobal_batch(c, P, tasks, with_cigar != 0, res, cells, W);
    int64_t o = 0;
    for (int32_t i = 0; i < n; ++i) {
        const SgResult &r = res[(size_t)i];
        out3[3 * i] = r.max;
        out3[3 * i + 1] = r.end1;
        out3[3 * i + 2] = r.end2;
        if (with_cigar) {
            cig_off[i] = o;
            if (o + (int64_t)r.cigar.size() > cig_cap) throw Error(DCNS_ERR_CAPACITY, "cigar capacity");
            if (!r.cigar.empty()) std::memcpy(cigar + o, r.cigar.data(), r.cigar.size() * 4);
            o += (int64_t)r.cigar.size();
        }
    }
    if (with_cigar) cig_off[n] = o;
    c->stats = dcns_stats{};
    c->stats.semiglobal_cells = cells;
}

}  // namespace

extern "C" {

int dcns_semiglobal_batch(dcns_ctx *c, const dcns_batch *b, const dcns_params *p, const int32_t *pair_of_task,
                          const int32_t *anchors, const int32_t *reverse, int32_t n, int32_t with_cigar,
                          int32_t *out3, uint32_t *cigar, int64_t *cig_off, int64_t cig_cap) {
    if (!c) return DCNS_ERR_ARG;
    return guarded(c, [&] {
        check_batch(b);
        if (!p) throw Error(DCNS_ERR_ARG, "params must be non-null");
        validate_params(*p);
        if (p->open2 > 0 || p->ext2 > 0 || p->zdrop < 0 || p->bandwidth < 2)
            throw Error(DCNS_ERR_ARG, "open2/ext2 must be <= 0, zdrop >= 0, bandwidth >= 2");
        hip_check(hipSetDevice(c->device), "hipSetDevice");
        Params P = to_params(*p);
        c->staged.reset();
        Staged S = stage_batch(c, *b, P);
        semiglobal_tasks(c, b, P, nullptr, pair_of_task, anchors, reverse, n, with_cigar, out3, cigar, cig_off,
                         cig_cap);
    });
}

int dcns_semiglobal_weighted_batch(dcns_ctx *c, const dcns_batch *b, const dcns_params *p, const dcns_profile *prof,
                                   int32_t nprof, const int16_t *cat, const int64_t *cat_off, int32_t two_piece,
                                   const int32_t *pair_of_task, const int32_t *anchors, const int32_t *reverse,
                                   int32_t n, int32_t with_cigar, int32_t *out3, uint32_t *cigar, int64_t *cig_off,
                                   int64_t cig_cap) {
    if (!c) return DCNS_ERR_ARG;
    return guarded(c, [&] {
        check_batch(b);
        if (!p) throw Error(DCNS_ERR_ARG, "params must be non-null");
        validate_params(*p);
        if (p->bandwidth < 1) throw Error(DCNS_ERR_ARG, "bandwidth must be >= 1");
        const std::vector<Profile> pv = profiles_of(prof, nprof);
        hip_check(hipSetDevice(c->device), "hipSetDevice");
        Params P = to_params(*p);
        c->staged.reset();
        Staged S = stage_batch(c, *b, P);
        const SgWeights W = stage_weights(c, *b, pv, PairCategories{cat, cat_off}, two_piece ? 1 : 2);
        semiglobal_tasks(c, b, P, &W, pair_of_task, anchors, reverse, n, with_cigar, out3, cigar, cig_off, cig_cap);
    });
}

int dcns_find_similar_weighted_batch(dcns_ctx *c, const dcns_batch *b, const dcns_params *p,
                                     const dcns_profile *prof, int32_t nprof, const int16_t *cat,
                                     const int64_t *cat_off, int32_t two_piece, dcns_frag_set **out) {
    if (!c || !out) return DCNS_ERR_ARG;
    *out = nullptr;
    dcns_frag_set *fs = new (std::nothrow) dcns_frag_set();
    if (!fs) return DCNS_ERR_INTERNAL;
    int rc = guarded(c, [&] {
        check_batch(b);
        if (!p) throw Error(DCNS_ERR_ARG, "params must be non-null");
        const std::vector<Profile> pv = profiles_of(prof, nprof);
        hip_check(hipSetDevice(c->device), "hipSetDevice");
        c->staged.reset();
        find_similar_weighted(c, *b, *p, pv, PairCategories{cat, cat_off}, two_piece != 0, fs->fs);
    });
    if (rc != DCNS_OK) {
        delete fs;
        return rc;
    }
    *out = fs;
    return DCNS_OK;
}

int dcns_find_similar_seeded_batch(dcns_ctx *c, const dcns_batch *b, const dcns_params *p, const int32_t *seeds,
                                   const int64_t *seed_first, int32_t max_align_length, dcns_frag_set **out) {
    if (!c || !out) return DCNS_ERR_ARG;
    *out = nullptr;
    dcns_frag_set *fs = new (std::nothrow) dcns_frag_set();
    if (!fs) return DCNS_ERR_INTERNAL;
    int rc = guarded(c, [&] {
        check_batch(b);
        if (!p) throw Error(DCNS_ERR_ARG, "params must be non-null");
        if (b->n > 0 && !seed_first) throw Error(DCNS_ERR_ARG, "seed_first must be non-null");
        std::vector<std::vector<Seed>> sd((size_t)b->n);
        for (int32_t i = 0; i < b->n; ++i) {
            if (seed_first[i + 1] < seed_first[i] || (seed_first[i + 1] > seed_first[i] && !seeds))
                throw Error(DCNS_ERR_ARG, "invalid seed ranges");
            for (int64_t k = seed_first[i]; k < seed_first[i + 1]; ++k)
                sd[(size_t)i].push_back(Seed(seeds[4 * k], seeds[4 * k + 1], seeds[4 * k + 2], seeds[4 * k + 3]));
        }
        hip_check(hipSetDevice(c->device), "hipSetDevice");
        c->staged.reset();
        find_similar_seeded(c, *b, *p, sd, max_align_length, fs->fs);
    });
    if (rc != DCNS_OK) {
        delete fs;
        return rc;
    }
    *out = fs;
    return DCNS_OK;
}

int dcns_link_seeds(const int32_t *seeds, int64_t n, int32_t insertion_order, int32_t *out, int64_t cap,
                    int64_t *n_out) {
    return guarded(nullptr, [&] {
        if (n < 0 || (n > 0 && !seeds) || !n_out || cap < 0 || (cap > 0 && !out))
            throw Error(DCNS_ERR_ARG, "invalid arguments");
        std::set<Seed> set;
        for (int64_t k = 0; k < n; ++k)
            set.insert(Seed(seeds[4 * k], seeds[4 * k + 1], seeds[4 * k + 2], seeds[4 * k + 3]));
        std::vector<Seed> list = link_seed_list(std::vector<Seed>(set.begin(), set.end()));
        if (!insertion_order) {
            set.clear();
            insert_in_start1_order(set, list);
            list.assign(set.begin(), set.end());
        }
        *n_out = (int64_t)list.size();
        if (*n_out > cap) throw Error(DCNS_ERR_CAPACITY, "more linked seeds than cap");
        for (size_t o = 0; o < list.size(); ++o) {
            out[4 * o] = list[o].getStart1();
            out[4 * o + 1] = list[o].getEnd1();
            out[4 * o + 2] = list[o].getStart2();
            out[4 * o + 3] = list[o].getEnd2();
        }
    });
}

int dcns_read_score_folder(const char *folder, dcns_profile *out, int32_t cap, int32_t *n) {
    return guarded(nullptr, [&] {
        if (!folder || !n || cap < 0 || (cap > 0 && !out)) throw Error(DCNS_ERR_ARG, "invalid arguments");
        const std::vector<Profile> v = read_score_folder(folder);
        *n = (int32_t)v.size();
        if ((int32_t)v.size() > cap) throw Error(DCNS_ERR_CAPACITY, "more profiles than cap");
        for (size_t i = 0; i < v.size(); ++i)
            out[i] = dcns_profile{v[i].category, v[i].match, v[i].mismatch, v[i].open1, v[i].ext1, v[i].open2,
                                  v[i].ext2, v[i].zdrop};
    });
}

int dcns_gff_categories(const char *gff_path, const char *chr, int64_t chr_size, int64_t from, int64_t count,
                        int16_t *out) {
    return guarded(nullptr, [&] {
        if (!gff_path || !chr || chr_size < 0 || count < 0 || (count > 0 && !out))
            throw Error(DCNS_ERR_ARG, "invalid arguments");
        const GffAnnotation g = read_gff(gff_path);
        gff_categories(g, chr, chr_size, from, count, out);
    });
}

int dcns_stage(dcns_ctx *c, const dcns_batch *b, const dcns_params *p) {
    if (!c) return DCNS_ERR_ARG;
    return guarded(c, [&] {
        check_batch(b);
        if (!p) throw Error(DCNS_ERR_ARG, "params must be non-null");
        validate_params(*p);
        hip_check(hipSetDevice(c->device), "hipSetDevice");
        c->staged = std::make_shared<Staged>(stage_batch(c, *b, to_params(*p)));
        c->sync();
    });
}

int dcns_find_similar_staged(dcns_ctx *c, const dcns_params *p, dcns_frag_set **out) {
    if (!c || !out) return DCNS_ERR_ARG;
    *out = nullptr;
    dcns_frag_set *fs = new (std::nothrow) dcns_frag_set();
    if (!fs) return DCNS_ERR_INTERNAL;
    int rc = guarded(c, [&] {
        if (!c->staged) throw Error(DCNS_ERR_ARG, "no staged batch: call dcns_stage first");
        if (!p) throw Error(DCNS_ERR_ARG, "params must be non-null");
        hip_check(hipSetDevice(c->device), "hipSetDevice");
        find_similar_staged(c, *c->staged, *p, fs->fs);
    });
    if (rc != DCNS_OK) {
        delete fs;
        return rc;
    }
    *out = fs;
    return DCNS_OK;
}

int64_t dcns_frag_set_total(const dcns_frag_set *fs) { return fs ? (int64_t)fs->fs.frags.size() : -1; }

int dcns_frag_set_pair(const dcns_frag_set *fs, int32_t pair, int64_t *first, int64_t *count, int64_t *seeds_size) {
    if (!fs || pair < 0 || (size_t)pair >= fs->fs.first.size()) return DCNS_ERR_ARG;
    if (first) *first = fs->fs.first[(size_t)pair];
    if (count) *count = fs->fs.count[(size_t)pair];
    if (seeds_size) *seeds_size = fs->fs.seeds_size[(size_t)pair];
    return DCNS_OK;
}

int dcns_frag_set_get(const dcns_frag_set *fs, int64_t idx, int32_t *ints5, double *dbls2, uint32_t *cigar,
                      int32_t cig_cap, int32_t *ncig) {
    if (!fs || idx < 0 || (size_t)idx >= fs->fs.frags.size()) return DCNS_ERR_ARG;
    const PairedSimilarFragment &f = fs->fs.frags[(size_t)idx];
    if (ints5) {
        ints5[0] = (int32_t)f.getStart1();
        ints5[1] = (int32_t)f.getEnd1();
        ints5[2] = (int32_t)f.getStart2();
        ints5[3] = (int32_t)f.getEnd2();
        ints5[4] = (int32_t)f.getScore();
    }
    if (dbls2) {
        dbls2[0] = f.getPValue();
        dbls2[1] = f.getEValue();
    }
    const auto &cg = f.getCigar();
    if (ncig) *ncig = (int32_t)cg.size();
    if (cigar && cig_cap > 0)
        std::memcpy(cigar, cg.data(), std::min<size_t>((size_t)cig_cap, cg.size()) * sizeof(uint32_t));
    return (int64_t)cg.size() > cig_cap && cigar ? DCNS_ERR_CAPACITY : DCNS_OK;
}

void dcns_frag_set_free(dcns_frag_set *fs) { delete fs; }

int dcns_kmer_set_read(const char *path, int32_t freq, dcns_kmer_set **out) {
    if (!out) return DCNS_ERR_ARG;
    *out = nullptr;
    dcns_kmer_set *s = nullptr;
    const int rc = guarded(nullptr, [&] {
        if (!path && freq > 0) throw Error(DCNS_ERR_ARG, "k-mer dump path must be non-null");
        s = new dcns_kmer_set{read_kmer_dump(path ? path : "", freq)};
    });
    if (rc == DCNS_OK) *out = s;
    return rc;
}

int dcns_kmer_set_info(const dcns_kmer_set *s, int32_t *k, int64_t *members, int64_t *records) {
    if (!s) return DCNS_ERR_ARG;
    if (k) *k = s->ks.k;
    if (members) *members = (int64_t)s->ks.codes.size() + s->ks.n_general();
    if (records) *records = s->ks.records;
    return DCNS_OK;
}

int dcns_kmer_set_get(const dcns_kmer_set *s, char *out, int64_t cap) {
    if (!s || cap < 0 || (cap > 0 && !out)) return DCNS_ERR_ARG;
    const int64_t k = s->ks.k, n = (int64_t)s->ks.codes.size() + s->ks.n_general();
    if (n * k > cap) return DCNS_ERR_CAPACITY;
    static const char letter[4] = {'A', 'C', 'T', 'G'};  // inverse of (b >> 1) & 3
    int64_t o = 0;
    for (uint64_t code : s->ks.codes) {
        for (int64_t j = 0; j < k; ++j) out[o + j] = letter[code & 3], code >>= 2;  // first base lowest
        o += k;
    }
    if (!s->ks.general.empty()) std::memcpy(out + o, s->ks.general.data(), s->ks.general.size());
    return DCNS_OK;
}

void dcns_kmer_set_free(dcns_kmer_set *s) { delete s; }

int dcns_cds_mask_create(const char *const *names, const int64_t *sizes, int32_t n, dcns_cds_mask **out) {
    if (!out) return DCNS_ERR_ARG;
    *out = nullptr;
    dcns_cds_mask *m = nullptr;
    const int rc = guarded(nullptr, [&] {
        if (n < 0 || (n > 0 && (!names || !sizes))) throw Error(DCNS_ERR_ARG, "invalid genome arrays");
        std::vector<std::string> nm;
        for (int32_t i = 0; i < n; ++i) {
            if (!names[i]) throw Error(DCNS_ERR_ARG, "null sequence name");
            nm.emplace_back(names[i]);
        }
        m = new dcns_cds_mask{CdsMask(nm, std::vector<int64_t>(sizes, sizes + n))};
    });
    if (rc == DCNS_OK) *out = m;
    return rc;
}

int dcns_cds_mask_gff(dcns_cds_mask *m, const char *gff_path, int32_t gene) {
    return guarded(nullptr, [&] {
        if (!m || !gff_path) throw Error(DCNS_ERR_ARG, "invalid arguments");
        m->m.gff(gff_path, gene != 0);
    });
}

int dcns_cds_mask_sam(dcns_cds_mask *m, const char *sam_path, const char *cds_fasta, double similarity,
                      int32_t *rebuilt) {
    return guarded(nullptr, [&] {
        if (!m || !sam_path || !cds_fasta) throw Error(DCNS_ERR_ARG, "invalid arguments");
        if (!std::ifstream(cds_fasta).good()) throw Error(DCNS_ERR_ARG, std::string("error in opening fasta file ") + cds_fasta);
        std::map<std::string, std::string> seqs;
        std::vector<std::string> names;
        cli::read_plain_fasta(cds_fasta, seqs, names);
        std::map<std::string, int64_t> sizes;
        for (const auto &kv : seqs) sizes[kv.first] = (int64_t)kv.second.size();
        const bool r = m->m.sam(sam_path, sizes, similarity);
        if (rebuilt) *rebuilt = r ? 1 : 0;
    });
}

int dcns_cds_mask_get(const dcns_cds_mask *m, int32_t s, uint8_t *out, int64_t cap) {
    return guarded(nullptr, [&] {
        if (!m || s < 0 || s >= m->m.size()) throw Error(DCNS_ERR_ARG, "invalid arguments");
        const int64_t n = m->m.seq_size(s);
        if (cap < n || (n > 0 && !out)) throw Error(DCNS_ERR_CAPACITY, "mask buffer too small");
        std::memset(out, 0, (size_t)n);
        for (const auto &r : m->m.ranges(s)) std::memset(out + r.first, 1, (size_t)(r.second - r.first));
    });
}

void dcns_cds_mask_free(dcns_cds_mask *m) { delete m; }

int dcns_mask_stage(dcns_ctx *c, const dcns_kmer_set *kmers, const char *seq, const int64_t *offsets, int32_t nseq) {
    if (!c) return DCNS_ERR_ARG;
    return guarded(c, [&] {
        hip_check(hipSetDevice(c->device), "hipSetDevice");
        static const KmerSet none;  // no dump: k = 10, no members
        mask_stage(c, kmers ? kmers->ks : none, (const uint8_t *)seq, offsets, nseq);
    });
}

int64_t dcns_mask_body_size(int64_t len, int32_t width) { return len < 0 ? -1 : mask_body_size(len, width); }

int dcns_mask_run(dcns_ctx *c, const dcns_cds_mask *cds, char alt, int32_t soft, int32_t width, uint8_t *mask_out,
                  char *body_out) {
    if (!c) return DCNS_ERR_ARG;
    return guarded(c, [&] {
        hip_check(hipSetDevice(c->device), "hipSetDevice");
        mask_run(c, cds ? &cds->m : nullptr, (uint8_t)alt, soft != 0, width, mask_out, (uint8_t *)body_out);
    });
}

}  // extern "C"
