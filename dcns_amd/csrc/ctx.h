// ctx.h -- device context, grow-only device buffers and launch timing.
#pragma once
#include <hip/hip_runtime.h>

#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include <algorithm>
#include <atomic>
#include <thread>

#include "../../include/dcns.h"
#include "dcns_types.h"

namespace dcns {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) throw Error(DCNS_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

// Device buffer that only grows.  keep=true preserves the old contents.
struct DBuf {
    void *p = nullptr;
    size_t cap = 0;
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    ~DBuf() {
        if (p) (void)hipFree(p);
    }
    void ensure(size_t bytes, hipStream_t st = nullptr, bool keep = false, size_t used = 0) {
        if (bytes <= cap) return;
        size_t n = bytes + bytes / 4 + 4096;
        void *q = nullptr;
        hip_check(hipMalloc(&q, n), "hipMalloc");
        if (keep && p && used) {
            hip_check(hipMemcpyAsync(q, p, used, hipMemcpyDeviceToDevice, st), "hipMemcpyAsync");
            hip_check(hipStreamSynchronize(st), "hipStreamSynchronize");
        }
        if (p) hip_check(hipFree(p), "hipFree");
        p = q;
        cap = n;
    }
    template <class T>
    T *as() const {
        return static_cast<T *>(p);
    }
};

// Pinned host buffer (hipHostMalloc) for device->host result transfers.
struct HBuf {
    void *p = nullptr;
    size_t cap = 0;
    HBuf() = default;
    HBuf(const HBuf &) = delete;
    HBuf &operator=(const HBuf &) = delete;
    ~HBuf() {
        if (p) (void)hipHostFree(p);
    }
    void ensure(size_t bytes) {
        if (bytes <= cap) return;
        if (p) hip_check(hipHostFree(p), "hipHostFree");
        p = nullptr;
        const size_t n = bytes + bytes / 4 + 4096;
        hip_check(hipHostMalloc(&p, n, hipHostMallocDefault), "hipHostMalloc");
        cap = n;
    }
    template <class T>
    T *as() const {
        return static_cast<T *>(p);
    }
};

// Runs fn(i) for i in [0, n) on up to `threads` host threads (dynamic
// schedule over chunks of consecutive indices, about 64 per thread, so that
// cheap bodies do not serialise on the shared counter).
template <class F>
void parallel_for(int64_t n, F &&fn, int threads = 0);

struct KernelTimer {
    double ms = 0;
    int64_t launches = 0;
};

template <class F>
void parallel_for(int64_t n, F &&fn, int threads) {
    if (threads <= 0) threads = (int)std::min<unsigned>(std::max(1u, std::thread::hardware_concurrency()), 16u);
    threads = (int)std::min<int64_t>(threads, n);
    if (threads <= 1) {
        for (int64_t i = 0; i < n; ++i) fn(i);
        return;
    }
    const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(n / ((int64_t)threads * 64), 4096));
    std::atomic<int64_t> next(0);
    auto work = [&]() {
        for (int64_t a = next.fetch_add(chunk); a < n; a = next.fetch_add(chunk)) {
            const int64_t b = std::min(n, a + chunk);
            for (int64_t i = a; i < b; ++i) fn(i);
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; ++t) pool.emplace_back(work);
    work();
    for (auto &t : pool) t.join();
}

}  // namespace dcns

namespace dcns {
struct Staged;
struct MaskStaged;
struct SeedPhase;
struct RandPlan;
}

struct dcns_ctx {
    int device = 0;
    std::shared_ptr<dcns::Staged> staged;  // sequences kept resident by dcns_stage
    std::shared_ptr<dcns::MaskStaged> mstaged;  // genome and k-mer tables kept resident by dcns_mask_stage
    std::shared_ptr<dcns::SeedPhase> seedphase;  // recursion pool sizes and the captured level graph
    std::shared_ptr<dcns::RandPlan> rnd_plan;     // the last random-pair launch plan (reused for the same lengths)
    hipStream_t stream = nullptr;
    bool timing = false;
    std::string err;
    dcns_stats stats{};
    dcns_tuning tune{};  // launch shapes (0 = default), see dcns_set_tuning
    std::map<std::string, dcns::KernelTimer> timers;
    std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;  // recorded, not yet read
    std::vector<hipEvent_t> event_pool;
    // reusable device buffers
    dcns::DBuf seq, pairs, wtasks, l0wins, pow, traj, tcount, hitlist, ctr, nodes, roots, tasks_a, tasks_b;
    dcns::DBuf lastrow;  // per window: the last level-0 row whose maximum reaches min_score
    // level-0 strong rows (sink chunks, then the sorted index) for the rescans' level-0 tails
    dcns::DBuf strong_win, strong_rv, strong_fill, strong_rv2, strong_off, strong_seg, strong_ng,
        strong_tmp, strong_total, strong_key, strong_key2, strong_fill64;
    dcns::DBuf seg_tasks, seg_groups, seg_traj, seg_tcount;  // level 0 in row segments
    dcns::DBuf xtasks, xout, tb_off, tb, cigar, cig_host_dummy;
    dcns::DBuf sw_off1, sw_len1, sw_off2, sw_len2, sw_order, sw_out, sw_bound, sw_woff;
    dcns::DBuf bas_cig, bas_st;  // SmithWaterman_bas walks: cigar entries, counts and status
    dcns::DBuf sg_tasks, sg_out, sg_scr, sg_tb, sg_cig, sg_strips, sg_order, sg_sync;  // two-piece semi-global tasks
    dcns::DBuf sg_wcat, sg_wcat_off, sg_wtab;  // weighted DP: per-position profile indices and the profiles
    dcns::DBuf xw_scr, xw_off, xw_wmax;                  // x-drop tasks with bands past the wave kernel's
    dcns::DBuf order_hist, task_order;                   // scan tasks ordered by row count
    dcns::DBuf xorder;                                   // x-drop pass-2 order
    dcns::DBuf xlist, xlist2;                            // x-drop tasks handed to the next wider kernel
    dcns::DBuf cig_scan, cig_packed;                     // cigar packing: counts / offsets / scan storage, runs
    dcns::DBuf rnd_thr, rnd_len1, rnd_len2, rnd_seed, rnd_out, rnd_order;  // random fragment pairs generated on the device
    dcns::DBuf rnd_ovf;    // random pairs whose packed maximum left f16's exact integers
    dcns::DBuf rnd_wlist;  // random pairs scored one per wave (the cached plan's)
    dcns::DBuf rnd_wtmp;   // random pairs scored one per wave (int32 launches, re-runs)
    dcns::DBuf mk_seq, mk_off, mk_keys, mk_desc, mk_gbytes, mk_ghash, mk_gtab;  // maskGenome: genome, k-mer set
    dcns::DBuf mk_mask, mk_ranges, mk_boff, mk_body;                          // maskGenome: mask, CDS, FASTA bodies
    dcns::HBuf h_hits, h_roots, h_nodes, h_xout, h_cigar;  // pinned download buffers

    int32_t cus = 0;  // compute units of the device (cached)
    int32_t cu_count() {
        if (!cus) {
            hipDeviceProp_t prop;
            dcns::hip_check(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
            cus = prop.multiProcessorCount;
        }
        return cus;
    }

    hipEvent_t take_event() {
        if (!event_pool.empty()) {
            hipEvent_t e = event_pool.back();
            event_pool.pop_back();
            return e;
        }
        hipEvent_t e;
        dcns::hip_check(hipEventCreate(&e), "hipEventCreate");
        return e;
    }
    // Wrap a launch: records HIP events on the context stream when timing is on.
    template <class F>
    void launch(const char *name, F &&f) {
        hipEvent_t a = nullptr, b = nullptr;
        if (timing) {
            a = take_event();
            b = take_event();
            dcns::hip_check(hipEventRecord(a, stream), "hipEventRecord");
        }
        dcns::hip_check(f(), name);
        if (timing) {
            dcns::hip_check(hipEventRecord(b, stream), "hipEventRecord");
            pending.push_back({name, {a, b}});
        }
    }
    // Reads the events of finished launches (call after a stream sync).
    void collect() {
        for (auto &pe : pending) {
            float ms = 0;
            dcns::hip_check(hipEventElapsedTime(&ms, pe.second.first, pe.second.second), "hipEventElapsedTime");
            auto &t = timers[pe.first];
            t.ms += ms;
            t.launches += 1;
            event_pool.push_back(pe.second.first);
            event_pool.push_back(pe.second.second);
        }
        pending.clear();
    }
    void sync() {
        dcns::hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        collect();
    }
};
