// pipeline.cpp -- host orchestration of dCNS's hot path over the gfx950
// kernels: staging, the recursion level loop, exact replay of the reference's
// std::set<Seed> insertion order, link_seeds, x-drop extension and fragment
// dedup (findSimilarFragmentsForPairedSequence4.cpp:433-524).
#include "pipeline.h"

#include <immintrin.h>

#include <algorithm>
#include <unordered_map>
#include <chrono>
#include <cmath>
#include <functional>

#include "dp_common.h"
#include "kernels.h"

namespace dcns {

namespace {
double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

template <class T>
void upload(dcns_ctx *c, DBuf &buf, const std::vector<T> &v) {
    buf.ensure(std::max<size_t>(v.size() * sizeof(T), 64));
    if (!v.empty())
        hip_check(hipMemcpyAsync(buf.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, c->stream), "upload");
}

template <class T>
void download(dcns_ctx *c, std::vector<T> &v, const void *src, size_t n) {
    v.resize(n);
    if (n) hip_check(hipMemcpyAsync(v.data(), src, n * sizeof(T), hipMemcpyDeviceToHost, c->stream), "download");
}

Counters read_ctr(dcns_ctx *c) {
    Counters h;
    hip_check(hipMemcpyAsync(&h, c->ctr.p, sizeof(h), hipMemcpyDeviceToHost, c->stream), "read counters");
    c->sync();
    return h;
}

void write_ctr(dcns_ctx *c, const Counters &h) {
    hip_check(hipMemcpyAsync(c->ctr.p, &h, sizeof(h), hipMemcpyHostToDevice, c->stream), "write counters");
}

// Window schedule of overload 1 (findSimilarFragmentsForPairedSequence4.cpp:461-480):
// sp = 0, step, 2*step, ...; stop after the first window that ends past L2,
// or when sp passes L2.
int32_t window_count(int32_t L2, int32_t W, int32_t step) {
    int32_t n = 1, sp = 0;
    while (true) {
        sp += step;
        if (sp > L2) break;
        ++n;
        if (sp + W > L2) break;
    }
    return n;
}
}  // namespace

void validate_params(const dcns_params &p) {
    auto bad = [](const char *m) { throw Error(DCNS_ERR_ARG, m); };
    if (p.match < 1 || p.match > 127) bad("match must be in [1, 127]");
    if (p.mismatch > 0 || p.mismatch < -128) bad("mismatch must be in [-128, 0]");
    if (p.open > 0 || p.open < -128 || p.ext > 0 || p.ext < -128) bad("gap open/extend must be in [-128, 0]");
    if (p.window < 1 || p.window > kMaxWindow) bad("window must be in [1, 64]");
    if (p.step < 1) bad("step must be >= 1");
    if (p.min_score < 1) bad("min_score must be >= 1");
    if (p.w < 1 || 2 * p.w + 1 > 161) bad("x-drop band half-width must be in [1, 80]");
    if (p.xdrop < 0) bad("xdrop must be >= 0");
    if ((int64_t)p.window * p.match >= 32767) bad("window * match must stay below 32767");
}

Params to_params(const dcns_params &p) {
    Params P;
    P.match = p.match;
    P.mismatch = p.mismatch;
    P.open = p.open;
    P.ext = p.ext;
    P.open2 = p.open2;
    P.ext2 = p.ext2;
    P.zdrop = p.zdrop;
    P.bandwidth = p.bandwidth;
    P.window = p.window;
    P.step = p.step;
    P.min_score = p.min_score;
    P.w = p.w;
    P.xdrop = p.xdrop;
    P.min_ref = p.min_score / p.match;
    return P;
}

Staged stage_batch(dcns_ctx *c, const dcns_batch &b, const Params &P) {
    Staged S;
    S.pd.resize((size_t)b.n);
    int64_t total = 0;
    auto span = [](int64_t len) { return ((len + 64) + 63) & ~(int64_t)63; };  // 64-byte aligned, >= 64 bytes slack
    for (int32_t i = 0; i < b.n; ++i) {
        if (b.len1[i] < 0 || b.len2[i] < 0) throw Error(DCNS_ERR_ARG, "negative sequence length");
        total += span(b.len1[i]) + span(b.len2[i]);
    }
    // N (code 2) around every sequence, 64 bytes before the first one too, so
    // that a DP reading one base past either end of a sequence reads N
    std::vector<int8_t> host((size_t)total + 128, 2);
    int64_t o = 64, nw = 0;
    for (int32_t i = 0; i < b.n; ++i) {
        PairDesc &d = S.pd[(size_t)i];
        d.len1 = b.len1[i];
        d.len2 = b.len2[i];
        d.off1 = o;
        if (d.len1) std::memcpy(&host[(size_t)o], b.seqs + b.off1[i], (size_t)d.len1);
        o += span(d.len1);
        d.off2 = o;
        if (d.len2) std::memcpy(&host[(size_t)o], b.seqs + b.off2[i], (size_t)d.len2);
        o += span(d.len2);
        d.nwin = window_count(d.len2, P.window, P.step);
        if (nw + d.nwin > INT32_MAX / 2) throw Error(DCNS_ERR_ARG, "batch has too many windows");
        d.win_base = (int32_t)nw;
        nw += d.nwin;
    }
    S.nwin = nw;
    S.window = P.window;
    S.step = P.step;
    upload(c, c->seq, host);
    upload_layout(c, S);
    return S;
}

// Issue slots (wave instructions) of the two level-0 paths, measured on the
// kernels: a row of the 128-window lane kernel (~337 VALU, ~115 s_nop on its
// critical path) and a step of the one-window wave kernel.
constexpr double kLaneRowIssue = 340.0, kLaneRowLatency = 450.0;
constexpr double kWaveStepIssue = 35.0, kWaveStepLatency = 40.0;

// Level-0 wave tasks (128 windows of one pair per wave, longest reference
// first), window -> pair map and per-pair window-scan cells; uploaded once
// per staged batch and window layout.
//
// A lane task's rows are serial, so a batch with few pairs is held to its
// longest reference's rows while most of the chip idles.  The longest tasks
// move to the one-window-per-wave path (R + width - 1 steps per window, ~10x
// the issue slots) for as long as that shortens the estimated launch: the
// larger of the longest remaining chain and the total issue slots over the
// chip's SIMDs.  A full batch keeps every task on the lane kernel.
void upload_layout(dcns_ctx *c, Staged &S) {
    S.wt.clear();
    S.pair_of_window.assign((size_t)S.nwin, 0);
    S.level0_cells.assign(S.pd.size(), 0);
    for (size_t p = 0; p < S.pd.size(); ++p) {
        const PairDesc &d = S.pd[p];
        for (int32_t g = 0; g < d.nwin; g += 128) S.wt.push_back(make_int2((int)p, g));
        int64_t widths = 0;
        for (int32_t w = 0; w < d.nwin; ++w) {
            S.pair_of_window[(size_t)(d.win_base + w)] = (int32_t)p;
            widths += std::min(S.window, d.len2 - w * S.step);
        }
        S.level0_cells[p] = (int64_t)d.len1 * widths;
    }
    std::stable_sort(S.wt.begin(), S.wt.end(), [&](const int2 &a, const int2 &b) {
        return S.pd[(size_t)a.x].len1 > S.pd[(size_t)b.x].len1;
    });
    const size_t n = S.wt.size();
    auto task_windows = [&](size_t i) {
        return std::min(128, S.pd[(size_t)S.wt[i].x].nwin - S.wt[i].y);
    };
    const double simds = 4.0 * c->cu_count();
    double lane_total = 0;
    for (size_t i = 0; i < n; ++i) lane_total += (double)S.pd[(size_t)S.wt[i].x].len1 * kLaneRowIssue;
    auto lane_chain = [&](size_t i) { return i < n ? (double)S.pd[(size_t)S.wt[i].x].len1 * kLaneRowLatency : 0.0; };
    double best = std::max(lane_chain(0), lane_total / simds), wave_total = 0, wave_chain = 0;
    size_t moved = 0;
    for (size_t k = 1; k <= n; ++k) {
        const double rows = (double)S.pd[(size_t)S.wt[k - 1].x].len1 + S.window - 1;
        lane_total -= (double)S.pd[(size_t)S.wt[k - 1].x].len1 * kLaneRowIssue;
        wave_total += task_windows(k - 1) * rows * kWaveStepIssue;
        wave_chain = std::max(wave_chain, rows * kWaveStepLatency);
        const double t = std::max(std::max(lane_chain(k), wave_chain), (lane_total + wave_total) / simds);
        if (t < best) {
            best = t;
            moved = k;
        }
    }
    S.ww.clear();
    for (size_t i = 0; i < moved; ++i)
        for (int32_t w = 0; w < task_windows(i); ++w) S.ww.push_back(make_int2(S.wt[i].x, S.wt[i].y + w));
    S.wt.erase(S.wt.begin(), S.wt.begin() + (std::ptrdiff_t)moved);
    upload(c, c->pairs, S.pd);
    upload(c, c->wtasks, S.wt);
    upload(c, c->l0wins, S.ww);
    upload(c, c->pow, S.pair_of_window);
}

void relayout(dcns_ctx *c, Staged &S, const Params &P) {
    if (S.window == P.window && S.step == P.step) return;
    int64_t nw = 0;
    for (PairDesc &d : S.pd) {
        d.nwin = window_count(d.len2, P.window, P.step);
        if (nw + d.nwin > INT32_MAX / 2) throw Error(DCNS_ERR_ARG, "batch has too many windows");
        d.win_base = (int32_t)nw;
        nw += d.nwin;
    }
    S.nwin = nw;
    S.window = P.window;
    S.step = P.step;
    upload_layout(c, S);
}

// ---------------------------------------------------------------------------
// Seed scan: device level loop + host replay.
// ---------------------------------------------------------------------------
namespace {

struct Forest {
    int32_t NW = 0, hits = 0;
    const int32_t *hit_window = nullptr;  // global window id of hit k
    const Node *roots = nullptr;          // root node of hit k
    const Node *dyn = nullptr;            // nodes allocated during the recursion (id >= NW)
    const Node &child(int32_t id) const { return dyn[(size_t)(id - NW)]; }
};

// Launch shapes of the seed phase with the defaults filled in.
struct SeedTuning {
    int32_t level0_waves;  // > 0: persistent level-0 waves (total); 0: one workgroup per task
    int32_t wave_tasks, wave_rows, graph_levels, pool_min;
};

SeedTuning seed_tuning(dcns_ctx *c) {
    const dcns_tuning &t = c->tune;
    SeedTuning T;
    const int32_t mode = t.level0_mode ? t.level0_mode : 1;
    T.level0_waves = mode == 2 ? (t.level0_waves ? t.level0_waves : 8) * c->cu_count() : 0;
    T.wave_tasks = t.wave_tasks ? t.wave_tasks : 1024;
    T.wave_rows = t.wave_rows ? t.wave_rows : 4096;
    T.graph_levels = t.graph_levels ? std::max(t.graph_levels, 0) : 4;
    T.pool_min = t.pool_min;
    return T;
}

}  // namespace

// Seed-phase state of a context: the recursion pools' capacities (grown when
// a level overflows them, kept for the next batch with the same window
// count) and the captured graph of the recursion levels.
struct SeedPhase {
    int64_t nw = -1;
    int32_t pool_min = -1;
    int64_t scan_cap = 0, node_cap = 0, next_cap = 0;
    hipGraphExec_t exec = nullptr;
    std::vector<int64_t> key;   // every value the captured launches bake in
    std::vector<int64_t> seen;  // the last layout run without a graph
    ~SeedPhase() {
        if (exec) (void)hipGraphExecDestroy(exec);
    }
};

namespace {

// Lane and resolve launches of the captured levels stride over the task list
// with at most this many waves (the task count is not known when they are
// enqueued).
constexpr int32_t kGraphLaneWaves = 4096;
// References from which the rescans' level-0 tails are on by default.
constexpr int32_t kTailMinRef = 65536;

void run_levels(dcns_ctx *c, const Staged &S, const Params &P, Forest &F, int64_t &gpu_cells, int32_t &levels) {
    const int32_t NW = (int32_t)S.nwin;
    F = Forest{};
    F.NW = NW;
    levels = 0;
    gpu_cells = 0;
    for (int64_t v : S.level0_cells) gpu_cells += v;
    if (NW == 0 || (int64_t)P.window * P.match < P.min_score) return;  // no window can reach min_score
    const int wm = P.window <= 38 ? 38 : 64;
    const int32_t cap = P.window * P.match - P.min_score + 1;
    const ScoreTab tab = make_score_tab(P.match, P.mismatch);
    const SeedTuning T = seed_tuning(c);
    // level-0 tails of the rescans (level0_tail_rows): on by request, or by
    // default where a reference is long enough for rescans to dominate
    int32_t delta = level0_tail_rows(P, wm);
    {
        int32_t longest = 0;
        for (const PairDesc &d : S.pd) longest = std::max(longest, d.len1);
        const int32_t want = c->tune.level0_tail;
        if (want < 0 || (want == 0 && longest < kTailMinRef) || !scan_uses_f16(P)) delta = 0;
    }
    if (!c->seedphase) c->seedphase = std::make_shared<SeedPhase>();
    SeedPhase &SP = *c->seedphase;
    if (SP.nw != NW || SP.pool_min != T.pool_min) {
        if (T.pool_min > 0) {  // pools that must grow (tests of the growth path)
            SP.scan_cap = (int64_t)NW + T.pool_min;
            SP.node_cap = (int64_t)NW + T.pool_min;
            SP.next_cap = T.pool_min;
        } else {
            SP.scan_cap = (int64_t)NW + std::max<int64_t>(NW / 4, 4096);
            SP.node_cap = (int64_t)NW + std::max<int64_t>(NW / 2, 8192);
            SP.next_cap = std::max<int64_t>(NW / 4, 4096);
        }
        SP.nw = NW;
        SP.pool_min = T.pool_min;
    }
    auto ensure_pools = [&]() {
        c->traj.ensure((size_t)SP.scan_cap * cap * sizeof(TrajEntry));
        c->tcount.ensure((size_t)SP.scan_cap * sizeof(int32_t));
        c->nodes.ensure((size_t)SP.node_cap * sizeof(Node));
        c->tasks_a.ensure((size_t)SP.next_cap * sizeof(ScanTask));
        c->tasks_b.ensure((size_t)SP.next_cap * sizeof(ScanTask));
        c->task_order.ensure((size_t)SP.next_cap * sizeof(int32_t) + 64);
    };
    ensure_pools();
    c->hitlist.ensure((size_t)NW * sizeof(int32_t) + 64);
    c->lastrow.ensure((size_t)NW * sizeof(int32_t) + 64);
    c->ctr.ensure(sizeof(Counters) + 64);  // before the sink takes the address of its chunk counter
    StrongSink sink{nullptr, nullptr, nullptr, nullptr, 0};
    StrongIndex SI{nullptr, nullptr, nullptr, 0};
    if (delta > 0) {
        // chunk budget: a quarter of the free device memory, at most 4 GiB
        size_t free_b = 0, total_b = 0;
        hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
        const int64_t bytes = std::min<int64_t>((int64_t)(free_b / 4), (int64_t)4 << 30);
        const int64_t chunks = std::max<int64_t>(bytes / (kStrongChunk * 12 + 4), 1024);
        c->strong_win.ensure((size_t)chunks * kStrongChunk * sizeof(uint32_t));
        c->strong_rv.ensure((size_t)chunks * kStrongChunk * sizeof(uint64_t));
        c->strong_fill.ensure((size_t)chunks * sizeof(int32_t));
        sink = StrongSink{c->strong_win.as<uint32_t>(), c->strong_rv.as<uint64_t>(), c->strong_fill.as<int32_t>(),
                          &c->ctr.as<Counters>()->strong_chunks, (int32_t)std::min<int64_t>(chunks, INT32_MAX / 2)};
    }
    c->ctr.ensure(sizeof(Counters) + 64);
    c->order_hist.ensure(2 * kOrderBuckets * sizeof(int32_t));
    Counters *ctr = c->ctr.as<Counters>();
    const int32_t *hits_dev = &ctr->hits, *lt_dev = &ctr->level_tasks;
    // level L >= 1 scans the tasks in buffer (L - 1) & 1 and queues the next
    // level's into buffer L & 1; level 0's resolve queues into buffer 0
    auto buf = [&](int L) { return (L & 1) ? c->tasks_b.as<ScanTask>() : c->tasks_a.as<ScanTask>(); };
    auto resolve_launch = [&](int L, int32_t n_max, int32_t blocks_max) {
        return launch_resolve(c->stream, wm, c->seq.as<int8_t>(), c->pairs.as<PairDesc>(), c->pow.as<int32_t>(), P, tab,
                              L == 0, c->hitlist.as<int32_t>(), L == 0 ? nullptr : buf(L - 1),
                              L == 0 ? hits_dev : lt_dev, n_max, blocks_max, c->nodes.as<Node>(),
                              (int32_t)SP.node_cap, c->traj.as<TrajEntry>(), c->tcount.as<int32_t>(), cap,
                              (int32_t)SP.scan_cap, buf(L), (int32_t)SP.next_cap, ctr, c->lastrow.as<int32_t>(), delta);
    };
    auto level_launches = [&](int L, int32_t n_max, int32_t lane_blocks_max, bool timed) {
        auto run = [&](const char *name, auto f) {
            if (timed) c->launch(name, f);
            else hip_check(f(), name);
        };
        run("level_begin", [&] { return launch_level_begin(c->stream, ctr, c->order_hist.as<int32_t>()); });
        run("order_tasks", [&] {
            return launch_order_tasks(c->stream, buf(L - 1), lt_dev, n_max, c->order_hist.as<int32_t>(),
                                      c->task_order.as<int32_t>());
        });
        run("scan_general", [&] {
            return launch_scan_general(c->stream, wm, c->seq.as<int8_t>(), c->pairs.as<PairDesc>(), buf(L - 1),
                                       c->task_order.as<int32_t>(), lt_dev, n_max, T.wave_tasks, T.wave_rows,
                                       lane_blocks_max, P, tab, c->traj.as<TrajEntry>(), c->tcount.as<int32_t>(), cap,
                                       ctr);
        });
        if (delta > 0)
            run("rescan_tail", [&] {
                return launch_rescan_tail(c->stream, wm, c->seq.as<int8_t>(), c->pairs.as<PairDesc>(), P, buf(L - 1),
                                          lt_dev, n_max, lane_blocks_max, c->nodes.as<Node>(),
                                          c->lastrow.as<int32_t>(), SI, c->traj.as<TrajEntry>(),
                                          c->tcount.as<int32_t>(), cap, ctr);
            });
    };

    c->launch("seed_init", [&] { return launch_seed_init(c->stream, ctr, NW); });
    // Level 0 in row segments when the batch's longest references would hold
    // the launch (its lane rows are serial): each 128-window group's rows in
    // segments of about seg rows, each started warm rows early from zeros,
    // as many waves as segments; the segments' trajectories then merge.
    SegL0 sg;
    std::vector<int4> seg_tasks, seg_groups;
    {
        const int32_t warm = level0_tail_rows(P, wm);
        const int32_t want = c->tune.level0_split;
        if (warm > 0 && scan_uses_f16(P) && want >= 0) {
            double total = 0;
            int32_t longest = 0;
            for (const PairDesc &d : S.pd) {
                total += (double)((d.nwin + 127) / 128) * d.len1;
                longest = std::max(longest, d.len1);
            }
            const double thr_time = total * kLaneRowIssue / (4.0 * c->cu_count());
            int32_t seg = 0;
            if (want == 1)
                seg = 4 * warm;
            else if (longest * kLaneRowLatency > 1.5 * thr_time)
                seg = std::max({8 * warm, 1024, (int32_t)(thr_time / (2.0 * kLaneRowLatency))});
            if (seg > 0) {
                std::vector<int32_t> byl((size_t)S.pd.size());
                for (size_t p = 0; p < byl.size(); ++p) byl[p] = (int32_t)p;
                std::stable_sort(byl.begin(), byl.end(),
                                 [&](int32_t a, int32_t b) { return S.pd[(size_t)a].len1 > S.pd[(size_t)b].len1; });
                for (int32_t p : byl) {
                    const PairDesc &d = S.pd[(size_t)p];
                    if (d.len1 <= 0) continue;
                    const int32_t nseg = d.len1 > seg + seg / 2 ? (d.len1 + seg - 1) / seg : 1;
                    for (int32_t g = 0; g < d.nwin; g += 128) {
                        seg_groups.push_back(make_int4(p, g, (int32_t)seg_tasks.size(), nseg));
                        for (int32_t k = 0; k < nseg; ++k) {
                            const int32_t r0 = (int32_t)((int64_t)d.len1 * k / nseg);
                            const int32_t r1 = (int32_t)((int64_t)d.len1 * (k + 1) / nseg);
                            seg_tasks.push_back(make_int4(p, g, r0, r1));
                        }
                    }
                }
                upload(c, c->seg_tasks, seg_tasks);
                upload(c, c->seg_groups, seg_groups);
                c->seg_traj.ensure(std::max<size_t>(seg_tasks.size(), 1) * 128 * (size_t)cap * sizeof(TrajEntry));
                c->seg_tcount.ensure(std::max<size_t>(seg_tasks.size(), 1) * 128 * sizeof(int32_t));
                sg.tasks = c->seg_tasks.as<int4>();
                sg.groups = c->seg_groups.as<int4>();
                sg.traj = c->seg_traj.as<TrajEntry>();
                sg.tcount = c->seg_tcount.as<int32_t>();
                sg.warm = warm;
            }
        }
    }
    if (!seg_tasks.empty()) {
        hip_check(launch_fill_i32(c->stream, c->lastrow.as<int32_t>(), NW, -1), "lastrow");
        c->launch("scan_uniform", [&] {
            return launch_scan_uniform_seg(c->stream, wm, c->seq.as<int8_t>(), c->pairs.as<PairDesc>(), sg,
                                           (int32_t)seg_tasks.size(), P, cap, ctr, c->lastrow.as<int32_t>(), sink);
        });
        c->launch("l0_merge", [&] {
            return launch_l0_merge(c->stream, c->pairs.as<PairDesc>(), sg, (int32_t)seg_groups.size(),
                                   c->traj.as<TrajEntry>(), c->tcount.as<int32_t>(), cap, P.min_score,
                                   c->hitlist.as<int32_t>(), ctr);
        });
    } else {
        c->launch("scan_uniform", [&] {
            return launch_scan_uniform(c->stream, wm, c->seq.as<int8_t>(), c->pairs.as<PairDesc>(),
                                       c->wtasks.as<int2>(), (int32_t)S.wt.size(), T.level0_waves, P, tab,
                                       c->traj.as<TrajEntry>(), c->tcount.as<int32_t>(), cap,
                                       c->hitlist.as<int32_t>(), ctr, c->l0wins.as<int2>(), (int32_t)S.ww.size(),
                                       c->lastrow.as<int32_t>(), sink);
        });
    }
    // the strong rows' index: chunk fills -> offsets and total, compaction,
    // a stable sort by window, per-window ranges and next-larger links
    if (delta > 0) {
        const int32_t nchunks = read_ctr(c).strong_chunks;
        if (nchunks > sink.cap_chunks) {
            delta = 0;  // entries were dropped: full rescans for this batch
        } else {
            c->strong_off.ensure((size_t)std::max(nchunks, 1) * sizeof(int64_t));
            c->strong_fill64.ensure((size_t)std::max(nchunks, 1) * sizeof(int64_t));
            c->strong_total.ensure(64);
            const size_t tmp0 = strong_index_scratch(nchunks, 1);
            c->strong_tmp.ensure(tmp0);
            hip_check(launch_strong_count(c->stream, sink, nchunks, c->strong_tmp.p, c->strong_tmp.cap,
                                          c->strong_fill64.as<int64_t>(), c->strong_off.as<int64_t>(),
                                          c->strong_total.as<int64_t>()),
                      "strong_count");
            int64_t n = 0;
            hip_check(hipMemcpyAsync(&n, c->strong_total.p, sizeof(n), hipMemcpyDeviceToHost, c->stream), "strong total");
            c->sync();
            c->strong_key2.ensure((size_t)std::max<int64_t>(n, 1) * sizeof(uint64_t));
            c->strong_key.ensure((size_t)std::max<int64_t>(n, 1) * sizeof(uint64_t));
            c->strong_rv2.ensure((size_t)std::max<int64_t>(n, 1) * sizeof(uint64_t));
            c->strong_seg.ensure((size_t)NW * sizeof(int2) + 64);
            c->strong_ng.ensure((size_t)std::max<int64_t>(n, 1) * sizeof(int32_t));
            c->strong_tmp.ensure(strong_index_scratch(nchunks, n));
            int32_t win_bits = 1;
            while (win_bits < 32 && ((int64_t)1 << win_bits) < (int64_t)NW) ++win_bits;
            c->launch("strong_index", [&] {
                return launch_strong_index(c->stream, sink, nchunks, n, NW, win_bits, c->strong_tmp.p, c->strong_tmp.cap,
                                           c->strong_key2.as<uint64_t>(), c->strong_rv2.as<uint64_t>(),
                                           c->strong_key.as<uint64_t>(), c->strong_off.as<int64_t>(),
                                           c->strong_seg.as<int2>(),
                                           c->strong_ng.as<int32_t>());
            });
            SI = StrongIndex{c->strong_seg.as<int2>(), c->strong_rv.as<uint64_t>(), c->strong_ng.as<int32_t>(), delta};
        }
    }

    Counters h{};
    int L = 1;  // the next level to run
    bool resolved0 = false;
    bool use_graph = T.graph_levels > 0;
    if (use_graph) {
        // level 0's resolve and graph_levels levels, enqueued with no host
        // read between them: every launch reads its task count from the
        // counter block; captured once per batch layout, pools and settings
        const PairDesc *pd = c->pairs.as<PairDesc>();
        std::vector<int64_t> key = {(int64_t)c->seq.p, (int64_t)pd, (int64_t)c->pow.p, (int64_t)c->hitlist.p,
                                    (int64_t)c->nodes.p, (int64_t)c->traj.p, (int64_t)c->tcount.p,
                                    (int64_t)c->tasks_a.p, (int64_t)c->tasks_b.p, (int64_t)c->task_order.p,
                                    (int64_t)c->order_hist.p, (int64_t)c->lastrow.p, (int64_t)ctr, NW,
                                    SP.scan_cap, SP.node_cap,
                                    SP.next_cap, T.wave_tasks, T.wave_rows, T.graph_levels, wm, cap,
                                    P.match, P.mismatch, P.open, P.ext, P.window, P.step, P.min_score, P.min_ref,
                                    delta, (int64_t)SI.seg, (int64_t)SI.rv, (int64_t)SI.ng};
        // a layout seen for the first time runs host-driven; one that recurs
        // (a staged batch stepped again) is captured and replayed from then on
        if ((!SP.exec || SP.key != key) && SP.seen != key) {
            SP.seen = key;
            use_graph = false;
        } else if (!SP.exec || SP.key != key) {
            if (SP.exec) {
                hip_check(hipGraphExecDestroy(SP.exec), "hipGraphExecDestroy");
                SP.exec = nullptr;
            }
            hip_check(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
            hipGraph_t g = nullptr;
            try {
                hip_check(resolve_launch(0, NW, kGraphLaneWaves), "resolve");
                for (int l = 1; l <= T.graph_levels; ++l) {
                    level_launches(l, (int32_t)SP.next_cap, kGraphLaneWaves, false);
                    hip_check(resolve_launch(l, (int32_t)SP.next_cap, kGraphLaneWaves), "resolve");
                }
            } catch (...) {
                (void)hipStreamEndCapture(c->stream, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            hip_check(hipStreamEndCapture(c->stream, &g), "hipStreamEndCapture");
            const hipError_t e = hipGraphInstantiate(&SP.exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            hip_check(e, "hipGraphInstantiate");
            SP.key = key;
        }
    }
    if (use_graph) {
        c->launch("seed_levels", [&] { return hipGraphLaunch(SP.exec, c->stream); });
        h = read_ctr(c);
        if (!h.overflow) {
            resolved0 = true;
            levels = 1 + h.levels;
            L = T.graph_levels + 1;
        } else {
            // a pool overflowed: the recursion runs again from level 0, host-driven,
            // growing the pools (the level-0 scan's trajectories are intact)
            Counters r{};
            r.nodes = NW;
            r.scans = NW;
            r.hits = h.hits;
            r.entries = h.entries;
            h = r;
            write_ctr(c, h);
        }
    } else {
        h = read_ctr(c);
    }

    // resolve with capacity retry: a walk stops at its first allocation that
    // does not fit, so a retry can need more than the last one requested;
    // pools at least double per attempt
    auto resolve = [&](int level) {
        Counters before = h;
        before.next_tasks = 0;
        before.overflow = 0;
        const int32_t n = level == 0 ? h.hits : h.level_tasks;
        for (int attempt = 0; attempt < 32; ++attempt) {
            c->nodes.ensure((size_t)SP.node_cap * sizeof(Node), c->stream, true, (size_t)before.nodes * sizeof(Node));
            c->traj.ensure((size_t)SP.scan_cap * cap * sizeof(TrajEntry), c->stream, true,
                           (size_t)before.scans * cap * sizeof(TrajEntry));
            c->tcount.ensure((size_t)SP.scan_cap * sizeof(int32_t), c->stream, true,
                             (size_t)before.scans * sizeof(int32_t));
            // the output task buffer (its contents are rewritten); the input one is untouched
            ((level & 1) ? c->tasks_b : c->tasks_a).ensure((size_t)SP.next_cap * sizeof(ScanTask));
            write_ctr(c, before);
            c->launch("resolve", [&] { return resolve_launch(level, n, 0); });
            h = read_ctr(c);
            if (!h.overflow) return;
            SP.node_cap = std::max<int64_t>(2 * SP.node_cap, (int64_t)h.nodes + h.nodes / 4 + 1024);
            SP.scan_cap = std::max<int64_t>(2 * SP.scan_cap, (int64_t)h.scans + h.scans / 4 + 1024);
            SP.next_cap = std::max<int64_t>(2 * SP.next_cap, (int64_t)h.next_tasks + h.next_tasks / 4 + 1024);
        }
        throw Error(DCNS_ERR_INTERNAL, "seed recursion pools could not be sized");
    };
    if (!resolved0) {
        resolve(0);
        levels = 1;
    }
    // levels past the captured ones (or all of them): host-driven, exact grids
    // every right child starts below its parent's first row, so a window's
    // recursion is at most its reference's length deep
    int32_t max_len1 = 0;
    for (const PairDesc &d : S.pd) max_len1 = std::max(max_len1, d.len1);
    while (h.next_tasks > 0) {
        if (levels > max_len1 + 2) throw Error(DCNS_ERR_INTERNAL, "seed recursion deeper than its reference");
        const int32_t n = h.next_tasks;
        c->task_order.ensure((size_t)n * sizeof(int32_t) + 64);
        level_launches(L, n, 0, true);
        h = read_ctr(c);
        resolve(L);
        ++levels;
        ++L;
    }
    gpu_cells += (int64_t)h.gen_cells;

    // bring the forest home (pinned buffers): hit windows, their roots, the dynamic nodes
    const int32_t hits = h.hits;
    const size_t ndyn = (size_t)(h.nodes - NW);
    c->roots.ensure((size_t)std::max(hits, 1) * sizeof(Node));
    c->launch("gather", [&] {
        return launch_gather_roots(c->stream, c->nodes.as<Node>(), c->hitlist.as<int32_t>(), hits_dev, hits,
                                   c->roots.as<Node>());
    });
    c->h_hits.ensure((size_t)std::max(hits, 1) * sizeof(int32_t));
    c->h_roots.ensure((size_t)std::max(hits, 1) * sizeof(Node));
    c->h_nodes.ensure(std::max<size_t>(ndyn, 1) * sizeof(Node));
    if (hits) {
        hip_check(hipMemcpyAsync(c->h_hits.p, c->hitlist.p, (size_t)hits * sizeof(int32_t), hipMemcpyDeviceToHost,
                                 c->stream), "download hits");
        hip_check(hipMemcpyAsync(c->h_roots.p, c->roots.p, (size_t)hits * sizeof(Node), hipMemcpyDeviceToHost,
                                 c->stream), "download roots");
    }
    if (ndyn)
        hip_check(hipMemcpyAsync(c->h_nodes.p, c->nodes.as<Node>() + NW, ndyn * sizeof(Node), hipMemcpyDeviceToHost,
                                 c->stream), "download nodes");
    c->sync();
    F.hits = hits;
    F.hit_window = c->h_hits.as<int32_t>();
    F.roots = c->h_roots.as<Node>();
    F.dyn = c->h_nodes.as<Node>();
}

Seed seed_of(const Node &n) { return Seed(n.s1, n.e1, n.s2, n.e2); }

// Reverse-SW lengths of getSeeds (:152-153) for a node, for cell accounting.
int64_t reverse_cells(const Node &n, int32_t sp, int32_t tw, int32_t W) {
    const int64_t r1 = std::min(1 + (n.e1 - n.a), tw + W);
    const int64_t r2 = std::min(1 + (n.e2 - sp), tw + W);
    return r1 * r2;
}

// Inner getSeeds (:61-128) replayed from the device's recursion tree: the
// set of the node's range is {seed}, then the left child's set, then the
// right child's set, each inserted in its own set order.
std::set<Seed> inner_set(const Forest &F, int32_t id, int32_t sp, int32_t tw, int32_t W, int64_t &cells) {
    std::set<Seed> out;
    const Node &n = F.child(id);
    cells += (int64_t)(n.b - n.a) * tw;
    if (n.e1 < 0) return out;
    cells += reverse_cells(n, sp, tw, W);
    out.insert(seed_of(n));
    if (n.left >= 0)
        for (const Seed &s : inner_set(F, n.left, sp, tw, W, cells)) out.insert(s);
    if (n.right >= 0)
        for (const Seed &s : inner_set(F, n.right, sp, tw, W, cells)) out.insert(s);
    return out;
}

}  // namespace

namespace {

// One pass of link_seeds (findSimilarFragmentsForPairedSequence4.cpp:298-334)
// as written: for each seed s1 still unused, in start1 order, the seeds
// after it are taken in order until s1 merges or the scan stops; a seed
// inside s1 is marked used on the way.  Returns whether any seed merged.
bool link_pass_scalar(const std::vector<Seed> &cur, std::vector<Seed> &next) {
    const size_t n = cur.size();
    std::vector<char> used(n, 0);
    bool changed = false;
    for (size_t i = 0; i < n; ++i) {
        if (used[i]) continue;
        const Seed s1 = cur[i];
        bool merged = false;
        for (size_t j = i + 1; j < n; ++j) {
            if (used[j]) continue;
            const Seed s2 = cur[j];
            if (s1.getStart2() <= s2.getStart2() && s1.getEnd1() >= s2.getEnd1() && s1.getEnd2() >= s2.getEnd2()) {
                used[j] = 1;  // s2 inside s1
            } else if (s1.getStart2() <= s2.getStart2() && s1.getEnd1() <= s2.getEnd1() &&
                       s1.getEnd2() <= s2.getEnd2()) {
                const int32_t d = s2.getStart1() - s1.getStart1();
                if (d == s2.getStart2() - s1.getStart2() && d == s2.getEnd1() - s1.getEnd1() &&
                    d == s2.getEnd2() - s1.getEnd2()) {  // same diagonal: merge
                    next.push_back(Seed(s1.getStart1(), s2.getEnd1(), s1.getStart2(), s2.getEnd2()));
                    used[i] = 1;
                    used[j] = 1;
                    merged = true;
                    break;
                }
            } else if (s2.getStart1() > s1.getEnd1() || s2.getStart2() > s1.getEnd2()) {
                break;
            }
        }
        if (merged) changed = true;
        else next.push_back(s1);
    }
    return changed;
}

// The same pass, 8 candidates at a time.  A candidate's fate depends only on
// s1, its own coordinates and whether it is still unused (marking one seed
// used changes no other's), so a block yields, per candidate, "inside s1",
// "merges" and "stops the scan" masks; the first unused merge-or-stop ends
// the scan, and the unused seeds inside s1 before it are marked used, which
// is what the one-at-a-time loop does.  Coordinates are copied to separate
// arrays, padded by 8 seeds that are marked used.
__attribute__((target("avx2"))) inline uint32_t mask(__m256i v) {
    return (uint32_t)_mm256_movemask_ps(_mm256_castsi256_ps(v));
}

__attribute__((target("avx2"))) bool link_pass_avx2(const std::vector<Seed> &cur, std::vector<Seed> &next) {
    const int64_t n = (int64_t)cur.size();
    std::vector<int32_t> S1((size_t)n + 8, 0), E1((size_t)n + 8, 0), S2((size_t)n + 8, 0), E2((size_t)n + 8, 0);
    for (int64_t i = 0; i < n; ++i) {
        S1[(size_t)i] = cur[(size_t)i].getStart1();
        E1[(size_t)i] = cur[(size_t)i].getEnd1();
        S2[(size_t)i] = cur[(size_t)i].getStart2();
        E2[(size_t)i] = cur[(size_t)i].getEnd2();
    }
    std::vector<uint8_t> used((size_t)n + 8, 1);
    std::fill(used.begin(), used.begin() + n, 0);
    const __m256i ones = _mm256_set1_epi32(-1);
    bool changed = false;
    for (int64_t i = 0; i < n; ++i) {
        if (used[(size_t)i]) continue;
        const int32_t a1 = S1[(size_t)i], ae1 = E1[(size_t)i], a2 = S2[(size_t)i], ae2 = E2[(size_t)i];
        const __m256i vs1 = _mm256_set1_epi32(a1), ve1 = _mm256_set1_epi32(ae1), vs2 = _mm256_set1_epi32(a2),
                      ve2 = _mm256_set1_epi32(ae2);
        bool merged = false;
        for (int64_t j = i + 1; j < n; j += 8) {
            const __m256i b1 = _mm256_loadu_si256((const __m256i *)(S1.data() + j));
            const __m256i be1 = _mm256_loadu_si256((const __m256i *)(E1.data() + j));
            const __m256i b2 = _mm256_loadu_si256((const __m256i *)(S2.data() + j));
            const __m256i be2 = _mm256_loadu_si256((const __m256i *)(E2.data() + j));
            const __m256i lo2 = _mm256_cmpgt_epi32(vs2, b2);  // s2.start2 < s1.start2
            const uint32_t inside = mask(_mm256_andnot_si256(
                _mm256_or_si256(lo2, _mm256_or_si256(_mm256_cmpgt_epi32(be1, ve1), _mm256_cmpgt_epi32(be2, ve2))), ones));
            const uint32_t after = mask(_mm256_andnot_si256(
                _mm256_or_si256(lo2, _mm256_or_si256(_mm256_cmpgt_epi32(ve1, be1), _mm256_cmpgt_epi32(ve2, be2))), ones));
            const __m256i d = _mm256_sub_epi32(b1, vs1);
            const uint32_t diag = mask(_mm256_and_si256(
                _mm256_cmpeq_epi32(d, _mm256_sub_epi32(b2, vs2)),
                _mm256_and_si256(_mm256_cmpeq_epi32(d, _mm256_sub_epi32(be1, ve1)),
                                 _mm256_cmpeq_epi32(d, _mm256_sub_epi32(be2, ve2)))));
            const uint32_t beyond = mask(_mm256_or_si256(_mm256_cmpgt_epi32(b1, ve1), _mm256_cmpgt_epi32(b2, ve2)));
            const uint32_t live = (uint32_t)_mm_movemask_epi8(
                                      _mm_cmpeq_epi8(_mm_loadl_epi64((const __m128i *)(used.data() + j)), _mm_setzero_si128())) &
                                  0xffu;
            const uint32_t merge = ~inside & after & diag;
            const uint32_t stop = live & (merge | (~inside & ~after & beyond)) & 0xffu;
            uint32_t absorb = live & inside & (stop ? (stop & (0u - stop)) - 1u : 0xffu);
            while (absorb) {
                used[(size_t)(j + __builtin_ctz(absorb))] = 1;
                absorb &= absorb - 1;
            }
            if (stop) {
                const int k = __builtin_ctz(stop);
                if ((merge >> k) & 1u) {
                    const int64_t q = j + k;
                    next.push_back(Seed(a1, E1[(size_t)q], a2, E2[(size_t)q]));
                    used[(size_t)i] = 1;
                    used[(size_t)q] = 1;
                    merged = true;
                }
                break;
            }
        }
        if (merged) changed = true;
        else next.push_back(Seed(a1, ae1, a2, ae2));
    }
    return changed;
}

// One seed against one candidate, in the order of the reference's tests:
// inside (absorbed), after on the same diagonal (merge, stop), beyond (stop).
enum { kLinkNone = 0, kLinkInside = 1, kLinkMerge = 2, kLinkStop = 3 };
inline int link_decide(int32_t a1, int32_t ae1, int32_t a2, int32_t ae2, int32_t b1, int32_t be1, int32_t b2,
                      int32_t be2) {
    if (a2 <= b2 && ae1 >= be1 && ae2 >= be2) return kLinkInside;
    if (a2 <= b2 && ae1 <= be1 && ae2 <= be2) {
        const int32_t d = b1 - a1;
        return (d == b2 - a2 && d == be1 - ae1 && d == be2 - ae2) ? kLinkMerge : kLinkNone;
    }
    if (b1 > ae1 || b2 > ae2) return kLinkStop;
    return kLinkNone;
}

// Link passes over fixed positions (a consumed seed dies in place, a merged
// seed keeps its first seed's position), each pass cut into chunks that run
// on all host threads.  A chunk starts where no earlier live seed's end1
// reaches its first seed's start1, so no earlier seed can absorb a seed of it
// (only a merge can cross).  A chunk's seeds read positions past the chunk
// from the pass's starting state, and the seeds they consume there are
// listed instead of written.  Then, in order, a chunk into which an earlier
// chunk consumed runs again from the pass's starting state with those seeds
// dead: every seed sees what the one-at-a-time pass shows it.
struct LinkArrays {
    std::vector<int32_t> S1, E1, S2, E2;
    std::vector<uint8_t> dead;
};

__attribute__((target("avx2"))) bool link_chunk(LinkArrays &L, const LinkArrays &P, int64_t b, int64_t e,
                                                int64_t n, std::vector<int64_t> &kills) {
    kills.clear();
    bool merged = false;
    const __m256i ones = _mm256_set1_epi32(-1);
    auto killed_past = [&](int64_t q) { return std::find(kills.begin(), kills.end(), q) != kills.end(); };
    for (int64_t i = b; i < e; ++i) {
        if (L.dead[(size_t)i]) continue;
        const int32_t a1 = L.S1[(size_t)i], ae1 = L.E1[(size_t)i], a2 = L.S2[(size_t)i], ae2 = L.E2[(size_t)i];
        const __m256i vs1 = _mm256_set1_epi32(a1), ve1 = _mm256_set1_epi32(ae1), vs2 = _mm256_set1_epi32(a2),
                      ve2 = _mm256_set1_epi32(ae2);
        int64_t mj = -1;
        bool done = false;
        int64_t j = i + 1;
        // the chunk's own positions 8 at a time (as link_pass_avx2)
        for (; j + 8 <= e; j += 8) {
            const __m256i b1 = _mm256_loadu_si256((const __m256i *)(L.S1.data() + j));
            const __m256i be1 = _mm256_loadu_si256((const __m256i *)(L.E1.data() + j));
            const __m256i b2 = _mm256_loadu_si256((const __m256i *)(L.S2.data() + j));
            const __m256i be2 = _mm256_loadu_si256((const __m256i *)(L.E2.data() + j));
            const __m256i lo2 = _mm256_cmpgt_epi32(vs2, b2);
            const uint32_t inside = mask(_mm256_andnot_si256(
                _mm256_or_si256(lo2, _mm256_or_si256(_mm256_cmpgt_epi32(be1, ve1), _mm256_cmpgt_epi32(be2, ve2))), ones));
            const uint32_t after = mask(_mm256_andnot_si256(
                _mm256_or_si256(lo2, _mm256_or_si256(_mm256_cmpgt_epi32(ve1, be1), _mm256_cmpgt_epi32(ve2, be2))), ones));
            const __m256i d = _mm256_sub_epi32(b1, vs1);
            const uint32_t diag = mask(_mm256_and_si256(
                _mm256_cmpeq_epi32(d, _mm256_sub_epi32(b2, vs2)),
                _mm256_and_si256(_mm256_cmpeq_epi32(d, _mm256_sub_epi32(be1, ve1)),
                                 _mm256_cmpeq_epi32(d, _mm256_sub_epi32(be2, ve2)))));
            const uint32_t beyond = mask(_mm256_or_si256(_mm256_cmpgt_epi32(b1, ve1), _mm256_cmpgt_epi32(b2, ve2)));
            const uint32_t live = (uint32_t)_mm_movemask_epi8(
                                      _mm_cmpeq_epi8(_mm_loadl_epi64((const __m128i *)(L.dead.data() + j)), _mm_setzero_si128())) &
                                  0xffu;
            const uint32_t merge = ~inside & after & diag;
            const uint32_t stop = live & (merge | (~inside & ~after & beyond)) & 0xffu;
            uint32_t absorb = live & inside & (stop ? (stop & (0u - stop)) - 1u : 0xffu);
            while (absorb) {
                L.dead[(size_t)(j + __builtin_ctz(absorb))] = 1;
                absorb &= absorb - 1;
            }
            if (stop) {
                const int k = __builtin_ctz(stop);
                if ((merge >> k) & 1u) mj = j + k;
                done = true;
                break;
            }
        }
        // the rest of the chunk, then past it from the starting state
        for (; !done && j < n; ++j) {
            const bool own = j < e;
            if (own ? L.dead[(size_t)j] != 0 : (P.dead[(size_t)j] != 0 || killed_past(j))) continue;
            const LinkArrays &A = own ? L : P;
            const int r = link_decide(a1, ae1, a2, ae2, A.S1[(size_t)j], A.E1[(size_t)j], A.S2[(size_t)j], A.E2[(size_t)j]);
            if (r == kLinkInside) {
                if (own) L.dead[(size_t)j] = 1;
                else kills.push_back(j);
            } else if (r == kLinkMerge) {
                mj = j;
                break;
            } else if (r == kLinkStop) {
                break;
            }
        }
        if (mj >= 0) {
            const bool own = mj < e;
            L.E1[(size_t)i] = own ? L.E1[(size_t)mj] : P.E1[(size_t)mj];
            L.E2[(size_t)i] = own ? L.E2[(size_t)mj] : P.E2[(size_t)mj];
            if (own) L.dead[(size_t)mj] = 1;
            else kills.push_back(mj);
            merged = true;
        }
    }
    return merged;
}

std::vector<Seed> link_parallel(const std::vector<Seed> &cur) {
    const int64_t n = (int64_t)cur.size();
    LinkArrays L, P;
    for (std::vector<int32_t> *v : {&L.S1, &L.E1, &L.S2, &L.E2}) v->assign((size_t)n + 8, 0);
    L.dead.assign((size_t)n + 8, 1);
    for (int64_t i = 0; i < n; ++i) {
        L.S1[(size_t)i] = cur[(size_t)i].getStart1();
        L.E1[(size_t)i] = cur[(size_t)i].getEnd1();
        L.S2[(size_t)i] = cur[(size_t)i].getStart2();
        L.E2[(size_t)i] = cur[(size_t)i].getEnd2();
        L.dead[(size_t)i] = 0;
    }
    P.S1 = L.S1;  // start coordinates never change
    P.S2 = L.S2;
    const int threads = (int)std::min<unsigned>(std::max(1u, std::thread::hardware_concurrency()), 16u);
    bool changed = true;
    while (changed) {
        P.E1 = L.E1;
        P.E2 = L.E2;
        P.dead = L.dead;
        const int64_t want = std::max<int64_t>(n / ((int64_t)threads * 8), 4096);
        std::vector<int64_t> starts = {0};
        int32_t reach = INT32_MIN;
        for (int64_t p = 0; p < n; ++p) {
            if (L.dead[(size_t)p]) continue;
            if (p - starts.back() >= want && reach < L.S1[(size_t)p]) starts.push_back(p);
            reach = std::max(reach, L.E1[(size_t)p]);
        }
        starts.push_back(n);
        const int64_t nc = (int64_t)starts.size() - 1;
        std::vector<std::vector<int64_t>> kills((size_t)nc);
        std::vector<uint8_t> merged((size_t)nc, 0);
        parallel_for(nc, [&](int64_t c) {
            merged[(size_t)c] = link_chunk(L, P, starts[(size_t)c], starts[(size_t)c + 1], n, kills[(size_t)c]);
        }, threads);
        std::vector<int64_t> incoming, keep;
        for (int64_t c = 0; c < nc; ++c) {
            const int64_t b = starts[(size_t)c], e = starts[(size_t)c + 1];
            bool hit = false;
            for (int64_t q : incoming) hit |= q >= b && q < e;
            if (hit) {
                for (int64_t p = b; p < e; ++p) {
                    L.E1[(size_t)p] = P.E1[(size_t)p];
                    L.E2[(size_t)p] = P.E2[(size_t)p];
                    L.dead[(size_t)p] = P.dead[(size_t)p];
                }
                for (int64_t q : incoming)
                    if (q >= b && q < e) L.dead[(size_t)q] = 1;
                merged[(size_t)c] = link_chunk(L, P, b, e, n, kills[(size_t)c]);
            }
            keep.clear();
            for (int64_t q : incoming)
                if (q >= e) keep.push_back(q);
            for (int64_t q : kills[(size_t)c]) keep.push_back(q);
            incoming.swap(keep);
        }
        changed = false;
        for (uint8_t m : merged) changed |= m != 0;
    }
    std::vector<Seed> out;
    for (int64_t i = 0; i < n; ++i)
        if (!L.dead[(size_t)i]) out.push_back(Seed(L.S1[(size_t)i], L.E1[(size_t)i], L.S2[(size_t)i], L.E2[(size_t)i]));
    return out;
}

// Lists from this size on link on all host threads.
constexpr size_t kLinkParallelMin = 65536;

}  // namespace

// link_seeds (findSimilarFragmentsForPairedSequence4.cpp:287-344) up to its
// last loop: seeds in start1 order (std::sort, as the reference), merge
// passes until none merges; the list in the order the reference inserts it
// back into the set.
std::vector<Seed> link_seed_list(std::vector<Seed> cur) {
    std::sort(cur.begin(), cur.end(), [](Seed a, Seed b) { return a.getStart1() < b.getStart1(); });
    static const bool avx2 = __builtin_cpu_supports("avx2");
    if (avx2 && cur.size() >= kLinkParallelMin) return link_parallel(cur);
    bool changed = true;
    while (changed) {
        std::vector<Seed> next;
        next.reserve(cur.size());
        changed = avx2 ? link_pass_avx2(cur, next) : link_pass_scalar(cur, next);
        cur.swap(next);
    }
    return cur;
}

// Inserting k into a std::set descends from the root, going left at every
// node x with k < x.  Under Seed's comparator, k < x needs x.start1 >
// k.start1, or equal start1 and x.start2 > k.start2.  When k's start1
// exceeds every earlier seed's, or equals the largest and k's start2 exceeds
// every earlier start2 of that start1, no node passes that test: the descent
// ends at the rightmost node, which is exactly where the end() hint puts k
// (and where both reject it when the rightmost node is equivalent to k).  So
// the hint changes no result, only the cost, for lists in start1 order.
template <class Set>
void insert_in_start1_order(Set &seeds, const std::vector<Seed> &list) {
    int32_t max1 = INT32_MIN, max2 = INT32_MIN;
    for (const Seed &s : list) {
        const int32_t a1 = s.getStart1(), a2 = s.getStart2();
        if (a1 > max1 || (a1 == max1 && a2 > max2)) seeds.insert(seeds.end(), s);
        else seeds.insert(s);
        if (a1 > max1) {
            max1 = a1;
            max2 = a2;
        } else if (a1 == max1) {
            max2 = std::max(max2, a2);
        }
    }
}

template void insert_in_start1_order(std::set<Seed> &, const std::vector<Seed> &);
template void insert_in_start1_order(SeedSet &, const std::vector<Seed> &);

void link_seeds(SeedSet &seeds) {
    const std::vector<Seed> cur = link_seed_list(std::vector<Seed>(seeds.begin(), seeds.end()));
    SeedSet fresh;  // the old nodes go with the old set's blocks
    insert_in_start1_order(fresh, cur);
    seeds.swap(fresh);
}

void seed_sets(dcns_ctx *c, const Staged &S, const Params &P, bool keep_unlinked, SeedResult &out) {
    Forest F;
    run_levels(c, S, P, F, out.gpu_cells, out.levels);
    const size_t np = S.pd.size();
    out.linked.assign(np, {});
    out.before_link.assign(np, 0);
    if (keep_unlinked) out.unlinked.assign(np, {});
    // hits bucketed by pair, then each pair's hits in window order
    std::vector<int64_t> first(np + 1, 0);
    for (int32_t k = 0; k < F.hits; ++k) ++first[(size_t)S.pair_of_window[(size_t)F.hit_window[k]] + 1];
    for (size_t p = 0; p < np; ++p) first[p + 1] += first[p];
    std::vector<int32_t> by_pair((size_t)F.hits);
    {
        std::vector<int64_t> fill(first.begin(), first.end() - 1);
        for (int32_t k = 0; k < F.hits; ++k)
            by_pair[(size_t)fill[(size_t)S.pair_of_window[(size_t)F.hit_window[k]]]++] = k;
    }
    std::vector<int64_t> cells(np, 0);
    parallel_for((int64_t)np, [&](int64_t p) {
        const PairDesc &d = S.pd[(size_t)p];
        int32_t *hk = by_pair.data() + first[(size_t)p];
        const int64_t nh = first[(size_t)p + 1] - first[(size_t)p];
        std::sort(hk, hk + nh, [&](int32_t a, int32_t b) { return F.hit_window[a] < F.hit_window[b]; });
        // The children's sets depend only on the recursion tree, not on the
        // pair's set: a pair with few windows in flight (one long pair) builds
        // them on all host threads first, in their own set order; the
        // window loop below then replays the gate and the insertions in the
        // reference's order.
        struct Kids {
            std::vector<Seed> left, right;
            int64_t cells = 0;
        };
        std::vector<Kids> kids;
        const bool prebuild = np < 16 && nh > 64;
        if (prebuild) {
            kids.resize((size_t)nh);
            parallel_for(nh, [&](int64_t q) {
                const Node &r = F.roots[hk[q]];
                if (r.e1 < 0) return;
                const int32_t sp = (F.hit_window[hk[q]] - d.win_base) * P.step;
                const int32_t tw = std::min(P.window, d.len2 - sp);
                Kids &k = kids[(size_t)q];
                if (r.left >= 0) {
                    std::set<Seed> x = inner_set(F, r.left, sp, tw, P.window, k.cells);
                    k.left.assign(x.begin(), x.end());
                }
                if (r.right >= 0) {
                    std::set<Seed> x = inner_set(F, r.right, sp, tw, P.window, k.cells);
                    k.right.assign(x.begin(), x.end());
                }
            });
        }
        SeedSet seeds;
        int64_t cl = S.level0_cells[(size_t)p];
        for (int64_t q = 0; q < nh; ++q) {
            const Node &r = F.roots[hk[q]];
            const int32_t w = F.hit_window[hk[q]] - d.win_base;
            const int32_t sp = w * P.step;
            const int32_t tw = std::min(P.window, d.len2 - sp);
            if (r.e1 < 0) continue;
            cl += reverse_cells(r, sp, tw, P.window);
            const Seed s0 = seed_of(r);
            if (seeds.find(s0) != seeds.end()) continue;  // the set gate of :170
            seeds.insert(s0);
            if (prebuild) {
                const Kids &k = kids[(size_t)q];
                for (const Seed &x : k.left) seeds.insert(x);
                for (const Seed &x : k.right) seeds.insert(x);
                cl += k.cells;
                continue;
            }
            if (r.left >= 0)
                for (const Seed &x : inner_set(F, r.left, sp, tw, P.window, cl)) seeds.insert(x);
            if (r.right >= 0)
                for (const Seed &x : inner_set(F, r.right, sp, tw, P.window, cl)) seeds.insert(x);
        }
        cells[(size_t)p] = cl;
        out.before_link[(size_t)p] = (int64_t)seeds.size();
        if (keep_unlinked) out.unlinked[(size_t)p].assign(seeds.begin(), seeds.end());
        link_seeds(seeds);
        out.linked[(size_t)p].assign(seeds.begin(), seeds.end());
    });
    out.scan_cells = 0;
    for (int64_t v : cells) out.scan_cells += v;
}

// ---------------------------------------------------------------------------
// x-drop extension.
// ---------------------------------------------------------------------------
void xdrop_extend(dcns_ctx *c, const Staged &S, const Params &P, const std::vector<XTask> &tasks,
                  std::vector<XResult> &out, int64_t &cells, bool score_only) {
    (void)S;
    const int32_t n = (int32_t)tasks.size();
    out.assign((size_t)n, XResult{});
    cells = 0;
    if (n == 0) return;
    upload(c, c->xtasks, tasks);
    c->xout.ensure((size_t)n * sizeof(XOut));
    c->ctr.ensure(sizeof(Counters) + 64);
    hip_check(hipMemsetAsync(c->xout.p, 0, (size_t)n * sizeof(XOut), c->stream), "hipMemsetAsync");
    int32_t *pending_dev = &c->ctr.as<Counters>()->pending;
    int32_t pending = 0;
    auto read_pending = [&]() {
        hip_check(hipMemcpyAsync(&pending, pending_dev, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream), "pending");
        c->sync();
    };
    // pass 1: one lane per task while the band fits the 21-cell lane
    // kernel, then one lane per task at 41 cells for the tasks that outgrew
    // it, then one wave per task for those that outgrew that; each kernel
    // lists its tasks that need a wider one (xlist, then xlist2).  A kernel
    // narrower than the initial band is skipped.
    int32_t nnext = n;
    const int32_t *in_list = nullptr;
    const int32_t B0 = 2 * P.w + 1;
    // rows a lane kernel's pass may run before the wave kernel takes the task
    const int32_t lane_rows = c->tune.xdrop_lane_rows > 0 ? c->tune.xdrop_lane_rows
                              : c->tune.xdrop_lane_rows < 0 ? INT32_MAX : kXdropLaneRows;
    c->xlist.ensure((size_t)n * sizeof(int32_t));
    c->xlist2.ensure((size_t)n * sizeof(int32_t));
    for (const int band : {kXdropLaneBand, kXdropLane2Band}) {
        if (B0 > band || nnext == 0) continue;
        int32_t *list = band == kXdropLaneBand ? c->xlist.as<int32_t>() : c->xlist2.as<int32_t>();
        hip_check(hipMemsetAsync(pending_dev, 0, sizeof(int32_t), c->stream), "hipMemsetAsync");
        c->launch(band == kXdropLaneBand ? "xdrop_pass1" : "xdrop_pass1_41", [&] {
            return launch_xdrop_pass1(c->stream, band, c->seq.as<int8_t>(), c->pairs.as<PairDesc>(),
                                      c->xtasks.as<XTask>(), in_list, nnext, P, c->xout.as<XOut>(), pending_dev, list,
                                      lane_rows);
        });
        read_pending();
        nnext = pending;
        in_list = list;
    }
    if (nnext > 0) {
        hip_check(hipMemsetAsync(pending_dev, 0, sizeof(int32_t), c->stream), "hipMemsetAsync");
        c->launch("xdrop_pass1_wave", [&] {
            return launch_xdrop_wave_pass1(c->stream, c->seq.as<int8_t>(), c->pairs.as<PairDesc>(),
                                           c->xtasks.as<XTask>(), in_list, nnext, P, c->xout.as<XOut>(), pending_dev);
        });
        read_pending();
    }
    c->h_xout.ensure((size_t)n * sizeof(XOut));
    XOut *xo = c->h_xout.as<XOut>();
    bool wide = false;
    if (pending) {
        // bands past the wave kernel's 192 cells: per-task scratch rows in global memory
        hip_check(hipMemcpyAsync(xo, c->xout.p, (size_t)n * sizeof(XOut), hipMemcpyDeviceToHost, c->stream), "xout");
        c->sync();
        std::vector<int64_t> off((size_t)n, -1);
        std::vector<int32_t> wmax((size_t)n, 0);
        int64_t total = 0;
        for (int32_t i = 0; i < n; ++i) {
            if (xo[i].status != 1 && xo[i].status != 2) continue;
            const PairDesc &pd = S.pd[(size_t)tasks[(size_t)i].pair];
            int32_t w = P.w;
            while (w < std::max(pd.len1, pd.len2)) w *= 2;
            wmax[(size_t)i] = w;
            off[(size_t)i] = total;
            total += 2 * (2 * (int64_t)w + 1);
        }
        upload(c, c->xw_off, off);
        upload(c, c->xw_wmax, wmax);
        c->xw_scr.ensure((size_t)std::max<int64_t>(total, 16) * 4);
        c->launch("xdrop_pass1", [&] {
            return launch_xdrop_wide_pass1(c->stream, c->seq.as<int8_t>(), c->pairs.as<PairDesc>(),
                                           c->xtasks.as<XTask>(), n, P, c->xout.as<XOut>(), c->xw_scr.as<int32_t>(),
                                           c->xw_off.as<int64_t>(), c->xw_wmax.as<int32_t>());
        });
        wide = true;
    }
    // traceback store and cigar arena sizes from pass 1
    hip_check(hipMemcpyAsync(xo, c->xout.p, (size_t)n * sizeof(XOut), hipMemcpyDeviceToHost, c->stream), "xout");
    c->sync();
    if (score_only) {
        // x_extend_seed without traceback (findSimilarFragmentsForPairedSequence4.cpp:205-241)
        for (int32_t i = 0; i < n; ++i) cells += xo[i].cells;
        parallel_for(n, [&](int64_t i) {
            const XOut &o = xo[i];
            const XTask &t = tasks[(size_t)i];
            XResult &r = out[(size_t)i];
            r.start1 = t.start1;
            r.start2 = t.start2;
            r.end1 = r.end2 = 0;
            r.ok = false;
            if (o.fe1 < 0 || o.fe2 < 0) {
                r.score = o.fmax;
                return;
            }
            r.score = o.rmax;
            if (o.re1 < 0 || o.re2 < 0) return;
            r.ok = true;
            r.end1 = t.start1 + o.fe1;
            r.end2 = t.start2 + o.fe2;
            r.start1 = t.start1 + o.fe1 - o.re1;
            r.start2 = t.start2 + o.fe2 - o.re2;
        });
        return;
    }
    // the band class of a task's pass 2: 0 and 1 the lane kernels, 2 the wave
    // kernel (also the lane bands' tasks of more than lane_rows rows), 3 the
    // scratch-band kernel
    auto band_class = [&](int32_t B, int32_t rows) {
        if (B > kXdropWaveBand) return 3;
        if (B > kXdropLane2Band || rows > lane_rows) return 2;
        return B <= kXdropLaneBand ? 0 : 1;
    };
    constexpr int NCLS = 4;
    std::vector<int64_t> tboff((size_t)n, 0);
    int64_t tbtotal = 0, cigtotal = 0;
    int32_t in_class[NCLS] = {0, 0, 0, 0};
    for (int32_t i = 0; i < n; ++i) {
        XOut &o = xo[i];
        cells += o.cells;
        o.cig_off = cigtotal;
        if (o.fe1 < 0 || o.fe2 < 0) continue;
        const int32_t B = 2 * o.rw + 1, BP = (B + 3) & ~3;
        tboff[(size_t)i] = tbtotal;
        tbtotal += (int64_t)(o.rrows + 1) * BP;
        cigtotal += std::max(0, o.re1) + std::max(0, o.re2) + 2;
        ++in_class[band_class(B, o.rrows)];
    }
    hip_check(hipMemcpyAsync(c->xout.p, xo, (size_t)n * sizeof(XOut), hipMemcpyHostToDevice, c->stream), "xout up");
    upload(c, c->tb_off, tboff);
    // pass 2 order: by band class (each kernel's tasks contiguous, the lane
    // classes first, then the wave class), then by traced cells, largest
    // first, on a quarter-octave scale (a counting sort; the order only
    // balances the launches)
    {
        constexpr int NB = 4 * 48;
        std::vector<int32_t> bucket((size_t)n);
        std::vector<int64_t> start(NCLS * NB + 2, 0);
        for (int32_t i = 0; i < n; ++i) {
            const XOut &o = xo[i];
            int32_t b = NCLS * NB;  // no traceback: last
            if (o.fe1 >= 0 && o.fe2 >= 0) {
                const int32_t B = 2 * o.rw + 1;
                const int cls = band_class(B, o.rrows);
                const uint64_t cl = (uint64_t)(o.rrows + 1) * (uint64_t)B;
                const int lg = 63 - __builtin_clzll(cl | 1);
                const int q = lg >= 2 ? (int)((cl >> (lg - 2)) & 3) : 0;
                b = cls * NB + (NB - 1 - std::min(NB - 1, 4 * lg + q));
            }
            bucket[(size_t)i] = b;
            ++start[(size_t)b + 1];
        }
        for (size_t b = 1; b < start.size(); ++b) start[b] += start[b - 1];
        std::vector<int32_t> order2((size_t)n);
        for (int32_t i = 0; i < n; ++i) order2[(size_t)start[(size_t)bucket[(size_t)i]]++] = i;
        upload(c, c->xorder, order2);
    }
    c->tb.ensure((size_t)std::max<int64_t>(tbtotal, 64));
    c->cigar.ensure((size_t)std::max<int64_t>(cigtotal, 16) * sizeof(uint32_t));
    if (wide) {
        c->launch("xdrop_pass2", [&] {
            return launch_xdrop_wide_pass2(c->stream, c->seq.as<int8_t>(), c->pairs.as<PairDesc>(),
                                           c->xtasks.as<XTask>(), n, P, c->xout.as<XOut>(), c->tb_off.as<int64_t>(),
                                           c->tb.as<uint8_t>(), c->cigar.as<uint32_t>(), c->xw_scr.as<int32_t>(),
                                           c->xw_off.as<int64_t>());
        });
    }
    if (in_class[0] > 0)
        c->launch("xdrop_pass2", [&] {
            return launch_xdrop_pass2(c->stream, kXdropLaneBand, c->seq.as<int8_t>(), c->pairs.as<PairDesc>(),
                                      c->xtasks.as<XTask>(), c->xorder.as<int32_t>(), in_class[0], P,
                                      c->xout.as<XOut>(), c->tb_off.as<int64_t>(), c->tb.as<uint8_t>(),
                                      c->cigar.as<uint32_t>());
        });
    if (in_class[1] > 0)
        c->launch("xdrop_pass2_41", [&] {
            return launch_xdrop_pass2(c->stream, kXdropLane2Band, c->seq.as<int8_t>(), c->pairs.as<PairDesc>(),
                                      c->xtasks.as<XTask>(), c->xorder.as<int32_t>() + in_class[0], in_class[1], P,
                                      c->xout.as<XOut>(), c->tb_off.as<int64_t>(), c->tb.as<uint8_t>(),
                                      c->cigar.as<uint32_t>());
        });
    if (in_class[2] > 0)
        c->launch("xdrop_pass2_wave", [&] {
            return launch_xdrop_wave_pass2(c->stream, c->seq.as<int8_t>(), c->pairs.as<PairDesc>(),
                                           c->xtasks.as<XTask>(), c->xorder.as<int32_t>() + in_class[0] + in_class[1],
                                           in_class[2], P, c->xout.as<XOut>(), c->tb_off.as<int64_t>(),
                                           c->tb.as<uint8_t>(), c->cigar.as<uint32_t>());
        });
    // the arena holds re1 + re2 + 2 entries per task, the walk writes a few
    // runs: pack them before the download
    const size_t scan_bytes = pack_cigars_scratch_bytes(n);
    c->cig_scan.ensure(scan_bytes);
    c->cig_packed.ensure((size_t)std::max<int64_t>(cigtotal, 16) * sizeof(uint32_t));
    hip_check(launch_pack_cigars(c->stream, c->xout.as<XOut>(), n, c->cig_scan.p, scan_bytes,
                                 c->cigar.as<uint32_t>(), c->cig_packed.as<uint32_t>()),
              "pack cigars");
    hip_check(hipMemcpyAsync(xo, c->xout.p, (size_t)n * sizeof(XOut), hipMemcpyDeviceToHost, c->stream), "xout");
    c->sync();
    int64_t packed = 0;
    for (int32_t i = 0; i < n; ++i)
        if (xo[i].fe1 >= 0 && xo[i].fe2 >= 0) packed += xo[i].ncig;
    c->h_cigar.ensure((size_t)std::max<int64_t>(packed, 16) * sizeof(uint32_t));
    const uint32_t *cig = c->h_cigar.as<uint32_t>();
    if (packed) {
        hip_check(hipMemcpyAsync(c->h_cigar.p, c->cig_packed.p, (size_t)packed * sizeof(uint32_t),
                                 hipMemcpyDeviceToHost, c->stream), "cigar");
        c->sync();
    }
    parallel_for(n, [&](int64_t i) {
        const XOut &o = xo[i];
        const XTask &t = tasks[(size_t)i];
        XResult &r = out[(size_t)i];
        r.start1 = t.start1;
        r.start2 = t.start2;
        r.end1 = r.end2 = 0;
        if (o.fe1 < 0 || o.fe2 < 0) {  // forward pass found no positive cell (:259-266)
            r.ok = false;
            r.score = o.fmax;
            return;
        }
        r.score = o.rmax;
        r.cig = cig + o.cig_off;
        r.ncig = o.ncig;
        if (o.re1 < 0 || o.re2 < 0) {
            r.ok = false;
            return;
        }
        r.ok = true;
        r.end1 = t.start1 + o.fe1 - o.iii;
        r.end2 = t.start2 + o.fe2 - o.jjj;
        r.start1 = t.start1 + o.fe1 - o.re1;
        r.start2 = t.start2 + o.fe2 - o.re2;
    });
}

// ---------------------------------------------------------------------------
// findSimilarFragmentsForPairedSequence, overload 1.
// ---------------------------------------------------------------------------

// Positions of the fragments of one pair by (start1, start2) and by (end1,
// end2), so that the dedup scan's first fragment with an equal start or an
// equal end is found without walking the list.
struct FragIndex {
    std::unordered_map<uint64_t, std::set<int32_t>> by_start, by_end;
    static uint64_t key(uint32_t a, uint32_t b) { return ((uint64_t)a << 32) | b; }
    void add(int32_t i, const PairedSimilarFragment &f) {
        by_start[key(f.getStart1(), f.getStart2())].insert(i);
        by_end[key(f.getEnd1(), f.getEnd2())].insert(i);
    }
    void remove(int32_t i, const PairedSimilarFragment &f) {
        by_start[key(f.getStart1(), f.getStart2())].erase(i);
        by_end[key(f.getEnd1(), f.getEnd2())].erase(i);
    }
    int32_t first_match(const PairedSimilarFragment &f) const {
        int32_t best = -1;
        auto a = by_start.find(key(f.getStart1(), f.getStart2()));
        if (a != by_start.end() && !a->second.empty()) best = *a->second.begin();
        auto b = by_end.find(key(f.getEnd1(), f.getEnd2()));
        if (b != by_end.end() && !b->second.empty() && (best < 0 || *b->second.begin() < best)) best = *b->second.begin();
        return best;
    }
};
void find_similar(dcns_ctx *c, const dcns_batch &b, const dcns_params &p, FragSet &out) {
    validate_params(p);
    Staged S = stage_batch(c, b, to_params(p));
    find_similar_staged(c, S, p, out);
}

void find_similar_staged(dcns_ctx *c, Staged &S, const dcns_params &p, FragSet &out) {
    validate_params(p);
    const Params P = to_params(p);
    relayout(c, S, P);
    const int32_t npairs = (int32_t)S.pd.size();
    const double t0 = now_ms();
    SeedResult sr;
    seed_sets(c, S, P, false, sr);
    const double t1 = now_ms();
    std::vector<XTask> tasks;
    {
        size_t nl = 0;
        for (const auto &v : sr.linked) nl += v.size();
        tasks.reserve(nl);
    }
    std::vector<int64_t> first_task((size_t)npairs + 1, 0);
    for (int32_t i = 0; i < npairs; ++i) {
        first_task[(size_t)i] = (int64_t)tasks.size();
        for (const Seed &s : sr.linked[(size_t)i]) {
            XTask t;
            t.pair = i;
            t.start1 = s.getStart1() + 1;  // :492-495
            t.start2 = s.getStart2() + 1;
            t.hint = s.getEnd1() - s.getStart1() + 1;
            tasks.push_back(t);
        }
    }
    first_task[(size_t)npairs] = (int64_t)tasks.size();
    std::vector<XResult> xr;
    int64_t xcells = 0;
    xdrop_extend(c, S, P, tasks, xr, xcells);
    const double t2 = now_ms();

    out.frags.clear();
    out.first.assign((size_t)npairs, 0);
    out.count.assign((size_t)npairs, 0);
    out.seeds_size = sr.before_link;
    std::vector<std::vector<PairedSimilarFragment>> per((size_t)npairs);
    parallel_for(npairs, [&](int64_t i) {
        const double L1d = (double)S.pd[(size_t)i].len1, L2d = (double)S.pd[(size_t)i].len2;
        std::vector<PairedSimilarFragment> &fr = per[(size_t)i];
        FragIndex idx;
        for (int64_t k = first_task[(size_t)i]; k < first_task[(size_t)i + 1]; ++k) {
            const XResult &r = xr[(size_t)k];
            if (!r.ok) continue;  // p = 100
            const double e = p.k * L1d * L2d * std::exp(-1.0 * p.lambda * (double)r.score);  // :282-283
            const double pv = 1.0 - std::exp(-1.0 * e);
            if (!(pv < p.pvalue)) continue;
            // dedup, :500-521: 1-based uint32 stored coordinates against 0-based int32 ones
            PairedSimilarFragment nf((uint32_t)(r.start1 + 1), (uint32_t)(r.end1 + 1), (uint32_t)(r.start2 + 1),
                                     (uint32_t)(r.end2 + 1), (uint32_t)r.score,
                                     std::vector<uint32_t>(r.cig, r.cig + r.ncig), pv, e);
            // the reference scans the fragments in order and stops at the first
            // one with the same start or the same end; the index finds that one
            const int32_t g_at = idx.first_match(nf);
            if (g_at < 0) {
                idx.add((int32_t)fr.size(), nf);
                fr.push_back(nf);
                continue;
            }
            PairedSimilarFragment &g = fr[(size_t)g_at];
            const bool by_start = g.getStart1() == nf.getStart1() && g.getStart2() == nf.getStart2();
            const bool take = by_start ? ((uint32_t)r.end1 >= g.getEnd1() && (uint32_t)r.end2 >= g.getEnd2())
                                       : ((uint32_t)r.start1 <= g.getStart1() && (uint32_t)r.start2 <= g.getStart2());
            if (take) {
                idx.remove(g_at, g);
                g = nf;
                idx.add(g_at, g);
            }
        }
    });
    for (int32_t i = 0; i < npairs; ++i) {
        out.first[(size_t)i] = (int64_t)out.frags.size();
        out.count[(size_t)i] = (int64_t)per[(size_t)i].size();
        for (auto &f : per[(size_t)i]) out.frags.push_back(std::move(f));
    }
    const double t3 = now_ms();
    dcns_stats &st = c->stats;
    st = dcns_stats{};
    st.scan_cells = sr.scan_cells;
    st.gpu_scan_cells = sr.gpu_cells;
    st.xdrop_cells = xcells;
    for (int64_t v : sr.before_link) st.seeds += v;
    for (auto &v : sr.linked) st.linked_seeds += (int64_t)v.size();
    st.fragments = (int64_t)out.frags.size();
    st.levels = sr.levels;
    st.ms_seed = t1 - t0;
    st.ms_xdrop = t2 - t1;
    st.ms_replay = t3 - t2;
    st.ms_total = t3 - t0;
}

}  // namespace dcns
