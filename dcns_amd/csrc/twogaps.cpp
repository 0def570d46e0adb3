// twogaps.cpp -- host orchestration of the two-piece-affine path:
//   findSimilarFragmentsForPairedSequence overload 4
//   (impl/findSimilarFragmentsForPairedSequence4.cpp:752-774) =
//   getAllExtendSeed (:348-429) followed by overload 2 (:532-616),
// with every SemiGlobal call (impl/sequenceAlignment3.cpp:932 -> :681) run as
// batched device tasks; and of the weighted paths _wighted (:830-925) and
// _wighted_1gap (:779-826), the same flow with the weighted DP
// (sequenceAlignment3.cpp:1483-1890).
//
// Order of work per batch: seed sets (device scan + host std::set replay),
// score-only x-drop of every linked seed (device), the x-seed filter and
// merge (host, per pair, in set order), the reverse SemiGlobal of every
// x-seed (device, one batch), the StartAlignment set (host), the forward
// SemiGlobal with traceback of every new start (device, one batch), and the
// fragment replace/append pass (host, in x-seed order).  The host passes keep
// the reference's sequential semantics; only independent DPs are batched.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <chrono>
#include <cmath>
#include <map>
#include <set>
#include <unordered_map>
#include <utility>

#include "kernels.h"
#include "pipeline.h"

namespace dcns {
namespace {

double ms_now() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

constexpr int32_t kNeg = -536870912;

void validate_two_piece(const dcns_params &p) {
    if (p.open2 > 0 || p.open2 < -1000000 || p.ext2 > 0 || p.ext2 < -10000)
        throw Error(DCNS_ERR_ARG, "second gap piece: open2 in [-1e6, 0], ext2 in [-1e4, 0]");
    if (p.zdrop < 0) throw Error(DCNS_ERR_ARG, "zdrop must be >= 0");
    if (p.bandwidth < 2) throw Error(DCNS_ERR_ARG, "bandwidth must be >= 2");
}

}  // namespace

// DCNS_PROGRESS in the environment: the two-piece path reports its task
// lists and every SemiGlobal launch chunk on stderr as they finish.
bool progress() {
    static const bool on = std::getenv("DCNS_PROGRESS") != nullptr;
    return on;
}

void semiglobal_batch(dcns_ctx *c, const Params &P, const std::vector<SgTask> &tasks, bool trace,
                      std::vector<SgResult> &out, int64_t &cells, const SgWeights *Wp) {
    const SgWeights W = Wp ? *Wp : SgWeights{nullptr, nullptr, nullptr, 0, 0};
    const int32_t n = (int32_t)tasks.size();
    out.assign((size_t)n, SgResult{0, -2, -2, {}, 0});
    cells = 0;
    std::vector<int32_t> idx;
    for (int32_t i = 0; i < n; ++i)
        if (tasks[(size_t)i].R > 0 && tasks[(size_t)i].C > 0) idx.push_back(i);
    if (idx.empty()) return;
    // largest first, so the strips of the long tasks start early
    std::stable_sort(idx.begin(), idx.end(), [&](int32_t a, int32_t b) {
        return (int64_t)tasks[(size_t)a].R * tasks[(size_t)a].C > (int64_t)tasks[(size_t)b].R * tasks[(size_t)b].C;
    });
    size_t free_b = 0, total_b = 0;
    hip_check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo");
    const int64_t budget = std::max<int64_t>(std::min<int64_t>((int64_t)(free_b / 2), (int64_t)96 << 30),
                                             (int64_t)1 << 30);
    // a chunk also stops at this many band cells (a few seconds of DP), so
    // that very large task lists finish chunk by chunk
    const int64_t cell_budget = (int64_t)2 << 40;
    int64_t band_total = 0;
    for (int32_t i : idx) band_total += sg_trace_bytes(tasks[(size_t)i].R, tasks[(size_t)i].C, P.bandwidth);
    if (progress())
        std::fprintf(stderr, "SemiGlobal%s: %zu tasks, %.3e band cells\n", trace ? " with traceback" : "", idx.size(),
                     (double)band_total);
    int64_t band_done = 0;
    size_t pos = 0;
    while (pos < idx.size()) {
        // one chunk: tasks whose traceback stores and scratch fit the budget
        std::vector<SgTask> ct;
        std::vector<int32_t> which;
        std::vector<SgOut> so;
        int64_t tb_total = 0, scr_total = 0, cig_total = 0, band_chunk = 0;
        const double tc0 = ms_now();
        while (pos < idx.size()) {
            SgTask t = tasks[(size_t)idx[pos]];
            const int64_t band = sg_trace_bytes(t.R, t.C, P.bandwidth);
            const int64_t tbb = trace ? (band + 255) / 256 * 256 : 0;
            const int64_t scr = sg_scratch_ints(t.R, t.C, P.bandwidth);
            if (!ct.empty() && (tb_total + tbb + 4 * (scr_total + scr) > budget || band_chunk + band > cell_budget))
                break;
            band_chunk += band;
            if (ct.empty() && tbb + 4 * scr > budget)
                throw Error(DCNS_ERR_CAPACITY, "a SemiGlobal task's traceback store exceeds device memory");
            t.tb = trace ? tb_total : -1;
            t.scr = scr_total;
            tb_total += tbb;
            scr_total += scr;
            SgOut o{};
            o.cig_off = cig_total;
            o.stale[0] = o.stale[1] = kNeg;
            if (trace) cig_total += (int64_t)t.R + t.C + 2;
            ct.push_back(t);
            so.push_back(o);
            which.push_back(idx[pos]);
            ++pos;
        }
        const int32_t m = (int32_t)ct.size();
        // the strips of every task, a task's strips consecutive: each strip's
        // boundary records follow its task's row results in the scratch
        std::vector<SgStrip> strips;
        for (int32_t k = 0; k < m; ++k) {
            SgTask &t = ct[(size_t)k];
            t.sched = (int64_t)strips.size();
            t.nstrips = sg_task_strips(t.R, t.C, P.bandwidth, k, t.scr + 2 * ((int64_t)t.R + 2), strips);
        }
        const int32_t ns = (int32_t)strips.size();
        int64_t max_rows = 0, max_cols = 0;
        for (const SgTask &t : ct) {
            max_rows = std::max<int64_t>(max_rows, t.R);
            max_cols = std::max<int64_t>(max_cols, t.C);
        }
        const bool keyed = sg_keyed_ok(P, max_rows, max_cols);
        // Launch order: longest remaining chain first.  A strip's chain is the
        // steps from its start to its task's end: its own rows, or the next
        // strip's chain one hand-over (two 64-step blocks) later, whichever
        // ends last.  The chain shrinks along a task, so every strip comes
        // after its predecessor; the long tasks' strips start first and the
        // short tasks fill the device around them.
        std::vector<int64_t> chain((size_t)ns);
        for (int32_t k = m - 1; k >= 0; --k) {
            const SgTask &t = ct[(size_t)k];
            int64_t next = 0;
            for (int64_t i = t.sched + t.nstrips - 1; i >= t.sched; --i) {
                const SgStrip &g = strips[(size_t)i];
                next = std::max<int64_t>((int64_t)g.rhi - g.rlo + 64, next + 128);
                chain[(size_t)i] = next;
            }
        }
        std::vector<int32_t> order((size_t)ns);
        for (int32_t i = 0; i < ns; ++i) order[(size_t)i] = i;
        std::stable_sort(order.begin(), order.end(),
                         [&](int32_t a, int32_t b) { return chain[(size_t)a] > chain[(size_t)b]; });
        // The strips whose chain is within a quarter of the longest set the
        // launch's end: they issue ahead of the waves that share their SIMD.
        if (ns > 0) {
            const int64_t crit = chain[(size_t)order[0]] - chain[(size_t)order[0]] / 4;
            for (int32_t &o : order)
                if (chain[(size_t)o] >= crit) o |= 0x40000000;
        }
        c->sg_strips.ensure(std::max<size_t>(strips.size(), 1) * sizeof(SgStrip));
        hip_check(hipMemcpyAsync(c->sg_strips.p, strips.data(), strips.size() * sizeof(SgStrip), hipMemcpyHostToDevice,
                                 c->stream), "sg strips");
        c->sg_order.ensure(std::max<size_t>(order.size(), 1) * sizeof(int32_t));
        hip_check(hipMemcpyAsync(c->sg_order.p, order.data(), order.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                                 c->stream), "sg order");
        c->sg_sync.ensure((size_t)(ns + 2) * 4);
        hip_check(hipMemsetAsync(c->sg_sync.p, 0, (size_t)(ns + 2) * 4, c->stream), "sg sync");
        c->sg_tasks.ensure((size_t)m * sizeof(SgTask));
        c->sg_out.ensure((size_t)m * sizeof(SgOut));
        c->sg_scr.ensure((size_t)std::max<int64_t>(scr_total, 16) * 4);
        if (trace) {
            c->sg_tb.ensure((size_t)std::max<int64_t>(tb_total, 256));
            c->sg_cig.ensure((size_t)std::max<int64_t>(cig_total, 16) * 4);
        }
        hip_check(hipMemcpyAsync(c->sg_tasks.p, ct.data(), (size_t)m * sizeof(SgTask), hipMemcpyHostToDevice,
                                 c->stream), "sg tasks");
        hip_check(hipMemcpyAsync(c->sg_out.p, so.data(), (size_t)m * sizeof(SgOut), hipMemcpyHostToDevice, c->stream),
                  "sg out");
        c->launch("semiglobal", [&] {
            return launch_sg_dp(c->stream, trace, keyed, c->seq.as<int8_t>(), c->pairs.as<PairDesc>(), c->sg_tasks.as<SgTask>(),
                                c->sg_strips.as<SgStrip>(), c->sg_order.as<int32_t>(), ns, P, c->sg_scr.as<int32_t>(),
                                trace ? c->sg_tb.as<uint8_t>() : nullptr, c->sg_out.as<SgOut>(), W,
                                c->sg_sync.as<int32_t>());
        });
        c->launch("semiglobal_finish", [&] {
            return launch_sg_finish(c->stream, c->seq.as<int8_t>(), c->pairs.as<PairDesc>(), c->sg_tasks.as<SgTask>(),
                                    c->sg_strips.as<SgStrip>(), m, P, c->sg_scr.as<int32_t>(),
                                    trace ? c->sg_tb.as<uint8_t>() : nullptr, c->sg_out.as<SgOut>(),
                                    trace ? c->sg_cig.as<uint32_t>() : nullptr, W, trace && keyed && W.mode == 0);
        });
        int32_t stalled = 0;
        hip_check(hipMemcpyAsync(&stalled, c->sg_sync.as<int32_t>() + 1, 4, hipMemcpyDeviceToHost, c->stream),
                  "sg error word");
        hip_check(hipMemcpyAsync(so.data(), c->sg_out.p, (size_t)m * sizeof(SgOut), hipMemcpyDeviceToHost, c->stream),
                  "sg out");
        std::vector<uint32_t> cig;
        if (trace && cig_total) {
            cig.resize((size_t)cig_total);
            hip_check(hipMemcpyAsync(cig.data(), c->sg_cig.p, (size_t)cig_total * 4, hipMemcpyDeviceToHost, c->stream),
                      "sg cigar");
        }
        c->sync();
        if (stalled) throw Error(DCNS_ERR_INTERNAL, "SemiGlobal strip pipeline: a strip waited too long for its predecessor");
        band_done += band_chunk;
        if (progress())
            std::fprintf(stderr,
                         "  chunk of %d tasks, %d strips (longest chain %lld steps): %.3e band cells in %.1f ms "
                         "(%.3e of %.3e done)\n",
                         m, ns, (long long)(ns ? chain[(size_t)(order[0] & 0x3fffffff)] : 0), (double)band_chunk,
                         ms_now() - tc0, (double)band_done, (double)band_total);
        for (int32_t k = 0; k < m; ++k) {
            const SgOut &o = so[(size_t)k];
            SgResult &r = out[(size_t)which[(size_t)k]];
            r.max = o.max;
            r.cells = o.cells;
            r.end1 = o.e1 - 1;  // the reference's final decrement (:921-922)
            r.end2 = o.e2 - 1;
            cells += o.cells;
            if (trace && o.ncig > 0) {
                r.cigar.assign(cig.begin() + o.cig_off, cig.begin() + o.cig_off + o.ncig);
                std::reverse(r.cigar.begin(), r.cigar.end());
            }
        }
    }
}

void find_similar_2gaps(dcns_ctx *c, const dcns_batch &b, const dcns_params &p, FragSet &out) {
    validate_params(p);
    validate_two_piece(p);
    Staged S = stage_batch(c, b, to_params(p));
    find_similar_2gaps_staged(c, S, p, out);
}

namespace {

// getAllExtendSeed, then overload 2's loop (W.mode 0) or the weighted loops
// (W.mode 1: _wighted, 2: _wighted_1gap).
void two_gaps_flow(dcns_ctx *c, Staged &S, const dcns_params &p, const SgWeights &W, FragSet &out) {
    const Params P = to_params(p);
    relayout(c, S, P);
    const int32_t npairs = (int32_t)S.pd.size();
    const double t0 = ms_now();
    // getAllExtendSeed: the same window loop and seed set as overload 1 (the
    // loop at :377-392 differs only by a window at sp > L2, whose negative
    // width yields no seed), then link_seeds
    SeedResult sr;
    seed_sets(c, S, P, false, sr);
    const double t1 = ms_now();
    std::vector<XTask> xt;
    std::vector<int64_t> first((size_t)npairs + 1, 0);
    for (int32_t i = 0; i < npairs; ++i) {
        first[(size_t)i] = (int64_t)xt.size();
        for (const Seed &s : sr.linked[(size_t)i])
            xt.push_back(XTask{i, s.getStart1() + 1, s.getStart2() + 1, s.getEnd1() - s.getStart1() + 1});
    }
    first[(size_t)npairs] = (int64_t)xt.size();
    std::vector<XResult> xr;
    int64_t xcells = 0;
    xdrop_extend(c, S, P, xt, xr, xcells, true);
    // x-seed filter and merge (:400-427): no break, every matching x-seed is visited
    std::vector<std::vector<Seed>> xs((size_t)npairs);
    parallel_for(npairs, [&](int64_t i) {
        const double L1d = (double)S.pd[(size_t)i].len1, L2d = (double)S.pd[(size_t)i].len2;
        std::vector<Seed> &v = xs[(size_t)i];
        for (int64_t k = first[(size_t)i]; k < first[(size_t)i + 1]; ++k) {
            const XResult &r = xr[(size_t)k];
            if (!r.ok) continue;  // p = 100
            const double e = p.k * L1d * L2d * std::exp(-1.0 * p.lambda * (double)r.score);
            const double pv = 1.0 - std::exp(-1.0 * e);
            if (!(pv < p.pvalue)) continue;
            bool never_used = true;
            for (Seed &x : v) {
                if (x.getStart1() == r.start1 && x.getStart2() == r.start2) {
                    never_used = false;
                    if (r.end1 > x.getEnd1() && r.end2 > x.getEnd2()) {
                        x.setEnd1(r.end1);
                        x.setEnd2(r.end2);
                    }
                } else if (x.getEnd1() == r.end1 && x.getEnd2() == r.end2) {
                    if (r.start1 < x.getStart1() && r.start2 < x.getStart2()) {
                        x.setStart1(r.start1);
                        x.setStart2(r.start2);
                    }
                    never_used = false;
                }
            }
            if (never_used) v.push_back(Seed(r.start1, r.end1, r.start2, r.end2));
        }
    });
    const double t2 = ms_now();
    if (progress()) {
        size_t nx = 0;
        for (const auto &v : xs) nx += v.size();
        std::fprintf(stderr, "two-piece: %zu linked seeds, %zu x-seeds after the p-value filter and merge (%.0f ms)\n",
                     xt.size(), nx, t2 - t0);
    }
    // overload 2: reverse SemiGlobal (no traceback) from every x-seed's end
    std::vector<SgTask> rt;
    std::vector<int64_t> xfirst((size_t)npairs + 1, 0);
    for (int32_t i = 0; i < npairs; ++i) {
        xfirst[(size_t)i] = (int64_t)rt.size();
        for (const Seed &x : xs[(size_t)i]) {
            SgTask t{};
            t.pair = i;
            t.rev = 1;
            t.a1 = x.getEnd1();
            t.a2 = x.getEnd2();
            t.R = x.getEnd1() + 1;
            t.C = x.getEnd2() + 1;
            rt.push_back(t);
        }
    }
    xfirst[(size_t)npairs] = (int64_t)rt.size();
    std::vector<SgResult> rr;
    int64_t sgcells = 0, c1 = 0, c2 = 0;
    semiglobal_batch(c, P, rt, false, rr, c1, &W);
    // StartAlignment set (:572-574), then the forward SemiGlobal with traceback
    std::vector<SgTask> ft;
    std::vector<int64_t> ffirst((size_t)npairs + 1, 0);
    for (int32_t i = 0; i < npairs; ++i) {
        ffirst[(size_t)i] = (int64_t)ft.size();
        std::set<std::pair<int32_t, int32_t>> starts;
        const int32_t L1 = S.pd[(size_t)i].len1, L2 = S.pd[(size_t)i].len2;
        for (int64_t k = xfirst[(size_t)i]; k < xfirst[(size_t)i + 1]; ++k) {
            const Seed &x = xs[(size_t)i][(size_t)(k - xfirst[(size_t)i])];
            const int32_t start1 = x.getEnd1() - rr[(size_t)k].end1, start2 = x.getEnd2() - rr[(size_t)k].end2;
            if (!starts.insert({start1, start2}).second) continue;
            SgTask t{};
            t.pair = i;
            t.rev = 0;
            t.a1 = start1;
            t.a2 = start2;
            t.R = L1 - start1;
            t.C = L2 - start2;
            ft.push_back(t);
        }
    }
    ffirst[(size_t)npairs] = (int64_t)ft.size();
    std::vector<SgResult> fr;
    semiglobal_batch(c, P, ft, true, fr, c2, &W);
    sgcells = c1 + c2;
    const double t3 = ms_now();
    // fragment replace / append (:584-611), p-values without a filter
    std::vector<std::vector<PairedSimilarFragment>> per((size_t)npairs);
    parallel_for(npairs, [&](int64_t i) {
        const double L1d = (double)S.pd[(size_t)i].len1, L2d = (double)S.pd[(size_t)i].len2;
        std::vector<PairedSimilarFragment> &v = per[(size_t)i];
        for (int64_t k = ffirst[(size_t)i]; k < ffirst[(size_t)i + 1]; ++k) {
            const SgResult &r = fr[(size_t)k];
            const int32_t start1 = ft[(size_t)k].a1, start2 = ft[(size_t)k].a2;
            // _wighted_1gap keeps every forward result (:795-822)
            if (W.mode != 2 && !(r.cigar.size() > 0 && r.end1 > 0 && r.end2 > 0)) continue;
            const int32_t end1 = r.end1 + start1, end2 = r.end2 + start2;
            int32_t to_remove = -1;
            bool found = false;
            for (size_t q = 0; q < v.size(); ++q) {
                if (v[q].getEnd1() == (uint32_t)(end1 + 1) && v[q].getEnd2() == (uint32_t)(end2 + 1)) {
                    found = true;
                    if ((uint32_t)(start1 + 1) < v[q].getStart1() && (uint32_t)(start2 + 1) < v[q].getStart2())
                        to_remove = (int32_t)q;
                }
            }
            if (found && to_remove < 0) continue;
            const double e = p.k * L1d * L2d * std::exp(-1.0 * p.lambda * (double)r.max);
            const double pv = 1.0 - std::exp(-1.0 * e);
            PairedSimilarFragment f((uint32_t)(start1 + 1), (uint32_t)(end1 + 1), (uint32_t)(start2 + 1),
                                    (uint32_t)(end2 + 1), (uint32_t)r.max, r.cigar, pv, e);
            if (to_remove >= 0)
                v[(size_t)to_remove] = f;
            else
                v.push_back(f);
        }
    });
    out.frags.clear();
    out.first.assign((size_t)npairs, 0);
    out.count.assign((size_t)npairs, 0);
    out.seeds_size = sr.before_link;
    for (int32_t i = 0; i < npairs; ++i) {
        out.first[(size_t)i] = (int64_t)out.frags.size();
        out.count[(size_t)i] = (int64_t)per[(size_t)i].size();
        for (auto &f : per[(size_t)i]) out.frags.push_back(std::move(f));
    }
    const double t4 = ms_now();
    dcns_stats &st = c->stats;
    st = dcns_stats{};
    st.scan_cells = sr.scan_cells;
    st.gpu_scan_cells = sr.gpu_cells;
    st.xdrop_cells = xcells;
    st.semiglobal_cells = sgcells;
    for (int64_t v : sr.before_link) st.seeds += v;
    for (auto &v : sr.linked) st.linked_seeds += (int64_t)v.size();
    st.fragments = (int64_t)out.frags.size();
    st.levels = sr.levels;
    st.ms_seed = t1 - t0;
    st.ms_xdrop = t2 - t1;
    st.ms_semiglobal = t3 - t2;
    st.ms_replay = t4 - t3;
    st.ms_total = t4 - t0;
}

}  // namespace

void find_similar_2gaps_staged(dcns_ctx *c, Staged &S, const dcns_params &p, FragSet &out) {
    validate_params(p);
    validate_two_piece(p);
    two_gaps_flow(c, S, p, SgWeights{nullptr, nullptr, nullptr, 0, 0}, out);
}

void find_similar_seeded(dcns_ctx *c, const dcns_batch &b, const dcns_params &p,
                         const std::vector<std::vector<Seed>> &seeds, int32_t max_len, FragSet &out) {
    validate_params(p);
    validate_two_piece(p);
    if (max_len < 1) throw Error(DCNS_ERR_ARG, "maximumAlignLength must be >= 1");
    if ((int32_t)seeds.size() != b.n) throw Error(DCNS_ERR_ARG, "one seed list per pair");
    const Params P = to_params(p);
    const double t0 = ms_now();
    Staged S = stage_batch(c, b, P);
    const int32_t npairs = b.n;
    // every seed's reverse SemiGlobal (:657-670), speculatively: whether a seed
    // is used depends on the fragments of the seeds before it, the DP does not
    std::vector<SgTask> rt;
    std::vector<std::vector<int64_t>> rtask((size_t)npairs);  // -1: seed outside the sequences
    std::vector<std::vector<std::pair<int32_t, int32_t>>> pend((size_t)npairs);  // patched ends
    for (int32_t i = 0; i < npairs; ++i) {
        const int32_t L1 = b.len1[i], L2 = b.len2[i];
        const int8_t *s1 = b.seqs + b.off1[i], *s2 = b.seqs + b.off2[i];
        for (const Seed &x : seeds[(size_t)i]) {
            int32_t e1 = x.getEnd1(), e2 = x.getEnd2();
            if (e1 < 0 || e1 >= L1 || e2 < 0 || e2 >= L2) {  // undefined in the reference
                rtask[(size_t)i].push_back(-1);
                pend[(size_t)i].push_back({0, 0});
                continue;
            }
            SgTask t{};
            t.pair = i;
            t.rev = 1;
            t.R = std::min(1 + e1, max_len);
            t.C = std::min(1 + e2, max_len);
            if (s1[e1] != s2[e2]) {  // the reference's patch (:661-665); the DP keeps R x C
                --e1;
                --e2;
            }
            t.a1 = e1;
            t.a2 = e2;
            rtask[(size_t)i].push_back((int64_t)rt.size());
            pend[(size_t)i].push_back({e1, e2});
            rt.push_back(t);
        }
    }
    std::vector<SgResult> rr;
    int64_t c1 = 0, c2 = 0;
    semiglobal_batch(c, P, rt, false, rr, c1);
    // the forward SemiGlobal with traceback of every distinct start
    std::vector<SgTask> ft;
    std::vector<std::map<std::pair<int32_t, int32_t>, int64_t>> ftask((size_t)npairs);
    for (int32_t i = 0; i < npairs; ++i) {
        const int32_t L1 = b.len1[i], L2 = b.len2[i];
        for (size_t k = 0; k < seeds[(size_t)i].size(); ++k) {
            const int64_t r = rtask[(size_t)i][k];
            if (r < 0) continue;
            const int32_t s1 = pend[(size_t)i][k].first - rr[(size_t)r].end1,
                          s2 = pend[(size_t)i][k].second - rr[(size_t)r].end2;
            if (ftask[(size_t)i].count({s1, s2})) continue;
            SgTask t{};
            t.pair = i;
            t.rev = 0;
            t.a1 = s1;
            t.a2 = s2;
            t.R = (s1 < 0 || s2 < 0) ? 0 : std::min(L1 - s1, max_len);  // outside: no DP
            t.C = (s1 < 0 || s2 < 0) ? 0 : std::min(L2 - s2, max_len);
            ftask[(size_t)i][{s1, s2}] = (int64_t)ft.size();
            ft.push_back(t);
        }
    }
    std::vector<SgResult> fr;
    semiglobal_batch(c, P, ft, true, fr, c2);
    const double t1 = ms_now();
    // the reference's loop (:649-742), per pair in seed order
    std::vector<std::vector<PairedSimilarFragment>> per((size_t)npairs);
    std::vector<int64_t> cells((size_t)npairs, 0);
    parallel_for(npairs, [&](int64_t i) {
        const double L1d = (double)b.len1[i], L2d = (double)b.len2[i];
        std::vector<PairedSimilarFragment> &v = per[(size_t)i];
        std::set<std::pair<int32_t, int32_t>> starts;
        for (size_t k = 0; k < seeds[(size_t)i].size(); ++k) {
            const Seed &x = seeds[(size_t)i][k];
            bool use = true;
            for (const PairedSimilarFragment &f : v)  // 1-based uint32 against 0-based seeds
                if (f.getStart1() <= (uint32_t)x.getStart1() && f.getEnd1() >= (uint32_t)x.getEnd1() &&
                    f.getStart2() <= (uint32_t)x.getStart2() && f.getEnd2() >= (uint32_t)x.getEnd2()) {
                    use = false;
                    break;
                }
            if (!use) continue;
            const int64_t r = rtask[(size_t)i][k];
            if (r < 0) continue;
            cells[(size_t)i] += rr[(size_t)r].cells;
            const int32_t start1 = pend[(size_t)i][k].first - rr[(size_t)r].end1,
                          start2 = pend[(size_t)i][k].second - rr[(size_t)r].end2;
            if (!starts.insert({start1, start2}).second) continue;
            const SgResult &g = fr[(size_t)ftask[(size_t)i].at({start1, start2})];
            cells[(size_t)i] += g.cells;
            if (!(g.cigar.size() > 0 && g.end1 > 0 && g.end2 > 0)) continue;  // the reference reports an error
            const int32_t end1 = g.end1 + start1, end2 = g.end2 + start2;
            int32_t to_remove = -1;
            bool found = false;
            for (size_t q = 0; q < v.size(); ++q) {
                if (v[q].getEnd1() == (uint32_t)(end1 + 1) && v[q].getEnd2() == (uint32_t)(end2 + 1)) {
                    found = true;
                    if ((uint32_t)(start1 + 1) < v[q].getStart1() && (uint32_t)(start2 + 1) < v[q].getStart2())
                        to_remove = (int32_t)q;
                }
            }
            if (found && to_remove < 0) continue;
            const double e = p.k * L1d * L2d * std::exp(-1.0 * p.lambda * (double)g.max);
            const double pv = 1.0 - std::exp(-1.0 * e);
            PairedSimilarFragment f((uint32_t)(start1 + 1), (uint32_t)(end1 + 1), (uint32_t)(start2 + 1),
                                    (uint32_t)(end2 + 1), (uint32_t)g.max, g.cigar, pv, e);
            if (to_remove >= 0)
                v[(size_t)to_remove] = f;
            else
                v.push_back(f);
        }
    });
    out.frags.clear();
    out.first.assign((size_t)npairs, 0);
    out.count.assign((size_t)npairs, 0);
    out.seeds_size.assign((size_t)npairs, 0);
    int64_t total = 0;
    for (int32_t i = 0; i < npairs; ++i) {
        out.first[(size_t)i] = (int64_t)out.frags.size();
        out.count[(size_t)i] = (int64_t)per[(size_t)i].size();
        out.seeds_size[(size_t)i] = (int64_t)seeds[(size_t)i].size();
        for (auto &f : per[(size_t)i]) out.frags.push_back(std::move(f));
        total += cells[(size_t)i];
    }
    const double t2 = ms_now();
    dcns_stats &st = c->stats;
    st = dcns_stats{};
    st.semiglobal_cells = total;
    st.fragments = (int64_t)out.frags.size();
    st.ms_semiglobal = t1 - t0;
    st.ms_replay = t2 - t1;
    st.ms_total = t2 - t0;
}

SgWeights stage_weights(dcns_ctx *c, const dcns_batch &b, const std::vector<Profile> &prof, const PairCategories &pc,
                        int mode) {
    if (prof.empty() || (int)prof.size() > kMaxProfiles)
        throw Error(DCNS_ERR_ARG, "1 to 64 score profiles are supported");
    if (!pc.cat || !pc.off) throw Error(DCNS_ERR_ARG, "categories are required");
    // dense profile indices; a category without a profile is an error (the
    // reference dereferences a null score matrix for it)
    std::unordered_map<int32_t, uint8_t> index;
    std::vector<int32_t> tab(prof.size() * 8, 0);
    for (size_t i = 0; i < prof.size(); ++i) {
        const Profile &f = prof[i];
        if (!index.emplace((int32_t)f.category, (uint8_t)i).second)
            throw Error(DCNS_ERR_ARG, "duplicate score profile category " + std::to_string(f.category));
        const int32_t v[8] = {f.match, f.mismatch, f.open1, f.ext1, f.open2, f.ext2, f.zdrop, 0};
        std::copy(v, v + 8, tab.begin() + 8 * (int64_t)i);
    }
    const int32_t np = b.n;
    std::vector<int64_t> off((size_t)std::max(np, 1), 0);
    int64_t total = 0;
    for (int32_t i = 0; i < np; ++i) {
        off[(size_t)i] = total;
        total += (int64_t)b.len1[i] + 2;
    }
    std::vector<uint8_t> dense((size_t)std::max<int64_t>(total, 1));
    for (int32_t i = 0; i < np; ++i) {
        const int16_t *src = pc.cat + pc.off[i];
        for (int64_t x = 0; x < (int64_t)b.len1[i] + 2; ++x) {
            auto it = index.find((int32_t)src[x]);
            if (it == index.end())
                throw Error(DCNS_ERR_ARG, "category " + std::to_string(src[x]) + " has no score profile");
            dense[(size_t)(off[(size_t)i] + x)] = it->second;
        }
    }
    c->sg_wcat.ensure(dense.size());
    c->sg_wcat_off.ensure(off.size() * sizeof(int64_t));
    c->sg_wtab.ensure(tab.size() * sizeof(int32_t));
    hip_check(hipMemcpyAsync(c->sg_wcat.p, dense.data(), dense.size(), hipMemcpyHostToDevice, c->stream), "categories");
    hip_check(hipMemcpyAsync(c->sg_wcat_off.p, off.data(), off.size() * sizeof(int64_t), hipMemcpyHostToDevice,
                             c->stream), "category offsets");
    hip_check(hipMemcpyAsync(c->sg_wtab.p, tab.data(), tab.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                             c->stream), "score profiles");
    c->sync();  // the host vectors end here
    return SgWeights{c->sg_wcat.as<uint8_t>(), c->sg_wcat_off.as<int64_t>(), c->sg_wtab.as<int32_t>(),
                     (int32_t)prof.size(), mode};
}

void find_similar_weighted(dcns_ctx *c, const dcns_batch &b, const dcns_params &p,
                           const std::vector<Profile> &prof, const PairCategories &pc, bool two_piece, FragSet &out) {
    validate_params(p);
    if (p.bandwidth < 1) throw Error(DCNS_ERR_ARG, "bandwidth must be >= 1");
    Staged S = stage_batch(c, b, to_params(p));
    const SgWeights W = stage_weights(c, b, prof, pc, two_piece ? 1 : 2);
    two_gaps_flow(c, S, p, W, out);
}

}  // namespace dcns
