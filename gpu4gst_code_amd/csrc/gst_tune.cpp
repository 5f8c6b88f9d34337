// Autotuner: picks the solver's launch knobs for one graph by timing a sample of queries.
//
// One-query solves: the frontier threshold at which a grow round pulls over every vertex
// (pull_permille), the arcs per frontier work item (chunk_arcs), the lanes serving one work item in
// push rounds (push_lanes: a wave serves 64/push_lanes items, each item's masks spread over its lanes)
// and the merge tile (merge_tile_kb: the LDS budget that sets the vertices per merge workgroup).
// Coordinate descent from the built-in defaults, two sweeps, a change kept only when it is at
// least 2% faster.  Batched engine: push lanes, then rounds per step, timed as a 64-slot stream
// over the sample.
// The choice is stored in the context and used by every later solve on the graph that leaves the
// knob at 0.  No counterpart in the reference (its launch shapes are fixed, src/DPQ.cu:597-610).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "gst_batch.h"
#include "gst_internal.h"

namespace {

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Knobs {
    int32_t pull_permille, chunk_arcs, push_lanes, merge_tile_kb;
};

}  // namespace

extern "C" gst_status gst_autotune(gst_ctx* ctx, const gst_graph* G, int32_t nq, const gst_groups* sample,
                                   int32_t flags, gst_tune_report* rep) {
    if (!ctx || !G || nq < 1 || !sample) return ctx ? gst_set_error(ctx, GST_INVALID_ARG, "autotune: null argument or empty sample") : GST_INVALID_ARG;
    if (G->ctx != ctx) return gst_set_error(ctx, GST_INVALID_ARG, "graph belongs to another context");
    if (rep) std::memset(rep, 0, sizeof(*rep));
    const bool bounded = (flags & GST_TUNE_BOUNDED) != 0;
    const bool batch = (flags & GST_TUNE_BATCH) != 0;
    const bool eager = (flags & GST_TUNE_EAGER) != 0;
    int g = sample[0].n_groups;
    for (int32_t q = 1; q < nq; ++q)
        if (sample[q].n_groups != g) return gst_set_error(ctx, GST_INVALID_ARG, "autotune: the sample must share one group count");
    // forget an earlier choice for this graph, so that the defaults are what is measured first
    ctx->tuned.erase(std::remove_if(ctx->tuned.begin(), ctx->tuned.end(),
                                    [&](const gst_tuned& t) { return t.uid == G->uid && t.g == g && t.ub == bounded; }),
                     ctx->tuned.end());
    // optima of the sample (the bounds of a bounded sweep: 5% above each optimum)
    std::vector<double> ubs((size_t)nq, 0.0);
    for (int32_t q = 0; q < nq; ++q) {
        gst_opts o{};
        o.prune_lb = 1;
        gst_result r;
        gst_status s = gst_solve(ctx, G, &sample[q], &o, &r);   // validates the query too
        if (s != GST_OK) return s;
        ubs[(size_t)q] = bounded ? r.cost * 1.05 : 0.0;
        gst_result_free(&r);
    }
    int tried = 0;
    // mean ms per query of the sample under knobs k (one untimed solve first: round budgets)
    auto eval = [&](const Knobs& k, double* out) -> gst_status {
        gst_opts o{};
        o.prune_lb = 1;
        o.want_tree = 1;
        o.pull_permille = k.pull_permille;
        o.chunk_arcs = k.chunk_arcs;
        o.push_lanes = k.push_lanes;
        o.merge_tile_kb = k.merge_tile_kb;
        gst_result r;
        o.ub = ubs[0];
        gst_status s = gst_dpbf_solve(ctx, G, &sample[0], &o, &r);
        if (s != GST_OK) return s;
        gst_result_free(&r);
        const double t0 = now_ms();
        for (int32_t q = 0; q < nq; ++q) {
            o.ub = ubs[(size_t)q];
            s = gst_dpbf_solve(ctx, G, &sample[q], &o, &r);
            if (s != GST_OK) return s;
            gst_result_free(&r);
        }
        *out = (now_ms() - t0) / nq;
        tried++;
        return GST_OK;
    };
    const std::vector<int32_t> c_pull = {450, -1, 250, 650, 850};
    const std::vector<int32_t> c_chunk = {32, 16, 64};
    const std::vector<int32_t> c_lanes = {64, 16, 32};
    const std::vector<int32_t> c_tile = {48, 32, 16, 8};
    Knobs best{450, 32, 64, 48};
    double best_ms = 0.0;
    gst_status s = eval(best, &best_ms);
    if (s != GST_OK) return s;
    const double default_ms = best_ms;
    for (int sweep = 0; sweep < 2; ++sweep) {
        bool changed = false;
        for (int knob = 0; knob < 4; ++knob) {
            const std::vector<int32_t>& cands = knob == 0 ? c_pull : knob == 1 ? c_chunk : knob == 2 ? c_lanes : c_tile;
            for (int32_t v : cands) {
                Knobs k = best;
                int32_t& slot = knob == 0 ? k.pull_permille : knob == 1 ? k.chunk_arcs : knob == 2 ? k.push_lanes : k.merge_tile_kb;
                if (slot == v) continue;
                slot = v;
                double ms = 0.0;
                s = eval(k, &ms);
                if (s != GST_OK) return s;
                if (ms < best_ms * 0.98) { best = k; best_ms = ms; changed = true; }
            }
        }
        if (!changed) break;
    }
    gst_tuned t;
    t.uid = G->uid; t.g = g; t.ub = bounded;
    t.chunk_arcs = best.chunk_arcs; t.pull_permille = best.pull_permille;
    t.push_lanes = best.push_lanes; t.merge_tile_kb = best.merge_tile_kb;
    ctx->tuned.push_back(t);
    if (rep) {
        rep->chunk_arcs = best.chunk_arcs; rep->pull_permille = best.pull_permille;
        rep->push_lanes = best.push_lanes; rep->merge_tile_kb = best.merge_tile_kb;
        rep->default_ms = default_ms; rep->tuned_ms = best_ms;
    }
    if (batch) {
        ctx->tuned_batch.erase(std::remove_if(ctx->tuned_batch.begin(), ctx->tuned_batch.end(),
                                              [&](const gst_tuned_batch& b) { return b.uid == G->uid; }),
                               ctx->tuned_batch.end());
        // the engine is timed as a stream: 64 slots (fewer when they do not fit) fed the sample
        // over and over, 128 jobs at least, so that slots refill while others run as in a long batch
        const int32_t njobs = std::max<int32_t>(nq, 128);
        std::vector<gst_opts> bo((size_t)nq);
        std::vector<gst_result> res((size_t)njobs);
        std::vector<gst_status> st((size_t)njobs);
        for (int32_t q = 0; q < nq; ++q) {
            bo[(size_t)q] = gst_opts();
            bo[(size_t)q].prune_lb = 1;
            bo[(size_t)q].want_tree = 1;
            bo[(size_t)q].ub = ubs[(size_t)q];
        }
        int32_t tune_slots = 64;
        auto beval = [&](int32_t R, int32_t lanes, double* qps) -> gst_status {
            for (auto& o : bo) o.push_lanes = lanes;
            gst_batch_opts bopt{};
            bopt.slots = tune_slots;
            bopt.rounds_per_step = R;
            bopt.use_graph = eager ? 0 : 1;
            double best_q = 0.0;
            for (int rep2 = 0; rep2 < (eager ? 1 : 2); ++rep2) {   // the first run builds the captured step
                int32_t next = 0;
                auto next_job = [&](gst_batch_job& j, bool) -> bool {
                    if (next >= njobs) return false;
                    const int32_t q = next % nq;
                    j = gst_batch_job();
                    j.G = G; j.groups = &sample[q]; j.opts = bo[(size_t)q]; j.res = &res[(size_t)next]; j.tag = next;
                    ++next;
                    return true;
                };
                auto job_done = [&](gst_batch_job& j) { st[(size_t)j.tag] = j.status; };
                const double t0 = now_ms();
                gst_status rs = gst_batch_run(ctx, next_job, job_done, (uint32_t)G->n, g, G->m, &bopt, nullptr, 0);
                const double ms = now_ms() - t0;
                for (auto& r : res) gst_result_free(&r);
                if (rs == GST_OOM && bopt.slots != 0) {   // 64 slots do not fit: the engine's own count
                    tune_slots = bopt.slots = 0;
                    --rep2;
                    continue;
                }
                if (rs != GST_OK) return rs;
                best_q = std::max(best_q, njobs * 1000.0 / ms);
            }
            *qps = best_q;
            tried++;
            return GST_OK;
        };
        // lanes first at the default rounds per step (the larger effect), then rounds per step at
        // the chosen lanes; a change must win by 3% (lanes) or 5% (rounds), so that timing noise
        // between near-equal settings leaves the defaults in place
        int32_t bR = 8, bL = 16;
        double bq = 0.0;
        s = beval(bR, bL, &bq);
        if (s != GST_OK) return s;
        const double bdef = bq;
        {
            int32_t L0 = bL;
            double q0 = bq;
            for (int32_t lanes : {8, 32}) {
                double q = 0.0;
                s = beval(bR, lanes, &q);
                if (s != GST_OK) return s;
                if (q > q0 * 1.03) { q0 = q; L0 = lanes; }
            }
            bL = L0; bq = q0;
        }
        {
            int32_t R0 = bR;
            double q0 = bq;
            for (int32_t R : {4, 16}) {
                double q = 0.0;
                s = beval(R, bL, &q);
                if (s != GST_OK) return s;
                if (q > q0 * 1.05) { q0 = q; R0 = R; }
            }
            bR = R0; bq = q0;
        }
        gst_tuned_batch tb;
        tb.uid = G->uid; tb.rounds_per_step = bR; tb.push_lanes = bL;
        ctx->tuned_batch.push_back(tb);
        if (rep) {
            rep->batch_rounds_per_step = bR; rep->batch_push_lanes = bL;
            rep->batch_default_qps = bdef; rep->batch_tuned_qps = bq;
        }
    }
    if (rep) rep->configs_tried = tried;
    return GST_OK;
}
