// Host side of the C-ABI: contexts, errors, graph ingestion, argument checks, tree validation
// and the DP level tables.  Device work lives in gst_dpbf.hip.
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <atomic>
#include <new>
#include <numeric>
#include <vector>

#include "gst_batch.h"
#include "gst_internal.h"
#include "gst_mpnn.h"

gst_status gst_set_error(gst_ctx* ctx, gst_status s, const char* fmt, ...) {
    if (ctx) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        std::vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        std::lock_guard<std::mutex> lk(ctx->err_mu);
        ctx->last_error = buf;
    }
    return s;
}

// Masks ordered by (popcount, value); merge pairs per level with the canonical orientation
// (X1 holds the lowest bit of X), which visits each unordered split once.  The reference keeps
// an O(4^g) list of all disjoint mask pairs and walks each pair from both sides
// (rucgraph/test_gpu.h:19-44, src/DPQ.cu:285-314).
void gst_levels::build(int g_) {
    g = g_;
    full = (1 << g) - 1;
    for (int k = 0; k <= GST_MAX_GROUPS + 1; ++k) { m[k] = 0; off[k] = 0; pair_off[k] = 0; pair_cnt[k] = 0; }
    masks.clear();
    pos_of.assign((size_t)full + 1, 0xFFFFFFFFu);
    rank_of.assign((size_t)full + 1, 0xFFFFFFFFu);
    for (int k = 1; k <= g; ++k) {
        off[k] = (int)masks.size();
        for (uint32_t X = 1; X <= (uint32_t)full; ++X) {
            if (__builtin_popcount(X) != k) continue;
            rank_of[X] = (uint32_t)(masks.size() - (size_t)off[k]);
            pos_of[X] = (uint32_t)masks.size();
            masks.push_back(X);
        }
        m[k] = (int)masks.size() - off[k];
    }
    off[g + 1] = (int)masks.size();
    pairs.clear();
    for (int k = 2; k <= g; ++k) {
        pair_off[k] = (int)pairs.size();
        for (int r = 0; r < m[k]; ++r) {
            const uint32_t X = masks[(size_t)(off[k] + r)];
            const uint32_t low = X & (0u - X);
            const uint32_t rest = X ^ low;
            // submasks s of rest, s != rest: X1 = low|s, X2 = X^X1 non-empty
            uint32_t sub = rest;
            while (true) {
                sub = (sub - 1) & rest;  // walks every proper submask of rest, ending at 0
                const uint32_t X1 = low | sub;
                const uint32_t X2 = X ^ X1;
                const uint64_t e = (uint64_t)(uint32_t)r | ((uint64_t)pos_of[X1] << 16) |
                                   ((uint64_t)pos_of[X2] << 32) | ((uint64_t)X1 << 48);
                pairs.push_back(e);
                if (sub == 0) break;
            }
        }
        pair_cnt[k] = (int)pairs.size() - pair_off[k];
    }
}

uint64_t gst_next_graph_uid() {
    static std::atomic<uint64_t> next{1};
    return next.fetch_add(1);
}

const gst_levels& gst_get_levels(gst_ctx* ctx, int g) {
    if (!ctx->lv_ready[g]) {
        ctx->lv_cache[g].build(g);
        ctx->lv_ready[g] = true;
    }
    return ctx->lv_cache[g];
}

namespace {

bool weight_ok(uint32_t bits, int32_t w_type) {
    if (w_type == GST_W_INT32) return (int32_t)bits >= 0;
    float f;
    std::memcpy(&f, &bits, 4);
    return std::isfinite(f) && f >= 0.0f;
}

// lighter-weight comparison for either weight type
bool w_less(uint32_t a, uint32_t b, int32_t w_type) {
    if (w_type == GST_W_INT32) return (int32_t)a < (int32_t)b;
    float fa, fb;
    std::memcpy(&fa, &a, 4);
    std::memcpy(&fb, &b, 4);
    return fa < fb;
}

double w_to_double(uint32_t bits, int32_t w_type) {
    if (w_type == GST_W_INT32) return (double)(int32_t)bits;
    float f;
    std::memcpy(&f, &bits, 4);
    return (double)f;
}

gst_status check_csr(gst_ctx* ctx, const gst_csr* c) {
    if (!c || c->n < 1 || c->m < 0 || !c->row_ptr || (c->m > 0 && (!c->col || !c->w)))
        return gst_set_error(ctx, GST_INVALID_ARG, "csr: null array or empty graph");
    if (c->w_type != GST_W_INT32 && c->w_type != GST_W_FLOAT32)
        return gst_set_error(ctx, GST_INVALID_ARG, "csr: unknown weight type %d", c->w_type);
    if (c->row_ptr[0] != 0 || c->row_ptr[c->n] != c->m)
        return gst_set_error(ctx, GST_INVALID_ARG, "csr: row_ptr[0] must be 0 and row_ptr[n] == m");
    if (c->m >= (int64_t)0x7FFFFFFF)
        return gst_set_error(ctx, GST_INVALID_ARG, "csr: %lld arcs exceed the 2^31-1 arc limit", (long long)c->m);
    for (int32_t v = 0; v < c->n; ++v)
        if (c->row_ptr[v + 1] < c->row_ptr[v])
            return gst_set_error(ctx, GST_INVALID_ARG, "csr: row_ptr decreases at vertex %d", v);
    const uint32_t* w = (const uint32_t*)c->w;
    for (int64_t e = 0; e < c->m; ++e) {
        if (c->col[e] < 0 || c->col[e] >= c->n)
            return gst_set_error(ctx, GST_INVALID_ARG, "csr: arc %lld points to vertex %d outside [0,%d)",
                                 (long long)e, c->col[e], c->n);
        if (!weight_ok(w[e], c->w_type))
            return gst_set_error(ctx, GST_INVALID_ARG, "csr: arc %lld has a negative or non-finite weight", (long long)e);
    }
    return GST_OK;
}

// group-list checks shared by every solve entry point
gst_status check_query(gst_ctx* ctx, const gst_graph* g, const gst_groups* groups) {
    const int ng = groups->n_groups;
    if (ng < 1) return gst_set_error(ctx, GST_INVALID_ARG, "need at least one group");
    if (ng > GST_MAX_GROUPS)
        return gst_set_error(ctx, GST_TOO_MANY_GROUPS, "%d groups exceed the maximum of %d", ng, GST_MAX_GROUPS);
    if (!groups->grp_ptr || (groups->grp_ptr[ng] > 0 && !groups->grp_vtx) || groups->grp_ptr[0] != 0)
        return gst_set_error(ctx, GST_INVALID_ARG, "groups: malformed grp_ptr");
    for (int i = 0; i < ng; ++i) {
        if (groups->grp_ptr[i + 1] < groups->grp_ptr[i])
            return gst_set_error(ctx, GST_INVALID_ARG, "groups: grp_ptr decreases at group %d", i);
        if (groups->grp_ptr[i + 1] == groups->grp_ptr[i])
            return gst_set_error(ctx, GST_INFEASIBLE, "group %d is empty", i);
    }
    for (int32_t j = 0; j < groups->grp_ptr[ng]; ++j)
        if (groups->grp_vtx[j] < 0 || groups->grp_vtx[j] >= g->n)
            return gst_set_error(ctx, GST_INVALID_ARG, "groups: member %d outside [0,%d)", groups->grp_vtx[j], g->n);
    return GST_OK;
}

}  // namespace

extern "C" {

const char* gst_status_str(gst_status s) {
    switch (s) {
        case GST_OK: return "GST_OK";
        case GST_INFEASIBLE: return "GST_INFEASIBLE";
        case GST_TOO_MANY_GROUPS: return "GST_TOO_MANY_GROUPS";
        case GST_OOM: return "GST_OOM";
        case GST_INVALID_ARG: return "GST_INVALID_ARG";
        case GST_HIP_ERROR: return "GST_HIP_ERROR";
        case GST_COST_OVERFLOW: return "GST_COST_OVERFLOW";
        case GST_NO_DEVICE: return "GST_NO_DEVICE";
        case GST_INTERNAL: return "GST_INTERNAL";
    }
    return "GST_UNKNOWN";
}

gst_status gst_ctx_create(int32_t device, gst_ctx** out) {
    if (!out) return GST_INVALID_ARG;
    *out = nullptr;
    gst_ctx* ctx = new (std::nothrow) gst_ctx();
    if (!ctx) return GST_OOM;
    ctx->device = device;
    const gst_status s = gst_ctx_device_init(ctx);
    if (s != GST_OK) {
        // keep the context so the caller can read the message, but report the failure
        *out = ctx;
        return s;
    }
    *out = ctx;
    return GST_OK;
}

void gst_ctx_destroy(gst_ctx* ctx) {
    if (!ctx) return;
    gst_ctx_device_fini(ctx);
    delete ctx;
}

const char* gst_last_error(const gst_ctx* ctx) { return ctx ? ctx->last_error.c_str() : "null context"; }

gst_status gst_graph_upload(gst_ctx* ctx, const gst_csr* c, gst_graph** out) {
    if (!ctx || !out) return GST_INVALID_ARG;
    *out = nullptr;
    if (!ctx->stream) return gst_set_error(ctx, GST_NO_DEVICE, "context has no device");
    gst_status s = check_csr(ctx, c);
    if (s != GST_OK) return s;
    gst_graph* g = new (std::nothrow) gst_graph();
    if (!g) return gst_set_error(ctx, GST_OOM, "graph: host allocation failed");
    g->ctx = ctx;
    g->uid = gst_next_graph_uid();
    g->n = c->n;
    g->w_type = c->w_type;
    const uint32_t* w = (const uint32_t*)c->w;
    // deduplicated rows without self-loops (the lightest parallel arc is kept), each row in
    // increasing (weight, neighbour) order: a relaxation that walks a row can stop at the first
    // arc whose weight alone already rules it out.  -0.0f weights are stored as +0.0f so that
    // weight bit patterns order like the weights.
    g->h_row.assign((size_t)c->n + 1, 0);
    g->h_col.reserve((size_t)c->m);
    g->h_w.reserve((size_t)c->m);
    std::vector<std::pair<int32_t, uint32_t>> tmp;
    std::vector<uint64_t> keyed;
    for (int32_t v = 0; v < c->n; ++v) {
        tmp.clear();
        for (int64_t e = c->row_ptr[v]; e < c->row_ptr[v + 1]; ++e)
            if (c->col[e] != v) tmp.push_back({c->col[e], w[e] == 0x80000000u && c->w_type == GST_W_FLOAT32 ? 0u : w[e]});
        std::sort(tmp.begin(), tmp.end(), [&](const std::pair<int32_t, uint32_t>& a, const std::pair<int32_t, uint32_t>& b) {
            if (a.first != b.first) return a.first < b.first;
            return w_less(a.second, b.second, c->w_type);
        });
        keyed.clear();
        int32_t last = -1;
        for (auto& p : tmp) {
            if (p.first == last) continue;  // parallel arc: the lightest one was kept
            keyed.push_back(arc_key(p.second, p.first));
            last = p.first;
        }
        std::sort(keyed.begin(), keyed.end());
        for (uint64_t k : keyed) {
            g->h_col.push_back((int32_t)(uint32_t)k);
            g->h_w.push_back((uint32_t)(k >> 32));
        }
        g->h_row[(size_t)v + 1] = (int64_t)g->h_col.size();
        g->max_deg = std::max<int32_t>(g->max_deg, (int32_t)(g->h_row[(size_t)v + 1] - g->h_row[(size_t)v]));
    }
    g->m = (int64_t)g->h_col.size();
    // symmetry: every arc needs its reverse with the same weight
    for (int32_t v = 0; v < g->n; ++v) {
        for (int64_t e = g->h_row[(size_t)v]; e < g->h_row[(size_t)v + 1]; ++e) {
            const int32_t u = g->h_col[(size_t)e];
            if (gst_find_arc(g->h_row.data(), g->h_col.data(), g->h_w.data(), u, v, g->h_w[(size_t)e]) < 0) {
                delete g;
                return gst_set_error(ctx, GST_INVALID_ARG,
                                     "csr: arc %d->%d has no reverse arc of the same weight (graph must be undirected)", v, u);
            }
        }
    }
    s = gst_graph_device_upload(ctx, g);
    if (s != GST_OK) { delete g; return s; }
    *out = g;
    return GST_OK;
}

void gst_graph_free(gst_graph* g) {
    if (!g) return;
    gst_graph_device_free(g);
    delete g;
}

int32_t gst_graph_num_vertices(const gst_graph* g) { return g ? g->n : 0; }
int64_t gst_graph_num_arcs(const gst_graph* g) { return g ? g->m : 0; }

gst_status gst_solve(gst_ctx* ctx, const gst_graph* g, const gst_groups* groups, const gst_opts* opts,
                     gst_result* result) {
    if (!ctx || !g || !groups || !result) return ctx ? gst_set_error(ctx, GST_INVALID_ARG, "null argument") : GST_INVALID_ARG;
    std::memset(result, 0, sizeof(*result));
    result->root = -1;
    if (!ctx->stream) return gst_set_error(ctx, GST_NO_DEVICE, "context has no device");
    if (g->ctx != ctx) return gst_set_error(ctx, GST_INVALID_ARG, "graph belongs to another context");
    const gst_status cs = check_query(ctx, g, groups);
    if (cs != GST_OK) return cs;
    return gst_dpbf_solve(ctx, g, groups, opts, result);
}

gst_status gst_mst_postprocess(gst_ctx* ctx, const gst_graph* g, const int32_t* vertices, int32_t nv, int32_t* eu,
                               int32_t* ev, double* ew, int32_t* ne, double* cost) {
    if (!ctx || !g || nv < 1 || !vertices || !eu || !ev || !ew || !ne || !cost)
        return ctx ? gst_set_error(ctx, GST_INVALID_ARG, "mst_postprocess: null argument or empty set") : GST_INVALID_ARG;
    *ne = 0;
    *cost = 0.0;
    if (!ctx->stream) return gst_set_error(ctx, GST_NO_DEVICE, "context has no device");
    if (g->ctx != ctx) return gst_set_error(ctx, GST_INVALID_ARG, "graph belongs to another context");
    if (nv > 4 * 512) return gst_set_error(ctx, GST_INVALID_ARG, "mst_postprocess: too many vertices (%d)", nv);
    for (int32_t i = 0; i < nv; ++i)
        if (vertices[i] < 0 || vertices[i] >= g->n)
            return gst_set_error(ctx, GST_INVALID_ARG, "mst_postprocess: vertex %d out of range", vertices[i]);
    return gst_dpbf_mst_postprocess(ctx, g, vertices, nv, eu, ev, ew, ne, cost);
}

gst_status gst_solve_noalloc(gst_ctx* ctx, const gst_graph* g, const gst_groups* groups, const gst_opts* opts,
                             int32_t edge_cap, int32_t* eu, int32_t* ev, double* ew, gst_result* result) {
    if (!ctx || !g || !groups || !result) return ctx ? gst_set_error(ctx, GST_INVALID_ARG, "null argument") : GST_INVALID_ARG;
    std::memset(result, 0, sizeof(*result));
    result->root = -1;
    if (!ctx->stream) return gst_set_error(ctx, GST_NO_DEVICE, "context has no device");
    if (g->ctx != ctx) return gst_set_error(ctx, GST_INVALID_ARG, "graph belongs to another context");
    const gst_status cs = check_query(ctx, g, groups);
    if (cs != GST_OK) return cs;
    return gst_dpbf_solve_noalloc(ctx, g, groups, opts, edge_cap, eu, ev, ew, result);
}

gst_status gst_solve_batch(gst_ctx* ctx, const gst_graph* g, int32_t nq, const gst_groups* groups,
                           const gst_opts* opts, const gst_batch_opts* bopts, gst_result* results,
                           gst_status* status, gst_batch_stats* stats) {
    if (!ctx || !g || nq < 0 || (nq > 0 && (!groups || !results || !status)))
        return ctx ? gst_set_error(ctx, GST_INVALID_ARG, "null argument") : GST_INVALID_ARG;
    if (stats) std::memset(stats, 0, sizeof(*stats));
    if (!ctx->stream) return gst_set_error(ctx, GST_NO_DEVICE, "context has no device");
    if (g->ctx != ctx) return gst_set_error(ctx, GST_INVALID_ARG, "graph belongs to another context");
    if (bopts && (bopts->slots < 0 || bopts->rounds_per_step < 0))
        return gst_set_error(ctx, GST_INVALID_ARG, "batch options: negative slots or rounds_per_step");
    gst_status first = GST_OK;
    int gmax = 1;
    for (int32_t q = 0; q < nq; ++q) {
        std::memset(&results[q], 0, sizeof(results[q]));
        results[q].root = -1;
        status[q] = check_query(ctx, g, &groups[q]);
        if (status[q] == GST_OK && opts && !gst_push_lanes_ok(opts[q].push_lanes))
            status[q] = gst_set_error(ctx, GST_INVALID_ARG, "push_lanes must be 0, 1, 8, 16, 32 or 64");
        if (status[q] == GST_OK && opts && opts[q].half_bound != 0 && opts[q].half_bound != -1)
            status[q] = gst_set_error(ctx, GST_INVALID_ARG, "half_bound must be 0 or -1");
        if (status[q] == GST_OK && opts && opts[q].tour_bound != 0 && opts[q].tour_bound != -1)
            status[q] = gst_set_error(ctx, GST_INVALID_ARG, "tour_bound must be 0 or -1");
        if (status[q] == GST_OK) gmax = std::max(gmax, groups[q].n_groups);
        else if (first == GST_OK) first = status[q];
    }
    int32_t next = 0;
    auto next_job = [&](gst_batch_job& j, bool) -> bool {   // every query is ready at once
        while (next < nq && status[next] != GST_OK) ++next;
        if (next >= nq) return false;
        j = gst_batch_job();
        j.G = g;
        j.groups = &groups[next];
        if (opts) j.opts = opts[next];
        else { j.opts = gst_opts(); j.opts.want_tree = 1; }
        j.res = &results[next];
        j.tag = next;
        ++next;
        return true;
    };
    std::vector<uint8_t> done((size_t)std::max(nq, 0), 0);
    auto job_done = [&](gst_batch_job& j) {
        status[j.tag] = j.status;
        done[(size_t)j.tag] = 1;
    };
    int64_t valid = 0;
    for (int32_t q = 0; q < nq; ++q) valid += status[q] == GST_OK ? 1 : 0;
    bool engine_error = false;
    const gst_status rs = gst_batch_run(ctx, next_job, job_done, (uint32_t)g->n, gmax, g->m, bopts, stats, g->uid,
                                        std::max<int64_t>(valid, 1), &engine_error);
    if (engine_error) {
        // the engine stopped early: every valid query without a result carries its error
        for (int32_t q = 0; q < nq; ++q)
            if (status[q] == GST_OK && !done[(size_t)q]) {
                gst_result_free(&results[q]);
                results[q].root = -1;
                status[q] = rs;
            }
    }
    if (first == GST_OK) first = rs;
    return first;
}


gst_status gst_solve_learned(gst_ctx* ctx, const gst_graph* g, const gst_groups* groups, gst_model* m,
                             double topk_frac, const gst_opts* opts, gst_result* result, gst_learned_stats* stats) {
    if (!m) return ctx ? gst_set_error(ctx, GST_INVALID_ARG, "null model") : GST_INVALID_ARG;
    if (!(topk_frac > 0.0 && topk_frac <= 1.0))
        return ctx ? gst_set_error(ctx, GST_INVALID_ARG, "topk_frac must be in (0, 1]") : GST_INVALID_ARG;
    // same argument checks as gst_solve (runs them and the plain solve path's validation)
    if (!ctx || !g || !groups || !result) return ctx ? gst_set_error(ctx, GST_INVALID_ARG, "null argument") : GST_INVALID_ARG;
    std::memset(result, 0, sizeof(*result));
    result->root = -1;
    if (!ctx->stream) return gst_set_error(ctx, GST_NO_DEVICE, "context has no device");
    if (g->ctx != ctx) return gst_set_error(ctx, GST_INVALID_ARG, "graph belongs to another context");
    const gst_status cs = check_query(ctx, g, groups);
    if (cs != GST_OK) return cs;
    return gst_learned_solve(ctx, g, groups, m, topk_frac, opts, result, stats);
}

gst_status gst_solve_batch_learned(gst_ctx* ctx, const gst_graph* g, int32_t nq, const gst_groups* groups,
                                   gst_model* m, double topk_frac, const gst_opts* opts, const gst_batch_opts* bopts,
                                   gst_result* results, gst_status* status, gst_learned_stats* lstats,
                                   gst_batch_stats* stats) {
    if (!ctx || !g || !m || nq < 0 || (nq > 0 && (!groups || !results || !status)))
        return ctx ? gst_set_error(ctx, GST_INVALID_ARG, "null argument") : GST_INVALID_ARG;
    if (!(topk_frac > 0.0 && topk_frac <= 1.0)) return gst_set_error(ctx, GST_INVALID_ARG, "topk_frac must be in (0, 1]");
    if (stats) std::memset(stats, 0, sizeof(*stats));
    if (!ctx->stream) return gst_set_error(ctx, GST_NO_DEVICE, "context has no device");
    if (g->ctx != ctx) return gst_set_error(ctx, GST_INVALID_ARG, "graph belongs to another context");
    if (bopts && (bopts->slots < 0 || bopts->rounds_per_step < 0))
        return gst_set_error(ctx, GST_INVALID_ARG, "batch options: negative slots or rounds_per_step");
    for (int32_t q = 0; q < nq; ++q) {
        std::memset(&results[q], 0, sizeof(results[q]));
        results[q].root = -1;
        if (lstats) { std::memset(&lstats[q], 0, sizeof(lstats[q])); lstats[q].ub = -1.0; }
        status[q] = check_query(ctx, g, &groups[q]);
        if (status[q] == GST_OK && opts && !gst_push_lanes_ok(opts[q].push_lanes))
            status[q] = gst_set_error(ctx, GST_INVALID_ARG, "push_lanes must be 0, 1, 8, 16, 32 or 64");
        if (status[q] == GST_OK && opts && opts[q].half_bound != 0 && opts[q].half_bound != -1)
            status[q] = gst_set_error(ctx, GST_INVALID_ARG, "half_bound must be 0 or -1");
        if (status[q] == GST_OK && opts && opts[q].tour_bound != 0 && opts[q].tour_bound != -1)
            status[q] = gst_set_error(ctx, GST_INVALID_ARG, "tour_bound must be 0 or -1");
    }
    return gst_learned_solve_batch(ctx, g, nq, groups, m, topk_frac, opts, bopts, results, status, lstats, stats);
}

void gst_result_free(gst_result* r) {
    if (!r) return;
    std::free(r->eu); std::free(r->ev); std::free(r->ew);
    r->eu = nullptr; r->ev = nullptr; r->ew = nullptr; r->n_edges = 0;
}

gst_status gst_validate_tree(const gst_csr* c, const gst_groups* groups, int32_t n_edges, const int32_t* eu,
                             const int32_t* ev, int32_t root, double* cost_out) {
    if (!c || !groups || n_edges < 0 || (n_edges > 0 && (!eu || !ev))) return GST_INVALID_ARG;
    const uint32_t* w = (const uint32_t*)c->w;
    std::vector<int32_t> verts;
    verts.reserve((size_t)n_edges * 2 + 1);
    if (n_edges == 0) {
        if (root < 0 || root >= c->n) return GST_INVALID_ARG;
        verts.push_back(root);
    }
    double cost = 0.0;
    std::vector<uint64_t> keys;
    keys.reserve((size_t)n_edges);
    for (int32_t i = 0; i < n_edges; ++i) {
        const int32_t x = eu[i], y = ev[i];
        if (x < 0 || y < 0 || x >= c->n || y >= c->n || x == y) return GST_INVALID_ARG;
        // the edge must exist; its weight is the lightest parallel arc
        bool found = false;
        uint32_t best = 0;
        for (int64_t e = c->row_ptr[x]; e < c->row_ptr[x + 1]; ++e) {
            if (c->col[e] == y && (!found || w_less(w[e], best, c->w_type))) { best = w[e]; found = true; }
        }
        if (!found) return GST_INVALID_ARG;
        cost += w_to_double(best, c->w_type);
        keys.push_back(((uint64_t)(uint32_t)std::min(x, y) << 32) | (uint32_t)std::max(x, y));
        verts.push_back(x);
        verts.push_back(y);
    }
    std::sort(keys.begin(), keys.end());
    if (std::adjacent_find(keys.begin(), keys.end()) != keys.end()) return GST_INVALID_ARG;  // repeated edge
    std::sort(verts.begin(), verts.end());
    verts.erase(std::unique(verts.begin(), verts.end()), verts.end());
    if ((int64_t)verts.size() != (int64_t)n_edges + 1) return GST_INVALID_ARG;  // not a tree
    // union-find over the tree vertices: n_edges unions without a cycle => connected tree
    std::vector<int32_t> parent(verts.size());
    std::iota(parent.begin(), parent.end(), 0);
    auto idx = [&](int32_t v) { return (int32_t)(std::lower_bound(verts.begin(), verts.end(), v) - verts.begin()); };
    auto find = [&](int32_t a) {
        while (parent[(size_t)a] != a) { parent[(size_t)a] = parent[(size_t)parent[(size_t)a]]; a = parent[(size_t)a]; }
        return a;
    };
    for (int32_t i = 0; i < n_edges; ++i) {
        const int32_t a = find(idx(eu[i])), b = find(idx(ev[i]));
        if (a == b) return GST_INVALID_ARG;  // cycle
        parent[(size_t)a] = b;
    }
    // every group touched
    for (int i = 0; i < groups->n_groups; ++i) {
        bool hit = false;
        for (int32_t j = groups->grp_ptr[i]; j < groups->grp_ptr[i + 1] && !hit; ++j)
            hit = std::binary_search(verts.begin(), verts.end(), groups->grp_vtx[j]);
        if (!hit) return GST_INFEASIBLE;
    }
    if (cost_out) *cost_out = cost;
    return GST_OK;
}

}  // extern "C"
