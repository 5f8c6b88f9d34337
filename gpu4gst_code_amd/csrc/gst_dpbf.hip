// DPBF group Steiner tree solve on one MI355X (gfx950): kernels and the per-query driver.
//
// Formulation (Ding et al. ICDE'07 DPBF, as restated from
// rucgraph/graph_v_of_v_idealID_DPBF_only_ec.h:133-324 and the dense shape of src/DP.cu:102-215):
//   T(v,X) = 0                                   if v belongs to every group in X   (leaf)
//   T(v,X) = min_{X1 ⊂ X} T(v,X1) + T(v,X\X1)                                        (merge)
//   T(v,X) = min_u T(u,X) + w(u,v)                                                    (grow)
// Masks are processed level by level (level k = popcount k).  Level k: one merge kernel fills
// T(.,X) for every X of the level from the final lower levels, then frontier grow rounds run the
// multi-source shortest-path closure of all C(g,k) masks of the level together until no cost
// drops.  The full mask needs no grow: min_v T(v,full) cannot drop by growing.  The tree is
// traced on the device from the packed predecessors.
//
// HBM layout of the state table: one u64 word per (v,X); level k is a block [V][m_k] (m_k =
// C(g,k) masks in increasing mask order), so the C(g,k) masks a grow relaxation touches at a
// neighbour are one contiguous run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>

#include "gst_internal.h"

#define HIPCHK(call)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return gst_set_error(ctx, GST_HIP_ERROR, "%s failed: %s (%s:%d)", #call,         \
                                 hipGetErrorString(e_), __FILE__, __LINE__);                 \
    } while (0)

#include "gst_batch.h"
#include "gst_dpbf_dev.h"


// =======================================================================================
// per-query device buffers (grown on demand, reused across queries)
// =======================================================================================
struct gst_query_bufs {
    uint64_t cap_T = 0, cap_V = 0, cap_items = 0, cap_W = 0, cap_grp = 0, cap_edges = 0;
    uint64_t* T = nullptr;
    uint32_t* gm = nullptr;
    uint64_t* chg[3] = {nullptr, nullptr, nullptr};
    uint32_t* inq[3] = {nullptr, nullptr, nullptr};
    uint64_t* lists[3] = {nullptr, nullptr, nullptr};
    RoundCtl* ctl = nullptr;
    unsigned long long* root = nullptr;
    unsigned long long* stats = nullptr;
    unsigned long long* spider = nullptr;  // [0] spider tree (cost << 32 | centre), [1] cheapest spider
    uint32_t* tour = nullptr;    // tour table, 2^g * (g + 1) words, then the walk scratch 2^g * g * g
    int tour_g = 0;              // group count tour is sized for
    uint32_t* edges = nullptr;   // 3 * cap_edges + 1 (count)
    int32_t* grp = nullptr;      // grp_ptr (g+1) then members
    // level tables for the current g
    int tab_g = -1;
    uint64_t* pairs = nullptr;
    uint32_t* masks = nullptr;
    uint32_t* pos_of = nullptr;
    bool dirty = true;           // bitmaps / flags need clearing
    unsigned long long* ring = nullptr;   // async task slots (4 words each)
    uint64_t cap_ring = 0;                // slots (power of two)
    AsyncCtl* actl = nullptr;
    bool ring_dirty = true;               // slots need their tags invalidated
    int level_rounds[GST_MAX_GROUPS + 2] = {0};   // grow rounds per level of the current solve
    bool budget_valid = false;                   // unused by chained lookups (see budgets)
    // round budgets learned per (graph, group count, bounded?): hints only, every chained solve
    // checks convergence and falls back to host-checked rounds
    struct Budget { uint64_t uid; int ng; bool ub; int rounds[GST_MAX_GROUPS + 2]; };
    std::vector<Budget> budgets;
    // every (vertex, chunk) item of the current graph, for pull rounds
    uint64_t* all_items = nullptr;
    uint64_t cap_all = 0;
    uint32_t n_all = 0;          // hub items (rows longer than HUB_DEG)
    uint64_t n_items_total = 0;  // work items of a full frontier
    uint32_t* all_cnt = nullptr;
    uint64_t all_uid = 0;
    int64_t all_m = -1;
    uint32_t all_chunk = 0;
    RoundCtl* h_ctl = nullptr;   // pinned
    std::vector<hipEvent_t> evs;
    uint32_t* rlog = nullptr;    // round log of every level (GST_TRACE_ROUNDS), NCTL * RLOG_ROUNDS * 4 words
    uint8_t* act = nullptr;      // activity bytes, act[k * n + v] (levels 2..g-1)
    uint64_t cap_act = 0;
    uint32_t* live = nullptr;    // live list of a bounded merge (n entries)
    uint64_t cap_live = 0;
};

static void free_bufs(gst_query_bufs* b) {
    if (!b) return;
    hipFree(b->T); hipFree(b->gm);
    for (int i = 0; i < 3; ++i) { hipFree(b->chg[i]); hipFree(b->inq[i]); hipFree(b->lists[i]); }
    hipFree(b->ctl); hipFree(b->root); hipFree(b->stats); hipFree(b->edges); hipFree(b->grp);
    hipFree(b->pairs); hipFree(b->masks); hipFree(b->pos_of);
    hipFree(b->all_items); hipFree(b->all_cnt);
    hipFree(b->ring); hipFree(b->actl); hipFree(b->rlog); hipFree(b->spider); hipFree(b->tour); hipFree(b->act); hipFree(b->live);
    if (b->h_ctl) hipHostFree(b->h_ctl);
    for (auto e : b->evs) hipEventDestroy(e);
    delete b;
}

void gst_query_bufs_free(gst_ctx* ctx) {
    free_bufs(ctx->qb);
    ctx->qb = nullptr;
}

gst_status gst_ctx_device_init(gst_ctx* ctx) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return gst_set_error(ctx, GST_NO_DEVICE, "no HIP device available");
    if (ctx->device < 0 || ctx->device >= ndev)
        return gst_set_error(ctx, GST_INVALID_ARG, "device %d out of range (%d devices)", ctx->device, ndev);
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t s;
    HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    ctx->stream = (void*)s;
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, ctx->device));
    ctx->cu_count = prop.multiProcessorCount;
    return gst_noalloc_init(ctx);
}

void gst_ctx_device_fini(gst_ctx* ctx) {
    if (!ctx) return;
    gst_batch_free(ctx);
    gst_noalloc_fini(ctx);
    hipSetDevice(ctx->device);
    gst_query_bufs_free(ctx);
    if (ctx->stream) hipStreamDestroy((hipStream_t)ctx->stream);
    if (ctx->stream2) hipStreamDestroy((hipStream_t)ctx->stream2);
    if (ctx->stream3) hipStreamDestroy((hipStream_t)ctx->stream3);
    ctx->stream = nullptr;
    ctx->stream2 = nullptr;
    ctx->stream3 = nullptr;
}

gst_status gst_graph_device_upload(gst_ctx* ctx, gst_graph* g) {
    HIPCHK(hipSetDevice(ctx->device));
    const size_t n = (size_t)g->n, m = (size_t)g->m;
    std::vector<uint32_t> row32(n + 1);
    for (size_t i = 0; i <= n; ++i) row32[i] = (uint32_t)g->h_row[i];
    std::vector<uint64_t> arc(std::max<size_t>(m, 1));
    for (size_t e = 0; e < m; ++e) arc[e] = ((uint64_t)g->h_w[e] << 32) | (uint32_t)g->h_col[e];
    const size_t mm = std::max<size_t>(m, 1);
    if (hipMalloc(&g->d_row, (n + 1) * 4) != hipSuccess || hipMalloc(&g->d_col, mm * 4) != hipSuccess ||
        hipMalloc(&g->d_w, mm * 4) != hipSuccess || hipMalloc(&g->d_arc, mm * 8) != hipSuccess) {
        gst_graph_device_free(g);
        return gst_set_error(ctx, GST_OOM, "graph upload: device allocation failed (%zu arcs)", m);
    }
    hipStream_t s = (hipStream_t)ctx->stream;
    HIPCHK(hipMemcpyAsync(g->d_row, row32.data(), (n + 1) * 4, hipMemcpyHostToDevice, s));
    if (m) {
        HIPCHK(hipMemcpyAsync(g->d_col, g->h_col.data(), m * 4, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(g->d_w, g->h_w.data(), m * 4, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(g->d_arc, arc.data(), m * 8, hipMemcpyHostToDevice, s));
    }
    HIPCHK(hipStreamSynchronize(s));
    return GST_OK;
}

void gst_graph_device_free(gst_graph* g) {
    if (!g) return;
    if (g->ctx) hipSetDevice(g->ctx->device);
    hipFree(g->d_row); hipFree(g->d_col); hipFree(g->d_w); hipFree(g->d_arc);
    g->d_row = nullptr; g->d_col = nullptr; g->d_w = nullptr; g->d_arc = nullptr;
}

namespace {

template <class T>
bool grow_alloc(T*& p, uint64_t& cap, uint64_t need, size_t elem) {
    if (p && cap >= need) return true;
    hipFree(p);
    p = nullptr;
    cap = 0;
    if (hipMalloc((void**)&p, std::max<uint64_t>(need, 1) * elem) != hipSuccess) { p = nullptr; return false; }
    cap = need;
    return true;
}

uint32_t default_chunk(int m) {
    // arcs per work item: small enough that hub rows split over many wavefronts, large enough
    // to amortise the item setup (swept on BA V=200k: 32 beats 64 and 128)
    (void)m;
    return 32;
}

}  // namespace

// Tree of a traced edge list: duplicates removed, then a spanning forest of the union (every
// edge on a cycle of the union has weight zero, so any spanning tree of it has the DP cost).
void gst_finish_tree(bool flt, uint32_t nout, const uint32_t* hu, const uint32_t* hv, const uint32_t* hw,
                     gst_result* res) {
    // deduplicate (a tie-free optimum is a tree; duplicates can only come from zero-cost ties)
    std::vector<std::pair<uint64_t, uint32_t>> ed;
    ed.reserve(nout);
    for (uint32_t i = 0; i < nout; ++i) {
        uint32_t x = hu[i], y = hv[i];
        if (x > y) std::swap(x, y);
        ed.push_back({((uint64_t)x << 32) | y, hw[i]});
    }
    std::sort(ed.begin(), ed.end());
    ed.erase(std::unique(ed.begin(), ed.end(),
                         [](const std::pair<uint64_t, uint32_t>& p, const std::pair<uint64_t, uint32_t>& q) {
                             return p.first == q.first;
                         }),
             ed.end());
    // Every edge on a cycle of the traced union has weight zero (dropping it would give a
    // feasible structure cheaper than the optimum), so any spanning tree of the union has the
    // DP cost: keep the edges that join two components.
    {
        std::vector<uint32_t> vs;
        vs.reserve(ed.size() * 2);
        for (auto& p2 : ed) { vs.push_back((uint32_t)(p2.first >> 32)); vs.push_back((uint32_t)p2.first); }
        std::sort(vs.begin(), vs.end());
        vs.erase(std::unique(vs.begin(), vs.end()), vs.end());
        std::vector<uint32_t> par(vs.size());
        for (size_t i = 0; i < par.size(); ++i) par[i] = (uint32_t)i;
        auto fnd = [&](uint32_t x) {
            while (par[x] != x) { par[x] = par[par[x]]; x = par[x]; }
            return x;
        };
        auto idx = [&](uint32_t v) { return (uint32_t)(std::lower_bound(vs.begin(), vs.end(), v) - vs.begin()); };
        std::vector<std::pair<uint64_t, uint32_t>> kept;
        kept.reserve(ed.size());
        for (auto& p2 : ed) {
            const uint32_t a2 = fnd(idx((uint32_t)(p2.first >> 32))), b2 = fnd(idx((uint32_t)p2.first));
            if (a2 == b2) continue;
            par[a2] = b2;
            kept.push_back(p2);
        }
        ed.swap(kept);
    }
    res->n_edges = (int32_t)ed.size();
    res->eu = (int32_t*)std::malloc(sizeof(int32_t) * std::max<size_t>(ed.size(), 1));
    res->ev = (int32_t*)std::malloc(sizeof(int32_t) * std::max<size_t>(ed.size(), 1));
    res->ew = (double*)std::malloc(sizeof(double) * std::max<size_t>(ed.size(), 1));
    double tc = 0;
    for (size_t i = 0; i < ed.size(); ++i) {
        res->eu[i] = (int32_t)(ed[i].first >> 32);
        res->ev[i] = (int32_t)(ed[i].first & 0xFFFFFFFFu);
        double wv;
        if (flt) { float f; std::memcpy(&f, &ed[i].second, 4); wv = (double)f; }
        else wv = (double)ed[i].second;
        res->ew[i] = wv;
        tc += wv;
    }
    res->tree_cost = tc;
}

gst_status gst_dpbf_solve(gst_ctx* ctx, const gst_graph* G, const gst_groups* groups,
                          const gst_opts* opts_in, gst_result* res) {
    const auto t_start = std::chrono::steady_clock::now();
    gst_opts opts{};
    opts.ub = 0; opts.prune_lb = 1; opts.want_tree = 1;
    if (opts_in) opts = *opts_in;
    if (!gst_push_lanes_ok(opts.push_lanes))
        return gst_set_error(ctx, GST_INVALID_ARG, "push_lanes must be 0, 1, 8, 16, 32 or 64 (got %d)", opts.push_lanes);
    if (opts.half_bound != 0 && opts.half_bound != -1)
        return gst_set_error(ctx, GST_INVALID_ARG, "half_bound must be 0 or -1 (got %d)", opts.half_bound);
    if (opts.tour_bound != 0 && opts.tour_bound != -1)
        return gst_set_error(ctx, GST_INVALID_ARG, "tour_bound must be 0 or -1 (got %d)", opts.tour_bound);
    if (opts.merge_tile_kb < 0 || opts.merge_tile_kb > 48)
        return gst_set_error(ctx, GST_INVALID_ARG, "merge_tile_kb must be 0..48 (got %d)", opts.merge_tile_kb);
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)ctx->stream;
    const int g = groups->n_groups;
    const uint32_t n = (uint32_t)G->n;
    const gst_levels& L = gst_get_levels(ctx, g);
    const bool flt = G->w_type == GST_W_FLOAT32;

    // bound in cost bits
    const bool has_ub = std::isfinite(opts.ub) && opts.ub > 0;
    gst_apply_tuned(ctx, G, g, has_ub, opts);
    gst_bound_bits bb = gst_make_bound(opts.ub, flt, opts.half_bound);
    uint32_t ub_bits = bb.ub;

    // ---- buffers ----
    if (!ctx->qb) ctx->qb = new gst_query_bufs();
    gst_query_bufs* B = ctx->qb;
    const uint64_t words = (uint64_t)n * (uint64_t)L.full;
    uint32_t Wmax = 1;
    for (int k = 1; k <= g; ++k) Wmax = std::max<uint32_t>(Wmax, (uint32_t)((L.m[k] + 63) / 64));
    const uint32_t chunk = opts.chunk_arcs > 0 ? (uint32_t)opts.chunk_arcs : default_chunk(0);
    const uint64_t items_cap = (uint64_t)n + (uint64_t)G->m / chunk + 64;
    const uint64_t edges_cap = (uint64_t)n + 2 * (uint64_t)g + 16;
    const uint64_t grp_need = (uint64_t)(g + 1) + (uint64_t)groups->grp_ptr[g];
    bool ok = true;
    const uint64_t oldV = B->cap_V, oldW = B->cap_W;
    ok = ok && grow_alloc(B->T, B->cap_T, words, 8);
    uint64_t capv = B->cap_V, capw = B->cap_W, capi = B->cap_items;
    if (!(B->gm && capv >= n && capw >= Wmax)) {
        uint64_t dummy;
        for (int i = 0; i < 3; ++i) {
            dummy = 0; hipFree(B->chg[i]); B->chg[i] = nullptr;
            ok = ok && grow_alloc(B->chg[i], dummy, (uint64_t)n * Wmax, 8);
            dummy = 0; hipFree(B->inq[i]); B->inq[i] = nullptr;
            ok = ok && grow_alloc(B->inq[i], dummy, (uint64_t)n, 4);
        }
        dummy = 0; hipFree(B->gm); B->gm = nullptr;
        ok = ok && grow_alloc(B->gm, dummy, (uint64_t)n, 4);
        B->cap_V = n; B->cap_W = Wmax;
        B->dirty = true;
    }
    (void)oldV; (void)oldW;
    if (!(B->lists[0] && capi >= items_cap)) {
        for (int i = 0; i < 3; ++i) {
            uint64_t dummy = 0; hipFree(B->lists[i]); B->lists[i] = nullptr;
            ok = ok && grow_alloc(B->lists[i], dummy, items_cap, 8);
        }
        B->cap_items = items_cap;
    }
    {
        uint64_t need = 1;
        while (need < 2 * items_cap) need <<= 1;
        if (!B->ring || B->cap_ring < need) {
            hipFree(B->ring);
            B->ring = nullptr;
            B->cap_ring = 0;
            if (hipMalloc((void**)&B->ring, need * 32) != hipSuccess) ok = false;
            else { B->cap_ring = need; B->ring_dirty = true; }
        }
        if (!B->actl) {
            uint64_t d = 0;
            ok = ok && grow_alloc(B->actl, d, 1, sizeof(AsyncCtl));
        }
    }
    if (!B->ctl) {
        uint64_t d = 0;
        ok = ok && grow_alloc(B->ctl, d, NCTL, sizeof(RoundCtl));
        d = 0; ok = ok && grow_alloc(B->root, d, 1, 8);
        d = 0; ok = ok && grow_alloc(B->stats, d, STAT_WORDS, 8);
        if (hipHostMalloc((void**)&B->h_ctl, sizeof(RoundCtl) * NCTL) != hipSuccess) ok = false;
    }
    ok = ok && grow_alloc(B->edges, B->cap_edges, 3 * edges_cap + 1, 4);
    if (ok && B->cap_edges != 3 * edges_cap + 1) {
        // capacity bookkeeping: cap_edges counts words; recover the per-array capacity
    }
    ok = ok && grow_alloc(B->grp, B->cap_grp, grp_need, 4);
    ok = ok && grow_alloc(B->act, B->cap_act, (uint64_t)n * (uint64_t)(g + 1), 1);
    ok = ok && grow_alloc(B->live, B->cap_live, (uint64_t)n, 4);
    if (!ok) {
        gst_query_bufs_free(ctx);
        return gst_set_error(ctx, GST_OOM, "solve: device allocation failed (V=%u, g=%d, table %.2f GB)",
                             n, g, words * 8.0 / 1e9);
    }
    const uint32_t ecap = (uint32_t)((B->cap_edges - 1) / 3);
    if (B->tab_g != g) {
        hipFree(B->pairs); hipFree(B->masks); hipFree(B->pos_of);
        B->pairs = nullptr; B->masks = nullptr; B->pos_of = nullptr;
        uint64_t d = 0;
        ok = grow_alloc(B->pairs, d, std::max<size_t>(L.pairs.size(), 1), 8);
        d = 0; ok = ok && grow_alloc(B->masks, d, L.masks.size(), 4);
        d = 0; ok = ok && grow_alloc(B->pos_of, d, L.pos_of.size(), 4);
        if (!ok) { gst_query_bufs_free(ctx); return gst_set_error(ctx, GST_OOM, "solve: table allocation failed"); }
        if (!L.pairs.empty()) HIPCHK(hipMemcpyAsync(B->pairs, L.pairs.data(), L.pairs.size() * 8, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(B->masks, L.masks.data(), L.masks.size() * 4, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(B->pos_of, L.pos_of.data(), L.pos_of.size() * 4, hipMemcpyHostToDevice, s));
        B->tab_g = g;
    }
    const int fill_grid = std::max(1, ctx->cu_count * 8);
    if (B->ring_dirty) {
        k_fill_u64<<<fill_grid, 256, 0, s>>>((uint64_t*)B->ring, ~0ull, B->cap_ring * 4);
        B->ring_dirty = false;
    }
    if (B->dirty) {
        for (int i = 0; i < 3; ++i) {
            k_fill_u64<<<fill_grid, 256, 0, s>>>(B->chg[i], 0ull, B->cap_V * B->cap_W);
            k_fill_u32<<<fill_grid, 256, 0, s>>>(B->inq[i], 0u, B->cap_V);
        }
        B->dirty = false;
    }

    // ---- all (vertex, chunk) items of the graph, rebuilt when the graph or chunk changes ----
    if (B->all_uid != G->uid || B->all_chunk != chunk || !B->all_items) {
        if (!grow_alloc(B->all_items, B->cap_all, items_cap, 8)) {
            gst_query_bufs_free(ctx);
            return gst_set_error(ctx, GST_OOM, "solve: item list allocation failed");
        }
        if (!B->all_cnt) {
            uint64_t d = 0;
            if (!grow_alloc(B->all_cnt, d, 1, 4)) { gst_query_bufs_free(ctx); return gst_set_error(ctx, GST_OOM, "solve: allocation failed"); }
        }
        k_fill_u32<<<1, 64, 0, s>>>(B->all_cnt, 0u, 1);
        k_compact<<<grid_for(n, 256, fill_grid), 256, 0, s>>>(nullptr, G->d_row, n, chunk, B->all_items, B->all_cnt, nullptr,
                                                              HUB_DEG);
        uint32_t na = 0;
        HIPCHK(hipMemcpyAsync(&na, B->all_cnt, 4, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        B->n_all = na; B->all_uid = G->uid; B->all_chunk = chunk;
        uint64_t tot = 0;
        for (uint32_t v = 0; v < n; ++v) tot += (uint64_t)((G->h_row[v + 1] - G->h_row[v] + chunk - 1) / chunk);
        B->n_items_total = tot;
    }

    // ---- group bits ----
    std::vector<int32_t> hgrp((size_t)grp_need);
    for (int i = 0; i <= g; ++i) hgrp[(size_t)i] = groups->grp_ptr[i];
    for (int32_t j = 0; j < groups->grp_ptr[g]; ++j) hgrp[(size_t)(g + 1 + j)] = groups->grp_vtx[j];
    HIPCHK(hipMemcpyAsync(B->grp, hgrp.data(), grp_need * 4, hipMemcpyHostToDevice, s));
    k_fill_u32<<<grid_for(n, 256, fill_grid), 256, 0, s>>>(B->gm, 0u, n);
    k_set_group_bits<<<64, 256, 0, s>>>(B->gm, B->grp, B->grp + g + 1, g);
    k_fill_u64<<<1, 256, 0, s>>>((uint64_t*)B->stats, 0ull, STAT_WORDS);
    k_fill_u64<<<1, 64, 0, s>>>((uint64_t*)B->root, GST_INF_WORD, 1);
    k_fill_u32<<<1, 256, 0, s>>>((uint32_t*)B->ctl, 0u, sizeof(RoundCtl) * NCTL / 4);
    k_fill_u32<<<1, 64, 0, s>>>((uint32_t*)B->actl, 0u, sizeof(AsyncCtl) / 4);

    std::vector<uint64_t> lvl_base(GST_MAX_GROUPS + 2, 0);
    for (int k = 1; k <= g; ++k) lvl_base[k] = (uint64_t)n * (uint64_t)L.off[k];

    // events: pairs (start,end) per phase; kind 0 = grow, 1 = merge, 2 = trace
    struct Ev { hipEvent_t a, b; int kind; };
    std::vector<Ev> evs;
    auto new_ev = [&](int kind) -> Ev* {
        Ev e; e.kind = kind;
        hipEventCreate(&e.a); hipEventCreate(&e.b);
        evs.push_back(e);
        return &evs.back();
    };
    auto cleanup_evs = [&]() { for (auto& e : evs) { hipEventDestroy(e.a); hipEventDestroy(e.b); } };

    // 5 workgroups x 4 waves per CU (occupancy 5); the arrival counters hold up to 64 x 32 groups
    const int grow_grid = std::max(1, std::min(ctx->cu_count * 5, 64 * 32));
    const int rounds_per_sync = opts.rounds_per_sync > 0 ? opts.rounds_per_sync : 8;
    uint32_t total_rounds = 0, grow_launches = 0, merge_launches = 0, async_launches = 0;
    const int mode = opts.grow_mode;   // 0 tuned, 1 rounds only, 2 work queue only, 3 rounds then queue
    // chained rounds once every level's round count is known for this graph and group count
    gst_query_bufs::Budget* bud = nullptr;
    for (auto& b : B->budgets)
        if (b.uid == (G->budget_key ? G->budget_key : G->uid) && b.ng == g && b.ub == has_ub) bud = &b;
    const bool chained = (mode == 0 || mode == 1) && opts.rounds_per_sync == 0 && bud != nullptr &&
                         !opts.reserved[0];
    for (int k = 0; k < GST_MAX_GROUPS + 2; ++k) B->level_rounds[k] = chained ? bud->rounds[k] : 0;
    const int async_grid = std::max(1, ctx->cu_count * 5);
    int* ctx_level_rounds = B->level_rounds;

    static const bool trace_rounds = std::getenv("GST_TRACE_ROUNDS") != nullptr;
    if (trace_rounds) {
        if (!B->rlog && hipMalloc(&B->rlog, sizeof(uint32_t) * NCTL * RLOG_ROUNDS * 4) != hipSuccess)
            return gst_set_error(ctx, GST_OOM, "round log allocation failed");
        k_fill_u32<<<64, 256, 0, s>>>(B->rlog, 0xFFFFFFFFu, (uint64_t)NCTL * RLOG_ROUNDS * 4);
    }
    // the automatic bound: the spider (g >= 3), and level 1 under its running spider (half bound on)
    const bool spider_on = opts.auto_bound && g >= 3;
    const bool run_spider = spider_on && opts.half_bound == 0 && GST_RUN_SPIDER;
    const uint32_t* tour_ptr = nullptr;   // set once level 1 is final
    auto run_grow = [&](int k) -> gst_status {
        GrowArgs ga;
        std::memset(&ga, 0, sizeof(ga));
        ga.Tk = B->T + lvl_base[k];
        ga.T1 = B->T + lvl_base[1];
        ga.masks = B->masks + L.off[k];
        ga.row = G->d_row; ga.arc = (const uint2*)G->d_arc;
        for (int i = 0; i < 3; ++i) { ga.list[i] = B->lists[i]; ga.chg[i] = B->chg[i]; ga.inq[i] = B->inq[i]; }
        ga.ctl = B->ctl + k; ga.stats = B->stats;
        ga.m = (uint32_t)L.m[k]; ga.W = (uint32_t)((L.m[k] + 63) / 64); ga.g = (uint32_t)g;
        ga.chunk = chunk; ga.ub = ub_bits; ga.gmax = bb.gmax; ga.lbcap = bb.lbcap;
        // the one-label lower bound reads the group distances of level 1, which are final only
        // after level 1 has converged: never prune with it inside level 1
        ga.use_lb = (opts.prune_lb && k >= 2) ? 1u : 0u;
        ga.tour = k >= 2 ? tour_ptr : nullptr;
        // activity bytes once a bound is known (after level 1 with the automatic bound)
        ga.act = (k >= 2 && GST_MERGE_SKIP && ub_bits != GST_INF_COST) ? B->act + (size_t)k * n : nullptr;
        // level 1 of an automatically bounded solve grows under its running spider (word 0)
        ga.run_sp = (k == 1 && run_spider) ? B->spider : nullptr;
        ga.all_items = B->all_items; ga.n_all = B->n_all; ga.n_all_dev = nullptr; ga.n = n;
        {
            const int pm = opts.pull_permille == 0 ? 450 : opts.pull_permille;
            ga.pull_thresh = pm < 0 ? 0xFFFFFFFFu
                                    : (uint32_t)std::max<uint64_t>(1, B->n_items_total * (uint64_t)pm / 1000);
        }
        ga.ring = B->ring; ga.ring_mask = (uint32_t)(B->cap_ring - 1); ga.actl = B->actl;
        ga.rlog = trace_rounds ? B->rlog + (size_t)k * RLOG_ROUNDS * 4 : nullptr;
        uint32_t launch_seq = 0;   // the level's control block starts with seq 0
        ga.seq = 0;
        ga.push_lanes = opts.push_lanes > 0 ? (uint32_t)opts.push_lanes
                        : itempar_level(GST_SOLO_ITEMPAR, k) ? PUSH_ITEM_PAR : default_push_lanes(false);
        // the library's item-parallel choice: sparse rounds in its default lane groups
        ga.push_lanes_sparse = (opts.push_lanes == 0 && ga.push_lanes == PUSH_ITEM_PAR) ? default_push_lanes(false) : 0u;
        auto launch_async = [&]() {
            Ev* e = new_ev(0);
            hipEventRecord(e->a, s);
            k_async_begin<<<1, 64, 0, s>>>(B->ctl + k, B->actl);
            if (flt) k_grow_async<true><<<async_grid, GROW_BLOCK, 0, s>>>(ga);
            else k_grow_async<false><<<async_grid, GROW_BLOCK, 0, s>>>(ga);
            k_async_end<<<1, 64, 0, s>>>(B->ctl + k, B->actl);
            hipEventRecord(e->b, s);
            grow_launches++;
            async_launches++;
        };
        // work queue for the whole level: bounded solves (sparse frontiers) and level 1 (the
        // frontier starts at the group members); no host round trip
        if (mode == 2) {
            launch_async();
            HIPCHK(hipGetLastError());
            return GST_OK;
        }
        // chained: a round budget learned from earlier solves, no host round trip; convergence
        // of every level is checked once at the end of the query
        if (chained) {
            const int budget = ctx_level_rounds[k] + std::max(3, ctx_level_rounds[k] / 3);
            Ev* e = new_ev(0);
            hipEventRecord(e->a, s);
            for (int rr = 0; rr < budget; ++rr) {
                ga.seq = ++launch_seq;
                if (flt) k_grow_round<true><<<grow_grid, GROW_BLOCK, 0, s>>>(ga);
                else k_grow_round<false><<<grow_grid, GROW_BLOCK, 0, s>>>(ga);
                grow_launches++;
            }
            hipEventRecord(e->b, s);
            HIPCHK(hipGetLastError());
            return GST_OK;
        }
        // dense phase in rounds (pull while the queue holds >= the pull threshold), then the
        // work queue finishes the level once the frontier is sparse
        int batch = opts.rounds_per_sync > 0 ? rounds_per_sync
                                             : (mode != 3 ? std::max(8, ctx_level_rounds[k] + 2)
                                                          : std::max(2, ctx_level_rounds[k]));
        for (int it = 0; it < 100000; ++it) {
            Ev* e = new_ev(0);
            hipEventRecord(e->a, s);
            for (int rr = 0; rr < batch; ++rr) {
                ga.seq = ++launch_seq;
                if (flt) k_grow_round<true><<<grow_grid, GROW_BLOCK, 0, s>>>(ga);
                else k_grow_round<false><<<grow_grid, GROW_BLOCK, 0, s>>>(ga);
                grow_launches++;
            }
            hipEventRecord(e->b, s);
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(B->h_ctl, B->ctl + k, sizeof(RoundCtl), hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            if (B->h_ctl->err & 1u) return gst_set_error(ctx, GST_COST_OVERFLOW, "integer path cost exceeds 32 bits");
            if (B->h_ctl->level_done) break;
            if (mode == 3 && B->h_ctl->cnt[B->h_ctl->round % 3] < ga.pull_thresh) {
                ctx_level_rounds[k] = (int)B->h_ctl->round;
                total_rounds += B->h_ctl->round;
                launch_async();
                HIPCHK(hipGetLastError());
                return GST_OK;
            }
            batch = opts.rounds_per_sync > 0 ? rounds_per_sync : (mode != 3 ? 4 : 1);
        }
        ctx_level_rounds[k] = (int)B->h_ctl->round + 1;
        if (!B->h_ctl->level_done) { B->dirty = true; return gst_set_error(ctx, GST_INTERNAL, "grow did not converge"); }
        total_rounds += B->h_ctl->round + 1;
        return GST_OK;
    };

    gst_status st = GST_OK;
    // the spider words: [0] the spider (tree) bound, [1] the cheapest spider's own cost, [2] the running
    // spider as level 1 ended
    if (spider_on) {
        if (!B->spider && hipMalloc(&B->spider, 24) != hipSuccess) return gst_set_error(ctx, GST_OOM, "solve: allocation failed");
        k_fill_u64<<<1, 64, 0, s>>>((uint64_t*)B->spider, GST_INF_WORD, 3);
    }
    // ---- level 1 ----
    {
        Ev* e = new_ev(1);
        hipEventRecord(e->a, s);
        k_init_level1<<<grid_for(n, 256, fill_grid), 256, 0, s>>>(B->T + lvl_base[1], B->gm, G->d_row, n, (uint32_t)g,
                                                                   B->chg[0], B->inq[0], g > 1 ? 1 : 0);
        if (g > 1)
            k_compact<<<grid_for(n, 256, fill_grid), 256, 0, s>>>(B->inq[0], G->d_row, n, chunk, B->lists[0],
                                                                  &B->ctl[1].cnt[0], B->stats, 0u);
        hipEventRecord(e->b, s);
        if (g > 1) st = run_grow(1);
        else k_root_scan_l1<<<grid_for(n, 256, fill_grid), 256, 0, s>>>(B->T, n, B->root);
    }
    // ---- automatic bound: the cheapest spider once the group distances are final ----
    double spider_costs[2] = {0.0, 0.0};   // the spider's sum of distances, its post-processed tree
    if (st == GST_OK && spider_on) {
        // word 0 holds the running spider of level 1 (or inf): keep it, then lower word 0 to the
        // cheapest spider of the final distances
        HIPCHK(hipMemcpyAsync(B->spider + 2, B->spider, 8, hipMemcpyDeviceToDevice, s));
        const uint64_t* T1 = B->T + lvl_base[1];
        if (flt) k_spider_scan<true><<<grid_for(n, 256, fill_grid), 256, 0, s>>>(T1, n, (uint32_t)g, B->spider);
        else k_spider_scan<false><<<grid_for(n, 256, fill_grid), 256, 0, s>>>(T1, n, (uint32_t)g, B->spider);
        // keep the spider's own cost, then MST post-process its paths (lowers word 0 when cheaper;
        // skipped on the device when the caller's bound is at or below 90% of the spider's cost)
        HIPCHK(hipMemcpyAsync(B->spider + 1, B->spider, 8, hipMemcpyDeviceToDevice, s));
        if (flt) k_spider_tree<true><<<1, MST_THREADS, 0, s>>>(T1, (uint32_t)g, B->gm, G->d_row, (const uint2*)G->d_arc, B->spider, nullptr, ub_bits);
        else k_spider_tree<false><<<1, MST_THREADS, 0, s>>>(T1, (uint32_t)g, B->gm, G->d_row, (const uint2*)G->d_arc, B->spider, nullptr, ub_bits);
        unsigned long long spw[3] = {GST_INF_WORD, GST_INF_WORD, GST_INF_WORD};
        HIPCHK(hipMemcpyAsync(spw, B->spider, 24, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        // level 1 grew under its running spider: its states up to half of the spider as it stood at
        // the level's end are exact, so the later levels' lower bounds read level 1 capped there
        if (run_spider && (uint32_t)(spw[2] >> 32) != GST_INF_COST)
            bb.lbcap = std::min(bb.lbcap, gst_spider_gmax_bits((uint32_t)(spw[2] >> 32), flt));
        const uint32_t sp = (uint32_t)(spw[0] >> 32);
        if (sp != GST_INF_COST) {
            const uint32_t sp1 = (uint32_t)(spw[1] >> 32);
            float f0, f1;
            std::memcpy(&f0, &sp, 4);
            std::memcpy(&f1, &sp1, 4);
            spider_costs[0] = flt ? (double)f1 : (double)sp1;
            spider_costs[1] = flt ? (double)f0 : (double)sp;
        }
        if (sp != GST_INF_COST) {
            double U;
            if (flt) { float f; std::memcpy(&f, &sp, 4); U = (double)f; } else U = (double)sp;
            const gst_bound_bits sb = gst_make_bound(U, flt, opts.half_bound);
            // level 1 keeps the cap it was grown under (bb.lbcap)
            if (sb.ub < ub_bits) ub_bits = sb.ub;
            if (sb.gmax < bb.gmax) bb.gmax = sb.gmax;
            bb.ub = ub_bits;
        }
    }
    // ---- tour table: group-to-group walks from the final level 1 ----
    const uint32_t* tour = nullptr;
    // (with three groups a level-2 state misses one group, where the tour bound is the one-label one)
    if (st == GST_OK && opts.prune_lb && opts.tour_bound == 0 && g >= 4 && ub_bits != GST_INF_COST) {
        const size_t words = ((size_t)1 << g) * (size_t)(g + 1) + ((size_t)1 << g) * (size_t)(g * g);
        if (B->tour_g < g) {
            hipFree(B->tour);
            B->tour = nullptr;
            B->tour_g = 0;
            if (hipMalloc(&B->tour, words * 4) != hipSuccess) return gst_set_error(ctx, GST_OOM, "solve: allocation failed");
            B->tour_g = g;
        }
        uint32_t* Pw = B->tour + ((size_t)1 << g) * (size_t)(g + 1);
        const size_t tl = tour_lds_bytes((uint32_t)g);
        const uint64_t* T1 = B->T + lvl_base[1];
        if (tl) {
            if (flt) k_tour_table<true, true><<<1, 256, tl, s>>>(T1, B->grp, (uint32_t)g, bb.lbcap, Pw, B->tour);
            else k_tour_table<false, true><<<1, 256, tl, s>>>(T1, B->grp, (uint32_t)g, bb.lbcap, Pw, B->tour);
        } else {
            if (flt) k_tour_table<true, false><<<1, 256, 0, s>>>(T1, B->grp, (uint32_t)g, bb.lbcap, Pw, B->tour);
            else k_tour_table<false, false><<<1, 256, 0, s>>>(T1, B->grp, (uint32_t)g, bb.lbcap, Pw, B->tour);
        }
        HIPCHK(hipGetLastError());
        tour = B->tour;
    }
    tour_ptr = tour;
    // ---- levels 2..g ----
    for (int k = 2; k <= g && st == GST_OK; ++k) {
        const bool full = (k == g);
        MergeArgs ma;
        ma.T = B->T; ma.pairs = B->pairs + L.pair_off[k]; ma.masks = B->masks; ma.gm = B->gm; ma.row = G->d_row;
        ma.list0 = B->lists[0]; ma.chg0 = B->chg[0]; ma.inq0 = B->inq[0]; ma.ctl = B->ctl + k; ma.root = B->root;
        ma.stats = B->stats;
        for (int j = 0; j < GST_MAX_GROUPS + 2; ++j) { ma.lvl_base[j] = lvl_base[j]; ma.m[j] = (uint32_t)L.m[j]; ma.off[j] = (uint32_t)L.off[j]; }
        ma.n = n; ma.g = (uint32_t)g; ma.k = (uint32_t)k; ma.P = (uint32_t)L.pair_cnt[k];
        ma.W = (uint32_t)((L.m[k] + 63) / 64); ma.chunk = chunk; ma.ub = ub_bits; ma.use_lb = opts.prune_lb ? 1u : 0u;
        ma.gmax = bb.gmax; ma.lbcap = bb.lbcap;
        ma.spider = nullptr;
        ma.tour = tour;
        ma.grp = nullptr;
        ma.tour_w = nullptr;
        ma.act = (GST_MERGE_SKIP && ub_bits != GST_INF_COST) ? B->act : nullptr;
        ma.live = k >= 3 ? B->live : nullptr;
        ma.is_full = full ? 1u : 0u;
        const uint32_t Lk = (uint32_t)L.off[k], mk = (uint32_t)L.m[k], Wk = (mk + 63) / 64;
        const size_t pair_bytes = (size_t)L.pair_cnt[k] * 8;
        const bool pin = pair_bytes <= MERGE_PAIRS_LDS;
        const uint32_t VB = merge_tile_vertices(Lk, Wk, (size_t)opts.merge_tile_kb * 1024);
        ma.VB = VB;
        ma.pairs_in_lds = pin ? 1u : 0u;
        const size_t lds = merge_lds_bytes(VB, Lk, Wk, pair_bytes, pin);
        const int mgrid = std::max(1, std::min<int>((int)((n + VB - 1) / VB), ctx->cu_count * 8));
        Ev* e = new_ev(1);
        hipEventRecord(e->a, s);
        if (merge_uses_list(ma))
            k_merge_live<<<std::max(1, std::min<int>((int)((n + 255) / 256), ctx->cu_count * 8)), 256, 0, s>>>(ma);
        if (flt) k_merge<true><<<mgrid, MERGE_THREADS, lds, s>>>(ma);
        else k_merge<false><<<mgrid, MERGE_THREADS, lds, s>>>(ma);
        hipEventRecord(e->b, s);
        if (!full)
            k_compact<<<grid_for(n, 256, fill_grid), 256, 0, s>>>(B->inq[0], G->d_row, n, chunk, B->lists[0],
                                                                  &B->ctl[k].cnt[0], B->stats, 0u);
        merge_launches++;
        HIPCHK(hipGetLastError());
        if (!full) st = run_grow(k);
    }
    if (st != GST_OK) { B->dirty = true; cleanup_evs(); return st; }

    // ---- root + traceback ----
    unsigned long long h_root = GST_INF_WORD;
    HIPCHK(hipMemcpyAsync(&h_root, B->root, 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(B->h_ctl, B->ctl, sizeof(RoundCtl) * NCTL, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    uint32_t all_err = 0;
    bool unconverged = false;
    for (int k = 0; k <= g; ++k) {
        all_err |= B->h_ctl[k].err;
        if (k >= 1 && k < g && g > 1 && !B->h_ctl[k].level_done) unconverged = true;
    }
    if (all_err & 1u) { cleanup_evs(); return gst_set_error(ctx, GST_COST_OVERFLOW, "integer path cost exceeds 32 bits"); }
    // learn the round budget of every level from this solve
    {
        if (!bud) {
            if (B->budgets.size() >= 16) B->budgets.erase(B->budgets.begin());
            B->budgets.push_back(gst_query_bufs::Budget{G->budget_key ? G->budget_key : G->uid, g, has_ub, {0}});
            bud = &B->budgets.back();
        }
        for (int k = 1; k < g; ++k) {
            const int used = (int)B->h_ctl[k].round + 1;
            bud->rounds[k] = chained ? std::max(bud->rounds[k], used) : used;
        }
    }
    if (chained) {
        total_rounds = 0;
        for (int k = 1; k < g; ++k) total_rounds += B->h_ctl[k].round + 1;
        if (unconverged) {
            // a level ran out of its round budget: redo the query with a host check per batch
            cleanup_evs();
            B->dirty = true;
            gst_opts o2 = opts;
            o2.reserved[0] = 1;   // internal: no chaining for the redo
            return gst_dpbf_solve(ctx, G, groups, &o2, res);
        }
    }
    if (all_err & 4u) {
        // the work queue gave up (wait bound or ring capacity): redo the query on the round path
        cleanup_evs();
        B->dirty = true;
        B->ring_dirty = true;
        gst_opts o2 = opts;
        o2.grow_mode = 1;
        return gst_dpbf_solve(ctx, G, groups, &o2, res);
    }

    std::memset(res, 0, sizeof(*res));
    res->n_groups = g;
    res->root = -1;
    res->used_ub = has_ub ? 1 : 0;
    if (h_root == GST_INF_WORD) {
        cleanup_evs();
        if (has_ub)
            return gst_set_error(ctx, GST_INFEASIBLE, "no tree of cost <= bound %.9g touches all %d groups", opts.ub, g);
        return gst_set_error(ctx, GST_INFEASIBLE, "no tree touches all %d groups", g);
    }
    const uint32_t root_cost_bits = (uint32_t)(h_root >> 32);
    const uint32_t root_v = (uint32_t)(h_root & 0xFFFFFFFFu);
    if (flt) { float f; std::memcpy(&f, &root_cost_bits, 4); res->cost = (double)f; }
    else res->cost = (double)root_cost_bits;
    res->root = (int32_t)root_v;

    if (opts.want_tree) {
        TraceArgs ta;
        ta.T = B->T; ta.pos_of = B->pos_of; ta.row = G->d_row; ta.arc = (const uint2*)G->d_arc; ta.root = B->root;
        ta.flt = flt ? 1u : 0u;
        ta.out_u = B->edges; ta.out_v = B->edges + ecap; ta.out_w = B->edges + 2 * ecap; ta.out_n = B->edges + 3 * ecap;
        ta.ostride = 1;
        ta.ctl = B->ctl;   // entry 0 carries query-level flags
        for (int j = 0; j < GST_MAX_GROUPS + 2; ++j) { ta.lvl_base[j] = lvl_base[j]; ta.m[j] = (uint32_t)L.m[j]; ta.off[j] = (uint32_t)L.off[j]; }
        ta.full = (uint32_t)L.full; ta.cap = ecap; ta.max_iter = n + 4u * (uint32_t)g + 64u;
        Ev* e = new_ev(2);
        hipEventRecord(e->a, s);
        k_traceback<<<1, WAVE, 0, s>>>(ta);
        hipEventRecord(e->b, s);
        HIPCHK(hipGetLastError());
        uint32_t nout = 0;
        HIPCHK(hipMemcpyAsync(&nout, B->edges + 3 * ecap, 4, hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(B->h_ctl, B->ctl, sizeof(RoundCtl), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if ((B->h_ctl[0].err & 2u) || nout > ecap) {
            cleanup_evs();
            return gst_set_error(ctx, GST_INTERNAL, "device traceback failed (edges=%u)", nout);
        }
        std::vector<uint32_t> hu(nout), hv(nout), hw(nout);
        if (nout) {
            HIPCHK(hipMemcpyAsync(hu.data(), B->edges, nout * 4, hipMemcpyDeviceToHost, s));
            HIPCHK(hipMemcpyAsync(hv.data(), B->edges + ecap, nout * 4, hipMemcpyDeviceToHost, s));
            HIPCHK(hipMemcpyAsync(hw.data(), B->edges + 2 * ecap, nout * 4, hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
        }
        gst_finish_tree(flt, nout, hu.data(), hv.data(), hw.data(), res);
    }

    if (trace_rounds) {
        std::vector<uint32_t> lg((size_t)NCTL * RLOG_ROUNDS * 4);
        HIPCHK(hipMemcpyAsync(lg.data(), B->rlog, lg.size() * 4, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        for (int k = 1; k < g; ++k) {
            std::fprintf(stderr, "[rounds] level %d (masks %d):", k, L.m[k]);
            for (uint32_t r = 0; r < RLOG_ROUNDS; ++r) {
                const uint32_t* e = &lg[((size_t)k * RLOG_ROUNDS + r) * 4];
                if (e[0] == 0xFFFFFFFFu) break;
                // wall clock runs at 100 MHz: 10 ns per tick
                const double us = e[3] == 0xFFFFFFFFu ? -1.0 : (double)(uint32_t)(e[3] - e[2]) * 0.01;
                std::fprintf(stderr, " %u%s/%u:%.1f", e[0] & 0x7FFFFFFFu, (e[0] >> 31) ? "P" : "", e[1], us);
            }
            std::fprintf(stderr, "\n");
        }
        std::fprintf(stderr, "[merges] us per launch, levels 2..%d:", g);
        for (auto& e : evs) {
            if (e.kind != 1) continue;
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, e.a, e.b));
            std::fprintf(stderr, " %.1f", ms * 1000.0);
        }
        std::fprintf(stderr, "\n");
    }
    // ---- stats ----
    unsigned long long hst[STAT_WORDS], hs[ST_NSTAT];
    HIPCHK(hipMemcpyAsync(hst, B->stats, sizeof(hst), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    stat_sum(hst, hs);
    gst_stats& S = res->stats;
    S.grow_items = hs[ST_ITEMS]; S.grow_arcs = hs[ST_ARCS]; S.grow_relax = hs[ST_RELAX];
    S.grow_updates = hs[ST_UPD]; S.grow_enqueued = hs[ST_ENQ]; S.grow_new_states = hs[ST_NEW]; S.grow_atomics = hs[ST_ATOM];
    S.merge_vertices = hs[ST_MVERT]; S.merge_pairs = hs[ST_MPAIRS];
    S.spider_cost = spider_costs[0]; S.spider_tree_cost = spider_costs[1];
    S.grow_rounds = total_rounds; S.grow_launches = grow_launches; S.merge_launches = merge_launches;
    S.async_launches = async_launches;
    for (auto& e : evs) {
        float ms = 0;
        hipEventElapsedTime(&ms, e.a, e.b);
        if (e.kind == 0) S.t_grow_ms += ms;
        else if (e.kind == 1) S.t_merge_ms += ms;
        else S.t_trace_ms += ms;
    }
    cleanup_evs();
    S.grow_bytes = gst_grow_bytes(S);
    S.merge_bytes = hs[ST_MBYTES];
    S.t_total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    return GST_OK;
}

// MST post-processing of a caller's vertex set (k_mst_induced): one workgroup on the context's
// stream; the caller has checked the arguments.
gst_status gst_dpbf_mst_postprocess(gst_ctx* ctx, const gst_graph* G, const int32_t* vertices, int32_t nv, int32_t* eu,
                                    int32_t* ev, double* ew, int32_t* ne, double* cost) {
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)ctx->stream;
    const bool flt = G->w_type == GST_W_FLOAT32;
    const size_t out_words = 1 + 3 * (size_t)MST_SMAX;
    uint32_t* d = nullptr;
    if (hipMalloc(&d, ((size_t)nv + out_words) * 4) != hipSuccess)
        return gst_set_error(ctx, GST_OOM, "mst_postprocess: allocation failed");
    std::vector<uint32_t> out(out_words);
    hipError_t e = hipMemcpyAsync(d, vertices, (size_t)nv * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        if (flt) k_mst_induced<true><<<1, MST_THREADS, 0, s>>>(d, (uint32_t)nv, G->d_row, (const uint2*)G->d_arc, d + nv);
        else k_mst_induced<false><<<1, MST_THREADS, 0, s>>>(d, (uint32_t)nv, G->d_row, (const uint2*)G->d_arc, d + nv);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out.data(), d + nv, out_words * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    hipFree(d);
    if (e != hipSuccess) return gst_set_error(ctx, GST_HIP_ERROR, "mst_postprocess: %s", hipGetErrorString(e));
    if (out[0] == MST_EMPTY)
        return gst_set_error(ctx, GST_INVALID_ARG, "mst_postprocess: more than %u distinct vertices or %u induced edges",
                             MST_SMAX, MST_EMAX);
    double c = 0.0;
    for (uint32_t t = 0; t < out[0]; ++t) {
        eu[t] = (int32_t)out[1 + 3 * t];
        ev[t] = (int32_t)out[2 + 3 * t];
        const uint32_t wb = out[3 + 3 * t];
        float f;
        std::memcpy(&f, &wb, 4);
        ew[t] = flt ? (double)f : (double)wb;
        c += ew[t];
    }
    *ne = (int32_t)out[0];
    *cost = c;
    return GST_OK;
}
