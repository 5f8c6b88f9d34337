// One exact GST query without any allocation in the call: gst_solve_noalloc.
//
// For callers that cannot allocate (code under HIP graph capture, tight serving loops, harnesses
// that forbid allocation).  The per-query device state of the one-query solve (state table,
// frontier bitmaps and lists, control blocks, group lists, level tables, tour table, traceback
// output) is carved out of one fixed file-scope device array, `scratch`, of
// GST_NOALLOC_SCRATCH_BYTES; the host-side work (readbacks, the tree's deduplication and spanning
// forest) uses buffers of fixed size made when the context is created; the tree is written into the
// caller's arrays.  The call launches the same kernels as gst_solve's round path (grow rounds checked
// by the host every few rounds, merges, spider bound, tour table, device traceback), so it returns
// the same exact optimum.
//
// The array is one per GPU and keeps the last call's contents: everything that must start at zero
// is cleared by each call, and calls that overlap on one GPU (two contexts on the same device used
// from two threads at once) race on it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>

#include "gst_dpbf_dev.h"

#define HIPCHK(call)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return gst_set_error(ctx, GST_HIP_ERROR, "%s failed: %s (%s:%d)", #call,         \
                                 hipGetErrorString(e_), __FILE__, __LINE__);                 \
    } while (0)

// The scratch array.  In host code its name is a host-side stand-in, not a device address: every
// host use goes through the address hipGetSymbolAddress returned for the current device.
static __device__ __align__(256) unsigned char scratch[GST_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int NA_MAX_DEVICES = 64;
std::mutex na_mu;
void* na_dev_ptr[NA_MAX_DEVICES] = {nullptr};   // scratch address per device, looked up once
hipError_t na_dev_err[NA_MAX_DEVICES] = {hipSuccess};
bool na_dev_done[NA_MAX_DEVICES] = {false};

// one traced edge, keyed by its endpoints (smaller id in the high half)
struct NaEdge {
    uint64_t key;
    uint32_t w;
};

// carves 256-byte aligned pieces off the scratch; with base == nullptr it only counts bytes
struct Carve {
    unsigned char* base;
    size_t off = 0;
    template <class T>
    T* take(size_t count) {
        off = (off + 255) & ~(size_t)255;
        T* p = base ? (T*)(base + off) : nullptr;
        off += std::max<size_t>(count, 1) * sizeof(T);
        return p;
    }
};

}  // namespace

// host working memory of one context (fixed size, made at set-up)
struct gst_noalloc_host {
    void* scratch = nullptr;                 // device address of this device's scratch
    gst_status init_status = GST_OK;
    char init_msg[256] = {0};
    RoundCtl ctl[NCTL];                      // control blocks read back
    unsigned long long stats[STAT_WORDS];    // striped counters, summed by stat_sum
    uint32_t tri[3 * GST_NOALLOC_MAX_EDGES + 1];   // traced (u, v, w) triples, then their count
    NaEdge edges[GST_NOALLOC_MAX_EDGES];
    uint32_t verts[2 * GST_NOALLOC_MAX_EDGES];
    uint32_t parent[2 * GST_NOALLOC_MAX_EDGES];
};

// Context set-up, with the context's device current: the scratch address of this device (looked up
// once per device and kept), the host working block, and the level tables of every group count
// (built here so that a call never builds one).
gst_status gst_noalloc_init(gst_ctx* ctx) {
    gst_noalloc_host* H = new (std::nothrow) gst_noalloc_host();
    if (!H) return gst_set_error(ctx, GST_OOM, "no-allocation solve: host working block allocation failed");
    ctx->na = H;
    for (int g = 1; g <= GST_MAX_GROUPS; ++g) (void)gst_get_levels(ctx, g);
    ctx->last_error.reserve(512);   // error messages of a call then fit without a reallocation
    const int d = ctx->device;
    if (d < 0 || d >= NA_MAX_DEVICES) {
        H->init_status = GST_INVALID_ARG;
        std::snprintf(H->init_msg, sizeof(H->init_msg), "no-allocation solve: device %d above the supported %d", d,
                      NA_MAX_DEVICES);
        return GST_OK;
    }
    std::lock_guard<std::mutex> lk(na_mu);
    if (!na_dev_done[d]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        na_dev_err[d] = e;
        na_dev_ptr[d] = (e == hipSuccess) ? p : nullptr;
        na_dev_done[d] = true;
    }
    if (na_dev_err[d] != hipSuccess || !na_dev_ptr[d]) {
        H->init_status = GST_HIP_ERROR;
        std::snprintf(H->init_msg, sizeof(H->init_msg), "no-allocation solve: hipGetSymbolAddress(scratch) failed: %s",
                      hipGetErrorString(na_dev_err[d]));
        return GST_OK;   // the context serves every other entry point
    }
    H->scratch = na_dev_ptr[d];
    return GST_OK;
}

void gst_noalloc_fini(gst_ctx* ctx) {
    delete ctx->na;
    ctx->na = nullptr;
}

namespace {

// the traced triples -> a spanning forest of their union in the caller's arrays (as
// gst_finish_tree: duplicates dropped; every edge on a cycle of the union weighs zero)
gst_status finish_tree_into(gst_ctx* ctx, gst_noalloc_host* H, bool flt, uint32_t nout, int32_t edge_cap, int32_t* eu,
                            int32_t* ev, double* ew, gst_result* res) {
    uint32_t ne = 0;
    for (uint32_t i = 0; i < nout; ++i) {
        uint32_t x = H->tri[3 * i], y = H->tri[3 * i + 1];
        if (x > y) std::swap(x, y);
        H->edges[ne++] = NaEdge{((uint64_t)x << 32) | y, H->tri[3 * i + 2]};
    }
    std::sort(H->edges, H->edges + ne, [](const NaEdge& a, const NaEdge& b) { return a.key < b.key; });
    ne = (uint32_t)(std::unique(H->edges, H->edges + ne, [](const NaEdge& a, const NaEdge& b) { return a.key == b.key; }) -
                    H->edges);
    uint32_t nv = 0;
    for (uint32_t i = 0; i < ne; ++i) {
        H->verts[nv++] = (uint32_t)(H->edges[i].key >> 32);
        H->verts[nv++] = (uint32_t)H->edges[i].key;
    }
    std::sort(H->verts, H->verts + nv);
    nv = (uint32_t)(std::unique(H->verts, H->verts + nv) - H->verts);
    for (uint32_t i = 0; i < nv; ++i) H->parent[i] = i;
    auto find = [&](uint32_t x) {
        while (H->parent[x] != x) { H->parent[x] = H->parent[H->parent[x]]; x = H->parent[x]; }
        return x;
    };
    auto idx = [&](uint32_t v) { return (uint32_t)(std::lower_bound(H->verts, H->verts + nv, v) - H->verts); };
    uint32_t kept = 0;
    double tc = 0.0;
    for (uint32_t i = 0; i < ne; ++i) {
        const uint32_t a = find(idx((uint32_t)(H->edges[i].key >> 32))), b = find(idx((uint32_t)H->edges[i].key));
        if (a == b) continue;
        H->parent[a] = b;
        if ((int64_t)kept >= (int64_t)edge_cap)
            return gst_set_error(ctx, GST_INVALID_ARG, "no-allocation solve: the tree has more than edge_cap = %d edges",
                                 edge_cap);
        double wv;
        if (flt) { float f; std::memcpy(&f, &H->edges[i].w, 4); wv = (double)f; }
        else wv = (double)H->edges[i].w;
        eu[kept] = (int32_t)(H->edges[i].key >> 32);
        ev[kept] = (int32_t)(H->edges[i].key & 0xFFFFFFFFu);
        ew[kept] = wv;
        tc += wv;
        ++kept;
    }
    res->n_edges = (int32_t)kept;
    res->tree_cost = tc;
    return GST_OK;
}

}  // namespace

gst_status gst_dpbf_solve_noalloc(gst_ctx* ctx, const gst_graph* G, const gst_groups* groups, const gst_opts* opts_in,
                                  int32_t edge_cap, int32_t* eu, int32_t* ev, double* ew, gst_result* res) {
    gst_noalloc_host* H = ctx->na;
    if (!H) return gst_set_error(ctx, GST_NO_DEVICE, "no-allocation solve: context has no working block");
    if (H->init_status != GST_OK) return gst_set_error(ctx, H->init_status, "%s", H->init_msg);
    gst_opts opts{};
    opts.prune_lb = 1;
    opts.want_tree = 1;
    if (opts_in) opts = *opts_in;
    if (!gst_push_lanes_ok(opts.push_lanes))
        return gst_set_error(ctx, GST_INVALID_ARG, "push_lanes must be 0, 1, 8, 16, 32 or 64 (got %d)", opts.push_lanes);
    if (opts.half_bound != 0 && opts.half_bound != -1)
        return gst_set_error(ctx, GST_INVALID_ARG, "half_bound must be 0 or -1 (got %d)", opts.half_bound);
    if (opts.tour_bound != 0 && opts.tour_bound != -1)
        return gst_set_error(ctx, GST_INVALID_ARG, "tour_bound must be 0 or -1 (got %d)", opts.tour_bound);
    if (opts.merge_tile_kb < 0 || opts.merge_tile_kb > 48)
        return gst_set_error(ctx, GST_INVALID_ARG, "merge_tile_kb must be 0..48 (got %d)", opts.merge_tile_kb);
    if (opts.grow_mode != 0 && opts.grow_mode != 1)
        return gst_set_error(ctx, GST_INVALID_ARG, "no-allocation solve: grow_mode must be 0 or 1 (rounds)");
    if (opts.want_tree && (edge_cap < 0 || (edge_cap > 0 && (!eu || !ev || !ew))))
        return gst_set_error(ctx, GST_INVALID_ARG, "no-allocation solve: edge arrays missing for edge_cap = %d", edge_cap);
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)ctx->stream;
    const int g = groups->n_groups;
    const uint32_t n = (uint32_t)G->n;
    const gst_levels& L = gst_get_levels(ctx, g);   // built at set-up
    const bool flt = G->w_type == GST_W_FLOAT32;
    const bool has_ub = std::isfinite(opts.ub) && opts.ub > 0;
    gst_apply_tuned(ctx, G, g, has_ub, opts);
    gst_bound_bits bb = gst_make_bound(opts.ub, flt, opts.half_bound);
    uint32_t ub_bits = bb.ub;

    // ---- the scratch layout of this query ----
    uint32_t Wmax = 1;
    for (int k = 1; k <= g; ++k) Wmax = std::max<uint32_t>(Wmax, (uint32_t)((L.m[k] + 63) / 64));
    const uint32_t chunk = opts.chunk_arcs > 0 ? (uint32_t)opts.chunk_arcs : 32u;
    const uint64_t words = (uint64_t)n * (uint64_t)L.full;
    const uint64_t items_cap = (uint64_t)n + (uint64_t)G->m / chunk + 64;
    const uint32_t ecap = (uint32_t)std::min<uint64_t>((uint64_t)n + 2 * (uint64_t)g + 16, GST_NOALLOC_MAX_EDGES);
    const uint32_t members = (uint32_t)groups->grp_ptr[g];
    const size_t tour_words = ((size_t)1 << g) * (size_t)(g + 1 + g * g);
    struct Bufs {
        uint64_t* T; uint32_t* gm; uint64_t* chg[3]; uint32_t* inq[3]; uint64_t* lists[3]; RoundCtl* ctl;
        unsigned long long* root; unsigned long long* stats; unsigned long long* spider; uint32_t* tour; uint32_t* edges;
        int32_t* grp; uint64_t* pairs; uint32_t* masks; uint32_t* pos_of; uint64_t* all_items; uint32_t* all_cnt;
        uint8_t* act;
    } B;
    auto layout = [&](Carve& c) {
        B.T = c.take<uint64_t>(words);
        B.gm = c.take<uint32_t>(n);
        for (int i = 0; i < 3; ++i) {
            B.chg[i] = c.take<uint64_t>((size_t)n * Wmax);
            B.inq[i] = c.take<uint32_t>(n);
            B.lists[i] = c.take<uint64_t>(items_cap);
        }
        B.ctl = c.take<RoundCtl>(NCTL);
        B.root = c.take<unsigned long long>(1);
        B.stats = c.take<unsigned long long>(STAT_WORDS);
        B.spider = c.take<unsigned long long>(2);
        B.tour = c.take<uint32_t>(tour_words);
        B.edges = c.take<uint32_t>(3 * (size_t)ecap + 1);
        B.grp = c.take<int32_t>((size_t)g + 1 + members);
        B.pairs = c.take<uint64_t>(L.pairs.size());
        B.masks = c.take<uint32_t>(L.masks.size());
        B.pos_of = c.take<uint32_t>(L.pos_of.size());
        B.all_items = c.take<uint64_t>(items_cap);
        B.all_cnt = c.take<uint32_t>(1);
        B.act = c.take<uint8_t>((size_t)n * (size_t)(g + 1));
    };
    {
        Carve dry{nullptr};
        layout(dry);
        if (dry.off > (size_t)GST_NOALLOC_SCRATCH_BYTES)
            return gst_set_error(ctx, GST_INVALID_ARG,
                                 "no-allocation solve: the query needs %.1f MiB of device scratch (V=%u, %lld arcs, g=%d), "
                                 "above the fixed %u MiB",
                                 dry.off / 1048576.0, n, (long long)G->m, g, (unsigned)(GST_NOALLOC_SCRATCH_BYTES >> 20));
        Carve c{(unsigned char*)H->scratch};
        layout(c);
    }

    // ---- per-call state: cleared or uploaded (the scratch keeps the previous call's contents) ----
    const int fill_grid = std::max(1, ctx->cu_count * 8);
    for (int i = 0; i < 3; ++i) {
        k_fill_u64<<<fill_grid, 256, 0, s>>>(B.chg[i], 0ull, (uint64_t)n * Wmax);
        k_fill_u32<<<fill_grid, 256, 0, s>>>(B.inq[i], 0u, n);
    }
    k_fill_u32<<<grid_for(n, 256, fill_grid), 256, 0, s>>>(B.gm, 0u, n);
    k_fill_u64<<<1, 256, 0, s>>>((uint64_t*)B.stats, 0ull, STAT_WORDS);
    k_fill_u64<<<1, 64, 0, s>>>((uint64_t*)B.root, GST_INF_WORD, 1);
    k_fill_u32<<<1, 256, 0, s>>>((uint32_t*)B.ctl, 0u, sizeof(RoundCtl) * NCTL / 4);
    k_fill_u32<<<1, 64, 0, s>>>(B.all_cnt, 0u, 1);
    k_fill_u64<<<1, 64, 0, s>>>((uint64_t*)B.spider, GST_INF_WORD, 2);
    HIPCHK(hipMemcpyAsync(B.grp, groups->grp_ptr, ((size_t)g + 1) * 4, hipMemcpyHostToDevice, s));
    if (members) HIPCHK(hipMemcpyAsync(B.grp + g + 1, groups->grp_vtx, (size_t)members * 4, hipMemcpyHostToDevice, s));
    if (!L.pairs.empty()) HIPCHK(hipMemcpyAsync(B.pairs, L.pairs.data(), L.pairs.size() * 8, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(B.masks, L.masks.data(), L.masks.size() * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(B.pos_of, L.pos_of.data(), L.pos_of.size() * 4, hipMemcpyHostToDevice, s));
    // hub rows' work items for pull rounds
    k_compact<<<grid_for(n, 256, fill_grid), 256, 0, s>>>(nullptr, G->d_row, n, chunk, B.all_items, B.all_cnt, nullptr,
                                                          HUB_DEG);
    uint32_t n_all = 0;
    HIPCHK(hipMemcpyAsync(&n_all, B.all_cnt, 4, hipMemcpyDeviceToHost, s));
    k_set_group_bits<<<64, 256, 0, s>>>(B.gm, B.grp, B.grp + g + 1, g);
    uint64_t n_items_total = 0;
    for (uint32_t v = 0; v < n; ++v) n_items_total += (uint64_t)((G->h_row[v + 1] - G->h_row[v] + chunk - 1) / chunk);
    HIPCHK(hipStreamSynchronize(s));

    uint64_t lvl_base[GST_MAX_GROUPS + 2] = {0};
    for (int k = 1; k <= g; ++k) lvl_base[k] = (uint64_t)n * (uint64_t)L.off[k];
    const int grow_grid = std::max(1, std::min(ctx->cu_count * 5, 64 * 32));
    const int batch = opts.rounds_per_sync > 0 ? opts.rounds_per_sync : 8;
    uint32_t total_rounds = 0, grow_launches = 0, merge_launches = 0;
    const uint32_t* tour = nullptr;

    // grow rounds of level k until no cost drops, the host checking every `batch` rounds
    auto run_grow = [&](int k) -> gst_status {
        GrowArgs ga;
        std::memset(&ga, 0, sizeof(ga));
        ga.Tk = B.T + lvl_base[k];
        ga.T1 = B.T + lvl_base[1];
        ga.masks = B.masks + L.off[k];
        ga.row = G->d_row; ga.arc = (const uint2*)G->d_arc;
        for (int i = 0; i < 3; ++i) { ga.list[i] = B.lists[i]; ga.chg[i] = B.chg[i]; ga.inq[i] = B.inq[i]; }
        ga.ctl = B.ctl + k; ga.stats = B.stats;
        ga.m = (uint32_t)L.m[k]; ga.W = (uint32_t)((L.m[k] + 63) / 64); ga.g = (uint32_t)g;
        ga.chunk = chunk; ga.ub = ub_bits; ga.gmax = bb.gmax; ga.lbcap = bb.lbcap;
        ga.use_lb = (opts.prune_lb && k >= 2) ? 1u : 0u;   // level 1's distances are final only after it
        ga.tour = k >= 2 ? tour : nullptr;
        // activity bytes once a bound is known (after level 1 with the automatic bound)
        ga.act = (k >= 2 && GST_MERGE_SKIP && ub_bits != GST_INF_COST) ? B.act + (size_t)k * n : nullptr;
        ga.all_items = B.all_items; ga.n_all = n_all; ga.n_all_dev = nullptr; ga.n = n;
        const int pm = opts.pull_permille == 0 ? 450 : opts.pull_permille;
        ga.pull_thresh = pm < 0 ? 0xFFFFFFFFu : (uint32_t)std::max<uint64_t>(1, n_items_total * (uint64_t)pm / 1000);
        ga.push_lanes = opts.push_lanes > 0 ? (uint32_t)opts.push_lanes : default_push_lanes(false);
        ga.push_lanes_sparse = 0;
        uint32_t seq = 0;
        for (int it = 0; it < 100000; ++it) {
            for (int rr = 0; rr < batch; ++rr) {
                ga.seq = ++seq;
                if (flt) k_grow_round<true><<<grow_grid, GROW_BLOCK, 0, s>>>(ga);
                else k_grow_round<false><<<grow_grid, GROW_BLOCK, 0, s>>>(ga);
                grow_launches++;
            }
            HIPCHK(hipGetLastError());
            HIPCHK(hipMemcpyAsync(&H->ctl[k], B.ctl + k, sizeof(RoundCtl), hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            if (H->ctl[k].err & 1u) return gst_set_error(ctx, GST_COST_OVERFLOW, "integer path cost exceeds 32 bits");
            if (H->ctl[k].level_done) {
                total_rounds += H->ctl[k].round + 1;
                return GST_OK;
            }
        }
        return gst_set_error(ctx, GST_INTERNAL, "grow did not converge");
    };

    gst_status st = GST_OK;
    // ---- level 1 ----
    k_init_level1<<<grid_for(n, 256, fill_grid), 256, 0, s>>>(B.T + lvl_base[1], B.gm, G->d_row, n, (uint32_t)g, B.chg[0],
                                                               B.inq[0], g > 1 ? 1 : 0);
    if (g > 1) {
        k_compact<<<grid_for(n, 256, fill_grid), 256, 0, s>>>(B.inq[0], G->d_row, n, chunk, B.lists[0], &B.ctl[1].cnt[0],
                                                              B.stats, 0u);
        st = run_grow(1);
    } else {
        k_root_scan_l1<<<grid_for(n, 256, fill_grid), 256, 0, s>>>(B.T, n, B.root);
    }
    // ---- automatic bound: the cheapest spider once the group distances are final ----
    double spider_costs[2] = {0.0, 0.0};   // the spider's sum of distances, its post-processed tree
    if (st == GST_OK && opts.auto_bound && g >= 3) {
        const uint64_t* T1s = B.T + lvl_base[1];
        if (flt) k_spider_scan<true><<<grid_for(n, 256, fill_grid), 256, 0, s>>>(T1s, n, (uint32_t)g, B.spider);
        else k_spider_scan<false><<<grid_for(n, 256, fill_grid), 256, 0, s>>>(T1s, n, (uint32_t)g, B.spider);
        // keep the spider's own cost, then MST post-process its paths (lowers word 0 when cheaper;
        // skipped on the device when the caller's bound is at or below 90% of the spider's cost)
        HIPCHK(hipMemcpyAsync(B.spider + 1, B.spider, 8, hipMemcpyDeviceToDevice, s));
        if (flt) k_spider_tree<true><<<1, MST_THREADS, 0, s>>>(T1s, (uint32_t)g, B.gm, G->d_row, (const uint2*)G->d_arc, B.spider, nullptr, ub_bits);
        else k_spider_tree<false><<<1, MST_THREADS, 0, s>>>(T1s, (uint32_t)g, B.gm, G->d_row, (const uint2*)G->d_arc, B.spider, nullptr, ub_bits);
        unsigned long long spw[2] = {GST_INF_WORD, GST_INF_WORD};
        HIPCHK(hipMemcpyAsync(spw, B.spider, 16, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
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
            if (sb.ub < ub_bits) ub_bits = sb.ub;
            if (sb.gmax < bb.gmax) bb.gmax = sb.gmax;   // level 1 keeps the cap it was grown under
            bb.ub = ub_bits;
        }
    }
    // ---- tour table from the final level 1 ----
    if (st == GST_OK && opts.prune_lb && opts.tour_bound == 0 && g >= 4 && ub_bits != GST_INF_COST) {
        uint32_t* Pw = B.tour + ((size_t)1 << g) * (size_t)(g + 1);
        const size_t tl = tour_lds_bytes((uint32_t)g);
        const uint64_t* T1 = B.T + lvl_base[1];
        if (tl) {
            if (flt) k_tour_table<true, true><<<1, 256, tl, s>>>(T1, B.grp, (uint32_t)g, bb.lbcap, Pw, B.tour);
            else k_tour_table<false, true><<<1, 256, tl, s>>>(T1, B.grp, (uint32_t)g, bb.lbcap, Pw, B.tour);
        } else {
            if (flt) k_tour_table<true, false><<<1, 256, 0, s>>>(T1, B.grp, (uint32_t)g, bb.lbcap, Pw, B.tour);
            else k_tour_table<false, false><<<1, 256, 0, s>>>(T1, B.grp, (uint32_t)g, bb.lbcap, Pw, B.tour);
        }
        HIPCHK(hipGetLastError());
        tour = B.tour;
    }
    // ---- levels 2..g ----
    for (int k = 2; k <= g && st == GST_OK; ++k) {
        const bool full = (k == g);
        MergeArgs ma;
        std::memset(&ma, 0, sizeof(ma));
        ma.T = B.T; ma.pairs = B.pairs + L.pair_off[k]; ma.masks = B.masks; ma.gm = B.gm; ma.row = G->d_row;
        ma.list0 = B.lists[0]; ma.chg0 = B.chg[0]; ma.inq0 = B.inq[0]; ma.ctl = B.ctl + k; ma.root = B.root;
        ma.stats = B.stats;
        for (int j = 0; j < GST_MAX_GROUPS + 2; ++j) { ma.lvl_base[j] = lvl_base[j]; ma.m[j] = (uint32_t)L.m[j]; ma.off[j] = (uint32_t)L.off[j]; }
        ma.n = n; ma.g = (uint32_t)g; ma.k = (uint32_t)k; ma.P = (uint32_t)L.pair_cnt[k];
        ma.W = (uint32_t)((L.m[k] + 63) / 64); ma.chunk = chunk; ma.ub = ub_bits; ma.use_lb = opts.prune_lb ? 1u : 0u;
        ma.gmax = bb.gmax; ma.lbcap = bb.lbcap;
        ma.tour = tour;
        ma.act = (GST_MERGE_SKIP && ub_bits != GST_INF_COST) ? B.act : nullptr;
        ma.is_full = full ? 1u : 0u;
        const uint32_t Lk = (uint32_t)L.off[k], mk = (uint32_t)L.m[k], Wk = (mk + 63) / 64;
        const size_t pair_bytes = (size_t)L.pair_cnt[k] * 8;
        const bool pin = pair_bytes <= MERGE_PAIRS_LDS;
        const uint32_t VB = merge_tile_vertices(Lk, Wk, (size_t)opts.merge_tile_kb * 1024);
        ma.VB = VB;
        ma.pairs_in_lds = pin ? 1u : 0u;
        const size_t lds = merge_lds_bytes(VB, Lk, Wk, pair_bytes, pin);
        const int mgrid = std::max(1, std::min<int>((int)((n + VB - 1) / VB), ctx->cu_count * 8));
        if (flt) k_merge<true><<<mgrid, MERGE_THREADS, lds, s>>>(ma);
        else k_merge<false><<<mgrid, MERGE_THREADS, lds, s>>>(ma);
        merge_launches++;
        if (!full)
            k_compact<<<grid_for(n, 256, fill_grid), 256, 0, s>>>(B.inq[0], G->d_row, n, chunk, B.lists[0],
                                                                  &B.ctl[k].cnt[0], B.stats, 0u);
        HIPCHK(hipGetLastError());
        if (!full) st = run_grow(k);
    }
    if (st != GST_OK) {
        (void)hipStreamSynchronize(s);
        return st;
    }

    // ---- root, flags, traceback ----
    unsigned long long h_root = GST_INF_WORD;
    HIPCHK(hipMemcpyAsync(&h_root, B.root, 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(H->ctl, B.ctl, sizeof(RoundCtl) * NCTL, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    uint32_t all_err = 0;
    for (int k = 0; k <= g; ++k) all_err |= H->ctl[k].err;
    if (all_err & 1u) return gst_set_error(ctx, GST_COST_OVERFLOW, "integer path cost exceeds 32 bits");
    std::memset(res, 0, sizeof(*res));
    res->n_groups = g;
    res->root = -1;
    res->used_ub = has_ub ? 1 : 0;
    if (h_root == GST_INF_WORD) {
        if (has_ub)
            return gst_set_error(ctx, GST_INFEASIBLE, "no tree of cost <= bound %.9g touches all %d groups", opts.ub, g);
        return gst_set_error(ctx, GST_INFEASIBLE, "no tree touches all %d groups", g);
    }
    const uint32_t cb = (uint32_t)(h_root >> 32);
    if (flt) { float f; std::memcpy(&f, &cb, 4); res->cost = (double)f; }
    else res->cost = (double)cb;
    res->root = (int32_t)(h_root & 0xFFFFFFFFu);
    if (opts.want_tree) {
        TraceArgs ta;
        std::memset(&ta, 0, sizeof(ta));
        ta.T = B.T; ta.pos_of = B.pos_of; ta.row = G->d_row; ta.arc = (const uint2*)G->d_arc; ta.root = B.root;
        ta.flt = flt ? 1u : 0u;
        ta.out_u = B.edges; ta.out_v = B.edges + 1; ta.out_w = B.edges + 2; ta.ostride = 3;
        ta.out_n = B.edges + 3 * (size_t)ecap;
        ta.ctl = B.ctl;
        for (int j = 0; j < GST_MAX_GROUPS + 2; ++j) { ta.lvl_base[j] = lvl_base[j]; ta.m[j] = (uint32_t)L.m[j]; ta.off[j] = (uint32_t)L.off[j]; }
        ta.full = (uint32_t)L.full; ta.cap = ecap; ta.max_iter = n + 4u * (uint32_t)g + 64u;
        k_traceback<<<1, WAVE, 0, s>>>(ta);
        HIPCHK(hipGetLastError());
        uint32_t nout = 0;
        HIPCHK(hipMemcpyAsync(&nout, B.edges + 3 * (size_t)ecap, 4, hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(H->ctl, B.ctl, sizeof(RoundCtl), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        if (nout > ecap)   // the traceback stops writing at ecap edges and reports more
            return gst_set_error(ctx, ecap == GST_NOALLOC_MAX_EDGES ? GST_INVALID_ARG : GST_INTERNAL,
                                 "no-allocation solve: the traced tree has more than %u edges (the limit is %u)", ecap,
                                 (unsigned)GST_NOALLOC_MAX_EDGES);
        if (H->ctl[0].err & 2u)
            return gst_set_error(ctx, GST_INTERNAL, "device traceback failed (edges=%u)", nout);
        if (nout) {
            HIPCHK(hipMemcpyAsync(H->tri, B.edges, 12ull * nout, hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
        }
        const gst_status fs = finish_tree_into(ctx, H, flt, nout, edge_cap, eu, ev, ew, res);
        if (fs != GST_OK) return fs;
    }
    // ---- stats (no HIP events here: creating them is the runtime's business, not this call's) ----
    HIPCHK(hipMemcpyAsync(H->stats, B.stats, sizeof(H->stats), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    unsigned long long hs[ST_NSTAT];
    stat_sum(H->stats, hs);
    gst_stats& S = res->stats;
    S.grow_items = hs[ST_ITEMS]; S.grow_arcs = hs[ST_ARCS]; S.grow_relax = hs[ST_RELAX];
    S.grow_updates = hs[ST_UPD]; S.grow_enqueued = hs[ST_ENQ]; S.grow_new_states = hs[ST_NEW]; S.grow_atomics = hs[ST_ATOM];
    S.merge_vertices = hs[ST_MVERT]; S.merge_pairs = hs[ST_MPAIRS];
    S.spider_cost = spider_costs[0]; S.spider_tree_cost = spider_costs[1];
    S.grow_rounds = total_rounds; S.grow_launches = grow_launches; S.merge_launches = merge_launches;
    S.grow_bytes = gst_grow_bytes(S);
    S.merge_bytes = hs[ST_MBYTES];
    return GST_OK;
}
