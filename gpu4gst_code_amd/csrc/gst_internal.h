// Internal declarations shared by the host driver (gst_api.cpp) and the HIP kernels
// (gst_dpbf.hip).  Not part of the public ABI.
#pragma once

#include <stdint.h>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gst.h"

// ---------------------------------------------------------------------------------------
// Packed DP state word: high 32 bits = cost, low 32 bits = predecessor.
//   cost  : uint32 for integer weights, IEEE-754 bits of a non-negative float for float32
//           weights (non-negative float bit patterns order like the floats); 0xFFFFFFFF = inf.
//   pred  : < 0x80000000  -> grow: the neighbour x whose state (x, X) grew into this one over the
//                            arc (u, x); the tree edge is (u, x)
//           0x80000000     -> leaf: v itself covers the mask
//           0x80000000|X1  -> merge of (v,X1) and (v,X^X1), X1 holds the lowest bit of X
// One 64-bit atomic-min keeps cost and predecessor consistent (the reference updates cost by
// atomicMin and the predecessor by plain stores, src/DPQ.cu:267-272) and breaks cost ties by
// the smaller predecessor.
// ---------------------------------------------------------------------------------------
#define GST_INF_COST   0xFFFFFFFFu
#define GST_INF_WORD   0xFFFFFFFFFFFFFFFFull
#define GST_PRED_LEAF  0x80000000u
#define GST_PRED_MERGE 0x80000000u

struct gst_levels {
    int g = 0;
    int full = 0;                       // 2^g - 1
    int m[GST_MAX_GROUPS + 2] = {0};    // masks per level (popcount k), k = 1..g
    int off[GST_MAX_GROUPS + 2] = {0};  // per-vertex row offset of level k: sum_{j<k} m_j
    std::vector<uint32_t> masks;        // masks ordered by (popcount, value); row position = index
    std::vector<uint32_t> pos_of;       // mask -> row position (index into masks)
    std::vector<uint32_t> rank_of;      // mask -> rank within its level
    // merge pairs per level k (k >= 2): packed (dst rank:16 | pos1:16 | pos2:16 | X1:16)
    std::vector<uint64_t> pairs;
    int pair_off[GST_MAX_GROUPS + 2] = {0};
    int pair_cnt[GST_MAX_GROUPS + 2] = {0};
    void build(int g);
};

struct gst_graph {
    gst_ctx* ctx = nullptr;
    uint64_t uid = 0;            // unique per upload: caches key on it, never on the address
    uint64_t budget_key = 0;     // key of the learned round budgets (0: uid)
    int32_t n = 0;
    int64_t m = 0;
    int32_t w_type = 0;
    uint32_t* d_row = nullptr;   // n+1 arc offsets (arcs < 2^31)
    int32_t*  d_col = nullptr;   // m
    uint32_t* d_w = nullptr;     // m (int32 or float bits)
    uint64_t* d_arc = nullptr;   // m arc records, (weight bits << 32) | neighbour: the grow and the
                                 // traceback read both in one 8-byte load
    // host copy for validation / traceback weight lookups
    std::vector<int64_t> h_row;
    std::vector<int32_t> h_col;
    std::vector<uint32_t> h_w;
    int32_t max_deg = 0;
};

// Algorithmic bytes of the grow kernels, from their counters: a work item reads its id, its row
// bounds and its change word (24 B) and was written once when enqueued (8 B); a scanned arc is one
// 8-byte record (neighbour, weight); a relaxation reads the neighbour's 8-byte state; an update
// writes that state and ORs the neighbour's change word (16 B).  The own costs an item reads
// (8 B per changed mask) are not counted, so the figure is a lower bound of the kernel's needs.
inline uint64_t gst_grow_bytes(const gst_stats& s) {
    return s.grow_items * 24ull + s.grow_enqueued * 8ull + s.grow_arcs * 8ull + s.grow_relax * 8ull +
           s.grow_updates * 16ull;
}

// Rows of an uploaded graph are ordered by (weight bits, neighbour); weights are non-negative
// int32 or non-negative float32 (never -0.0f), so the bit patterns order like the weights.
inline uint64_t arc_key(uint32_t w, int32_t col) { return ((uint64_t)w << 32) | (uint32_t)col; }

// index of arc u->v of weight w in a (weight, neighbour)-ordered row, -1 when absent
inline int64_t gst_find_arc(const int64_t* row, const int32_t* col, const uint32_t* w, int32_t u, int32_t v,
                            uint32_t wv) {
    int64_t lo = row[u], hi = row[u + 1];
    const uint64_t key = arc_key(wv, v);
    while (lo < hi) {
        const int64_t mid = (lo + hi) / 2;
        if (arc_key(w[mid], col[mid]) < key) lo = mid + 1; else hi = mid;
    }
    return (lo < row[u + 1] && col[lo] == v && w[lo] == wv) ? lo : -1;
}

// Cost bits of a bound U for the kernels:
//   ub    : states of cost >= ub are dropped (ints: floor(U)+1, so cost <= U stays; floats: U with a
//           relative slack of 1e-4, because the DP sums fp32 costs in its own order and rounding must
//           never drop a structure whose exact cost is <= U)
//   gmax  : with half_bound == 0, only states of cost <= U/2 are grown (ints: floor(floor(U)/2);
//           floats: half the slackened bound); otherwise every finite state below ub
//   lbcap : cap on level-1 distances in the lower bound: gmax when level 1 is cut at gmax, else inf
struct gst_bound_bits {
    uint32_t ub = 0xFFFFFFFFu, gmax = 0xFFFFFFFEu, lbcap = 0xFFFFFFFFu;
};
// The grow cap (gmax bits) the device derives from a spider cost sp (spider_bound in the device code,
// the same float operations: single precision, round to nearest).
inline uint32_t gst_spider_gmax_bits(uint32_t sp, bool flt) {
    if (!flt) return sp / 2u;
    float f;
    std::memcpy(&f, &sp, 4);
    const float u = f * 1.0001f;
    uint32_t ub;
    std::memcpy(&ub, &u, 4);
    ub += 1u;
    float uf;
    std::memcpy(&uf, &ub, 4);
    const float h = uf * 0.5f;
    uint32_t gm;
    std::memcpy(&gm, &h, 4);
    return gm + 1u;
}

inline gst_bound_bits gst_make_bound(double U, bool flt, int32_t half_bound) {
    gst_bound_bits b;
    if (!(std::isfinite(U) && U > 0)) return b;
    if (flt) {
        float fu = (float)(U * (1.0 + 1e-4));
        fu = std::nextafter(fu, INFINITY);
        std::memcpy(&b.ub, &fu, 4);
        if (half_bound == 0) {
            float fh = std::nextafter(fu * 0.5f, INFINITY);
            std::memcpy(&b.gmax, &fh, 4);
        }
    } else {
        const double c = std::floor(U) + 1.0;
        b.ub = c >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)c;
        if (half_bound == 0 && b.ub != 0xFFFFFFFFu) b.gmax = (b.ub - 1u) / 2u;
    }
    if (half_bound == 0 && b.gmax != 0xFFFFFFFEu) b.lbcap = b.gmax;
    return b;
}

// 1 selects the item-parallel push round (one item per lane); 8..64 lanes per item otherwise
inline bool gst_push_lanes_ok(int32_t p) { return p == 0 || p == 1 || p == 8 || p == 16 || p == 32 || p == 64; }

struct gst_query_bufs;  // device buffers of one query (defined in gst_dpbf.hip)

// knobs chosen by gst_autotune for one graph (and group count / bound presence), applied to
// solves whose options leave them at 0
struct gst_tuned {
    uint64_t uid = 0;
    int g = 0;
    bool ub = false;
    int32_t chunk_arcs = 0, pull_permille = 0, push_lanes = 0, merge_tile_kb = 0;
};
struct gst_tuned_batch {
    uint64_t uid = 0;
    int32_t rounds_per_step = 0, push_lanes = 0, pull_permille = 0;
};

struct gst_ctx {
    std::vector<gst_tuned> tuned;
    std::vector<gst_tuned_batch> tuned_batch;
    int device = 0;
    std::string last_error;
    std::mutex err_mu;          // last_error is set from the learned batch's selection thread too
    void* stream = nullptr;     // hipStream_t
    void* stream2 = nullptr;    // hipStream_t (default priority): learned-bound selection beside the engine
    void* stream3 = nullptr;    // hipStream_t (default priority): a second selection stream
    gst_query_bufs* qb = nullptr;
    struct gst_batch_engine* be = nullptr;   // batched engine (gst_batch.hip), made on first use
    struct gst_noalloc_host* na = nullptr;   // no-allocation solve (gst_noalloc.hip): made at set-up
    gst_levels lv_cache[GST_MAX_GROUPS + 1];
    bool lv_ready[GST_MAX_GROUPS + 1] = {false};
    int cu_count = 0;
};

// implemented in gst_dpbf.hip
gst_status gst_dpbf_solve(gst_ctx* ctx, const gst_graph* g, const gst_groups* groups,
                          const gst_opts* opts, gst_result* res);
void gst_query_bufs_free(gst_ctx* ctx);
gst_status gst_dpbf_mst_postprocess(gst_ctx* ctx, const gst_graph* G, const int32_t* vertices, int32_t nv, int32_t* eu,
                                    int32_t* ev, double* ew, int32_t* ne, double* cost);
// tree of a traced edge list (duplicates removed, spanning forest of the union) into res
void gst_finish_tree(bool flt, uint32_t nout, const uint32_t* hu, const uint32_t* hv, const uint32_t* hw,
                     gst_result* res);
gst_status gst_ctx_device_init(gst_ctx* ctx);
// implemented in gst_noalloc.hip: set-up at context creation, the allocation-free one-query solve
gst_status gst_noalloc_init(gst_ctx* ctx);
void gst_noalloc_fini(gst_ctx* ctx);
gst_status gst_dpbf_solve_noalloc(gst_ctx* ctx, const gst_graph* g, const gst_groups* groups, const gst_opts* opts,
                                  int32_t edge_cap, int32_t* eu, int32_t* ev, double* ew, gst_result* res);
void gst_ctx_device_fini(gst_ctx* ctx);
gst_status gst_graph_device_upload(gst_ctx* ctx, gst_graph* g);
void gst_graph_device_free(gst_graph* g);

gst_status gst_set_error(gst_ctx* ctx, gst_status s, const char* fmt, ...);
// fill the knobs an opts leaves at 0 from the autotuner's choice for (graph, g, bound present)
inline void gst_apply_tuned(const gst_ctx* ctx, const gst_graph* G, int g, bool has_ub, gst_opts& o) {
    for (const gst_tuned& t : ctx->tuned) {
        if (t.uid != G->uid || t.g != g || t.ub != has_ub) continue;
        if (o.chunk_arcs == 0) o.chunk_arcs = t.chunk_arcs;
        if (o.pull_permille == 0) o.pull_permille = t.pull_permille;
        if (o.push_lanes == 0) o.push_lanes = t.push_lanes;
        if (o.merge_tile_kb == 0) o.merge_tile_kb = t.merge_tile_kb;
        return;
    }
}
inline const gst_tuned_batch* gst_find_tuned_batch(const gst_ctx* ctx, uint64_t uid) {
    for (const gst_tuned_batch& t : ctx->tuned_batch)
        if (t.uid == uid) return &t;
    return nullptr;
}
uint64_t gst_next_graph_uid();
const gst_levels& gst_get_levels(gst_ctx* ctx, int g);
