// Batched DPBF engine: many exact queries in flight on one GPU.
//
// The one-query driver (gst_dpbf.hip) is bound by round latency on sparse frontiers: a grow round
// over a few thousand items still costs a launch and a chain of dependent memory accesses.  Here up
// to MAX_SLOTS queries ("slots") hold their own state tables, frontiers and round-control blocks,
// and every launch serves all of them:
//   k_plan_step   one wave: per slot, a finished grow level becomes a merge of the next level
//                 (or of the full mask), a merged full mask becomes a traceback; lists the work
//   k_batch_merge the merges of every listed slot, the grid split evenly among them
//   k_batch_compact the new frontier of every slot merged below the full mask
//   k_batch_trace one wave per slot whose full mask was merged one step earlier
//   k_plan_grow   one wave: per growing slot, the workgroups its round needs (as the one-query
//                 round sizes itself), scaled down to the grid when the sum exceeds it
//   k_batch_grow  one grow round of every growing slot, each slot on its own slice of the grid
// A step is k_plan_step, the three batch kernels, then rounds_per_step (k_plan_grow, k_batch_grow)
// pairs; it is captured once as a HIP graph and replayed.  The host reads the slot phases after
// each step, collects finished queries and refills their slots.  All kernel bodies are the
// one-query bodies (gst_dpbf_dev.h) run on a slot's slice of the grid, so each query computes
// exactly what gst_solve computes for it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "gst_batch.h"
#include "gst_dpbf_dev.h"

#define HIPCHK(call)                                                                         \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return gst_set_error(ctx, GST_HIP_ERROR, "%s failed: %s (%s:%d)", #call,         \
                                 hipGetErrorString(e_), __FILE__, __LINE__);                 \
    } while (0)

namespace {

constexpr uint32_t MAX_SLOTS = 256;
// 1: a new query's set-up (g >= 2) runs batched in the next step's first part; 0: at once, per query
#ifndef GST_BATCH_DEFER_FILL
#define GST_BATCH_DEFER_FILL 1
#endif
// workgroups a batched grow round may use, per CU.  The grid exceeds what is resident at once: each
// slot's round gets the workgroups its work needs (at most GROW_MAX_PART), and the dispatcher hands
// the CUs that slots with short rounds free to the workgroups still waiting, so a slot with a long
// round no longer runs on a scaled-down share while the others idle.  Measured on BA V=200k g=6,
// 128 slots: x5 / 8 / 12 / 20 give the spider-bounded stream 821 / 904 / 991 / 1076 queries/s and the
// learned-bound stream 1526 / 1536 / 1525 / 1472.
#ifndef GST_BATCH_GRID_MULT
#define GST_BATCH_GRID_MULT 12
#endif
constexpr uint32_t GROW_MAX_PART = 64 * 32;   // 64 arrival groups of 32 workgroups (RoundCtl::sub)
constexpr uint32_t PLAN_THREADS = MAX_SLOTS;          // the plan kernels run one thread per slot
constexpr uint32_t PLAN_WAVES = PLAN_THREADS / WAVE;
enum : uint32_t { PH_EMPTY = 0, PH_GROW = 1, PH_TRACE = 2, PH_DONE = 3, PH_FILL = 4 };

struct SlotDev {
    uint32_t phase;
    uint32_t k;     // level being grown (PH_GROW)
    uint32_t g;
    uint32_t pad;
};

struct BatchPlan {
    uint32_t seq;                          // grow launch number (the round-control seq)
    uint32_t n_grow, n_merge, n_compact, n_trace;
    uint32_t grow_slot[MAX_SLOTS];
    uint32_t grow_level[MAX_SLOTS];
    uint32_t grow_start[MAX_SLOTS + 1];    // first workgroup of each growing slot; [n_grow] = total
    uint32_t merge_slot[MAX_SLOTS];
    uint32_t merge_level[MAX_SLOTS];
    uint32_t compact_slot[MAX_SLOTS];
    uint32_t compact_level[MAX_SLOTS];
    uint32_t trace_slot[MAX_SLOTS];
    uint32_t n_fill;
    uint32_t fill_slot[MAX_SLOTS];         // slots whose new query is set up by this step's part A
    SlotDev sd_out[MAX_SLOTS];             // the slots' phases after the step's plan (read by the host)
};

// set-up of a slot's new query (g >= 2), run by the next step's part A: group bits, level 1 and
// its frontier, the per-job counters cleared
struct FillArgs {
    uint32_t* gm;
    unsigned long long* root;
    unsigned long long* stats;
    unsigned long long* spider;
    uint32_t* ctl;            // the slot's round-control blocks, as words
    uint32_t* hub_cnt;
    const int32_t* grp;       // grp_ptr[0..g] then the members
    uint64_t* T1;
    const uint32_t* row;
    uint64_t* chg0;
    uint32_t* inq0;
    uint64_t* list0;
    uint32_t* cnt0;           // level 1's first frontier count
    uint64_t* hub_items;      // a job's own graph: its hub row items are listed too (nullptr: none)
    uint32_t n, g, chunk, pad;
};

// a slot's argument blocks, contiguous so that filling a slot is one copy
struct SlotBlock {
    GrowArgs g[NCTL];      // by level
    MergeArgs m[NCTL];     // by level
    TraceArgs t;
    FillArgs f;
    SlotDev sd;
};

struct BatchArgs {
    SlotBlock* sb;            // [slot]
    BatchPlan* plan;
    uint32_t S;               // slots (<= MAX_SLOTS)
    uint32_t grid;            // workgroups a grow round may use in total
};

// exclusive prefix sum of x over the plan workgroup (one thread per slot); tot receives the sum
__device__ __forceinline__ uint32_t plan_scan(uint32_t x, uint32_t& tot) {
    __shared__ uint32_t wsum[PLAN_WAVES];
    uint32_t incl = x;
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        const uint32_t y = __shfl_up(incl, d, WAVE);
        if (lane_id() >= (uint32_t)d) incl += y;
    }
    const uint32_t w = threadIdx.x / WAVE;
    if (lane_id() == WAVE - 1) wsum[w] = incl;
    __syncthreads();
    uint32_t off = 0;
    tot = 0;
    for (uint32_t i = 0; i < PLAN_WAVES; ++i) {
        if (i < w) off += wsum[i];
        tot += wsum[i];
    }
    __syncthreads();   // wsum is reused by the next call
    return off + incl - x;
}

// a slot's per-job state before a fill: group bits cleared, root = inf, statistics, spider bound,
// control blocks and hub item count reset (one launch instead of six)
__global__ void k_slot_reset(uint32_t* gm, uint32_t n, unsigned long long* root, unsigned long long* stats,
                             unsigned long long* spider, uint32_t* ctl, uint32_t* hub_cnt) {
    const uint32_t tid = blockIdx.x * blockDim.x + threadIdx.x;
    for (uint32_t v = tid; v < n; v += gridDim.x * blockDim.x) gm[v] = 0u;
    if (blockIdx.x == 0) {
        for (uint32_t i = threadIdx.x; i < (uint32_t)(sizeof(RoundCtl) * NCTL / 4); i += blockDim.x) ctl[i] = 0u;
        for (uint32_t i = threadIdx.x; i < (uint32_t)STAT_WORDS; i += blockDim.x) stats[i] = 0ull;
        if (threadIdx.x == 0) {
            *root = GST_INF_WORD;
            *spider = GST_INF_WORD;
            *hub_cnt = 0u;
        }
    }
}

// The slots set up by a step share the launch's workgroups: slot i of the fill list takes
// workgroups [i * per, (i + 1) * per) (per >= 1 since the grid has at least MAX_SLOTS workgroups).
__device__ __forceinline__ const FillArgs* fill_share(const BatchArgs& b, uint32_t& bid, uint32_t& nblk) {
    const BatchPlan* p = b.plan;
    const uint32_t nf = p->n_fill;
    if (nf == 0) return nullptr;
    const uint32_t per = max(1u, gridDim.x / nf);
    const uint32_t i = blockIdx.x / per;
    if (i >= nf) return nullptr;
    bid = blockIdx.x - i * per;
    nblk = per;
    return &b.sb[p->fill_slot[i]].f;
}

// fill, first pass: group bits cleared, root = inf, statistics, spider bound, control blocks and hub
// item count reset
__global__ void __launch_bounds__(256) k_batch_fill_reset(BatchArgs b) {
    uint32_t bid, nblk;
    const FillArgs* f = fill_share(b, bid, nblk);
    if (!f) return;
    for (uint32_t v = bid * blockDim.x + threadIdx.x; v < f->n; v += nblk * blockDim.x) f->gm[v] = 0u;
    if (bid == 0) {
        for (uint32_t i = threadIdx.x; i < (uint32_t)(sizeof(RoundCtl) * NCTL / 4); i += blockDim.x) f->ctl[i] = 0u;
        for (uint32_t i = threadIdx.x; i < (uint32_t)STAT_WORDS; i += blockDim.x) f->stats[i] = 0ull;
        if (threadIdx.x == 0) {
            *f->root = GST_INF_WORD;
            *f->spider = GST_INF_WORD;
            *f->hub_cnt = 0u;
        }
    }
}

// fill, second pass: the members' group bits
__global__ void __launch_bounds__(256) k_batch_fill_bits(BatchArgs b) {
    uint32_t bid, nblk;
    const FillArgs* f = fill_share(b, bid, nblk);
    if (!f) return;
    for (uint32_t i = 0; i < f->g; ++i) {
        const int32_t s0 = f->grp[i], e = f->grp[i + 1];
        for (int32_t j = s0 + (int32_t)(bid * blockDim.x + threadIdx.x); j < e; j += (int32_t)(nblk * blockDim.x))
            atomicOr(&f->gm[f->grp[f->g + 1 + j]], 1u << i);
    }
}

// fill, third pass: level 1 (T(v,{i}) = 0 for members of group i, inf elsewhere), members with arcs
// flagged as the first frontier
__global__ void __launch_bounds__(256) k_batch_fill_level1(BatchArgs b) {
    uint32_t bid, nblk;
    const FillArgs* f = fill_share(b, bid, nblk);
    if (!f) return;
    const uint32_t g = f->g;
    for (uint32_t v = bid * blockDim.x + threadIdx.x; v < f->n; v += nblk * blockDim.x) {
        const uint32_t gmv = f->gm[v];
        for (uint32_t i = 0; i < g; ++i)
            f->T1[(uint64_t)v * g + i] = ((gmv >> i) & 1u) ? (uint64_t)GST_PRED_LEAF : GST_INF_WORD;
        if (gmv && f->row[v + 1] > f->row[v]) { f->chg0[v] = gmv; f->inq0[v] = 1u; }
    }
}

// fill, last pass: level 1's work-item list, and a job's own graph's hub row items
__global__ void __launch_bounds__(256) k_batch_fill_compact(BatchArgs b) {
    uint32_t bid, nblk;
    const FillArgs* f = fill_share(b, bid, nblk);
    if (!f) return;
    compact_body(f->inq0, f->row, f->n, f->chunk, f->list0, f->cnt0, f->stats, 0u, bid, nblk);
    if (f->hub_items)
        compact_body(nullptr, f->row, f->n, f->chunk, f->hub_items, f->hub_cnt, nullptr, HUB_DEG, bid, nblk);
}

__global__ void k_reset_phases(SlotBlock* sb) {
    if (threadIdx.x < MAX_SLOTS) sb[threadIdx.x].sd = SlotDev{PH_EMPTY, 0u, 0u, 0u};
}

// a slot whose level 1 (grown under its running spider) has ended: the spider's bound bounds the later
// levels, and their lower bounds read level 1 capped at its half (every state at or below it is
// exact: the last round's cap was at least that, and the caps before it were higher)
template <bool FLT>
__device__ __forceinline__ void level1_bound_to_later(SlotBlock& sb, uint32_t g) {
    const unsigned long long sw = *(volatile unsigned long long*)sb.g[1].run_sp;
    if ((uint32_t)(sw >> 32) == GST_INF_COST) return;
    uint32_t ub = GST_INF_COST, gmax = GST_INF_COST;
    running_spider_bound<FLT>(sw, ub, gmax);
    for (uint32_t k = 2; k <= g; ++k) {
        if (k < g) {
            GrowArgs* ga = &sb.g[k];
            ga->ub = min(ga->ub, ub);
            ga->gmax = min(ga->gmax, gmax);
            ga->lbcap = min(ga->lbcap, gmax);
        }
        MergeArgs* ma = &sb.m[k];
        ma->ub = min(ma->ub, ub);
        ma->gmax = min(ma->gmax, gmax);
        ma->lbcap = min(ma->lbcap, gmax);
    }
}

template <bool FLT>
__global__ void __launch_bounds__(PLAN_THREADS) k_plan_step(BatchArgs b) {
    const uint32_t s = threadIdx.x;
    bool dm = false, dc = false, dt = false, df = false;
    uint32_t lvl = 0;
    if (s < b.S) {
        SlotDev* sd = &b.sb[s].sd;
        const uint32_t ph = sd->phase;
        if (ph == PH_FILL) {
            df = true;
            sd->phase = PH_GROW;          // set up by this part A, grown from part B on
        } else if (ph == PH_TRACE) {
            dt = true;
            sd->phase = PH_DONE;
        } else if (ph == PH_GROW && b.sb[s].g[sd->k].ctl->level_done) {
            if (sd->k == 1 && b.sb[s].g[1].run_sp) level1_bound_to_later<FLT>(b.sb[s], sd->g);
            lvl = sd->k + 1;
            dm = true;
            if (lvl == sd->g) {
                sd->phase = PH_TRACE;     // traced in the next step, after this merge found the root
            } else {
                dc = true;                // the merge flags the new frontier; compaction lists it
                sd->k = lvl;
            }
        }
    }
    BatchPlan* p = b.plan;
    uint32_t nm, nc, nt, nf;
    const uint32_t im = plan_scan(dm ? 1u : 0u, nm);
    const uint32_t ic = plan_scan(dc ? 1u : 0u, nc);
    const uint32_t it = plan_scan(dt ? 1u : 0u, nt);
    const uint32_t jf = plan_scan(df ? 1u : 0u, nf);
    if (dm) { p->merge_slot[im] = s; p->merge_level[im] = lvl; }
    if (dc) { p->compact_slot[ic] = s; p->compact_level[ic] = lvl; }
    if (dt) p->trace_slot[it] = s;
    if (df) p->fill_slot[jf] = s;
    if (s < b.S) p->sd_out[s] = b.sb[s].sd;
    if (s == 0) {
        p->n_merge = nm;
        p->n_compact = nc;
        p->n_trace = nt;
        p->n_fill = nf;
    }
}

// Merges of every slot whose level finished, and the tour tables of the slots entering level 2.
// The first MAX_SLOTS workgroups build the tables (one per slot, from its final level 1): the grow
// rounds are their first readers (the merge applies the one-label bound only), so the table builds,
// short chains of dependent phases, run beside the merges instead of before them.  The other
// workgroups (at least MAX_SLOTS) split among the merges.
template <bool FLT, bool PW_LDS>
__global__ void __launch_bounds__(MERGE_THREADS) __attribute__((amdgpu_waves_per_eu(GST_MERGE_WPE))) k_batch_merge(BatchArgs b) {
    extern __shared__ unsigned long long lds[];
    const BatchPlan* p = b.plan;
    const uint32_t nm = p->n_merge;
    if (nm == 0) return;
    if (blockIdx.x < MAX_SLOTS) {
        const uint32_t i = blockIdx.x;
        if (i >= nm || p->merge_level[i] != 2) return;
        const MergeArgs& m2 = b.sb[p->merge_slot[i]].m[2];
        if (!m2.tour) return;
        tour_table_body<FLT, PW_LDS>(m2.T + m2.lvl_base[1], m2.grp, m2.grp + m2.g + 1, m2.g, m2.lbcap, m2.tour_w,
                                     const_cast<uint32_t*>(m2.tour));
        // a table switched off (walks too short to prune) leaves the slot's grow rounds: they then
        // read no table at all
        __syncthreads();
        if (threadIdx.x == 0 && m2.tour[0] == 0u) {
            SlotBlock& sb = b.sb[p->merge_slot[i]];
            for (uint32_t k = 2; k < m2.g; ++k) sb.g[k].tour = nullptr;
        }
        return;
    }
    const uint32_t bid = blockIdx.x - MAX_SLOTS, nblk = gridDim.x - MAX_SLOTS;
    const uint32_t per = max(1u, nblk / nm);   // nblk >= MAX_SLOTS, so per >= 1 covers all
    const uint32_t i = bid / per;
    if (i >= nm) return;
    const MergeArgs& ma = b.sb[p->merge_slot[i]].m[p->merge_level[i]];
    if (merge_uses_list(ma)) merge_body<FLT, true>(ma, lds, bid - i * per, per);
    else merge_body<FLT, false>(ma, lds, bid - i * per, per);
}

// before the merges: the pre-pass of every merge that walks its live list (dead vertices written,
// live ones listed), the grid split evenly among the listed merges
__global__ void __launch_bounds__(256) k_batch_merge_live(BatchArgs b) {
    const BatchPlan* p = b.plan;
    const uint32_t nm = p->n_merge;
    if (nm == 0) return;
    const uint32_t per = max(1u, gridDim.x / nm);
    const uint32_t i = blockIdx.x / per;
    if (i >= nm) return;
    const MergeArgs& ma = b.sb[p->merge_slot[i]].m[p->merge_level[i]];
    if (merge_uses_list(ma)) merge_live_body(ma, blockIdx.x - i * per, per);
}

// after the merges: one workgroup per slot whose level-2 merge found a spider (automatic bound)
// post-processes the spider's paths into a tree (MST of their vertices, unneeded leaves removed) and
// lowers the spider's cost to the tree's when cheaper
template <bool FLT>
__global__ void __launch_bounds__(MST_THREADS) k_batch_spider_tree(BatchArgs b) {
    const BatchPlan* p = b.plan;
    const uint32_t i = blockIdx.x;
    if (i >= p->n_merge || p->merge_level[i] != 2) return;
    const SlotBlock& sb = b.sb[p->merge_slot[i]];
    const MergeArgs& m2 = sb.m[2];
    if (!m2.spider) return;
    // a job whose own bound (the learned one, typically within a few per cent of the optimum) is
    // already well below the spider needs no tree (its walk costs ~50 us of the step)
    if (spider_tree_useless<FLT>(m2.ub, (uint32_t)(*m2.spider >> 32))) return;
    __shared__ MstLds L;
    spider_tree_body<FLT>(L, m2.T + m2.lvl_base[1], m2.g, m2.gm, sb.g[1].row, sb.g[1].arc, m2.spider, nullptr);
}

// after the merges: a slot whose level-2 merge found a spider (automatic bound) tightens the bound
// of its later levels in the argument blocks (read by every later launch)
template <bool FLT>
__global__ void __launch_bounds__(PLAN_THREADS) k_batch_bound(BatchArgs b) {
    const BatchPlan* p = b.plan;
    const uint32_t i = threadIdx.x;
    if (i >= p->n_merge || p->merge_level[i] != 2) return;
    const uint32_t s = p->merge_slot[i];
    const MergeArgs& m2 = b.sb[s].m[2];
    if (!m2.spider) return;
    const uint32_t sp = (uint32_t)(*m2.spider >> 32);
    if (sp == GST_INF_COST) return;
    uint32_t ub, gmax;
    spider_bound<FLT>(sp, ub, gmax);
    const uint32_t g = m2.g;
    for (uint32_t k = 2; k <= g; ++k) {
        if (k < g) {
            GrowArgs* ga = &b.sb[s].g[k];
            if (ub < ga->ub) ga->ub = ub;
            if (gmax < ga->gmax) ga->gmax = gmax;
        }
        if (k >= 3) {
            MergeArgs* ma = &b.sb[s].m[k];
            if (ub < ma->ub) ma->ub = ub;
            if (gmax < ma->gmax) ma->gmax = gmax;
        }
    }
}

__global__ void __launch_bounds__(256) k_batch_compact(BatchArgs b) {
    const BatchPlan* p = b.plan;
    const uint32_t nc = p->n_compact;
    if (nc == 0) return;
    const uint32_t per = max(1u, gridDim.x / nc);
    const uint32_t i = blockIdx.x / per;
    if (i >= nc) return;
    const GrowArgs& a = b.sb[p->compact_slot[i]].g[p->compact_level[i]];
    compact_body(a.inq[0], a.row, a.n, a.chunk, a.list[0], &a.ctl->cnt[0], a.stats, 0u, blockIdx.x - i * per, per);
}

__global__ void __launch_bounds__(WAVE) k_batch_trace(BatchArgs b) {
    const BatchPlan* p = b.plan;
    if (blockIdx.x >= p->n_trace) return;
    trace_body(b.sb[p->trace_slot[blockIdx.x]].t);
}

__global__ void __launch_bounds__(PLAN_THREADS) k_plan_grow(BatchArgs b) {
    const uint32_t s = threadIdx.x;
    uint32_t need = 0, lvl = 0;
    if (s < b.S) {
        const SlotDev sd = b.sb[s].sd;
        if (sd.phase == PH_GROW) {
            const GrowArgs& a = b.sb[s].g[sd.k];
            const RoundCtl* c = a.ctl;
            if (!c->level_done) {
                const uint32_t r = c->round;
                const uint32_t n_cur = c->cnt[r % 3], n_prv = c->cnt[(r + 2) % 3];
                const bool pull = n_cur > 0 && n_cur >= a.pull_thresh && a.all_items != nullptr;
                // workgroups in proportion to the round's work, so that when the grid is shared out
                // every slot's round takes about as long: a pull round sweeps every light row
                // (64/min(m,64) rows per wave) and every hub item (one per wave); a push round
                // serves 64/push_lanes items per wave; the list of the previous round is cleared
                // with one thread per item
                const uint32_t mch = a.m < (uint32_t)WAVE ? a.m : (uint32_t)WAVE;
                const uint32_t rows_per_wave = (uint32_t)WAVE / mch;
                const uint32_t items_per_wave = push_items_per_wave(round_push_lanes(a.push_lanes, a.push_lanes_sparse, n_cur));
                const uint32_t n_all = a.n_all_dev ? *a.n_all_dev : a.n_all;
                const uint32_t waves = pull ? (a.n + rows_per_wave - 1) / rows_per_wave + n_all
                                            : (n_cur + items_per_wave - 1) / items_per_wave;
                need = max(1u, max((waves + GROW_WAVES - 1) / GROW_WAVES, (n_prv + GROW_BLOCK - 1) / GROW_BLOCK));
                lvl = sd.k;
            }
        }
    }
    uint32_t total, n_act, all;
    plan_scan(min(need, 1u << 24), total);
    // a slot's round takes at most GROW_MAX_PART workgroups (its arrival counters count 64 groups of 32)
    uint32_t alloc = min(need, min(b.grid, GROW_MAX_PART));
    if (total > b.grid && need) alloc = min(GROW_MAX_PART, max(1u, (uint32_t)((uint64_t)min(need, 1u << 24) * b.grid / total)));
    const uint32_t start = plan_scan(alloc, all);
    const uint32_t rank = plan_scan(need != 0 ? 1u : 0u, n_act);
    BatchPlan* p = b.plan;
    if (need) {
        p->grow_slot[rank] = s;
        p->grow_level[rank] = lvl;
        p->grow_start[rank] = start;
    }
    if (s == 0) {
        p->n_grow = n_act;
        p->grow_start[n_act] = all;
        p->seq = p->seq + 1;
    }
}

// one grow round of every growing slot; the grid is b.grid + MAX_SLOTS workgroups (the plan's
// per-slot minimum of one workgroup can exceed b.grid by at most one per slot)
template <bool FLT>
__global__ void __launch_bounds__(GROW_BLOCK) __attribute__((amdgpu_waves_per_eu(GST_GROW_WPE))) k_batch_grow(BatchArgs b) {
    const BatchPlan* p = b.plan;
    const uint32_t ng = p->n_grow;
    const uint32_t bid = blockIdx.x;
    if (ng == 0 || bid >= p->grow_start[ng]) return;
    uint32_t lo = 0, hi = ng;   // grow_start[lo] <= bid < grow_start[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (p->grow_start[mid] <= bid) lo = mid; else hi = mid;
    }
    const uint32_t st = p->grow_start[lo];
    grow_round_body<FLT>(b.sb[p->grow_slot[lo]].g[p->grow_level[lo]], bid - st, p->grow_start[lo + 1] - st,
                         p->seq);
}

template <class T>
bool dalloc(T*& p, size_t count) {
    p = nullptr;
    return hipMalloc((void**)&p, std::max<size_t>(count, 1) * sizeof(T)) == hipSuccess;
}

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

// ---------------------------------------------------------------------------------------
// engine state (one per context)
// ---------------------------------------------------------------------------------------
struct gst_batch_engine {
    struct Slot {
        uint64_t* T = nullptr;
        uint32_t* gm = nullptr;
        uint64_t* chg[3] = {nullptr, nullptr, nullptr};
        uint32_t* inq[3] = {nullptr, nullptr, nullptr};
        uint64_t* lists[3] = {nullptr, nullptr, nullptr};
        // readback block (one copy per finished query): root | ctl[NCTL] | stats | edge count |
        // edges as (u, v, w) triples, ecap of them
        unsigned char* back = nullptr;
        RoundCtl* ctl = nullptr;
        unsigned long long* root = nullptr;
        unsigned long long* stats = nullptr;
        uint32_t* out_n = nullptr;
        uint32_t* edges = nullptr;
        unsigned long long* spider = nullptr;   // cheapest spider of the level-2 merge (automatic bound):
                                                // (cost << 32 | centre), lowered by its post-processed tree
        uint32_t* tour = nullptr;           // tour table 2^g * (g + 1) words, then walk scratch 2^g * g * g
        uint8_t* act = nullptr;             // activity bytes act[k * n + v] of the job's levels
        uint32_t* live = nullptr;           // the live list of the slot's current merge (n entries)
        int tour_g = 0;
        uint64_t* hub_items = nullptr;      // hub row items of a job's own graph (pull rounds)
        uint32_t* hub_cnt = nullptr;
        int32_t* grp = nullptr;
        size_t grp_cap = 0;
        // pinned host staging (args written at fill) and readback
        int32_t* h_grp = nullptr;
        SlotBlock* h_sb = nullptr;          // argument blocks
        unsigned char* h_back = nullptr;    // header and the first EDGE_PREFIX edges of back
        uint32_t* h_edges = nullptr;        // 3 * ecap (trees longer than the prefix)
        // host bookkeeping
        bool busy = false;
        gst_batch_job job;
        int g = 0;
        bool flt = false;
    };
    int S = 0;
    uint32_t n_cap = 0;
    int g_cap = 0;
    uint64_t items_cap = 0;
    uint32_t W_cap = 1;
    uint32_t ecap = 0;
    std::vector<Slot> slots;
    SlotBlock* d_sb = nullptr;              // [MAX_SLOTS]
    BatchPlan* d_plan = nullptr;
    SlotDev* h_slots = nullptr;             // pinned phase snapshot (the plan's sd_out)
    // level tables per group count
    uint64_t* pairs[GST_MAX_GROUPS + 1] = {nullptr};
    uint32_t* masks[GST_MAX_GROUPS + 1] = {nullptr};
    uint32_t* pos_of[GST_MAX_GROUPS + 1] = {nullptr};
    // hub work items of the last persistent graph (pull rounds)
    uint64_t gi_uid = 0;
    uint32_t gi_chunk = 0;
    uint64_t* all_items = nullptr;
    uint64_t all_cap = 0;
    uint32_t n_all = 0;
    uint64_t n_items_total = 0;
    uint32_t* all_cnt = nullptr;
    // captured step
    hipGraphExec_t gexec = nullptr;     // part A of a step
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec_b = nullptr;   // part B
    hipGraph_t graph_b = nullptr;
    hipStream_t cstream = nullptr;      // readbacks of finished queries
    hipEvent_t a_done = nullptr;        // end of the last step's part A
    int g_S = -1, g_R = -1, g_flt = -1;
};

// layout of a slot's readback block
constexpr size_t BK_CTL = 8;
constexpr size_t BK_STATS = (BK_CTL + sizeof(RoundCtl) * NCTL + 7) / 8 * 8;
constexpr size_t BK_OUTN = BK_STATS + 8 * STAT_WORDS;
constexpr size_t BK_EDGES = BK_OUTN + 8;
constexpr uint32_t EDGE_PREFIX = 1024;   // edges read back with the header (12 KiB)
static size_t back_bytes(uint32_t ecap) { return BK_EDGES + 12 * (size_t)ecap; }

static void free_slot(gst_batch_engine::Slot& sl) {
    hipFree(sl.T); hipFree(sl.gm);
    for (int i = 0; i < 3; ++i) { hipFree(sl.chg[i]); hipFree(sl.inq[i]); hipFree(sl.lists[i]); }
    hipFree(sl.back); hipFree(sl.grp); hipFree(sl.spider); hipFree(sl.tour); hipFree(sl.act); hipFree(sl.live);
    hipFree(sl.hub_items); hipFree(sl.hub_cnt);
    if (sl.h_grp) hipHostFree(sl.h_grp);
    if (sl.h_sb) hipHostFree(sl.h_sb);
    if (sl.h_back) hipHostFree(sl.h_back);
    if (sl.h_edges) hipHostFree(sl.h_edges);
    sl = gst_batch_engine::Slot();
}

static void free_engine_slots(gst_batch_engine* E) {
    for (auto& sl : E->slots) free_slot(sl);
    E->slots.clear();
    hipFree(E->d_sb); hipFree(E->d_plan);
    E->d_sb = nullptr; E->d_plan = nullptr;
    if (E->h_slots) hipHostFree(E->h_slots);
    E->h_slots = nullptr;
    if (E->gexec) hipGraphExecDestroy(E->gexec);
    if (E->gexec_b) hipGraphExecDestroy(E->gexec_b);
    if (E->graph_b) hipGraphDestroy(E->graph_b);
    E->gexec_b = nullptr; E->graph_b = nullptr;
    if (E->graph) hipGraphDestroy(E->graph);
    E->gexec = nullptr; E->graph = nullptr; E->g_S = -1;
    E->S = 0; E->n_cap = 0; E->g_cap = 0;
}

void gst_batch_free(gst_ctx* ctx) {
    gst_batch_engine* E = ctx->be;
    if (!E) return;
    hipSetDevice(ctx->device);
    free_engine_slots(E);
    for (int g = 0; g <= GST_MAX_GROUPS; ++g) { hipFree(E->pairs[g]); hipFree(E->masks[g]); hipFree(E->pos_of[g]); }
    hipFree(E->all_items); hipFree(E->all_cnt);
    if (E->cstream) hipStreamDestroy(E->cstream);
    if (E->a_done) hipEventDestroy(E->a_done);
    delete E;
    ctx->be = nullptr;
}

// bytes of device memory one slot needs
static size_t slot_bytes(uint32_t n, int g, uint64_t items_cap, uint32_t W, uint32_t ecap) {
    const size_t words = (size_t)n * (((size_t)1 << g) - 1);
    return words * 8 + (size_t)n * 8 + (size_t)n * (size_t)(g + 1) + 3 * ((size_t)n * W * 8 + (size_t)n * 4 + items_cap * 8) +
           sizeof(RoundCtl) * NCTL + 8 + 8 * STAT_WORDS + (3 * (size_t)ecap + 1) * 4 + 64 * 1024;
}

static gst_status ensure_slots(gst_ctx* ctx, gst_batch_engine* E, int S, uint32_t n, int g, int64_t m, uint32_t chunk) {
    uint32_t W = 1;
    const gst_levels& L = gst_get_levels(ctx, g);
    for (int k = 1; k <= g; ++k) W = std::max<uint32_t>(W, (uint32_t)((L.m[k] + 63) / 64));
    const uint64_t items_cap = (uint64_t)n + (uint64_t)m / chunk + 64;
    const uint32_t ecap = n + 2u * (uint32_t)g + 16u;
    if (E->S >= S && E->n_cap >= n && E->g_cap >= g && E->items_cap >= items_cap && E->W_cap >= W && !E->slots.empty())
        return GST_OK;
    hipStream_t s = (hipStream_t)ctx->stream;
    HIPCHK(hipStreamSynchronize(s));
    free_engine_slots(E);
    E->n_cap = n; E->g_cap = g; E->items_cap = items_cap; E->W_cap = W; E->ecap = ecap;
    E->slots.resize((size_t)S);
    bool ok = dalloc(E->d_sb, MAX_SLOTS) && dalloc(E->d_plan, 1) &&
              hipHostMalloc((void**)&E->h_slots, sizeof(SlotDev) * MAX_SLOTS) == hipSuccess;
    const size_t words = (size_t)n * (((size_t)1 << g) - 1);
    for (int i = 0; ok && i < S; ++i) {
        auto& sl = E->slots[(size_t)i];
        ok = ok && dalloc(sl.T, words) && dalloc(sl.gm, n) && dalloc(sl.act, (size_t)n * (size_t)(g + 1)) &&
             dalloc(sl.live, n);
        for (int j = 0; ok && j < 3; ++j)
            ok = dalloc(sl.chg[j], (size_t)n * W) && dalloc(sl.inq[j], n) && dalloc(sl.lists[j], items_cap);
        ok = ok && dalloc(sl.back, back_bytes(ecap)) && dalloc(sl.spider, 1) &&
             dalloc(sl.hub_items, items_cap) && dalloc(sl.hub_cnt, 1);
        if (ok) {
            sl.root = (unsigned long long*)sl.back;
            sl.ctl = (RoundCtl*)(sl.back + BK_CTL);
            sl.stats = (unsigned long long*)(sl.back + BK_STATS);
            sl.out_n = (uint32_t*)(sl.back + BK_OUTN);
            sl.edges = (uint32_t*)(sl.back + BK_EDGES);
        }
        ok = ok && hipHostMalloc((void**)&sl.h_sb, sizeof(SlotBlock)) == hipSuccess &&
             hipHostMalloc((void**)&sl.h_back, back_bytes(std::min(ecap, EDGE_PREFIX))) == hipSuccess &&
             hipHostMalloc((void**)&sl.h_edges, sizeof(uint32_t) * 3 * (size_t)ecap) == hipSuccess;
        if (ok) {
            const int fg = std::max(1, ctx->cu_count * 4);
            for (int j = 0; j < 3; ++j) {
                k_fill_u64<<<fg, 256, 0, s>>>(sl.chg[j], 0ull, (uint64_t)n * W);
                k_fill_u32<<<fg, 256, 0, s>>>(sl.inq[j], 0u, n);
            }
        }
    }
    if (!ok) {
        free_engine_slots(E);
        return gst_set_error(ctx, GST_OOM, "batch: device allocation of %d slots failed (V=%u, g=%d)", S, n, g);
    }
    k_fill_u32<<<64, 256, 0, s>>>((uint32_t*)E->d_sb, 0u, sizeof(SlotBlock) * MAX_SLOTS / 4);
    k_fill_u32<<<4, 256, 0, s>>>((uint32_t*)E->d_plan, 0u, sizeof(BatchPlan) / 4);
    E->S = S;
    HIPCHK(hipGetLastError());
    return GST_OK;
}

static gst_status ensure_tables(gst_ctx* ctx, gst_batch_engine* E, int g) {
    if (E->masks[g]) return GST_OK;
    const gst_levels& L = gst_get_levels(ctx, g);
    hipStream_t s = (hipStream_t)ctx->stream;
    if (!dalloc(E->pairs[g], std::max<size_t>(L.pairs.size(), 1)) || !dalloc(E->masks[g], L.masks.size()) ||
        !dalloc(E->pos_of[g], L.pos_of.size()))
        return gst_set_error(ctx, GST_OOM, "batch: level table allocation failed");
    if (!L.pairs.empty()) HIPCHK(hipMemcpyAsync(E->pairs[g], L.pairs.data(), L.pairs.size() * 8, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(E->masks[g], L.masks.data(), L.masks.size() * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(E->pos_of[g], L.pos_of.data(), L.pos_of.size() * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));   // the level vectors are host memory
    return GST_OK;
}

// hub items of a persistent graph, for pull rounds (rebuilt when the graph or chunk changes)
static gst_status ensure_hub_items(gst_ctx* ctx, gst_batch_engine* E, const gst_graph* G, uint32_t chunk) {
    if (E->gi_uid == G->uid && E->gi_chunk == chunk && E->all_items) return GST_OK;
    hipStream_t s = (hipStream_t)ctx->stream;
    const uint64_t cap = (uint64_t)G->n + (uint64_t)G->m / chunk + 64;
    if (E->all_cap < cap) {
        hipFree(E->all_items);
        E->all_items = nullptr;
        E->all_cap = 0;
        if (!dalloc(E->all_items, cap)) return gst_set_error(ctx, GST_OOM, "batch: item list allocation failed");
        E->all_cap = cap;
    }
    if (!E->all_cnt && !dalloc(E->all_cnt, 1)) return gst_set_error(ctx, GST_OOM, "batch: allocation failed");
    k_fill_u32<<<1, 64, 0, s>>>(E->all_cnt, 0u, 1);
    const uint32_t n = (uint32_t)G->n;
    k_compact<<<std::max(1, std::min<int>((int)((n + 255) / 256), ctx->cu_count * 8)), 256, 0, s>>>(
        nullptr, G->d_row, n, chunk, E->all_items, E->all_cnt, nullptr, HUB_DEG);
    uint32_t na = 0;
    HIPCHK(hipMemcpyAsync(&na, E->all_cnt, 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    uint64_t tot = 0;
    for (uint32_t v = 0; v < n; ++v) tot += (uint64_t)((G->h_row[v + 1] - G->h_row[v] + chunk - 1) / chunk);
    E->n_all = na; E->n_items_total = tot; E->gi_uid = G->uid; E->gi_chunk = chunk;
    return GST_OK;
}

// Enqueue the set-up of one query in slot `si`: group bits, level 1, its frontier, and the
// argument blocks of every level.
static gst_status fill_slot(gst_ctx* ctx, gst_batch_engine* E, int si, gst_batch_job& job, uint32_t chunk,
                            uint32_t tuned_lanes, int32_t tuned_pull) {
    hipStream_t s = (hipStream_t)ctx->stream;
    auto& sl = E->slots[(size_t)si];
    const gst_graph* G = job.G;
    const gst_groups* gr = job.groups;
    const int g = gr->n_groups;
    const uint32_t n = (uint32_t)G->n;
    const bool flt = G->w_type == GST_W_FLOAT32;
    const gst_levels& L = gst_get_levels(ctx, g);
    gst_status st = ensure_tables(ctx, E, g);
    if (st != GST_OK) return st;
    const bool persistent = !job.temp_graph;
    if (persistent) {
        st = ensure_hub_items(ctx, E, G, chunk);
        if (st != GST_OK) return st;
    }
    // groups
    const size_t gneed = (size_t)(g + 1) + (size_t)gr->grp_ptr[g];
    if (sl.grp_cap < gneed) {
        hipFree(sl.grp);
        if (sl.h_grp) hipHostFree(sl.h_grp);
        sl.grp = nullptr; sl.h_grp = nullptr; sl.grp_cap = 0;
        const size_t cap = std::max<size_t>(gneed, 4096);
        if (!dalloc(sl.grp, cap) || hipHostMalloc((void**)&sl.h_grp, cap * 4) != hipSuccess)
            return gst_set_error(ctx, GST_OOM, "batch: group buffer allocation failed");
        sl.grp_cap = cap;
    }
    for (int i = 0; i <= g; ++i) sl.h_grp[i] = gr->grp_ptr[i];
    std::memcpy(sl.h_grp + g + 1, gr->grp_vtx, sizeof(int32_t) * (size_t)gr->grp_ptr[g]);
    HIPCHK(hipMemcpyAsync(sl.grp, sl.h_grp, gneed * 4, hipMemcpyHostToDevice, s));

    // tour table (built on the device when the slot enters level 2)
    const bool use_tour = job.opts.prune_lb && job.opts.tour_bound == 0 && g >= 4 &&
                          (job.opts.auto_bound || (std::isfinite(job.opts.ub) && job.opts.ub > 0));
    if (use_tour && sl.tour_g < g) {
        hipFree(sl.tour);
        sl.tour = nullptr;
        sl.tour_g = 0;
        if (!dalloc(sl.tour, ((size_t)1 << g) * (size_t)(g + 1 + g * g)))
            return gst_set_error(ctx, GST_OOM, "batch: tour table allocation failed");
        sl.tour_g = g;
    }
    const uint32_t* tour = use_tour ? sl.tour : nullptr;

    uint64_t lvl_base[GST_MAX_GROUPS + 2] = {0};
    for (int k = 1; k <= g; ++k) lvl_base[k] = (uint64_t)n * (uint64_t)L.off[k];
    const gst_bound_bits bb = gst_make_bound(job.opts.ub, flt, job.opts.half_bound);
    const uint32_t ubb = bb.ub;
    // activity bytes for bounded jobs (a caller's bound or the automatic one)
    const bool use_act = GST_MERGE_SKIP && (ubb != GST_INF_COST || job.opts.auto_bound);
    // g >= 2: the set-up runs in the next step's part A, batched with the other slots filled this
    // step (four launches per step instead of five per query); g = 1 is set up here
    const bool defer = GST_BATCH_DEFER_FILL && g > 1;
    FillArgs& fa = sl.h_sb->f;
    std::memset(&fa, 0, sizeof(fa));
    if (defer) {
        fa.gm = sl.gm; fa.root = sl.root; fa.stats = sl.stats; fa.spider = sl.spider; fa.ctl = (uint32_t*)sl.ctl;
        fa.hub_cnt = sl.hub_cnt; fa.grp = sl.grp; fa.T1 = sl.T + lvl_base[1]; fa.row = G->d_row;
        fa.chg0 = sl.chg[0]; fa.inq0 = sl.inq[0]; fa.list0 = sl.lists[0]; fa.cnt0 = &sl.ctl[1].cnt[0];
        // a job's own graph gets its hub row items on the device, for pull rounds on dense frontiers
        fa.hub_items = persistent ? nullptr : sl.hub_items;
        fa.n = n; fa.g = (uint32_t)g; fa.chunk = chunk;
    } else {
        const int fg = std::max(1, std::min<int>((int)((n + 255) / 256), ctx->cu_count * 8));
        k_slot_reset<<<fg, 256, 0, s>>>(sl.gm, n, sl.root, sl.stats, sl.spider, (uint32_t*)sl.ctl, sl.hub_cnt);
        k_set_group_bits<<<64, 256, 0, s>>>(sl.gm, sl.grp, sl.grp + g + 1, g);
        k_init_level1<<<fg, 256, 0, s>>>(sl.T + lvl_base[1], sl.gm, G->d_row, n, (uint32_t)g, sl.chg[0], sl.inq[0],
                                        g > 1 ? 1 : 0);
        if (g > 1)
            k_compact<<<fg, 256, 0, s>>>(sl.inq[0], G->d_row, n, chunk, sl.lists[0], &sl.ctl[1].cnt[0], sl.stats, 0u);
        else
            k_root_scan_l1<<<fg, 256, 0, s>>>(sl.T, n, sl.root);
        if (!persistent) {
            // a job's own graph gets its hub row items on the device, for pull rounds on dense
            // frontiers (the count was cleared by k_slot_reset)
            k_compact<<<fg, 256, 0, s>>>(nullptr, G->d_row, n, chunk, sl.hub_items, sl.hub_cnt, nullptr, HUB_DEG);
        }
    }

    // argument blocks
    for (int k = 1; k < g; ++k) {
        GrowArgs& a = sl.h_sb->g[k];
        std::memset(&a, 0, sizeof(a));
        a.Tk = sl.T + lvl_base[k];
        a.T1 = sl.T + lvl_base[1];
        a.masks = E->masks[g] + L.off[k];
        a.row = G->d_row; a.arc = (const uint2*)G->d_arc;
        for (int i = 0; i < 3; ++i) { a.list[i] = sl.lists[i]; a.chg[i] = sl.chg[i]; a.inq[i] = sl.inq[i]; }
        a.ctl = sl.ctl + k;
        a.stats = sl.stats;
        a.m = (uint32_t)L.m[k];
        a.W = (uint32_t)((L.m[k] + 63) / 64);
        a.g = (uint32_t)g;
        a.chunk = chunk;
        a.ub = ubb;
        a.gmax = bb.gmax;
        a.lbcap = bb.lbcap;
        a.use_lb = (job.opts.prune_lb && k >= 2) ? 1u : 0u;
        a.tour = k >= 2 ? tour : nullptr;
        a.act = (k >= 2 && use_act) ? sl.act + (size_t)k * n : nullptr;
        // level 1 of an automatically bounded job is bounded by its running spider (the same word the
        // level-2 merge lowers to the cheapest spider); the plan of the step that ends level 1 moves
        // that bound to the later levels
        a.run_sp = (k == 1 && job.opts.auto_bound && g >= 3 && job.opts.half_bound == 0 && GST_RUN_SPIDER) ? sl.spider
                                                                                                          : nullptr;
        a.all_items = persistent ? E->all_items : sl.hub_items;
        a.n_all = persistent ? E->n_all : 0u;
        a.n_all_dev = persistent ? nullptr : sl.hub_cnt;
        a.n = n;
        {
            // the job's pull threshold (0: the tuned or built-in 450 per mille, < 0: never pull)
            const int32_t pm = job.opts.pull_permille != 0 ? job.opts.pull_permille
                                                           : (tuned_pull != 0 ? tuned_pull : 450);
            // a job's own graph: its work items are at most n + m / chunk
            const uint64_t items_total = persistent ? E->n_items_total : (uint64_t)n + (uint64_t)G->m / chunk;
            a.pull_thresh = pm > 0 ? (uint32_t)std::max<uint64_t>(1, items_total * (uint64_t)pm / 1000ull) : 0xFFFFFFFFu;
        }
        a.rlog = nullptr;
        a.seq = 0;
        a.push_lanes = job.opts.push_lanes > 0 ? (uint32_t)job.opts.push_lanes
                                               : (tuned_lanes ? tuned_lanes : default_push_lanes(true));
        a.push_lanes_sparse = 0;
        if (job.opts.push_lanes == 0 && a.push_lanes != PUSH_ITEM_PAR) {
            // a level of m masks gets at least the power of two >= m lanes per item (8..64): an
            // item's changed masks then spread over its lanes instead of looping (measured: +5% on
            // the g=6 batch, +27% on mixed g = 3..8, whose middle levels hold up to 70 masks)
            uint32_t want = 8;
            while (want < (uint32_t)L.m[k] && want < 64) want <<= 1;
            if (want > a.push_lanes) a.push_lanes = want;
            // dense rounds of the chosen levels one item per lane, the others in these lane groups
            if (itempar_level(GST_BATCH_ITEMPAR, k)) {
                a.push_lanes_sparse = a.push_lanes;
                a.push_lanes = PUSH_ITEM_PAR;
            }
        }
    }
    for (int k = 2; k <= g; ++k) {
        MergeArgs& ma = sl.h_sb->m[k];
        std::memset(&ma, 0, sizeof(ma));
        ma.T = sl.T; ma.pairs = E->pairs[g] + L.pair_off[k]; ma.masks = E->masks[g]; ma.gm = sl.gm; ma.row = G->d_row;
        ma.list0 = sl.lists[0]; ma.chg0 = sl.chg[0]; ma.inq0 = sl.inq[0]; ma.ctl = sl.ctl + k; ma.root = sl.root;
        ma.stats = sl.stats;
        for (int j = 0; j < GST_MAX_GROUPS + 2; ++j) { ma.lvl_base[j] = lvl_base[j]; ma.m[j] = (uint32_t)L.m[j]; ma.off[j] = (uint32_t)L.off[j]; }
        ma.n = n; ma.g = (uint32_t)g; ma.k = (uint32_t)k; ma.P = (uint32_t)L.pair_cnt[k];
        ma.W = (uint32_t)((L.m[k] + 63) / 64); ma.chunk = chunk; ma.ub = ubb; ma.use_lb = job.opts.prune_lb ? 1u : 0u;
        ma.gmax = bb.gmax; ma.lbcap = bb.lbcap;
        ma.spider = (job.opts.auto_bound && k == 2 && g >= 3) ? sl.spider : nullptr;
        ma.tour = tour;
        ma.grp = sl.grp;
        ma.tour_w = tour ? sl.tour + ((size_t)1 << g) * (size_t)(g + 1) : nullptr;
        ma.act = use_act ? sl.act : nullptr;
        ma.live = (use_act && k >= 3) ? sl.live : nullptr;
        ma.is_full = (k == g) ? 1u : 0u;
        const uint32_t Lk = (uint32_t)L.off[k], Wk = (uint32_t)((L.m[k] + 63) / 64);
        ma.VB = merge_tile_vertices(Lk, Wk);
        ma.pairs_in_lds = ((size_t)L.pair_cnt[k] * 8 <= MERGE_PAIRS_LDS) ? 1u : 0u;
    }
    {
        TraceArgs& ta = sl.h_sb->t;
        std::memset(&ta, 0, sizeof(ta));
        ta.T = sl.T; ta.pos_of = E->pos_of[g]; ta.row = G->d_row; ta.arc = (const uint2*)G->d_arc; ta.root = sl.root;
        ta.flt = flt ? 1u : 0u;
        ta.out_u = sl.edges; ta.out_v = sl.edges + 1; ta.out_w = sl.edges + 2; ta.ostride = 3;
        ta.out_n = sl.out_n;
        ta.ctl = sl.ctl;
        for (int j = 0; j < GST_MAX_GROUPS + 2; ++j) { ta.lvl_base[j] = lvl_base[j]; ta.m[j] = (uint32_t)L.m[j]; ta.off[j] = (uint32_t)L.off[j]; }
        ta.full = (uint32_t)L.full; ta.cap = E->ecap; ta.max_iter = n + 4u * (uint32_t)g + 64u;
    }
    sl.h_sb->sd.phase = g > 1 ? (defer ? PH_FILL : PH_GROW) : PH_TRACE;
    sl.h_sb->sd.k = 1;
    sl.h_sb->sd.g = (uint32_t)g;
    sl.h_sb->sd.pad = 0;
    HIPCHK(hipMemcpyAsync(E->d_sb + si, sl.h_sb, sizeof(SlotBlock), hipMemcpyHostToDevice, s));
    HIPCHK(hipGetLastError());
    sl.busy = true;
    sl.g = g;
    sl.flt = flt;
    return GST_OK;
}

// A step is two parts.  Part A: plan, tour tables, merges, bounds, compaction, traceback, then the
// slots' phase snapshot to the host; part B: R grow rounds.  The host reads part A's snapshot and
// the finished queries' results while part B runs.  ev (2 + 2R events, or nullptr) brackets the
// merge launch and every grow launch.
static void enqueue_step_a(gst_ctx* ctx, gst_batch_engine* E, const BatchArgs& ba, bool flt,
                           const hipEvent_t* ev = nullptr) {
    hipStream_t s = (hipStream_t)ctx->stream;
    const int mgrid = std::max<int>((int)MAX_SLOTS, ctx->cu_count * 8);
    // one LDS size for every (g <= g_cap, level) a slot may merge: tile rows, bitmaps, and the
    // largest split table that is staged in LDS (sized for the engine's group count, so that small
    // group counts leave room for more workgroups per CU)
    size_t pairs_lds = 0;
    for (int gg = 2; gg <= E->g_cap; ++gg) {
        const gst_levels& L = gst_get_levels(ctx, gg);
        for (int k = 2; k <= gg; ++k) {
            const size_t pb = (size_t)L.pair_cnt[k] * 8;
            if (pb <= MERGE_PAIRS_LDS) pairs_lds = std::max(pairs_lds, pb);
        }
    }
    const size_t mlds = MERGE_TILE_BYTES + 8 * 256 + 8 + pairs_lds;
    if (flt) k_plan_step<true><<<1, PLAN_THREADS, 0, s>>>(ba);
    else k_plan_step<false><<<1, PLAN_THREADS, 0, s>>>(ba);
    // the queries that entered slots since the last step: set-up in four passes over their slots
    const int fgrid = std::max<int>((int)MAX_SLOTS, ctx->cu_count * 4);
    k_batch_fill_reset<<<fgrid, 256, 0, s>>>(ba);
    k_batch_fill_bits<<<fgrid, 256, 0, s>>>(ba);
    k_batch_fill_level1<<<fgrid, 256, 0, s>>>(ba);
    k_batch_fill_compact<<<fgrid, 256, 0, s>>>(ba);
    // the tour tables' walk table sits in the merge launch's LDS when it fits there (the launch's LDS,
    // and so the merges' workgroups per CU, stays as the merges need), else in the slot's scratch
    const size_t tl = tour_lds_bytes((uint32_t)E->g_cap);
    const bool pw_lds = tl != 0 && tl <= mlds;
    if (ev) hipEventRecord(ev[0], s);
    k_batch_merge_live<<<mgrid, 256, 0, s>>>(ba);
    if (pw_lds) {
        if (flt) k_batch_merge<true, true><<<mgrid + MAX_SLOTS, MERGE_THREADS, mlds, s>>>(ba);
        else k_batch_merge<false, true><<<mgrid + MAX_SLOTS, MERGE_THREADS, mlds, s>>>(ba);
    } else {
        if (flt) k_batch_merge<true, false><<<mgrid + MAX_SLOTS, MERGE_THREADS, mlds, s>>>(ba);
        else k_batch_merge<false, false><<<mgrid + MAX_SLOTS, MERGE_THREADS, mlds, s>>>(ba);
    }
    if (ev) hipEventRecord(ev[1], s);
    if (flt) k_batch_spider_tree<true><<<MAX_SLOTS, MST_THREADS, 0, s>>>(ba);
    else k_batch_spider_tree<false><<<MAX_SLOTS, MST_THREADS, 0, s>>>(ba);
    if (flt) k_batch_bound<true><<<1, PLAN_THREADS, 0, s>>>(ba);
    else k_batch_bound<false><<<1, PLAN_THREADS, 0, s>>>(ba);
    k_batch_compact<<<mgrid, 256, 0, s>>>(ba);
    k_batch_trace<<<MAX_SLOTS, WAVE, 0, s>>>(ba);
    hipMemcpyAsync(E->h_slots, E->d_plan->sd_out, sizeof(SlotDev) * (size_t)ba.S, hipMemcpyDeviceToHost, s);
}

static void enqueue_step_b(gst_ctx* ctx, const BatchArgs& ba, int R, bool flt, const hipEvent_t* ev = nullptr) {
    hipStream_t s = (hipStream_t)ctx->stream;
    for (int r = 0; r < R; ++r) {
        k_plan_grow<<<1, PLAN_THREADS, 0, s>>>(ba);
        if (ev) hipEventRecord(ev[2 + 2 * r], s);
        if (flt) k_batch_grow<true><<<ba.grid + MAX_SLOTS, GROW_BLOCK, 0, s>>>(ba);
        else k_batch_grow<false><<<ba.grid + MAX_SLOTS, GROW_BLOCK, 0, s>>>(ba);
        if (ev) hipEventRecord(ev[3 + 2 * r], s);
    }
}

// The engine loop.  Returns the first failing job status once every job has finished (completed =
// true), or an engine error at once (completed = false: jobs may still be in slots).
static gst_status batch_run_impl(gst_ctx* ctx, const std::function<bool(gst_batch_job&, bool)>& next_job,
                                 const std::function<void(gst_batch_job&)>& job_done, uint32_t n_cap, int g_cap,
                                 int64_t m_cap, const gst_batch_opts* bo_in, gst_batch_stats* bst, uint64_t tune_uid,
                                 int64_t max_jobs, bool& completed) {
    completed = false;
    gst_batch_opts bo{};
    if (bo_in) bo = *bo_in;
    if (!ctx->be) ctx->be = new gst_batch_engine();
    gst_batch_engine* E = ctx->be;
    hipStream_t s = (hipStream_t)ctx->stream;
    HIPCHK(hipSetDevice(ctx->device));
    const uint32_t chunk = 32;
    const double t0 = now_ms();
    // slots: as asked, else as many as fit in half of the free device memory, at most 32
    int S = bo.slots > 0 ? std::min<int>(bo.slots, (int)MAX_SLOTS) : 0;
    {
        uint32_t W = 1;
        const gst_levels& L = gst_get_levels(ctx, g_cap);
        for (int k = 1; k <= g_cap; ++k) W = std::max<uint32_t>(W, (uint32_t)((L.m[k] + 63) / 64));
        const size_t per = slot_bytes(n_cap, g_cap, (uint64_t)n_cap + (uint64_t)m_cap / chunk + 64, W,
                                      n_cap + 2u * (uint32_t)g_cap + 16u);
        size_t fre = 0, tot = 0;
        HIPCHK(hipMemGetInfo(&fre, &tot));
        if (E->S > 0) fre += (size_t)E->S * per;   // the engine's own slots would be reused
        const int fit = (int)std::min<size_t>(MAX_SLOTS, fre / 2 / std::max<size_t>(per, 1));
        if (S == 0) S = std::min(32, fit);
        if (bo.slots <= 0 && max_jobs > 0 && S > max_jobs) S = (int)max_jobs;   // no slot would ever be used
        if (S < 1) return gst_set_error(ctx, GST_OOM, "batch: one slot needs %.2f GB of device memory", per / 1e9);
    }
    gst_status st = ensure_slots(ctx, E, S, n_cap, g_cap, m_cap, chunk);
    if (st != GST_OK) return st;   // an engine with more slots is reused; its extra slots stay idle
    const gst_tuned_batch* tb = tune_uid ? gst_find_tuned_batch(ctx, tune_uid) : nullptr;
    const int R = bo.rounds_per_step > 0 ? bo.rounds_per_step : (tb && tb->rounds_per_step > 0 ? tb->rounds_per_step : 8);
    const uint32_t tuned_lanes = tb ? (uint32_t)tb->push_lanes : 0u;
    const int32_t tuned_pull = tb ? tb->pull_permille : 0;
    const bool timed = bo.time_kernels != 0;
    const bool use_graph = bo.use_graph != 0 && !timed;
    std::vector<hipEvent_t> tev;
    double t_grow = 0.0, t_merge = 0.0;
    uint64_t n_merge_launch = 0;
    const size_t nev = 2 + 2 * (size_t)R;   // per step; two sets by step parity (a step's part B
                                            // may still run when the next step is enqueued)
    if (timed) {
        tev.resize(2 * nev);
        for (auto& e : tev) HIPCHK(hipEventCreate(&e));
    }
    if (!E->cstream) HIPCHK(hipStreamCreateWithFlags(&E->cstream, hipStreamNonBlocking));
    if (!E->a_done) HIPCHK(hipEventCreateWithFlags(&E->a_done, hipEventDisableTiming));
    hipStream_t cs = E->cstream;
    struct EvFree {
        std::vector<hipEvent_t>& v;
        ~EvFree() { for (auto& e : v) hipEventDestroy(e); }
    } ev_free{tev};

    BatchArgs ba;
    ba.sb = E->d_sb; ba.plan = E->d_plan;
    ba.S = (uint32_t)S;
    ba.grid = (uint32_t)std::max(1, std::min(ctx->cu_count * GST_BATCH_GRID_MULT, 16384));
    k_reset_phases<<<1, MAX_SLOTS, 0, s>>>(E->d_sb);

    std::vector<double> lat;
    uint64_t steps = 0, n_done = 0, relax = 0, gbytes = 0, mbytes = 0, busy_steps = 0;
    gst_status first_err = GST_OK;
    bool flt_set = false, flt = false;
    int busy = 0;
    // next_job may have nothing now and more later (a finished job or the caller's own pipeline can
    // make new work), so free slots are offered again after every step; the run ends when no slot is
    // busy and a waiting call has no job
    auto try_fill = [&](int si, bool wait) -> gst_status {
        for (;;) {
            gst_batch_job job;
            if (!next_job(job, wait)) break;
            const bool jf = job.G->w_type == GST_W_FLOAT32;
            if (!flt_set) { flt = jf; flt_set = true; }
            if (jf != flt) {
                job.status = gst_set_error(ctx, GST_INVALID_ARG, "batch: int and float graphs cannot share a batch");
                job_done(job);
                continue;
            }
            if ((uint32_t)job.G->n > E->n_cap || job.groups->n_groups > E->g_cap || job.groups->n_groups < 1) {
                job.status = gst_set_error(ctx, GST_INVALID_ARG, "batch: query exceeds the engine's capacity");
                job_done(job);
                continue;
            }
            auto& sl = E->slots[(size_t)si];
            sl.job = job;
            sl.job.t_fill = now_ms();
            gst_status fs = fill_slot(ctx, E, si, sl.job, chunk, tuned_lanes, tuned_pull);
            if (fs != GST_OK) return fs;
            busy++;
            return GST_OK;
        }
        return GST_OK;
    };
    // free slots take the jobs that are ready (stop at the first that gets none); with no slot busy,
    // wait for one job
    auto fill_free = [&]() -> gst_status {
        for (int i = 0; i < S; ++i) {
            if (E->slots[(size_t)i].busy) continue;
            const gst_status fs = try_fill(i, false);
            if (fs != GST_OK) return fs;
            if (!E->slots[(size_t)i].busy) break;
        }
        for (int i = 0; i < S && busy == 0; ++i) {
            if (E->slots[(size_t)i].busy) continue;
            const gst_status fs = try_fill(i, true);
            if (fs != GST_OK) return fs;
            if (!E->slots[(size_t)i].busy) break;
        }
        return GST_OK;
    };
    st = fill_free();
    if (st != GST_OK) return st;
    bool graph_ready = false;
    if (use_graph && busy > 0) {
        if (E->gexec && E->g_S == S && E->g_R == R && E->g_flt == (int)flt) {
            graph_ready = true;
        } else {
            if (E->gexec) { hipGraphExecDestroy(E->gexec); E->gexec = nullptr; }
            if (E->graph) { hipGraphDestroy(E->graph); E->graph = nullptr; }
            if (E->gexec_b) { hipGraphExecDestroy(E->gexec_b); E->gexec_b = nullptr; }
            if (E->graph_b) { hipGraphDestroy(E->graph_b); E->graph_b = nullptr; }
            HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            enqueue_step_a(ctx, E, ba, flt);
            HIPCHK(hipStreamEndCapture(s, &E->graph));
            HIPCHK(hipGraphInstantiate(&E->gexec, E->graph, nullptr, nullptr, 0));
            HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            enqueue_step_b(ctx, ba, R, flt);
            HIPCHK(hipStreamEndCapture(s, &E->graph_b));
            HIPCHK(hipGraphInstantiate(&E->gexec_b, E->graph_b, nullptr, nullptr, 0));
            E->g_S = S; E->g_R = R; E->g_flt = (int)flt;
            graph_ready = true;
        }
    }
    // a step runs rounds_per_step grow rounds of every query in flight; every level converges, so
    // this bound (minutes of work) only turns an engine defect into an error instead of a spin
    const uint64_t max_steps = bo.max_steps > 0 ? (uint64_t)bo.max_steps : 2ull * 1000 * 1000;
    // kernel times of one step (its events are complete once the next step's part A is)
    auto add_times = [&](uint64_t step) -> gst_status {
        const hipEvent_t* te = tev.data() + (size_t)(step & 1) * nev;
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, te[0], te[1]));
        t_merge += ms;
        n_merge_launch++;
        for (int r = 0; r < R; ++r) {
            HIPCHK(hipEventElapsedTime(&ms, te[2 + 2 * (size_t)r], te[3 + 2 * (size_t)r]));
            t_grow += ms;
        }
        return GST_OK;
    };
    // Per step: enqueue part A, its completion event, part B; wait for part A; collect the queries
    // it finished (their results are final after its traceback) and refill their slots while part
    // B runs: the refills queue behind part B on the engine stream, so the next step sees them.
    while (busy > 0) {
        const hipEvent_t* te = timed ? tev.data() + (size_t)(steps & 1) * nev : nullptr;
        if (graph_ready) HIPCHK(hipGraphLaunch(E->gexec, s));
        else enqueue_step_a(ctx, E, ba, flt, te);
        HIPCHK(hipEventRecord(E->a_done, s));
        if (graph_ready) HIPCHK(hipGraphLaunch(E->gexec_b, s));
        else enqueue_step_b(ctx, ba, R, flt, te);
        HIPCHK(hipGetLastError());
        steps++;
        busy_steps += (uint64_t)busy;
        HIPCHK(hipEventSynchronize(E->a_done));
        if (timed && steps >= 2) {
            st = add_times(steps - 2);
            if (st != GST_OK) return st;
        }
        if (steps >= max_steps) return gst_set_error(ctx, GST_INTERNAL, "batch: queries did not finish in %llu steps",
                                                     (unsigned long long)steps);
        // finished queries: read root, control blocks, statistics and edge count, then the edges
        std::vector<int> fin;
        for (int i = 0; i < S; ++i)
            if (E->slots[(size_t)i].busy && E->h_slots[i].phase == PH_DONE) fin.push_back(i);
        if (fin.empty()) {
            // jobs that became ready since the last step take the free slots (they join the next step)
            st = fill_free();
            if (st != GST_OK) return st;
            continue;
        }
        // one copy per finished query: the readback block's header, plus the first EDGE_PREFIX
        // edges when the tree is wanted; longer trees take a second copy
        const uint32_t pre = std::min(E->ecap, EDGE_PREFIX);
        HIPCHK(hipStreamWaitEvent(cs, E->a_done, 0));
        for (int i : fin) {
            auto& sl = E->slots[(size_t)i];
            const size_t nb = sl.job.opts.want_tree ? back_bytes(pre) : BK_EDGES;
            HIPCHK(hipMemcpyAsync(sl.h_back, sl.back, nb, hipMemcpyDeviceToHost, cs));
        }
        HIPCHK(hipStreamSynchronize(cs));
        bool more = false;
        for (int i : fin) {
            auto& sl = E->slots[(size_t)i];
            uint32_t nout = 0;
            std::memcpy(&nout, sl.h_back + BK_OUTN, 4);
            if (sl.job.opts.want_tree && nout > pre && nout <= E->ecap) {
                HIPCHK(hipMemcpyAsync(sl.h_edges, sl.edges, 12ull * nout, hipMemcpyDeviceToHost, cs));
                more = true;
            }
        }
        if (more) HIPCHK(hipStreamSynchronize(cs));
        std::vector<uint32_t> eu, ev, ew;
        for (int i : fin) {
            auto& sl = E->slots[(size_t)i];
            gst_batch_job& job = sl.job;
            const unsigned char* hb = sl.h_back;
            unsigned long long rw;
            std::memcpy(&rw, hb, 8);
            const RoundCtl* c = (const RoundCtl*)(hb + BK_CTL);
            unsigned long long hst[STAT_WORDS], hs[ST_NSTAT];
            std::memcpy(hst, hb + BK_STATS, sizeof(hst));
            stat_sum(hst, hs);
            uint32_t nout = 0;
            std::memcpy(&nout, hb + BK_OUTN, 4);
            uint32_t err = 0, rounds = 0;
            for (int k = 0; k <= sl.g; ++k) err |= c[k].err;
            for (int k = 1; k < sl.g; ++k) rounds += c[k].round + 1;
            gst_result* res = job.res;
            std::memset(res, 0, sizeof(*res));
            res->root = -1;
            res->n_groups = sl.g;
            const bool has_ub = gst_make_bound(job.opts.ub, sl.flt, 0).ub != GST_INF_COST;
            res->used_ub = has_ub ? 1 : 0;
            gst_stats& S2 = res->stats;
            S2.grow_items = hs[ST_ITEMS]; S2.grow_arcs = hs[ST_ARCS]; S2.grow_relax = hs[ST_RELAX];
            S2.grow_updates = hs[ST_UPD]; S2.grow_enqueued = hs[ST_ENQ]; S2.grow_new_states = hs[ST_NEW]; S2.grow_atomics = hs[ST_ATOM];
            S2.merge_vertices = hs[ST_MVERT]; S2.merge_pairs = hs[ST_MPAIRS];
            S2.grow_rounds = rounds;
            S2.grow_bytes = gst_grow_bytes(S2);
            S2.merge_bytes = hs[ST_MBYTES];
            relax += S2.grow_relax; gbytes += S2.grow_bytes; mbytes += S2.merge_bytes;
            if (err & 1u) {
                job.status = gst_set_error(ctx, GST_COST_OVERFLOW, "integer path cost exceeds 32 bits");
            } else if (rw == GST_INF_WORD) {
                job.status = gst_set_error(ctx, GST_INFEASIBLE, has_ub ? "no tree of cost <= bound touches all groups"
                                                                        : "no tree touches all groups");
            } else if ((err & 2u) || (job.opts.want_tree && nout > E->ecap)) {
                job.status = gst_set_error(ctx, GST_INTERNAL, "device traceback failed (edges=%u)", nout);
            } else {
                const uint32_t cb = (uint32_t)(rw >> 32);
                if (sl.flt) { float f; std::memcpy(&f, &cb, 4); res->cost = (double)f; }
                else res->cost = (double)cb;
                res->root = (int32_t)(rw & 0xFFFFFFFFu);
                if (job.opts.want_tree) {
                    // edges come as (u, v, w) triples: from the prefix, or from the second copy
                    const uint32_t* tri = nout <= pre ? (const uint32_t*)(hb + BK_EDGES) : sl.h_edges;
                    eu.resize(nout); ev.resize(nout); ew.resize(nout);
                    for (uint32_t e = 0; e < nout; ++e) { eu[e] = tri[3 * e]; ev[e] = tri[3 * e + 1]; ew[e] = tri[3 * e + 2]; }
                    gst_finish_tree(sl.flt, nout, eu.data(), ev.data(), ew.data(), res);
                }
                job.status = GST_OK;
            }
            if (job.status != GST_OK && first_err == GST_OK) first_err = job.status;
            if (job.status != GST_OK && (err & 1u)) {
                // an aborted level can leave change flags behind: clear the slot's frontier state
                const int fg = std::max(1, ctx->cu_count * 4);
                for (int j = 0; j < 3; ++j) {
                    k_fill_u64<<<fg, 256, 0, s>>>(sl.chg[j], 0ull, (uint64_t)E->n_cap * E->W_cap);
                    k_fill_u32<<<fg, 256, 0, s>>>(sl.inq[j], 0u, E->n_cap);
                }
            }
            job.t_done = now_ms();
            lat.push_back(job.t_done - job.t_fill);
            n_done++;
            sl.busy = false;
            busy--;
            job_done(job);
            // refill the slot with a ready job, or park it (a parked slot's phase block is written
            // again only by a later fill, which copies the whole block behind this copy)
            st = try_fill(i, false);
            if (st != GST_OK) return st;
            if (!sl.busy) {
                sl.h_sb->sd.phase = PH_EMPTY;
                HIPCHK(hipMemcpyAsync(&E->d_sb[i].sd, &sl.h_sb->sd, sizeof(SlotDev), hipMemcpyHostToDevice, s));
            }
        }
        st = fill_free();
        if (st != GST_OK) return st;
    }
    HIPCHK(hipStreamSynchronize(s));
    if (timed && steps >= 1) {
        st = add_times(steps - 1);
        if (st != GST_OK) return st;
    }
    if (bst) {
        std::memset(bst, 0, sizeof(*bst));
        bst->t_total_ms = now_ms() - t0;
        bst->queries_per_s = bst->t_total_ms > 0 ? n_done * 1000.0 / bst->t_total_ms : 0.0;
        if (!lat.empty()) {
            std::sort(lat.begin(), lat.end());
            auto pct = [&](double q) { return lat[std::min(lat.size() - 1, (size_t)std::floor(q * (lat.size() - 1) + 0.5))]; };
            bst->latency_p50_ms = pct(0.50);
            bst->latency_p99_ms = pct(0.99);
        }
        bst->steps = steps;
        bst->grow_launches = steps * (uint64_t)R;
        bst->grow_relax = relax;
        bst->grow_bytes = gbytes;
        bst->merge_bytes = mbytes;
        bst->slots = S;
        bst->rounds_per_step = R;
        bst->used_graph = graph_ready ? 1 : 0;
        bst->timed_kernels = timed ? 1 : 0;
        bst->t_grow_ms = t_grow;
        bst->t_merge_ms = t_merge;
        bst->merge_launches = n_merge_launch;
        bst->busy_slot_steps = busy_steps;
        bst->queries = (int64_t)n_done;
    }
    completed = true;
    return first_err;
}

gst_status gst_batch_run(gst_ctx* ctx, const std::function<bool(gst_batch_job&, bool)>& next_job,
                         const std::function<void(gst_batch_job&)>& job_done, uint32_t n_cap, int g_cap,
                         int64_t m_cap, const gst_batch_opts* bo_in, gst_batch_stats* bst, uint64_t tune_uid,
                         int64_t max_jobs, bool* engine_error) {
    bool completed = false;
    const gst_status st =
        batch_run_impl(ctx, next_job, job_done, n_cap, g_cap, m_cap, bo_in, bst, tune_uid, max_jobs, completed);
    if (engine_error) *engine_error = !completed;
    if (completed) return st;
    // An engine error left the loop early: wait until nothing of this call runs on the engine's
    // streams (the caller frees what the jobs point at), then free every slot.  Jobs still in slots
    // and the job a failed fill took are not reported to job_done: the caller marks every query whose
    // job never finished with the returned status.  An aborted level can leave change flags and queue
    // flags behind, so every slot's frontier state is cleared for the next call.
    gst_batch_engine* E = ctx->be;
    (void)hipStreamSynchronize((hipStream_t)ctx->stream);
    if (E && E->cstream) (void)hipStreamSynchronize(E->cstream);
    if (E) {
        hipStream_t s = (hipStream_t)ctx->stream;
        const int fg = std::max(1, ctx->cu_count * 4);
        for (auto& sl : E->slots) {
            sl.busy = false;
            for (int j = 0; j < 3; ++j) {
                if (sl.chg[j]) k_fill_u64<<<fg, 256, 0, s>>>(sl.chg[j], 0ull, (uint64_t)E->n_cap * E->W_cap);
                if (sl.inq[j]) k_fill_u32<<<fg, 256, 0, s>>>(sl.inq[j], 0u, E->n_cap);
            }
        }
        (void)hipStreamSynchronize(s);
    }
    return st;
}
