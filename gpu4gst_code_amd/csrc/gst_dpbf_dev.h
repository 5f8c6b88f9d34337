// Device code of the DPBF solve shared by the one-query driver (gst_dpbf.hip) and the batched
// engine (gst_batch.hip): the packed-state helpers, the control blocks and the kernel bodies.
// Every body takes its workgroup index and workgroup count as arguments, so a batched launch can
// hand each query slot its own slice of the grid.
#pragma once

#include <hip/hip_runtime.h>

#include "gst_internal.h"

namespace {

constexpr int WAVE = 64;
#ifndef GST_GROW_BLOCK
#define GST_GROW_BLOCK 256
#endif
constexpr int GROW_BLOCK = GST_GROW_BLOCK;   // 4 waves
constexpr int GROW_WAVES = GROW_BLOCK / WAVE;
constexpr int TRACE_STACK = 512;
constexpr int NCTL = GST_MAX_GROUPS + 2;     // one round-control block per level (entry 0: query flags)

// device-side control block of the grow rounds of one level
struct RoundCtl {
    union {
        struct {
            uint32_t round;
            uint32_t seq;   // launch number of the last finished round
        };
        unsigned long long rsq;   // (seq << 32) | round, read and written as one word
    };
    uint32_t cnt[3];        // items in list r%3
    uint32_t level_done;
    uint32_t err;           // bit0: int cost overflow, bit1: traceback failure
    uint32_t top;           // arrival counter over block groups
    uint32_t live;          // the level's merge: vertices its pre-pass listed as live
    uint32_t sub[64];       // arrival counters of groups of 32 workgroups
};

enum StatIdx {
    ST_ITEMS = 0, ST_ARCS, ST_RELAX, ST_UPD, ST_ENQ, ST_NEW, ST_ATOM, ST_MVERT, ST_MPAIRS, ST_MBYTES, ST_NSTAT
};
// The counters are striped: workgroup b adds into stripe b % STAT_STRIPES (ST_NSTAT words each), so
// that the one atomic per workgroup and counter at the end of a launch spreads over several
// addresses instead of queueing on one (measured: up to 15 us per merge launch of 1000+
// workgroups); readers sum the stripes (stat_sum).
#ifndef GST_STAT_STRIPES
#define GST_STAT_STRIPES 16
#endif
constexpr int STAT_STRIPES = GST_STAT_STRIPES;
constexpr int STAT_WORDS = ST_NSTAT * STAT_STRIPES;
__device__ __forceinline__ unsigned long long* stat_stripe(unsigned long long* stats, uint32_t bid) {
    return stats + (size_t)(bid % STAT_STRIPES) * ST_NSTAT;
}
inline void stat_sum(const unsigned long long* striped, unsigned long long* out) {
    for (int i = 0; i < ST_NSTAT; ++i) out[i] = 0;
    for (int s = 0; s < STAT_STRIPES; ++s)
        for (int i = 0; i < ST_NSTAT; ++i) out[i] += striped[(size_t)s * ST_NSTAT + i];
}


struct GrowArgs {
    uint64_t* Tk;          // level-k block
    const uint64_t* T1;    // level-1 block (group distances, for the lower bound)
    const uint32_t* masks; // masks of level k by rank
    const uint32_t* row;
    const uint2* arc;      // arc records: .x = neighbour, .y = weight bits
    uint64_t* list[3];
    uint64_t* chg[3];
    uint32_t* inq[3];
    RoundCtl* ctl;
    unsigned long long* stats;
    uint32_t m;            // masks at this level
    uint32_t W;            // 64-bit bitmap words per vertex
    uint32_t g;
    uint32_t chunk;        // arcs per work item
    uint32_t ub;           // cost bits of the bound (GST_INF_COST: none)
    uint32_t use_lb;
    const uint64_t* all_items;  // (vertex, chunk) items of the hub rows, for pull rounds
    uint32_t n_all;
    const uint32_t* n_all_dev;  // when set, the hub item count is read here (built on the device)
    uint32_t n;            // vertices
    uint32_t pull_thresh;  // a round with at least this many queued items pulls over all items
    // asynchronous (work-queue) mode
    unsigned long long* ring;   // 4 words per task slot
    uint32_t ring_mask;         // slots - 1 (power of two)
    struct AsyncCtl* actl;
    uint32_t* rlog;             // optional round log (RLOG_ROUNDS x 4 words), nullptr: off
    uint32_t seq;               // launch number within the level (1, 2, ...), set per launch
    uint32_t push_lanes;        // lanes per work item in push rounds (8, 16, 32 or 64; PUSH_ITEM_PAR: one
                                // item per lane)
    uint32_t push_lanes_sparse; // with PUSH_ITEM_PAR: lanes per item in rounds below GST_IP_MIN_ITEMS items
                                // (0: the item-parallel shape in every round)
    uint32_t gmax;              // a state is grown only while its cost is <= gmax (half the bound)
    uint32_t lbcap;             // level-1 distances above gmax are only known to exceed it: the lower
                                // bound reads min(distance, lbcap) (GST_INF_COST: no cap)
    const uint32_t* tour;       // tour table of the query (tour_table_body), nullptr: one-label only
    uint8_t* act;               // level k's activity bytes (act[v] = 1 once some (v, X) of level k is
                                // finite; read by the merges above), nullptr at level 1
    unsigned long long* run_sp; // level 1 of a solve with the automatic bound: the running spider word
                                // (cost << 32 | centre), lowered by the push items and read as a bound by
                                // every round (see run_spider_item); nullptr: off
};

// Activity bytes (and with them the merges' skipping of vertices with no finite lower state) are
// kept for bounded solves only: an unbounded solve makes almost every state finite, so there the
// checks would cost without skipping anything.  0 switches them off everywhere.
#ifndef GST_MERGE_SKIP
#define GST_MERGE_SKIP 1
#endif

// Level 1 of a solve with the automatic bound grows under its running spider (run_spider_item);
// 0 grows it unbounded until the level-2 merge finds the spider.
#ifndef GST_RUN_SPIDER
#define GST_RUN_SPIDER 1
#endif

// A state of the grow's level was lowered from inf (old word `old`): its vertex now holds a finite
// state at this level.  Every writer stores the same byte, so a plain store serves.
__device__ __forceinline__ uint32_t mark_active(uint8_t* act, uint32_t u, unsigned long long old) {
    const bool fresh = (uint32_t)(old >> 32) == 0xFFFFFFFFu;
    if (act && fresh) act[u] = 1;
    return fresh ? 1u : 0u;   // counted: states made finite (the rest of the updates lower a finite cost)
}

// Lanes per push item unless the caller chose (measured on BA V=200k, g=6): one query alone is
// latency-bound and fastest with a whole wave per item; a batch of queries is throughput-bound and
// fastest with four items per wave, whose chains of dependent loads then overlap.
#ifndef GST_BATCH_PUSH_LANES
#define GST_BATCH_PUSH_LANES 16
#endif
#ifndef GST_SOLO_PUSH_LANES
#define GST_SOLO_PUSH_LANES 64
#endif
__host__ __device__ __forceinline__ uint32_t default_push_lanes(bool batched) {
    return batched ? (uint32_t)GST_BATCH_PUSH_LANES : (uint32_t)GST_SOLO_PUSH_LANES;
}

// push_lanes == PUSH_ITEM_PAR selects the item-parallel push round (push_items_body): a wave takes up
// to 64 items, one per lane, and spreads their (item, mask) pairs and then their (pair, arc)
// relaxations over its lanes.  Round sizing counts ITEMPAR_PLAN items per wave for it.
constexpr uint32_t PUSH_ITEM_PAR = 1;
// relaxations in flight per lane in the item-parallel push round (IP_UNR x 64 per wave)
#ifndef GST_IP_UNR
#define GST_IP_UNR 2
#endif
constexpr uint32_t IP_UNR = GST_IP_UNR;
#ifndef GST_ITEMPAR_PLAN
#define GST_ITEMPAR_PLAN 4
#endif
// Levels whose dense rounds take the item-parallel shape when the caller leaves push_lanes to the
// library (batched engine / one-query solves): 0 = none, 1 = level 1 (the group distances, whose
// rounds are the densest), 2 = every level.  Measured on BA V=200k g=6 with 128 slots: level 1
// item-parallel takes the batched grow of 64 spider- and learned-bounded full solves from 245 to
// 272 GB/s and the spider-bounded stream up 6%, but the learned-bound stream down 6% (1429 vs 1522
// queries/s), and every level item-parallel loses on both; so the default is none.
#ifndef GST_BATCH_ITEMPAR
#define GST_BATCH_ITEMPAR 0
#endif
#ifndef GST_SOLO_ITEMPAR
#define GST_SOLO_ITEMPAR 0
#endif
__host__ __device__ __forceinline__ bool itempar_level(int mode, int k) { return mode == 2 || (mode == 1 && k == 1); }
// The item-parallel shape serves dense rounds: when the library chose it (sparse_lanes != 0), a
// round of fewer than GST_IP_MIN_ITEMS items, which is latency-bound, runs in lane groups of
// sparse_lanes lanes per item instead.
#ifndef GST_IP_MIN_ITEMS
#define GST_IP_MIN_ITEMS 8192
#endif
__host__ __device__ __forceinline__ uint32_t round_push_lanes(uint32_t push_lanes, uint32_t sparse_lanes, uint32_t n_cur) {
    return (push_lanes == PUSH_ITEM_PAR && sparse_lanes != 0 && n_cur < (uint32_t)GST_IP_MIN_ITEMS) ? sparse_lanes
                                                                                                  : push_lanes;
}
__host__ __device__ __forceinline__ uint32_t push_items_per_wave(uint32_t push_lanes) {
    return push_lanes == PUSH_ITEM_PAR ? (uint32_t)GST_ITEMPAR_PLAN : 64u / push_lanes;
}

// round log (GST_TRACE_ROUNDS=1): per round {queued items | pull<<31, workgroups taking part,
// wall clock at start (block 0), wall clock when the last workgroup finished}
constexpr uint32_t RLOG_ROUNDS = 512;

// device-side control of an asynchronous grow launch
struct AsyncCtl {
    uint32_t head;      // next task index to claim
    uint32_t tail;      // next task index to reserve
    uint32_t pending;   // tasks reserved (or initial) and not finished
    uint32_t err;       // bit0: int cost overflow, bit2: wait bound exceeded, bit3: ring overflow
    uint32_t pad[4];
};

struct MergeArgs {
    uint64_t* T;           // whole table
    const uint64_t* pairs; // pairs of this level
    const uint32_t* masks; // all masks by row position
    const uint32_t* gm;    // group mask per vertex
    const uint32_t* row;
    uint64_t* list0;
    uint64_t* chg0;
    uint32_t* inq0;
    RoundCtl* ctl;
    unsigned long long* root;
    unsigned long long* stats;
    uint64_t lvl_base[GST_MAX_GROUPS + 2];  // word offset of each level block
    uint32_t m[GST_MAX_GROUPS + 2];
    uint32_t off[GST_MAX_GROUPS + 2];
    uint32_t n;
    uint32_t g;
    uint32_t k;
    uint32_t P;            // pairs at this level
    uint32_t W;            // bitmap words at this level
    uint32_t chunk;
    uint32_t ub;
    uint32_t use_lb;
    uint32_t is_full;
    uint32_t VB;           // vertices per workgroup tile
    uint32_t pairs_in_lds;
    uint32_t gmax;         // merged states above this cost are not put on the frontier
    uint32_t lbcap;        // cap on level-1 distances in the lower bound (see GrowArgs)
    unsigned long long* spider;  // k == 2 with an automatic bound: min over the tile's vertices v of
                                 // (sum_i T(v,{i}) << 32 | v)
    const uint32_t* tour;  // tour table of the query (the merge itself applies the one-label bound only)
    // batched engine, level 2: the slot's groups and the walk scratch the tour table is built from
    const int32_t* grp;    // grp_ptr[0..g] then the members
    uint32_t* tour_w;      // 2^g * g * g words
    uint8_t* act;          // activity bytes of every level, act[j * n + v] (nullptr: no skipping)
    uint32_t* live;        // with act and k >= 3: the list of live vertices (n entries) that the
                           // pre-pass (merge_live_body) builds and the merge walks; nullptr: the
                           // merge walks every vertex in tiles and writes the dead ones itself
};

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

template <bool FLT>
__device__ __forceinline__ uint32_t add_cost(uint32_t a, uint32_t b, uint32_t& ovf) {
    if (a == GST_INF_COST || b == GST_INF_COST) return GST_INF_COST;
    if (FLT) {
        return __float_as_uint(__fadd_rn(__uint_as_float(a), __uint_as_float(b)));
    } else {
        const uint64_t s = (uint64_t)a + (uint64_t)b;
        if (s >= (uint64_t)GST_INF_COST) { ovf = 1; return GST_INF_COST; }
        return (uint32_t)s;
    }
}

// a + b >= ub, exact for ints, in float arithmetic for floats
template <bool FLT>
__device__ __forceinline__ bool reaches_ub(uint32_t a, uint32_t b, uint32_t ub) {
    if (ub == GST_INF_COST) return a == GST_INF_COST || b == GST_INF_COST;
    if (a == GST_INF_COST || b == GST_INF_COST) return true;
    if (FLT) return __fadd_rn(__uint_as_float(a), __uint_as_float(b)) >= __uint_as_float(ub);
    return (uint64_t)a + (uint64_t)b >= (uint64_t)ub;
}

// Pull: can an arc of weight w still lower (or tie with a smaller predecessor) a state of cost
// ocost and lower bound lb under the bound ub?  Every candidate through the arc costs at least w,
// so a false answer holds for every heavier arc of the row as well.
template <bool FLT>
__device__ __forceinline__ bool pull_arc_live(uint32_t w, uint32_t ocost, uint32_t lb, uint32_t ub) {
    if (w > ocost) return false;
    return ub == GST_INF_COST || !reaches_ub<FLT>(w, lb, ub);
}

// position of the t-th (0-based) set bit of x; t < popcount(x)
__device__ __forceinline__ uint32_t select_bit(uint64_t x, uint32_t t) {
    uint32_t lo = 0;
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) {
        const uint32_t mid = lo + s - 1;
        const uint64_t msk = (mid >= 63) ? ~0ull : ((2ull << mid) - 1ull);
        if ((uint32_t)__popcll(x & msk) <= t) lo += s;
    }
    return lo;
}

// OR over aligned groups of pl lanes (pl a power of two dividing the wave)
__device__ __forceinline__ uint64_t group_or64(uint64_t v, uint32_t pl) {
    for (uint32_t d = 1; d < pl; d <<= 1) v |= __shfl_xor(v, d, WAVE);
    return v;
}

__device__ __forceinline__ uint32_t wave_max32(uint32_t v) {
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) v = max(v, (uint32_t)__shfl_xor(v, d, WAVE));
    return v;
}

__device__ __forceinline__ uint64_t wave_or64(uint64_t v) {
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) v |= __shfl_xor(v, d, WAVE);
    return v;
}

__device__ __forceinline__ unsigned long long wave_min64(unsigned long long v) {
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        unsigned long long o = __shfl_xor(v, d, WAVE);
        v = o < v ? o : v;
    }
    return v;
}

__device__ __forceinline__ uint32_t wave_min32(uint32_t v) {
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) v = min(v, (uint32_t)__shfl_xor(v, d, WAVE));
    return v;
}

// Bound bits of a spider cost sp (a feasible tree's cost), by the rule of the host's bound set-up:
// ints keep cost <= sp and grow cost <= sp/2; floats get a relative slack of about 1e-4 and one ulp
// (bit patterns of positive floats step by one ulp per unit).
template <bool FLT>
__device__ __forceinline__ void spider_bound(uint32_t sp, uint32_t& ub, uint32_t& gmax) {
    if (FLT) {
        const float u = __fmul_rn(__uint_as_float(sp), 1.0001f);
        ub = __float_as_uint(u) + 1u;
        gmax = __float_as_uint(__fmul_rn(__uint_as_float(ub), 0.5f)) + 1u;
    } else {
        ub = sp + 1u;
        gmax = sp / 2u;
    }
}

__device__ __forceinline__ uint32_t wave_sum32(uint32_t v) {
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) v += __shfl_xor(v, d, WAVE);
    return v;
}

// Work-item queue of one workgroup in LDS.  Waves reserve slots with an LDS atomic (one per
// wave and step), and the workgroup appends its items to the global list with a single global
// atomicAdd at the end, so the global counter sees one atomic per workgroup instead of one per
// wave-step.  A reservation that does not fit goes straight to the global list.
constexpr uint32_t QCAP = 2048;
struct BlockQueue {
    uint64_t items[QCAP];
    uint32_t n;       // slots reserved
    uint32_t valid;   // items[0, min(n, valid)) hold queued items
    uint32_t gbase;
};

__device__ __forceinline__ void bq_init(BlockQueue& q) {
    if (threadIdx.x == 0) { q.n = 0; q.valid = QCAP; }
}

// Append n items (v, chunk 0..n-1) per lane.  Convergent: every lane of the wave calls it.
__device__ __forceinline__ void bq_push(BlockQueue& q, uint64_t* list, uint32_t* cnt, uint32_t v,
                                        uint32_t n, uint32_t& enq_stat) {
    if (__ballot(n != 0) == 0) return;
    const uint32_t lane = lane_id();
    uint32_t incl = n;
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        const uint32_t y = __shfl_up(incl, d, WAVE);
        if (lane >= (uint32_t)d) incl += y;
    }
    const uint32_t total = __shfl(incl, WAVE - 1, WAVE);
    uint32_t b = 0, gb = 0xFFFFFFFFu;
    if (lane == 0) {
        b = atomicAdd(&q.n, total);
        if (b + total > QCAP) {
            atomicMin(&q.valid, b);
            gb = atomicAdd(cnt, total);
        }
    }
    b = __shfl(b, 0, WAVE);
    gb = __shfl(gb, 0, WAVE);
    const uint32_t off = incl - n;
    for (uint32_t c = 0; c < n; ++c) {
        const uint64_t item = (uint64_t)v | ((uint64_t)c << 32);
        if (gb == 0xFFFFFFFFu) q.items[b + off + c] = item;
        else list[gb + off + c] = item;
    }
    enq_stat += n;
}

// Append, per lane, n[j] items (v[j], chunk 0..n[j]-1) for j < N with one reservation for the
// wave.  Convergent: every lane of the wave calls it.
template <int N>
__device__ __forceinline__ void bq_push_multi(BlockQueue& q, uint64_t* list, uint32_t* cnt, const uint32_t (&v)[N],
                                              const uint32_t (&n)[N], uint32_t& enq_stat) {
    uint32_t tot = 0;
#pragma unroll
    for (int j = 0; j < N; ++j) tot += n[j];
    if (__ballot(tot != 0) == 0) return;
    const uint32_t lane = lane_id();
    uint32_t incl = tot;
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        const uint32_t y = __shfl_up(incl, d, WAVE);
        if (lane >= (uint32_t)d) incl += y;
    }
    const uint32_t total = __shfl(incl, WAVE - 1, WAVE);
    uint32_t b = 0, gb = 0xFFFFFFFFu;
    if (lane == 0) {
        b = atomicAdd(&q.n, total);
        if (b + total > QCAP) {
            atomicMin(&q.valid, b);
            gb = atomicAdd(cnt, total);
        }
    }
    b = __shfl(b, 0, WAVE);
    gb = __shfl(gb, 0, WAVE);
    uint32_t off = incl - tot;
#pragma unroll
    for (int j = 0; j < N; ++j) {
        for (uint32_t c = 0; c < n[j]; ++c, ++off) {
            const uint64_t item = (uint64_t)v[j] | ((uint64_t)c << 32);
            if (gb == 0xFFFFFFFFu) q.items[b + off] = item;
            else list[gb + off] = item;
        }
    }
    enq_stat += tot;
}

// Publish the workgroup's queued items (all threads of the block call it).
__device__ __forceinline__ void bq_flush(BlockQueue& q, uint64_t* list, uint32_t* cnt) {
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t k = min(q.n, q.valid);
        q.gbase = k ? atomicAdd(cnt, k) : 0u;
        q.n = k;
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < q.n; i += blockDim.x) list[q.gbase + i] = q.items[i];
}

__device__ __forceinline__ uint32_t num_chunks(const uint32_t* row, uint32_t v, uint32_t chunk) {
    const uint32_t d = row[v + 1] - row[v];
    return (d + chunk - 1) / chunk;
}

// One-label lower bound of state (v, X): max over groups i not in X of dist(v, G_i), read from the
// (final) level-1 row of v; inf when some missing group is unreachable within the bound.  When level
// 1 grew only states of cost <= gmax, a distance above gmax (or inf) says only that the true distance
// exceeds gmax: each is capped at lbcap = gmax (bit patterns of non-negative floats order like the
// floats, so the cap is a u32 min for both weight types).
//
// With a tour table (tour != nullptr) the bound also takes the tour bounds of PrunedDP++
// (rucgraph/graph_v_of_v_idealID_PrunedDPPlusPlus.h:435-494).  Any tree that completes (v, X) holds v
// and touches every missing group (the set U); walking around it from v passes every group of U and
// costs at most twice its weight.  With W(i, U) the cheapest walk between groups that starts at
// group i and passes every group of U (group-to-group steps priced by the closest pair of members)
// and Wmin(U) = min_i W(i, U):
//   lb2 = (max_i (d_i + W(i, U)) + min_i d_i) / 2          (the reference's lb2, :471-476)
//   lb1 = min_i d_i + Wmin(U) / 2                          (a relaxation of the reference's lb1,
//                                                           :463-470, which takes the walk's ends)
// The table holds W(i, U) at tour[U * (g + 1) + i] and Wmin(U) at tour[U * (g + 1) + g]; an
// entry of GST_INF_COST switches the tour terms off for that state, tour[0] == 0 for every state.  Int costs round the halves up
// (the completing tree's cost is an integer); float bounds are scaled by 0.9998, far beyond the
// rounding of the few float additions behind them.
template <bool FLT, class DistF>
__device__ __forceinline__ uint32_t state_lb(DistF dist, const uint32_t* tour, uint32_t g, uint32_t X, uint32_t lbcap) {
    uint32_t one = 0, dmin = GST_INF_COST;
    const uint32_t U = ((1u << g) - 1u) & ~X;
    // tour[0] (the row of the empty set, never a missing set) says whether the table is in use
    const uint32_t* tu = (tour && (U & (U - 1u)) && tour[0]) ? tour + (size_t)U * (g + 1u) : nullptr;
    bool tour_ok = tu != nullptr;
    uint64_t tmax_i = 0;
    float tmax_f = 0.0f;
    for (uint32_t gi = 0; gi < g; ++gi) {
        if ((X >> gi) & 1u) continue;
        const uint32_t di = min(dist(gi), lbcap);
        if (di == GST_INF_COST) return GST_INF_COST;
        one = max(one, di);
        if (tour_ok) {
            dmin = min(dmin, di);
            const uint32_t w2 = tu[gi];
            if (w2 == GST_INF_COST) {
                tour_ok = false;
            } else if (FLT) {
                tmax_f = fmaxf(tmax_f, __fadd_rn(__uint_as_float(di), __uint_as_float(w2)));
            } else {
                tmax_i = max(tmax_i, (uint64_t)di + w2);
            }
        }
    }
    if (!tour_ok) return one;
    const uint32_t wmin = tu[g];
    if (wmin == GST_INF_COST) return one;
    uint32_t lb2, lb1;
    if (FLT) {
        const float dm = __uint_as_float(dmin);
        lb2 = __float_as_uint(__fmul_rn(__fmul_rn(__fadd_rn(tmax_f, dm), 0.5f), 0.9998f));
        lb1 = __float_as_uint(__fmul_rn(__fadd_rn(dm, __fmul_rn(__uint_as_float(wmin), 0.5f)), 0.9998f));
    } else {
        const uint64_t a = (tmax_i + dmin + 1u) / 2u, b = (uint64_t)dmin + ((uint64_t)wmin + 1u) / 2u;
        lb2 = (uint32_t)min<uint64_t>(a, GST_INF_COST - 1u);
        lb1 = (uint32_t)min<uint64_t>(b, GST_INF_COST - 1u);
    }
    return max(one, max(lb2, lb1));
}

template <bool FLT>
__device__ __forceinline__ uint32_t label_lb(const uint64_t* T1, const uint32_t* tour, uint32_t v, uint32_t g,
                                             uint32_t X, uint32_t lbcap) {
    const uint64_t* r1 = T1 + (uint64_t)v * g;
    return state_lb<FLT>([&](uint32_t gi) { return (uint32_t)(r1[gi] >> 32); }, tour, g, X, lbcap);
}

// Builds the tour table of one query with one workgroup, once level 1 is final:
//   D(a, b) = min over members u of group b of min(T(u, {a}), lbcap)  (closest pair of members, as
//             level 1 knows it: a distance above the cap is only known to exceed it)
//   metric closure of D over intermediate groups, then a Held-Karp DP over group sets:
//   P(m, i, j) = cheapest walk from group i to group j passing every group of m (i, j in m),
//   one popcount layer of m at a time (open walks, then closed ones i == j of the same layer)
// into tour[m * (g + 1) + i] = min_j P(m, i, j) and tour[m * (g + 1) + g] = min_i of those.
// The tour terms exceed the one-label bound by at most half a walk, and a grown state has cost
// <= gmax, so walks much shorter than the grow limit cannot prune: when the cheapest walk through
// every group is below 2/5 of lbcap (about U/5; groups that overlap or lie close together), tour[0]
// = 0 switches the table off and the grow rounds skip its loads.
// The walk table P (2^g * g * g words) lives in LDS when it fits (TOUR_LDS_MAX_G groups: 64 KiB,
// given to the launch as dynamic LDS), else in the global scratch buffer Pw: each popcount layer
// reads the previous one, so its latency repeats once per layer.  Sums that overflow the int range
// saturate to GST_INF_COST, which only ever weakens the bound.
constexpr uint32_t TOUR_LDS_MAX_G = 8;
// the table is used when the walk through all groups is at least this many tenths of lbcap
#ifndef GST_TOUR_MIN_TENTHS
#define GST_TOUR_MIN_TENTHS 4
#endif
__host__ __device__ constexpr size_t tour_lds_bytes(uint32_t g) {
    return g <= TOUR_LDS_MAX_G ? ((size_t)1 << g) * g * g * 4 : 0;
}
template <bool FLT, bool PW_LDS>
__device__ void tour_table_body(const uint64_t* T1, const int32_t* grp_ptr, const int32_t* grp_vtx, uint32_t g,
                                uint32_t lbcap, uint32_t* Pw_global, uint32_t* tour) {
    __shared__ uint32_t D[GST_MAX_GROUPS * GST_MAX_GROUPS];
    extern __shared__ uint32_t tour_pw_lds[];
    uint32_t* Pw = PW_LDS ? tour_pw_lds : Pw_global;
    const uint32_t tid = threadIdx.x, nt = blockDim.x;
    const uint32_t G2 = g * g, S = 1u << g;
    uint32_t o = 0;
    for (uint32_t i = tid; i < G2; i += nt) D[i] = GST_INF_COST;
    __syncthreads();
    const uint32_t tot = (uint32_t)grp_ptr[g];
    for (uint32_t e = tid; e < tot; e += nt) {
        uint32_t b = 0;
        while (b + 1 < g && (uint32_t)grp_ptr[b + 1] <= e) ++b;
        const uint32_t u = (uint32_t)grp_vtx[e];
        const uint64_t* r1 = T1 + (uint64_t)u * g;
        for (uint32_t a = 0; a < g; ++a) atomicMin(&D[a * g + b], min((uint32_t)(r1[a] >> 32), lbcap));
    }
    __syncthreads();
    // row and column kk do not change in round kk (D(kk, kk) = 0), so one round runs in parallel
    for (uint32_t kk = 0; kk < g; ++kk) {
        for (uint32_t i = tid; i < G2; i += nt) {
            const uint32_t a = i / g, b = i - a * g;
            const uint32_t c = add_cost<FLT>(D[a * g + kk], D[kk * g + b], o);
            if (c < D[i]) D[i] = c;
        }
        __syncthreads();
    }
    for (uint32_t x = tid; x < S * G2; x += nt) Pw[x] = GST_INF_COST;
    __syncthreads();
    for (uint32_t i = tid; i < g; i += nt) Pw[(1u << i) * G2 + i * g + i] = 0u;
    __syncthreads();
    for (uint32_t c = 2; c <= g; ++c) {
        for (uint32_t x = tid; x < S * G2; x += nt) {
            const uint32_t m = x / G2, rem = x - m * G2, i = rem / g, j = rem - i * g;
            if ((uint32_t)__popc(m) != c || i == j || !((m >> i) & 1u) || !((m >> j) & 1u)) continue;
            const uint32_t prev = m ^ (1u << j);
            uint32_t best = GST_INF_COST;
            for (uint32_t kk = 0; kk < g; ++kk)
                if ((prev >> kk) & 1u) best = min(best, add_cost<FLT>(Pw[prev * G2 + i * g + kk], D[kk * g + j], o));
            Pw[x] = best;
        }
        __syncthreads();
        for (uint32_t x = tid; x < S * g; x += nt) {
            const uint32_t m = x / g, i = x - m * g;
            if ((uint32_t)__popc(m) != c || !((m >> i) & 1u)) continue;
            uint32_t best = GST_INF_COST;
            for (uint32_t kk = 0; kk < g; ++kk)
                if (kk != i && ((m >> kk) & 1u)) best = min(best, add_cost<FLT>(Pw[m * G2 + i * g + kk], D[kk * g + i], o));
            Pw[m * G2 + i * g + i] = best;
        }
        __syncthreads();
    }
    for (uint32_t x = tid; x < S * (g + 1); x += nt) {
        const uint32_t m = x / (g + 1), i = x - m * (g + 1);
        uint32_t best = GST_INF_COST;
        for (uint32_t a = 0; a < g; ++a) {
            if (!((m >> a) & 1u) || (i < g && a != i)) continue;
            for (uint32_t j = 0; j < g; ++j)
                if ((m >> j) & 1u) best = min(best, Pw[m * G2 + a * g + j]);
        }
        tour[x] = best;
    }
    __syncthreads();
    if (tid == 0) {
        const uint32_t wall = tour[(size_t)(S - 1) * (g + 1) + g];
        bool use = wall != GST_INF_COST && wall != 0u;
        if (use && lbcap != GST_INF_COST) {
            if (FLT) use = __uint_as_float(wall) * 10.0f >= (float)GST_TOUR_MIN_TENTHS * __uint_as_float(lbcap);
            else use = (uint64_t)wall * 10u >= (uint64_t)lbcap * GST_TOUR_MIN_TENTHS;
        }
        tour[0] = use ? 1u : 0u;
    }
}

template <bool FLT, bool PW_LDS>
__global__ void __launch_bounds__(256) k_tour_table(const uint64_t* T1, const int32_t* grp, uint32_t g, uint32_t lbcap,
                                                    uint32_t* Pw, uint32_t* tour) {
    tour_table_body<FLT, PW_LDS>(T1, grp, grp + g + 1, g, lbcap, Pw, tour);
}

// push rounds: 1 = level 1 relaxes its changed masks without re-checking them against the bound
#ifndef GST_LEVEL1_NOFILTER
#define GST_LEVEL1_NOFILTER 1
#endif
// push rounds: 1 = read a neighbour's change word before the atomic OR that sets its bit (skips the
// atomic when the bit is set already, at the cost of one more dependent load when it is not)
#ifndef GST_CHG_PREREAD
#define GST_CHG_PREREAD 0
#endif
// push rounds: 1 = the atomics of a lane's UNR arcs are issued together (the change bit beside the
// atomic-min), 0 = each arc's atomic-min, then its change bit, then its queue push, one arc at a time
#ifndef GST_PUSH_BATCH_ATOMICS
#define GST_PUSH_BATCH_ATOMICS 1
#endif
// arc steps per lane issued together.  With the atomics of a step batched, 4 steps spill 38-45
// registers per lane into scratch memory, whose traffic doubled the grow's HBM bytes; 2 spill 7
// (measured on BA V=200k g=6, 128 slots: 1 / 2 / 3 / 4 steps give the learned-bound stream 1508 /
// 1616 / 1586 / 1530 queries/s)
#ifndef GST_GROW_UNR
#define GST_GROW_UNR 2
#endif
#ifndef GST_GROW_WPE
#define GST_GROW_WPE 5
#endif
constexpr uint32_t UNR = GST_GROW_UNR;   // arc steps issued together (independent gathers in flight per wave)
// Rows longer than this are "hubs": pull rounds split them into chunk items combined with atomics;
// shorter rows are pulled whole by one lane group with plain stores.
constexpr uint32_t HUB_DEG = 128;

// ---------------------------------------------------------------------------------------
// small utility kernels (fills instead of hipMemset so that every step can be graph-captured)
// ---------------------------------------------------------------------------------------
__global__ void k_fill_u32(uint32_t* p, uint32_t val, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = val;
}

__global__ void k_fill_u64(uint64_t* p, uint64_t val, uint64_t n) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = val;
}

__global__ void k_set_group_bits(uint32_t* gm, const int32_t* grp_ptr, const int32_t* grp_vtx, int g) {
    for (int i = 0; i < g; ++i) {
        const int s = grp_ptr[i], e = grp_ptr[i + 1];
        for (int j = s + blockIdx.x * blockDim.x + threadIdx.x; j < e; j += gridDim.x * blockDim.x)
            atomicOr(&gm[grp_vtx[j]], 1u << i);
    }
}

__global__ void k_ctl_reset(RoundCtl* ctl) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        ctl->round = 0; ctl->seq = 0;
        ctl->cnt[0] = 0; ctl->cnt[1] = 0; ctl->cnt[2] = 0;
        ctl->level_done = 0;
    }
}

// Level 1: T(v,{i}) = 0 for members of group i, inf elsewhere; members with arcs are flagged as
// the first frontier (bitmap = their groups).
__global__ void k_init_level1(uint64_t* T1, const uint32_t* gm, const uint32_t* row, uint32_t n,
                              uint32_t g, uint64_t* chg0, uint32_t* inq0, int seed_frontier) {
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n;
         v += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t gmv = gm[v];
        for (uint32_t i = 0; i < g; ++i)
            T1[v * g + i] = ((gmv >> i) & 1u) ? (uint64_t)GST_PRED_LEAF : GST_INF_WORD;
        if (gmv && seed_frontier && row[v + 1] > row[v]) { chg0[v] = gmv; inq0[v] = 1; }
    }
}

// Work-item list from per-vertex flags (flags == nullptr: every vertex with arcs), one item per
// chunk of the vertex's row; one global atomic per workgroup.
__device__ __forceinline__ void compact_body(const uint32_t* flags, const uint32_t* row, uint32_t n, uint32_t chunk,
                                             uint64_t* list, uint32_t* cnt_out, unsigned long long* stats,
                                             uint32_t min_deg, uint32_t bid, uint32_t nblk) {
    __shared__ uint32_t s_wtot[4];
    __shared__ uint32_t s_base;
    const uint32_t lane = lane_id(), wid = threadIdx.x / WAVE;
    for (uint64_t base = (uint64_t)bid * blockDim.x; base < n; base += (uint64_t)nblk * blockDim.x) {
        const uint64_t v = base + threadIdx.x;
        uint32_t k = 0;
        if (v < n && (flags == nullptr || flags[v]) && row[v + 1] - row[v] > min_deg)
            k = (row[v + 1] - row[v] + chunk - 1) / chunk;
        uint32_t incl = k;
#pragma unroll
        for (int d = 1; d < WAVE; d <<= 1) {
            const uint32_t y = __shfl_up(incl, d, WAVE);
            if (lane >= (uint32_t)d) incl += y;
        }
        if (lane == WAVE - 1) s_wtot[wid] = incl;
        __syncthreads();
        if (threadIdx.x == 0) {
            const uint32_t tot = s_wtot[0] + s_wtot[1] + s_wtot[2] + s_wtot[3];
            s_base = tot ? atomicAdd(cnt_out, tot) : 0u;
            if (tot && stats) atomicAdd(&stat_stripe(stats, bid)[ST_ENQ], (unsigned long long)tot);
        }
        __syncthreads();
        uint32_t off = s_base;
        for (uint32_t w = 0; w < wid; ++w) off += s_wtot[w];
        off += incl - k;
        for (uint32_t c = 0; c < k; ++c) list[off + c] = v | ((uint64_t)c << 32);
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256) k_compact(const uint32_t* flags, const uint32_t* row, uint32_t n,
                                                 uint32_t chunk, uint64_t* list, uint32_t* cnt_out,
                                                 unsigned long long* stats, uint32_t min_deg) {
    compact_body(flags, row, n, chunk, list, cnt_out, stats, min_deg, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------------------------------
// grow, item-parallel push round.  A wave takes up to 64 queued items at once, one per lane
// (item loads, row bounds, change words: one round trip for all of them), spreads their changed
// masks over its lanes as (item, mask) pairs, up to 64 at a time (own costs and lower bounds: one
// round trip), keeps the pairs under the bound, and spreads their arcs over its lanes as
// (pair, arc) relaxations, UNR x 64 at a time (arc records, neighbour states, then the atomics:
// three round trips).  Each dependent step of the relaxation chain thus serves up to 64
// independent accesses of the wave, where a lane group per item serves one.
// ---------------------------------------------------------------------------------------
// one wave's staging: the pairs of the current batch, then the kept ones
struct PushWaveLds {
    uint32_t src[WAVE];    // batch pair q: source lane | bit << 8
    uint32_t v[WAVE];      // kept pair k: vertex
    uint32_t rk[WAVE];     // mask rank | (first kept pair of its item) << 31
    uint32_t cv[WAVE];     // own cost
    uint32_t e0[WAVE];     // first arc of the item's chunk
    uint32_t incl[WAVE];   // inclusive prefix of the kept pairs' arc counts
    uint32_t cut[WAVE];    // arcs of the pair that may still be live (rows run in increasing weight)
};

// orders the LDS accesses of one wave's lanes (the hardware runs a wave's LDS operations in order;
// this keeps the compiler from moving them across each other)
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// GST_ITEMPAR_CODE 0 leaves this body out of the grow kernels (push_lanes = 1 then runs 64-lane
// groups).  Inlined: out of line, the call and its scratch frame made the one-query grow 2x slower.
#ifndef GST_ITEMPAR_CODE
#define GST_ITEMPAR_CODE 1
#endif
// Running spider of level 1 (automatic bound).  While level 1 grows, every tentative group distance
// d_i(v) is the cost of a real path from v to group i (its predecessor chain), so for a vertex v that
// holds a distance to every group, sum_i d_i(v) is the cost of a feasible structure (the union of
// those paths, which contains a tree touching every group): an upper bound of the optimum.  The push
// items of level 1 offer their vertex's sum (the lanes of the item's group add the row; chunk 0 of
// an item only) to the slot's spider word by atomic-min, and every later round grows only states of
// cost <= half of it (and records only states below it).  The bound only falls, so the level grows
// a superset of what one run under the final bound grows, and every state of cost <= half the
// final bound is exact: the lower bounds of the later levels read level 1 capped there.
template <bool FLT>
__device__ __forceinline__ void run_spider_item(const GrowArgs& a, bool on, uint32_t v, uint32_t pl, uint32_t PL) {
    uint32_t acc = 0, o = 0;   // o: an int sum past 32 bits is only an infinite candidate here
    for (uint32_t j = pl; on && j < a.g; j += PL) acc = add_cost<FLT>(acc, (uint32_t)(a.T1[(uint64_t)v * a.g + j] >> 32), o);
    if (!on) acc = GST_INF_COST;
    for (uint32_t d = 1; d < PL; d <<= 1) acc = add_cost<FLT>(acc, (uint32_t)__shfl_xor(acc, d, WAVE), o);
    if (on && pl == 0 && acc != GST_INF_COST) {
        const unsigned long long w = ((unsigned long long)acc << 32) | v;
        if (w < *(volatile unsigned long long*)a.run_sp) atomicMin(a.run_sp, w);
    }
}
// the bound bits of the running spider word sw, applied to (ub, gmax)
template <bool FLT>
__device__ __forceinline__ void running_spider_bound(unsigned long long sw, uint32_t& ub, uint32_t& gmax) {
    const uint32_t sp = (uint32_t)(sw >> 32);
    if (sp == GST_INF_COST) return;
    uint32_t u, gm;
    spider_bound<FLT>(sp, u, gm);
    ub = min(ub, u);
    gmax = min(gmax, gm);
}

template <bool FLT>
__device__ __forceinline__ void push_items_body(const GrowArgs& a, BlockQueue& q, PushWaveLds& S, const uint64_t* lcur,
                                                const uint64_t* ccur, uint64_t* cnxt, uint32_t* inxt, uint64_t* lnxt,
                                                uint32_t* cntn, uint32_t n_cur, uint32_t gwave, uint32_t nwaves,
                                                uint32_t rub, uint32_t rgmax, bool prune_lb, uint32_t& st_items, uint32_t& st_arcs,
                                                uint32_t& st_relax, uint32_t& st_upd, uint32_t& st_enq,
                                                uint32_t& st_new, uint32_t& st_atom, uint32_t& ovf) {
    const uint32_t lane = lane_id();
    const uint64_t lt = (1ull << lane) - 1ull;
    const uint32_t W = a.W, m = a.m;
    // Level 1 skips the bound check of its masks: its bits are set only by its own rounds (and the
    // members' leaves), for costs <= the cap of the round that set them, which only fall; a running
    // spider can lower the cap meanwhile, and such a state is then grown once more, which is harmless
    const bool filter = rub != GST_INF_COST && !(GST_LEVEL1_NOFILTER && a.Tk == a.T1);
    // items per wave and step: the round's items spread over every wave taking part
    const uint32_t ipw = min((uint32_t)WAVE, max(1u, (n_cur + nwaves - 1) / nwaves));
    for (uint32_t base = gwave * ipw; base < n_cur; base += nwaves * ipw) {
        const uint32_t i = base + lane;
        const bool has = lane < ipw && i < n_cur;
        uint32_t v = 0, e0 = 0, len = 0;
        if (has) {
            const uint64_t it = lcur[i];
            v = (uint32_t)it;
            const uint32_t c = (uint32_t)(it >> 32);
            const uint32_t rs = a.row[v], re = a.row[v + 1];
            e0 = rs + c * a.chunk;
            len = min(re, e0 + a.chunk) - e0;
        }
        st_items += wave_sum32(has ? 1u : 0u);
        for (uint32_t wd = 0; wd < W; ++wd) {
            const uint64_t bm = has ? ccur[(uint64_t)v * W + wd] : 0ull;
            const uint32_t nb = (uint32_t)__popcll(bm);
            uint32_t pin = nb;
#pragma unroll
            for (int d = 1; d < WAVE; d <<= 1) {
                const uint32_t y = __shfl_up(pin, d, WAVE);
                if (lane >= (uint32_t)d) pin += y;
            }
            const uint32_t tot = __shfl(pin, WAVE - 1, WAVE);
            const uint32_t pofs = pin - nb;
            for (uint32_t p0 = 0; p0 < tot; p0 += WAVE) {
                // batch pairs p0 .. p0+63: each lane writes those of its own changed masks
                for (uint32_t t = p0 > pofs ? p0 - pofs : 0u; t < nb && pofs + t < p0 + WAVE; ++t)
                    S.src[pofs + t - p0] = lane | (select_bit(bm, t) << 8);
                wave_lds_sync();
                const uint32_t np = min((uint32_t)WAVE, tot - p0);
                const bool pk = lane < np;
                const uint32_t src = pk ? S.src[lane] : 0u;
                const uint32_t sl = src & 0xFFu;
                const uint32_t pv = __shfl(v, sl, WAVE), pe0 = __shfl(e0, sl, WAVE), plen = __shfl(len, sl, WAVE);
                const uint32_t prk = wd * 64u + (src >> 8);
                uint32_t pcv = GST_INF_COST, lb = 0;
                if (pk) {
                    pcv = (uint32_t)(a.Tk[(uint64_t)pv * m + prk] >> 32);
                    if (filter && prune_lb) lb = label_lb<FLT>(a.T1, a.tour, pv, a.g, a.masks[prk], a.lbcap);
                }
                bool keep = pk && plen != 0 && pcv != GST_INF_COST;
                if (filter) keep = keep && pcv <= rgmax && !reaches_ub<FLT>(pcv, lb, rub);
                const uint64_t kb = __ballot(keep);
                wave_lds_sync();   // every lane has read its src entry before the next batch's writes
                if (kb == 0) continue;
                const uint64_t below = kb & lt;
                const uint32_t kpos = (uint32_t)__popcll(below);
                const uint32_t nk = (uint32_t)__popcll(kb);
                // the pairs of one item are adjacent in batch order: the first kept one counts the
                // item's arcs (the others read the same records again)
                const uint32_t qprev = below ? 63u - (uint32_t)__clzll(below) : 0u;
                const uint32_t prev_sl = __shfl(sl, qprev, WAVE);
                const bool first = below == 0 || prev_sl != sl;
                uint32_t ain = keep ? plen : 0u;
#pragma unroll
                for (int d = 1; d < WAVE; d <<= 1) {
                    const uint32_t y = __shfl_up(ain, d, WAVE);
                    if (lane >= (uint32_t)d) ain += y;
                }
                const uint32_t tot_arcs = __shfl(ain, WAVE - 1, WAVE);
                if (keep) {
                    S.v[kpos] = pv;
                    S.rk[kpos] = prk | (first ? 0x80000000u : 0u);
                    S.cv[kpos] = pcv;
                    S.e0[kpos] = pe0;
                    S.incl[kpos] = ain;
                    S.cut[kpos] = plen;
                }
                wave_lds_sync();
                for (uint32_t r0 = 0; r0 < tot_arcs; r0 += IP_UNR * WAVE) {
                    uint32_t uu[IP_UNR], wt[IP_UNR], xv[IP_UNR], xrk[IP_UNR], xcv[IP_UNR], xp[IP_UNR], xoff[IP_UNR];
                    uint2 rec[IP_UNR];
#pragma unroll
                    for (uint32_t j = 0; j < IP_UNR; ++j) {
                        const uint32_t r = r0 + j * WAVE + lane;
                        xp[j] = 0xFFFFFFFFu;
                        xoff[j] = 0;
                        rec[j] = make_uint2(0xFFFFFFFFu, 0u);
                        xv[j] = 0; xrk[j] = 0; xcv[j] = GST_INF_COST;
                        if (r < tot_arcs) {
                            // the pair of relaxation r: the first kept pair whose prefix exceeds r
                            uint32_t lo = 0, hi = nk - 1;
                            while (lo < hi) {
                                const uint32_t mid = (lo + hi) >> 1;
                                if (S.incl[mid] > r) hi = mid; else lo = mid + 1;
                            }
                            const uint32_t off = r - (lo ? S.incl[lo - 1] : 0u);
                            if (off < S.cut[lo]) {
                                xp[j] = lo;
                                xoff[j] = off;
                                const uint32_t rkf = S.rk[lo];
                                xv[j] = S.v[lo];
                                xrk[j] = rkf & 0x7FFFFFFFu;
                                xcv[j] = S.cv[lo];
                                st_arcs += rkf >> 31;
                                rec[j] = a.arc[S.e0[lo] + off];
                            }
                        }
                    }
                    // arcs of a pair that the bound cuts: every heavier arc of the row is cut too
#pragma unroll
                    for (uint32_t j = 0; j < IP_UNR; ++j) {
                        uu[j] = 0xFFFFFFFFu;
                        wt[j] = rec[j].y;
                        if (xp[j] == 0xFFFFFFFFu) continue;
                        if (rub == GST_INF_COST || !reaches_ub<FLT>(xcv[j], wt[j], rub)) uu[j] = rec[j].x;
                        else atomicMin(&S.cut[xp[j]], xoff[j]);
                    }
                    // the neighbour's state and its change word, loaded together: a bit already set
                    // in this round needs no atomic OR (bits of the next round's words are only ever
                    // set during this round, so a stale read can miss a bit but never show one)
                    unsigned long long cw0[IP_UNR], cwp[IP_UNR];
#pragma unroll
                    for (uint32_t j = 0; j < IP_UNR; ++j) {
                        cw0[j] = uu[j] != 0xFFFFFFFFu ? *(const unsigned long long*)(a.Tk + (uint64_t)uu[j] * m + xrk[j]) : 0ull;
                        cwp[j] = uu[j] != 0xFFFFFFFFu ? *(const unsigned long long*)&cnxt[(uint64_t)uu[j] * W + (xrk[j] >> 6)] : 0ull;
                    }
                    uint32_t cand[IP_UNR];
                    uint32_t mpos = 0, mzero = 0, mgrow = 0;
#pragma unroll
                    for (uint32_t j = 0; j < IP_UNR; ++j) {
                        cand[j] = GST_INF_COST;
                        if (uu[j] == 0xFFFFFFFFu) continue;
                        st_relax++;
                        cand[j] = add_cost<FLT>(xcv[j], wt[j], ovf);
                        const unsigned long long curw = cw0[j];
                        if (cand[j] == GST_INF_COST || cand[j] > (uint32_t)(curw >> 32)) continue;
                        if (rub != GST_INF_COST && reaches_ub<FLT>(cand[j], 0u, rub)) continue;
                        if ((((unsigned long long)cand[j] << 32) | xv[j]) >= curw) continue;
                        if (cand[j] == xcv[j]) mzero |= 1u << j; else mpos |= 1u << j;
                        if (cand[j] <= rgmax && !((cwp[j] >> (xrk[j] & 63u)) & 1ull)) mgrow |= 1u << j;
                    }
                    // the atomics of all IP_UNR relaxations back to back; the change bit beside the
                    // atomic-min (a bit set for a min that lost to a concurrent lower cost only
                    // makes the next round relax the state from the cost it holds)
                    const uint32_t mbit = mpos & mgrow;
                    unsigned long long ob[IP_UNR], old[IP_UNR];
                    uint32_t ru0[IP_UNR], ru1[IP_UNR];
#pragma unroll
                    for (uint32_t j = 0; j < IP_UNR; ++j) {
                        ob[j] = 0ull;
                        ru0[j] = ru1[j] = 0u;
                        if ((mbit >> j) & 1u) {
                            ru0[j] = a.row[uu[j]];
                            ru1[j] = a.row[uu[j] + 1];
                            ob[j] = atomicOr((unsigned long long*)&cnxt[(uint64_t)uu[j] * W + (xrk[j] >> 6)],
                                             1ull << (xrk[j] & 63u));
                            st_atom++;
                        }
                    }
#pragma unroll
                    for (uint32_t j = 0; j < IP_UNR; ++j) {
                        old[j] = 0ull;
                        if ((mpos >> j) & 1u)
                        {
                            old[j] = atomicMin((unsigned long long*)(a.Tk + (uint64_t)uu[j] * m + xrk[j]),
                                               ((unsigned long long)cand[j] << 32) | xv[j]);
                            st_atom++;
                        }
                    }
                    uint32_t enq_n[IP_UNR];
#pragma unroll
                    for (uint32_t j = 0; j < IP_UNR; ++j) {
                        enq_n[j] = 0u;
                        if (((mpos >> j) & 1u) && cand[j] < (uint32_t)(old[j] >> 32)) {
                            st_upd++;
                            st_new += mark_active(a.act, uu[j], old[j]);
                        }
                        const uint64_t bitv = 1ull << (xrk[j] & 63u);
                        if (((mbit >> j) & 1u) && W > 1 && !(ob[j] & bitv)) st_atom++;
                        if (((mbit >> j) & 1u) &&
                            ((W == 1) ? (ob[j] == 0ull) : (!(ob[j] & bitv) && atomicExch(&inxt[uu[j]], 1u) == 0u)))
                            enq_n[j] = (ru1[j] - ru0[j] + a.chunk - 1) / a.chunk;
                    }
                    if (mzero) {
                        // zero-effect edges: strict cost decreases only, by compare-and-swap, so that
                        // equal-cost predecessors can never close a cycle; the bit after a success
#pragma unroll
                        for (uint32_t j = 0; j < IP_UNR; ++j) {
                            if (!((mzero >> j) & 1u)) continue;
                            const uint32_t u = uu[j];
                            unsigned long long* addr = (unsigned long long*)(a.Tk + (uint64_t)u * m + xrk[j]);
                            const unsigned long long nw = ((unsigned long long)cand[j] << 32) | xv[j];
                            unsigned long long curw = cw0[j];
                            bool improved = false;
                            while (cand[j] < (uint32_t)(curw >> 32)) {
                                st_atom++;
                                const unsigned long long o = atomicCAS(addr, curw, nw);
                                if (o == curw) { improved = true; st_new += mark_active(a.act, u, o); break; }
                                curw = o;
                            }
                            if (!improved) continue;
                            st_upd++;
                            if (cand[j] > rgmax) continue;
                            const uint64_t bitv = 1ull << (xrk[j] & 63u);
                            const uint32_t r0u = a.row[u], r1u = a.row[u + 1];
                            const unsigned long long o2 =
                                atomicOr((unsigned long long*)&cnxt[(uint64_t)u * W + (xrk[j] >> 6)], bitv);
                            st_atom += (W > 1 && !(o2 & bitv)) ? 2u : 1u;
                            if ((W == 1) ? (o2 == 0ull) : (!(o2 & bitv) && atomicExch(&inxt[u], 1u) == 0u))
                                enq_n[j] = (r1u - r0u + a.chunk - 1) / a.chunk;
                        }
                    }
                    bq_push_multi<IP_UNR>(q, lnxt, cntn, uu, enq_n, st_enq);
                }
                wave_lds_sync();   // the kept pairs are read before the next batch rewrites them
            }
        }
    }
}

// ---------------------------------------------------------------------------------------
// grow: one round of frontier relaxation for all masks of one level
// ---------------------------------------------------------------------------------------
template <bool FLT>
__device__ __forceinline__ void grow_round_body(const GrowArgs& a, uint32_t bid, uint32_t nblk, uint32_t seq) {
    __shared__ BlockQueue q;
    __shared__ PushWaveLds s_pw[GROW_WAVES];
    __shared__ unsigned long long s_stats[ST_NSTAT];
    __shared__ uint32_t s_ovf;
    RoundCtl* ctl = a.ctl;
    if (ctl->level_done) return;
    // the round's bound: the level's own, lowered on level 1 of an automatically bounded solve by the
    // running spider as it stood when the round started
    uint32_t rub = a.ub, rgmax = a.gmax;
    if (a.run_sp) running_spider_bound<FLT>(*a.run_sp, rub, rgmax);
    // (round, seq) is one 8-byte word: a workgroup that starts only after this launch's round has
    // ended (its participants all arrived and advanced the round) sees its own seq and stays out
    const unsigned long long rsq =
        __hip_atomic_load(&ctl->rsq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if ((uint32_t)(rsq >> 32) == seq) return;
    const uint32_t r = (uint32_t)rsq;
    const uint32_t cur = r % 3, nxt = (r + 1) % 3, prv = (r + 2) % 3;
    const uint32_t n_cur = ctl->cnt[cur];
    const uint32_t n_prv = ctl->cnt[prv];
    const uint32_t lane = lane_id();
    const uint32_t W = a.W, m = a.m;
    const bool pull = n_cur > 0 && n_cur >= a.pull_thresh && a.all_items != nullptr;
    // only as many workgroups as the round needs take part (a sparse round costs a few blocks)
    uint32_t part = nblk;
    if (!pull) {
        const uint32_t per_blk = GROW_WAVES * push_items_per_wave(round_push_lanes(a.push_lanes, a.push_lanes_sparse, n_cur));   // items per workgroup step
        const uint32_t need_b = (n_cur + per_blk - 1) / per_blk;
        const uint32_t need_a = (n_prv + GROW_BLOCK - 1) / GROW_BLOCK;
        part = max(1u, min(nblk, max(need_b, need_a)));
    }
    if (a.rlog && bid == 0 && threadIdx.x == 0 && r < RLOG_ROUNDS) {
        a.rlog[4 * r + 0] = n_cur | (pull ? 0x80000000u : 0u);
        a.rlog[4 * r + 1] = part;
        a.rlog[4 * r + 2] = (uint32_t)wall_clock64();
    }
    if (bid >= part) return;
    bq_init(q);
    if (threadIdx.x < ST_NSTAT) s_stats[threadIdx.x] = 0;
    if (threadIdx.x == 0) s_ovf = 0;

    // phase A: clear the change bitmaps / queue flags of round r-1's items
    {
        const uint64_t* lp = a.list[prv];
        uint64_t* cp = a.chg[prv];
        uint32_t* ip = a.inq[prv];
        for (uint32_t i = bid * blockDim.x + threadIdx.x; i < n_prv; i += part * blockDim.x) {
            const uint64_t it = lp[i];
            if ((it >> 32) == 0) {
                const uint32_t v = (uint32_t)it;
                for (uint32_t w = 0; w < W; ++w) cp[(uint64_t)v * W + w] = 0;
                ip[v] = 0;
            }
        }
    }
    __syncthreads();

    // phase B: relax the items of round r
    uint32_t st_items = 0, st_arcs = 0, st_relax = 0, st_upd = 0, st_enq = 0, st_new = 0, st_atom = 0;
    uint32_t ovf = 0;
    const uint32_t gwave = bid * GROW_WAVES + (threadIdx.x / WAVE);
    const uint32_t nwaves = part * GROW_WAVES;
    const uint64_t* lcur = a.list[cur];
    const uint64_t* ccur = a.chg[cur];
    uint64_t* cnxt = a.chg[nxt];
    uint32_t* inxt = a.inq[nxt];
    uint64_t* lnxt = a.list[nxt];
    uint32_t* cntn = &ctl->cnt[nxt];
    const bool prune_lb = a.use_lb && rub != GST_INF_COST;

    // ---- pull round, light rows: each lane group owns one row and each lane one mask of it ----
    // (tropical SpMM row-split: the lane scans the whole row, so no cross-lane reduction, and the
    // row has a single writer, so plain stores and a ballot-built change word replace atomics)
    if (pull) {
        const uint32_t mchunk = m < (uint32_t)WAVE ? m : (uint32_t)WAVE;
        const uint32_t G = WAVE / mchunk;                 // rows per wave
        const uint32_t gslot = lane / mchunk, t0 = lane % mchunk;
        const bool slot_ok = gslot < G;
        for (uint32_t base = gwave * G; base < a.n; base += nwaves * G) {
            const uint32_t u = base + gslot;
            const bool row_ok = slot_ok && u < a.n;
            uint32_t rs = 0, re = 0;
            if (row_ok) { rs = a.row[u]; re = a.row[u + 1]; }
            const bool light = row_ok && re > rs && re - rs <= HUB_DEG;
            const uint32_t deg = light ? re - rs : 0u;
            st_items += wave_sum32(light && t0 == 0 ? 1u : 0u);
            uint32_t maxdeg = deg;
#pragma unroll
            for (int d = 1; d < WAVE; d <<= 1) maxdeg = max(maxdeg, (uint32_t)__shfl_xor(maxdeg, d, WAVE));
            for (uint32_t mb = 0; mb < m; mb += mchunk) {
                const uint32_t rk = mb + t0;
                const bool lane_ok = light && rk < m;
                unsigned long long* addr = (unsigned long long*)(a.Tk + (uint64_t)u * m + rk);
                const unsigned long long own = lane_ok ? *addr : 0ull;
                const uint32_t ocost = (uint32_t)(own >> 32);
                const uint32_t lb = (prune_lb && lane_ok) ? label_lb<FLT>(a.T1, a.tour, u, a.g, a.masks[rk], a.lbcap) : 0u;
                unsigned long long bs = GST_INF_WORD, bz = GST_INF_WORD;
                for (uint32_t s = 0; s < maxdeg; s += UNR) {
                    uint32_t nb[UNR], wt[UNR], cn[UNR];
                    bool live[UNR];
#pragma unroll
                    for (uint32_t j = 0; j < UNR; ++j) {
                        const bool ok = lane_ok && s + j < deg;
                        const uint2 r = ok ? a.arc[rs + s + j] : make_uint2(0xFFFFFFFFu, 0u);
                        wt[j] = r.y;
                        live[j] = ok && pull_arc_live<FLT>(wt[j], ocost, lb, rub);
                        nb[j] = live[j] ? r.x : 0xFFFFFFFFu;
                        st_arcs += (ok && t0 == 0) ? 1u : 0u;   // arcs read (once per row, lane t0 = 0)
                    }
                    // rows run in increasing weight: when the step's first arc is dead for every
                    // lane, the rest of every row is dead too
                    if (__ballot(live[0]) == 0) break;
#pragma unroll
                    for (uint32_t j = 0; j < UNR; ++j)
                        cn[j] = nb[j] != 0xFFFFFFFFu ? (uint32_t)(a.Tk[(uint64_t)nb[j] * m + rk] >> 32) : GST_INF_COST;
#pragma unroll
                    for (uint32_t j = 0; j < UNR; ++j) {
                        if (nb[j] == 0xFFFFFFFFu) continue;
                        st_relax++;
                        if (cn[j] < ocost && cn[j] <= rgmax) {
                            const uint32_t cand = add_cost<FLT>(cn[j], wt[j], ovf);
                            if (cand != GST_INF_COST && !(rub != GST_INF_COST && reaches_ub<FLT>(cand, lb, rub))) {
                                const unsigned long long nw = ((unsigned long long)cand << 32) | nb[j];
                                if (cand > cn[j]) bs = nw < bs ? nw : bs;
                                else if (cand < ocost) bz = nw < bz ? nw : bz;
                            }
                        }
                    }
                }
                // single writer: zero-effect candidates only on a strict decrease, strict ones by min
                unsigned long long nwv = own;
                if (lane_ok) {
                    if (bz != GST_INF_WORD && (uint32_t)(bz >> 32) < (uint32_t)(bs >> 32)) nwv = bz;
                    else if (bs < own) nwv = bs;
                }
                const bool improved = lane_ok && (uint32_t)(nwv >> 32) < ocost;
                if (lane_ok && nwv != own) *addr = nwv;
                if (improved) st_new += mark_active(a.act, u, own);
                const uint64_t lmask = mchunk == 64 ? ~0ull : ((1ull << mchunk) - 1);
                const uint64_t imp = __ballot(improved);
                // only states that may still grow go on the frontier
                const uint64_t fl = __ballot(improved && (uint32_t)(nwv >> 32) <= rgmax);
                if (a.run_sp) {
                    // level 1 (one mask per group: the whole row in the lane group): a row that
                    // changed offers its vertex's sum of group distances to the running spider
                    const uint32_t cst = lane_ok ? (uint32_t)(nwv >> 32) : GST_INF_COST;
                    uint32_t acc = 0, o = 0;
                    for (uint32_t j = 0; j < mchunk; ++j)
                        acc = add_cost<FLT>(acc, (uint32_t)__shfl(cst, (int)(gslot * mchunk + j), WAVE), o);
                    const bool row_imp = slot_ok && ((imp >> (gslot * mchunk)) & lmask) != 0;
                    if (row_imp && t0 == 0 && acc != GST_INF_COST) {
                        const unsigned long long w = ((unsigned long long)acc << 32) | u;
                        if (w < *(volatile unsigned long long*)a.run_sp) atomicMin(a.run_sp, w);
                    }
                }
                uint32_t enq_v = 0, enq_n = 0;
                if (lane_ok && t0 == 0) {
                    st_upd += (uint32_t)__popcll((imp >> (gslot * mchunk)) & lmask);
                    const uint64_t word = (fl >> (gslot * mchunk)) & lmask;
                    if (word) {
                        // single writer of u's change word in a pull round; the flag marks the
                        // first masks block that changed
                        cnxt[(uint64_t)u * W + (mb >> 6)] = word;
                        st_atom += W > 1 ? 1u : 0u;
                        const bool first = (W == 1) ? true : (atomicExch(&inxt[u], 1u) == 0u);
                        if (first) { enq_v = u; enq_n = (deg + a.chunk - 1) / a.chunk; }
                    }
                }
                bq_push(q, lnxt, cntn, enq_v, enq_n, st_enq);
            }
        }
    }

    // ---- pull round, hub rows: (row, chunk) items, lanes = (arc slot, mask), combined by atomics ----
    const uint32_t n_all = pull ? (a.n_all_dev ? *a.n_all_dev : a.n_all) : 0u;
    for (uint32_t i = gwave; pull && i < n_all; i += nwaves) {
        const uint64_t it = a.all_items[i];
        const uint32_t u = (uint32_t)it;
        const uint32_t c = (uint32_t)(it >> 32);
        const uint32_t rs = a.row[u], re = a.row[u + 1];
        const uint32_t e0 = rs + c * a.chunk;
        const uint32_t e1 = min(re, e0 + a.chunk);
        st_items++;
        for (uint32_t mb = 0; mb < m; mb += WAVE) {
            const uint32_t na = min((uint32_t)WAVE, m - mb);
            const uint32_t A = WAVE / na;
            const uint32_t t = lane % na, s_arc = lane / na;
            const bool lane_ok = s_arc < A;
            const uint32_t rk = mb + t;
            unsigned long long* addr = (unsigned long long*)(a.Tk + (uint64_t)u * m + rk);
            const unsigned long long own = *addr;
            const uint32_t ocost = (uint32_t)(own >> 32);
            const uint32_t lb = prune_lb ? label_lb<FLT>(a.T1, a.tour, u, a.g, a.masks[rk], a.lbcap) : 0u;
            unsigned long long bs = GST_INF_WORD, bz = GST_INF_WORD;  // best via positive / zero-effect edges
            const uint32_t len = e1 - e0;
            const uint32_t steps = (len + A - 1) / A;
            for (uint32_t s = 0; s < steps; s += UNR) {
                uint32_t nb[UNR], wt[UNR], cn[UNR];
                bool live[UNR];
#pragma unroll
                for (uint32_t j = 0; j < UNR; ++j) {
                    const uint32_t e = e0 + (s + j) * A + s_arc;
                    const bool ok = lane_ok && e < e1;
                    const uint2 r = ok ? a.arc[e] : make_uint2(0xFFFFFFFFu, 0u);
                    wt[j] = r.y;
                    live[j] = ok && pull_arc_live<FLT>(wt[j], ocost, lb, rub);
                    nb[j] = live[j] ? r.x : 0xFFFFFFFFu;
                    st_arcs += (ok && t == 0) ? 1u : 0u;   // arcs read (once per arc slot, mask lane 0)
                }
                if (__ballot(live[0]) == 0) break;   // each lane's arcs run in increasing weight
#pragma unroll
                for (uint32_t j = 0; j < UNR; ++j)
                    cn[j] = nb[j] != 0xFFFFFFFFu ? (uint32_t)(a.Tk[(uint64_t)nb[j] * m + rk] >> 32) : GST_INF_COST;
#pragma unroll
                for (uint32_t j = 0; j < UNR; ++j) {
                    if (nb[j] == 0xFFFFFFFFu) continue;
                    st_relax++;
                    // only a cheaper neighbour can lower (or tie) our cost, and only one still grown
                    if (cn[j] < ocost && cn[j] <= rgmax) {
                        const uint32_t cand = add_cost<FLT>(cn[j], wt[j], ovf);
                        if (cand != GST_INF_COST && !(rub != GST_INF_COST && reaches_ub<FLT>(cand, lb, rub))) {
                            const unsigned long long nw = ((unsigned long long)cand << 32) | nb[j];
                            if (cand > cn[j]) bs = nw < bs ? nw : bs;
                            else if (cand < ocost) bz = nw < bz ? nw : bz;
                        }
                    }
                }
            }
            for (uint32_t k2 = 1; k2 < A; k2 <<= 1) {
                const unsigned long long y1 = __shfl_down(bs, k2 * na, WAVE);
                const unsigned long long y2 = __shfl_down(bz, k2 * na, WAVE);
                if ((s_arc % (2 * k2)) == 0 && s_arc + k2 < A) {
                    bs = y1 < bs ? y1 : bs;
                    bz = y2 < bz ? y2 : bz;
                }
            }
            uint32_t enq_v = 0, enq_n = 0;
            if (s_arc == 0) {
                bool improved = false;
                uint32_t ncost = GST_INF_COST;
                if (bz != GST_INF_WORD && (uint32_t)(bz >> 32) < (uint32_t)(bs >> 32)) {
                    // a zero-effect edge wins: accept only a strict cost decrease
                    const uint32_t cz = (uint32_t)(bz >> 32);
                    unsigned long long curw = own;
                    while (cz < (uint32_t)(curw >> 32)) {
                        st_atom++;
                        const unsigned long long old = atomicCAS(addr, curw, bz);
                        if (old == curw) { improved = true; st_new += mark_active(a.act, u, old); break; }
                        curw = old;
                    }
                    ncost = cz;
                } else if (bs < own) {
                    st_atom++;
                    const unsigned long long old = atomicMin(addr, bs);
                    improved = (uint32_t)(bs >> 32) < (uint32_t)(old >> 32);
                    if (improved) st_new += mark_active(a.act, u, old);
                    ncost = (uint32_t)(bs >> 32);
                }
                if (improved) st_upd++;
                if (improved && ncost <= rgmax) {
                    const uint64_t bitv = 1ull << (rk & 63);
                    unsigned long long* cw = (unsigned long long*)&cnxt[(uint64_t)u * W + (rk >> 6)];
                    const unsigned long long ob = atomicOr(cw, (unsigned long long)bitv);
                    st_atom += (W > 1 && !(ob & bitv)) ? 2u : 1u;
                    const bool first = (W == 1) ? (ob == 0) : (!(ob & bitv) && atomicExch(&inxt[u], 1u) == 0u);
                    if (first) {
                        enq_v = u;
                        enq_n = (re - rs + a.chunk - 1) / a.chunk;
                    }
                }
            }
            bq_push(q, lnxt, cntn, enq_v, enq_n, st_enq);
        }
    }

    // ---- push round: queued (vertex, chunk) items relax their changed masks into their
    // neighbours.  A wave serves PG items at once, one group of PL lanes each, so that the chains
    // of dependent loads of several items are in flight together; inside a group the lanes are
    // (mask, arc slot) pairs over blocks of up to PL masks. ----
    const uint32_t rpl = round_push_lanes(a.push_lanes, a.push_lanes_sparse, n_cur);
    const bool item_par = rpl == PUSH_ITEM_PAR;
    if (GST_ITEMPAR_CODE && !pull && item_par)
        push_items_body<FLT>(a, q, s_pw[threadIdx.x / WAVE], lcur, ccur, cnxt, inxt, lnxt, cntn, n_cur, gwave, nwaves,
                             rub, rgmax, prune_lb, st_items, st_arcs, st_relax, st_upd, st_enq, st_new, st_atom, ovf);
    const uint32_t PL = item_par ? (uint32_t)WAVE : rpl, PG = WAVE / PL;   // PL: 8, 16, 32 or 64
    const uint32_t pgrp = lane / PL, pl = lane % PL;
    for (uint32_t base = gwave * PG; !pull && !item_par && base < n_cur; base += nwaves * PG) {
        const uint32_t i = base + pgrp;
        const bool has = i < n_cur;
        uint32_t v = 0, e0 = 0, e1 = 0, c = 0;
        if (has) {
            const uint64_t it = lcur[i];
            v = (uint32_t)it;
            c = (uint32_t)(it >> 32);
            const uint32_t rs = a.row[v], re = a.row[v + 1];
            e0 = rs + c * a.chunk;
            e1 = min(re, e0 + a.chunk);
        }
        st_items += wave_sum32(has && pl == 0 ? 1u : 0u);
        if (a.run_sp) run_spider_item<FLT>(a, has && c == 0, v, pl, PL);
        for (uint32_t wd = 0; wd < W; ++wd) {
            uint64_t bm = has ? ccur[(uint64_t)v * W + wd] : 0ull;
            // drop masks whose cost (plus lower bound) reaches the bound.  Level 1 skips the check:
            // its bits are set only by its own rounds (and the members' leaves), for costs <= the cap
            // of the round that set them, which only fall (no lower bound there; a running spider can
            // lower the cap meanwhile, and such a state is then grown once more, which is harmless)
            if (rub != GST_INF_COST && !(GST_LEVEL1_NOFILTER && a.Tk == a.T1)) {
                const uint32_t na0 = (uint32_t)__popcll(bm);
                uint64_t keep = 0;
                for (uint32_t t0 = pl; t0 < na0; t0 += PL) {
                    const uint32_t bit = select_bit(bm, t0);
                    const uint32_t rk = wd * 64 + bit;
                    const uint32_t cv = (uint32_t)(a.Tk[(uint64_t)v * m + rk] >> 32);
                    const uint32_t lb = prune_lb ? label_lb<FLT>(a.T1, a.tour, v, a.g, a.masks[rk], a.lbcap) : 0u;
                    if (cv <= rgmax && !reaches_ub<FLT>(cv, lb, rub)) keep |= 1ull << bit;
                }
                bm = group_or64(keep, PL);
            }
            const uint32_t na = (uint32_t)__popcll(bm);
            const uint32_t na_max = wave_max32(na);
            for (uint32_t mb = 0; mb < na_max; mb += PL) {
                const uint32_t nm = na > mb ? min(PL, na - mb) : 0u;
                const uint32_t A = nm ? PL / nm : 0u;   // arc slots per step
                const uint32_t t = nm ? pl % nm : 0u, s_arc = nm ? pl / nm : 0u;
                const bool lane_ok = nm != 0 && s_arc < A;
                const uint32_t rk = wd * 64 + (lane_ok ? select_bit(bm, mb + t) : 0u);
                const uint32_t cv = lane_ok ? (uint32_t)(a.Tk[(uint64_t)v * m + rk] >> 32) : GST_INF_COST;
                const uint64_t bitv = 1ull << (rk & 63);
                const uint32_t wsel = rk >> 6;
                const uint32_t len = e1 - e0;
                const uint32_t steps = wave_max32(lane_ok ? (len + A - 1) / A : 0u);
                for (uint32_t s = 0; s < steps; s += UNR) {
                    uint32_t uu[UNR], wt[UNR];
                    unsigned long long cw0[UNR];
                    bool live[UNR];
#pragma unroll
                    for (uint32_t j = 0; j < UNR; ++j) {
                        const uint32_t e = e0 + (s + j) * A + s_arc;
                        // the arc load waits only for the item, not for the own cost (both in flight)
                        const bool in_row = lane_ok && e < e1;
                        const bool ok = in_row && cv != GST_INF_COST;
                        const uint2 r = in_row ? a.arc[e] : make_uint2(0xFFFFFFFFu, 0u);
                        wt[j] = r.y;
                        // under a bound, cv + w >= ub holds for every heavier arc of the row as well
                        live[j] = ok && (rub == GST_INF_COST || !reaches_ub<FLT>(cv, wt[j], rub));
                        uu[j] = live[j] ? r.x : 0xFFFFFFFFu;
                        st_arcs += (in_row && t == 0) ? 1u : 0u;   // arcs read (once per arc slot, mask lane 0)
                    }
                    if (__ballot(live[0]) == 0) break;
#pragma unroll
                    for (uint32_t j = 0; j < UNR; ++j)
                        cw0[j] = uu[j] != 0xFFFFFFFFu ? *(const unsigned long long*)(a.Tk + (uint64_t)uu[j] * m + rk) : 0ull;
#if GST_PUSH_BATCH_ATOMICS
                    // Every arc's update is decided first and the atomics of all UNR arcs are then
                    // issued back to back, so that their round trips overlap instead of queueing
                    // one behind the other.  A neighbour's change bit is set beside its atomic-min,
                    // before the outcome is known: when the min loses to a concurrent lower cost,
                    // the bit only makes the next round relax (u, X) from the cost it holds, which
                    // is always a valid step; every state that was lowered still gets its bit.
                    uint32_t cand[UNR];
                    uint32_t mpos = 0, mzero = 0, mgrow = 0;
#pragma unroll
                    for (uint32_t j = 0; j < UNR; ++j) {
                        cand[j] = GST_INF_COST;
                        if (uu[j] == 0xFFFFFFFFu) continue;
                        st_relax++;
                        cand[j] = add_cost<FLT>(cv, wt[j], ovf);
                        const unsigned long long curw = cw0[j];
                        if (cand[j] == GST_INF_COST || cand[j] > (uint32_t)(curw >> 32)) continue;
                        if (rub != GST_INF_COST && reaches_ub<FLT>(cand[j], 0u, rub)) continue;
                        if ((((unsigned long long)cand[j] << 32) | v) >= curw) continue;
                        // cand == cv: an effective zero-weight edge (w == 0, or a float sum that
                        // rounds back to cv), taken only on a strict cost decrease below
                        if (cand[j] == cv) mzero |= 1u << j; else mpos |= 1u << j;
                        if (cand[j] <= rgmax) mgrow |= 1u << j;
                    }
                    const uint32_t mbit = mpos & mgrow;
                    unsigned long long ob[UNR], old[UNR];
                    uint32_t ru0[UNR], ru1[UNR];
#pragma unroll
                    for (uint32_t j = 0; j < UNR; ++j) {
                        ob[j] = 0ull;
                        ru0[j] = ru1[j] = 0u;
                        if ((mbit >> j) & 1u) {
                            ru0[j] = a.row[uu[j]];
                            ru1[j] = a.row[uu[j] + 1];
                            ob[j] = atomicOr((unsigned long long*)&cnxt[(uint64_t)uu[j] * W + wsel], (unsigned long long)bitv);
                            st_atom++;
                        }
                    }
#pragma unroll
                    for (uint32_t j = 0; j < UNR; ++j) {
                        old[j] = 0ull;
                        if ((mpos >> j) & 1u) {
                            old[j] = atomicMin((unsigned long long*)(a.Tk + (uint64_t)uu[j] * m + rk),
                                               ((unsigned long long)cand[j] << 32) | v);
                            st_atom++;
                        }
                    }
                    uint32_t enq_n[UNR];
#pragma unroll
                    for (uint32_t j = 0; j < UNR; ++j) {
                        enq_n[j] = 0u;
                        if (((mpos >> j) & 1u) && cand[j] < (uint32_t)(old[j] >> 32)) {
                            st_upd++;
                            st_new += mark_active(a.act, uu[j], old[j]);
                        }
                        // first change of u in this round: queue it (one word per vertex: the word
                        // itself was empty; several words: u's queue flag)
                        if (((mbit >> j) & 1u) && W > 1 && !(ob[j] & bitv)) st_atom++;
                        if (((mbit >> j) & 1u) &&
                            ((W == 1) ? (ob[j] == 0ull) : (!(ob[j] & bitv) && atomicExch(&inxt[uu[j]], 1u) == 0u)))
                            enq_n[j] = (ru1[j] - ru0[j] + a.chunk - 1) / a.chunk;
                    }
                    if (mzero) {
                        // zero-effect edges: strict cost decreases only, by compare-and-swap, so that
                        // equal-cost predecessors can never close a cycle; the bit after a success
#pragma unroll
                        for (uint32_t j = 0; j < UNR; ++j) {
                            if (!((mzero >> j) & 1u)) continue;
                            const uint32_t u = uu[j];
                            unsigned long long* addr = (unsigned long long*)(a.Tk + (uint64_t)u * m + rk);
                            const unsigned long long nw = ((unsigned long long)cand[j] << 32) | v;
                            unsigned long long curw = cw0[j];
                            bool improved = false;
                            while (cand[j] < (uint32_t)(curw >> 32)) {
                                st_atom++;
                                const unsigned long long o = atomicCAS(addr, curw, nw);
                                if (o == curw) { improved = true; st_new += mark_active(a.act, u, o); break; }
                                curw = o;
                            }
                            if (!improved) continue;
                            st_upd++;
                            if (cand[j] > rgmax) continue;
                            const uint32_t r0 = a.row[u], r1 = a.row[u + 1];
                            const unsigned long long o2 = atomicOr((unsigned long long*)&cnxt[(uint64_t)u * W + wsel],
                                                                   (unsigned long long)bitv);
                            st_atom += (W > 1 && !(o2 & bitv)) ? 2u : 1u;
                            if ((W == 1) ? (o2 == 0ull) : (!(o2 & bitv) && atomicExch(&inxt[u], 1u) == 0u))
                                enq_n[j] = (r1 - r0 + a.chunk - 1) / a.chunk;
                        }
                    }
                    bq_push_multi<UNR>(q, lnxt, cntn, uu, enq_n, st_enq);
#else
#pragma unroll
                    for (uint32_t j = 0; j < UNR; ++j) {
                        uint32_t enq_v = 0, enq_n = 0;
                        const uint32_t u = uu[j];
                        if (u != 0xFFFFFFFFu) {
                            st_relax++;
                            const uint32_t cand = add_cost<FLT>(cv, wt[j], ovf);
                            unsigned long long curw = cw0[j];
                            if (cand != GST_INF_COST && cand <= (uint32_t)(curw >> 32) &&
                                (rub == GST_INF_COST || !reaches_ub<FLT>(cand, 0u, rub))) {
                                unsigned long long* addr = (unsigned long long*)(a.Tk + (uint64_t)u * m + rk);
                                const unsigned long long nw = ((unsigned long long)cand << 32) | v;
                                bool improved = false;
                                if (nw < curw) {
                                    if (cand == cv) {
                                        // effective zero-weight edge: only strict cost decreases, so
                                        // equal-cost predecessors can never close a cycle
                                        while (cand < (uint32_t)(curw >> 32)) {
                                            st_atom++;
                                            const unsigned long long old = atomicCAS(addr, curw, nw);
                                            if (old == curw) { improved = true; st_new += mark_active(a.act, u, old); break; }
                                            curw = old;
                                        }
                                    } else {
                                        const unsigned long long old = atomicMin(addr, nw);
                                        st_atom++;
                                        improved = cand < (uint32_t)(old >> 32);
                                        if (improved) st_new += mark_active(a.act, u, old);
                                    }
                                }
                                if (improved) st_upd++;
                                if (improved && cand <= rgmax) {
                                    unsigned long long* cwp = (unsigned long long*)&cnxt[(uint64_t)u * W + wsel];
                                    // u's row bounds are read beside the atomic OR, not after it
                                    const uint32_t ru0 = a.row[u], ru1 = a.row[u + 1];
                                    if (!GST_CHG_PREREAD || !(*cwp & bitv)) {
                                        const unsigned long long ob = atomicOr(cwp, (unsigned long long)bitv);
                                        st_atom += (W > 1 && !(ob & bitv)) ? 2u : 1u;
                                        // first change of u in this round: queue it (one word per
                                        // vertex: the word itself was empty; several words: flag)
                                        const bool first = (W == 1) ? (ob == 0) : (!(ob & bitv) && atomicExch(&inxt[u], 1u) == 0u);
                                        if (first) {
                                            enq_v = u;
                                            enq_n = (ru1 - ru0 + a.chunk - 1) / a.chunk;
                                        }
                                    }
                                }
                            }
                        }
                        bq_push(q, lnxt, cntn, enq_v, enq_n, st_enq);
                    }
#endif
                }
            }
        }
    }
    bq_flush(q, lnxt, cntn);

    // statistics: LDS reduction, one global atomic per workgroup and counter
    {
        const uint32_t si = wave_sum32(lane == 0 ? st_items : 0u);
        const uint32_t sa = wave_sum32(st_arcs);
        const uint32_t sr = wave_sum32(st_relax);
        const uint32_t su = wave_sum32(st_upd);
        const uint32_t se = wave_sum32(st_enq);
        const uint32_t sn = wave_sum32(st_new);
        const uint32_t sx = wave_sum32(st_atom);
        const uint32_t so = wave_sum32(ovf);
        if (lane == 0) {
            if (sn) atomicAdd(&s_stats[ST_NEW], (unsigned long long)sn);
            if (sx) atomicAdd(&s_stats[ST_ATOM], (unsigned long long)sx);
            if (si) atomicAdd(&s_stats[ST_ITEMS], (unsigned long long)si);
            if (sa) atomicAdd(&s_stats[ST_ARCS], (unsigned long long)sa);
            if (sr) atomicAdd(&s_stats[ST_RELAX], (unsigned long long)sr);
            if (su) atomicAdd(&s_stats[ST_UPD], (unsigned long long)su);
            if (se) atomicAdd(&s_stats[ST_ENQ], (unsigned long long)se);
            if (so) s_ovf = 1;
        }
        __syncthreads();
        if (threadIdx.x < ST_MVERT && s_stats[threadIdx.x]) atomicAdd(&stat_stripe(a.stats, bid)[threadIdx.x], s_stats[threadIdx.x]);
        if (threadIdx.x == 0 && s_ovf) atomicOr(&ctl->err, 1u);
    }

    // end of round: the last workgroup to arrive (two-level count: groups of 32, then groups)
    // resets the counter of the list cleaned in phase A and advances the round, or marks the level
    // finished when this round had no items.  The decision uses only values every workgroup read
    // at kernel start, and the next launch sees these stores at the kernel boundary, so no
    // device-scope fence (an L2 write-back on this chip) is needed here.
    if (threadIdx.x == 0) {
        const uint32_t grp = bid >> 5;
        const uint32_t gsize = min(32u, part - (grp << 5));
        if (atomicAdd(&ctl->sub[grp], 1u) == gsize - 1) {
            ctl->sub[grp] = 0;
            const uint32_t ngrp = (part + 31) >> 5;
            if (atomicAdd(&ctl->top, 1u) == ngrp - 1) {
                ctl->top = 0;
                if (a.rlog && r < RLOG_ROUNDS) a.rlog[4 * r + 3] = (uint32_t)wall_clock64();
                if (n_cur == 0) {
                    ctl->level_done = 1;
                    __hip_atomic_store(&ctl->rsq, ((unsigned long long)seq << 32) | r,
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                } else {
                    ctl->cnt[prv] = 0;
                    __hip_atomic_store(&ctl->rsq, ((unsigned long long)seq << 32) | (r + 1),
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
            }
        }
    }
}

// 5 workgroups of 4 waves per CU must fit at once (the grid is sized for it): at most 96 VGPRs
template <bool FLT>
__global__ void __launch_bounds__(GROW_BLOCK) __attribute__((amdgpu_waves_per_eu(GST_GROW_WPE))) k_grow_round(GrowArgs a) {
    grow_round_body<FLT>(a, blockIdx.x, gridDim.x, a.seq);
}

// ---------------------------------------------------------------------------------------
// grow, asynchronous: one persistent launch finishes a level from its current frontier with a
// device work queue instead of rounds.  Tasks:
//   VERTEX v          take v's change bitmap (atomic exchange) and relax its arcs for those masks,
//                     or, for a row longer than one chunk, split it into CHUNK tasks
//   CHUNK v,c,wd,bm   relax arcs [c*chunk, ...) of v for the masks in word wd of bm
// Task index i < n0 is item i of the current round list (chunk-0 items are VERTEX tasks, the rest
// are no-ops: the vertex task covers the row); later tasks live in a ring of 4-word slots, every
// word tagged with the task index, so a reader knows a slot is complete without any fence.
// pending counts tasks reserved and not finished: it is raised before a task becomes visible
// and lowered after the task's own enqueues, so pending == 0 means the level is finished.
// Every wait is bounded: past the bound the launch reports an error instead of spinning on.
// ---------------------------------------------------------------------------------------
constexpr uint32_t TASK_VERTEX = 0, TASK_CHUNK = 1;
constexpr uint32_t ASYNC_BATCH = 4;           // tasks claimed per wave and claim
constexpr uint32_t ASYNC_MAX_WAIT = 1u << 21; // polls of one slot before giving up (~seconds)

__device__ __forceinline__ unsigned long long ld_agent(const unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(unsigned long long* p, unsigned long long v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t ld_agent32(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// write one task into the ring (caller reserved index idx >= n0)
__device__ __forceinline__ void task_write(const GrowArgs& a, uint32_t n0, uint32_t idx, uint32_t kind, uint32_t v,
                                           uint32_t c, uint32_t wd, uint64_t bm) {
    unsigned long long* sl = a.ring + 4ull * ((idx - n0) & a.ring_mask);
    const unsigned long long tag = (unsigned long long)idx << 32;
    st_agent(sl + 0, tag | v);
    st_agent(sl + 1, tag | ((unsigned long long)c << 9) | ((unsigned long long)wd << 1) | kind);
    if (kind == TASK_CHUNK) {
        st_agent(sl + 2, tag | (bm & 0xFFFFFFFFull));
        st_agent(sl + 3, tag | (bm >> 32));
    }
}

// relax arcs [e0,e1) of v for the masks set in word wd of bm; queue improved neighbours.
// Convergent: all lanes of the wave call it with the same arguments.
template <bool FLT>
__device__ void async_relax(const GrowArgs& a, uint32_t n0, uint32_t v, uint32_t e0, uint32_t e1, uint32_t wd,
                           uint64_t bm, uint32_t& st_relax, uint32_t& st_upd, uint32_t& st_enq, uint32_t& ovf,
                           uint32_t& st_new, uint32_t& st_arcs) {
    const uint32_t lane = lane_id();
    const uint32_t m = a.m, W = a.W;
    uint64_t* chg = a.chg[a.ctl->round % 3];
    uint32_t* inq = a.inq[a.ctl->round % 3];
    const bool prune_lb = a.use_lb && a.ub != GST_INF_COST;
    // drop masks at or above the bound
    if (a.ub != GST_INF_COST) {
        const uint32_t na0 = (uint32_t)__popcll(bm);
        uint64_t keep = 0;
        if (lane < na0) {
            const uint32_t bit = select_bit(bm, lane);
            const uint32_t rk = wd * 64 + bit;
            const uint32_t cv = (uint32_t)(ld_agent((const unsigned long long*)(a.Tk + (uint64_t)v * m + rk)) >> 32);
            const uint32_t lb = prune_lb ? label_lb<FLT>(a.T1, a.tour, v, a.g, a.masks[rk], a.lbcap) : 0u;
            if (cv <= a.gmax && !reaches_ub<FLT>(cv, lb, a.ub)) keep = 1ull << bit;
        }
        bm = wave_or64(keep);
    }
    const uint32_t na = (uint32_t)__popcll(bm);
    if (na == 0) return;
    const uint32_t A = WAVE / na;
    const uint32_t t = lane % na, s_arc = lane / na;
    const bool lane_ok = s_arc < A;
    const uint32_t rk = wd * 64 + select_bit(bm, t);
    // own cost, read at the coherence point (it may have just been lowered on another XCD)
    const uint32_t cv = (uint32_t)(ld_agent((const unsigned long long*)(a.Tk + (uint64_t)v * m + rk)) >> 32);
    const uint64_t bitv = 1ull << (rk & 63);
    const uint32_t wsel = rk >> 6;
    const uint32_t steps = (e1 - e0 + A - 1) / A;
    for (uint32_t s = 0; s < steps; s += UNR) {
        uint32_t uu[UNR], wt[UNR];
        unsigned long long cw0[UNR];
        bool live[UNR];
#pragma unroll
        for (uint32_t j = 0; j < UNR; ++j) {
            const uint32_t e = e0 + (s + j) * A + s_arc;
            const bool in_row = lane_ok && e < e1;
            const bool ok = in_row && cv != GST_INF_COST;
            const uint2 r = in_row ? a.arc[e] : make_uint2(0xFFFFFFFFu, 0u);
            wt[j] = r.y;
            live[j] = ok && (a.ub == GST_INF_COST || !reaches_ub<FLT>(cv, wt[j], a.ub));
            uu[j] = live[j] ? r.x : 0xFFFFFFFFu;
            st_arcs += (in_row && t == 0) ? 1u : 0u;   // arcs read (once per arc slot, mask lane 0)
        }
        if (__ballot(live[0]) == 0) break;   // arcs run in increasing weight
#pragma unroll
        for (uint32_t j = 0; j < UNR; ++j)
            cw0[j] = uu[j] != 0xFFFFFFFFu ? *(const unsigned long long*)(a.Tk + (uint64_t)uu[j] * m + rk) : 0ull;
        uint32_t enq_mask = 0;
        uint32_t enq_u[UNR];
#pragma unroll
        for (uint32_t j = 0; j < UNR; ++j) {
            enq_u[j] = 0;
            const uint32_t u = uu[j];
            if (u == 0xFFFFFFFFu) continue;
            st_relax++;
            const uint32_t cand = add_cost<FLT>(cv, wt[j], ovf);
            unsigned long long curw = cw0[j];
            if (cand == GST_INF_COST || cand > (uint32_t)(curw >> 32)) continue;
            if (a.ub != GST_INF_COST && reaches_ub<FLT>(cand, 0u, a.ub)) continue;
            unsigned long long* addr = (unsigned long long*)(a.Tk + (uint64_t)u * m + rk);
            const unsigned long long nw = ((unsigned long long)cand << 32) | v;
            bool improved = false;
            if (nw < curw) {
                if (cand == cv) {
                    while (cand < (uint32_t)(curw >> 32)) {
                        const unsigned long long old = atomicCAS(addr, curw, nw);
                        if (old == curw) { improved = true; st_new += mark_active(a.act, u, old); break; }
                        curw = old;
                    }
                } else {
                    const unsigned long long old = atomicMin(addr, nw);
                    improved = cand < (uint32_t)(old >> 32);
                    if (improved) st_new += mark_active(a.act, u, old);
                }
            }
            if (!improved) continue;
            st_upd++;
            if (cand > a.gmax) continue;   // above half the bound: kept, never grown
            unsigned long long* cwp = (unsigned long long*)&chg[(uint64_t)u * W + wsel];
            const unsigned long long ob = atomicOr(cwp, (unsigned long long)bitv);
            const bool first = (W == 1) ? (ob == 0) : (!(ob & bitv) && atomicExch(&inq[u], 1u) == 0u);
            if (first) { enq_mask |= 1u << j; enq_u[j] = u; }
        }
        // queue the vertices whose first change this was: one reservation per wave and step
#pragma unroll
        for (uint32_t j = 0; j < UNR; ++j) {
            const bool want = (enq_mask >> j) & 1u;
            const uint64_t bal = __ballot(want);
            if (bal == 0) continue;
            const uint32_t cntw = (uint32_t)__popcll(bal);
            uint32_t base = 0;
            if (lane == 0) {
                const uint32_t p = atomicAdd(&a.actl->pending, cntw);
                if (p + cntw > a.ring_mask + 1) atomicOr(&a.actl->err, 8u);
                base = atomicAdd(&a.actl->tail, cntw);
            }
            base = __shfl(base, 0, WAVE);
            if (want) {
                const uint32_t pos = base + (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
                task_write(a, n0, pos, TASK_VERTEX, enq_u[j], 0, 0, 0);
            }
            st_enq += want ? 1u : 0u;
        }
    }
}

template <bool FLT>
__global__ void __launch_bounds__(GROW_BLOCK) k_grow_async(GrowArgs a) {
    RoundCtl* ctl = a.ctl;
    AsyncCtl* ac = a.actl;
    if (ctl->level_done) return;
    const uint32_t r = ctl->round;
    const uint32_t cur = r % 3, prv = (r + 2) % 3;
    const uint32_t n0 = ctl->cnt[cur];
    const uint32_t lane = lane_id();
    const uint32_t W = a.W;
    // clear the previous round's bitmaps and flags (the round scheme's phase A)
    {
        const uint32_t n_prv = ctl->cnt[prv];
        const uint64_t* lp = a.list[prv];
        for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_prv; i += gridDim.x * blockDim.x) {
            const uint64_t it = lp[i];
            if ((it >> 32) == 0) {
                const uint32_t v = (uint32_t)it;
                for (uint32_t w = 0; w < W; ++w) a.chg[prv][(uint64_t)v * W + w] = 0;
                a.inq[prv][v] = 0;
            }
        }
    }
    uint64_t* chg = a.chg[cur];
    uint32_t* inq = a.inq[cur];
    const uint64_t* lcur = a.list[cur];
    uint32_t st_items = 0, st_arcs = 0, st_relax = 0, st_upd = 0, st_enq = 0, ovf = 0, st_new = 0;
    uint32_t done_local = 0;
    bool stop = false;
    while (!stop) {
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(&ac->head, ASYNC_BATCH);
        base = __shfl(base, 0, WAVE);
        for (uint32_t j = 0; j < ASYNC_BATCH && !stop; ++j) {
            const uint32_t idx = base + j;
            uint32_t kind, v, c = 0, wd = 0;
            uint64_t bm = 0;
            if (idx < n0) {
                const uint64_t it = lcur[idx];
                v = (uint32_t)it;
                c = (uint32_t)(it >> 32);
                kind = c == 0 ? TASK_VERTEX : 0xFFu;  // later chunks of an initial row: no-op
            } else {
                // wait for the slot of task idx (lane 0 polls, then broadcasts)
                unsigned long long w0 = 0, w1 = 0, w2 = 0, w3 = 0;
                uint32_t outcome = 0;  // 1 = task ready, 2 = level finished, 3 = gave up
                if (lane == 0) {
                    if (done_local) { atomicSub(&ac->pending, done_local); done_local = 0; }
                    const unsigned long long* sl = a.ring + 4ull * ((idx - n0) & a.ring_mask);
                    for (uint32_t it = 0; it < ASYNC_MAX_WAIT; ++it) {
                        w0 = ld_agent(sl + 0);
                        w1 = ld_agent(sl + 1);
                        if ((uint32_t)(w0 >> 32) == idx && (uint32_t)(w1 >> 32) == idx) {
                            if ((w1 & 1ull) == TASK_CHUNK) {
                                w2 = ld_agent(sl + 2);
                                w3 = ld_agent(sl + 3);
                                if ((uint32_t)(w2 >> 32) != idx || (uint32_t)(w3 >> 32) != idx) continue;
                            }
                            outcome = 1;
                            // invalidate the slot: tags restart at 0 in every launch
                            const unsigned long long dead = 0xFFFFFFFF00000000ull;
                            unsigned long long* sw = a.ring + 4ull * ((idx - n0) & a.ring_mask);
                            st_agent(sw + 0, dead);
                            st_agent(sw + 1, dead);
                            if ((w1 & 1ull) == TASK_CHUNK) { st_agent(sw + 2, dead); st_agent(sw + 3, dead); }
                            break;
                        }
                        if (ld_agent32(&ac->pending) == 0) { outcome = 2; break; }
                        // exponential back-off: idle waves must not flood the coherence point
                        if (it < 2) continue;
                        else if (it < 8) __builtin_amdgcn_s_sleep(8);
                        else if (it < 32) __builtin_amdgcn_s_sleep(32);
                        else __builtin_amdgcn_s_sleep(127);
                    }
                    if (outcome == 0) { outcome = 3; atomicOr(&ac->err, 4u); }
                }
                outcome = __shfl(outcome, 0, WAVE);
                if (outcome != 1) { stop = true; break; }
                w0 = __shfl(w0, 0, WAVE); w1 = __shfl(w1, 0, WAVE);
                w2 = __shfl(w2, 0, WAVE); w3 = __shfl(w3, 0, WAVE);
                v = (uint32_t)w0;
                kind = (uint32_t)(w1 & 1ull);
                wd = (uint32_t)((w1 >> 1) & 0xFFull);
                c = (uint32_t)((w1 & 0xFFFFFFFFull) >> 9);
                bm = (w2 & 0xFFFFFFFFull) | ((w3 & 0xFFFFFFFFull) << 32);
            }
            if (kind == TASK_VERTEX) {
                // take the change bitmap of v
                uint64_t bms[16];
                uint32_t nz = 0;
                if (lane == 0) atomicExch(&inq[v], 0u);
                for (uint32_t w = 0; w < W && w < 16; ++w) {
                    uint64_t x = 0;
                    if (lane == 0) x = atomicExch((unsigned long long*)&chg[(uint64_t)v * W + w], 0ull);
                    bms[w] = __shfl(x, 0, WAVE);
                    nz += bms[w] != 0;
                }
                const uint32_t rs = a.row[v], re = a.row[v + 1];
                const uint32_t nc = (re - rs + a.chunk - 1) / a.chunk;
                if (nz && nc == 1) {
                    st_items++;
                    for (uint32_t w = 0; w < W && w < 16; ++w)
                        if (bms[w]) async_relax<FLT>(a, n0, v, rs, re, w, bms[w], st_relax, st_upd, st_enq, ovf, st_new, st_arcs);
                } else if (nz && nc > 1) {
                    // split the row: one CHUNK task per (chunk, non-empty word)
                    const uint32_t ntask = nc * nz;
                    uint32_t tb = 0;
                    if (lane == 0) {
                        const uint32_t p = atomicAdd(&ac->pending, ntask);
                        if (p + ntask > a.ring_mask + 1) atomicOr(&ac->err, 8u);
                        tb = atomicAdd(&ac->tail, ntask);
                    }
                    tb = __shfl(tb, 0, WAVE);
                    uint32_t k2 = 0;
                    for (uint32_t w = 0; w < W && w < 16; ++w) {
                        if (!bms[w]) continue;
                        for (uint32_t cc = lane; cc < nc; cc += WAVE)
                            task_write(a, n0, tb + k2 * nc + cc, TASK_CHUNK, v, cc, w, bms[w]);
                        k2++;
                    }
                    st_enq += ntask;
                }
            } else if (kind == TASK_CHUNK) {
                const uint32_t rs = a.row[v], re = a.row[v + 1];
                const uint32_t e0 = rs + c * a.chunk;
                const uint32_t e1 = min(re, e0 + a.chunk);
                st_items++;
                async_relax<FLT>(a, n0, v, e0, e1, wd, bm, st_relax, st_upd, st_enq, ovf, st_new, st_arcs);
            }
            done_local++;
        }
        if (lane == 0 && done_local) { atomicSub(&ac->pending, done_local); done_local = 0; }
    }
    if (lane == 0 && done_local) atomicSub(&ac->pending, done_local);
    const uint32_t si = wave_sum32(lane == 0 ? st_items : 0u);
    const uint32_t sa = wave_sum32(st_arcs);
    const uint32_t sr = wave_sum32(st_relax);
    const uint32_t su = wave_sum32(st_upd);
    const uint32_t se = wave_sum32(st_enq);
    const uint32_t sn = wave_sum32(st_new);
    const uint32_t so = wave_sum32(ovf);
    if (lane == 0) {
        unsigned long long* sp = stat_stripe(a.stats, blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE);
        if (sn) atomicAdd(&sp[ST_NEW], (unsigned long long)sn);
        if (si) atomicAdd(&sp[ST_ITEMS], (unsigned long long)si);
        if (sa) atomicAdd(&sp[ST_ARCS], (unsigned long long)sa);
        if (sr) atomicAdd(&sp[ST_RELAX], (unsigned long long)sr);
        if (su) atomicAdd(&sp[ST_UPD], (unsigned long long)su);
        if (se) atomicAdd(&sp[ST_ENQ], (unsigned long long)se);
        if (so) atomicOr(&ac->err, 1u);
    }
}

// after an asynchronous launch: the level is finished (counters of the round scheme reset)
__global__ void k_async_end(RoundCtl* ctl, AsyncCtl* ac) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (ac->err & 1u) ctl->err |= 1u;
    if (ac->err & 12u) ctl->err |= 4u;   // lost work: the host re-runs the query on the round path
    ctl->level_done = 1;
}

// prepare an asynchronous launch from the current round's list
__global__ void k_async_begin(const RoundCtl* ctl, AsyncCtl* ac) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const uint32_t n0 = ctl->cnt[ctl->round % 3];
    ac->head = 0;
    ac->tail = n0;
    ac->pending = n0;
}

// ---------------------------------------------------------------------------------------
// merge: T(v,X) for all X of level k from the final lower levels of v.
// A workgroup takes a tile of VB consecutive vertices: their lower-level rows (contiguous runs of
// the level blocks) are staged in LDS with coalesced loads, the split table of the level sits in
// LDS, and each thread owns (vertex, mask) outputs: min over the 2^(k-1)-1 splits in registers,
// then a coalesced store of the level-k rows.  No atomics except the LDS bitmap of finite masks.
// ---------------------------------------------------------------------------------------
// merge workgroup: 8 waves (measured on BA V=200k g=6 against 4 and 16, with the tile below)
#ifndef GST_MERGE_THREADS
#define GST_MERGE_THREADS 512
#endif
constexpr int MERGE_THREADS = GST_MERGE_THREADS;
// default LDS budget of a merge tile's cost rows (measured on BA V=200k g=6 with 512-thread
// workgroups: 48 KiB beats 32 and 64; three workgroups per CU)
#ifndef GST_MERGE_TILE_KB
#define GST_MERGE_TILE_KB 48
#endif
constexpr size_t MERGE_TILE_BYTES = GST_MERGE_TILE_KB * 1024;
// loads in flight per thread while a merge tile's rows are staged
#ifndef GST_MERGE_UNR
#define GST_MERGE_UNR 4
#endif
constexpr int MERGE_UNR = GST_MERGE_UNR;
// at most this many vertices per merge tile: low levels (few lower-level words per vertex) would
// otherwise take tiles far below the LDS budget (measured on BA V=200k g=6: 512 takes the one-query
// merge from 272 to 246 us per query and the batched merge 2% down against 256; 1024 is no better,
// 128 is 15% slower)
#ifndef GST_MERGE_VB_MAX
#define GST_MERGE_VB_MAX 512
#endif
// split tables up to this size are staged in LDS
constexpr size_t MERGE_PAIRS_LDS = 16 * 1024;
// levels of at most this many splits per mask (k <= 4) take their splits from registers; the others
// read the staged table (the full level, with one mask, as LDS broadcasts: every lane reads the same
// entry).  Measured on BA V=200k g=6: 1, 4 or 8 give the same merge time, 16 spills registers (-20%)
#ifndef GST_MERGE_REG_SPLITS
#define GST_MERGE_REG_SPLITS 8
#endif
constexpr uint32_t MERGE_REG_SPLITS = GST_MERGE_REG_SPLITS;
// merge waves per SIMD the register allocation must allow (three 48 KiB-tile workgroups of 8 waves
// per CU: 6)
#ifndef GST_MERGE_WPE
#define GST_MERGE_WPE 6
#endif

// A tile's staged rows are VB rows of L 4-byte costs at an odd stride, so that lanes reading the
// same position of different vertices' rows hit different LDS banks
__host__ __device__ __forceinline__ uint32_t merge_row_stride(uint32_t L) { return L | 1u; }
// 8-byte words taken by VB rows of L 4-byte costs
__host__ __device__ __forceinline__ size_t merge_rows_words(uint32_t VB, uint32_t L) {
    return ((size_t)VB * merge_row_stride(L) + 1) / 2;
}
// vertices per merge tile for L lower-level costs and W bitmap words per vertex within a budget of
// tile_bytes (0: the default)
__host__ __device__ __forceinline__ uint32_t merge_tile_vertices(uint32_t L, uint32_t W, size_t tile_bytes = 0) {
    const size_t per = (size_t)merge_row_stride(L) * 4 + (size_t)W * 8 + 12;
    size_t vb = (tile_bytes ? tile_bytes : MERGE_TILE_BYTES) / per;
    if (vb < 1) vb = 1;
    if (vb > GST_MERGE_VB_MAX) vb = GST_MERGE_VB_MAX;
    return (uint32_t)vb;
}
// dynamic LDS bytes of a merge launch: rows, bitmaps, group masks, live list and finite flags of the
// tile, split table
__host__ __device__ __forceinline__ size_t merge_lds_bytes(uint32_t VB, uint32_t L, uint32_t W, size_t pair_bytes,
                                                         bool pin) {
    return merge_rows_words(VB, L) * 8 + (size_t)VB * W * 8 + 3 * (((size_t)VB + 1) / 2 * 8) + (pin ? pair_bytes : 0);
}

// i / d for i < 2^20 and 1 <= d < 2^12 without an integer division: (i + 0.5) / d in float is
// off by less than 2^-3 / d, while its fractional part stays at least 0.5 / d away from an integer
__device__ __forceinline__ uint32_t div_small(uint32_t i, float inv_d) {
    return (uint32_t)(((float)i + 0.5f) * inv_d);
}

// a bounded merge of level k >= 3 with activity bytes walks the list of its live vertices
__host__ __device__ __forceinline__ bool merge_uses_list(const MergeArgs& a) {
    return a.live != nullptr && a.act != nullptr && a.k >= 3;
}

// Pre-pass of a merge that walks its live list.  A vertex is live when it belongs to at least k groups
// or holds a finite state at some level 2..k-1 (no other vertex can hold a finite level-k state: no
// split pair of it is finite and it is no leaf).  Workgroups of 256 threads take 256-vertex tiles: a
// dead vertex's level-k words are written inf (coalesced, one thread per word) and its level-k
// activity byte cleared; a live vertex is appended to the list (one counter atomic per wave).
__device__ __forceinline__ void merge_live_body(const MergeArgs& a, uint32_t bid, uint32_t nblk) {
    __shared__ uint32_t s_dead[256];
    const uint32_t k = a.k, mk = a.m[k], n = a.n;
    const float inv_mk = 1.0f / (float)mk;
    uint64_t* Tk = a.T + a.lvl_base[k];
    uint8_t* act_out = a.is_full ? nullptr : a.act + (size_t)k * n;
    uint64_t st_bytes = 0;
    for (uint32_t v0 = bid * 256u; v0 < n; v0 += nblk * 256u) {
        const uint32_t nv = min(256u, n - v0);
        const uint32_t v = v0 + threadIdx.x;
        bool live = false;
        if (threadIdx.x < nv) {
            live = (uint32_t)__popc(a.gm[v]) >= k;
            for (uint32_t j = 2; j < k && !live; ++j) live = a.act[(size_t)j * n + v] != 0;
            if (!live && act_out) act_out[v] = 0;
            // this vertex's group mask and activity bytes read, its activity byte cleared
            st_bytes += 4 + (k - 2) + ((!live && act_out) ? 1 : 0);
        }
        const uint64_t bal = __ballot(live);
        uint32_t base = 0;
        if (lane_id() == 0 && bal) base = atomicAdd(&a.ctl->live, (uint32_t)__popcll(bal));
        base = __shfl(base, 0, WAVE);
        if (live) a.live[base + (uint32_t)__popcll(bal & ((1ull << lane_id()) - 1ull))] = v;
        s_dead[threadIdx.x] = (threadIdx.x < nv && !live) ? 1u : 0u;
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < nv * mk; i += 256u) {
            const uint32_t b = div_small(i, inv_mk);
            if (s_dead[b]) {
                Tk[(uint64_t)v0 * mk + i] = GST_INF_WORD;
                st_bytes += 8;   // a dead vertex's word written
            }
        }
        __syncthreads();
    }
    __shared__ unsigned long long s_b;
    if (threadIdx.x == 0) s_b = 0;
    __syncthreads();
    if (st_bytes) atomicAdd(&s_b, (unsigned long long)st_bytes);
    __syncthreads();
    if (threadIdx.x == 0 && s_b) atomicAdd(&stat_stripe(a.stats, bid)[ST_MBYTES], s_b);
}

template <bool FLT, bool LIST>
__device__ __forceinline__ void merge_body(const MergeArgs& a, unsigned long long* lds, uint32_t bid, uint32_t nblk) {
    __shared__ uint32_t s_nl;
    const uint32_t k = a.k;
    const uint32_t L = a.off[k];             // lower-level words per vertex
    const uint32_t Lp = merge_row_stride(L); // their stride in LDS
    const uint32_t mk = a.m[k];
    const uint32_t S = (1u << (k - 1)) - 1;  // splits per mask
    const uint32_t VB = a.VB;
    const uint32_t W = a.W;
    const uint32_t g = a.g;
    // LDS: the live vertices' lower-level costs (the merge reads no predecessors), the bitmaps of
    // finite outputs, the group masks of the tile's vertices (bit 31: live), the list of live tile
    // vertices, their finite-output flags, then the split table
    const uint32_t VB2 = (VB + 1) & ~1u;
    uint32_t* rows = (uint32_t*)lds;                                            // VB * L costs
    unsigned long long* bm = lds + merge_rows_words(VB, L);                     // VB * W
    uint32_t* gms = (uint32_t*)(bm + (size_t)VB * W);                           // VB
    uint32_t* lst = gms + VB2;                                                  // VB
    uint32_t* fin = lst + VB2;                                                  // VB
    // split table: in LDS transposed to [split][mask], so that the lanes of one wave (consecutive
    // masks) read consecutive entries; else the global [mask][split] table
    const unsigned long long* pt = (const unsigned long long*)a.pairs;
    uint32_t pr_stride = 1, ps_stride = S;   // entry (r, s2) at pt[r * ps_stride + s2 * pr_stride]
    if (a.pairs_in_lds) {
        unsigned long long* pl = (unsigned long long*)(fin + VB2);
        for (uint32_t i = threadIdx.x; i < a.P; i += blockDim.x) {
            const uint32_t r = i / S, s2 = i - r * S;
            pl[s2 * mk + r] = a.pairs[i];
        }
        pt = pl;
        pr_stride = mk;
        ps_stride = 1;
    }
    const bool prune = a.ub != GST_INF_COST;
    const bool prune_lb = prune && a.use_lb && !a.is_full;
    const float inv_mk = 1.0f / (float)mk;
    uint32_t ovf = 0;
    unsigned long long best_root = GST_INF_WORD;
    uint64_t st_pairs = 0, st_vert = 0, st_bytes = 0;
    const uint64_t* Tk_base = a.T + a.lvl_base[k];
    const uint32_t* mk_masks = a.masks + a.off[k];
    // Levels k >= 3 with activity bytes: a vertex whose levels 2..k-1 hold no finite state and which
    // is not a member of k groups has no finite pair and no leaf at level k, so its level-k words are
    // all inf and its lower rows are never read ("live" vertices are the others).  Level 2 merges
    // every vertex (its pairs are level-1 distances).
    const bool skip = a.act != nullptr && k >= 3;
    uint8_t* act_out = (a.act && !a.is_full) ? a.act + (size_t)k * a.n : nullptr;

    // Splits in registers: when a level has at most MERGE_REG_SPLITS splits per mask and at most a
    // workgroup's worth of masks, thread t serves the (vertex, mask) items t, t + istride, ... with
    // istride a multiple of the mask count, so its mask is fixed; its splits are enumerated once into
    // registers as (position of X1, position of X \ X1, X1), 10 bits each (g <= 10), and the split
    // loop reads only the two staged costs per split.
    const bool reg_splits = S <= MERGE_REG_SPLITS && mk <= blockDim.x && g <= 10;
    const uint32_t istride = reg_splits ? (blockDim.x / mk) * mk : blockDim.x;
    uint32_t sp[MERGE_REG_SPLITS];
#pragma unroll
    for (uint32_t s2 = 0; s2 < MERGE_REG_SPLITS; ++s2) sp[s2] = 0u;
    if (reg_splits) {
        __syncthreads();   // the split table is staged
        const unsigned long long* pr = pt + (size_t)(threadIdx.x % mk) * ps_stride;
#pragma unroll
        for (uint32_t s2 = 0; s2 < MERGE_REG_SPLITS; ++s2) {
            if (s2 < S) {
                const unsigned long long e = pr[(size_t)s2 * pr_stride];
                sp[s2] = (uint32_t)((e >> 16) & 0x3FFu) | ((uint32_t)((e >> 32) & 0x3FFu) << 10) |
                         ((uint32_t)(e >> 48) << 20);
            }
        }
    }

    // LIST: tile t is entries t * VB ... of the live list (absolute vertex ids in lst, the per-vertex
    // LDS arrays indexed by list position, v0 = 0); the pre-pass has written the dead vertices
    const uint32_t n_tile = LIST ? a.ctl->live : a.n;
    for (uint32_t t = bid; t * VB < n_tile; t += nblk) {
        const uint32_t v0 = LIST ? 0u : t * VB;
        const uint32_t nv = min(VB, n_tile - t * VB);
        __syncthreads();
        if (threadIdx.x == 0) s_nl = (skip && !LIST) ? 0u : nv;
        for (uint32_t b = threadIdx.x; b < nv; b += blockDim.x) {
            if (LIST) {
                const uint32_t v = a.live[t * VB + b];
                lst[b] = v;
                gms[b] = a.gm[v] | 0x80000000u;
            } else {
                gms[b] = a.gm[v0 + b];
                if (!skip) lst[b] = b;
            }
            fin[b] = 0u;
        }
        __syncthreads();
        if (skip && !LIST) {
            for (uint32_t b = threadIdx.x; b < nv; b += blockDim.x) {
                bool live = (uint32_t)__popc(gms[b]) >= k;
                for (uint32_t j = 2; j < k && !live; ++j) live = a.act[(size_t)j * a.n + v0 + b] != 0;
                if (live) {
                    lst[atomicAdd(&s_nl, 1u)] = b;
                    gms[b] |= 0x80000000u;
                }
            }
            __syncthreads();
        }
        const uint32_t nl = s_nl;
        if (skip && !LIST && nl < nv) {
            // the non-live vertices' level-k words: inf, written without reading their rows
            for (uint32_t i = threadIdx.x; i < nv * mk; i += blockDim.x) {
                const uint32_t b = div_small(i, inv_mk);
                if (!(gms[b] & 0x80000000u)) ((uint64_t*)Tk_base)[(uint64_t)v0 * mk + i] = GST_INF_WORD;
            }
        }
        // stage the cost halves of the live vertices' lower-level rows (each a contiguous run of m_j
        // words; without skipping, the tile's run of each level is one contiguous block of nv * m_j
        // words), MERGE_UNR loads in flight per thread
        for (uint32_t j = 1; j < k; ++j) {
            const uint32_t mj = a.m[j];
            const float inv_mj = 1.0f / (float)mj;
            const uint64_t* src = a.T + a.lvl_base[j] + (uint64_t)v0 * mj;
            const uint32_t cnt = nl * mj;
            const uint32_t oj = a.off[j];
            for (uint32_t i0 = threadIdx.x; i0 < cnt; i0 += MERGE_UNR * blockDim.x) {
                // only the cost half of each word is read (4-byte loads of the high halves): the
                // merge reads no predecessors, and each load in flight holds one register
                const uint32_t* src_hi = (const uint32_t*)src + 1;
                uint32_t x[MERGE_UNR];
#pragma unroll
                for (int u = 0; u < MERGE_UNR; ++u) {
                    const uint32_t i = i0 + u * blockDim.x;
                    uint32_t val = 0u;
                    if (i < cnt) {
                        if (skip) {
                            const uint32_t li = div_small(i, inv_mj);
                            val = src_hi[2 * ((uint64_t)lst[li] * mj + (i - li * mj))];
                        } else {
                            val = src_hi[2 * (uint64_t)i];
                        }
                    }
                    x[u] = val;
                }
#pragma unroll
                for (int u = 0; u < MERGE_UNR; ++u) {
                    const uint32_t i = i0 + u * blockDim.x;
                    if (i < cnt) {
                        const uint32_t b = div_small(i, inv_mj), t = i - b * mj;
                        rows[b * Lp + oj + t] = x[u];
                    }
                }
            }
        }
        for (uint32_t i = threadIdx.x; i < nv * W; i += blockDim.x) bm[i] = 0;
        __syncthreads();
        if (a.spider) {
            // cheapest spider of the tile (level 2: every vertex is staged): sum of the group
            // distances of one vertex (positions 0..g-1 of the staged rows are the masks 1<<i), with
            // the vertex (the spider's centre)
            unsigned long long sm = GST_INF_WORD;
            for (uint32_t b = threadIdx.x; b < nl; b += blockDim.x) {
                uint32_t acc = 0, o2 = 0;
                for (uint32_t gi = 0; gi < g; ++gi) acc = add_cost<FLT>(acc, rows[b * Lp + gi], o2);
                if (!o2 && acc != GST_INF_COST) {
                    const unsigned long long w = ((unsigned long long)acc << 32) | (v0 + lst[b]);
                    sm = w < sm ? w : sm;
                }
            }
            sm = wave_min64(sm);
            if (lane_id() == 0 && sm != GST_INF_WORD) atomicMin(a.spider, sm);
        }
        const uint32_t items = nl * mk;
        for (uint32_t i = threadIdx.x; threadIdx.x < istride && i < items; i += istride) {
            const uint32_t li = div_small(i, inv_mk), r = i - li * mk;
            const uint32_t b = LIST ? li : lst[li];
            const uint32_t v = LIST ? lst[li] : v0 + b;
            const uint32_t X = mk_masks[r];
            const uint32_t gmv = gms[b] & 0x7FFFFFFFu;
            const uint32_t* rv = rows + (size_t)li * Lp;
            unsigned long long best = ((X & gmv) == X) ? (unsigned long long)GST_PRED_LEAF : GST_INF_WORD;
            const unsigned long long* pr = pt + (size_t)r * ps_stride;
            if (reg_splits) {
#pragma unroll
                for (uint32_t s2 = 0; s2 < MERGE_REG_SPLITS; ++s2) {
                    if (s2 < S) {
                        const uint32_t e = sp[s2];
                        const uint32_t sum = add_cost<FLT>(rv[e & 0x3FFu], rv[(e >> 10) & 0x3FFu], ovf);
                        if (sum != GST_INF_COST) {
                            const unsigned long long wv = ((unsigned long long)sum << 32) | (GST_PRED_MERGE | (e >> 20));
                            best = wv < best ? wv : best;
                        }
                    }
                }
            }
            for (uint32_t s2 = 0; !reg_splits && s2 < S; ++s2) {
                const unsigned long long e = pr[(size_t)s2 * pr_stride];
                const uint32_t p1 = (uint32_t)((e >> 16) & 0xFFFF);
                const uint32_t p2 = (uint32_t)((e >> 32) & 0xFFFF);
                const uint32_t c1 = rv[p1];
                const uint32_t c2 = rv[p2];
                const uint32_t sum = add_cost<FLT>(c1, c2, ovf);
                if (sum != GST_INF_COST) {
                    const unsigned long long wv = ((unsigned long long)sum << 32) | (GST_PRED_MERGE | (uint32_t)(e >> 48));
                    best = wv < best ? wv : best;
                }
            }
            if (best != GST_INF_WORD && prune) {
                // one-label bound only (level-1 position gi of the staged row is the mask 1<<gi):
                // a merged state is grown only after the grow rounds' own check, which adds the
                // tour bounds, so repeating those here would only cost the merge time
                const uint32_t lb = prune_lb ? state_lb<FLT>([&](uint32_t gi) { return rv[gi]; }, nullptr, g, X, a.lbcap) : 0u;
                if (reaches_ub<FLT>((uint32_t)(best >> 32), lb, a.ub)) best = GST_INF_WORD;
            }
            ((uint64_t*)Tk_base)[(uint64_t)v * mk + r] = best;
            bool flag = false;
            if (best != GST_INF_WORD) {
                if (act_out) fin[b] = 1u;
                if (a.is_full) {
                    const unsigned long long rw = ((best >> 32) << 32) | v;
                    best_root = rw < best_root ? rw : best_root;
                } else {
                    flag = (uint32_t)(best >> 32) <= a.gmax;
                }
            }
            // frontier bits: lanes i, i+1, ... of one wave hold consecutive masks, so the bits of
            // one (vertex, bitmap word) are a contiguous run of the wave's ballot; the first lane of
            // each run ORs the run into the LDS bitmap (one LDS atomic per run, not per mask)
            const uint64_t bal = __ballot(flag);
            const uint32_t lane = lane_id();
            const bool leader = lane == 0 || (r & 63u) == 0;
            if (leader && bal) {
                const uint32_t cnt = min(min(64u - (r & 63u), mk - r), 64u - lane);
                const uint64_t run = (bal >> lane) & (cnt == 64u ? ~0ull : ((1ull << cnt) - 1ull));
                if (run) atomicOr(&bm[b * W + (r >> 6)], (unsigned long long)(run << (r & 63u)));
            }
        }
        st_pairs += (uint64_t)items * S;
        st_vert += nl;
        // bytes: live rows read, every level-k word written, group masks, activity bytes (with the
        // live list: the live vertices' share and the list read; the pre-pass counts the rest)
        st_bytes += (uint64_t)nl * L * 8 + (uint64_t)nv * mk * 8 + (uint64_t)nv * 4 +
                    ((skip && !LIST) ? (uint64_t)nv * (k - 2) : 0) + (act_out ? nv : 0) + (LIST ? (uint64_t)nv * 4 : 0);
        __syncthreads();
        if (act_out)
            for (uint32_t b = threadIdx.x; b < nv; b += blockDim.x) act_out[LIST ? lst[b] : v0 + b] = (uint8_t)fin[b];
        if (!a.is_full) {
            for (uint32_t b = threadIdx.x; b < nv; b += blockDim.x) {
                const uint32_t v = LIST ? lst[b] : v0 + b;
                unsigned long long any = 0;
                for (uint32_t w = 0; w < W; ++w) any |= bm[b * W + w];
                if (!any || a.row[v + 1] == a.row[v]) continue;   // nothing to grow, or no arcs
                for (uint32_t w = 0; w < W; ++w) {
                    const unsigned long long x = bm[b * W + w];
                    if (x) a.chg0[(uint64_t)v * W + w] = x;
                }
                a.inq0[v] = 1;
            }
        }
    }
    if (a.is_full) {
        best_root = wave_min64(best_root);
        if (lane_id() == 0 && best_root != GST_INF_WORD) atomicMin(a.root, best_root);
    }
    const uint32_t so = wave_sum32(ovf);
    if (threadIdx.x == 0) {
        unsigned long long* sp = stat_stripe(a.stats, bid);
        atomicAdd(&sp[ST_MPAIRS], (unsigned long long)st_pairs);
        atomicAdd(&sp[ST_MVERT], (unsigned long long)st_vert);
        atomicAdd(&sp[ST_MBYTES], (unsigned long long)st_bytes);
    }
    if (lane_id() == 0 && so) atomicOr(&a.ctl->err, 1u);
}

__global__ void __launch_bounds__(256) k_merge_live(MergeArgs a) { merge_live_body(a, blockIdx.x, gridDim.x); }

template <bool FLT>
__global__ void __launch_bounds__(MERGE_THREADS) __attribute__((amdgpu_waves_per_eu(GST_MERGE_WPE))) k_merge(MergeArgs a) {
    extern __shared__ unsigned long long lds[];
    if (merge_uses_list(a)) merge_body<FLT, true>(a, lds, blockIdx.x, gridDim.x);
    else merge_body<FLT, false>(a, lds, blockIdx.x, gridDim.x);
}

// cheapest spider: min over v of sum_i T(v,{i}) (the union of v's shortest paths to every group
// holds a tree of at most that cost), as the word (cost << 32 | v): the smallest cost, then the
// smallest centre
template <bool FLT>
__global__ void k_spider_scan(const uint64_t* T1, uint32_t n, uint32_t g, unsigned long long* out) {
    unsigned long long sm = GST_INF_WORD;
    for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t acc = 0, o2 = 0;
        for (uint32_t gi = 0; gi < g; ++gi) acc = add_cost<FLT>(acc, (uint32_t)(T1[v * g + gi] >> 32), o2);
        if (!o2 && acc != GST_INF_COST) {
            const unsigned long long w = ((unsigned long long)acc << 32) | (uint32_t)v;
            sm = w < sm ? w : sm;
        }
    }
    sm = wave_min64(sm);
    if (lane_id() == 0 && sm != GST_INF_WORD) atomicMin(out, sm);
}

// ---------------------------------------------------------------------------------------
// MST post-processing (rucgraph/graph_hash_of_mixed_weighted_MST_postprocessing.h:73-103 with the
// Prim of graph_hash_of_mixed_weighted/common_algorithms/graph_hash_of_mixed_weighted_minimum_spanning_tree.h:29-132):
// the minimum spanning tree of the subgraph of the input graph induced by a vertex set.  One
// workgroup: the set in an LDS hash, the induced edges gathered by waves scanning the members' rows,
// sorted by weight (bitonic, in LDS), then Kruskal (one thread, union-find).  A disconnected set
// gives a spanning forest (the reference's Prim stops after the component of its first vertex;
// every set the solver passes is connected, where the two agree).
//   spider tree (the automatic bound): the set is the union of the spider centre's shortest paths
//   to every group, walked along the level-1 predecessors; after the MST, leaves that no group needs
//   are removed.  The tree is feasible (connected, every group touched), so its cost is a bound no
//   lower than the optimum, and no higher than the spider's sum of distances.
// ---------------------------------------------------------------------------------------
constexpr uint32_t MST_SMAX = 512;     // vertices of the set
constexpr uint32_t MST_HCAP = 1024;    // hash slots (power of two)
constexpr uint32_t MST_EMAX = 2048;    // induced edges (power of two)
constexpr uint32_t MST_WALK = 256;     // hops of one shortest path
constexpr uint32_t MST_EMPTY = 0xFFFFFFFFu;
constexpr int MST_THREADS = 256;

struct MstLds {
    uint32_t hkey[MST_HCAP];           // vertex id, MST_EMPTY when free
    uint32_t hval[MST_HCAP];           // its index in vs
    uint32_t vs[MST_SMAX];             // the set, by index
    uint32_t nb[MST_SMAX];             // tree adjacency (pruning): first incident tree edge
    uint32_t deg[MST_SMAX];
    unsigned long long e[MST_EMAX];    // (weight bits << 32) | (a << 16) | b, a < b indices; bit 63: in the tree
    uint32_t par[MST_SMAX];
    uint32_t enext[2 * MST_SMAX];      // tree adjacency lists: entry 2j / 2j+1 for the two ends of tree edge j
    uint32_t tedge[MST_SMAX];          // the tree's edges, as indices into e
    uint32_t nv, ne, nt, fail;
};

__device__ __forceinline__ uint32_t mst_hash(uint32_t v) { return (v * 2654435761u) & (MST_HCAP - 1); }

// inserts v (any thread, concurrently); false when the table is full
__device__ __forceinline__ bool mst_insert(MstLds& L, uint32_t v) {
    uint32_t h = mst_hash(v);
    for (uint32_t p = 0; p < MST_HCAP; ++p, h = (h + 1) & (MST_HCAP - 1)) {
        const uint32_t old = atomicCAS(&L.hkey[h], MST_EMPTY, v);
        if (old == MST_EMPTY || old == v) return true;
    }
    return false;
}

// index of v in the set, MST_EMPTY when absent (after mst_number)
__device__ __forceinline__ uint32_t mst_find(const MstLds& L, uint32_t v) {
    uint32_t h = mst_hash(v);
    for (uint32_t p = 0; p < MST_HCAP; ++p, h = (h + 1) & (MST_HCAP - 1)) {
        const uint32_t k = L.hkey[h];
        if (k == v) return L.hval[h];
        if (k == MST_EMPTY) return MST_EMPTY;
    }
    return MST_EMPTY;
}

__device__ __forceinline__ uint32_t mst_root(MstLds& L, uint32_t x) {
    while (L.par[x] != x) {
        L.par[x] = L.par[L.par[x]];
        x = L.par[x];
    }
    return x;
}

__device__ __forceinline__ void mst_clear(MstLds& L) {
    for (uint32_t i = threadIdx.x; i < MST_HCAP; i += blockDim.x) L.hkey[i] = MST_EMPTY;
    if (threadIdx.x == 0) { L.nv = 0; L.ne = 0; L.nt = 0; L.fail = 0; }
    __syncthreads();
}

// After the inserts: numbers the set (vs, hval), gathers the induced edges, sorts them and keeps a
// minimum spanning forest (bit 63 of the kept edges, their indices in tedge[0..nt)).  All threads.
template <bool FLT>
__device__ void mst_build(MstLds& L, const uint32_t* row, const uint2* arc) {
    __syncthreads();
    for (uint32_t h = threadIdx.x; h < MST_HCAP; h += blockDim.x) {
        const uint32_t k = L.hkey[h];
        if (k == MST_EMPTY) continue;
        const uint32_t i = atomicAdd(&L.nv, 1u);
        if (i < MST_SMAX) { L.vs[i] = k; L.hval[h] = i; }
        else L.fail = 1;
    }
    __syncthreads();
    if (L.fail) return;
    const uint32_t nv = L.nv;
    // induced edges: wave w scans the rows of members w, w + waves, ...; each undirected edge is taken
    // from the row of its smaller endpoint
    const uint32_t lane = lane_id(), wv = threadIdx.x / WAVE, nw = blockDim.x / WAVE;
    for (uint32_t i = wv; i < nv; i += nw) {
        const uint32_t v = L.vs[i];
        const uint32_t r0 = row[v], r1 = row[v + 1];
        for (uint32_t e = r0 + lane; e < r1; e += WAVE) {
            const uint2 rc = arc[e];
            if (rc.x <= v) continue;
            const uint32_t j = mst_find(L, rc.x);
            if (j == MST_EMPTY) continue;
            const uint32_t slot = atomicAdd(&L.ne, 1u);
            if (slot < MST_EMAX) L.e[slot] = ((unsigned long long)rc.y << 32) | ((unsigned long long)i << 16) | j;
            else L.fail = 1;
        }
    }
    __syncthreads();
    if (L.fail) return;
    const uint32_t ne = L.ne;
    uint32_t P = 1;
    while (P < ne) P <<= 1;
    for (uint32_t i = ne + threadIdx.x; i < P; i += blockDim.x) L.e[i] = ~0ull >> 1;   // after every real edge
    __syncthreads();
    // bitonic sort of e[0..P) by (weight, a, b)
    for (uint32_t k = 2; k <= P; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < P; i += blockDim.x) {
                const uint32_t l = i ^ j;
                if (l > i) {
                    const unsigned long long x = L.e[i], y = L.e[l];
                    const bool up = (i & k) == 0;
                    if ((x > y) == up) { L.e[i] = y; L.e[l] = x; }
                }
            }
            __syncthreads();
        }
    }
    // Kruskal over the sorted edges
    if (threadIdx.x == 0) {
        for (uint32_t i = 0; i < nv; ++i) L.par[i] = i;
        uint32_t nt = 0;
        for (uint32_t i = 0; i < ne && nt + 1 < nv; ++i) {
            const unsigned long long x = L.e[i];
            const uint32_t ra = mst_root(L, (uint32_t)(x >> 16) & 0xFFFFu), rb = mst_root(L, (uint32_t)x & 0xFFFFu);
            if (ra == rb) continue;
            L.par[ra] = rb;
            L.e[i] = x | (1ull << 63);
            L.tedge[nt++] = i;
        }
        L.nt = nt;
    }
    __syncthreads();
}

// Removes, one thread, tree leaves whose groups (group bits gbit(v)) are all touched by another
// kept vertex, until none is left; returns the kept edges' cost (GST_INF_COST on overflow).
template <bool FLT, class GroupF>
__device__ uint32_t mst_prune_cost(MstLds& L, uint32_t g, GroupF gbit) {
    const uint32_t nv = L.nv, nt = L.nt;
    uint32_t cover[GST_MAX_GROUPS];
    for (uint32_t i = 0; i < g; ++i) cover[i] = 0;
    for (uint32_t i = 0; i < nv; ++i) {
        L.deg[i] = 0;
        L.nb[i] = MST_EMPTY;
        const uint32_t b = gbit(L.vs[i]);
        for (uint32_t q = 0; q < g; ++q) cover[q] += (b >> q) & 1u;
    }
    for (uint32_t t = 0; t < nt; ++t) {
        const unsigned long long x = L.e[L.tedge[t]];
        const uint32_t a = (uint32_t)(x >> 16) & 0xFFFFu, b = (uint32_t)x & 0xFFFFu;
        L.deg[a]++; L.deg[b]++;
        L.enext[2 * t] = L.nb[a]; L.nb[a] = 2 * t;
        L.enext[2 * t + 1] = L.nb[b]; L.nb[b] = 2 * t + 1;
    }
    // par[] is reused as the stack of leaves to look at
    uint32_t sp = 0;
    for (uint32_t i = 0; i < nv; ++i) if (L.deg[i] == 1) L.par[sp++] = i;
    while (sp) {
        const uint32_t v = L.par[--sp];
        if (L.deg[v] != 1) continue;
        const uint32_t b = gbit(L.vs[v]);
        bool need = false;
        for (uint32_t q = 0; q < g; ++q) need = need || (((b >> q) & 1u) && cover[q] <= 1u);
        if (need) continue;
        // its one live tree edge
        for (uint32_t h = L.nb[v]; h != MST_EMPTY; h = L.enext[h]) {
            const uint32_t t = h >> 1;
            unsigned long long& x = L.e[L.tedge[t]];
            if (!(x >> 63)) continue;
            x &= ~(1ull << 63);
            const uint32_t a = (uint32_t)(x >> 16) & 0xFFFFu, c = (uint32_t)x & 0xFFFFu;
            const uint32_t o = a == v ? c : a;
            L.deg[v] = 0;
            for (uint32_t q = 0; q < g; ++q) cover[q] -= (b >> q) & 1u;
            if (--L.deg[o] == 1) L.par[sp++] = o;
            break;
        }
    }
    uint32_t cost = 0, ovf = 0;
    for (uint32_t t = 0; t < nt; ++t) {
        const unsigned long long x = L.e[L.tedge[t]];
        if (x >> 63) cost = add_cost<FLT>(cost, (uint32_t)((x >> 32) & 0x7FFFFFFFu), ovf);
    }
    return ovf ? GST_INF_COST : cost;
}

// The spider's tree: the centre c and its cost come from *spw (cost << 32 | c); the union of c's
// shortest paths (level-1 predecessors) is MST post-processed and pruned, and *spw's cost is lowered
// to the tree's when that is cheaper.  When tree != nullptr its edges go there as (u, v, weight
// bits) triples after a count word (at most MST_SMAX - 1 edges); nothing is written when the set or
// its edges overflow the tables.  One workgroup of MST_THREADS threads.
template <bool FLT>
__device__ void spider_tree_body(MstLds& L, const uint64_t* T1, uint32_t g, const uint32_t* gm, const uint32_t* row,
                                 const uint2* arc, unsigned long long* spw, uint32_t* tree) {
    const unsigned long long w0 = *spw;
    if (w0 == GST_INF_WORD || g < 2) return;
    const uint32_t c = (uint32_t)w0, sp0 = (uint32_t)(w0 >> 32);
    mst_clear(L);
    if (threadIdx.x < g) {
        // thread i walks group i's path from c: level-1 predecessors lead to a member (leaf word)
        const uint32_t i = threadIdx.x;
        uint32_t v = c;
        bool ok = mst_insert(L, v);
        for (uint32_t hop = 0; ok; ++hop) {
            const uint64_t w = T1[(uint64_t)v * g + i];
            const uint32_t pred = (uint32_t)w;
            if ((uint32_t)(w >> 32) == GST_INF_COST || hop == MST_WALK) { ok = false; break; }
            if (pred & GST_PRED_LEAF) break;   // v is a member of group i
            v = pred;
            ok = mst_insert(L, v);
        }
        if (!ok) L.fail = 1;
    }
    mst_build<FLT>(L, row, arc);
    if (L.fail) return;
    if (threadIdx.x == 0) {
        const uint32_t full = (1u << g) - 1u;
        const uint32_t cost = mst_prune_cost<FLT>(L, g, [&](uint32_t v) { return gm[v] & full; });
        if (cost != GST_INF_COST && cost < sp0) atomicMin(spw, ((unsigned long long)cost << 32) | c);
        if (tree) {
            uint32_t m = 0;
            for (uint32_t t = 0; t < L.nt; ++t) {
                const unsigned long long x = L.e[L.tedge[t]];
                if (!(x >> 63)) continue;
                tree[1 + 3 * m] = L.vs[(x >> 16) & 0xFFFFu];
                tree[2 + 3 * m] = L.vs[x & 0xFFFFu];
                tree[3 + 3 * m] = (uint32_t)((x >> 32) & 0x7FFFFFFFu);
                ++m;
            }
            tree[0] = m;
        }
    }
}

// a job bound (cost bits ub, GST_INF_COST: none) at or below 90% of the spider's cost leaves the spider
// no use: post-processing lowers it by a few per cent
template <bool FLT>
__device__ __forceinline__ bool spider_tree_useless(uint32_t ub, uint32_t sp) {
    if (ub == GST_INF_COST || sp == GST_INF_COST) return false;
    return FLT ? __uint_as_float(ub) <= 0.9f * __uint_as_float(sp) : (uint64_t)ub * 10u <= (uint64_t)sp * 9u;
}

template <bool FLT>
__global__ void __launch_bounds__(MST_THREADS) k_spider_tree(const uint64_t* T1, uint32_t g, const uint32_t* gm,
                                                           const uint32_t* row, const uint2* arc,
                                                           unsigned long long* spw, uint32_t* tree, uint32_t ub) {
    if (spider_tree_useless<FLT>(ub, (uint32_t)(*spw >> 32))) return;
    __shared__ MstLds L;
    spider_tree_body<FLT>(L, T1, g, gm, row, arc, spw, tree);
}

// The reference's post-processing of a caller's vertex set: out[0] = edge count (MST_EMPTY when the
// set or its induced edges overflow the tables), then (u, v, weight bits) per forest edge.
template <bool FLT>
__global__ void __launch_bounds__(MST_THREADS) k_mst_induced(const uint32_t* verts, uint32_t nverts, const uint32_t* row,
                                                           const uint2* arc, uint32_t* out) {
    __shared__ MstLds L;
    mst_clear(L);
    for (uint32_t i = threadIdx.x; i < nverts; i += blockDim.x)
        if (!mst_insert(L, verts[i])) L.fail = 1;
    mst_build<FLT>(L, row, arc);
    if (threadIdx.x != 0) return;
    if (L.fail) { out[0] = MST_EMPTY; return; }
    for (uint32_t t = 0; t < L.nt; ++t) {
        const unsigned long long x = L.e[L.tedge[t]];
        out[1 + 3 * t] = L.vs[(x >> 16) & 0xFFFFu];
        out[2 + 3 * t] = L.vs[x & 0xFFFFu];
        out[3 + 3 * t] = (uint32_t)((x >> 32) & 0x7FFFFFFFu);
    }
    out[0] = L.nt;
}

// g == 1: the full mask is level 1; its minimum is any member (cost 0)
__global__ void k_root_scan_l1(const uint64_t* T1, uint32_t n, unsigned long long* root) {
    unsigned long long best = GST_INF_WORD;
    const uint64_t tid0 = (uint64_t)blockIdx.x * blockDim.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t base = tid0; base < n; base += stride) {
        const uint64_t v = base + threadIdx.x;
        if (v < n) {
            const unsigned long long wv = T1[v];
            if (wv != GST_INF_WORD) {
                const unsigned long long rw = ((wv >> 32) << 32) | v;
                best = rw < best ? rw : best;
            }
        }
    }
    best = wave_min64(best);
    if (lane_id() == 0 && best != GST_INF_WORD) atomicMin(root, best);
}

// ---------------------------------------------------------------------------------------
// traceback on the device: one wavefront walks the predecessor structure from (root, full)
// ---------------------------------------------------------------------------------------
struct TraceArgs {
    const uint64_t* T;
    const uint32_t* pos_of;   // mask -> row position
    const uint32_t* row;
    const uint2* arc;         // arc records: .x = neighbour, .y = weight bits
    uint32_t flt;             // weights are float32 bits
    const unsigned long long* root;
    uint32_t* out_u;
    uint32_t* out_v;
    uint32_t* out_w;
    uint32_t* out_n;
    RoundCtl* ctl;
    uint64_t lvl_base[GST_MAX_GROUPS + 2];
    uint32_t m[GST_MAX_GROUPS + 2];
    uint32_t off[GST_MAX_GROUPS + 2];
    uint32_t full;
    uint32_t cap;             // capacity of the edge output
    uint32_t max_iter;
    uint32_t ostride;         // edge i is out_u[i * ostride], out_v[i * ostride], out_w[i * ostride]
};

// Weight bits of the arc (v, x) for a grown state (v, X) of cost bits cv whose predecessor is x with
// cost bits cx.  The grow set cv = cx + w(v, x) (ints exactly; floats rounded once), and rows are
// ordered by (weight, neighbour): a binary search for the first arc of weight >= the estimate (less a
// rounding margin for floats) and a short forward scan find it; parallel arcs were collapsed, so the
// neighbour is met once.  A full scan of the row is the fallback.  found = false: no such arc.
__device__ __forceinline__ uint32_t trace_arc_weight(const TraceArgs& a, uint32_t v, uint32_t x, uint32_t cv,
                                                     uint32_t cx, bool& found) {
    const uint32_t rs = a.row[v], re = a.row[v + 1];
    uint32_t wlo, whi;
    if (a.flt) {
        const float fv = __uint_as_float(cv), fx = __uint_as_float(cx);
        const float d = fv - fx, eps = fv * 1e-5f + 1e-30f;
        wlo = __float_as_uint(fmaxf(d - eps, 0.0f));
        whi = __float_as_uint(d + eps);
    } else {
        wlo = cv >= cx ? cv - cx : 0u;
        whi = wlo;
    }
    uint32_t lo = rs, hi = re;   // first arc of weight >= wlo
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a.arc[mid].y < wlo) lo = mid + 1; else hi = mid;
    }
    for (uint32_t e = lo; e < re; ++e) {
        const uint2 r = a.arc[e];
        if (r.y > whi) break;
        if (r.x == x) { found = true; return r.y; }
    }
    for (uint32_t e = rs; e < re; ++e) {
        const uint2 r = a.arc[e];
        if (r.x == x) { found = true; return r.y; }
    }
    found = false;
    return 0u;
}

__device__ __forceinline__ void trace_body(const TraceArgs& a) {
    __shared__ uint64_t stk[TRACE_STACK];
    __shared__ uint32_t sp_s, nout_s, fail_s;
    const uint32_t lane = threadIdx.x;
    if (lane == 0) {
        const unsigned long long rw = *a.root;
        sp_s = 0; nout_s = 0; fail_s = 0;
        if (rw != GST_INF_WORD) {
            stk[0] = ((uint64_t)a.full << 32) | (uint32_t)(rw & 0xFFFFFFFFu);
            sp_s = 1;
        }
    }
    __syncthreads();
    uint32_t iter = 0;
    while (true) {
        const uint32_t sp = sp_s;
        if (sp == 0 || fail_s) break;
        if (++iter > a.max_iter) { if (lane == 0) fail_s = 1; break; }
        const uint32_t take = min(sp, (uint32_t)WAVE);
        uint64_t item = 0;
        const bool act = lane < take;
        if (act) item = stk[sp - 1 - lane];
        __syncthreads();
        if (lane == 0) sp_s = sp - take;
        __syncthreads();
        uint32_t npush = 0;
        uint64_t push0 = 0, push1 = 0;
        uint32_t eu = 0, ev = 0, ew = 0;
        bool emit = false;
        if (act) {
            const uint32_t v = (uint32_t)item;
            const uint32_t X = (uint32_t)(item >> 32);
            const uint32_t pos = a.pos_of[X];
            const uint32_t k = (uint32_t)__popc(X);
            const uint32_t rk = pos - a.off[k];
            const uint64_t wv = a.T[a.lvl_base[k] + (uint64_t)v * a.m[k] + rk];
            const uint32_t pred = (uint32_t)wv;
            if (wv == GST_INF_WORD) {
                fail_s = 1;
            } else if (pred == GST_PRED_LEAF) {
                // v covers X by itself
            } else if (pred & 0x80000000u) {
                const uint32_t X1 = pred & 0x7FFFFFFFu;
                if (X1 == 0 || (X1 & ~X) || X1 == X) { fail_s = 1; }
                else {
                    push0 = ((uint64_t)X1 << 32) | v;
                    push1 = ((uint64_t)(X ^ X1) << 32) | v;
                    npush = 2;
                }
            } else {
                const uint32_t u = pred;   // the neighbour this state grew from
                const uint64_t wu = a.T[a.lvl_base[k] + (uint64_t)u * a.m[k] + rk];
                bool found = false;
                if (a.flt) {
                    ew = trace_arc_weight(a, v, u, (uint32_t)(wv >> 32), (uint32_t)(wu >> 32), found);
                } else {
                    // ints: the grow set cost(v) = cost(u) + w(v, u) exactly, and a later drop of
                    // cost(u) (<= gmax, so grown again) would have lowered cost(v) or moved its
                    // predecessor: the weight is the difference
                    const uint32_t cv = (uint32_t)(wv >> 32), cu = (uint32_t)(wu >> 32);
                    found = cu <= cv;
                    ew = cv - cu;
                }
                if (!found || wu == GST_INF_WORD) fail_s = 1;
                eu = v; ev = u;
                emit = found;
                push0 = ((uint64_t)X << 32) | u;
                npush = found ? 1 : 0;
            }
        }
        __syncthreads();
        if (act) {
            if (npush) {
                const uint32_t p = atomicAdd(&sp_s, npush);
                if (p + npush > TRACE_STACK) fail_s = 1;
                else {
                    stk[p] = push0;
                    if (npush == 2) stk[p + 1] = push1;
                }
            }
            if (emit) {
                const uint32_t o = atomicAdd(&nout_s, 1u);
                if (o < a.cap) {
                    const uint64_t oo = (uint64_t)o * a.ostride;
                    a.out_u[oo] = eu; a.out_v[oo] = ev; a.out_w[oo] = ew;
                }
                else fail_s = 1;
            }
        }
        __syncthreads();
    }
    __syncthreads();
    if (lane == 0) {
        *a.out_n = nout_s;
        if (fail_s) atomicOr(&a.ctl->err, 2u);
    }
}

__global__ void __launch_bounds__(WAVE) k_traceback(TraceArgs a) { trace_body(a); }

int grid_for(uint64_t work, int block, int max_blocks) {
    uint64_t b = (work + block - 1) / block;
    if (b < 1) b = 1;
    if (b > (uint64_t)max_blocks) b = max_blocks;
    return (int)b;
}

}  // namespace
