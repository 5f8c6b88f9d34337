/*
 * gst.h — C-ABI of the MI355X group Steiner tree (DPBF) solver.
 *
 * The reference exposes its solver as plain C++ free functions:
 *   DPBF_GPU(...)                              rucgraph/DPQ.cuh:57, src/DPQ.cu:501-768
 *   graph_v_of_v_idealID_DPBF_only_ec(...)     rucgraph/graph_v_of_v_idealID_DPBF_only_ec.h:133-134
 * with the input graph converted to CSR by toCSR (rucgraph/graph_v_of_v_idealID/csr_graph.hpp:29-65)
 * and groups given as "group vertices" whose adjacency lists hold the members
 * (rucgraph/graph_hash_of_mixed_weighted_generate_random_groups_of_vertices.h:42-47).
 *
 * This header replaces that boundary with a C ABI: plain pointers and sizes, status codes, no
 * exceptions, no C++ or torch types.  The C++ wrapper in gst.hpp rebuilds the reference's host
 * API (graph_v_of_v_idealID in, tree + cost out) on top of it.
 *
 * Ownership: the caller owns every input array; arrays inside a gst_result are owned by the
 * library and released by gst_result_free.  A gst_ctx is single-threaded (one host thread at a
 * time); separate contexts may be used concurrently.
 */
#ifndef GST_H
#define GST_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GST_MAX_GROUPS 12      /* compile-time maximum number of groups g (state table holds 2^g-1 words/vertex) */

typedef enum {
    GST_OK = 0,
    GST_INFEASIBLE = 1,        /* no connected tree touches every group (reference: empty tree, src/DPQ.cu:620-624,723-726) */
    GST_TOO_MANY_GROUPS = 2,   /* g > GST_MAX_GROUPS (reference: exit(1) for g>=20, ...DPBF_only_ec.h:139-142) */
    GST_OOM = 3,               /* device allocation failed */
    GST_INVALID_ARG = 4,       /* malformed CSR, negative / non-finite weight, vertex id out of range, ... */
    GST_HIP_ERROR = 5,         /* a HIP runtime call or kernel failed */
    GST_COST_OVERFLOW = 6,     /* an integer path cost exceeded the 32-bit cost field */
    GST_NO_DEVICE = 7,         /* no usable GPU */
    GST_INTERNAL = 8           /* internal consistency check failed (traceback did not close, ...) */
} gst_status;

typedef enum { GST_W_INT32 = 0, GST_W_FLOAT32 = 1 } gst_weight_type;

/* Undirected graph as symmetric CSR (both arcs of every edge present).  Replaces CSR_graph /
 * toCSR (rucgraph/graph_v_of_v_idealID/csr_graph.hpp:11-65), which truncates weights to int;
 * here int32 weights stay int32 and float32 weights stay float32. */
typedef struct {
    int32_t        n;          /* vertices 0..n-1 */
    int64_t        m;          /* arcs = row_ptr[n] */
    const int64_t* row_ptr;    /* n+1 */
    const int32_t* col;        /* m */
    const void*    w;          /* m weights, int32_t or float, >= 0 */
    int32_t        w_type;     /* gst_weight_type */
} gst_csr;

/* g groups as a CSR of member lists.  Group i (bit i of the DP masks) holds
 * grp_vtx[grp_ptr[i] .. grp_ptr[i+1]).  Members may repeat across groups. */
typedef struct {
    int32_t        n_groups;
    const int32_t* grp_ptr;    /* n_groups+1 */
    const int32_t* grp_vtx;
} gst_groups;

typedef struct {
    double  ub;                /* known feasible-tree cost U: states whose cost (+ lower bound) exceeds U
                                  are dropped, so the solve returns the optimum when it is <= U and
                                  GST_INFEASIBLE otherwise.  <= 0 or not finite: no bound (plain DPBF) */
    int32_t prune_lb;          /* with a bound: also drop (v,X) when cost + max_{i not in X} dist(v,G_i) >= ub */
    int32_t want_tree;         /* 1: run the device traceback and return the edge list */
    int32_t rounds_per_sync;   /* grow rounds launched per host check (0 = tuned default) */
    int32_t chunk_arcs;        /* arcs per frontier work item (0 = tuned default) */
    int32_t pull_permille;     /* a grow round whose queue holds at least this many per mille of all work
                                  items pulls over every vertex instead of pushing (0 = tuned default,
                                  < 0 = never pull) */
    int32_t grow_mode;         /* 0 = tuned (pull/push rounds), 1 = rounds only, 2 = asynchronous device work
                                  queue for every level, 3 = rounds while the frontier is dense, then the
                                  work queue */
    int32_t push_lanes;        /* lanes per frontier item in push rounds: 8, 16, 32 or 64; 1 = one item per lane, its
                                  masks and arcs spread over the wave (0 = tuned) */
    int32_t half_bound;        /* with a bound U: 0 = grow only states of cost <= U/2 (exact: every optimal
                                  tree splits at a point of its weighted centre into parts of weight
                                  <= OPT/2, whose states are then all exact, and merges plus one last
                                  grow across the centre edge rebuild it; the tree-decomposition rule
                                  of PrunedDP, rucgraph/graph_v_of_v_idealID_PrunedDPPlusPlus.h:795,
                                  src/DPQ.cu:255-258); -1 = grow every state below U */
    int32_t merge_tile_kb;     /* LDS budget of a merge tile's cost rows in KiB, 1..48 (0 = tuned default) */
    int32_t auto_bound;        /* 1: once the group distances (level 1) are final, bound the rest of the solve
                                  by the cheapest spider, min over vertices c of sum_i dist(c, G_i): the union
                                  of those shortest paths holds a tree touching every group, so the bound is
                                  feasible (with g >= 3; the half_bound rule applies to it too) */
    int32_t tour_bound;        /* with prune_lb, a bound and g >= 4: 0 = also use the tour bounds of PrunedDP++
                                  once level 1 is final (walks around the completing tree through the
                                  missing groups; rucgraph/graph_v_of_v_idealID_PrunedDPPlusPlus.h:435-494,
                                  268-385, 565-584), unless the walk through all groups is too short to
                                  prune (below ~U/5: overlapping or close groups); -1 = one-label only */
    int32_t reserved[1];
} gst_opts;

typedef struct {
    uint64_t grow_items;       /* frontier work items dequeued */
    uint64_t grow_arcs;        /* arcs scanned by those items */
    uint64_t grow_relax;       /* (arc, mask) relaxations attempted */
    uint64_t grow_updates;     /* relaxations that lowered a cost */
    uint64_t grow_enqueued;    /* work items enqueued */
    uint64_t merge_vertices;   /* vertex rows merged */
    uint64_t merge_pairs;      /* (submask, complement) pairs evaluated */
    uint64_t grow_bytes;       /* algorithmic bytes of the grow kernel: 24 per item, 8 per enqueue,
                                  8 per arc, 8 per relaxation, 16 per update */
    uint64_t merge_bytes;      /* algorithmic bytes of the merge kernel: rows read, words written, flags */
    uint32_t grow_rounds;
    uint32_t grow_launches;
    uint32_t merge_launches;
    uint32_t async_launches;   /* grow levels finished by the asynchronous work queue */
    double   t_total_ms;       /* wall time of the solve, host clock */
    double   t_grow_ms;        /* device time of grow kernels (HIP events) */
    double   t_merge_ms;       /* device time of merge kernels (HIP events) */
    double   t_trace_ms;       /* device time of the traceback */
    uint64_t grow_new_states;  /* grow updates that made an infinite state finite (the others lowered a finite
                                  cost again: grow_updates / grow_new_states is the rounds' redundancy) */
    uint64_t grow_atomics;     /* device-scope atomic operations of the grow kernels (min, compare-and-swap,
                                  change-bit OR, queue-flag exchange): each is a random 64-byte line
                                  read-modify-write beyond the L2, the grow's costliest operation */
    double   spider_cost;      /* automatic bound (one-query solves): min over centres c of sum_i dist(c, G_i),
                                  0 when not computed */
    double   spider_tree_cost; /* the same centre's shortest paths MST post-processed and pruned
                                  (<= spider_cost; the bound used), 0 when not computed */
} gst_stats;

typedef struct {
    double   cost;             /* optimal DP cost (exact for int32, fp32 DP value for float32) */
    double   tree_cost;        /* sum of returned edge weights, recomputed in double */
    int32_t  root;             /* root vertex of the returned tree (-1 when none) */
    int32_t  n_edges;
    int32_t* eu;               /* n_edges, library-owned */
    int32_t* ev;
    double*  ew;
    uint8_t  used_ub;          /* a finite bound was applied */
    uint8_t  fell_back;        /* learned path: the bounded solve was replaced by plain DPBF */
    uint8_t  pad[2];
    int32_t  n_groups;
    gst_stats stats;
} gst_result;

typedef struct gst_ctx   gst_ctx;
typedef struct gst_graph gst_graph;

/* --- context -------------------------------------------------------------------------- */
/* Replaces the implicit device/global state of src/DPQ.cu:21 (global vectors) and the per-call
 * allocations of src/DPQ.cu:525-542; one context per device and host thread. */
gst_status  gst_ctx_create(int32_t device, gst_ctx** out);
void        gst_ctx_destroy(gst_ctx* ctx);
const char* gst_last_error(const gst_ctx* ctx);
const char* gst_status_str(gst_status s);

/* --- graph ---------------------------------------------------------------------------- */
/* Device-resident CSR reused across queries.  Replaces toCSR
 * (rucgraph/graph_v_of_v_idealID/csr_graph.hpp:29-65): each row is ordered by (weight, neighbour id)
 * so that relaxations can stop at the first arc that is too heavy, parallel arcs collapse to their
 * lightest weight, self-loops are dropped, and each arc gets its reverse-arc index. */
gst_status gst_graph_upload(gst_ctx* ctx, const gst_csr* csr, gst_graph** out);
void       gst_graph_free(gst_graph* g);
int32_t    gst_graph_num_vertices(const gst_graph* g);
int64_t    gst_graph_num_arcs(const gst_graph* g);

/* --- solve ---------------------------------------------------------------------------- */
/* One exact GST query.  Replaces DPBF_GPU (rucgraph/DPQ.cuh:57, src/DPQ.cu:501-768): the
 * returned edge list is the tree, result->cost is *real_cost, stats replace *pointer1 / *rt.
 * opts may be NULL (plain DPBF, tree returned). */
gst_status gst_solve(gst_ctx* ctx, const gst_graph* g, const gst_groups* groups,
                     const gst_opts* opts, gst_result* result);
void       gst_result_free(gst_result* r);

/* Tree check used by the tests and the C++ wrapper; extends this_is_a_feasible_solution_gpu
 * (rucgraph/test_gpu.h:45-77) and graph_hash_of_mixed_weighted_sum_of_ec
 * (rucgraph/graph_hash_of_mixed_weighted_sum_of_nw_ec.h:58-87): the edge set must be a tree
 * (connected, |E| = |V|-1, every edge present in the graph with that weight) touching every
 * group.  *cost_out receives the double sum of the edge weights.  A single vertex touching
 * every group is a valid tree of cost 0 (pass n_edges = 0 and root). */
gst_status gst_validate_tree(const gst_csr* csr, const gst_groups* groups, int32_t n_edges,
                             const int32_t* eu, const int32_t* ev, int32_t root,
                             double* cost_out);

/* MST post-processing of a vertex set (rucgraph/graph_hash_of_mixed_weighted_MST_postprocessing.h:73-103,
 * graph_hash_of_mixed_weighted_MST_postprocessing_hash_of_vertices, whose spanning tree is the Prim of
 * graph_hash_of_mixed_weighted/common_algorithms/graph_hash_of_mixed_weighted_minimum_spanning_tree.h:29-132):
 * the minimum spanning forest of the subgraph of `graph` induced by vertices[0..nv) (repeats ignored),
 * computed on the device by one workgroup.  On return *ne edges (eu, ev, ew; arrays of nv entries)
 * and *cost = the double sum of their weights.  At most 512 distinct vertices with at most 2048
 * induced edges (GST_INVALID_ARG beyond).  Where the set is disconnected the reference's Prim stops
 * after its first vertex's component; this returns a spanning forest of every component (on the
 * connected sets the reference passes, the two agree). */
gst_status gst_mst_postprocess(gst_ctx* ctx, const gst_graph* graph, const int32_t* vertices, int32_t nv,
                               int32_t* eu, int32_t* ev, double* ew, int32_t* ne, double* cost);

/* --- one query without allocation ------------------------------------------------------- */
/* No counterpart in the reference.  For callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid allocation): the same exact solve as gst_solve (same
 * kernels, the round path of grow_mode 0/1; 2 and 3 are rejected), with every per-query device buffer
 * carved out of one fixed device array of GST_NOALLOC_SCRATCH_BYTES per GPU (held from the first
 * context on that GPU until the process ends) and host work in buffers made by gst_ctx_create.  The
 * call makes no device or host allocation: it rejects a query whose buffers exceed the array
 * (GST_INVALID_ARG, with the size it needs in gst_last_error) or whose tree exceeds
 * GST_NOALLOC_MAX_EDGES edges.  The graph is uploaded beforehand (gst_graph_upload allocates).
 * With want_tree, the tree's edges go into the caller's eu/ev/ew (edge_cap entries each) and
 * result->n_edges counts them; result->eu/ev/ew stay NULL (gst_result_free is not needed).  No
 * HIP events are made: the t_*_ms stats stay 0.  The array is shared by every call on one GPU and
 * keeps the last call's contents (each call clears what must start at zero): calls that overlap on
 * one GPU, from two contexts of one device used by two threads at once, race on it. */
#define GST_NOALLOC_SCRATCH_BYTES (128u << 20)
#define GST_NOALLOC_MAX_EDGES 16384
gst_status gst_solve_noalloc(gst_ctx* ctx, const gst_graph* g, const gst_groups* groups, const gst_opts* opts,
                             int32_t edge_cap, int32_t* eu, int32_t* ev, double* ew, gst_result* result);

/* --- batched queries --------------------------------------------------------------------- */
/* Many exact queries on one graph in flight together.  Replaces calling DPBF_GPU once per query
 * (rucgraph/test_gpu.h:188-283 loops over queries; src/DPQ.cu:501-768 syncs with the host every
 * round).  Up to `slots` queries hold their own state tables; one launch per grow round relaxes
 * the current round of every query in flight, the grid split over the queries by their frontier
 * sizes; the merges, frontier compactions and tracebacks of all queries that reached them run in
 * one launch each; a step (those launches plus rounds_per_step grow rounds) is replayed as a HIP
 * graph.  Each query gets what gst_solve returns for it: the same exact cost (the tree may be
 * another tree of that cost). */
typedef struct {
    int32_t slots;             /* queries in flight, at most 256 (0: as many as fit in half of the free device
                                  memory, at most 32 and at most the number of valid queries) */
    int32_t rounds_per_step;   /* grow rounds per step between host checks (0: tuned, else 8) */
    int32_t use_graph;         /* 1: replay each step as a HIP graph (default), 0: launch kernel by kernel */
    int32_t time_kernels;      /* 1: launch kernel by kernel with HIP events around every batched grow and
                                  merge launch (t_grow_ms / t_merge_ms in the stats) */
    int32_t max_steps;         /* > 0: stop with GST_INTERNAL after this many steps (0: a bound of minutes of
                                  work that only a defect reaches); every query not finished by then
                                  gets that status */
    int32_t reserved[3];
} gst_batch_opts;

typedef struct {
    double   t_total_ms;       /* wall time of the call, host clock */
    double   queries_per_s;
    double   latency_p50_ms;   /* per query: from entering a slot to its result being read */
    double   latency_p99_ms;
    uint64_t steps;
    uint64_t grow_launches;
    uint64_t grow_relax;       /* (arc, mask) relaxations over all queries */
    uint64_t grow_bytes;       /* algorithmic bytes (same model as gst_stats) */
    uint64_t merge_bytes;
    int64_t  queries;          /* queries finished (solved or failed) */
    int32_t  slots;
    int32_t  rounds_per_step;
    int32_t  used_graph;
    int32_t  timed_kernels;
    double   t_grow_ms;        /* with time_kernels: device time of the batched grow launches (HIP events) */
    double   t_merge_ms;       /* with time_kernels: device time of the batched merge launches */
    uint64_t merge_launches;
    uint64_t busy_slot_steps;  /* slots holding a query, summed over steps (/ steps: mean slots busy) */
    int32_t  reserved[2];
} gst_batch_stats;

/* groups[nq]; opts: NULL (plain DPBF, trees returned) or nq entries, of which ub, prune_lb and
 * want_tree are used; bopts may be NULL.  results[nq] and status[nq] receive each query's outcome:
 * a failed query does not stop the others, and the call returns the first failing status (or
 * GST_OK).  stats may be NULL. */
gst_status gst_solve_batch(gst_ctx* ctx, const gst_graph* g, int32_t nq, const gst_groups* groups,
                           const gst_opts* opts, const gst_batch_opts* bopts, gst_result* results,
                           gst_status* status, gst_batch_stats* stats);

/* --- learned upper bound ----------------------------------------------------------------- */
/* No counterpart in the reference.  A two-layer mean-aggregation message-passing network (hidden
 * 64, BF16 weights, FP32 accumulation, BF16 MFMA kernels) scores vertices; the DPBF optimum of the
 * subgraph induced by the top-scored vertices plus every group member is a feasible tree whose
 * cost bounds the full solve (cost + lower bound >= bound prunes states, as PrunedDP++ does with
 * its bounds, rucgraph/graph_v_of_v_idealID_PrunedDPPlusPlus.h:716-730, 795-957).  The bounded
 * full solve returns the exact optimum; without a feasible subgraph tree the plain solve runs
 * (result->fell_back = 1).  Parameters are NPARAM fp32 values (gst_model_num_params). */
typedef struct gst_model gst_model;
typedef struct {
    uint32_t sub_vertices;     /* vertices of the induced subgraph */
    uint32_t sub_arcs;
    double   ub;               /* subgraph optimum used as the bound (-1: none) */
    double   t_score_select_ms;/* network forward + top-k + subgraph build */
    double   t_sub_solve_ms;   /* DPBF on the subgraph */
    double   t_full_solve_ms;  /* bounded DPBF on the full graph */
} gst_learned_stats;

int64_t    gst_model_num_params(void);
gst_status gst_model_create(gst_ctx* ctx, uint64_t seed, gst_model** out);
void       gst_model_free(gst_model* m);
gst_status gst_model_set_params(gst_model* m, const float* params);
gst_status gst_model_get_params(gst_model* m, float* params);
/* scores[n] (host) for one query */
gst_status gst_model_scores(gst_model* m, const gst_graph* g, const gst_groups* groups, float* scores);
/* forward + backward on one query with labels[n] in {0,1} (1 = vertex of the optimal tree):
 * weighted logistic loss (weight pos_weight on positives, mean over vertices).  Gradients add
 * into the model's accumulator; grad (NPARAM floats, may be NULL) receives this query's gradient. */
gst_status gst_model_loss_grad(gst_model* m, const gst_graph* g, const gst_groups* groups, const uint8_t* labels,
                               float pos_weight, float* loss, float* grad);
/* Adam step with the mean of the accumulated gradients, then clears the accumulator */
gst_status gst_model_adam_step(gst_model* m, float lr);
/* exact solve with the learned bound; topk_frac = fraction of vertices kept (e.g. 0.05) */
gst_status gst_solve_learned(gst_ctx* ctx, const gst_graph* g, const gst_groups* groups, gst_model* m,
                             double topk_frac, const gst_opts* opts, gst_result* result, gst_learned_stats* stats);

/* Many exact queries with the learned bound, pipelined through the batched engine (see
 * gst_solve_batch): for each query the network scores, the top-k selection and the induced subgraph
 * are computed on the device; the subgraph DPBF runs as one engine job and its optimum bounds the
 * full solve, which runs as a second engine job (plain DPBF when there is no subgraph tree).  So the
 * subgraph solves and bounded full solves of many queries share launches.  Each result is the
 * exact optimum, as from gst_solve.  opts (NULL or nq entries): prune_lb, want_tree, push_lanes and
 * half_bound are used (ub is set from the subgraph).  lstats (NULL or nq entries) receives each
 * query's subgraph size, bound and times (t_sub_solve_ms / t_full_solve_ms: job latency in the
 * engine).  The call runs the selections on a host thread of its own (device set to the context's,
 * joined before the call returns) beside the calling thread, which runs the engine; the model keeps
 * its subgraph buffers for later calls.  No counterpart in the reference. */
gst_status gst_solve_batch_learned(gst_ctx* ctx, const gst_graph* g, int32_t nq, const gst_groups* groups,
                                   gst_model* m, double topk_frac, const gst_opts* opts,
                                   const gst_batch_opts* bopts, gst_result* results, gst_status* status,
                                   gst_learned_stats* lstats, gst_batch_stats* stats);

/* --- autotuner ----------------------------------------------------------------------- */
/* No counterpart in the reference (its launch shapes are fixed, src/DPQ.cu:597-610).  Times the
 * sample (queries of one group count on graph g) under candidate launch knobs and keeps the fastest:
 * for one-query solves the pull threshold (frontier size at which a grow round pulls over every
 * vertex), the arcs per frontier work item, the lanes per work item in push rounds and the merge
 * tile (vertices per merge workgroup); with GST_TUNE_BATCH also the batched engine's rounds per
 * step and lanes per item.  Later solves on this graph that leave a knob at 0 use the choice.
 * GST_TUNE_BOUNDED tunes bounded solves (bound = 1.05 x each sample query's optimum). */
#define GST_TUNE_BOUNDED 1
#define GST_TUNE_BATCH   2
#define GST_TUNE_EAGER   4   /* time the engine launching kernel by kernel (as with time_kernels), not by graph replay */
typedef struct {
    int32_t chunk_arcs, pull_permille, push_lanes, merge_tile_kb;   /* chosen for one-query solves */
    int32_t batch_rounds_per_step, batch_push_lanes;                /* chosen for the engine (GST_TUNE_BATCH) */
    int32_t configs_tried;
    int32_t pad;
    double  default_ms, tuned_ms;          /* mean ms per sample query: built-in defaults, choice */
    double  batch_default_qps, batch_tuned_qps;
} gst_tune_report;

gst_status gst_autotune(gst_ctx* ctx, const gst_graph* g, int32_t nq, const gst_groups* sample, int32_t flags,
                        gst_tune_report* report);

/* --- synthetic instances (host, deterministic) ------------------------------------------ */
/* Host CSR owned by the library.  The generators use the library's own seeded RNG
 * (splitmix64 / xoshiro256**), so a seed names the same instance on every machine. */
typedef struct {
    int32_t  n;
    int64_t  m;
    int64_t* row_ptr;
    int32_t* col;
    void*    w;
    int32_t  w_type;
} gst_host_csr;

void gst_host_csr_free(gst_host_csr* h);

/* Random connected graph: random spanning tree (vertex i attaches to a uniform earlier vertex)
 * plus uniform extra edges without duplicates or self-loops, int weights uniform in [wlo,whi].
 * Same recipe as graph_v_of_v_idealID_generate_random_connected_graph
 * (rucgraph/graph_v_of_v_idealID/random_graph/graph_v_of_v_idealID_generate_random_connected_graph.h:41-101). */
gst_status gst_gen_random_connected(int32_t n, int64_t n_edges, int32_t wlo, int32_t whi,
                                    uint64_t seed, gst_host_csr* out);
/* Barabasi-Albert preferential attachment: an m-clique seed, then each new vertex links to m
 * distinct earlier vertices chosen with probability proportional to degree.  Int weights
 * uniform in [wlo,whi]. */
gst_status gst_gen_barabasi_albert(int32_t n, int32_t m, int32_t wlo, int32_t whi,
                                   uint64_t seed, gst_host_csr* out);
/* R-MAT (a,b,c,d) with 2^scale vertices and edge_factor*2^scale generated edges, symmetrised,
 * deduplicated, self-loops removed.  w_type GST_W_INT32: uniform [wlo,whi];
 * GST_W_FLOAT32: lognormal(mu, sigma). */
gst_status gst_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b, double c,
                        int32_t w_type, int32_t wlo, int32_t whi, double mu, double sigma,
                        uint64_t seed, gst_host_csr* out);
/* g groups with sizes uniform in [smin,smax] (clamped to n), members uniform without
 * replacement within a group (groups may overlap), as in
 * graph_hash_of_mixed_weighted_generate_random_groups_of_vertices
 * (rucgraph/graph_hash_of_mixed_weighted_generate_random_groups_of_vertices.h:27-50).
 * grp_ptr must hold g+1 entries; grp_vtx must hold g*min(smax,n) entries. */
gst_status gst_gen_groups(int32_t n, int32_t g, int32_t smin, int32_t smax, uint64_t seed,
                          int32_t* grp_ptr, int32_t* grp_vtx);

#ifdef __cplusplus
}
#endif
#endif /* GST_H */
