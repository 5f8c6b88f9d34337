/*
 * dpbf_oracle.c -- CPU oracle for the group Steiner tree solver.  TEST INFRASTRUCTURE ONLY:
 * it is linked by tests/, by __graft_entry__.smoke() as the checker, and by bench.py's CPU
 * baseline leg; the product path never calls it.
 *
 * Contents
 *   orc_dpbf         label-setting DPBF (Ding et al. ICDE'07), restating
 *                    rucgraph/graph_v_of_v_idealID_DPBF_only_ec.h:133-324:
 *                      seeding  T(v,p)=0 for every non-empty p within v's groups      (:175-197)
 *                      pop the cheapest state; the first full-mask pop is optimal      (:202-223)
 *                      grow     T(u,p) <- T(v,p) + w(v,u)                              (:227-264)
 *                      merge    T(v,p|p2) <- T(v,p) + T(v,p2), p2 disjoint from p      (:267-316)
 *                      traceback from (root, full) through grow/merge records          (:90-129)
 *                    Differences, on purpose: costs are doubles with +inf for an absent
 *                    state (the reference adds INT_MAX for an absent T(v,p2) at :279-293,
 *                    which overflows); weights are not truncated to int (:231 truncates);
 *                    the heap is an indexed binary heap with ties broken by state id.
 *   orc_dijkstra     multi-source Dijkstra with initial potentials (textbook), used for the
 *                    single-mask levels and for the closed forms of g <= 4 in the tests.
 *   orc_bruteforce   exhaustive search over connected vertex subsets touching every group,
 *                    cost = minimum spanning tree of the induced subgraph (n <= 20).
 *   orc_dpbf_batch   orc_dpbf over many group sets with a pool of threads (CPU baseline).
 *   orc_dpbf_pruned  orc_dpbf under a known feasible cost U, with the bounds of PrunedDP++
 *                    (rucgraph/graph_v_of_v_idealID_PrunedDPPlusPlus.h:497-964):
 *                      a state is kept only while cost + LB(v,p) <= U, LB the one-label bound
 *                      max over groups i not in p of dist(v, G_i)                      (:471-488)
 *                      and optionally the tour bounds lb1/lb2                          (:435-494)
 *                      over a walk table W built from group-to-group distances    (:268-385, 565-584)
 *                      a popped state is grown only while cost <= U/2 (optional)       (:795)
 *                    it must return the same optimum as orc_dpbf whenever U >= OPT.  The walk
 *                    table is a Held-Karp DP over the metric closure of the group distances
 *                    instead of the reference's heap search over (i, j, set) with subset
 *                    propagation; both give lower bounds of the same walks, so only the amount
 *                    of pruning can differ, never the optimum.
 *   orc_state_bounds the bound of every state, for the per-state tests.
 *
 * Pinning (tests/test_oracle.py): brute force on tiny graphs, the g=1 (cost 0) and g=2
 * (group-to-group shortest path) special cases, the g=3 spider and g=4 two-centre closed forms
 * built from orc_dijkstra, and trees re-validated edge by edge.  The bounds: every state's bound
 * is at most the brute-force cost of the cheapest tree holding v and touching the missing groups,
 * the bounded search keeps the brute-force optimum, and a planted max-for-min mistake in lb1
 * loses it on some instance.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_OK 0
#define ORC_INFEASIBLE 1
#define ORC_OOM 3
#define ORC_INVALID 4

/* ---------------------------------------------------------------- indexed binary heap */
typedef struct {
    int64_t* h;      /* heap array of state ids */
    int32_t* pos;    /* state -> heap index, -1 absent, -2 popped */
    const double* key;
    int64_t n;
} heap_t;

static int heap_less(const heap_t* H, int64_t a, int64_t b) {
    const double ka = H->key[a], kb = H->key[b];
    if (ka != kb) return ka < kb;
    return a < b;
}

static void heap_up(heap_t* H, int64_t i) {
    const int64_t x = H->h[i];
    while (i > 0) {
        const int64_t p = (i - 1) / 2;
        if (!heap_less(H, x, H->h[p])) break;
        H->h[i] = H->h[p];
        H->pos[H->h[i]] = (int32_t)i;
        i = p;
    }
    H->h[i] = x;
    H->pos[x] = (int32_t)i;
}

static void heap_down(heap_t* H, int64_t i) {
    const int64_t x = H->h[i];
    for (;;) {
        int64_t c = 2 * i + 1;
        if (c >= H->n) break;
        if (c + 1 < H->n && heap_less(H, H->h[c + 1], H->h[c])) c++;
        if (!heap_less(H, H->h[c], x)) break;
        H->h[i] = H->h[c];
        H->pos[H->h[i]] = (int32_t)i;
        i = c;
    }
    H->h[i] = x;
    H->pos[x] = (int32_t)i;
}

static void heap_push_or_decrease(heap_t* H, int64_t s) {
    if (H->pos[s] >= 0) {
        heap_up(H, H->pos[s]);
    } else {
        H->h[H->n] = s;
        H->pos[s] = (int32_t)H->n;
        H->n++;
        heap_up(H, H->n - 1);
    }
}

static int64_t heap_pop(heap_t* H) {
    const int64_t top = H->h[0];
    H->n--;
    if (H->n > 0) {
        H->h[0] = H->h[H->n];
        H->pos[H->h[0]] = 0;
        heap_down(H, 0);
    }
    H->pos[top] = -2;
    return top;
}

/*
 * Tour bounds of PrunedDP++ (rucgraph/graph_v_of_v_idealID_PrunedDPPlusPlus.h:435-494, 268-385,
 * 565-584).  D[a][b] = distance between groups a and b = min over members u of G_b of dist(u, G_a);
 * Dc = its closure over intermediate groups (Floyd-Warshall), so that walks through extra groups
 * never beat the table; W[S][i][j] = shortest walk from group i to group j through every group of S
 * (Held-Karp over Dc; W[S][i][i] closes the walk).  For a state (v, X) with uncovered set U:
 *   lb1 = min over i, j in U of (d(v,i) + W[U][i][j] + d(v,j)) / 2     (a tour from v around the
 *         rest of the tree passes every group of U and is at most twice its weight)
 *   lb2 = max over i in U of (d(v,i) + min_j W[U][i][j] + min_j d(v,j)) / 2
 * and the bound is max(lb1, lb2, one-label).  tabs: W of 2^g * g * g doubles.
 */
static void orc_tour_tables(int32_t g, const double* D, double* W) {
    double Dc[20 * 20];
    for (int32_t a = 0; a < g * g; ++a) Dc[a] = D[a];
    for (int32_t k = 0; k < g; ++k)
        for (int32_t a = 0; a < g; ++a)
            for (int32_t b = 0; b < g; ++b)
                if (Dc[a * g + k] + Dc[k * g + b] < Dc[a * g + b]) Dc[a * g + b] = Dc[a * g + k] + Dc[k * g + b];
    const int64_t S = (int64_t)1 << g;
    for (int64_t x = 0; x < S * g * g; ++x) W[x] = INFINITY;
    for (int32_t i = 0; i < g; ++i) W[((int64_t)1 << i) * g * g + i * g + i] = 0.0;
    for (int64_t m = 1; m < S; ++m) {
        /* open walks i -> j (i != j) through all of m: the last step enters j from k */
        for (int32_t i = 0; i < g; ++i) {
            if (!((m >> i) & 1)) continue;
            for (int32_t j = 0; j < g; ++j) {
                if (j == i || !((m >> j) & 1)) continue;
                const int64_t prev = m ^ ((int64_t)1 << j);
                double best = INFINITY;
                for (int32_t k = 0; k < g; ++k) {
                    if (!((prev >> k) & 1)) continue;
                    const double c = W[prev * g * g + i * g + k] + Dc[k * g + j];
                    if (c < best) best = c;
                }
                W[m * g * g + i * g + j] = best;
            }
        }
        /* closed walks i -> i through all of m (at least two groups) */
        for (int32_t i = 0; i < g; ++i) {
            if (!((m >> i) & 1) || m == ((int64_t)1 << i)) continue;
            double best = INFINITY;
            for (int32_t k = 0; k < g; ++k) {
                if (k == i || !((m >> k) & 1)) continue;
                const double c = W[m * g * g + i * g + k] + Dc[k * g + i];
                if (c < best) best = c;
            }
            W[m * g * g + i * g + i] = best;
        }
    }
}

static double orc_tour_lb(const double* dist, const double* W, int32_t n, int32_t g, int32_t v, uint32_t p,
                          int32_t planted) {
    const uint32_t full = ((uint32_t)1 << g) - 1, U = full ^ p;
    if (!U) return 0.0;
    double lb1 = INFINITY, lb2 = 0.0, dmin = INFINITY;
    for (int32_t j = 0; j < g; ++j)
        if ((U >> j) & 1u && dist[(int64_t)j * n + v] < dmin) dmin = dist[(int64_t)j * n + v];
    for (int32_t i = 0; i < g; ++i) {
        if (!((U >> i) & 1u)) continue;
        const double di = dist[(int64_t)i * n + v];
        double w2 = INFINITY;
        for (int32_t j = 0; j < g; ++j) {
            if (!((U >> j) & 1u)) continue;
            const double w = W[(int64_t)U * g * g + i * g + j];
            if (w < w2) w2 = w;
            const double c = (di + w + dist[(int64_t)j * n + v]) / 2.0;
            /* planted mistake for the tests: max instead of min over the tour's end groups */
            if (planted ? (lb1 == INFINITY || c > lb1) : (c < lb1)) lb1 = c;
        }
        const double c2 = (di + w2 + dmin) / 2.0;
        if (c2 > lb2) lb2 = c2;
    }
    return lb1 > lb2 ? lb1 : lb2;
}

/* max over groups i not in p of dist[i][v] (dist: g x n) */
static double orc_label_lb(const double* dist, int32_t n, int32_t g, int32_t v, uint32_t p) {
    double lb = 0.0;
    for (int32_t i = 0; i < g; ++i)
        if (!((p >> i) & 1u) && dist[(int64_t)i * n + v] > lb) lb = dist[(int64_t)i * n + v];
    return lb;
}

/* use_lb: 1 one-label, 2 also the tour bounds, 3 the tour bounds with a planted mistake (tests) */
static double orc_lb(const double* dist, const double* W, int32_t n, int32_t g, int32_t v, uint32_t p, int32_t use_lb) {
    double lb = orc_label_lb(dist, n, g, v, p);
    if (use_lb >= 2) {
        const double t = orc_tour_lb(dist, W, n, g, v, p, use_lb == 3);
        if (t > lb) lb = t;
    }
    return lb;
}

/* The bound of every state (v, p) as orc_dpbf_pruned applies it, into out[v * 2^g + p]
 * (tests compare each with the cost of the cheapest tree that holds v and touches the missing groups). */
int orc_state_bounds(int32_t n, int32_t g, const uint32_t* gm, const double* dist, int32_t use_lb, double* out) {
    if (n < 1 || g < 1 || g > 12 || !dist || use_lb < 1 || use_lb > 3) return ORC_INVALID;
    const int64_t S = (int64_t)1 << g;
    double* W = NULL;
    if (use_lb >= 2) {
        W = (double*)malloc(sizeof(double) * (size_t)(S * g * g));
        if (!W) return ORC_OOM;
        double D[20 * 20];
        for (int32_t a = 0; a < g; ++a)
            for (int32_t b = 0; b < g; ++b) {
                double m = INFINITY;
                for (int32_t u = 0; u < n; ++u)
                    if (((gm[u] >> b) & 1u) && dist[(int64_t)a * n + u] < m) m = dist[(int64_t)a * n + u];
                D[a * g + b] = m;
            }
        orc_tour_tables(g, D, W);
    }
    for (int32_t v = 0; v < n; ++v)
        for (int64_t p = 0; p < S; ++p) out[(int64_t)v * S + p] = orc_lb(dist, W, n, g, v, (uint32_t)p, use_lb);
    free(W);
    return ORC_OK;
}

/* ---------------------------------------------------------------- DPBF */
/*
 * n, row(n+1), col(m), w(m): symmetric CSR; g groups with per-vertex bitmask gm(n).
 * Outputs: *cost (inf if infeasible), *root, edges (eu, ev, ew) up to cap, *n_edges;
 * stats[0] = states popped, stats[1] = grow relaxations, stats[2] = merge evaluations.
 */
int orc_dpbf(int32_t n, const int64_t* row, const int32_t* col, const double* w, int32_t g,
             const uint32_t* gm, double* cost_out, int32_t* root_out, int32_t* eu, int32_t* ev,
             double* ew, int32_t cap, int32_t* n_edges, int64_t* stats) {
    if (n < 1 || g < 1 || g > 20) return ORC_INVALID;
    const int64_t S = (int64_t)1 << g;
    const uint32_t full = (uint32_t)(S - 1);
    const int64_t NS = (int64_t)n * S;
    double* cost = (double*)malloc(sizeof(double) * (size_t)NS);
    int8_t* ty = (int8_t*)malloc((size_t)NS);
    int32_t* pa = (int32_t*)malloc(sizeof(int32_t) * (size_t)NS);
    int32_t* pos = (int32_t*)malloc(sizeof(int32_t) * (size_t)NS);
    int64_t* hp = (int64_t*)malloc(sizeof(int64_t) * (size_t)NS);
    if (!cost || !ty || !pa || !pos || !hp) {
        free(cost); free(ty); free(pa); free(pos); free(hp);
        return ORC_OOM;
    }
    for (int64_t s = 0; s < NS; ++s) { cost[s] = INFINITY; ty[s] = -1; pos[s] = -1; }
    heap_t H = {hp, pos, cost, 0};
    int64_t pops = 0, relax = 0, merges = 0;

    for (int32_t v = 0; v < n; ++v) {
        const uint32_t gv = gm[v] & full;
        if (!gv) continue;
        /* every non-empty p inside v's group set is a single-vertex tree */
        for (uint32_t p = gv; p; p = (p - 1) & gv) {
            const int64_t s = (int64_t)v * S + p;
            cost[s] = 0.0; ty[s] = 0;
            heap_push_or_decrease(&H, s);
        }
    }

    int32_t root = -1;
    while (H.n > 0) {
        const int64_t s = heap_pop(&H);
        pops++;
        const int32_t v = (int32_t)(s / S);
        const uint32_t p = (uint32_t)(s % S);
        const double cvp = cost[s];
        if (p == full) { root = v; break; }
        /* grow */
        for (int64_t e = row[v]; e < row[v + 1]; ++e) {
            const int32_t u = col[e];
            const int64_t su = (int64_t)u * S + p;
            if (pos[su] == -2) continue;
            relax++;
            const double c = cvp + w[e];
            if (c < cost[su]) {
                cost[su] = c; ty[su] = 1; pa[su] = (int32_t)e;   /* arc v->u in v's row */
                heap_push_or_decrease(&H, su);
            }
        }
        /* merge with every disjoint p2 that has a tree at v */
        const uint32_t comp = full ^ p;
        for (uint32_t p2 = comp; p2; p2 = (p2 - 1) & comp) {
            const int64_t s2 = (int64_t)v * S + p2;
            if (!(cost[s2] < INFINITY)) continue;
            merges++;
            const int64_t sx = (int64_t)v * S + (p | p2);
            if (pos[sx] == -2) continue;
            const double c = cvp + cost[s2];
            if (c < cost[sx]) {
                cost[sx] = c; ty[sx] = 2; pa[sx] = (int32_t)p;
                heap_push_or_decrease(&H, sx);
            }
        }
    }
    if (stats) { stats[0] = pops; stats[1] = relax; stats[2] = merges; }
    int rc = ORC_OK;
    int32_t ne = 0;
    if (root < 0) {
        *cost_out = INFINITY;
        rc = ORC_INFEASIBLE;
    } else {
        *cost_out = cost[(int64_t)root * S + full];
        /* traceback with an explicit stack of states; reuse hp as the stack */
        int64_t sp = 0;
        hp[sp++] = (int64_t)root * S + full;
        while (sp > 0) {
            const int64_t s = hp[--sp];
            const int32_t v = (int32_t)(s / S);
            const uint32_t p = (uint32_t)(s % S);
            if (ty[s] == 0) continue;
            if (ty[s] == 1) {
                const int64_t e = pa[s];
                /* arc e lies in the row of the source vertex x with col[e] == v */
                int32_t lo = 0, hi = n;  /* find x: row[x] <= e < row[x+1] */
                while (hi - lo > 1) {
                    const int32_t mid = (lo + hi) / 2;
                    if (row[mid] <= e) lo = mid; else hi = mid;
                }
                const int32_t x = lo;
                if (ne < cap && eu) { eu[ne] = x; ev[ne] = v; ew[ne] = w[e]; }
                ne++;
                hp[sp++] = (int64_t)x * S + p;
            } else if (ty[s] == 2) {
                const uint32_t p1 = (uint32_t)pa[s];
                hp[sp++] = (int64_t)v * S + p1;
                hp[sp++] = (int64_t)v * S + (p ^ p1);
            } else {
                rc = ORC_INVALID;
                break;
            }
        }
    }
    if (root_out) *root_out = root;
    if (n_edges) *n_edges = ne;
    free(cost); free(ty); free(pa); free(pos); free(hp);
    return rc;
}

/* ---------------------------------------------------------------- pruned DPBF */
/*
 * As orc_dpbf (cost only), under a bound U >= OPT: states with cost + lb > U are never stored,
 * and with grow_frac > 0 only states with cost <= grow_frac * U are grown (PrunedDP++: 1/2).  dist (g x n, may be NULL when
 * use_lb == 0) holds the exact group distances for the one-label lower bound.
 * use_lb & ORC_POP_UB (with dist): PrunedDP++'s feasible tree at every pop
 * (rucgraph/graph_v_of_v_idealID_PrunedDPPlusPlus.h:668-711): the popped tree T(v,p) plus a shortest
 * path from v to every missing group touches every group, so U = min(U, cost + sum_{i not in p}
 * dist(v, G_i)) (the cost of that union, an upper bound of the tree inside it); and the exact form of
 * its early exit (:716-730): once the popped cost reaches U, every later tree costs at least U, so
 * OPT = U.  (The reference's own exit tests an int ratio, best / popped < 2, so it can stop up to 2x
 * above the optimum.)
 * stats[0] = states popped, stats[1] = grow relaxations, stats[2] = merge evaluations.
 */
#define ORC_POP_UB 16
int orc_dpbf_pruned(int32_t n, const int64_t* row, const int32_t* col, const double* w, int32_t g,
                    const uint32_t* gm, double U, double grow_frac, int32_t use_lb, const double* dist,
                    double* cost_out, int64_t* stats) {
    const int pop_ub = (use_lb & ORC_POP_UB) != 0;
    use_lb &= ~ORC_POP_UB;
    if (n < 1 || g < 1 || g > 12 || (use_lb && !dist) || (pop_ub && !dist)) return ORC_INVALID;
    const int64_t S = (int64_t)1 << g;
    const uint32_t full = (uint32_t)(S - 1);
    const int64_t NS = (int64_t)n * S;
    double* cost = (double*)malloc(sizeof(double) * (size_t)NS);
    int32_t* pos = (int32_t*)malloc(sizeof(int32_t) * (size_t)NS);
    int64_t* hp = (int64_t*)malloc(sizeof(int64_t) * (size_t)NS);
    double* W = use_lb >= 2 ? (double*)malloc(sizeof(double) * (size_t)(S * g * g)) : NULL;
    if (!cost || !pos || !hp || (use_lb >= 2 && !W)) {
        free(cost); free(pos); free(hp); free(W);
        return ORC_OOM;
    }
    if (use_lb >= 2) {
        double D[20 * 20];
        for (int32_t a = 0; a < g; ++a)
            for (int32_t b = 0; b < g; ++b) {
                double m = INFINITY;
                for (int32_t u = 0; u < n; ++u)
                    if (((gm[u] >> b) & 1u) && dist[(int64_t)a * n + u] < m) m = dist[(int64_t)a * n + u];
                D[a * g + b] = m;
            }
        orc_tour_tables(g, D, W);
    }
    for (int64_t s = 0; s < NS; ++s) { cost[s] = INFINITY; pos[s] = -1; }
    heap_t H = {hp, pos, cost, 0};
    int64_t pops = 0, relax = 0, merges = 0;
    /* one-label bound of (v, p) */
#define ORC_LB(v, p) (use_lb ? orc_lb(dist, W, n, g, (v), (p), use_lb) : 0.0)
    for (int32_t v = 0; v < n; ++v) {
        const uint32_t gv = gm[v] & full;
        if (!gv) continue;
        for (uint32_t p = gv; p; p = (p - 1) & gv) {
            const int64_t s = (int64_t)v * S + p;
            if (ORC_LB(v, p) > U) continue;
            cost[s] = 0.0;
            heap_push_or_decrease(&H, s);
        }
    }
    double best = INFINITY;
    while (H.n > 0) {
        const int64_t s = heap_pop(&H);
        pops++;
        const int32_t v = (int32_t)(s / S);
        const uint32_t p = (uint32_t)(s % S);
        const double cvp = cost[s];
        if (p == full) { best = cvp; break; }
        if (pop_ub) {
            if (cvp >= U) { best = U; break; }   /* nothing left costs less than the feasible U */
            double f = cvp;
            for (int32_t i = 0; i < g; ++i)
                if (!((p >> i) & 1u)) f += dist[(int64_t)i * n + v];
            if (f < U) U = f;
        }
        if (!(grow_frac > 0.0) || cvp <= grow_frac * U) {
            for (int64_t e = row[v]; e < row[v + 1]; ++e) {
                const int32_t u = col[e];
                const int64_t su = (int64_t)u * S + p;
                if (pos[su] == -2) continue;
                relax++;
                const double c = cvp + w[e];
                if (c < cost[su] && c + ORC_LB(u, p) <= U) {
                    cost[su] = c;
                    heap_push_or_decrease(&H, su);
                }
            }
        }
        const uint32_t comp = full ^ p;
        for (uint32_t p2 = comp; p2; p2 = (p2 - 1) & comp) {
            const int64_t s2 = (int64_t)v * S + p2;
            if (!(cost[s2] < INFINITY) || pos[s2] != -2) continue;   /* only settled partners */
            merges++;
            const int64_t sx = (int64_t)v * S + (p | p2);
            if (pos[sx] == -2) continue;
            const double c = cvp + cost[s2];
            if (c < cost[sx] && c + ORC_LB(v, p | p2) <= U) {
                cost[sx] = c;
                heap_push_or_decrease(&H, sx);
            }
        }
    }
#undef ORC_LB
    if (stats) { stats[0] = pops; stats[1] = relax; stats[2] = merges; }
    *cost_out = best;
    free(cost); free(pos); free(hp); free(W);
    return best < INFINITY ? ORC_OK : ORC_INFEASIBLE;
}

/* ---------------------------------------------------------------- Dijkstra */
/* dist[v] = min_s init[s] + d(s,v); init[v] = INFINITY marks a non-source */
int orc_dijkstra(int32_t n, const int64_t* row, const int32_t* col, const double* w, const double* init,
                 double* dist) {
    int64_t* hp = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
    int32_t* pos = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    if (!hp || !pos) { free(hp); free(pos); return ORC_OOM; }
    heap_t H = {hp, pos, dist, 0};
    for (int32_t v = 0; v < n; ++v) {
        dist[v] = init[v];
        pos[v] = -1;
        if (init[v] < INFINITY) heap_push_or_decrease(&H, v);
    }
    while (H.n > 0) {
        const int64_t v = heap_pop(&H);
        for (int64_t e = row[v]; e < row[v + 1]; ++e) {
            const int32_t u = col[e];
            if (pos[u] == -2) continue;
            const double c = dist[v] + w[e];
            if (c < dist[u]) { dist[u] = c; heap_push_or_decrease(&H, u); }
        }
    }
    free(hp); free(pos);
    return ORC_OK;
}

/* ---------------------------------------------------------------- brute force */
/* min over connected vertex subsets touching every group of the MST weight of the induced
 * subgraph; n <= 20.  Returns INFINITY when infeasible. */
double orc_bruteforce(int32_t n, const int64_t* row, const int32_t* col, const double* w, int32_t g,
                      const uint32_t* gm) {
    if (n < 1 || n > 20) return NAN;
    const uint32_t full = (uint32_t)((1u << g) - 1);
    double adj[20][20];
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) adj[i][j] = INFINITY;
    for (int v = 0; v < n; ++v)
        for (int64_t e = row[v]; e < row[v + 1]; ++e)
            if (w[e] < adj[v][col[e]]) adj[v][col[e]] = w[e];
    double best = INFINITY;
    for (uint32_t sub = 1; sub < (1u << n); ++sub) {
        uint32_t cover = 0;
        for (int v = 0; v < n; ++v)
            if (sub >> v & 1u) cover |= gm[v];
        if ((cover & full) != full) continue;
        /* Prim on the induced subgraph */
        int verts[20], k = 0;
        for (int v = 0; v < n; ++v)
            if (sub >> v & 1u) verts[k++] = v;
        double key[20];
        int in[20];
        for (int i = 0; i < k; ++i) { key[i] = INFINITY; in[i] = 0; }
        key[0] = 0.0;
        double total = 0.0;
        int ok = 1;
        for (int it = 0; it < k; ++it) {
            int bi = -1;
            for (int i = 0; i < k; ++i)
                if (!in[i] && (bi < 0 || key[i] < key[bi])) bi = i;
            if (!(key[bi] < INFINITY)) { ok = 0; break; }
            in[bi] = 1;
            total += key[bi];
            for (int i = 0; i < k; ++i)
                if (!in[i] && adj[verts[bi]][verts[i]] < key[i]) key[i] = adj[verts[bi]][verts[i]];
        }
        if (ok && total < best) best = total;
    }
    return best;
}

/* ---------------------------------------------------------------- batch (CPU baseline) */
typedef struct {
    int32_t n; const int64_t* row; const int32_t* col; const double* w;
    int32_t nq; const int32_t* gq; const uint32_t* gms;  /* gms: nq x n group masks */
    double* costs;
    int32_t next;
    pthread_mutex_t mu;
} batch_t;

static void* batch_worker(void* arg) {
    batch_t* B = (batch_t*)arg;
    for (;;) {
        pthread_mutex_lock(&B->mu);
        const int32_t q = B->next++;
        pthread_mutex_unlock(&B->mu);
        if (q >= B->nq) break;
        double c = INFINITY;
        int32_t root = -1, ne = 0;
        orc_dpbf(B->n, B->row, B->col, B->w, B->gq[q], B->gms + (int64_t)q * B->n, &c, &root, NULL, NULL, NULL, 0,
                 &ne, NULL);
        B->costs[q] = c;
    }
    return NULL;
}

int orc_dpbf_batch(int32_t nthreads, int32_t n, const int64_t* row, const int32_t* col, const double* w,
                   int32_t nq, const int32_t* gq, const uint32_t* gms, double* costs) {
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    batch_t B;
    B.n = n; B.row = row; B.col = col; B.w = w; B.nq = nq; B.gq = gq; B.gms = gms; B.costs = costs; B.next = 0;
    pthread_mutex_init(&B.mu, NULL);
    pthread_t th[256];
    for (int i = 0; i < nthreads; ++i) pthread_create(&th[i], NULL, batch_worker, &B);
    for (int i = 0; i < nthreads; ++i) pthread_join(th[i], NULL);
    pthread_mutex_destroy(&B.mu);
    return ORC_OK;
}
