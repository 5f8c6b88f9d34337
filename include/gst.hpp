// gst.hpp -- header-only C++ host API over the C-ABI in gst.h, shaped like the reference's
// host API so that a caller of the reference can switch with the same types and argument order:
//
//   reference                                                       here
//   graph_v_of_v_idealID (rucgraph/graph_v_of_v_idealID/           gst::graph_v_of_v_idealID (same type)
//     graph_v_of_v_idealID.h:50)
//   toCSR(graph) (rucgraph/graph_v_of_v_idealID/csr_graph.hpp:29)  gst::device_graph(ctx, graph)
//   DPBF_GPU(host_tree, host_tree_one_d, csr, cumpulsory_group_     gst::DPBF_GPU(dg, cumpulsory_group_vertices,
//            vertices, group_graph, input_graph, int* pointer1,                   group_graph, int* pointer1,
//            int* real_cost, community, c_size, s, double* rt)                    int* real_cost, double* rt)
//     (rucgraph/DPQ.cuh:57)
//   DPBF_GPU_T(..., int* pointer1, int* real_cost, ..., double* rt) gst::DPBF_GPU_T(dg, cumpulsory_group_vertices,
//     (rucgraph/DPQ.cuh:58)                                                        group_graph, int* pointer1,
//                                                                                  int* real_cost, double* rt)
//   graph_hash_of_mixed_weighted (returned tree,                    gst::graph_hash_of_mixed_weighted (= gst::tree:
//     rucgraph/graph_hash_of_mixed_weighted/                           hash_of_vectors[v].adj_vertices, hash_of_hashs,
//     graph_hash_of_mixed_weighted.h:10-126)                           add_vertex / add_edge / edge_weight / degree)
//   graph_hash_of_mixed_weighted_sum_of_ec                          gst::graph_hash_of_mixed_weighted_sum_of_ec
//     (rucgraph/graph_hash_of_mixed_weighted_sum_of_nw_ec.h:58)
//   this_is_a_feasible_solution_gpu (rucgraph/test_gpu.h:45-77)     gst::this_is_a_feasible_solution_gpu
//   graph_hash_of_mixed_weighted_MST_postprocessing_hash_of_vertices gst::graph_hash_of_mixed_weighted_MST_
//     (rucgraph/graph_hash_of_mixed_weighted_MST_postprocessing.h:73)  postprocessing_hash_of_vertices(dg, set)
//
// The arguments the reference needs only for its own engine (the V x 2^g host table, the CSR, the
// input graph again, the disjoint-set list, the community arrays) have no counterpart: the table
// lives in HBM and the graph was uploaded once by device_graph.
//
// Group convention as in the reference: every id in cumpulsory_group_vertices is a group vertex
// whose group_graph row lists its members (src/DPQ.cu:27-31).  Group bit i is the i-th id in the
// set's iteration order, as in set_max_ID (src/DPQ.cu:22-34); the optimal cost does not depend on
// the order.  Weights: a graph whose weights are all integral is solved with exact int32
// weights, otherwise with float32 weights (the reference truncates every weight to int,
// csr_graph.hpp:20,46).  Errors are reported by gst::error exceptions in this header only; no
// exception crosses the C ABI.
#pragma once

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <utility>
#include <vector>

#include "gst.h"

namespace gst {

typedef std::vector<std::vector<std::pair<int, double>>> graph_v_of_v_idealID;

class error : public std::runtime_error {
public:
    error(gst_status s, const std::string& msg) : std::runtime_error(msg), status(s) {}
    gst_status status;
};

// One vertex of a returned tree, as graph_hash_of_mixed_weighted_vectors
// (rucgraph/graph_hash_of_mixed_weighted/graph_hash_of_mixed_weighted.h:10-20).
struct graph_hash_of_mixed_weighted_vectors {
    double vertex_weight = 0.0;
    std::vector<std::pair<int, double>> adj_vertices;   // (neighbour, edge weight)
};

// The returned tree: the part of graph_hash_of_mixed_weighted that callers of DPBF_GPU use.  Every
// vertex sits in hash_of_vectors; hash_of_hashs stays empty (the reference keeps only high-degree
// vertices there, and its readers fall back to hash_of_vectors).
struct tree {
    std::unordered_map<int, graph_hash_of_mixed_weighted_vectors> hash_of_vectors;
    std::unordered_map<int, std::unordered_map<int, double>> hash_of_hashs;
    void add_vertex(int v, double w = 0.0) { hash_of_vectors[v].vertex_weight = w; }
    void add_edge(int u, int v, double w) {
        hash_of_vectors[u].adj_vertices.push_back({v, w});
        hash_of_vectors[v].adj_vertices.push_back({u, w});
    }
    double edge_weight(int u, int v) const {
        auto it = hash_of_vectors.find(u);
        if (it == hash_of_vectors.end()) return INFINITY;
        for (auto& p : it->second.adj_vertices)
            if (p.first == v) return p.second;
        return INFINITY;
    }
    bool contain_edge(int u, int v) const { return edge_weight(u, v) != INFINITY; }
    int degree(int v) const {
        auto it = hash_of_vectors.find(v);
        return it == hash_of_vectors.end() ? 0 : (int)it->second.adj_vertices.size();
    }
    std::vector<int> adj_v(int v) const {
        std::vector<int> out;
        auto it = hash_of_vectors.find(v);
        if (it != hash_of_vectors.end())
            for (auto& p : it->second.adj_vertices) out.push_back(p.first);
        return out;
    }
    size_t num_vertices() const { return hash_of_vectors.size(); }
    size_t num_edges() const {
        size_t s = 0;
        for (auto& kv : hash_of_vectors) s += kv.second.adj_vertices.size();
        return s / 2;
    }
    double cost() const {
        double c = 0;
        for (auto& kv : hash_of_vectors)
            for (auto& p : kv.second.adj_vertices)
                if (kv.first < p.first) c += p.second;
        return c;
    }
};
typedef tree graph_hash_of_mixed_weighted;

// Sum of the edge weights in int, each weight truncated as the reference adds them
// (graph_hash_of_mixed_weighted_sum_of_ec, rucgraph/graph_hash_of_mixed_weighted_sum_of_nw_ec.h:58-87).
inline int graph_hash_of_mixed_weighted_sum_of_ec(const tree& t) {
    int sum = 0;
    for (auto& kv : t.hash_of_vectors)
        for (auto& p : kv.second.adj_vertices)
            if (kv.first < p.first) sum += (int)p.second;
    return sum;
}

// this_is_a_feasible_solution_gpu (rucgraph/test_gpu.h:45-77): the tree is one connected component
// and some vertex of it is a member of every group (a group vertex gv's group_graph row lists its
// members).
inline bool this_is_a_feasible_solution_gpu(const tree& solu, const graph_v_of_v_idealID& group_graph,
                                            const std::unordered_set<int>& group_vertices) {
    if (solu.hash_of_vectors.empty()) return false;
    std::unordered_set<int> seen;
    std::vector<int> stack{solu.hash_of_vectors.begin()->first};
    seen.insert(stack.back());
    while (!stack.empty()) {
        const int v = stack.back();
        stack.pop_back();
        for (int u : solu.adj_v(v))
            if (seen.insert(u).second) stack.push_back(u);
    }
    if (seen.size() != solu.hash_of_vectors.size()) return false;
    for (int gv : group_vertices) {
        bool covered = false;
        if (gv >= 0 && (size_t)gv < group_graph.size())
            for (auto& p : group_graph[(size_t)gv])
                if (solu.hash_of_vectors.count(p.first)) { covered = true; break; }
        if (!covered) return false;
    }
    return true;
}

class context {
public:
    explicit context(int device = 0) {
        gst_status s = gst_ctx_create(device, &ctx_);
        if (s != GST_OK) {
            std::string m = ctx_ ? gst_last_error(ctx_) : gst_status_str(s);
            gst_ctx_destroy(ctx_);
            ctx_ = nullptr;
            throw error(s, m);
        }
    }
    ~context() { gst_ctx_destroy(ctx_); }
    context(const context&) = delete;
    context& operator=(const context&) = delete;
    gst_ctx* get() const { return ctx_; }
    void check(gst_status s) const {
        if (s != GST_OK) throw error(s, std::string(gst_status_str(s)) + ": " + gst_last_error(ctx_));
    }

private:
    gst_ctx* ctx_ = nullptr;
};

// Device-resident CSR of a graph_v_of_v_idealID (the toCSR step, done once per graph).
class device_graph {
public:
    device_graph(context& ctx, const graph_v_of_v_idealID& g) : ctx_(ctx) {
        const int32_t n = (int32_t)g.size();
        row_.assign((size_t)n + 1, 0);
        bool integral = true;
        for (int32_t v = 0; v < n; ++v) {
            row_[(size_t)v + 1] = row_[(size_t)v] + (int64_t)g[(size_t)v].size();
            for (auto& p : g[(size_t)v])
                if (p.second != std::floor(p.second) || p.second > 2147483647.0) integral = false;
        }
        col_.reserve((size_t)row_.back());
        for (int32_t v = 0; v < n; ++v)
            for (auto& p : g[(size_t)v]) col_.push_back(p.first);
        w_type_ = integral ? GST_W_INT32 : GST_W_FLOAT32;
        if (integral) {
            wi_.reserve(col_.size());
            for (auto& r : g) for (auto& p : r) wi_.push_back((int32_t)p.second);
        } else {
            wf_.reserve(col_.size());
            for (auto& r : g) for (auto& p : r) wf_.push_back((float)p.second);
        }
        gst_csr c = csr();
        ctx_.check(gst_graph_upload(ctx_.get(), &c, &g_));
    }
    ~device_graph() { gst_graph_free(g_); }
    device_graph(const device_graph&) = delete;
    device_graph& operator=(const device_graph&) = delete;
    const gst_graph* get() const { return g_; }
    gst_csr csr() const {
        gst_csr c;
        c.n = (int32_t)row_.size() - 1;
        c.m = row_.back();
        c.row_ptr = row_.data();
        c.col = col_.data();
        c.w = w_type_ == GST_W_INT32 ? (const void*)wi_.data() : (const void*)wf_.data();
        c.w_type = w_type_;
        return c;
    }
    context& ctx() const { return ctx_; }

private:
    context& ctx_;
    gst_graph* g_ = nullptr;
    std::vector<int64_t> row_;
    std::vector<int32_t> col_;
    std::vector<int32_t> wi_;
    std::vector<float> wf_;
    int32_t w_type_ = GST_W_INT32;
};

// Solve with explicit member lists (group i = bit i).
inline tree solve(device_graph& dg, const std::vector<std::vector<int>>& groups, double* cost,
                  const gst_opts* opts = nullptr, gst_stats* stats = nullptr) {
    std::vector<int32_t> ptr(groups.size() + 1, 0), vtx;
    for (size_t i = 0; i < groups.size(); ++i) {
        ptr[i + 1] = ptr[i] + (int32_t)groups[i].size();
        for (int v : groups[i]) vtx.push_back(v);
    }
    gst_groups gg{(int32_t)groups.size(), ptr.data(), vtx.data()};
    gst_result r;
    dg.ctx().check(gst_solve(dg.ctx().get(), dg.get(), &gg, opts, &r));
    tree t;
    if (r.n_edges == 0 && r.root >= 0) t.add_vertex(r.root);
    for (int32_t i = 0; i < r.n_edges; ++i) t.add_edge(r.eu[i], r.ev[i], r.ew[i]);
    if (cost) *cost = r.cost;
    if (stats) *stats = r.stats;
    gst_result_free(&r);
    return t;
}

// group vertices -> member lists, in the set's iteration order (group bit i = i-th id)
inline std::vector<std::vector<int>> member_lists(const std::unordered_set<int>& cumpulsory_group_vertices,
                                                  const graph_v_of_v_idealID& group_graph) {
    std::vector<std::vector<int>> groups;
    for (int gv : cumpulsory_group_vertices) {
        std::vector<int> mem;
        if (gv >= 0 && (size_t)gv < group_graph.size())
            for (auto& p : group_graph[(size_t)gv]) mem.push_back(p.first);
        groups.push_back(std::move(mem));
    }
    return groups;
}

// DPBF_GPU with this build's own output types: real_cost <- optimal cost as a double (exact for
// int weights, the fp32 DP value for float weights), rt <- seconds of device work,
// pointer1 += work items processed (reference: dequeued states).
inline tree DPBF_GPU(device_graph& dg, const std::unordered_set<int>& cumpulsory_group_vertices,
                     const graph_v_of_v_idealID& group_graph, double* real_cost, double* rt = nullptr,
                     long long* pointer1 = nullptr) {
    gst_stats st;
    tree t = solve(dg, member_lists(cumpulsory_group_vertices, group_graph), real_cost, nullptr, &st);
    if (rt) *rt = (st.t_grow_ms + st.t_merge_ms + st.t_trace_ms) / 1e3;
    if (pointer1) *pointer1 += (long long)st.grow_items;
    return t;
}

namespace detail {
inline tree dpbf_int(device_graph& dg, const std::unordered_set<int>& cumpulsory_group_vertices,
                     const graph_v_of_v_idealID& group_graph, int* pointer1, int* real_cost, double* rt,
                     const gst_opts* opts) {
    gst_stats st;
    double c = 0.0;
    tree t = solve(dg, member_lists(cumpulsory_group_vertices, group_graph), &c, opts, &st);
    if (real_cost) *real_cost = (int)c;   // truncated, as the reference's int costs of truncated weights
    if (rt) *rt = (st.t_grow_ms + st.t_merge_ms + st.t_trace_ms) / 1e3;
    if (pointer1) *pointer1 += (int)st.grow_items;
    return t;
}
}  // namespace detail

// DPBF_GPU with the reference's own scalar types and argument order (rucgraph/DPQ.cuh:57):
// pointer1 += work items processed, real_cost <- optimal cost as int, rt <- seconds of device work.
// For float weights the int cost is the optimum truncated toward zero; the double overload above
// returns it whole.
inline graph_hash_of_mixed_weighted DPBF_GPU(device_graph& dg, const std::unordered_set<int>& cumpulsory_group_vertices,
                                             const graph_v_of_v_idealID& group_graph, int* pointer1, int* real_cost,
                                             double* rt) {
    return detail::dpbf_int(dg, cumpulsory_group_vertices, group_graph, pointer1, real_cost, rt, nullptr);
}

// DPBF_GPU_T (rucgraph/DPQ.cuh:58, src/DPQ.cu:769-1040): the reference first runs a few unbounded
// rounds to find a feasible cost, then prunes with it and the one-label bound (and stops after 20
// rounds, so its answer may be above the optimum).  Here: the automatic spider bound (the cheapest
// union of shortest paths from one vertex to every group, a feasible cost once the group distances
// are final) with the one-label and tour lower bounds; the answer is the exact optimum.
inline graph_hash_of_mixed_weighted DPBF_GPU_T(device_graph& dg, const std::unordered_set<int>& cumpulsory_group_vertices,
                                               const graph_v_of_v_idealID& group_graph, int* pointer1, int* real_cost,
                                               double* rt) {
    gst_opts o{};
    o.prune_lb = 1;
    o.want_tree = 1;
    o.auto_bound = 1;
    return detail::dpbf_int(dg, cumpulsory_group_vertices, group_graph, pointer1, real_cost, rt, &o);
}

// graph_hash_of_mixed_weighted_MST_postprocessing_hash_of_vertices
// (rucgraph/graph_hash_of_mixed_weighted_MST_postprocessing.h:73-103): the minimum spanning tree of the
// subgraph induced by hash_of_v, every vertex of the set included (a disconnected set gives a
// spanning forest; the reference's Prim keeps only its first vertex's component).  On the device, by
// gst_mst_postprocess: at most 512 distinct vertices.
inline graph_hash_of_mixed_weighted graph_hash_of_mixed_weighted_MST_postprocessing_hash_of_vertices(
    device_graph& dg, const std::unordered_set<int>& hash_of_v) {
    std::vector<int32_t> vs(hash_of_v.begin(), hash_of_v.end());
    tree t;
    if (vs.empty()) return t;
    std::vector<int32_t> eu(vs.size()), ev(vs.size());
    std::vector<double> ew(vs.size());
    int32_t ne = 0;
    double cost = 0.0;
    dg.ctx().check(gst_mst_postprocess(dg.ctx().get(), dg.get(), vs.data(), (int32_t)vs.size(), eu.data(), ev.data(),
                                       ew.data(), &ne, &cost));
    for (int v : vs) t.add_vertex(v);
    for (int32_t i = 0; i < ne; ++i) t.add_edge(eu[(size_t)i], ev[(size_t)i], ew[(size_t)i]);
    return t;
}

// this_is_a_feasible_solution_gpu (rucgraph/test_gpu.h:45-77), strengthened: the edge set must be
// a tree of the graph touching every group; returns its cost or throws.
inline double check_tree(device_graph& dg, const std::vector<std::vector<int>>& groups, const tree& t) {
    std::vector<int32_t> eu, ev;
    int32_t root = -1;
    for (auto& kv : t.hash_of_vectors) {
        root = kv.first;
        for (auto& p : kv.second.adj_vertices)
            if (kv.first < p.first) { eu.push_back(kv.first); ev.push_back(p.first); }
    }
    std::vector<int32_t> ptr(groups.size() + 1, 0), vtx;
    for (size_t i = 0; i < groups.size(); ++i) {
        ptr[i + 1] = ptr[i] + (int32_t)groups[i].size();
        for (int v : groups[i]) vtx.push_back(v);
    }
    gst_groups gg{(int32_t)groups.size(), ptr.data(), vtx.data()};
    gst_csr c = dg.csr();
    double cost = 0;
    gst_status s = gst_validate_tree(&c, &gg, (int32_t)eu.size(), eu.data(), ev.data(), root, &cost);
    if (s != GST_OK) throw error(s, "not a feasible tree");
    return cost;
}

}  // namespace gst
