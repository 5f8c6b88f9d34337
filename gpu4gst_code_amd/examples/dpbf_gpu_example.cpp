// Caller written against the reference-shaped C++ API (include/gst.hpp): builds a
// graph_v_of_v_idealID and a group graph the way the reference's harness does
// (rucgraph/test_gpu.h:115-196: input graph, group vertices V..V+g-1 whose rows list members),
// calls gst::DPBF_GPU and gst::DPBF_GPU_T with the reference's scalar types, checks the trees
// with the reference's own check names and MST post-processes the optimal tree's vertex set.  Usage: dpbf_gpu_example <V> <E> <g> <seed>
#include <cstdio>
#include <cstdlib>

#include "gst.hpp"

int main(int argc, char** argv) {
    const int V = argc > 1 ? std::atoi(argv[1]) : 1000;
    const long long E = argc > 2 ? std::atoll(argv[2]) : 5000;
    const int g = argc > 3 ? std::atoi(argv[3]) : 3;
    const unsigned long long seed = argc > 4 ? std::strtoull(argv[4], nullptr, 10) : 1;
    try {
        gst_host_csr h;
        if (gst_gen_random_connected(V, E, 1, 100, seed, &h) != GST_OK) return 2;
        gst::graph_v_of_v_idealID input_graph((size_t)V);
        const int32_t* w = (const int32_t*)h.w;
        for (int v = 0; v < V; ++v)
            for (int64_t e = h.row_ptr[v]; e < h.row_ptr[v + 1]; ++e) input_graph[(size_t)v].push_back({h.col[e], (double)w[e]});
        gst_host_csr_free(&h);

        std::vector<int32_t> gp((size_t)g + 1), gvtx((size_t)g * 20);
        if (gst_gen_groups(V, g, 5, 20, seed + 1, gp.data(), gvtx.data()) != GST_OK) return 2;
        gst::graph_v_of_v_idealID group_graph((size_t)(V + g));
        std::vector<int> order;  // group vertex ids, inserted in order
        std::vector<std::vector<int>> lists;
        std::unordered_set<int> cumpulsory_group_vertices;
        for (int i = 0; i < g; ++i) {
            const int gv = V + i;
            cumpulsory_group_vertices.insert(gv);
            for (int32_t j = gp[(size_t)i]; j < gp[(size_t)i + 1]; ++j) {
                group_graph[(size_t)gv].push_back({gvtx[(size_t)j], 1.0});
                group_graph[(size_t)gvtx[(size_t)j]].push_back({gv, 1.0});
            }
        }
        for (int gv : cumpulsory_group_vertices) {
            std::vector<int> mem;
            for (auto& p : group_graph[(size_t)gv]) mem.push_back(p.first);
            lists.push_back(mem);
        }

        gst::context ctx(0);
        gst::device_graph dg(ctx, input_graph);
        // the reference's call (rucgraph/test_gpu.h:203) with its own types: int* pointer1,
        // int* real_cost, double* rt
        int pointer1 = 0, cost = 0;
        double rt = 0;
        int* real_cost = &cost;
        gst::graph_hash_of_mixed_weighted solu =
            gst::DPBF_GPU(dg, cumpulsory_group_vertices, group_graph, &pointer1, real_cost, &rt);
        const int sum = gst::graph_hash_of_mixed_weighted_sum_of_ec(solu);   // rucgraph/test_gpu.h:207
        if (!gst::this_is_a_feasible_solution_gpu(solu, group_graph, cumpulsory_group_vertices)) return 4;
        const double tc = gst::check_tree(dg, lists, solu);
        // DPBF_GPU_T (rucgraph/DPQ.cuh:58): bounded by a feasible cost found on the way, same optimum
        int pointer1_t = 0, cost_t = 0;
        double rt_t = 0;
        gst::graph_hash_of_mixed_weighted solu_t =
            gst::DPBF_GPU_T(dg, cumpulsory_group_vertices, group_graph, &pointer1_t, &cost_t, &rt_t);
        if (!gst::this_is_a_feasible_solution_gpu(solu_t, group_graph, cumpulsory_group_vertices)) return 4;
        // MST post-processing of the optimal tree's vertices (graph_hash_of_mixed_weighted_MST_postprocessing.h:73):
        // an optimal tree is a minimum spanning tree of its own vertex set, so the cost stays
        std::unordered_set<int> hash_of_v;
        for (auto& kv : solu.hash_of_vectors) hash_of_v.insert(kv.first);
        gst::graph_hash_of_mixed_weighted mst =
            gst::graph_hash_of_mixed_weighted_MST_postprocessing_hash_of_vertices(dg, hash_of_v);
        const int mst_sum = gst::graph_hash_of_mixed_weighted_sum_of_ec(mst);
        std::printf("cost %d sum_of_ec %d tree_cost %.1f vertices %zu edges %zu rt %.6f items %d; DPBF_GPU_T cost %d "
                    "sum_of_ec %d items %d; MST post-processing %d\n",
                    *real_cost, sum, tc, solu.num_vertices(), solu.num_edges(), rt, pointer1, cost_t,
                    gst::graph_hash_of_mixed_weighted_sum_of_ec(solu_t), pointer1_t, mst_sum);
        return (tc == *real_cost && sum == *real_cost && cost_t == *real_cost &&
                gst::graph_hash_of_mixed_weighted_sum_of_ec(solu_t) == cost_t && mst_sum == *real_cost &&
                mst.num_vertices() == solu.num_vertices()) ? 0 : 1;
    } catch (const gst::error& e) {
        std::fprintf(stderr, "gst error %d: %s\n", (int)e.status, e.what());
        return 3;
    }
}
