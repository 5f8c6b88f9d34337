// Internal: learned-bound solve (gst_mpnn.hip), called by the C ABI in gst_api.cpp.
#pragma once
#include "gst_internal.h"

gst_status gst_learned_solve(gst_ctx* ctx, const gst_graph* g, const gst_groups* groups, gst_model* m, double frac,
                             const gst_opts* opts, gst_result* res, gst_learned_stats* stats);
// Internal: batched learned-bound solve (gst_mpnn.hip); status[] holds each query's validation
// result on entry and its outcome on return.
gst_status gst_learned_solve_batch(gst_ctx* ctx, const gst_graph* g, int32_t nq, const gst_groups* groups,
                                   gst_model* m, double frac, const gst_opts* opts, const gst_batch_opts* bopts,
                                   gst_result* results, gst_status* status, gst_learned_stats* lstats,
                                   gst_batch_stats* bstats);
