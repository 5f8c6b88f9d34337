// Internal: learned-bound solve (gst_mpnn.hip), called by the C ABI in gst_api.cpp.
#pragma once
#include "gst_internal.h"

gst_status gst_learned_solve(gst_ctx* ctx, const gst_graph* g, const gst_groups* groups, gst_model* m, double frac,
                             const gst_opts* opts, gst_result* res, gst_learned_stats* stats);
