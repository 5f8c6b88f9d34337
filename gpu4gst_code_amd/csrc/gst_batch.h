// Batched engine interface shared by the C API (gst_api.cpp), the learned driver (gst_mpnn.hip)
// and the engine (gst_batch.hip).  Not part of the public ABI.
#pragma once

#include <functional>

#include "gst_internal.h"

// One DPBF solve run by the engine.  The groups pointer must stay valid until job_done sees it.
struct gst_batch_job {
    const gst_graph* G = nullptr;
    const gst_groups* groups = nullptr;
    gst_opts opts{};             // ub, prune_lb and want_tree are honoured
    bool temp_graph = false;     // graph made for this job only: its hub row items are built in the slot
    gst_result* res = nullptr;   // filled when the job finishes
    gst_status status = GST_OK;
    double t_fill = 0.0;         // host ms when the job entered a slot
    double t_done = 0.0;         // host ms when its result was read
    int64_t tag = 0;             // caller's bookkeeping
    int kind = 0;
};

// Runs jobs until every slot is idle and next_job has no job.  next_job(job, wait) is asked for a job
// whenever a slot is free, after every step (also after it once returned false): with wait false it
// returns a job only if one is ready now, so the engine steps on with the queries it holds; with wait
// true (asked only when no slot is busy) it blocks until a job is ready, and false then means no job
// will come.  job_done receives each finished job (status and result set) and may make more jobs
// available to next_job.  n_cap / g_cap / m_cap bound every job's graph size and group count.  Returns the first non-OK job status (the other jobs still run) or an engine error.
// tune_uid names the graph whose autotuned engine knobs apply (0: none).  max_jobs (0: unknown)
// bounds the jobs in flight at once: without an explicit slot count the engine sizes no more slots.
// *engine_error (may be null) tells the two kinds of non-OK return apart: false, a job's own status
// (every job finished and went through job_done); true, an engine error (a HIP call failed, a fill
// failed, the step bound was hit): the run stopped early, nothing of it still runs on the device,
// and the jobs that were in slots (or being filled) never reach job_done, so the caller must mark
// every job it handed out and did not see finish with the returned status.
gst_status gst_batch_run(gst_ctx* ctx, const std::function<bool(gst_batch_job&, bool wait)>& next_job,
                         const std::function<void(gst_batch_job&)>& job_done, uint32_t n_cap, int g_cap,
                         int64_t m_cap, const gst_batch_opts* bopts, gst_batch_stats* stats, uint64_t tune_uid = 0,
                         int64_t max_jobs = 0, bool* engine_error = nullptr);
void gst_batch_free(gst_ctx* ctx);
