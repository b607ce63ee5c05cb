// The per-pair pipeline (census of both images, matching costs, path
// aggregation, sum + winner-take-all, optional LR check) as one host call,
// shared by the plain entry points (sgm_runtime.hip) and the batched form
// (sgm_batch.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "sgm_internal.h"
#include "vxh/sgm.h"

#include <string>

namespace vxh {

// Sets the calling thread's vxh_last_error() message; returns st.
vxh_status report_error(vxh_status st, const std::string &msg);

// Enqueues the pipeline of one pair on `st`.  The config must be valid for
// (w, h) and `ws` must hold workspace_layout(w, h, D, num_paths, mode,
// lr_check).total bytes.  With `marks` (three events) records them after the
// census and cost passes, after the aggregation and at the end.  Returns
// VXH_OK or a launch failure (message in vxh_last_error()).
vxh_status run_pipeline(const uint8_t *d_left, const uint8_t *d_right, uint32_t w, uint32_t h,
                        const vxh_sgm_config *c, float *d_out, char *ws, int mode,
                        Schedule sc, hipStream_t st, hipEvent_t const *marks = nullptr);

}  // namespace vxh
