// The per-pair pipeline (census of both images, matching costs, path
// aggregation, sum + winner-take-all, optional LR check) as one host call,
// shared by the plain entry points (sgm_runtime.hip), the batched form
// (sgm_batch.hip) and pipeline encodes (pipeline.hip); the context's device
// pool and the hooks encoders use on an open pipeline.
#pragma once

#include <hip/hip_runtime.h>

#include "sgm_internal.h"
#include "vxh/pipeline.h"
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

// Lets the calling thread make allocating and synchronous HIP calls while a
// stream capture is open (a pipeline between begin and commit): switches the
// thread's capture-interaction mode to relaxed for the guard's lifetime.  The
// calls it guards touch no capturing stream.
struct CaptureRelaxed {
    hipStreamCaptureMode prev = hipStreamCaptureModeRelaxed;
    CaptureRelaxed() { (void)hipThreadExchangeStreamCaptureMode(&prev); }
    ~CaptureRelaxed() { (void)hipThreadExchangeStreamCaptureMode(&prev); }
    CaptureRelaxed(const CaptureRelaxed &) = delete;
    CaptureRelaxed &operator=(const CaptureRelaxed &) = delete;
};

// The context's device pool: scratch blocks (smallest free block of at least
// `bytes`, else a new allocation of exactly `bytes` rounded up to 4 KiB) and
// non-blocking streams, handed out to pipelines and returned when they end.
vxh_status ctx_take_block(vxh_context *ctx, size_t bytes, void **ptr, size_t *got);
void ctx_give_block(vxh_context *ctx, void *ptr);
vxh_status ctx_take_stream(vxh_context *ctx, hipStream_t *st);
void ctx_give_stream(vxh_context *ctx, hipStream_t st);

// What encoders need from an open pipeline: its capturing stream (with the
// pipeline's device made current; INVALID_CONFIG once committed), scratch that
// lives as long as the pipeline, and a device error word to fold into the
// pipeline's (read by vxh_pipeline_wait).
vxh_status pipeline_stream(vxh_pipeline *p, hipStream_t *st);
vxh_status pipeline_scratch(vxh_pipeline *p, size_t bytes, void **ptr);
vxh_status pipeline_watch_word(vxh_pipeline *p, const uint32_t *d_word);
vxh_context *pipeline_context(vxh_pipeline *p);

// Textures in flight (the reference's GpuGuard, vx-core/src/buffer.rs:59-66,
// 102-122): every encode records the textures it reads or writes
// (pipeline_use).  While a pipeline that used a texture has launched work that
// has not finished (its done event, recorded after each commit and relaunch,
// is pending), a host write into the texture is refused (INVALID_CONFIG) and
// a host read waits for that work.  pipeline_forget drops a destroyed
// pipeline's records.
void pipeline_use(vxh_pipeline *p, const vxh_texture *t);
void pipeline_forget(vxh_pipeline *p);
// The pipeline's done event (nullptr before its first launch).
hipEvent_t pipeline_done_event(vxh_pipeline *p);

// Count of pipelines between begin and commit in the process (texture copies
// avoid the legacy null stream while it is non-zero).
void capture_opened();
void capture_closed();

}  // namespace vxh
