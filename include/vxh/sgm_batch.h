/*
 * vxh batched SGM: many stereo pairs of uneven resolution and disparity range
 * in one persistent-kernel launch, captured in a hipGraph.
 *
 * No single reference entry point corresponds: the reference computes one
 * pair per SGMStereo::compute_disparity call
 * (vx-vision/src/kernels/sgm.rs:108-265) and batches different kernels into
 * one command buffer with Pipeline (vx-vision/src/pipeline.rs:12-80).  Each
 * pair's disparity map is bit-identical to vxh_sgm_compute_disparity on the
 * same inputs and config.
 */
#ifndef VXH_SGM_BATCH_H
#define VXH_SGM_BATCH_H

#include <stddef.h>
#include <stdint.h>

#include "vxh/sgm.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One pair of a batch: caller-owned device buffers (left/right w*h bytes,
 * out w*h floats, all resident on the batch's device), its config, and the
 * aggregation schedule (lanes per scanline and disparities per lane, as for
 * vxh_sgm_aggregate_layout; 0, 0 = the default for its D). */
typedef struct {
    const uint8_t *d_left;
    const uint8_t *d_right;
    float *d_out;
    uint32_t width;
    uint32_t height;
    vxh_sgm_config cfg;
    uint32_t lanes;
    uint32_t per_lane;
} vxh_sgm_pair_desc;

typedef struct vxh_sgm_batch vxh_sgm_batch;

/* Per-pair plan summary. */
typedef struct {
    uint32_t items;      /* work items of this pair in the batch's list */
    uint32_t lanes;      /* schedule used */
    uint32_t per_lane;
} vxh_sgm_batch_stats;

/* Validates every pair, allocates the per-pair scratch (census, summed
 * volume), builds the work list and captures the persistent launch in a
 * hipGraph.  Buffers named in `pairs` are bound at creation. */
vxh_status vxh_sgm_batch_create(int device, const vxh_sgm_pair_desc *pairs,
                                uint32_t n, vxh_sgm_batch **out);
/* Replays the graph on `stream` (NULL = the batch's own stream). */
vxh_status vxh_sgm_batch_launch(vxh_sgm_batch *batch, void *stream);
/* Launch on the batch's own stream, wait, and check the error flag. */
vxh_status vxh_sgm_batch_run(vxh_sgm_batch *batch);
/* After a launch has completed: VXH_ERR_GPU if any item's bounded dependency
 * wait expired (the batch's maps are then incomplete), else VXH_OK. */
vxh_status vxh_sgm_batch_check(vxh_sgm_batch *batch);
vxh_status vxh_sgm_batch_stats_get(const vxh_sgm_batch *batch, uint32_t pair,
                                   vxh_sgm_batch_stats *out);
uint32_t vxh_sgm_batch_items(const vxh_sgm_batch *batch);
void vxh_sgm_batch_destroy(vxh_sgm_batch *batch);
/* Message of the last batch failure on this thread. */
const char *vxh_sgm_batch_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
