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
 * out w*h floats, all resident on the batch's device), its config, the
 * aggregation layouts as for vxh_sgm_set_schedule ((lanes, per_lane) for the
 * non-horizontal directions, (lanes_h, per_lane_h) for the horizontal ones,
 * 0, 0 = default; the batched kernel runs 8 disparities per lane: 8x8,
 * 16x8 or 32x8 for D <= 64, <= 128, <= 256, and rejects other layouts), and
 * volume_format as for vxh_sgm_set_volume_format (0 = automatic). */
typedef struct {
    const uint8_t *d_left;
    const uint8_t *d_right;
    float *d_out;
    uint32_t width;
    uint32_t height;
    vxh_sgm_config cfg;
    uint32_t lanes;
    uint32_t per_lane;
    uint32_t lanes_h;
    uint32_t per_lane_h;
    uint32_t volume_format;
} vxh_sgm_pair_desc;

typedef struct vxh_sgm_batch vxh_sgm_batch;

/* Schedule of the persistent launch (the autotuner's knobs; 0 = default):
 *   blocks_per_cu  resident 256-thread workgroups per CU, 1 or 2 (default 2);
 *   wta_lag        pairs of aggregation items placed between a pair's last
 *                  aggregation item and its sum + winner items (default 2);
 *   wta_pixels     pixels per sum + winner item, a multiple of 4 (default
 *                  1024). */
typedef struct {
    uint32_t blocks_per_cu;
    uint32_t wta_lag;
    uint32_t wta_pixels;
} vxh_sgm_batch_options;

/* Per-pair plan summary. */
typedef struct {
    uint32_t items;      /* work items of this pair in the batch's list */
    uint32_t lanes;      /* schedule used for the non-horizontal directions */
    uint32_t per_lane;
} vxh_sgm_batch_stats;

/* Validates every pair, allocates the per-pair scratch (census, direction
 * volumes; vxh_sgm_workspace_bytes per pair), builds the work list and captures the persistent launch in a
 * hipGraph.  Buffers named in `pairs` are bound at creation. */
vxh_status vxh_sgm_batch_create(int device, const vxh_sgm_pair_desc *pairs,
                                uint32_t n, vxh_sgm_batch **out);
/* vxh_sgm_batch_create with an explicit launch schedule (NULL = defaults). */
vxh_status vxh_sgm_batch_create_ex(int device, const vxh_sgm_pair_desc *pairs,
                                   uint32_t n, const vxh_sgm_batch_options *options,
                                   vxh_sgm_batch **out);
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
