/*
 * vxh batched SGM: many stereo pairs of uneven resolution and disparity range
 * replayed from one hipGraph, by default as multi-pair launches around one
 * persistent aggregation launch per lane-layout class (see
 * vxh_sgm_batch_options for the forms).
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
 * 0, 0 = default), volume_format as for vxh_sgm_set_volume_format
 * (0 = automatic) and split as for vxh_sgm_set_split (0 = off). */
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
    uint32_t split;
} vxh_sgm_pair_desc;

typedef struct vxh_sgm_batch vxh_sgm_batch;

/* Launch options (the autotuner's knobs; 0 = default):
 *   form           0 = persistent (default): pairs in chunks of at most
 *                  budget_bytes of workspace, each chunk one census launch,
 *                  one cost launch, one persistent aggregation launch per
 *                  lane-layout class and volume format present (waves take
 *                  scanline groups of every pair by ticket), one sum +
 *                  winner launch; split scanlines are not used.
 *                  1 = streams: the per-pair pipelines dealt to `streams`
 *                  streams, largest pairs first, pairs on one stream in turn
 *                  sharing one workspace.
 *                  2 = sweeps: chunks as the persistent form; after the
 *                  census and cost launches, one forward sweep per
 *                  lane-layout class aggregates the four directions
 *                  (+1,0) (0,+1) (+1,+1) (-1,+1) together in raster order
 *                  and stores their sum once (u16), and one backward sweep
 *                  runs the other four over the image turned by 180 degrees,
 *                  adds that sum and takes each pixel's winner, so no
 *                  per-direction volume and no summed volume is stored.
 *                  Every pair must have num_paths 8; lane layouts, volume
 *                  format and split settings of the pairs do not apply.
 *   streams        streams form: 1..16 (default 4).
 *   blocks_per_cu  persistent form: blocks of 256 threads of the aggregation
 *                  per CU (default: as many as fit).
 *   budget_bytes   persistent and sweep forms: workspace bytes of one chunk
 *                  (default: 60 % of the device memory free at creation, at
 *                  least 1 GiB; a pair larger than the budget forms its own
 *                  chunk). */
typedef struct {
    uint32_t streams;
    uint32_t form;
    uint32_t blocks_per_cu;
    uint32_t reserved;
    uint64_t budget_bytes;
} vxh_sgm_batch_options;

/* Per-pair plan summary. */
typedef struct {
    uint32_t stream;          /* streams form: stream the pair runs on */
    uint32_t lanes;           /* layout of the non-horizontal directions */
    uint32_t per_lane;
    uint32_t lanes_h;         /* layout of the horizontal directions */
    uint32_t per_lane_h;
    uint32_t volume_format;   /* 1 = u8, 2 = u16 */
    uint64_t workspace_bytes; /* scratch the pair needs */
    uint32_t chunk;           /* persistent and sweep forms: chunk the pair runs in */
    uint32_t form;            /* 0 persistent, 1 streams, 2 sweeps */
} vxh_sgm_batch_stats;

/* Validates every pair, allocates the workspaces (persistent form: one shared
 * by the chunks, the largest chunk's; streams form: one per stream, the
 * largest of the pairs dealt to it) and captures every launch in a hipGraph.
 * Buffers named in `pairs` are bound at creation. */
vxh_status vxh_sgm_batch_create(int device, const vxh_sgm_pair_desc *pairs,
                                uint32_t n, vxh_sgm_batch **out);
/* vxh_sgm_batch_create with explicit options (NULL = defaults). */
vxh_status vxh_sgm_batch_create_ex(int device, const vxh_sgm_pair_desc *pairs,
                                   uint32_t n, const vxh_sgm_batch_options *options,
                                   vxh_sgm_batch **out);
/* Replays the graph on `stream` (NULL = the batch's own stream). */
vxh_status vxh_sgm_batch_launch(vxh_sgm_batch *batch, void *stream);
/* Launch on the batch's own stream, wait, and check the error flag. */
vxh_status vxh_sgm_batch_run(vxh_sgm_batch *batch);
/* After a launch has completed: VXH_ERR_GPU if any bounded wait expired (a
 * split scanline's head for its tail, or a sweep wave for the wave above it;
 * that pair's map is then incomplete), else VXH_OK. */
vxh_status vxh_sgm_batch_check(vxh_sgm_batch *batch);
vxh_status vxh_sgm_batch_stats_get(const vxh_sgm_batch *batch, uint32_t pair,
                                   vxh_sgm_batch_stats *out);
/* Kernel launches in the batch's graph. */
uint32_t vxh_sgm_batch_items(const vxh_sgm_batch *batch);
/* Chunks of the persistent and sweep forms (1 for the streams form). */
uint32_t vxh_sgm_batch_chunks(const vxh_sgm_batch *batch);
void vxh_sgm_batch_destroy(vxh_sgm_batch *batch);
/* Message of the last batch failure on this thread. */
const char *vxh_sgm_batch_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
