/*
 * vxh SGM serving loop: admits stereo requests as they arrive, sizes
 * admission by each request's cost-volume bytes against an HBM budget, keeps
 * device buffers resident in a slot pool, overlaps admitted requests on
 * several HIP streams, and reports latency percentiles beside throughput.
 *
 * No reference counterpart: the reference computes one pair per synchronous
 * SGMStereo::compute_disparity call (vx-vision/src/kernels/sgm.rs:108-265);
 * its TexturePool (vx-vision/src/pool.rs:17-147) is the pattern the slot pool
 * follows (reuse keyed by size, hit counters).  Every result is bit-identical
 * to vxh_sgm_compute_disparity on the same inputs and config.
 */
#ifndef VXH_SGM_SERVER_H
#define VXH_SGM_SERVER_H

#include <stddef.h>
#include <stdint.h>

#include "vxh/sgm.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    uint64_t hbm_budget_bytes;  /* sum of admitted requests' volume bytes   */
    uint32_t max_inflight;      /* admitted requests at once (>= 1)         */
    uint32_t num_streams;       /* HIP streams the admitted work spreads on */
    uint64_t work_budget_pxd;   /* sum of admitted requests' W*H*D while any
                                   runs (a larger request runs alone); 0: no
                                   cap.  Bounds how many requests share the
                                   GPU: a large request shares it with little,
                                   which keeps the tail latency down. */
    uint64_t busy_work_budget_pxd;  /* the cap instead while at least
                                       backlog_requests wait (0: unchanged):
                                       more overlap, more throughput, once a
                                       backlog shows the GPU is the limit */
    uint32_t backlog_requests;      /* 0: no busy state.  Otherwise requests
                                       admitted without a backlog queue on the
                                       first stream (pipelined, never sharing
                                       the GPU) and only those admitted with
                                       one spread over every stream. */
    uint32_t reserved;
} vxh_server_config;

typedef struct {
    uint64_t submitted;
    uint64_t completed;
    uint64_t rejected;               /* larger than the whole budget        */
    uint64_t peak_admitted_bytes;    /* max of the admitted volume bytes    */
    uint64_t peak_inflight;
    uint64_t slot_hits;              /* requests served from a pooled slot  */
    uint64_t slot_misses;            /* requests that allocated a new slot  */
    double p50_ms, p99_ms, max_ms;   /* latency: submit -> result on host   */
    double mean_queue_ms;            /* submit -> admission                 */
    double pxd_total;                /* sum of W*H*D over completed         */
    double first_submit_s, last_complete_s;  /* steady-clock seconds       */
    double peak_inflight_pxd;        /* max of the admitted W*H*D           */
} vxh_server_stats;

typedef struct vxh_server vxh_server;

/* Device bytes a request is admitted against: its pipeline workspace
 * (vxh_sgm_workspace_bytes: the per-direction cost volumes, their high-byte
 * volumes and the census). */
uint64_t vxh_server_request_bytes(uint32_t w, uint32_t h, const vxh_sgm_config *cfg);

vxh_status vxh_server_create(int device, const vxh_server_config *cfg,
                             vxh_server **out);
/* Pre-allocates `count` idle device slots (and pinned staging buffers) sized
 * for w x h pairs with `cfg`, so that the first requests of that shape (or any
 * smaller one) find their buffers resident instead of allocating them.
 * Slots are reused by capacity: any request of at most w*h pixels and at most
 * that workspace takes one. */
vxh_status vxh_server_reserve(vxh_server *srv, uint32_t w, uint32_t h,
                              const vxh_sgm_config *cfg, uint32_t count);
/* Queues a request.  left/right (w*h bytes) are copied into pinned staging
 * memory before the call returns; out (w*h floats) must stay valid until the
 * request completes.
 * A request whose volume alone exceeds the budget is rejected with
 * VXH_ERR_BUFFER_ALLOC; an invalid config with VXH_ERR_INVALID_CONFIG. */
vxh_status vxh_server_submit(vxh_server *srv, const uint8_t *left,
                             const uint8_t *right, uint32_t w, uint32_t h,
                             const vxh_sgm_config *cfg, float *out,
                             uint64_t *request_id);
/* Blocks until request `id` has completed; returns its latency in ms.
 * VXH_ERR_GPU when its device work failed, including an expired bounded wait
 * of a split scanline (its map is then not written to out). */
vxh_status vxh_server_wait(vxh_server *srv, uint64_t id, double *latency_ms);
/* Blocks until every submitted request has completed. */
vxh_status vxh_server_drain(vxh_server *srv);
vxh_status vxh_server_get_stats(vxh_server *srv, vxh_server_stats *out);
/* Clears latency samples and counters (not the slot pool). */
void vxh_server_reset_stats(vxh_server *srv);
void vxh_server_destroy(vxh_server *srv);
/* Message of the last server failure on this thread. */
const char *vxh_server_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
