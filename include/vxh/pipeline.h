/*
 * vxh pipelines and texture pools: several kernels chained into one unit of
 * GPU work, and reuse of device textures between frames.
 *
 *   Pipeline     vx-vision/src/pipeline.rs:12-80 (one Metal command buffer:
 *                begin, encode any number of kernels, commit, wait)
 *   TexturePool  vx-vision/src/pool.rs:17-147 (textures cached per
 *                (width, height, format), at most max_per_bucket per key)
 *   encode()     the second entry point every reference kernel has beside
 *                its synchronous one (docs/src/adding-a-kernel.md:101-138);
 *                here vxh_sgm_encode and vxh_depth_filter_encode_*.
 *
 * A pipeline records, it does not run: vxh_pipeline_begin starts a stream
 * capture on a stream of the context's pool, every encode (and any
 * stream-ordered work the caller puts on vxh_pipeline_stream) is captured,
 * and vxh_pipeline_commit turns the capture into one HIP graph and launches
 * it.  Nothing encoded runs before the commit, as with a Metal command buffer
 * that has not been committed.  The scratch memory of the encoded SGM passes
 * comes from the context's device pool and stays with the pipeline until it
 * is destroyed, so a committed pipeline can be launched again
 * (vxh_pipeline_relaunch) on the same textures without re-encoding.
 */
#ifndef VXH_PIPELINE_H
#define VXH_PIPELINE_H

#include <stddef.h>
#include <stdint.h>

#include "vxh/depth.h"
#include "vxh/sgm.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- Pipeline (vx-vision/src/pipeline.rs:12-80) ---- */
/* One pipeline is used from one thread at a time (like the reference's,
 * which is neither Send nor Sync); separate pipelines may be encoded from
 * separate threads at once. */
typedef struct vxh_pipeline vxh_pipeline;

/* Pipeline::begin (pipeline.rs:21-31): a new, empty pipeline on ctx's
 * device.  VXH_ERR_GPU when the capture cannot start. */
vxh_status vxh_pipeline_begin(vxh_context *ctx, vxh_pipeline **out);

/* Pipeline::cmd_buf (pipeline.rs:34-36): the hipStream_t being captured.
 * Stream-ordered calls made on it before the commit become part of the
 * pipeline (device work only: kernels, device-to-device copies, events).
 * NULL once committed. */
void *vxh_pipeline_stream(const vxh_pipeline *p);

/* Pipeline::retain / retain_all (pipeline.rs:39-46): the pipeline takes
 * ownership of `tex` and keeps it alive until it is taken back
 * (vxh_pipeline_take_retained) or the pipeline is destroyed, which destroys
 * it. */
vxh_status vxh_pipeline_retain(vxh_pipeline *p, vxh_texture *tex);

/* Pipeline::commit (pipeline.rs:57-62): ends the capture, instantiates the
 * graph and launches it; a second commit does nothing.  An encode that failed
 * to launch during the capture makes the commit fail with VXH_ERR_GPU (the
 * pipeline then runs nothing). */
vxh_status vxh_pipeline_commit(vxh_pipeline *p);

/* Pipeline::wait (pipeline.rs:65-69): blocks until the committed work is
 * done; returns at once when nothing was committed.  VXH_ERR_GPU when the
 * device reported an error, or when a split scanline's bounded wait in one of
 * the encoded SGM passes expired. */
vxh_status vxh_pipeline_wait(vxh_pipeline *p);

/* Pipeline::commit_and_wait (pipeline.rs:49-54): commit, then wait; the
 * retained textures are then taken back with vxh_pipeline_take_retained. */
vxh_status vxh_pipeline_commit_and_wait(vxh_pipeline *p);

/* Pipeline::take_retained (pipeline.rs:72-74): moves up to `cap` retained
 * textures into out (in retain order), returns how many were moved; the
 * caller owns them afterwards.  With out == NULL returns the number held. */
size_t vxh_pipeline_take_retained(vxh_pipeline *p, vxh_texture **out, size_t cap);

/* Pipeline::is_committed (pipeline.rs:77-79): 1 after the commit, else 0. */
int vxh_pipeline_is_committed(const vxh_pipeline *p);

/* Number of kernel launches and copies the pipeline recorded (0 before the
 * commit).  No reference counterpart. */
size_t vxh_pipeline_node_count(const vxh_pipeline *p);

/* Launches the committed graph again on the pipeline's stream (after the
 * previous launch, in stream order): the same kernels on the same textures
 * and scratch, reading the textures' current contents.  Wait with
 * vxh_pipeline_wait.  VXH_ERR_INVALID_CONFIG before the commit.  No
 * reference counterpart (a Metal command buffer runs once). */
vxh_status vxh_pipeline_relaunch(vxh_pipeline *p);

/* Drop of a Pipeline: an uncommitted pipeline is discarded without running;
 * a committed one is waited for.  Its scratch and stream go back to the
 * context's pool and retained textures are destroyed.  The context must
 * outlive its pipelines. */
void vxh_pipeline_destroy(vxh_pipeline *p);

/* SGMStereo as a pipeline stage: the encode() form of
 * vxh_sgm_compute_disparity (same validation and result), captured into `p`
 * with scratch from the context's device pool.  VXH_ERR_INVALID_CONFIG once
 * `p` is committed. */
vxh_status vxh_sgm_encode(vxh_sgm *sgm, vxh_pipeline *p, const vxh_texture *left,
                          const vxh_texture *right, vxh_texture *output,
                          const vxh_sgm_config *cfg);

/* DepthFilter::encode_bilateral / encode_median (depth_filter.rs:183-246) on
 * textures: validated as the apply_* calls, captured into `p`. */
vxh_status vxh_depth_filter_encode_bilateral(vxh_depth_filter *f, vxh_pipeline *p,
                                             const vxh_texture *input, vxh_texture *output,
                                             const vxh_depth_filter_config *cfg);
vxh_status vxh_depth_filter_encode_median(vxh_depth_filter *f, vxh_pipeline *p,
                                          const vxh_texture *input, vxh_texture *output,
                                          const vxh_depth_median_config *cfg);

/* Context scratch pool: out3 = blocks held, bytes held, bytes in use by live
 * pipelines.  No reference counterpart (Metal allocates per command). */
vxh_status vxh_context_pool_stats(vxh_context *ctx, size_t *out3);
/* Frees the context's idle scratch blocks and streams. */
void vxh_context_pool_trim(vxh_context *ctx);

/* ---- TexturePool (vx-vision/src/pool.rs:17-147) ---- */
typedef struct vxh_texture_pool vxh_texture_pool;

/* TexturePool::new (max_per_bucket 8) / with_capacity (pool.rs:26-43). */
vxh_status vxh_texture_pool_create(size_t max_per_bucket, vxh_texture_pool **out);
/* Destroys the pool and every texture cached in it (acquired textures are
 * the caller's).  As for any texture, the cached ones must be destroyed
 * before the context they were made on. */
void vxh_texture_pool_destroy(vxh_texture_pool *pool);

/* TexturePool::acquire (pool.rs:61-81): a cached texture of that size and
 * format (1 gray8, 2 R32Float, 3 RGBA8) if one is cached, else a new one on
 * ctx's device.  Contents are not cleared. */
vxh_status vxh_texture_pool_acquire(vxh_texture_pool *pool, vxh_context *ctx, uint32_t w,
                                    uint32_t h, uint32_t format, vxh_texture **out);
/* acquire_gray8 / acquire_r32float / acquire_rgba8 (pool.rs:46-58). */
vxh_status vxh_texture_pool_acquire_gray8(vxh_texture_pool *pool, vxh_context *ctx,
                                          uint32_t w, uint32_t h, vxh_texture **out);
vxh_status vxh_texture_pool_acquire_r32float(vxh_texture_pool *pool, vxh_context *ctx,
                                             uint32_t w, uint32_t h, vxh_texture **out);
vxh_status vxh_texture_pool_acquire_rgba8(vxh_texture_pool *pool, vxh_context *ctx,
                                          uint32_t w, uint32_t h, vxh_texture **out);
/* TexturePool::release (pool.rs:86-98): caches `tex` under its key, or
 * destroys it when that key already holds max_per_bucket textures. */
void vxh_texture_pool_release(vxh_texture_pool *pool, vxh_texture *tex);
/* TexturePool::clear (pool.rs:107-109): destroys every cached texture. */
void vxh_texture_pool_clear(vxh_texture_pool *pool);
/* cached_count / total_acquires / total_hits / hit_rate (pool.rs:112-133). */
size_t vxh_texture_pool_cached_count(const vxh_texture_pool *pool);
uint64_t vxh_texture_pool_total_acquires(const vxh_texture_pool *pool);
uint64_t vxh_texture_pool_total_hits(const vxh_texture_pool *pool);
double vxh_texture_pool_hit_rate(const vxh_texture_pool *pool);

#ifdef __cplusplus
}
#endif

#endif
