/*
 * vxh depth post-processing: the consumers of the SGM disparity map.
 *
 *   DepthFilter (vx-vision/src/kernels/depth_filter.rs:22-319,
 *   vx-vision/shaders/DepthFilter.metal:1-120): edge-preserving bilateral
 *   smoothing and a median filter of an R32Float depth (or disparity) map;
 *   disparity to depth (vx-vision/src/kernels/sgm.rs:284-286, where the
 *   reference leaves `depth = fx * baseline / disparity` to the caller).
 *
 * Results: the median is bit-identical to the reference's definition (the
 * value at index count / 2 of the first 121 values > 0 of the window in scan
 * order, sorted ascending; 0 when there is none); the conversion is
 * (fx * baseline) / d in IEEE fp32 for d > 0 and 0 otherwise; the bilateral
 * filter follows the shader's fp32 expression order and differs from an IEEE
 * CPU evaluation only by the last-ulp error of expf in each weight.
 */
#ifndef VXH_DEPTH_H
#define VXH_DEPTH_H

#include <stddef.h>
#include <stdint.h>

#include "vxh/sgm.h"

#ifdef __cplusplus
extern "C" {
#endif

/* DepthFilterConfig (depth_filter.rs:22-50): window radius, spatial and
 * depth sigmas of the bilateral filter; default 3, 5.0, 0.05. */
typedef struct {
    int32_t radius;
    float sigma_spatial;
    float sigma_depth;
} vxh_depth_filter_config;
void vxh_depth_filter_config_default(vxh_depth_filter_config *cfg);

/* DepthMedianConfig (depth_filter.rs:54-63): window radius, default 1. */
typedef struct {
    int32_t radius;
} vxh_depth_median_config;
void vxh_depth_median_config_default(vxh_depth_median_config *cfg);

typedef struct vxh_depth_filter vxh_depth_filter;

/* DepthFilter::new (depth_filter.rs:75-99). */
vxh_status vxh_depth_filter_create(vxh_context *ctx, vxh_depth_filter **out);
void vxh_depth_filter_destroy(vxh_depth_filter *f);

/* DepthFilter::apply_bilateral / apply_median (depth_filter.rs:101-181):
 * R32Float input and output textures of the same size, synchronous.  A
 * non-R32Float texture -> VXH_ERR_INVALID_CONFIG, different sizes ->
 * VXH_ERR_TEXTURE_SIZE_MISMATCH. */
vxh_status vxh_depth_filter_apply_bilateral(vxh_depth_filter *f, vxh_context *ctx,
                                            const vxh_texture *input, vxh_texture *output,
                                            const vxh_depth_filter_config *cfg);
vxh_status vxh_depth_filter_apply_median(vxh_depth_filter *f, vxh_context *ctx,
                                         const vxh_texture *input, vxh_texture *output,
                                         const vxh_depth_median_config *cfg);

/* DepthFilter::encode_bilateral / encode_median (depth_filter.rs:183-249):
 * stream-ordered on device buffers of w*h floats (in and out must not
 * overlap).  A radius above 255 -> VXH_ERR_INVALID_CONFIG (a negative radius
 * is an empty window, as in the shader's loops). */
vxh_status vxh_depth_bilateral_enqueue(const float *d_in, float *d_out, uint32_t w,
                                       uint32_t h, const vxh_depth_filter_config *cfg,
                                       void *stream);
vxh_status vxh_depth_median_enqueue(const float *d_in, float *d_out, uint32_t w, uint32_t h,
                                    const vxh_depth_median_config *cfg, void *stream);

/* Disparity map to depth: (fx * baseline) / d where d > 0, else 0 (invalid
 * pixels and the LR-check sentinel -1 become 0).  Stream-ordered, w*h floats;
 * d_depth may equal d_disp. */
vxh_status vxh_disparity_to_depth(const float *d_disp, float *d_depth, uint32_t w, uint32_t h,
                                  float fx, float baseline, void *stream);

#ifdef __cplusplus
}
#endif

#endif
