/*
 * vxh rectification: the input side of the stereo path.
 *
 *   Undistorter   vx-vision/src/kernels/undistort.rs:14-137 and
 *                 vx-vision/shaders/undistort.metal:4-20: remaps an image
 *                 through per-pixel coordinate maps (OpenCV-style remap) with
 *                 bilinear filtering and clamp-to-edge addressing.
 *   rectify maps  no reference counterpart (the reference consumes maps but
 *                 builds none): the maps of a pinhole camera with radial /
 *                 tangential distortion (k1, k2, p1, p2, k3), a rectifying
 *                 rotation R and a new projection P, in the coordinates the
 *                 remap reads, so a calibrated stereo pair can be rectified on
 *                 the device before SGM.
 *
 * Remap, per output pixel (x, y), in IEEE fp32 with no contraction:
 *   sx = map_x[y][x], sy = map_y[y][x]   (pixel coordinates: texel i spans
 *                                          [i, i+1), its centre is i + 0.5)
 *   u = min(max(sx - 0.5, -1), W),  v = min(max(sy - 0.5, -1), H)
 *                                         (IEEE min/max: a NaN coordinate
 *                                          acts as -1; the clamp changes no
 *                                          finite result, edges repeat)
 *   i = floor(u), j = floor(v), a = u - i, b = v - j
 *   t(i, j) = input texel at (clamp(i, 0, W-1), clamp(j, 0, H-1)), k / 255
 *             for a gray8 texel k
 *   top = t(i,j) + a * (t(i+1,j) - t(i,j)),  bot = likewise at j + 1
 *   value = top + b * (bot - top)
 *   R32Float output: value; gray8 output: rint(min(max(value, 0), 1) * 255)
 *   (round half to even).
 * Metal's sampler filters with the GPU's own fixed-point weights, which no
 * document states; this definition is exact fp32 and identical on the
 * device and in the oracle.
 */
#ifndef VXH_RECTIFY_H
#define VXH_RECTIFY_H

#include <stddef.h>
#include <stdint.h>

#include "vxh/pipeline.h"
#include "vxh/sgm.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vxh_undistorter vxh_undistorter;

/* Undistorter::new (undistort.rs:20-35). */
vxh_status vxh_undistorter_create(vxh_context *ctx, vxh_undistorter **out);
void vxh_undistorter_destroy(vxh_undistorter *u);

/* Undistorter::apply (undistort.rs:37-73): synchronous.  input gray8 or
 * R32Float of any size; map_x, map_y R32Float of the output's size; output
 * gray8 or R32Float.  A map of another size -> VXH_ERR_TEXTURE_SIZE_MISMATCH,
 * a map or texture of another format -> VXH_ERR_INVALID_CONFIG. */
vxh_status vxh_undistorter_apply(vxh_undistorter *u, vxh_context *ctx,
                                 const vxh_texture *input, const vxh_texture *map_x,
                                 const vxh_texture *map_y, vxh_texture *output);

/* Undistorter::encode (undistort.rs:75-97) into a pipeline. */
vxh_status vxh_undistorter_encode(vxh_undistorter *u, vxh_pipeline *p,
                                  const vxh_texture *input, const vxh_texture *map_x,
                                  const vxh_texture *map_y, vxh_texture *output);

/* Stream-ordered remap on device buffers: input in_w x in_h of in_format
 * (1 gray8, 2 R32Float), maps and output out_w x out_h (output of
 * out_format).  Buffers must not overlap the output. */
vxh_status vxh_remap_enqueue(const void *d_in, uint32_t in_format, uint32_t in_w,
                             uint32_t in_h, const float *d_map_x, const float *d_map_y,
                             void *d_out, uint32_t out_format, uint32_t out_w, uint32_t out_h,
                             void *stream);

/* Camera model of the rectify maps. */
typedef struct {
    double fx, fy, cx, cy;          /* camera matrix K of the distorted image */
    double k1, k2, p1, p2, k3;      /* distortion coefficients                */
    double R[9];                    /* rectifying rotation, row-major         */
    double new_fx, new_fy, new_cx, new_cy;   /* projection of the output     */
} vxh_rectify_params;

/* With fx, fy, cx, cy already set: R = identity and new_* = K, i.e. plain
 * undistortion into the same camera (distortion fields are left as they are). */
void vxh_rectify_params_default(vxh_rectify_params *p);

/* The 3 x 3 matrix the maps use: inverse(Knew * R), row-major, computed in
 * double and rounded to float (m9).  VXH_ERR_INVALID_CONFIG when singular. */
vxh_status vxh_rectify_inverse(const vxh_rectify_params *p, float *m9);

/* Builds map_x / map_y (w x h floats each) on `stream`.  Per output pixel
 * (x, y), in fp32 with M = vxh_rectify_inverse(p) and the coefficients
 * rounded to float:
 *   X = M0 x + M1 y + M2,  Y = M3 x + M4 y + M5,  Z = M6 x + M7 y + M8
 *   q = 1 / Z,  x' = X q,  y' = Y q,  r2 = x'x' + y'y'
 *   kr = 1 + ((k3 r2 + k2) r2 + k1) r2
 *   xd = x' kr + p1 (2 x' y') + p2 (r2 + 2 x'x')
 *   yd = y' kr + p1 (r2 + 2 y'y') + p2 (2 x' y')
 *   map_x = fx xd + cx + 0.5,  map_y = fy yd + cy + 0.5
 * (OpenCV's undistort/rectify model; + 0.5 turns its pixel-centre
 * coordinates into the remap's).  Each product and sum is rounded in the
 * order written, left to right. */
vxh_status vxh_rectify_maps_enqueue(const vxh_rectify_params *p, uint32_t w, uint32_t h,
                                    float *d_map_x, float *d_map_y, void *stream);

/* Texture form: map_x / map_y are R32Float textures of one size; synchronous. */
vxh_status vxh_rectify_maps(vxh_context *ctx, const vxh_rectify_params *p,
                            vxh_texture *map_x, vxh_texture *map_y);

#ifdef __cplusplus
}
#endif

#endif
