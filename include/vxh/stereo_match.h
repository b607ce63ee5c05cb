/*
 * vxh sparse stereo: ORB descriptor matching between rectified images.
 *
 *   StereoMatcher  vx-vision/src/kernels/stereomatch.rs:17-339 and
 *                  vx-vision/shaders/StereoMatch.metal:32-115: for every left
 *                  keypoint, the right keypoint of least Hamming distance
 *                  among those on its epipolar band (|dy| <= max_epipolar)
 *                  with a disparity lx - rx in [min_disparity, max_disparity];
 *                  accepted when that distance <= max_hamming, then
 *                  triangulated.
 *
 * Results, defined as the shader computes them:
 *   - distance matrix n_left x n_right u16, 0xFFFF where the band or the
 *     disparity range excludes the pair (a NaN position excludes it too);
 *   - best = the first (lowest right index) of the least distances, over all
 *     right keypoints including the 0xFFFF entries; accepted iff
 *     best <= (max_hamming mod 2^16) (the shader's ushort conversion, :96);
 *     the Lowe ratio test is commented out in the shader (:98-101), so
 *     ratio_thresh is accepted and unused;
 *   - disparity = lx - rx, Z = (fx * baseline) / disparity,
 *     X = (Z * (lx - cx)) / fx, Y = (Z * (ly - cy)) / fy in IEEE fp32.
 * Matches come out in ascending left index (the shader appends with an
 * atomic counter, in no defined order).
 */
#ifndef VXH_STEREO_MATCH_H
#define VXH_STEREO_MATCH_H

#include <stddef.h>
#include <stdint.h>

#include "vxh/pipeline.h"
#include "vxh/sgm.h"

#ifdef __cplusplus
extern "C" {
#endif

/* CornerPoint (vx-vision/src/types.rs:6-12): 16 bytes. */
typedef struct {
    float x, y;
    float response;
    uint32_t pyramid_level;
} vxh_corner_point;

/* ORBOutput (types.rs:34-39): a 256-bit descriptor and its angle, 36 bytes. */
typedef struct {
    uint32_t desc[8];
    float angle;
} vxh_orb_output;

/* StereoMatchResult (types.rs:62-70): 28 bytes, the Rust struct's layout
 * (the shader's own struct puts point_3d at byte 16, a float3). */
typedef struct {
    uint32_t left_idx;
    uint32_t right_idx;
    float disparity;
    float point_3d[3];
    float pad;
} vxh_stereo_match;

/* StereoConfig and its Default (stereomatch.rs:19-69). */
typedef struct {
    float max_epipolar;     /* 2.0  */
    float min_disparity;    /* 1.0  */
    float max_disparity;    /* 120.0 */
    uint32_t max_hamming;   /* 50   */
    float ratio_thresh;     /* 0.8 (unused, as in the shader) */
    float fx, fy, cx, cy;   /* 600, 600, 320, 240 */
    float baseline;         /* 0.12 */
} vxh_stereo_config;
void vxh_stereo_config_default(vxh_stereo_config *cfg);

typedef struct vxh_stereo_matcher vxh_stereo_matcher;

/* StereoMatcher::new (stereomatch.rs:86-112). */
vxh_status vxh_stereo_matcher_create(vxh_context *ctx, vxh_stereo_matcher **out);
void vxh_stereo_matcher_destroy(vxh_stereo_matcher *m);

/* StereoMatcher::run (stereomatch.rs:114-247): host keypoints and
 * descriptors in, matches out (capacity n_left; *n_out receives the count).
 * Synchronous.  n_left or n_right 0 -> no matches. */
vxh_status vxh_stereo_matcher_run(vxh_stereo_matcher *m, vxh_context *ctx,
                                  const vxh_corner_point *left_kpts, size_t n_left,
                                  const vxh_corner_point *right_kpts, size_t n_right,
                                  const vxh_orb_output *left_descs,
                                  const vxh_orb_output *right_descs,
                                  const vxh_stereo_config *cfg, vxh_stereo_match *out,
                                  size_t *n_out);

/* Stream-ordered form on device buffers: keypoints as vxh_corner_point,
 * descriptors as 8 uint32 words per keypoint (ORBOutput.desc, flattened as
 * stereomatch.rs:144-145 does), d_matches with room for n_left entries,
 * *d_count receives the number of matches.  n_left, n_right < 2^31. */
vxh_status vxh_stereo_match_enqueue(const vxh_corner_point *d_left_kpts,
                                    const vxh_corner_point *d_right_kpts,
                                    const uint32_t *d_left_desc, const uint32_t *d_right_desc,
                                    uint32_t n_left, uint32_t n_right,
                                    const vxh_stereo_config *cfg, vxh_stereo_match *d_matches,
                                    uint32_t *d_count, void *stream);

/* The shader's first kernel alone (StereoMatch.metal:32-62): the distance
 * matrix, row-major n_left x n_right u16.  For tests and for callers that
 * want the whole matrix; the matcher itself never writes it. */
vxh_status vxh_stereo_hamming_matrix_enqueue(const vxh_corner_point *d_left_kpts,
                                             const vxh_corner_point *d_right_kpts,
                                             const uint32_t *d_left_desc,
                                             const uint32_t *d_right_desc, uint32_t n_left,
                                             uint32_t n_right, const vxh_stereo_config *cfg,
                                             uint16_t *d_matrix, void *stream);

#ifdef __cplusplus
}
#endif

#endif
