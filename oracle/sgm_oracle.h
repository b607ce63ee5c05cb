/*
 * Scalar CPU oracle for vx-vision's SGMStereo path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path is
 * compared against.  Only the test suite, the smoke check and the CPU-baseline
 * leg of the benchmark may load it; the product library never links it and
 * never falls back to it.
 *
 * Every function restates the reference's Metal shader
 * (vx-vision/shaders/SGMStereo.metal) and host orchestration
 * (vx-vision/src/kernels/sgm.rs) in plain scalar C++; the citations are on the
 * definitions in sgm_oracle.cpp.  Extensions that the reference lacks
 * (reverse/diagonal paths beyond its two, the left-right check) follow the
 * definitions written next to their functions.
 */
#ifndef VX_SGM_ORACLE_H
#define VX_SGM_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Census transform of one gray8 image (row-major, w*h bytes) into
 * 2*w*h uint32 words: out[2*i] = lo bits 0..31, out[2*i+1] = hi bits 32..63. */
void oracle_census(const uint8_t *img, uint32_t w, uint32_t h, uint32_t rx,
                   uint32_t ry, uint32_t *out);

/* Matching cost C(x, y, d) for one pixel (Hamming distance of census words). */
uint32_t oracle_cost(const uint32_t *lc, const uint32_t *rc, uint32_t w,
                     uint32_t x, uint32_t y, uint32_t d);

/* Aggregates one scan direction (dx, dy in {-1,0,1}, not both 0) and adds its
 * per-path costs into the summed volume S (u16, (y*w+x)*nd+d), mod 2^16. */
void oracle_aggregate_dir(const uint32_t *lc, const uint32_t *rc, uint32_t w,
                          uint32_t h, uint32_t nd, uint32_t p1, uint32_t p2,
                          int dx, int dy, uint16_t *S);

/* Winner-take-all with parabola subpixel refinement into a w*h float map. */
void oracle_wta(const uint16_t *S, uint32_t w, uint32_t h, uint32_t nd,
                float *out);

/* Integer argmin disparity for every left pixel (the WTA's best_d). */
void oracle_wta_index(const uint16_t *S, uint32_t w, uint32_t h, uint32_t nd,
                      int32_t *out);

/* Left-right consistency check applied to a WTA map in place. */
void oracle_lr_check(const uint16_t *S, uint32_t w, uint32_t h, uint32_t nd,
                     uint32_t tol, float *disp);

/* Direction table: index -> (dx, dy).  Paths 2, 4 and 8 use the first
 * num_paths entries. */
int oracle_direction(uint32_t index, int *dx, int *dy);

/* Full pipeline.  Returns 0 on success, a non-zero code for an invalid config
 * (the same codes as the product's vxh_status).  nthreads <= 1 runs the
 * single-core scalar path; more threads split independent scanlines and give
 * identical bits. */
int oracle_sgm(const uint8_t *left, const uint8_t *right, uint32_t w,
               uint32_t h, uint32_t nd, float p1, float p2, uint32_t rx,
               uint32_t ry, uint32_t num_paths, uint32_t lr_check,
               uint32_t lr_tol, uint32_t nthreads, float *out);

/* Same, but also returns the summed volume (w*h*nd u16) for volume-level
 * comparisons. */
int oracle_sgm_volume(const uint8_t *left, const uint8_t *right, uint32_t w,
                      uint32_t h, uint32_t nd, float p1, float p2, uint32_t rx,
                      uint32_t ry, uint32_t num_paths, uint32_t nthreads,
                      uint16_t *S);

/* Single-scanline recurrence on caller-supplied costs, for hand-derived
 * known-answer vectors: costs is steps*nd bytes; out receives steps*nd u16
 * per-path aggregated values. */
void oracle_scanline(const uint8_t *costs, uint32_t steps, uint32_t nd,
                     uint32_t p1, uint32_t p2, uint16_t *out);

/* Depth filters and disparity-to-depth (depth_oracle.cpp). */
void oracle_depth_bilateral(const float *in, uint32_t w, uint32_t h, int radius, float ss,
                            float sd, float *out);
void oracle_depth_median(const float *in, uint32_t w, uint32_t h, int radius, float *out);
void oracle_disparity_to_depth(const float *disp, size_t n, float fx, float baseline,
                               float *out);

/* Undistorter remap (rectify_oracle.cpp): formats 1 = gray8, 2 = R32Float. */
void oracle_remap(const void *in, uint32_t in_fmt, uint32_t in_w, uint32_t in_h,
                  const float *map_x, const float *map_y, void *out, uint32_t out_fmt,
                  uint32_t out_w, uint32_t out_h);
/* Rectify maps from 22 doubles (fx fy cx cy k1 k2 p1 p2 k3 R[9] nfx nfy ncx
 * ncy); m_out (may be NULL) receives the float inverse matrix.  0 or 6. */
int oracle_rectify_maps(const double *p, uint32_t w, uint32_t h, float *map_x, float *map_y,
                        float *m_out);

/* Sparse stereo matcher (stereo_match_oracle.cpp): keypoints 16 bytes each,
 * descriptors 8 words each, p = max_epipolar, min_disparity, max_disparity,
 * ratio_thresh, fx, fy, cx, cy, baseline. */
void oracle_stereo_matrix(const void *lk, const void *rk, const uint32_t *ld,
                          const uint32_t *rd, uint32_t nl, uint32_t nr, const float *p,
                          uint16_t *out);
uint32_t oracle_stereo_match(const void *lk, const void *rk, const uint32_t *ld,
                             const uint32_t *rd, uint32_t nl, uint32_t nr, const float *p,
                             uint32_t max_hamming, uint32_t *out);

#ifdef __cplusplus
}
#endif

#endif
