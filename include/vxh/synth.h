/*
 * Seeded, integer-only synthetic rectified stereo pairs.
 *
 * The reference generates its benchmark pair on the host
 * (vx-vision/benches/kernel_benchmarks.rs:14-24 and :199-208: a gradient image
 * and a copy shifted by 8 px).  This generator replaces that with a pair whose
 * ground truth has many disparities, so that every path, occlusion and the
 * left image border are exercised.  It uses integer arithmetic only, so any
 * host reproduces the same bytes.
 *
 * Algorithm (all draws from one SplitMix64 stream, in this order):
 *   1. raw[i] = next() & 0xFF for i in raster order;
 *      left = 3x3 box filter of raw, (sum + 4) / 9, clamp-to-edge borders.
 *   2. disparity field: background min(8, nd-1); n_rects rectangles, each
 *      drawn as x0 = u(w), y0 = u(h), rw = 1 + u(max(1, w/4)),
 *      rh = 1 + u(max(1, h/4)), d = u(nd), painted in draw order;
 *      n_planes slanted planes, each x0 = u(w), y0 = u(h),
 *      rw = 1 + u(max(1, w/3)), rh = 1 + u(max(1, h/3)), a = u(nd),
 *      bx = u(65) - 32, by = u(65) - 32, painting
 *      clamp(floor((1024 a + bx (x-x0) + by (y-y0) + 512) / 1024), 0, nd-1).
 *      u(n) = next() % n.
 *   3. right[y][x - d] = left[y][x] for x - d >= 0, the larger d winning a
 *      collision; pixels never written get next() & 0xFF in raster order.
 *   4. right += round(2 * (sum of 12 uniforms - 6)) per pixel in raster
 *      order, computed as floor((s - 393216 + 16384) / 32768) where s is the
 *      sum of 12 draws of next() & 0xFFFF; clamped to [0, 255].
 */
#ifndef VXH_SYNTH_H
#define VXH_SYNTH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Fills left and right (w*h bytes each) and, if truth is not NULL, the integer
 * ground-truth disparity field (w*h int32).  Returns 0, or 6 for w, h or nd of
 * zero. */
int vxh_synth_pair(uint32_t w, uint32_t h, uint32_t nd, uint32_t n_rects,
                   uint32_t n_planes, uint64_t seed, uint8_t *left,
                   uint8_t *right, int32_t *truth);

#ifdef __cplusplus
}
#endif

#endif
