/*
 * Scalar CPU oracle for vx-vision's depth filters and the disparity-to-depth
 * conversion (TEST INFRASTRUCTURE ONLY: the checker the HIP kernels are
 * compared against, never part of the product).
 *
 * Restates:
 *   bilateral   vx-vision/shaders/DepthFilter.metal:17-60  (depth_bilateral)
 *   median      vx-vision/shaders/DepthFilter.metal:66-120 (depth_median)
 *   to depth    vx-vision/src/kernels/sgm.rs:284-286 (depth = fx * baseline /
 *               disparity, left to the caller there), defined here for
 *               disparity > 0 as (fx * baseline) / disparity in fp32 and 0
 *               for every other value (invalid pixels, the LR sentinel -1).
 *
 * Semantics kept:
 *   - all arithmetic in fp32, loops dy outer / dx inner, clamp-to-edge reads;
 *   - bilateral: a centre <= 0 writes 0; neighbours <= 0 are skipped; the
 *     weight is exp(-(dx^2+dy^2) * inv_2ss) * exp(-(n-c)^2 * inv_2sd) with
 *     inv_2ss = 1 / (2 * ss * ss), inv_2sd = 1 / (2 * sd * sd); the result is
 *     sum / weight_sum when weight_sum > 1e-8, else the centre;
 *   - median: the first 121 values > 0 in scan order, sorted ascending, the
 *     one at index count / 2; 0 when none.
 */
#include <cmath>
#include <cstdint>

#include "sgm_oracle.h"

extern "C" {

void oracle_depth_bilateral(const float *in, uint32_t w, uint32_t h, int radius, float ss,
                            float sd, float *out) {
    const float inv_2ss = 1.0f / (2.0f * ss * ss);
    const float inv_2sd = 1.0f / (2.0f * sd * sd);
    const int wm = (int)w - 1, hm = (int)h - 1;
    for (int cy = 0; cy < (int)h; ++cy)
        for (int cx = 0; cx < (int)w; ++cx) {
            const float center = in[(size_t)cy * w + cx];
            float &o = out[(size_t)cy * w + cx];
            if (center <= 0.0f) {
                o = 0.0f;
                continue;
            }
            float sum = 0.0f, weight_sum = 0.0f;
            for (int dy = -radius; dy <= radius; ++dy)
                for (int dx = -radius; dx <= radius; ++dx) {
                    const int sx = cx + dx < 0 ? 0 : (cx + dx > wm ? wm : cx + dx);
                    const int sy = cy + dy < 0 ? 0 : (cy + dy > hm ? hm : cy + dy);
                    const float n = in[(size_t)sy * w + sx];
                    if (n <= 0.0f) continue;
                    const float sd2 = (float)(dx * dx + dy * dy);
                    const float ws = std::exp(-sd2 * inv_2ss);
                    const float diff = n - center;
                    const float wd = std::exp(-diff * diff * inv_2sd);
                    const float wgt = ws * wd;
                    sum += n * wgt;
                    weight_sum += wgt;
                }
            o = weight_sum > 1e-8f ? sum / weight_sum : center;
        }
}

void oracle_depth_median(const float *in, uint32_t w, uint32_t h, int radius, float *out) {
    const int wm = (int)w - 1, hm = (int)h - 1;
    float vals[121];
    for (int cy = 0; cy < (int)h; ++cy)
        for (int cx = 0; cx < (int)w; ++cx) {
            int count = 0;
            for (int dy = -radius; dy <= radius; ++dy)
                for (int dx = -radius; dx <= radius; ++dx) {
                    const int sx = cx + dx < 0 ? 0 : (cx + dx > wm ? wm : cx + dx);
                    const int sy = cy + dy < 0 ? 0 : (cy + dy > hm ? hm : cy + dy);
                    const float d = in[(size_t)sy * w + sx];
                    if (d > 0.0f && count < 121) vals[count++] = d;
                }
            for (int i = 0; i < count - 1; ++i)
                for (int j = 0; j < count - 1 - i; ++j)
                    if (vals[j] > vals[j + 1]) {
                        const float t = vals[j];
                        vals[j] = vals[j + 1];
                        vals[j + 1] = t;
                    }
            out[(size_t)cy * w + cx] = count > 0 ? vals[count / 2] : 0.0f;
        }
}

void oracle_disparity_to_depth(const float *disp, size_t n, float fx, float baseline,
                               float *out) {
    const float fb = fx * baseline;
    for (size_t i = 0; i < n; ++i) out[i] = disp[i] > 0.0f ? fb / disp[i] : 0.0f;
}

}  // extern "C"
