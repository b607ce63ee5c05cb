/*
 * Scalar CPU oracle for vx-vision's Undistorter remap and for the rectify
 * maps that feed it (TEST INFRASTRUCTURE ONLY: the checker the HIP kernels
 * are compared against, never part of the product).
 *
 * Restates:
 *   remap         vx-vision/shaders/undistort.metal:4-20: per output pixel,
 *                 read (map_x, map_y), sample the input there with
 *                 coord::pixel, clamp_to_edge, filter::linear, write.  The
 *                 linear filter is defined in exact fp32 (texel centres at
 *                 i + 0.5, weights a = u - floor(u), lerp x + t (y - x),
 *                 columns first); gray8 texels read as k / 255 and gray8
 *                 output is rint(clamp(v, 0, 1) * 255) (round half to even).
 *                 Coordinates are clamped to [-1, W] / [-1, H] after the
 *                 -0.5 shift, which changes no finite result and maps a NaN
 *                 to the first edge (IEEE fmax returns the number).
 *   rectify maps  no reference counterpart: OpenCV's undistort/rectify model
 *                 (K, k1 k2 p1 p2 k3, R, Knew) with M = inverse(Knew R) in
 *                 double rounded to float, then per pixel in fp32 in the order
 *                 written below, plus 0.5 into the remap's coordinates.
 *
 * Parity pins (tests/test_oracle_rectify.py): identity maps reproduce the
 * input bit for bit; integer shifts move it with repeated edges; half-texel
 * maps give the exact midpoints; a float64 numpy bilinear agrees to rounding;
 * unit-K zero-distortion maps are exactly x + 0.5; the map formula agrees to
 * rounding with a float64 restatement of OpenCV's initUndistortRectifyMap
 * loop, and with the closed-form inverse of pure radial k1 distortion.
 */
#include <cmath>
#include <cstddef>
#include <cstdint>

#include "sgm_oracle.h"

namespace {

float texel(const void *in, uint32_t fmt, uint32_t w, int x, int y) {
    const size_t i = (size_t)y * w + (size_t)x;
    if (fmt == 1) return (float)((const uint8_t *)in)[i] / 255.0f;
    return ((const float *)in)[i];
}

int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

}  // namespace

extern "C" {

void oracle_remap(const void *in, uint32_t in_fmt, uint32_t in_w, uint32_t in_h,
                  const float *map_x, const float *map_y, void *out, uint32_t out_fmt,
                  uint32_t out_w, uint32_t out_h) {
    for (uint32_t y = 0; y < out_h; ++y)
        for (uint32_t x = 0; x < out_w; ++x) {
            const size_t o = (size_t)y * out_w + x;
            const float u = std::fmin(std::fmax(map_x[o] - 0.5f, -1.0f), (float)in_w);
            const float v = std::fmin(std::fmax(map_y[o] - 0.5f, -1.0f), (float)in_h);
            const float fi = std::floor(u), fj = std::floor(v);
            const float a = u - fi, b = v - fj;
            const int i0 = (int)fi, j0 = (int)fj;
            const int xa = clampi(i0, 0, (int)in_w - 1), xb = clampi(i0 + 1, 0, (int)in_w - 1);
            const int ya = clampi(j0, 0, (int)in_h - 1), yb = clampi(j0 + 1, 0, (int)in_h - 1);
            const float t00 = texel(in, in_fmt, in_w, xa, ya);
            const float t10 = texel(in, in_fmt, in_w, xb, ya);
            const float t01 = texel(in, in_fmt, in_w, xa, yb);
            const float t11 = texel(in, in_fmt, in_w, xb, yb);
            const float top = t00 + a * (t10 - t00);
            const float bot = t01 + a * (t11 - t01);
            const float val = top + b * (bot - top);
            if (out_fmt == 1)
                ((uint8_t *)out)[o] =
                    (uint8_t)std::rint(std::fmin(std::fmax(val, 0.0f), 1.0f) * 255.0f);
            else
                ((float *)out)[o] = val;
        }
}

/* p: fx fy cx cy k1 k2 p1 p2 k3 R[9] nfx nfy ncx ncy (22 doubles).  Returns
 * 0, or 6 when Knew * R is singular. */
int oracle_rectify_maps(const double *p, uint32_t w, uint32_t h, float *map_x, float *map_y,
                        float *m_out) {
    const double *R = p + 9;
    const double nfx = p[18], nfy = p[19], ncx = p[20], ncy = p[21];
    double A[3][3];
    for (int c = 0; c < 3; ++c) {
        A[0][c] = nfx * R[c] + ncx * R[6 + c];
        A[1][c] = nfy * R[3 + c] + ncy * R[6 + c];
        A[2][c] = R[6 + c];
    }
    // Cofactor expansion along the first row; inverse = adjugate / det.
    const double c00 = A[1][1] * A[2][2] - A[1][2] * A[2][1];
    const double c01 = A[1][2] * A[2][0] - A[1][0] * A[2][2];
    const double c02 = A[1][0] * A[2][1] - A[1][1] * A[2][0];
    const double det = A[0][0] * c00 + A[0][1] * c01 + A[0][2] * c02;
    if (!(std::fabs(det) > 1e-300) || !std::isfinite(det)) return 6;
    const double adj[9] = {c00,
                           A[0][2] * A[2][1] - A[0][1] * A[2][2],
                           A[0][1] * A[1][2] - A[0][2] * A[1][1],
                           c01,
                           A[0][0] * A[2][2] - A[0][2] * A[2][0],
                           A[0][2] * A[1][0] - A[0][0] * A[1][2],
                           c02,
                           A[0][1] * A[2][0] - A[0][0] * A[2][1],
                           A[0][0] * A[1][1] - A[0][1] * A[1][0]};
    float m[9];
    for (int i = 0; i < 9; ++i) m[i] = (float)(adj[i] / det);
    if (m_out)
        for (int i = 0; i < 9; ++i) m_out[i] = m[i];
    const float fx = (float)p[0], fy = (float)p[1], cx = (float)p[2], cy = (float)p[3];
    const float k1 = (float)p[4], k2 = (float)p[5], p1 = (float)p[6], p2 = (float)p[7];
    const float k3 = (float)p[8];
    for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x) {
            const float px = (float)x, py = (float)y;
            const float X = m[0] * px + m[1] * py + m[2];
            const float Y = m[3] * px + m[4] * py + m[5];
            const float Z = m[6] * px + m[7] * py + m[8];
            const float q = 1.0f / Z;
            const float xp = X * q, yp = Y * q;
            const float x2 = xp * xp, y2 = yp * yp;
            const float r2 = x2 + y2;
            const float two_xy = 2.0f * xp * yp;
            const float kr = 1.0f + ((k3 * r2 + k2) * r2 + k1) * r2;
            const float xd = xp * kr + p1 * two_xy + p2 * (r2 + 2.0f * x2);
            const float yd = yp * kr + p1 * (r2 + 2.0f * y2) + p2 * two_xy;
            const size_t o = (size_t)y * w + x;
            map_x[o] = fx * xd + cx + 0.5f;
            map_y[o] = fy * yd + cy + 0.5f;
        }
    return 0;
}

}  // extern "C"
