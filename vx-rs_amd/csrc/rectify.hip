// Rectification kernels behind include/vxh/rectify.h: the remap of
// vx-vision's Undistorter (vx-vision/src/kernels/undistort.rs:14-137,
// vx-vision/shaders/undistort.metal:4-20) and a builder for its maps.
//
// Both are per-pixel maps, HBM-bound: the remap reads 8 B of map and
// gathers 4 texels (adjacent pixels read overlapping texels, which stay in
// L1/L2) and writes 1 or 4 B per pixel; the map builder writes 8 B per pixel
// from a handful of registers.  Blocks are 64 x 4: one wave per row segment,
// so map reads and output writes are coalesced 256-byte rows.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "sgm_pipeline.h"
#include "vxh/rectify.h"

namespace {

constexpr uint32_t kGray8 = 1, kR32F = 2;

__device__ __forceinline__ float texel(const void *in, uint32_t fmt, int w, int x, int y) {
    const size_t i = (size_t)y * w + x;
    if (fmt == kGray8) return (float)((const uint8_t *)in)[i] / 255.0f;
    return ((const float *)in)[i];
}

template <uint32_t IN, uint32_t OUT>
__global__ void __launch_bounds__(256)
remap_kernel(const void *__restrict__ in, int in_w, int in_h, const float *__restrict__ mx,
             const float *__restrict__ my, void *__restrict__ out, int out_w, int out_h) {
    const int x = blockIdx.x * 64 + threadIdx.x;
    const int y = blockIdx.y * 4 + threadIdx.y;
    if (x >= out_w || y >= out_h) return;
    const size_t o = (size_t)y * out_w + x;
    // Clamp before floor: a NaN becomes -1 (fmaxf returns the number), and
    // beyond [-1, W] both taps clamp to the same edge texel anyway.
    const float u = fminf(fmaxf(mx[o] - 0.5f, -1.0f), (float)in_w);
    const float v = fminf(fmaxf(my[o] - 0.5f, -1.0f), (float)in_h);
    const float fi = floorf(u), fj = floorf(v);
    const float a = u - fi, b = v - fj;
    const int i0 = (int)fi, j0 = (int)fj;
    const int xa = min(max(i0, 0), in_w - 1), xb = min(max(i0 + 1, 0), in_w - 1);
    const int ya = min(max(j0, 0), in_h - 1), yb = min(max(j0 + 1, 0), in_h - 1);
    const float t00 = texel(in, IN, in_w, xa, ya), t10 = texel(in, IN, in_w, xb, ya);
    const float t01 = texel(in, IN, in_w, xa, yb), t11 = texel(in, IN, in_w, xb, yb);
    const float top = t00 + a * (t10 - t00);
    const float bot = t01 + a * (t11 - t01);
    const float val = top + b * (bot - top);
    if (OUT == kGray8)
        ((uint8_t *)out)[o] = (uint8_t)rintf(fminf(fmaxf(val, 0.0f), 1.0f) * 255.0f);
    else
        ((float *)out)[o] = val;
}

struct MapCoeffs {
    float m[9];
    float fx, fy, cx, cy, k1, k2, p1, p2, k3;
};

__global__ void __launch_bounds__(256)
rectify_maps_kernel(MapCoeffs c, int w, int h, float *__restrict__ mx, float *__restrict__ my) {
    const int x = blockIdx.x * 64 + threadIdx.x;
    const int y = blockIdx.y * 4 + threadIdx.y;
    if (x >= w || y >= h) return;
    const float px = (float)x, py = (float)y;
    const float X = c.m[0] * px + c.m[1] * py + c.m[2];
    const float Y = c.m[3] * px + c.m[4] * py + c.m[5];
    const float Z = c.m[6] * px + c.m[7] * py + c.m[8];
    const float q = 1.0f / Z;
    const float xp = X * q, yp = Y * q;
    const float x2 = xp * xp, y2 = yp * yp;
    const float r2 = x2 + y2;
    const float two_xy = 2.0f * xp * yp;
    const float kr = 1.0f + ((c.k3 * r2 + c.k2) * r2 + c.k1) * r2;
    const float xd = xp * kr + c.p1 * two_xy + c.p2 * (r2 + 2.0f * x2);
    const float yd = yp * kr + c.p1 * (r2 + 2.0f * y2) + c.p2 * two_xy;
    const size_t o = (size_t)y * w + x;
    mx[o] = c.fx * xd + c.cx + 0.5f;
    my[o] = c.fy * yd + c.cy + 0.5f;
}

dim3 grid_of(uint32_t w, uint32_t h) { return dim3((w + 63) / 64, (h + 3) / 4); }

vxh_status invalid(const char *msg) {
    return vxh::report_error(VXH_ERR_INVALID_CONFIG, std::string("invalid config: ") + msg);
}

vxh_status launched(const char *what) {
    const hipError_t e = hipGetLastError();
    if (e == hipSuccess) return VXH_OK;
    return vxh::report_error(VXH_ERR_GPU, std::string("GPU error: ") + what + ": " +
                                              hipGetErrorString(e));
}

// The texture checks of the remap: gray8 or R32Float input and output,
// R32Float maps of the output's size.
vxh_status check_remap(const vxh_texture *in, const vxh_texture *mx, const vxh_texture *my,
                       const vxh_texture *out) {
    if (!in || !mx || !my || !out) return invalid("null texture");
    const uint32_t fi = vxh_texture_format(in), fo = vxh_texture_format(out);
    if ((fi != kGray8 && fi != kR32F) || (fo != kGray8 && fo != kR32F) ||
        vxh_texture_format(mx) != kR32F || vxh_texture_format(my) != kR32F)
        return invalid("remap textures: gray8/R32Float image, R32Float maps");
    const uint32_t w = vxh_texture_width(out), h = vxh_texture_height(out);
    if (vxh_texture_width(mx) != w || vxh_texture_height(mx) != h ||
        vxh_texture_width(my) != w || vxh_texture_height(my) != h)
        return vxh::report_error(VXH_ERR_TEXTURE_SIZE_MISMATCH, "texture size mismatch");
    return VXH_OK;
}

vxh_status remap_textures(const vxh_texture *in, const vxh_texture *mx, const vxh_texture *my,
                          vxh_texture *out, hipStream_t st) {
    return vxh_remap_enqueue(vxh_texture_device_ptr(in), vxh_texture_format(in),
                             vxh_texture_width(in), vxh_texture_height(in),
                             (const float *)vxh_texture_device_ptr(mx),
                             (const float *)vxh_texture_device_ptr(my),
                             vxh_texture_device_ptr(out), vxh_texture_format(out),
                             vxh_texture_width(out), vxh_texture_height(out), st);
}

}  // namespace

struct vxh_undistorter {
    int device;
    hipStream_t stream;
};

extern "C" {

vxh_status vxh_remap_enqueue(const void *d_in, uint32_t in_format, uint32_t in_w,
                             uint32_t in_h, const float *d_map_x, const float *d_map_y,
                             void *d_out, uint32_t out_format, uint32_t out_w, uint32_t out_h,
                             void *stream) {
    if (!d_in || !d_map_x || !d_map_y || !d_out) return invalid("null device buffer");
    if (in_w == 0 || in_h == 0 || out_w == 0 || out_h == 0) return invalid("empty image");
    if (in_w > (1u << 24) || in_h > (1u << 24)) return invalid("image over 2^24 pixels a side");
    const hipStream_t s = (hipStream_t)stream;
    const dim3 g = grid_of(out_w, out_h), b(64, 4);
    const int iw = (int)in_w, ih = (int)in_h, ow = (int)out_w, oh = (int)out_h;
    if (in_format == kGray8 && out_format == kGray8)
        hipLaunchKernelGGL((remap_kernel<kGray8, kGray8>), g, b, 0, s, d_in, iw, ih, d_map_x,
                           d_map_y, d_out, ow, oh);
    else if (in_format == kGray8 && out_format == kR32F)
        hipLaunchKernelGGL((remap_kernel<kGray8, kR32F>), g, b, 0, s, d_in, iw, ih, d_map_x,
                           d_map_y, d_out, ow, oh);
    else if (in_format == kR32F && out_format == kGray8)
        hipLaunchKernelGGL((remap_kernel<kR32F, kGray8>), g, b, 0, s, d_in, iw, ih, d_map_x,
                           d_map_y, d_out, ow, oh);
    else if (in_format == kR32F && out_format == kR32F)
        hipLaunchKernelGGL((remap_kernel<kR32F, kR32F>), g, b, 0, s, d_in, iw, ih, d_map_x,
                           d_map_y, d_out, ow, oh);
    else
        return invalid("remap formats: 1 (gray8) or 2 (R32Float)");
    return launched("remap");
}

void vxh_rectify_params_default(vxh_rectify_params *p) {
    if (!p) return;
    const double fx = p->fx, fy = p->fy, cx = p->cx, cy = p->cy;
    for (int i = 0; i < 9; ++i) p->R[i] = (i % 4 == 0) ? 1.0 : 0.0;
    p->new_fx = fx;
    p->new_fy = fy;
    p->new_cx = cx;
    p->new_cy = cy;
}

vxh_status vxh_rectify_inverse(const vxh_rectify_params *p, float *m9) {
    if (!p || !m9) return invalid("null argument");
    // A = Knew * R, Knew = [[nfx, 0, ncx], [0, nfy, ncy], [0, 0, 1]].
    double A[9];
    for (int c = 0; c < 3; ++c) {
        A[0 + c] = p->new_fx * p->R[0 + c] + p->new_cx * p->R[6 + c];
        A[3 + c] = p->new_fy * p->R[3 + c] + p->new_cy * p->R[6 + c];
        A[6 + c] = p->R[6 + c];
    }
    // Inverse by cofactors: inv = adj(A) / det(A).
    const double c00 = A[4] * A[8] - A[5] * A[7];
    const double c01 = A[5] * A[6] - A[3] * A[8];
    const double c02 = A[3] * A[7] - A[4] * A[6];
    const double det = A[0] * c00 + A[1] * c01 + A[2] * c02;
    if (!(std::fabs(det) > 1e-300) || !std::isfinite(det)) return invalid("singular Knew * R");
    const double inv[9] = {c00,
                           A[2] * A[7] - A[1] * A[8],
                           A[1] * A[5] - A[2] * A[4],
                           c01,
                           A[0] * A[8] - A[2] * A[6],
                           A[2] * A[3] - A[0] * A[5],
                           c02,
                           A[1] * A[6] - A[0] * A[7],
                           A[0] * A[4] - A[1] * A[3]};
    for (int i = 0; i < 9; ++i) m9[i] = (float)(inv[i] / det);
    return VXH_OK;
}

vxh_status vxh_rectify_maps_enqueue(const vxh_rectify_params *p, uint32_t w, uint32_t h,
                                    float *d_map_x, float *d_map_y, void *stream) {
    if (!p || !d_map_x || !d_map_y) return invalid("null argument");
    if (w == 0 || h == 0) return invalid("empty map");
    MapCoeffs c;
    vxh_status v = vxh_rectify_inverse(p, c.m);
    if (v != VXH_OK) return v;
    c.fx = (float)p->fx;
    c.fy = (float)p->fy;
    c.cx = (float)p->cx;
    c.cy = (float)p->cy;
    c.k1 = (float)p->k1;
    c.k2 = (float)p->k2;
    c.p1 = (float)p->p1;
    c.p2 = (float)p->p2;
    c.k3 = (float)p->k3;
    hipLaunchKernelGGL(rectify_maps_kernel, grid_of(w, h), dim3(64, 4), 0, (hipStream_t)stream,
                       c, (int)w, (int)h, d_map_x, d_map_y);
    return launched("rectify_maps");
}

vxh_status vxh_rectify_maps(vxh_context *ctx, const vxh_rectify_params *p, vxh_texture *map_x,
                            vxh_texture *map_y) {
    if (!ctx || !map_x || !map_y) return invalid("null argument");
    if (vxh_texture_format(map_x) != kR32F || vxh_texture_format(map_y) != kR32F)
        return invalid("maps must be R32Float");
    const uint32_t w = vxh_texture_width(map_x), h = vxh_texture_height(map_x);
    if (vxh_texture_width(map_y) != w || vxh_texture_height(map_y) != h)
        return vxh::report_error(VXH_ERR_TEXTURE_SIZE_MISMATCH, "texture size mismatch");
    vxh::CaptureRelaxed relaxed;
    if (hipSetDevice(vxh_context_device(ctx)) != hipSuccess)
        return vxh::report_error(VXH_ERR_DEVICE_NOT_FOUND, "no HIP device found");
    hipStream_t s = nullptr;
    vxh_status v = vxh::ctx_take_stream(ctx, &s);
    if (v != VXH_OK) return v;
    v = vxh_rectify_maps_enqueue(p, w, h, (float *)vxh_texture_device_ptr(map_x),
                                 (float *)vxh_texture_device_ptr(map_y), s);
    const hipError_t e = hipStreamSynchronize(s);
    vxh::ctx_give_stream(ctx, s);
    if (v != VXH_OK) return v;
    if (e != hipSuccess)
        return vxh::report_error(VXH_ERR_GPU, std::string("GPU error: rectify maps: ") +
                                                  hipGetErrorString(e));
    return VXH_OK;
}

vxh_status vxh_undistorter_create(vxh_context *ctx, vxh_undistorter **out) {
    if (!ctx || !out) return invalid("null argument");
    *out = nullptr;
    const int dev = vxh_context_device(ctx);
    if (hipSetDevice(dev) != hipSuccess)
        return vxh::report_error(VXH_ERR_DEVICE_NOT_FOUND, "no HIP device found");
    hipStream_t s = nullptr;
    const hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e != hipSuccess)
        return vxh::report_error(VXH_ERR_GPU, std::string("GPU error: stream: ") +
                                                  hipGetErrorString(e));
    *out = new vxh_undistorter{dev, s};
    return VXH_OK;
}

void vxh_undistorter_destroy(vxh_undistorter *u) {
    if (!u) return;
    vxh::CaptureRelaxed relaxed;
    (void)hipSetDevice(u->device);
    (void)hipStreamSynchronize(u->stream);
    (void)hipStreamDestroy(u->stream);
    delete u;
}

vxh_status vxh_undistorter_apply(vxh_undistorter *u, vxh_context *ctx, const vxh_texture *input,
                                 const vxh_texture *map_x, const vxh_texture *map_y,
                                 vxh_texture *output) {
    if (!u || !ctx) return invalid("null handle");
    vxh_status v = check_remap(input, map_x, map_y, output);
    if (v != VXH_OK) return v;
    vxh::CaptureRelaxed relaxed;
    (void)hipSetDevice(u->device);
    if ((v = remap_textures(input, map_x, map_y, output, u->stream)) != VXH_OK) return v;
    const hipError_t e = hipStreamSynchronize(u->stream);
    if (e != hipSuccess)
        return vxh::report_error(VXH_ERR_GPU, std::string("GPU error: undistort: ") +
                                                  hipGetErrorString(e));
    return VXH_OK;
}

vxh_status vxh_undistorter_encode(vxh_undistorter *u, vxh_pipeline *p, const vxh_texture *input,
                                  const vxh_texture *map_x, const vxh_texture *map_y,
                                  vxh_texture *output) {
    if (!u || !p) return invalid("null handle");
    vxh_status v = check_remap(input, map_x, map_y, output);
    if (v != VXH_OK) return v;
    hipStream_t s = nullptr;
    if ((v = vxh::pipeline_stream(p, &s)) != VXH_OK) return v;
    for (const vxh_texture *t : {input, map_x, map_y, (const vxh_texture *)output})
        vxh::pipeline_use(p, t);
    return remap_textures(input, map_x, map_y, output, s);
}

}  // extern "C"
