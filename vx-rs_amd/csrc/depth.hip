// Depth post-processing behind include/vxh/depth.h: the bilateral and median
// filters of vx-vision's DepthFilter (vx-vision/shaders/DepthFilter.metal:
// 17-120, vx-vision/src/kernels/depth_filter.rs:22-319) and the
// disparity-to-depth conversion the reference leaves to its caller
// (vx-vision/src/kernels/sgm.rs:284-286).
//
// All three are per-pixel maps over w*h fp32 values; each thread owns one
// pixel (blocks of 64 x 4: one wave per row segment, so every read of a
// window row is coalesced and the window's rows stay in L1).  The median and
// the bilateral filter above radius 7 read their clamp-to-edge window straight
// from global memory; the bilateral filter up to radius 7 stages it in LDS.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "sgm_pipeline.h"
#include "vxh/depth.h"

namespace {

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// DepthFilter.metal:17-60 in the shader's fp32 expression order (loops dy
// outer, dx inner; weights exp(-(dx^2+dy^2) * inv_2ss) * exp(-(n-c)^2 *
// inv_2sd); neighbours <= 0 skipped; a centre <= 0 writes 0).
__global__ void __launch_bounds__(256)
bilateral_kernel(const float *__restrict__ in, float *__restrict__ out, int w, int h,
                 int radius, float inv_2ss, float inv_2sd) {
    const int x = blockIdx.x * 64 + (int)threadIdx.x;
    const int y = blockIdx.y * 4 + (int)threadIdx.y;
    if (x >= w || y >= h) return;
    const float center = in[(size_t)y * w + x];
    if (center <= 0.0f) {
        out[(size_t)y * w + x] = 0.0f;
        return;
    }
    float sum = 0.0f, weight_sum = 0.0f;
    for (int dy = -radius; dy <= radius; ++dy) {
        const float *row = in + (size_t)clampi(y + dy, 0, h - 1) * w;
        for (int dx = -radius; dx <= radius; ++dx) {
            const float n = row[clampi(x + dx, 0, w - 1)];
            if (n <= 0.0f) continue;
            const float sd2 = (float)(dx * dx + dy * dy);
            const float ws = expf(-sd2 * inv_2ss);
            const float diff = n - center;
            const float wd = expf(-diff * diff * inv_2sd);
            const float wgt = ws * wd;
            sum += n * wgt;
            weight_sum += wgt;
        }
    }
    out[(size_t)y * w + x] = weight_sum > 1e-8f ? sum / weight_sum : center;
}

// bilateral_kernel for radius R <= 7 (the default is 3): the block's
// (4 + 2R) x (64 + 2R) window tile, clamp-to-edge, is read once into LDS, and
// the spatial weights exp(-(dx^2+dy^2) * inv_2ss), which depend only on the
// offset, are computed once per block into an LDS table with the same
// expression.  Each pixel then reads its taps from LDS at constant offsets and
// evaluates one exp per tap, accumulating in the same order: the results are
// bit-identical to bilateral_kernel's.
template <int R>
__global__ void __launch_bounds__(256)
bilateral_tile_kernel(const float *__restrict__ in, float *__restrict__ out, int w, int h,
                      float inv_2ss, float inv_2sd) {
    constexpr int S = 2 * R + 1, TW = 64 + 2 * R, TH = 4 + 2 * R;
    __shared__ float tile[TH * TW];
    __shared__ float spatial[S * S];
    const int tx = (int)threadIdx.x, ty = (int)threadIdx.y, t = ty * 64 + tx;
    const int x0 = (int)blockIdx.x * 64 - R, y0 = (int)blockIdx.y * 4 - R;
    for (int i = t; i < TH * TW; i += 256) {
        const int r = i / TW, c = i - r * TW;
        tile[i] = in[(size_t)clampi(y0 + r, 0, h - 1) * w + clampi(x0 + c, 0, w - 1)];
    }
    for (int i = t; i < S * S; i += 256) {
        const int dy = i / S - R, dx = i % S - R;
        const float sd2 = (float)(dx * dx + dy * dy);
        spatial[i] = expf(-sd2 * inv_2ss);
    }
    __syncthreads();
    const int x = (int)blockIdx.x * 64 + tx, y = (int)blockIdx.y * 4 + ty;
    if (x >= w || y >= h) return;
    const float *win = tile + ty * TW + tx;   // the window's top-left tap
    const float center = win[R * TW + R];
    if (center <= 0.0f) {
        out[(size_t)y * w + x] = 0.0f;
        return;
    }
    float sum = 0.0f, weight_sum = 0.0f;
#pragma unroll
    for (int dy = 0; dy < S; ++dy) {
#pragma unroll
        for (int dx = 0; dx < S; ++dx) {
            const float n = win[dy * TW + dx];
            if (n <= 0.0f) continue;
            const float ws = spatial[dy * S + dx];
            const float diff = n - center;
            const float wd = expf(-diff * diff * inv_2sd);
            const float wgt = ws * wd;
            sum += n * wgt;
            weight_sum += wgt;
        }
    }
    out[(size_t)y * w + x] = weight_sum > 1e-8f ? sum / weight_sum : center;
}

// The median of DepthFilter.metal:66-120: the values > 0 of the window (the
// first 121 in scan order), sorted ascending, the one at index count / 2; 0
// when there is none.  Radius R <= 2 (at most 25 values, so the cap never
// binds): the window lives in registers, an invalid entry as +inf, which
// sorts after every valid value; a sorting network orders the window and the
// entry at count / 2 is the answer.  Larger radii (median_kernel) replace the
// sort by rank selection: v is the answer iff #(values < v) <= count / 2 <
// #(values <= v), which picks the same value.
template <int R>
__global__ void __launch_bounds__(256)
median_small_kernel(const float *__restrict__ in, float *__restrict__ out, int w, int h) {
    constexpr int S = 2 * R + 1, N = S * S;
    const int x = blockIdx.x * 64 + (int)threadIdx.x;
    const int y = blockIdx.y * 4 + (int)threadIdx.y;
    if (x >= w || y >= h) return;
    float v[N];
    int count = 0;
#pragma unroll
    for (int dy = 0; dy < S; ++dy) {
        const float *row = in + (size_t)clampi(y + dy - R, 0, h - 1) * w;
#pragma unroll
        for (int dx = 0; dx < S; ++dx) {
            const float d = row[clampi(x + dx - R, 0, w - 1)];
            const bool ok = d > 0.0f;
            v[dy * S + dx] = ok ? d : INFINITY;
            count += ok ? 1 : 0;
        }
    }
    // Odd-even transposition sort (N rounds sort N values; every value is
    // positive or +inf, so min / max are exact), then the entry at count / 2.
#pragma unroll
    for (int r = 0; r < N; ++r) {
#pragma unroll
        for (int i = r & 1; i + 1 < N; i += 2) {
            const float a = v[i], b = v[i + 1];
            v[i] = fminf(a, b);
            v[i + 1] = fmaxf(a, b);
        }
    }
    const int k = count / 2;
    float res = v[0];
#pragma unroll
    for (int i = 1; i <= N / 2; ++i) res = k == i ? v[i] : res;
    out[(size_t)y * w + x] = count > 0 ? res : 0.0f;
}

// Any radius: the window is scanned in the shader's order to find where the
// 121st value > 0 falls (later values are ignored, as the shader's fixed
// array does), then rank selection over that prefix reading the window again
// (from L1).
constexpr int kMedianCap = 121;

__global__ void __launch_bounds__(256)
median_kernel(const float *__restrict__ in, float *__restrict__ out, int w, int h, int radius) {
    const int x = blockIdx.x * 64 + (int)threadIdx.x;
    const int y = blockIdx.y * 4 + (int)threadIdx.y;
    if (x >= w || y >= h) return;
    const int S = 2 * radius + 1;
    auto at = [&](int p) {
        const int dy = p / S - radius, dx = p % S - radius;
        return in[(size_t)clampi(y + dy, 0, h - 1) * w + clampi(x + dx, 0, w - 1)];
    };
    const int total = radius >= 0 ? S * S : 0;
    int count = 0, stop = total;
    for (int p = 0; p < total; ++p)
        if (at(p) > 0.0f && ++count == kMedianCap) {
            stop = p + 1;
            break;
        }
    float res = 0.0f;
    if (count > 0) {
        const int k = count / 2;
        for (int i = 0; i < stop; ++i) {
            const float vi = at(i);
            if (!(vi > 0.0f)) continue;
            int lt = 0, le = 0;
            for (int j = 0; j < stop; ++j) {
                const float vj = at(j);
                if (!(vj > 0.0f)) continue;
                lt += vj < vi ? 1 : 0;
                le += vj <= vi ? 1 : 0;
            }
            if (lt <= k && k < le) {
                res = vi;
                break;
            }
        }
    }
    out[(size_t)y * w + x] = res;
}

__global__ void __launch_bounds__(256)
to_depth_kernel(const float *disp, float *depth, size_t n, float fb) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float d = disp[i];
    depth[i] = d > 0.0f ? fb / d : 0.0f;
}

dim3 grid_of(uint32_t w, uint32_t h) { return dim3((w + 63) / 64, (h + 3) / 4); }

vxh_status launched(const char *what) {
    const hipError_t e = hipGetLastError();
    if (e == hipSuccess) return VXH_OK;
    return vxh::report_error(VXH_ERR_GPU, std::string("GPU error: ") + what + ": " +
                                              hipGetErrorString(e));
}

vxh_status check_map(const void *a, const void *b, uint32_t w, uint32_t h) {
    if (!a || !b || w == 0 || h == 0)
        return vxh::report_error(VXH_ERR_INVALID_CONFIG, "invalid config: depth map arguments");
    return VXH_OK;
}

// Window radius bound of both filters (a 511 x 511 window per pixel).
constexpr int kMaxRadius = 255;

vxh_status check_radius(int r) {
    if (r > kMaxRadius)
        return vxh::report_error(VXH_ERR_INVALID_CONFIG,
                                 "invalid config: filter radius above 255");
    return VXH_OK;
}

// Input and output textures of a filter call: R32Float, same size.
vxh_status check_textures(const vxh_texture *in, const vxh_texture *out) {
    if (!in || !out || vxh_texture_format(in) != 2 || vxh_texture_format(out) != 2)
        return vxh::report_error(VXH_ERR_INVALID_CONFIG,
                                 "invalid config: depth filters need R32Float textures");
    if (vxh_texture_width(in) != vxh_texture_width(out) ||
        vxh_texture_height(in) != vxh_texture_height(out))
        return vxh::report_error(VXH_ERR_TEXTURE_SIZE_MISMATCH, "texture size mismatch");
    return VXH_OK;
}

}  // namespace

struct vxh_depth_filter {
    int device;
    hipStream_t stream;
};

extern "C" {

void vxh_depth_filter_config_default(vxh_depth_filter_config *cfg) {
    if (!cfg) return;
    cfg->radius = 3;
    cfg->sigma_spatial = 5.0f;
    cfg->sigma_depth = 0.05f;
}

void vxh_depth_median_config_default(vxh_depth_median_config *cfg) {
    if (cfg) cfg->radius = 1;
}

vxh_status vxh_depth_bilateral_enqueue(const float *d_in, float *d_out, uint32_t w, uint32_t h,
                                       const vxh_depth_filter_config *cfg, void *stream) {
    vxh_status st = check_map(d_in, d_out, w, h);
    if (st != VXH_OK) return st;
    if (!cfg) return vxh::report_error(VXH_ERR_INVALID_CONFIG, "invalid config: null config");
    if ((st = check_radius(cfg->radius)) != VXH_OK) return st;
    // The shader's per-thread constants, in fp32 (DepthFilter.metal:33-34).
    const float inv_2ss = 1.0f / (2.0f * cfg->sigma_spatial * cfg->sigma_spatial);
    const float inv_2sd = 1.0f / (2.0f * cfg->sigma_depth * cfg->sigma_depth);
    const hipStream_t s = (hipStream_t)stream;
    switch (cfg->radius) {
#define VXH_BIL(RR)                                                                        \
    case RR:                                                                               \
        hipLaunchKernelGGL(bilateral_tile_kernel<RR>, grid_of(w, h), dim3(64, 4), 0, s, d_in, \
                           d_out, (int)w, (int)h, inv_2ss, inv_2sd);                       \
        break;
        VXH_BIL(1) VXH_BIL(2) VXH_BIL(3) VXH_BIL(4) VXH_BIL(5) VXH_BIL(6) VXH_BIL(7)
#undef VXH_BIL
    default:
        hipLaunchKernelGGL(bilateral_kernel, grid_of(w, h), dim3(64, 4), 0, s, d_in, d_out,
                           (int)w, (int)h, (int)cfg->radius, inv_2ss, inv_2sd);
    }
    return launched("depth_bilateral");
}

vxh_status vxh_depth_median_enqueue(const float *d_in, float *d_out, uint32_t w, uint32_t h,
                                    const vxh_depth_median_config *cfg, void *stream) {
    vxh_status st = check_map(d_in, d_out, w, h);
    if (st != VXH_OK) return st;
    if (!cfg) return vxh::report_error(VXH_ERR_INVALID_CONFIG, "invalid config: null config");
    if ((st = check_radius(cfg->radius)) != VXH_OK) return st;
    const hipStream_t s = (hipStream_t)stream;
    if (cfg->radius == 1)
        hipLaunchKernelGGL(median_small_kernel<1>, grid_of(w, h), dim3(64, 4), 0, s, d_in,
                           d_out, (int)w, (int)h);
    else if (cfg->radius == 2)
        hipLaunchKernelGGL(median_small_kernel<2>, grid_of(w, h), dim3(64, 4), 0, s, d_in,
                           d_out, (int)w, (int)h);
    else
        hipLaunchKernelGGL(median_kernel, grid_of(w, h), dim3(64, 4), 0, s, d_in, d_out,
                           (int)w, (int)h, (int)cfg->radius);
    return launched("depth_median");
}

vxh_status vxh_disparity_to_depth(const float *d_disp, float *d_depth, uint32_t w, uint32_t h,
                                  float fx, float baseline, void *stream) {
    vxh_status st = check_map(d_disp, d_depth, w, h);
    if (st != VXH_OK) return st;
    const size_t n = (size_t)w * h;
    hipLaunchKernelGGL(to_depth_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, d_disp, d_depth, n, fx * baseline);
    return launched("disparity_to_depth");
}

vxh_status vxh_depth_filter_create(vxh_context *ctx, vxh_depth_filter **out) {
    if (!ctx || !out) return vxh::report_error(VXH_ERR_INVALID_CONFIG, "invalid config: null argument");
    *out = nullptr;
    const int dev = vxh_context_device(ctx);
    if (hipSetDevice(dev) != hipSuccess)
        return vxh::report_error(VXH_ERR_DEVICE_NOT_FOUND, "no HIP device found");
    hipStream_t s = nullptr;
    const hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e != hipSuccess)
        return vxh::report_error(VXH_ERR_GPU, std::string("GPU error: stream: ") + hipGetErrorString(e));
    *out = new vxh_depth_filter{dev, s};
    return VXH_OK;
}

void vxh_depth_filter_destroy(vxh_depth_filter *f) {
    if (!f) return;
    (void)hipSetDevice(f->device);
    (void)hipStreamSynchronize(f->stream);
    (void)hipStreamDestroy(f->stream);
    delete f;
}

static vxh_status finish(vxh_depth_filter *f, vxh_status st) {
    if (st != VXH_OK) return st;
    const hipError_t e = hipStreamSynchronize(f->stream);
    if (e != hipSuccess)
        return vxh::report_error(VXH_ERR_GPU, std::string("GPU error: depth filter: ") +
                                                  hipGetErrorString(e));
    return VXH_OK;
}

vxh_status vxh_depth_filter_apply_bilateral(vxh_depth_filter *f, vxh_context *ctx,
                                            const vxh_texture *input, vxh_texture *output,
                                            const vxh_depth_filter_config *cfg) {
    if (!f || !ctx) return vxh::report_error(VXH_ERR_INVALID_CONFIG, "invalid config: null handle");
    vxh_status st = check_textures(input, output);
    if (st != VXH_OK) return st;
    (void)hipSetDevice(f->device);
    return finish(f, vxh_depth_bilateral_enqueue(
                         (const float *)vxh_texture_device_ptr(input),
                         (float *)vxh_texture_device_ptr(output), vxh_texture_width(input),
                         vxh_texture_height(input), cfg, f->stream));
}

vxh_status vxh_depth_filter_apply_median(vxh_depth_filter *f, vxh_context *ctx,
                                         const vxh_texture *input, vxh_texture *output,
                                         const vxh_depth_median_config *cfg) {
    if (!f || !ctx) return vxh::report_error(VXH_ERR_INVALID_CONFIG, "invalid config: null handle");
    vxh_status st = check_textures(input, output);
    if (st != VXH_OK) return st;
    (void)hipSetDevice(f->device);
    return finish(f, vxh_depth_median_enqueue((const float *)vxh_texture_device_ptr(input),
                                              (float *)vxh_texture_device_ptr(output),
                                              vxh_texture_width(input),
                                              vxh_texture_height(input), cfg, f->stream));
}

vxh_status vxh_depth_filter_encode_bilateral(vxh_depth_filter *f, vxh_pipeline *p,
                                             const vxh_texture *input, vxh_texture *output,
                                             const vxh_depth_filter_config *cfg) {
    if (!f || !p) return vxh::report_error(VXH_ERR_INVALID_CONFIG, "invalid config: null handle");
    vxh_status st = check_textures(input, output);
    if (st != VXH_OK) return st;
    hipStream_t s = nullptr;
    if ((st = vxh::pipeline_stream(p, &s)) != VXH_OK) return st;
    vxh::pipeline_use(p, input);
    vxh::pipeline_use(p, output);
    return vxh_depth_bilateral_enqueue((const float *)vxh_texture_device_ptr(input),
                                       (float *)vxh_texture_device_ptr(output),
                                       vxh_texture_width(input), vxh_texture_height(input), cfg,
                                       s);
}

vxh_status vxh_depth_filter_encode_median(vxh_depth_filter *f, vxh_pipeline *p,
                                          const vxh_texture *input, vxh_texture *output,
                                          const vxh_depth_median_config *cfg) {
    if (!f || !p) return vxh::report_error(VXH_ERR_INVALID_CONFIG, "invalid config: null handle");
    vxh_status st = check_textures(input, output);
    if (st != VXH_OK) return st;
    hipStream_t s = nullptr;
    if ((st = vxh::pipeline_stream(p, &s)) != VXH_OK) return st;
    vxh::pipeline_use(p, input);
    vxh::pipeline_use(p, output);
    return vxh_depth_median_enqueue((const float *)vxh_texture_device_ptr(input),
                                    (float *)vxh_texture_device_ptr(output),
                                    vxh_texture_width(input), vxh_texture_height(input), cfg, s);
}

}  // extern "C"
