// Per-pair SGM kernels for gfx950: thin grid wrappers around the device
// functions in sgm_device.h, and their host-side launchers.
#include "sgm_device.h"

namespace vxh {
namespace {

using namespace dev;

__global__ void __launch_bounds__(256)
census_kernel(const uint8_t *__restrict__ img, int w, int h, int rx, int ry,
              uint2 *__restrict__ out) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= w || y >= h) return;
    out[(size_t)y * w + x] = census_pixel(img, w, h, rx, ry, x, y);
}

template <int L, int K>
__global__ void __launch_bounds__(256)
aggregate_kernel(AggParams p, int dx, int dy) {
    const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    agg_wave<L, K>(p, dx, dy, wave * (64 / L));
}

__global__ void __launch_bounds__(256)
wta_kernel(const uint16_t *__restrict__ vol, int w, int h, int D, int Dp,
           float *__restrict__ out, int16_t *__restrict__ best) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)w * h) return;
    wta_pixel(vol, D, Dp, i, out, best);
}

__global__ void __launch_bounds__(256)
lr_kernel(const uint16_t *__restrict__ vol, int w, int h, int D, int Dp, int tol,
          const int16_t *__restrict__ best, float *__restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)w * h) return;
    lr_pixel(vol, w, D, Dp, tol, i, best, out);
}

__global__ void __launch_bounds__(256)
copy_kernel(uint4 *__restrict__ dst, const uint4 *__restrict__ src, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

template <int L, int K>
hipError_t launch_agg_lk(const AggParams &p, hipStream_t s) {
    constexpr int G = 64 / L;
    const int dx = kDirX[p.dir], dy = kDirY[p.dir];
    const long waves = (p.nscan + G - 1) / G;
    const long threads = waves * 64;
    const int block = 256;
    const unsigned grid = (unsigned)((threads + block - 1) / block);
    hipLaunchKernelGGL((aggregate_kernel<L, K>), dim3(grid), dim3(block), 0, s,
                       p, dx, dy);
    return hipGetLastError();
}

}  // namespace

Layout default_layout(int D) {
    if (D <= 64) return {64, 1};
    if (D <= 128) return {64, 2};
    return {64, 4};
}

// Instantiated lane layouts (L lanes per scanline, K disparities per lane).
static const Layout kLayouts[] = {{64, 1}, {64, 2}, {64, 4}, {32, 1}, {32, 2},
                                  {32, 4}, {32, 8}, {16, 4}, {16, 8}, {16, 16}};

bool layout_supported(Layout l, int D) {
    for (const Layout &k : kLayouts)
        if (k.L == l.L && k.K == l.K) return l.L * l.K >= D && D >= 1;
    return false;
}

hipError_t launch_census(const uint8_t *img, int w, int h, int rx, int ry,
                         uint2 *out, hipStream_t s) {
    dim3 block(64, 4);
    dim3 grid((w + 63) / 64, (h + 3) / 4);
    hipLaunchKernelGGL(census_kernel, grid, block, 0, s, img, w, h, rx, ry, out);
    return hipGetLastError();
}

hipError_t launch_aggregate(const AggParams &p, Layout l, hipStream_t s) {
    if (!layout_supported(l, p.D)) return hipErrorInvalidValue;
#define VXH_AGG_CASE(LL, KK) \
    if (l.L == LL && l.K == KK) return launch_agg_lk<LL, KK>(p, s);
    VXH_AGG_CASE(64, 1) VXH_AGG_CASE(64, 2) VXH_AGG_CASE(64, 4)
    VXH_AGG_CASE(32, 1) VXH_AGG_CASE(32, 2) VXH_AGG_CASE(32, 4) VXH_AGG_CASE(32, 8)
    VXH_AGG_CASE(16, 4) VXH_AGG_CASE(16, 8) VXH_AGG_CASE(16, 16)
#undef VXH_AGG_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_copy(void *dst, const void *src, size_t bytes, hipStream_t s) {
    const size_t n = bytes / 16;
    const unsigned grid = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(copy_kernel, dim3(grid ? grid : 1), dim3(256), 0, s, (uint4 *)dst,
                       (const uint4 *)src, n);
    return hipGetLastError();
}

hipError_t launch_wta(const uint16_t *vol, int w, int h, int D, int Dp, float *out,
                      int16_t *best, hipStream_t s) {
    const size_t n = (size_t)w * h;
    hipLaunchKernelGGL(wta_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256),
                       0, s, vol, w, h, D, Dp, out, best);
    return hipGetLastError();
}

hipError_t launch_lr_check(const uint16_t *vol, int w, int h, int D, int Dp, int tol,
                           const int16_t *best, float *out, hipStream_t s) {
    const size_t n = (size_t)w * h;
    hipLaunchKernelGGL(lr_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256),
                       0, s, vol, w, h, D, Dp, tol, best, out);
    return hipGetLastError();
}

}  // namespace vxh
