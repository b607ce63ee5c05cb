// Per-pair SGM kernels for gfx950: thin grid wrappers around the device
// functions in sgm_device.h, and their host-side launchers.
#include <algorithm>

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

// Field `f` of pass i (i wave-uniform): unrolled selects, so the kernel
// argument block is read with scalar loads instead of being copied to private
// memory for a dynamic index.
#define VXH_PASS_FIELD(m, i, f)                                   \
    ({                                                            \
        auto v_ = (m).pass[0].f;                                  \
        _Pragma("unroll") for (int j_ = 1; j_ < kMaxPasses; ++j_) \
            if (j_ == (i)) v_ = (m).pass[j_].f;                   \
        v_;                                                       \
    })

// m.npass independent direction passes in one launch.  Block b belongs to the
// pass whose [first_block, first_block + nblocks) holds it; inside the pass the
// block index is permuted so that the blocks the hardware places on one XCD
// (b = x mod 8) take a contiguous range of scanlines, whose census rows then
// stay in that XCD's L2.
template <int L, int K, int ACC>
__global__ void __launch_bounds__(256)
aggregate_kernel(MultiAgg m) {
    const int b = blockIdx.x;
    int i = 0;
#pragma unroll
    for (int j = 1; j < kMaxPasses; ++j)
        if (j < m.npass && b >= m.pass[j].first_block) i = j;
    const int local = b - VXH_PASS_FIELD(m, i, first_block);
    const int per = VXH_PASS_FIELD(m, i, nblocks) >> 3;
    const int lb = (local & 7) * per + (local >> 3);
    AggParams p = m.base;
    p.vol = VXH_PASS_FIELD(m, i, vol);
    p.dir = VXH_PASS_FIELD(m, i, dir);
    p.nscan = VXH_PASS_FIELD(m, i, nscan);
    p.accumulate = ACC;
    const int wave = lb * 4 + (int)(threadIdx.x >> 6);
    agg_wave<L, K, ACC>(p, VXH_PASS_FIELD(m, i, dx), VXH_PASS_FIELD(m, i, dy),
                        wave * (64 / L));
}

// Sum of `NV` volumes and winner-take-all with parabola refinement
// (SGMStereo.metal:179-214), 16 lanes per pixel: in chunk c, lane j owns
// entries 128c + 8j .. 128c + 8j + 7 of the pixel's slice and reads them with
// one 16-byte load per volume, so a pixel's slice is one contiguous 256-byte
// access.  The first strict minimum from 0xFFFF is the smallest key
// (value << 16 | d) over d < D; a slice of all 0xFFFF keeps d = 0.
template <int NV, int NC>
__global__ void __launch_bounds__(256)
wta_sum_kernel(const uint16_t *__restrict__ vol0, size_t vstride, int npx, int D, int Dp,
               float *__restrict__ out, int16_t *__restrict__ best,
               uint16_t *__restrict__ sum_out) {
    const int j = (int)(threadIdx.x & 15);
    const int pix = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 4);
    const bool valid = pix < npx;
    const size_t base = (size_t)(valid ? pix : 0) * (size_t)Dp;
    u16x2 v[NC][4];
    uint32_t key = 0xFFFFFFFFu;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        const int d0 = 128 * c + 8 * j;
#pragma unroll
        for (int q = 0; q < 4; ++q) v[c][q] = (u16x2){0, 0};
        if (d0 < Dp) {
#pragma unroll
            for (int k = 0; k < NV; ++k) {
                const uint4 x = *(const uint4 *)(vol0 + k * vstride + base + d0);
                v[c][0] += __builtin_bit_cast(u16x2, x.x);
                v[c][1] += __builtin_bit_cast(u16x2, x.y);
                v[c][2] += __builtin_bit_cast(u16x2, x.z);
                v[c][3] += __builtin_bit_cast(u16x2, x.w);
            }
            if (sum_out != nullptr && valid)
                *(uint4 *)(sum_out + base + d0) = make_uint4(
                    __builtin_bit_cast(uint32_t, v[c][0]), __builtin_bit_cast(uint32_t, v[c][1]),
                    __builtin_bit_cast(uint32_t, v[c][2]), __builtin_bit_cast(uint32_t, v[c][3]));
        }
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const uint32_t val = v[c][e / 2][e & 1];
            if (d0 + e < D) key = min(key, (val << 16) | (uint32_t)(d0 + e));
        }
    }
    key = group_min<16>(key);
    const int bd = (int)(key & 0xFFFFu);
    const uint32_t bc = key >> 16;
    uint32_t nb2 = 0;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        const int d0 = 128 * c + 8 * j;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const uint32_t val = v[c][e / 2][e & 1];
            if (d0 + e == bd - 1) nb2 |= val << 16;
            if (d0 + e == bd + 1) nb2 |= val;
        }
    }
    nb2 = group_or<16>(nb2);
    if (valid && j == 0) {
        const int bdw = bc == 0xFFFFu ? 0 : bd;
        float disp = (float)bdw;
        if (bdw > 0 && bdw < D - 1) {
            const float c0 = (float)(nb2 >> 16);
            const float c1 = (float)bc;
            const float c2 = (float)(nb2 & 0xFFFFu);
            const float den = 2.0f * (c0 + c2 - 2.0f * c1);
            if (fabsf(den) > 1e-6f) disp = disp + (c0 - c2) / den;
        }
        out[pix] = disp;
        if (best) best[pix] = (int16_t)bdw;
    }
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

// Scanline length of a direction: the passes with the longest scanlines are
// placed first in a launch so that their waves start first.
int scan_length(int dir, int w, int h) {
    if (kDirY[dir] == 0) return w;
    if (kDirX[dir] == 0) return h;
    return std::min(w, h);
}

template <int L, int K>
hipError_t launch_multi_lk(MultiAgg &m, hipStream_t s) {
    constexpr int G = 64 / L;
    int blocks = 0;
    for (int j = 0; j < m.npass; ++j) {
        const int waves = (m.pass[j].nscan + G - 1) / G;
        const int nb = (((waves + 3) / 4) + 7) & ~7;
        m.pass[j].first_block = blocks;
        m.pass[j].nblocks = nb;
        blocks += nb;
    }
    for (int j = m.npass; j < kMaxPasses; ++j) m.pass[j] = m.pass[0];
    // Every pass of one launch either stores or accumulates.
    if (m.pass[0].accumulate)
        hipLaunchKernelGGL((aggregate_kernel<L, K, 1>), dim3(blocks), dim3(256), 0, s, m);
    else
        hipLaunchKernelGGL((aggregate_kernel<L, K, 0>), dim3(blocks), dim3(256), 0, s, m);
    return hipGetLastError();
}

hipError_t launch_multi(MultiAgg &m, Layout l, hipStream_t s) {
#define VXH_AGG_CASE(LL, KK) \
    if (l.L == LL && l.K == KK) return launch_multi_lk<LL, KK>(m, s);
    VXH_AGG_CASE(64, 1) VXH_AGG_CASE(64, 2) VXH_AGG_CASE(64, 4)
    VXH_AGG_CASE(32, 1) VXH_AGG_CASE(32, 2) VXH_AGG_CASE(32, 4) VXH_AGG_CASE(32, 8)
    VXH_AGG_CASE(16, 4) VXH_AGG_CASE(16, 8) VXH_AGG_CASE(16, 16)
#undef VXH_AGG_CASE
    return hipErrorInvalidValue;
}

PassDesc make_pass(const AggParams &p, int dir, uint16_t *vol, int accumulate) {
    PassDesc d{};
    d.vol = vol;
    d.dir = dir;
    d.dx = kDirX[dir];
    d.dy = kDirY[dir];
    d.nscan = num_scanlines(dir, p.w, p.h);
    d.accumulate = accumulate;
    return d;
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
    if (!layout_supported(l, p.D) || p.dir < 0 || p.dir > 7) return hipErrorInvalidValue;
    MultiAgg m{};
    m.base = p;
    m.npass = 1;
    m.pass[0] = make_pass(p, p.dir, p.vol, p.accumulate);
    return launch_multi(m, l, s);
}

hipError_t launch_aggregate_multi(const AggParams &base, const int *dirs,
                                  uint16_t *const *vols, const int *accumulate, int n,
                                  Layout l, hipStream_t s) {
    if (!layout_supported(l, base.D) || n < 1 || n > kMaxPasses) return hipErrorInvalidValue;
    MultiAgg m{};
    m.base = base;
    m.base.wta_out = nullptr;
    m.npass = n;
    int order[kMaxPasses];
    for (int j = 0; j < n; ++j) {
        if (dirs[j] < 0 || dirs[j] > 7 || (accumulate[j] != 0) != (accumulate[0] != 0))
            return hipErrorInvalidValue;
        order[j] = j;
    }
    std::stable_sort(order, order + n, [&](int a, int b) {
        return scan_length(dirs[a], base.w, base.h) > scan_length(dirs[b], base.w, base.h);
    });
    for (int j = 0; j < n; ++j) {
        const int k = order[j];
        m.pass[j] = make_pass(base, dirs[k], vols[k], accumulate[k]);
    }
    return launch_multi(m, l, s);
}

hipError_t launch_wta_sum(const uint16_t *vol0, size_t vstride, int nvol, int npx, int D,
                          int Dp, float *out, int16_t *best, uint16_t *sum_out,
                          hipStream_t s) {
    if (Dp > 256 || Dp % 16 || npx < 1) return hipErrorInvalidValue;
    const unsigned grid = (unsigned)(((size_t)npx * 16 + 255) / 256);
#define VXH_WTA_CASE(NV, NC)                                                              \
    if (nvol == NV && (Dp > 128) == (NC == 2)) {                                         \
        hipLaunchKernelGGL((wta_sum_kernel<NV, NC>), dim3(grid), dim3(256), 0, s, vol0,  \
                           vstride, npx, D, Dp, out, best, sum_out);                     \
        return hipGetLastError();                                                         \
    }
    VXH_WTA_CASE(1, 1) VXH_WTA_CASE(1, 2) VXH_WTA_CASE(2, 1) VXH_WTA_CASE(2, 2)
    VXH_WTA_CASE(4, 1) VXH_WTA_CASE(4, 2) VXH_WTA_CASE(8, 1) VXH_WTA_CASE(8, 2)
#undef VXH_WTA_CASE
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
    return launch_wta_sum(vol, 0, 1, w * h, D, Dp, out, best, nullptr, s);
}

hipError_t launch_lr_check(const uint16_t *vol, int w, int h, int D, int Dp, int tol,
                           const int16_t *best, float *out, hipStream_t s) {
    const size_t n = (size_t)w * h;
    hipLaunchKernelGGL(lr_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256),
                       0, s, vol, w, h, D, Dp, tol, best, out);
    return hipGetLastError();
}

}  // namespace vxh
