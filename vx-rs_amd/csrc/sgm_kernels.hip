// SGM stereo kernels for gfx950 (wave64).
//
// Stages and the reference code whose results they reproduce bit for bit:
//   census_kernel     vx-vision/shaders/SGMStereo.metal:25-66
//   aggregate_kernel  vx-vision/shaders/SGMStereo.metal:72-173 (one launch per
//                     direction; diagonals per the start rule in sgm_internal.h)
//   wta_kernel        vx-vision/shaders/SGMStereo.metal:179-214
//   lr_kernel         left-right check (no reference counterpart)
//
// Aggregation layout: one scanline is owned by a group of L lanes of a wave;
// lane `sub` of the group owns disparities d = sub*K .. sub*K+K-1, so the
// per-pixel disparity slice lives in registers and the summed volume row
// (d innermost) is read and written as one contiguous segment per step.
//
// The reference recurrence updates prev[] in place over ascending d, so the
// d-1 term sees the value already produced in the same step
// (SGMStereo.metal:155,167) and every sum wraps mod 2^16.  The part that does
// not depend on d-1 (the d, d+1 and P2 terms) is computed for all d in
// parallel; the d-1 chain runs sequentially inside each lane and is carried
// across lanes by fixed-point rounds: each lane first assumes no carry-in,
// then re-evaluates whenever the carry-in from the lane below changes the
// value of min(A[first], o1).  Lane 0's carry-in is fixed, so after r rounds
// lanes 0..r hold final values; the loop ends when no lane changes, which is
// exactly the sequential result including every wrap.
#include "sgm_internal.h"

namespace vxh {
namespace {

constexpr uint32_t kNone = 0xFFFFFFFFu;

__device__ __forceinline__ void scan_start(int s, int dx, int dy, int w, int h,
                                           int &sx, int &sy, int &len) {
    if (dy == 0) {
        sy = s;
        sx = dx > 0 ? 0 : w - 1;
    } else if (dx == 0) {
        sx = s;
        sy = dy > 0 ? 0 : h - 1;
    } else if (s < w) {
        sx = s;
        sy = dy > 0 ? 0 : h - 1;
    } else {
        int j = s - w;
        sx = dx > 0 ? 0 : w - 1;
        sy = dy > 0 ? j + 1 : h - 2 - j;
    }
    int lx = dx > 0 ? w - sx : (dx < 0 ? sx + 1 : 0x7FFFFFFF);
    int ly = dy > 0 ? h - sy : (dy < 0 ? sy + 1 : 0x7FFFFFFF);
    len = lx < ly ? lx : ly;
}

__global__ void __launch_bounds__(256)
census_kernel(const uint8_t *__restrict__ img, int w, int h, int rx, int ry,
              uint2 *__restrict__ out) {
    int x = blockIdx.x * blockDim.x + threadIdx.x;
    int y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= w || y >= h) return;
    const uint32_t c = img[(size_t)y * w + x];
    uint32_t lo = 0, hi = 0;
    int bit = 0;
    for (int dy = -ry; dy <= ry; ++dy) {
        int sy = min(max(y + dy, 0), h - 1);
        const uint8_t *row = img + (size_t)sy * w;
        for (int dx = -rx; dx <= rx; ++dx) {
            if (dx == 0 && dy == 0) continue;
            int sx = min(max(x + dx, 0), w - 1);
            if ((uint32_t)row[sx] < c) {
                if (bit < 32) lo |= 1u << bit;
                else hi |= 1u << (bit - 32);
            }
            ++bit;
        }
    }
    out[(size_t)y * w + x] = make_uint2(lo, hi);
}

template <int L, int K>
__global__ void __launch_bounds__(256)
aggregate_kernel(AggParams p, int dx, int dy) {
    constexpr int G = 64 / L;
    const int lane = threadIdx.x & 63;
    const int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int g = lane / L, sub = lane % L;
    const int s = wave * G + g;
    const int w = p.w, h = p.h, D = p.D;
    const uint32_t P1 = p.p1, P2 = p.p2;

    int sx = 0, sy = 0, len = 0;
    if (s < p.nscan) scan_start(s, dx, dy, w, h, sx, sy, len);
    int maxlen = len;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) maxlen = max(maxlen, __shfl_xor(maxlen, o));
    if (maxlen == 0) return;

    uint32_t prev[K];
#pragma unroll
    for (int i = 0; i < K; ++i) prev[i] = 0;
    uint32_t pm = 0;

    for (int t = 0; t < maxlen; ++t) {
        const bool act = t < len;
        const int x = sx + t * dx, y = sy + t * dy;
        const size_t pix = act ? (size_t)y * w + x : 0;
        const uint2 lv = act ? p.lc[pix] : make_uint2(0u, 0u);
        uint32_t C[K];
#pragma unroll
        for (int i = 0; i < K; ++i) {
            const int d = sub * K + i;
            uint32_t c = 0;
            if (act && d < D) {
                if (x - d < 0) {
                    c = 64;
                } else {
                    const uint2 rv = p.rc[pix - d];
                    c = __popc(lv.x ^ rv.x) + __popc(lv.y ^ rv.y);
                }
            }
            C[i] = c;
        }

        uint32_t agg[K];
        if (t == 0) {
#pragma unroll
            for (int i = 0; i < K; ++i) agg[i] = C[i];
        } else {
            const uint32_t o3 = (pm + P2) & 0xFFFFu;
            const uint32_t nb = __shfl_down(prev[0], 1, L);
            uint32_t A[K], B[K];
#pragma unroll
            for (int i = 0; i < K; ++i) {
                const int d = sub * K + i;
                const uint32_t nx = (i + 1 < K) ? prev[(i + 1) % K] : nb;
                const uint32_t o2 = (d + 1 < D) ? ((nx + P1) & 0xFFFFu) : 0xFFFFu;
                A[i] = min(min(prev[i], o2), o3);
                B[i] = (C[i] - pm) & 0xFFFFu;
            }
            // Speculative pass: no carry-in from the lane below.
            {
                uint32_t o1 = 0xFFFFu;
#pragma unroll
                for (int i = 0; i < K; ++i) {
                    agg[i] = (B[i] + min(A[i], o1)) & 0xFFFFu;
                    o1 = (agg[i] + P1) & 0xFFFFu;
                }
            }
            // Fixed-point rounds over the carry-in.
            uint32_t e_used = A[0];
            while (true) {
                uint32_t up = __shfl_up(agg[K - 1], 1, L);
                if (sub == 0) up = kNone;
                const uint32_t o1in = up == kNone ? 0xFFFFu : ((up + P1) & 0xFFFFu);
                const uint32_t e = min(A[0], o1in);
                const bool ch = act && sub * K < D && (e != e_used);
                if (!__any(ch)) break;
                if (ch) {
                    e_used = e;
                    uint32_t o1 = o1in;
#pragma unroll
                    for (int i = 0; i < K; ++i) {
                        agg[i] = (B[i] + min(A[i], o1)) & 0xFFFFu;
                        o1 = (agg[i] + P1) & 0xFFFFu;
                    }
                }
            }
        }

        uint32_t m = 0xFFFFu;
#pragma unroll
        for (int i = 0; i < K; ++i)
            if (sub * K + i < D) m = min(m, agg[i]);
#pragma unroll
        for (int o = L / 2; o >= 1; o >>= 1) m = min(m, (uint32_t)__shfl_xor((int)m, o, L));
        pm = m;
#pragma unroll
        for (int i = 0; i < K; ++i) prev[i] = agg[i];

        if (act) {
            uint16_t *row = p.vol + pix * (size_t)D;
#pragma unroll
            for (int i = 0; i < K; ++i) {
                const int d = sub * K + i;
                if (d < D) {
                    uint32_t v = agg[i];
                    if (p.accumulate) v += row[d];
                    row[d] = (uint16_t)v;
                }
            }
        }
    }
}

__global__ void __launch_bounds__(256)
wta_kernel(const uint16_t *__restrict__ vol, int w, int h, int D,
           float *__restrict__ out, int16_t *__restrict__ best) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)w * h) return;
    const uint16_t *v = vol + i * D;
    uint32_t bc = 0xFFFFu;
    int bd = 0;
    for (int d = 0; d < D; ++d) {
        const uint32_t c = v[d];
        if (c < bc) {
            bc = c;
            bd = d;
        }
    }
    float disp = (float)bd;
    if (bd > 0 && bd < D - 1) {
        const float c0 = (float)v[bd - 1];
        const float c1 = (float)bc;
        const float c2 = (float)v[bd + 1];
        const float den = 2.0f * (c0 + c2 - 2.0f * c1);
        if (fabsf(den) > 1e-6f) disp = disp + (c0 - c2) / den;
    }
    out[i] = disp;
    if (best) best[i] = (int16_t)bd;
}

__global__ void __launch_bounds__(256)
lr_kernel(const uint16_t *__restrict__ vol, int w, int h, int D, int tol,
          const int16_t *__restrict__ best, float *__restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)w * h) return;
    const int x = (int)(i % w), y = (int)(i / w);
    const int d = best[i];
    const int xr = x - d;
    bool ok = false;
    if (xr >= 0) {
        uint32_t bc = 0xFFFFu;
        int bd = 0;
        for (int k = 0; k < D && xr + k < w; ++k) {
            const uint32_t c = vol[((size_t)y * w + xr + k) * D + k];
            if (c < bc) {
                bc = c;
                bd = k;
            }
        }
        ok = abs(d - bd) <= tol;
    }
    if (!ok) out[i] = -1.0f;
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

hipError_t launch_census(const uint8_t *img, int w, int h, int rx, int ry,
                         uint2 *out, hipStream_t s) {
    dim3 block(64, 4);
    dim3 grid((w + 63) / 64, (h + 3) / 4);
    hipLaunchKernelGGL(census_kernel, grid, block, 0, s, img, w, h, rx, ry, out);
    return hipGetLastError();
}

hipError_t launch_aggregate(const AggParams &p, hipStream_t s) {
    if (p.D <= 64) return launch_agg_lk<64, 1>(p, s);
    if (p.D <= 128) return launch_agg_lk<64, 2>(p, s);
    if (p.D <= 192) return launch_agg_lk<64, 3>(p, s);
    return launch_agg_lk<64, 4>(p, s);
}

hipError_t launch_wta(const uint16_t *vol, int w, int h, int D, float *out,
                      int16_t *best, hipStream_t s) {
    const size_t n = (size_t)w * h;
    hipLaunchKernelGGL(wta_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256),
                       0, s, vol, w, h, D, out, best);
    return hipGetLastError();
}

hipError_t launch_lr_check(const uint16_t *vol, int w, int h, int D, int tol,
                           const int16_t *best, float *out, hipStream_t s) {
    const size_t n = (size_t)w * h;
    hipLaunchKernelGGL(lr_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256),
                       0, s, vol, w, h, D, tol, best, out);
    return hipGetLastError();
}

}  // namespace vxh
