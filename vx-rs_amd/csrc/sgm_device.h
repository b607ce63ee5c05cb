// Device helpers shared by the per-pair kernels (sgm_kernels.hip) and the
// batched persistent kernel (sgm_batch.hip): wave-level DPP / permlane moves
// and reductions, scanline geometry, the census of one pixel
// (vx-vision/shaders/SGMStereo.metal:25-66) and the left-right check (no
// reference counterpart).  The path aggregation and the winner-take-all are in
// sgm_paths.h.
#pragma once

#include <type_traits>

#include "sgm_internal.h"

namespace vxh {
namespace dev {

typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

// DPP controls (gfx9 encoding).
constexpr int kQuadXor1 = 0xB1;      // quad_perm [1,0,3,2]
constexpr int kQuadXor2 = 0x4E;      // quad_perm [2,3,0,1]
constexpr int kRowShl1 = 0x101;      // lane i <- lane i+1 within a row of 16
constexpr int kRowShr1 = 0x111;      // lane i <- lane i-1 within a row of 16
constexpr int kWaveShl1 = 0x130;     // lane i <- lane i+1 across the wave
constexpr int kWaveShr1 = 0x138;     // lane i <- lane i-1 across the wave
constexpr int kRowMirror = 0x140;    // lane i <- lane 15-i within a row
constexpr int kRowHalfMirror = 0x141;// lane i <- lane 7-i within 8 lanes

// Calls f(std::integral_constant<int, i>) for i = 0 .. N-1.
template <int N, typename F, int I = 0>
__device__ __forceinline__ void static_for(F &&f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<N, F, I + 1>(static_cast<F &&>(f));
    }
}

// Lanes whose DPP source lies outside the wave (the ends of the wave shifts)
// read 0 (bound_ctrl); every caller masks those lanes.  With no `old` operand
// the compiler folds the move into the consuming ALU op (v_min_u32_dpp ...).
template <int CTRL>
__device__ __forceinline__ uint32_t dpp_mov(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, CTRL, 0xF, 0xF, true);
}

// Wave mask of lanes with a != b: one v_cmp whose result the branch reads
// directly.
__device__ __forceinline__ uint64_t lanes_ne(uint32_t a, uint32_t b) {
    return __builtin_amdgcn_uicmp(a, b, 33);
}

// Reduction with `op` over the L lanes of a group, broadcast to all of them.
// Each step pairs a lane with one partner (xor 1, xor 2, the mirrored lane in
// 8 and in 16, then the other 16- and 32-lane halves).
template <int L, typename Op>
__device__ __forceinline__ uint32_t group_reduce(uint32_t v, Op op) {
    if constexpr (L >= 2) v = op(v, dpp_mov<kQuadXor1>(v));
    if constexpr (L >= 4) v = op(v, dpp_mov<kQuadXor2>(v));
    if constexpr (L >= 8) v = op(v, dpp_mov<kRowHalfMirror>(v));
    if constexpr (L >= 16) v = op(v, dpp_mov<kRowMirror>(v));
    if constexpr (L >= 32) {
        auto p = __builtin_amdgcn_permlane16_swap(v, v, false, false);
        v = op(v, op((uint32_t)p[0], (uint32_t)p[1]));
    }
    if constexpr (L >= 64) {
        auto p = __builtin_amdgcn_permlane32_swap(v, v, false, false);
        v = op(v, op((uint32_t)p[0], (uint32_t)p[1]));
    }
    return v;
}

template <int L>
__device__ __forceinline__ uint32_t group_min(uint32_t v) {
    return group_reduce<L>(v, [](uint32_t a, uint32_t b) { return min(a, b); });
}

template <int L>
__device__ __forceinline__ uint32_t group_or(uint32_t v) {
    return group_reduce<L>(v, [](uint32_t a, uint32_t b) { return a | b; });
}

// Value held by lane sub-1 / sub+1 of the same group (lanes at the group
// edge receive a neighbour group's value, which callers mask).
template <int L>
__device__ __forceinline__ uint32_t from_lower_lane(uint32_t v) {
    if constexpr (L <= 16) return dpp_mov<kRowShr1>(v);
    else return dpp_mov<kWaveShr1>(v);
}
template <int L>
__device__ __forceinline__ uint32_t from_upper_lane(uint32_t v) {
    if constexpr (L <= 16) return dpp_mov<kRowShl1>(v);
    else return dpp_mov<kWaveShl1>(v);
}

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

// Census word of pixel (x, y) with a compile-time window of
// (2 RX + 1) x (2 RY + 1) (SGMStereo.metal:41-65): pixels at least RX / RY
// inside the image read their window without clamping.
template <int RX, int RY>
__device__ __forceinline__ uint2 census_pixel_fixed(const uint8_t *__restrict__ img, int w,
                                                    int h, int x, int y) {
    const uint32_t c = img[(size_t)y * w + x];
    uint64_t bits = 0;
    if (x >= RX && x + RX < w && y >= RY && y + RY < h) {
        const uint8_t *base = img + (size_t)(y - RY) * w + (x - RX);
        int bit = 0;
#pragma unroll
        for (int dy = 0; dy <= 2 * RY; ++dy) {
            const uint8_t *row = base + (size_t)dy * w;
#pragma unroll
            for (int dx = 0; dx <= 2 * RX; ++dx) {
                if (dx == RX && dy == RY) continue;
                bits |= (uint64_t)((uint32_t)row[dx] < c) << bit;
                ++bit;
            }
        }
    } else {
        int bit = 0;
#pragma unroll
        for (int dy = -RY; dy <= RY; ++dy) {
            const uint8_t *row = img + (size_t)min(max(y + dy, 0), h - 1) * w;
#pragma unroll
            for (int dx = -RX; dx <= RX; ++dx) {
                if (dx == 0 && dy == 0) continue;
                bits |= (uint64_t)((uint32_t)row[min(max(x + dx, 0), w - 1)] < c) << bit;
                ++bit;
            }
        }
    }
    return make_uint2((uint32_t)bits, (uint32_t)(bits >> 32));
}

// Census word of pixel (x, y) (SGMStereo.metal:31-65).
__device__ __forceinline__ uint2 census_pixel(const uint8_t *__restrict__ img, int w,
                                              int h, int rx, int ry, int x, int y) {
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
    return make_uint2(lo, hi);
}

// census_pixel with the default 9x7 window compiled in.
__device__ __forceinline__ uint2 census_at(const uint8_t *__restrict__ img, int w, int h,
                                           int rx, int ry, int x, int y) {
    if (rx == 4 && ry == 3) return census_pixel_fixed<4, 3>(img, w, h, x, y);
    return census_pixel(img, w, h, rx, ry, x, y);
}

// Target of the stores of lanes with nothing to store (past the end of their
// scanline, or padding lanes past the volume stride).  Every memory operation
// of the scan loop is issued unconditionally, so each step issues the same
// number of them on every path and the compiler can wait for exactly the
// loads a step consumes instead of draining the prefetch ring.  32 bytes per
// lane (the widest segment store).
static __device__ uint4 g_store_sink[64 * 2];

template <typename T>
__device__ __forceinline__ T *store_sink() {
    return (T *)&g_store_sink[(threadIdx.x & 63) * 2];
}

}  // namespace dev
}  // namespace vxh
