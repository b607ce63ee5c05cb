// Device functions shared by the per-pair kernels (sgm_kernels.hip) and the
// batched persistent kernel (sgm_batch.hip).
//
// SGM stereo device code for gfx950 (wave64).
//
// Stages and the reference code whose results they reproduce bit for bit:
//   census_kernel     vx-vision/shaders/SGMStereo.metal:25-66
//   aggregate_kernel  vx-vision/shaders/SGMStereo.metal:72-173 (one launch per
//                     direction; diagonals start wherever the predecessor
//                     pixel lies outside the image)
//   wta_kernel        vx-vision/shaders/SGMStereo.metal:179-214
//   lr_kernel         left-right check (no reference counterpart)
//
// Aggregation layout: a group of L lanes of one wave owns one scanline; lane
// `sub` of the group owns disparities d = sub*K .. sub*K+K-1, so the pixel's
// disparity slice lives in registers, the summed volume segment of a lane is
// one packed 2K-byte vector access, and every cross-lane step (neighbour
// values, the min over d) is a DPP / permlane move inside the wave.
//
// The reference recurrence updates prev[] in place over ascending d, so the
// d-1 term sees the value already produced in the same step
// (SGMStereo.metal:155,167), and every sum wraps mod 2^16.  The d, d+1 and P2
// terms do not depend on d-1 and are evaluated for all d in parallel (A[d]);
// the d-1 chain runs sequentially inside each lane and is carried across lanes
// by fixed-point rounds: every lane first assumes no carry-in, then
// re-evaluates whenever the carry-in from the lane below changes
// min(A[first], o1).  Lane 0's carry-in is fixed, so after r rounds lanes
// 0..r are final; the loop ends when no lane changes, which is exactly the
// sequential result including every wrap.
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

// Wave mask of lanes with a < b (signed) / a != b: one v_cmp whose result the
// branch reads directly.
__device__ __forceinline__ uint64_t lanes_lt(int a, int b) {
    return __builtin_amdgcn_sicmp(a, b, 40);
}
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

// K consecutive u16 volume entries of one lane, packed two per dword.
template <int K>
struct Seg {
    static constexpr int N = K >= 2 ? K / 2 : 1;
    uint32_t w[N];
    __device__ __forceinline__ void load(const uint16_t *p) {
        if constexpr (K == 1) w[0] = *p;
        else if constexpr (K == 2) w[0] = *(const uint32_t *)p;
        else if constexpr (K == 4) {
            uint2 v = *(const uint2 *)p;
            w[0] = v.x; w[1] = v.y;
        } else {
#pragma unroll
            for (int j = 0; j < K / 8; ++j) {
                uint4 v = ((const uint4 *)p)[j];
                w[4 * j] = v.x; w[4 * j + 1] = v.y; w[4 * j + 2] = v.z; w[4 * j + 3] = v.w;
            }
        }
    }
    __device__ __forceinline__ void store(uint16_t *p) const {
        if constexpr (K == 1) *p = (uint16_t)w[0];
        else if constexpr (K == 2) *(uint32_t *)p = w[0];
        else if constexpr (K == 4) *(uint2 *)p = make_uint2(w[0], w[1]);
        else {
#pragma unroll
            for (int j = 0; j < K / 8; ++j)
                ((uint4 *)p)[j] = make_uint4(w[4 * j], w[4 * j + 1], w[4 * j + 2], w[4 * j + 3]);
        }
    }
    // w = pack(agg) (+ w if accumulate), each half mod 2^16.
    __device__ __forceinline__ void set(const uint32_t *agg, bool accumulate) {
        if constexpr (K == 1) {
            w[0] = accumulate ? ((w[0] + agg[0]) & 0xFFFFu) : agg[0];
        } else {
#pragma unroll
            for (int j = 0; j < N; ++j) {
                uint32_t pk = agg[2 * j] | (agg[2 * j + 1] << 16);
                if (accumulate) {
                    u16x2 a = __builtin_bit_cast(u16x2, pk);
                    u16x2 b = __builtin_bit_cast(u16x2, w[j]);
                    pk = __builtin_bit_cast(uint32_t, (u16x2)(a + b));
                }
                w[j] = pk;
            }
        }
    }
};

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

// Census word at byte offset `off` (a pixel index inside the image times 8;
// images have fewer than 2^29 pixels, so the offset fits 32 bits and the load
// takes the scalar-base + 32-bit-offset form).  Plain global loads, not buffer
// loads: the compiler counts these exactly when placing its waits, so a step
// waits only for its own prefetched loads.
template <int IMM = 0>
__device__ __forceinline__ uint2 load_census(const uint2 *__restrict__ c, uint32_t off) {
    return *(const uint2 *)((const char *)c + off + IMM);
}

// Loads of one step: left census of the pixel, right census at x-d for the
// lane's K disparities, and (when accumulating) the lane's volume segment.
template <int K>
struct StepBuf {
    uint2 lv;
    uint2 rv[K];
    Seg<K> sv;
    bool edge;   // some active lane has x - d < 0 for one of its d (wave-uniform)
};

// One wave aggregates scanlines s_first .. s_first + 64/L - 1 of direction
// (dx, dy) (those below p.nscan), adding into the summed volume.  Wave-level:
// every lane of the wave must call it with the same arguments.
//
// ACC: 1 = the path adds into the volume, 0 = it stores (first path), -1 =
// read p.accumulate at run time.
template <int L, int K, int ACC = -1>
__device__ __forceinline__ void agg_wave(const AggParams &p, int dx, int dy, int s_first) {
    const int lane = threadIdx.x & 63;
    const int g = lane / L, sub = lane % L;
    const int s = s_first + g;
    const int w = p.w, h = p.h, D = p.D, Dp = p.Dp;
    const uint16_t P1 = (uint16_t)p.p1, P2 = (uint16_t)p.p2;
    const bool accumulate = ACC < 0 ? p.accumulate != 0 : ACC == 1;
    const int d0 = sub * K;
    const bool lane_real = d0 < D;
    const bool lane_store = d0 < Dp;
    const uint2 *__restrict__ lc = p.lc;
    const uint2 *__restrict__ rcg = (const uint2 *)((const char *)p.rc - kCensusGuard);

    int sx = 0, sy = 0, len = 0;
    if (s < p.nscan) scan_start(s, dx, dy, w, h, sx, sy, len);
    int maxlen = len;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) maxlen = max(maxlen, __shfl_xor(maxlen, o));
    if (maxlen == 0) return;
    const int pstep = dy * w + dx;
    const int pix0 = sy * w + sx;
    // Volume element of this lane at step t: voff0 + t * vstep, kept in 32-bit
    // arithmetic (mod 2^32; the volume has fewer than 2^32 entries).
    const uint32_t voff0 = (uint32_t)pix0 * (uint32_t)Dp + (uint32_t)d0;
    const uint32_t vstep = (uint32_t)pstep * (uint32_t)Dp;

    // o2 is 0xFFFF where d + 1 >= D (SGMStereo.metal:156).
    uint16_t om[K];
#pragma unroll
    for (int i = 0; i < K; ++i) om[i] = (d0 + i + 1 < D) ? 0 : 0xFFFF;

    uint16_t prev[K];
#pragma unroll
    for (int i = 0; i < K; ++i) prev[i] = 0;
    uint16_t pm = 0;

    auto issue = [&](StepBuf<K> &b, int t) {
        const bool a = t < len;
        const uint32_t pix = a ? (uint32_t)(pix0 + t * pstep) : 0u;
        const int x = sx + t * dx;
        b.lv = load_census(lc, pix << 3);
        // Right census at x - d for the lane's K disparities: one base, constant
        // offsets.  Reads with x - d < 0 fall into the previous row or the
        // guard bytes before rc and are replaced by 64 in step().  Inactive
        // lanes read around pixel 0.
        const uint32_t rb = ((pix - (uint32_t)(d0 + K - 1)) << 3) + (uint32_t)kCensusGuard;
        static_for<K>([&](auto I) {
            constexpr int i = decltype(I)::value;
            b.rv[i] = load_census<8 * (K - 1 - i)>(rcg, rb);
        });
        // Inactive lanes count as far from the edge.
        b.edge = lanes_lt(a ? x : 0x7FFFFFFF, d0 + K - 1) != 0;
        // Lanes without a volume entry here read entry 0 (always in range)
        // and discard it.
        const uint32_t off = (a && lane_store) ? voff0 + (uint32_t)t * vstep : 0u;
        if (accumulate) b.sv.load(p.vol + off);
    };

    auto step = [&](StepBuf<K> &cur, int t) {
        const bool act = t < len;
        const int pix = act ? pix0 + t * pstep : 0;
        const int x = sx + t * dx;

        uint16_t C[K];
        const uint2 lv = cur.lv;
        if (cur.edge) {
#pragma unroll
            for (int i = 0; i < K; ++i) {
                const uint32_t c = __popc(lv.x ^ cur.rv[i].x) + __popc(lv.y ^ cur.rv[i].y);
                C[i] = (uint16_t)(x >= d0 + i ? c : 64u);
            }
        } else {
#pragma unroll
            for (int i = 0; i < K; ++i)
                C[i] = (uint16_t)(__popc(lv.x ^ cur.rv[i].x) + __popc(lv.y ^ cur.rv[i].y));
        }

        uint16_t agg[K];
        if (t == 0) {
#pragma unroll
            for (int i = 0; i < K; ++i) agg[i] = C[i];
        } else {
            const uint16_t o3 = (uint16_t)(pm + P2);
            const uint16_t nb = (uint16_t)from_upper_lane<L>((uint32_t)prev[0]);
            uint16_t A[K], B[K];
#pragma unroll
            for (int i = 0; i < K; ++i) {
                const uint16_t nx = (i + 1 < K) ? prev[(i + 1) % K] : nb;
                const uint16_t o2 = (uint16_t)((uint16_t)(nx + P1) | om[i]);
                A[i] = min(min(prev[i], o2), o3);
                B[i] = (uint16_t)(C[i] - pm);
            }
            // Speculative pass: no carry-in from the lane below.
            {
                uint16_t o1 = 0xFFFF;
#pragma unroll
                for (int i = 0; i < K; ++i) {
                    agg[i] = (uint16_t)(B[i] + min(A[i], o1));
                    o1 = (uint16_t)(agg[i] + P1);
                }
            }
            // Fixed-point rounds over the carry-in (at most L).
            const bool live = act && lane_real && sub > 0;
            uint16_t e_used = A[0];
            for (int r = 0; r < L; ++r) {
                const uint16_t up = (uint16_t)from_lower_lane<L>((uint32_t)agg[K - 1]);
                // Lanes that take no carry keep e == e_used.
                const uint16_t e = live ? min(A[0], (uint16_t)(up + P1)) : e_used;
                if (lanes_ne(e, e_used) == 0) break;
                if (e != e_used) {
                    e_used = e;
                    agg[0] = (uint16_t)(B[0] + e);
                    uint16_t o1 = (uint16_t)(agg[0] + P1);
#pragma unroll
                    for (int i = 1; i < K; ++i) {
                        agg[i] = (uint16_t)(B[i] + min(A[i], o1));
                        o1 = (uint16_t)(agg[i] + P1);
                    }
                }
            }
        }

        uint16_t m = 0xFFFF;
#pragma unroll
        for (int i = 0; i < K; ++i) m = min(m, (uint16_t)(agg[i] | (uint16_t)(d0 + i < D ? 0 : 0xFFFF)));
        pm = (uint16_t)group_min<L>((uint32_t)m);
#pragma unroll
        for (int i = 0; i < K; ++i) prev[i] = agg[i];

        uint32_t a32[K];
#pragma unroll
        for (int i = 0; i < K; ++i) a32[i] = agg[i];
        if (p.wta_out == nullptr) {
            cur.sv.set(a32, accumulate);
            uint16_t *dst = (act && lane_store) ? p.vol + (voff0 + (uint32_t)t * vstep)
                                                : store_sink<uint16_t>();
            cur.sv.store(dst);
        } else {
            // Last path: the sums are final here, so pick the winner from
            // registers (SGMStereo.metal:190-211) instead of storing them.
            cur.sv.set(a32, accumulate);
            uint32_t v[K];
            if constexpr (K == 1) {
                v[0] = cur.sv.w[0];
            } else {
#pragma unroll
                for (int i = 0; i < K; ++i) v[i] = (cur.sv.w[i / 2] >> (16 * (i & 1))) & 0xFFFFu;
            }
            // First strict minimum: key = value << 16 | d, smallest key wins.
            uint32_t key = 0xFFFFFFFFu;
#pragma unroll
            for (int i = 0; i < K; ++i)
                if (d0 + i < D) key = min(key, (v[i] << 16) | (uint32_t)(d0 + i));
            key = group_min<L>(key);
            const int bd = (int)(key & 0xFFFFu);
            const uint32_t bc = key >> 16;
            // Neighbours bd-1 and bd+1, each held by exactly one lane.
            uint32_t nb2 = 0;
#pragma unroll
            for (int i = 0; i < K; ++i) {
                if (d0 + i == bd - 1) nb2 |= v[i] << 16;
                if (d0 + i == bd + 1) nb2 |= v[i];
            }
            nb2 = group_or<L>(nb2);
            // All-0xFFFF slices keep d = 0 (strict compare from 0xFFFF).
            const int bdw = bc == 0xFFFFu ? 0 : bd;
            float disp = (float)bdw;
            if (bdw > 0 && bdw < D - 1) {
                const float c0 = (float)(nb2 >> 16);
                const float c1 = (float)bc;
                const float c2 = (float)(nb2 & 0xFFFFu);
                const float den = 2.0f * (c0 + c2 - 2.0f * c1);
                if (fabsf(den) > 1e-6f) disp = disp + (c0 - c2) / den;
            }
            float *dst = (act && sub == 0) ? p.wta_out + pix : store_sink<float>();
            *dst = disp;
        }
    };

    // Loads run PF steps ahead through a ring of PF + 1 buffers; the loop is
    // unrolled by the ring size so every buffer stays in fixed registers, and
    // runs a whole number of rounds (the steps past maxlen are inactive on
    // every lane: their loads read in-range entries, their stores go to the
    // sink).  With no branch around any load or store, every step issues the
    // same memory operations and waits only for its own PF-old loads.
    constexpr int PF = K <= 2 ? 8 : (K <= 4 ? 6 : (K <= 8 ? 3 : 1));
    constexpr int NB = PF + 1;
    const int rounds = __builtin_amdgcn_readfirstlane((maxlen + NB - 1) / NB);
    StepBuf<K> buf[NB];
    // The prologue keeps its issue order (buffer 0 first): the waits at the
    // top of the loop are derived from it.
#pragma unroll
    for (int j = 0; j < PF; ++j) {
        issue(buf[j], j);
        __builtin_amdgcn_sched_barrier(0);
    }
    for (int r = 0; r < rounds; ++r) {
        const int t = r * NB;
#pragma unroll
        for (int u = 0; u < NB; ++u) {
            issue(buf[(u + PF) % NB], t + u + PF);
            step(buf[u], t + u);
        }
    }
}


// Winner-take-all and parabola refinement of pixel i (SGMStereo.metal:185-213).
__device__ __forceinline__ void wta_pixel(const uint16_t *__restrict__ vol, int D, int Dp,
                                          size_t i, float *__restrict__ out,
                                          int16_t *__restrict__ best) {
    const uint16_t *v = vol + i * Dp;
    uint32_t bc = 0xFFFFu;
    int bd = 0;
    // The slice is 32-byte aligned (Dp is a multiple of 16): read 8 entries
    // per 16-byte load, keeping the first strict minimum.
    for (int d8 = 0; d8 < D; d8 += 8) {
        const uint4 q = *(const uint4 *)(v + d8);
        const uint32_t wv[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const uint32_t c = (wv[j / 2] >> (16 * (j & 1))) & 0xFFFFu;
            if (d8 + j < D && c < bc) {
                bc = c;
                bd = d8 + j;
            }
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

// Left-right check of pixel i: D_R(x') is the lowest d minimising
// S(x'+d, y, d) over x'+d < w; the pixel is kept iff x - best >= 0 and
// |best - D_R(x - best)| <= tol, otherwise it becomes -1.0f.
__device__ __forceinline__ void lr_pixel(const uint16_t *__restrict__ vol, int w, int D,
                                         int Dp, int tol, size_t i,
                                         const int16_t *__restrict__ best,
                                         float *__restrict__ out) {
    const int x = (int)(i % w), y = (int)(i / w);
    const int d = best[i];
    const int xr = x - d;
    bool ok = false;
    if (xr >= 0) {
        uint32_t bc = 0xFFFFu;
        int bd = 0;
        for (int k = 0; k < D && xr + k < w; ++k) {
            const uint32_t c = vol[((size_t)y * w + xr + k) * Dp + k];
            if (c < bc) {
                bc = c;
                bd = k;
            }
        }
        ok = abs(d - bd) <= tol;
    }
    if (!ok) out[i] = -1.0f;
}

}  // namespace dev
}  // namespace vxh
