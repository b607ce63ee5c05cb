// The aggregation kernels of one layout class: every direction pass of a
// pair in one launch (paths_kernel), or with split scanlines and ticketed
// waves (paths_split_kernel).  Each layout class is compiled in its own
// translation unit (sgm_paths_c0.hip .. sgm_paths_c2.hip) through
// VXH_DEFINE_PATHS_LAUNCHER, so the three compile side by side.
#pragma once

#include "sgm_paths.h"

namespace vxh {
namespace {

using namespace dev;

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

// m.npass independent direction passes in one launch, each with one of the
// three lane layouts of class C.  Block b belongs to the pass whose
// [first_block, first_block + nblocks) holds it; inside the pass the block
// index is permuted so that the blocks the hardware places on one XCD
// (b = x mod 8) take a contiguous range of scanlines, whose census rows then
// stay in that XCD's L2.
template <int C, int MODE, bool SPLIT>
__device__ __forceinline__ void run_layout(const PassParams &p, int layout, int wave, int role,
                                           const uint32_t *tail_done, uint32_t *tail_early,
                                           uint32_t *error) {
    constexpr int L0 = kClassLayouts[C][0][0], K0 = kClassLayouts[C][0][1];
    constexpr int L1 = kClassLayouts[C][1][0], K1 = kClassLayouts[C][1][1];
    constexpr int L2 = kClassLayouts[C][2][0], K2 = kClassLayouts[C][2][1];
#define VXH_RUN(LL, KK)                                                                       \
    if (p.D % KK == 0)                                                                        \
        agg_pair_wave<LL, KK, MODE, true, SPLIT>(p, wave * (128 / LL), role, tail_done,       \
                                                 tail_early, error);                          \
    else                                                                                      \
        agg_pair_wave<LL, KK, MODE, false, SPLIT>(p, wave * (128 / LL), role, tail_done,      \
                                                  tail_early, error)
    if (layout == 0) {
        VXH_RUN(L0, K0);
    } else if (layout == 1 || kClassCount[C] < 3) {
        VXH_RUN(L1, K1);
    } else {
        VXH_RUN(L2, K2);
    }
#undef VXH_RUN
}

// Pass i's per-pass fields into p (i wave-uniform).
__device__ __forceinline__ int pass_params(const MultiPass &m, int i, PassParams &p) {
    p = m.base;
    p.vol = VXH_PASS_FIELD(m, i, vol);
    p.hi = VXH_PASS_FIELD(m, i, hi);
    p.dir = VXH_PASS_FIELD(m, i, dir);
    p.dx = VXH_PASS_FIELD(m, i, dx);
    p.dy = VXH_PASS_FIELD(m, i, dy);
    p.nscan = VXH_PASS_FIELD(m, i, nscan);
    return VXH_PASS_FIELD(m, i, layout);
}

// Split launch.  Waves take tickets from one counter per XCD (block b from
// counter b mod 8, the XCD the hardware places it on), in the order they
// start.  Counter k hands out the k-th eighth of every pass's scanline
// groups, passes longest first: for a split pass first the tail roles of the
// chunk, then its head roles, for an unsplit pass whole scanlines.  Each XCD
// thus works on contiguous scanlines (rows of one page), and a head starts
// only after its tail has started, so its waits on the tail's flags always
// end.
template <int C, int MODE>
__global__ void __launch_bounds__(256)
paths_split_kernel(MultiPass m) {
    const int lane = threadIdx.x & 63;
    uint32_t *const sync = m.sync;
    const int k = (int)(blockIdx.x & 7);
    const uint32_t tk = __hip_atomic_fetch_add(sync + 2 + k, lane == 0 ? 1u : 0u,
                                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int ticket = __builtin_amdgcn_readfirstlane((int)tk);
    int i = -1, local = 0, c0 = 0, cnt = 0, base = 0;
#pragma unroll
    for (int j = 0; j < kMaxPasses; ++j) {
        if (j < m.npass && i < 0) {
            const int nw = m.pass[j].nwaves;
            const int a = nw * k / 8, b = nw * (k + 1) / 8;
            const int n = (b - a) * (m.pass[j].split ? 2 : 1);
            if (ticket < base + n) {
                i = j;
                local = ticket - base;
                c0 = a;
                cnt = b - a;
            }
            base += n;
        }
    }
    if (i < 0) return;   // wave-uniform: no work left for this XCD
    PassParams p;
    const int layout = pass_params(m, i, p);
    const bool split = VXH_PASS_FIELD(m, i, split) != 0;
    const int role = !split ? kRoleWhole : (local < cnt ? kRoleTail : kRoleHead);
    const int wave = c0 + (local < cnt ? local : local - cnt);
    // Flags of this scanline group's tail: done, then early.
    const int gw = VXH_PASS_FIELD(m, i, first_wave) + wave;
    uint32_t *const done = sync + kSyncFlags + gw;
    run_layout<C, MODE, true>(p, layout, wave, role, done, done + m.total_waves, sync + 1);
    if (role == kRoleTail) {
        // Every store of this tail is visible before its flag.
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_store(done, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

template <int C, int MODE>
__global__ void __launch_bounds__(256)
paths_kernel(MultiPass m) {
    const int b = blockIdx.x;
    int i = 0;
#pragma unroll
    for (int j = 1; j < kMaxPasses; ++j)
        if (j < m.npass && b >= m.pass[j].first_block) i = j;
    const int local = b - VXH_PASS_FIELD(m, i, first_block);
    const int per = VXH_PASS_FIELD(m, i, nblocks) >> 3;
    const int lb = (local & 7) * per + (local >> 3);
    PassParams p;
    const int layout = pass_params(m, i, p);
    const int wave = lb * 4 + (int)(threadIdx.x >> 6);
    run_layout<C, MODE, false>(p, layout, wave, kRoleWhole, nullptr, nullptr, nullptr);
}

// The persistent aggregation of the batched form (see BatchPaths): every
// wave takes items from its XCD's list until the list is exhausted.
template <int C, int MODE>
__global__ void __launch_bounds__(256)
batch_paths_kernel(BatchPaths b) {
    const int k = (int)(blockIdx.x & 7);
    const uint32_t n = b.item_cnt[k];
    const BatchItem *__restrict__ items = b.items + b.item_off[k];
    const int lane = threadIdx.x & 63;
    for (;;) {
        uint32_t tk = 0;
        if (lane == 0)
            tk = __hip_atomic_fetch_add(b.counters + k, 1u, __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t ticket = (uint32_t)__builtin_amdgcn_readfirstlane((int)tk);
        if (ticket >= n) break;   // wave-uniform
        const BatchItem it = items[ticket];
        const BatchPass &d = b.passes[it.pass];
        const PassParams p = d.p;
        run_layout<C, MODE, false>(p, d.layout, (int)it.wave, kRoleWhole, nullptr, nullptr,
                                   nullptr);
    }
}

}  // namespace

}  // namespace vxh

// Defines launch_paths_c<CC>(), the host launcher of class CC's kernels.
#define VXH_DEFINE_PATHS_LAUNCHER(CC)                                                      \
    namespace vxh {                                                                        \
    hipError_t launch_paths_c##CC(int mode, bool split, const MultiPass &m, int blocks,    \
                                  hipStream_t s) {                                         \
        if (mode == kOutU8) {                                                              \
            if (split)                                                                     \
                hipLaunchKernelGGL((paths_split_kernel<CC, kOutU8>), dim3(blocks),         \
                                   dim3(256), 0, s, m);                                    \
            else                                                                           \
                hipLaunchKernelGGL((paths_kernel<CC, kOutU8>), dim3(blocks), dim3(256), 0, \
                                   s, m);                                                  \
        } else if (mode == kOutU16) {                                                      \
            if (split)                                                                     \
                hipLaunchKernelGGL((paths_split_kernel<CC, kOutU16>), dim3(blocks),        \
                                   dim3(256), 0, s, m);                                    \
            else                                                                           \
                hipLaunchKernelGGL((paths_kernel<CC, kOutU16>), dim3(blocks), dim3(256),   \
                                   0, s, m);                                               \
        } else {                                                                           \
            return hipErrorInvalidValue;                                                   \
        }                                                                                  \
        return hipGetLastError();                                                          \
    }                                                                                      \
    hipError_t launch_batch_paths_c##CC(int mode, const BatchPaths &b, int blocks,         \
                                        hipStream_t s) {                                   \
        if (mode == kOutU8)                                                                \
            hipLaunchKernelGGL((batch_paths_kernel<CC, kOutU8>), dim3(blocks), dim3(256),  \
                               0, s, b);                                                   \
        else if (mode == kOutU16)                                                          \
            hipLaunchKernelGGL((batch_paths_kernel<CC, kOutU16>), dim3(blocks), dim3(256), \
                               0, s, b);                                                   \
        else                                                                               \
            return hipErrorInvalidValue;                                                   \
        return hipGetLastError();                                                          \
    }                                                                                      \
    int batch_paths_occupancy_c##CC(int mode) {                                            \
        int n = 0;                                                                         \
        const hipError_t e =                                                               \
            mode == kOutU8                                                                 \
                ? hipOccupancyMaxActiveBlocksPerMultiprocessor(                            \
                      &n, batch_paths_kernel<CC, kOutU8>, 256, 0)                          \
                : hipOccupancyMaxActiveBlocksPerMultiprocessor(                            \
                      &n, batch_paths_kernel<CC, kOutU16>, 256, 0);                        \
        return e == hipSuccess ? n : 0;                                                    \
    }                                                                                      \
    }
