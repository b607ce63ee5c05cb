// Batched SGM: many pairs of uneven size and disparity range in one
// persistent-kernel launch, captured in a hipGraph.
//
// The reference runs one pair per command buffer (vx-vision/src/kernels/
// sgm.rs:135-261) and its Pipeline batches encodes of different kernels into
// one command buffer (vx-vision/src/pipeline.rs:12-80).  Here a batch is a
// work list of wave-sized items, each a slice of one pair's stage:
//   phase 0   census rows of the left or right image
//   phase 1   matching-cost rows (64-pixel tiles, as the per-pair kernel)
//   phase 2   scanline groups of any path direction (the directions of a
//             pair are independent: each writes its own volume)
//   phase 3   sum of the direction volumes and winner-take-all, pixel ranges
//   phase 4   left-right check pixels (only with lr_check)
// A wave takes the next item from a global counter, waits until every item of
// the same pair's previous phase has signalled completion, runs the slice with
// the same device code as the per-pair kernels, and signals.  Every item
// depends only on items earlier in the list, and an item is taken only by a
// running wave, so the waits always end.  Items are ordered so that the
// memory-bound sums of finished pairs run beside the aggregation of later
// pairs.  Per pair the arithmetic is exactly the per-pair path's, so the
// output is bit-identical.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <string>
#include <vector>

#include "sgm_paths.h"
#include "vxh/sgm.h"
#include "vxh/sgm_batch.h"

namespace vxh {
namespace {

using namespace dev;

constexpr int kMaxPhase = 5;      // census, cost, paths, wta, lr
constexpr int kLaunchKinds = 4;   // (u8, u16) x (D % 8 == 0, else)

enum ItemKind : int { kCensusL = 0, kCensusR = 1, kCost = 2, kAgg = 3, kWta = 4, kLr = 5 };

struct PairDesc {
    const uint8_t *left, *right;
    float *out;
    uint2 *lc, *rc, *flags;  // flags: kOutU8 only
    uint8_t *cost;           // matching-cost volume
    uint8_t *vol, *hi;       // direction k's volume at vol + k * vol_bytes (hi likewise)
    size_t vol_bytes;
    uint16_t *sum;           // lr_check only: summed volume
    int16_t *best;           // lr_check only: integer winners
    int w, h, D, Dp, P, mode;
    uint32_t p1x2, p2x2;
    int rx, ry;
    int cls;                 // layout class
    int lr_tol;
};

struct Item {
    int pair;
    int kind;
    int dir;      // path index for kAgg
    int first;    // first row / scanline / pixel
    int count;    // rows / pixels
    int phase;
};

__device__ __forceinline__ int dir_x(int i) {
    return i == 0 || i == 4 || i == 6 ? 1 : (i == 2 || i == 5 || i == 7 ? -1 : 0);
}
__device__ __forceinline__ int dir_y(int i) {
    return i == 1 || i == 4 || i == 5 ? 1 : (i == 3 || i == 6 || i == 7 ? -1 : 0);
}
__device__ __forceinline__ int dir_nscan(int i, int w, int h) {
    return dir_y(i) == 0 ? h : (dir_x(i) == 0 ? w : w + h - 1);
}

__global__ void batch_reset_kernel(uint32_t *counters, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) counters[i] = 0;
}

// Every item-level step below is wave-uniform: all 64 lanes issue the
// counter atomics (only lane 0 adds 1) and all poll the same counter, so the
// item loop has a single, uniform exit.  (Lane-0-only branches around these
// steps let the compiler build a loop in which only some lanes leave after an
// item while the rest re-enter with a stale index.)
__device__ __forceinline__ uint32_t wave_take(uint32_t *ctr, int lane) {
    const uint32_t v = __hip_atomic_fetch_add(ctr, lane == 0 ? 1u : 0u, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT);
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}

// Waits until *ctr >= target.  Bounded: after kSpinLimit polls (seconds) the
// wave raises *error and returns, so a planning mistake ends the kernel with
// an error instead of hanging the device.
constexpr uint32_t kSpinLimit = 1u << 22;

__device__ __forceinline__ bool wave_wait(const uint32_t *ctr, uint32_t target,
                                          uint32_t *error) {
    for (uint32_t n = 0; n < kSpinLimit; ++n) {
        const uint32_t v = __hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if ((uint32_t)__builtin_amdgcn_readfirstlane((int)v) >= target) return true;
        const uint32_t e = __hip_atomic_load(error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (__builtin_amdgcn_readfirstlane((int)e) != 0) return false;
        __builtin_amdgcn_s_sleep(2);
    }
    __hip_atomic_store(error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return false;
}

// The batch runs the layout with 8 disparities per lane in every class (the
// most disparities per lane within the register budget; a batch is bound by
// instruction throughput, not by its longest scanline).  The persistent kernel
// carries one aggregation body per class, so that register use stays at two
// waves per SIMD.
constexpr int kBatchLayout[kLayoutClasses] = {2, 2, 1};

template <int C, int MODE, bool ALIGNED>
__device__ __forceinline__ void agg_item(const PassParams &p, int first) {
    constexpr int L = kClassLayouts[C][kBatchLayout[C]][0];
    constexpr int K = kClassLayouts[C][kBatchLayout[C]][1];
    static_assert(K == 8, "batch layouts hold 8 disparities per lane");
    agg_pair_wave<L, K, MODE, ALIGNED>(p, first);
}

// One persistent launch runs the pairs of one volume format (MODE) whose D
// is (ALIGNED) or is not a multiple of 8, so that it carries a single
// aggregation body per layout class and keeps two waves per SIMD resident
// (the persistent grid is sized for them).
template <int MODE, bool ALIGNED>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
batch_kernel(const PairDesc *__restrict__ pairs, const Item *__restrict__ items,
             int nitems, uint32_t *next, uint32_t *done,
             const uint32_t *__restrict__ need, uint32_t *error) {
    __shared__ uint2 span[4][kCostSlots];
    const int lane = threadIdx.x & 63;
    for (uint32_t idx = wave_take(next, lane); idx < (uint32_t)nitems;
         idx = wave_take(next, lane)) {
        const Item it = items[idx];
        const PairDesc pd = pairs[it.pair];
        if (it.phase > 0) {
            const int c = it.pair * kMaxPhase + it.phase - 1;
            if (!wave_wait(done + c, need[c], error)) continue;
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
        if (it.kind == kCensusL || it.kind == kCensusR) {
            const uint8_t *img = it.kind == kCensusL ? pd.left : pd.right;
            uint2 *out = it.kind == kCensusL ? pd.lc : pd.rc;
            uint2 *fl = it.kind == kCensusL ? pd.flags : nullptr;
            for (int y = it.first; y < it.first + it.count; ++y)
                for (int x = lane; x < pd.w; x += 64) {
                    const size_t i = (size_t)y * pd.w + x;
                    out[i] = census_at(img, pd.w, pd.h, pd.rx, pd.ry, x, y);
                    if (fl) fl[i] = make_uint2(0, 0);
                }
        } else if (it.kind == kCost) {
            for (int y = it.first; y < it.first + it.count; ++y)
                for (int x0 = 0; x0 < pd.w; x0 += kCostTile)
                    cost_tile(pd.lc, pd.rc, pd.w, pd.D, pd.Dp, y, x0, pd.cost,
                              span[threadIdx.x >> 6]);
        } else if (it.kind == kAgg) {
            PassParams p;
            p.cost = pd.cost;
            p.vol = pd.vol + (size_t)it.dir * pd.vol_bytes;
            p.hi = pd.hi ? pd.hi + (size_t)it.dir * pd.vol_bytes : nullptr;
            p.flags = (uint8_t *)pd.flags;
            p.w = pd.w;
            p.h = pd.h;
            p.D = pd.D;
            p.Dp = pd.Dp;
            p.p1x2 = pd.p1x2;
            p.p2x2 = pd.p2x2;
            p.dir = it.dir;
            p.dx = dir_x(it.dir);
            p.dy = dir_y(it.dir);
            p.nscan = dir_nscan(it.dir, pd.w, pd.h);
            p.mode = MODE;
            if (pd.cls == 0) agg_item<0, MODE, ALIGNED>(p, it.first);
            else if (pd.cls == 1) agg_item<1, MODE, ALIGNED>(p, it.first);
            else agg_item<2, MODE, ALIGNED>(p, it.first);
        } else if (it.kind == kWta) {
            // Four pixels per wave step, 16 lanes each.
            const int npx = pd.w * pd.h;
            const int end = it.first + it.count;
            const int nc = pd.Dp > 128 ? 2 : 1;
            auto px = [&](int b) {
                const int q = b + (lane >> 4);
                return q < end ? q : npx;
            };
            for (int base = it.first; base < end; base += 4)
                wta_sum_pixel(pd.vol, pd.vol_bytes, pd.hi, pd.flags, pd.P, nc, MODE, px(base),
                              npx, pd.D, pd.Dp, pd.out, pd.best, pd.sum);
        } else {
            for (int i = it.first + lane; i < it.first + it.count; i += 64)
                lr_pixel(pd.sum, pd.w, pd.D, pd.Dp, pd.lr_tol, (size_t)i, pd.best, pd.out);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_fetch_add(done + it.pair * kMaxPhase + it.phase, lane == 0 ? 1u : 0u,
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Index of (L, K) among the layouts of class c, -1 if none.
int layout_index(int c, int L, int K) {
    for (int i = 0; i < kClassCount[c]; ++i)
        if (kClassLayouts[c][i][0] == L && kClassLayouts[c][i][1] == K) return i;
    return -1;
}

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

}  // namespace
}  // namespace vxh

struct vxh_sgm_batch {
    int device = 0;
    int npairs = 0;
    int nitems = 0;
    int grid = 0;
    int nlaunch = 0;             // persistent launches (one per format and alignment)
    void *slab = nullptr;       // per-pair workspaces
    void *meta = nullptr;       // pairs, items, counters
    vxh::PairDesc *d_pairs = nullptr;
    vxh::Item *d_items = nullptr;
    uint32_t *d_need = nullptr;
    uint32_t *d_counters = nullptr;   // [0..4) = next per launch, [4] = error, [5..] = done
    int ncounters = 0;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<vxh_sgm_batch_stats> stats;
};

namespace {
thread_local std::string g_batch_error;
vxh_status bfail(vxh_status st, const std::string &m) {
    g_batch_error = m;
    return st;
}
}  // namespace

extern "C" {

const char *vxh_sgm_batch_last_error(void) { return g_batch_error.c_str(); }

vxh_status vxh_sgm_batch_create(int device, const vxh_sgm_pair_desc *pairs, uint32_t n,
                                vxh_sgm_batch **out) {
    return vxh_sgm_batch_create_ex(device, pairs, n, nullptr, out);
}

vxh_status vxh_sgm_batch_create_ex(int device, const vxh_sgm_pair_desc *pairs, uint32_t n,
                                   const vxh_sgm_batch_options *opt, vxh_sgm_batch **out) {
    using namespace vxh;
    vxh_sgm_batch_options o{2, 2, 1024};
    if (opt) {
        if (opt->blocks_per_cu > 2 || opt->wta_pixels % 4 != 0 || opt->wta_pixels > (1u << 20) ||
            opt->wta_lag > 64)
            return bfail(VXH_ERR_INVALID_CONFIG, "invalid config: batch options");
        if (opt->blocks_per_cu) o.blocks_per_cu = opt->blocks_per_cu;
        if (opt->wta_lag) o.wta_lag = opt->wta_lag;
        if (opt->wta_pixels) o.wta_pixels = opt->wta_pixels;
    }
    if (!out || (!pairs && n)) return bfail(VXH_ERR_INVALID_CONFIG, "invalid config: null argument");
    *out = nullptr;
    if (n == 0 || n > 65535) return bfail(VXH_ERR_INVALID_CONFIG, "invalid config: batch size");
    if (hipSetDevice(device) != hipSuccess)
        return bfail(VXH_ERR_DEVICE_NOT_FOUND, "no HIP device found");

    std::vector<PairDesc> pd(n);
    std::vector<size_t> off(n);
    size_t slab = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const vxh_sgm_pair_desc &q = pairs[i];
        if (vxh_sgm_config_validate(&q.cfg, q.width, q.height) != VXH_OK)
            return bfail(VXH_ERR_INVALID_CONFIG,
                         std::string("pair ") + std::to_string(i) + ": " + vxh_last_error());
        if (!q.d_left || !q.d_right || !q.d_out)
            return bfail(VXH_ERR_INVALID_CONFIG, "invalid config: null pair buffer");
        if (q.volume_format > 2)
            return bfail(VXH_ERR_INVALID_CONFIG, "invalid config: volume format");
        const int D = (int)q.cfg.num_disparities;
        const int cls = layout_class(D);
        const int li = kBatchLayout[cls];
        const bool sched_ok =
            ((q.lanes == 0 && q.per_lane == 0) || layout_index(cls, (int)q.lanes, (int)q.per_lane) == li) &&
            ((q.lanes_h == 0 && q.per_lane_h == 0) ||
             layout_index(cls, (int)q.lanes_h, (int)q.per_lane_h) == li);
        if (!sched_ok)
            return bfail(VXH_ERR_INVALID_CONFIG,
                         "invalid config: unsupported lane layout (the batch runs " +
                             std::to_string(kClassLayouts[cls][li][0]) + "x" +
                             std::to_string(kClassLayouts[cls][li][1]) + " for D=" +
                             std::to_string(D) + ")");
        const int mode = q.volume_format == 1 ? kOutU8
                         : q.volume_format == 2 ? kOutU16
                                                : auto_volume_mode(q.cfg.p2);
        const WsLayout L = workspace_layout(q.width, q.height, (uint32_t)D, (int)q.cfg.num_paths,
                                            mode, q.cfg.lr_check != 0);
        off[i] = slab;
        slab += ws_align(L.total);
        PairDesc &d = pd[i];
        d.left = q.d_left;
        d.right = q.d_right;
        d.out = q.d_out;
        d.vol_bytes = L.vol_bytes;
        d.w = (int)q.width;
        d.h = (int)q.height;
        d.D = D;
        d.Dp = volume_stride(D);
        d.P = (int)q.cfg.num_paths;
        d.mode = mode;
        const uint32_t p1 = (uint32_t)q.cfg.p1 & 0xFFFFu, p2 = (uint32_t)q.cfg.p2 & 0xFFFFu;
        d.p1x2 = p1 | (p1 << 16);
        d.p2x2 = p2 | (p2 << 16);
        d.rx = (int)q.cfg.census_radius_x;
        d.ry = (int)q.cfg.census_radius_y;
        d.cls = cls;
        d.lr_tol = (int)std::min<uint32_t>(q.cfg.lr_tolerance, 1024u);
    }

    vxh_sgm_batch *b = new vxh_sgm_batch();
    b->device = device;
    b->npairs = (int)n;
    if (hipMalloc(&b->slab, slab) != hipSuccess) {
        delete b;
        return bfail(VXH_ERR_BUFFER_ALLOC, "buffer allocation failed (" +
                                               std::to_string(slab) + " bytes)");
    }
    for (uint32_t i = 0; i < n; ++i) {
        char *base = (char *)b->slab + off[i];
        const WsLayout L = workspace_layout(pairs[i].width, pairs[i].height,
                                            (uint32_t)pd[i].D, pd[i].P, pd[i].mode,
                                            pairs[i].cfg.lr_check != 0);
        const bool u8 = pd[i].mode == kOutU8, lr = pairs[i].cfg.lr_check != 0;
        pd[i].lc = (uint2 *)(base + L.lc);
        pd[i].rc = (uint2 *)(base + L.rc);
        pd[i].cost = (uint8_t *)(base + L.cost);
        pd[i].flags = u8 ? (uint2 *)(base + L.flags) : nullptr;
        pd[i].vol = (uint8_t *)(base + L.vol);
        pd[i].hi = u8 ? (uint8_t *)(base + L.hi) : nullptr;
        pd[i].sum = lr ? (uint16_t *)(base + L.sum) : nullptr;
        pd[i].best = lr ? (int16_t *)(base + L.best) : nullptr;
    }

    // Work list.  Census of every pair first; then, pair by pair (largest
    // first), the next pair's cost rows and this pair's aggregation items
    // (horizontal directions, the longest scanlines, first), each pair's
    // sum + winner items placed wta_lag pairs later so that they find its
    // directions finished and run beside the next pairs' aggregation; the LR
    // checks last.
    std::vector<int> order(n);
    for (uint32_t i = 0; i < n; ++i) order[i] = (int)i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) {
        return (size_t)pd[a].w * pd[a].h * pd[a].D > (size_t)pd[c].w * pd[c].h * pd[c].D;
    });
    std::vector<Item> items;
    std::vector<uint32_t> need((size_t)n * kMaxPhase, 0);
    b->stats.resize(n);
    auto push = [&](int p, int kind, int dir, int first, int count, int phase) {
        items.push_back(Item{p, kind, dir, first, count, phase});
        need[(size_t)p * kMaxPhase + phase]++;
    };
    const int px_item = (int)o.wta_pixels;
    auto push_wta = [&](int p) {
        const int npx = pd[p].w * pd[p].h;
        for (int i = 0; i < npx; i += px_item)
            push(p, kWta, 0, i, std::min(px_item, npx - i), 3);
    };
    for (int p : order) {
        const PairDesc &d = pd[p];
        const int rows = std::max(1, 4096 / d.w);
        for (int y = 0; y < d.h; y += rows) {
            push(p, kCensusL, 0, y, std::min(rows, d.h - y), 0);
            push(p, kCensusR, 0, y, std::min(rows, d.h - y), 0);
        }
    }
    auto push_cost = [&](int p) {
        const int rows = std::max(1, 2048 / pd[p].w);
        for (int y = 0; y < pd[p].h; y += rows)
            push(p, kCost, 0, y, std::min(rows, pd[p].h - y), 1);
    };
    push_cost(order[0]);
    const size_t kWtaLag = o.wta_lag;
    for (size_t k = 0; k < order.size(); ++k) {
        const int p = order[k];
        const PairDesc &d = pd[p];
        if (k + 1 < order.size()) push_cost(order[k + 1]);
        for (int pass = 0; pass < 2; ++pass)
            for (int dir = 0; dir < d.P; ++dir) {
                const bool horiz = kDirY[dir] == 0;
                if (horiz != (pass == 0)) continue;
                const int per = 128 / kClassLayouts[d.cls][kBatchLayout[d.cls]][0];
                const int ns = num_scanlines(dir, d.w, d.h);
                for (int s0 = 0; s0 < ns; s0 += per) push(p, kAgg, dir, s0, per, 2);
            }
        if (k >= kWtaLag) push_wta(order[k - kWtaLag]);
    }
    for (size_t k = order.size() >= kWtaLag ? order.size() - kWtaLag : 0; k < order.size(); ++k)
        push_wta(order[k]);
    for (int p : order) {
        if (!pairs[p].cfg.lr_check) continue;
        const int npx = pd[p].w * pd[p].h, chunk = 64 * 64;
        for (int i = 0; i < npx; i += chunk) push(p, kLr, 0, i, std::min(chunk, npx - i), 4);
    }
    for (uint32_t i = 0; i < n; ++i) {
        b->stats[i].items = 0;
        for (int ph = 0; ph < kMaxPhase; ++ph) b->stats[i].items += need[(size_t)i * kMaxPhase + ph];
        b->stats[i].lanes = (uint32_t)kClassLayouts[pd[i].cls][kBatchLayout[pd[i].cls]][0];
        b->stats[i].per_lane = (uint32_t)kClassLayouts[pd[i].cls][kBatchLayout[pd[i].cls]][1];
    }
    // One persistent launch per (volume format, D % 8 == 0) class of pairs:
    // the list is partitioned stably, so each launch keeps the order above.
    auto launch_of = [&](int p) {
        return (pd[p].mode == kOutU8 ? 0 : 2) + (pd[p].D % 8 == 0 ? 0 : 1);
    };
    std::vector<Item> sorted;
    int lstart[kLaunchKinds + 1] = {};
    for (int k = 0; k < kLaunchKinds; ++k) {
        lstart[k] = (int)sorted.size();
        for (const Item &it : items)
            if (launch_of(it.pair) == k) sorted.push_back(it);
    }
    lstart[kLaunchKinds] = (int)sorted.size();
    items.swap(sorted);
    b->nitems = (int)items.size();
    b->ncounters = kLaunchKinds + 1 + (int)n * kMaxPhase;
    const size_t meta = align256(n * sizeof(PairDesc)) + align256(items.size() * sizeof(Item)) +
                        align256(need.size() * 4) + align256((size_t)b->ncounters * 4);
    if (hipMalloc(&b->meta, meta) != hipSuccess) {
        vxh_sgm_batch_destroy(b);
        return bfail(VXH_ERR_BUFFER_ALLOC, "buffer allocation failed (" +
                                               std::to_string(meta) + " bytes)");
    }
    char *m = (char *)b->meta;
    b->d_pairs = (PairDesc *)m;
    m += align256(n * sizeof(PairDesc));
    b->d_items = (Item *)m;
    m += align256(items.size() * sizeof(Item));
    b->d_need = (uint32_t *)m;
    m += align256(need.size() * 4);
    b->d_counters = (uint32_t *)m;
    hipError_t e = hipMemcpy(b->d_pairs, pd.data(), n * sizeof(PairDesc), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(b->d_items, items.data(), items.size() * sizeof(Item), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(b->d_need, need.data(), need.size() * 4, hipMemcpyHostToDevice);
    int cus = 0;
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        vxh_sgm_batch_destroy(b);
        return bfail(VXH_ERR_GPU, std::string("GPU error: batch setup: ") + hipGetErrorString(e));
    }
    // Persistent grid: blocks_per_cu 256-thread workgroups per CU (the kernel's
    // registers allow two).
    b->grid = std::max(1, cus) * (int)o.blocks_per_cu;

    // Capture [reset counters, persistent kernel] once; replays reuse it.
    e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(batch_reset_kernel, dim3((b->ncounters + 255) / 256), dim3(256), 0,
                           b->stream, b->d_counters, b->ncounters);
        uint32_t *const err = b->d_counters + kLaunchKinds;
        uint32_t *const done = b->d_counters + kLaunchKinds + 1;
        for (int k = 0; k < kLaunchKinds; ++k) {
            const int cnt = lstart[k + 1] - lstart[k];
            if (cnt == 0) continue;
            const Item *its = b->d_items + lstart[k];
            uint32_t *next = b->d_counters + k;
#define VXH_BATCH_LAUNCH(MM, AL)                                                             \
    hipLaunchKernelGGL((batch_kernel<MM, AL>), dim3(b->grid), dim3(256), 0, b->stream,       \
                       b->d_pairs, its, cnt, next, done, b->d_need, err)
            if (k == 0) VXH_BATCH_LAUNCH(kOutU8, true);
            else if (k == 1) VXH_BATCH_LAUNCH(kOutU8, false);
            else if (k == 2) VXH_BATCH_LAUNCH(kOutU16, true);
            else VXH_BATCH_LAUNCH(kOutU16, false);
#undef VXH_BATCH_LAUNCH
            ++b->nlaunch;
        }
        e = hipStreamEndCapture(b->stream, &b->graph);
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        vxh_sgm_batch_destroy(b);
        return bfail(VXH_ERR_GPU, std::string("GPU error: batch graph: ") + hipGetErrorString(e));
    }
    *out = b;
    return VXH_OK;
}

vxh_status vxh_sgm_batch_launch(vxh_sgm_batch *b, void *stream) {
    if (!b) return bfail(VXH_ERR_INVALID_CONFIG, "invalid config: null batch");
    if (hipSetDevice(b->device) != hipSuccess) return bfail(VXH_ERR_DEVICE_NOT_FOUND, "no HIP device found");
    hipError_t e = hipGraphLaunch(b->exec, stream ? (hipStream_t)stream : b->stream);
    if (e != hipSuccess)
        return bfail(VXH_ERR_GPU, std::string("GPU error: graph launch: ") + hipGetErrorString(e));
    return VXH_OK;
}

vxh_status vxh_sgm_batch_run(vxh_sgm_batch *b) {
    vxh_status st = vxh_sgm_batch_launch(b, nullptr);
    if (st != VXH_OK) return st;
    hipError_t e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess)
        return bfail(VXH_ERR_GPU, std::string("GPU error: batch sync: ") + hipGetErrorString(e));
    return vxh_sgm_batch_check(b);
}

vxh_status vxh_sgm_batch_check(vxh_sgm_batch *b) {
    if (!b) return bfail(VXH_ERR_INVALID_CONFIG, "invalid config: null batch");
    uint32_t err = 0;
    hipError_t e = hipMemcpy(&err, b->d_counters + vxh::kLaunchKinds, 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess)
        return bfail(VXH_ERR_GPU, std::string("GPU error: batch check: ") + hipGetErrorString(e));
    if (err)
        return bfail(VXH_ERR_GPU, "GPU error: a batch item's dependency wait timed out");
    return VXH_OK;
}

vxh_status vxh_sgm_batch_stats_get(const vxh_sgm_batch *b, uint32_t pair,
                                   vxh_sgm_batch_stats *out) {
    if (!b || !out || pair >= (uint32_t)b->npairs)
        return bfail(VXH_ERR_INVALID_CONFIG, "invalid config: batch stats");
    *out = b->stats[pair];
    return VXH_OK;
}

uint32_t vxh_sgm_batch_items(const vxh_sgm_batch *b) { return b ? (uint32_t)b->nitems : 0; }

void vxh_sgm_batch_destroy(vxh_sgm_batch *b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    if (b->meta) (void)hipFree(b->meta);
    if (b->slab) (void)hipFree(b->slab);
    delete b;
}

}  // extern "C"
