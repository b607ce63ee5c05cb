// Batched SGM: many pairs of uneven size and disparity range replayed from
// one hipGraph.
//
// The reference runs one pair per command buffer (vx-vision/src/kernels/
// sgm.rs:135-261) and its Pipeline batches encodes of different kernels into
// one command buffer (vx-vision/src/pipeline.rs:12-80).  Here a batch is the
// per-pair pipeline (census, costs, aggregation, sum + winner, LR check) of
// every pair, captured once into a graph over `streams` streams: pairs are
// ordered by work (px * disparities, largest first) and dealt round-robin to
// the streams, which run side by side, so that one pair's HBM-bound passes
// (census, costs, sums) overlap another's instruction-bound aggregation.
// Pairs on one stream run one after the other and share that stream's
// workspace, sized for the largest of them.  Every pair runs exactly the
// plain path's kernels with its own schedule and volume format, so its map
// is bit-identical to the plain path's.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <string>
#include <vector>

#include "sgm_internal.h"
#include "sgm_pipeline.h"
#include "vxh/sgm.h"
#include "vxh/sgm_batch.h"

namespace {

thread_local std::string g_batch_error;

vxh_status bfail(vxh_status st, const std::string &m) {
    g_batch_error = m;
    return st;
}

constexpr uint32_t kDefaultStreams = 4;
constexpr uint32_t kMaxStreams = 16;

// Index of (L, K) among the layouts of class c, -1 if none.
int layout_index(int c, uint32_t L, uint32_t K) {
    for (int i = 0; i < vxh::kClassCount[c]; ++i)
        if ((uint32_t)vxh::kClassLayouts[c][i][0] == L &&
            (uint32_t)vxh::kClassLayouts[c][i][1] == K)
            return i;
    return -1;
}

struct PairPlan {
    vxh::Schedule sched;
    int mode;
    size_t ws_bytes;
    uint32_t stream;
};

}  // namespace

struct vxh_sgm_batch {
    int device = 0;
    std::vector<vxh_sgm_pair_desc> pairs;
    std::vector<PairPlan> plan;
    std::vector<void *> ws;              // one workspace per stream
    uint32_t *d_error = nullptr;         // OR of every pair's error word
    hipStream_t stream = nullptr;        // the batch's own stream (capture origin)
    std::vector<hipStream_t> sub;        // the streams the pairs run on
    std::vector<hipEvent_t> events;      // fork + one join per stream
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint32_t launches = 0;
};

extern "C" {

const char *vxh_sgm_batch_last_error(void) { return g_batch_error.c_str(); }

vxh_status vxh_sgm_batch_create(int device, const vxh_sgm_pair_desc *pairs, uint32_t n,
                                vxh_sgm_batch **out) {
    return vxh_sgm_batch_create_ex(device, pairs, n, nullptr, out);
}

vxh_status vxh_sgm_batch_create_ex(int device, const vxh_sgm_pair_desc *pairs, uint32_t n,
                                   const vxh_sgm_batch_options *opt, vxh_sgm_batch **out) {
    using namespace vxh;
    if (!out || (!pairs && n)) return bfail(VXH_ERR_INVALID_CONFIG, "invalid config: null argument");
    *out = nullptr;
    uint32_t streams = kDefaultStreams;
    if (opt) {
        if (opt->streams > kMaxStreams)
            return bfail(VXH_ERR_INVALID_CONFIG, "invalid config: batch options");
        if (opt->streams) streams = opt->streams;
    }
    if (n == 0 || n > 65535) return bfail(VXH_ERR_INVALID_CONFIG, "invalid config: batch size");
    if (hipSetDevice(device) != hipSuccess)
        return bfail(VXH_ERR_DEVICE_NOT_FOUND, "no HIP device found");
    streams = std::min(streams, n);

    std::vector<PairPlan> plan(n);
    for (uint32_t i = 0; i < n; ++i) {
        const vxh_sgm_pair_desc &q = pairs[i];
        if (vxh_sgm_config_validate(&q.cfg, q.width, q.height) != VXH_OK)
            return bfail(VXH_ERR_INVALID_CONFIG,
                         std::string("pair ") + std::to_string(i) + ": " + vxh_last_error());
        if (!q.d_left || !q.d_right || !q.d_out)
            return bfail(VXH_ERR_INVALID_CONFIG, "invalid config: null pair buffer");
        if (q.volume_format > 2 || q.split > 1)
            return bfail(VXH_ERR_INVALID_CONFIG, "invalid config: pair options");
        const int D = (int)q.cfg.num_disparities;
        const int cls = layout_class(D);
        Schedule sc = default_schedule(D);
        if (q.lanes || q.per_lane) sc.other = layout_index(cls, q.lanes, q.per_lane);
        if (q.lanes_h || q.per_lane_h) sc.horiz = layout_index(cls, q.lanes_h, q.per_lane_h);
        if (sc.other < 0 || sc.horiz < 0)
            return bfail(VXH_ERR_INVALID_CONFIG,
                         "invalid config: unsupported lane layout (pair " + std::to_string(i) +
                             ", D=" + std::to_string(D) + ")");
        sc.split = (int)q.split;
        plan[i].sched = sc;
        plan[i].mode = q.volume_format == 1 ? kOutU8
                       : q.volume_format == 2 ? kOutU16
                                              : auto_volume_mode(q.cfg.p2);
        plan[i].ws_bytes = workspace_layout(q.width, q.height, (uint32_t)D,
                                            (int)q.cfg.num_paths, plan[i].mode,
                                            q.cfg.lr_check != 0)
                               .total;
    }
    // Largest pairs first, dealt round-robin to the streams.
    std::vector<uint32_t> order(n);
    for (uint32_t i = 0; i < n; ++i) order[i] = i;
    auto work = [&](uint32_t i) {
        return (size_t)pairs[i].width * pairs[i].height * pairs[i].cfg.num_disparities;
    };
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t b) { return work(a) > work(b); });
    std::vector<size_t> ws_bytes(streams, 0);
    for (uint32_t k = 0; k < n; ++k) {
        PairPlan &p = plan[order[k]];
        p.stream = k % streams;
        ws_bytes[p.stream] = std::max(ws_bytes[p.stream], p.ws_bytes);
    }

    vxh_sgm_batch *b = new vxh_sgm_batch();
    b->device = device;
    b->pairs.assign(pairs, pairs + n);
    b->plan = plan;
    auto gpu_fail = [&](const char *what, hipError_t e) {
        vxh_sgm_batch_destroy(b);
        if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
            return bfail(VXH_ERR_BUFFER_ALLOC, std::string("buffer allocation failed: ") + what);
        return bfail(VXH_ERR_GPU, std::string("GPU error: ") + what + ": " + hipGetErrorString(e));
    };
    hipError_t e = hipSuccess;
    for (uint32_t s = 0; s < streams && e == hipSuccess; ++s) {
        void *p = nullptr;
        e = hipMalloc(&p, ws_bytes[s]);
        if (e == hipSuccess) b->ws.push_back(p);
    }
    if (e != hipSuccess) return gpu_fail("batch workspaces", e);
    if ((e = hipMalloc((void **)&b->d_error, 256)) != hipSuccess)
        return gpu_fail("batch error word", e);
    if ((e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess)
        return gpu_fail("stream", e);
    for (uint32_t s = 0; s < streams; ++s) {
        hipStream_t st = nullptr;
        if ((e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) != hipSuccess)
            return gpu_fail("stream", e);
        b->sub.push_back(st);
    }
    for (uint32_t s = 0; s <= streams; ++s) {
        hipEvent_t ev = nullptr;
        if ((e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess)
            return gpu_fail("event", e);
        b->events.push_back(ev);
    }

    // Capture: clear the error word, fork to every stream, each pair's
    // pipeline on its stream followed by folding its error word into the
    // batch's, join.
    e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return gpu_fail("graph capture", e);
    std::string what;
    vxh_status st = VXH_OK;
    e = launch_zero(b->d_error, 1, b->stream);
    if (e == hipSuccess) e = hipEventRecord(b->events[0], b->stream);
    for (uint32_t s = 0; s < streams && e == hipSuccess; ++s)
        e = hipStreamWaitEvent(b->sub[s], b->events[0], 0);
    for (uint32_t k = 0; k < n && e == hipSuccess && st == VXH_OK; ++k) {
        const uint32_t i = order[k];
        const vxh_sgm_pair_desc &q = pairs[i];
        const PairPlan &p = plan[i];
        char *ws = (char *)b->ws[p.stream];
        st = run_pipeline(q.d_left, q.d_right, q.width, q.height, &q.cfg, q.d_out, ws, p.mode,
                          p.sched, b->sub[p.stream]);
        if (st != VXH_OK) {
            what = vxh_last_error();
            break;
        }
        e = launch_or_word(b->d_error, (const uint32_t *)(ws + kErrorWordOffset),
                           b->sub[p.stream]);
        b->launches += 6 + (q.cfg.lr_check ? 1 : 0);
    }
    for (uint32_t s = 0; s < streams && e == hipSuccess && st == VXH_OK; ++s) {
        e = hipEventRecord(b->events[1 + s], b->sub[s]);
        if (e == hipSuccess) e = hipStreamWaitEvent(b->stream, b->events[1 + s], 0);
    }
    const hipError_t e2 = hipStreamEndCapture(b->stream, &b->graph);
    if (st != VXH_OK) {
        vxh_sgm_batch_destroy(b);
        return bfail(st, "batch capture: " + what);
    }
    if (e == hipSuccess) e = e2;
    if (e == hipSuccess) e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) return gpu_fail("batch graph", e);
    *out = b;
    return VXH_OK;
}

vxh_status vxh_sgm_batch_launch(vxh_sgm_batch *b, void *stream) {
    if (!b) return bfail(VXH_ERR_INVALID_CONFIG, "invalid config: null batch");
    if (hipSetDevice(b->device) != hipSuccess)
        return bfail(VXH_ERR_DEVICE_NOT_FOUND, "no HIP device found");
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
    hipError_t e = hipMemcpy(&err, b->d_error, 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess)
        return bfail(VXH_ERR_GPU, std::string("GPU error: batch check: ") + hipGetErrorString(e));
    if (err) return bfail(VXH_ERR_GPU, "GPU error: a split scanline's wait timed out");
    return VXH_OK;
}

vxh_status vxh_sgm_batch_stats_get(const vxh_sgm_batch *b, uint32_t pair,
                                   vxh_sgm_batch_stats *out) {
    if (!b || !out || pair >= (uint32_t)b->pairs.size())
        return bfail(VXH_ERR_INVALID_CONFIG, "invalid config: batch stats");
    const vxh_sgm_pair_desc &q = b->pairs[pair];
    const PairPlan &p = b->plan[pair];
    const int c = vxh::layout_class((int)q.cfg.num_disparities);
    out->stream = p.stream;
    out->lanes = (uint32_t)vxh::kClassLayouts[c][p.sched.other][0];
    out->per_lane = (uint32_t)vxh::kClassLayouts[c][p.sched.other][1];
    out->lanes_h = (uint32_t)vxh::kClassLayouts[c][p.sched.horiz][0];
    out->per_lane_h = (uint32_t)vxh::kClassLayouts[c][p.sched.horiz][1];
    out->volume_format = p.mode == vxh::kOutU8 ? 1u : 2u;
    out->workspace_bytes = p.ws_bytes;
    return VXH_OK;
}

uint32_t vxh_sgm_batch_items(const vxh_sgm_batch *b) { return b ? b->launches : 0; }

void vxh_sgm_batch_destroy(vxh_sgm_batch *b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (hipStream_t s : b->sub) (void)hipStreamSynchronize(s);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (hipEvent_t ev : b->events) (void)hipEventDestroy(ev);
    for (hipStream_t s : b->sub) (void)hipStreamDestroy(s);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    for (void *p : b->ws) (void)hipFree(p);
    if (b->d_error) (void)hipFree(b->d_error);
    delete b;
}

}  // extern "C"
