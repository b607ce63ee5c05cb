// Pipelines and the texture pool behind include/vxh/pipeline.h.
//
//   vxh_pipeline      vx-vision/src/pipeline.rs:12-80.  The reference records
//                     dispatches into one Metal command buffer and runs them on
//                     commit.  Here begin() opens a stream capture on a stream
//                     of the context's pool; encodes launch onto that stream,
//                     so they are recorded, not run; commit() ends the capture,
//                     instantiates one HIP graph and launches it.  The graph's
//                     first node clears the pipeline's error word and every
//                     encoded SGM pass folds its own error word into it, so
//                     wait() reports an expired bounded wait of any of them.
//   vxh_texture_pool  vx-vision/src/pool.rs:17-147: cached textures per (width,
//                     height, format), plus the device, so a pool shared by
//                     contexts on several GPUs never hands out another GPU's
//                     memory.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "sgm_internal.h"
#include "sgm_pipeline.h"
#include "vxh/pipeline.h"

namespace {

vxh_status gpu_error(const char *what, hipError_t e) {
    return vxh::report_error(VXH_ERR_GPU,
                             std::string("GPU error: ") + what + ": " + hipGetErrorString(e));
}

vxh_status invalid(const char *msg) {
    return vxh::report_error(VXH_ERR_INVALID_CONFIG, std::string("invalid config: ") + msg);
}

}  // namespace

struct vxh_pipeline {
    vxh_context *ctx = nullptr;
    int device = 0;
    hipStream_t stream = nullptr;
    bool capturing = false;    // between begin and commit
    bool committed = false;
    bool launched = false;     // the graph was instantiated and launched
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    size_t nodes = 0;
    uint32_t *d_error = nullptr;       // the start of scratch[0]
    std::vector<void *> scratch;       // context pool blocks, held until destroy
    std::vector<vxh_texture *> retained;
    // Recorded on the stream after every launch of the graph (the textures the
    // encodes used are in flight until it completes); null before the first.
    std::atomic<hipEvent_t> done{nullptr};
};

namespace vxh {

vxh_status pipeline_stream(vxh_pipeline *p, hipStream_t *st) {
    if (!p) return invalid("null pipeline");
    if (!p->capturing) return invalid("pipeline already committed");
    if (hipSetDevice(p->device) != hipSuccess)
        return report_error(VXH_ERR_DEVICE_NOT_FOUND, "no HIP device found");
    *st = p->stream;
    return VXH_OK;
}

vxh_status pipeline_scratch(vxh_pipeline *p, size_t bytes, void **ptr) {
    if (!p->capturing) return invalid("pipeline already committed");
    vxh_status v = ctx_take_block(p->ctx, bytes, ptr, nullptr);
    if (v == VXH_OK) p->scratch.push_back(*ptr);
    return v;
}

vxh_status pipeline_watch_word(vxh_pipeline *p, const uint32_t *d_word) {
    const hipError_t e = launch_or_word(p->d_error, d_word, p->stream);
    return e == hipSuccess ? VXH_OK : gpu_error("pipeline error word", e);
}

vxh_context *pipeline_context(vxh_pipeline *p) { return p ? p->ctx : nullptr; }

hipEvent_t pipeline_done_event(vxh_pipeline *p) { return p ? p->done.load() : nullptr; }

}  // namespace vxh

extern "C" {

void vxh_pipeline_destroy(vxh_pipeline *p) {
    if (!p) return;
    vxh::CaptureRelaxed relaxed;
    (void)hipSetDevice(p->device);
    if (p->capturing) {
        // Never committed: discard what was recorded; nothing of it ran.
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(p->stream, &g);
        vxh::capture_closed();
        if (g) (void)hipGraphDestroy(g);
    } else if (p->launched) {
        (void)hipStreamSynchronize(p->stream);
    }
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    vxh::pipeline_forget(p);
    if (hipEvent_t ev = p->done.exchange(nullptr)) (void)hipEventDestroy(ev);
    for (void *b : p->scratch) vxh::ctx_give_block(p->ctx, b);
    if (p->stream) vxh::ctx_give_stream(p->ctx, p->stream);
    for (vxh_texture *t : p->retained) vxh_texture_destroy(t);
    delete p;
}

vxh_status vxh_pipeline_begin(vxh_context *ctx, vxh_pipeline **out) {
    if (!ctx || !out) return invalid("null argument");
    *out = nullptr;
    vxh_pipeline *p = new vxh_pipeline();
    p->ctx = ctx;
    p->device = vxh_context_device(ctx);
    if (hipSetDevice(p->device) != hipSuccess) {
        delete p;
        return vxh::report_error(VXH_ERR_DEVICE_NOT_FOUND, "no HIP device found");
    }
    vxh_status v = vxh::ctx_take_stream(ctx, &p->stream);
    if (v != VXH_OK) {
        delete p;
        return v;
    }
    void *blk = nullptr;
    if ((v = vxh::ctx_take_block(ctx, 256, &blk, nullptr)) != VXH_OK) {
        vxh_pipeline_destroy(p);
        return v;
    }
    p->scratch.push_back(blk);
    p->d_error = (uint32_t *)blk;
    // Relaxed: the capture may be ended from another thread, and allocating
    // calls elsewhere in the process stay legal while it is open.
    hipError_t e = hipStreamBeginCapture(p->stream, hipStreamCaptureModeRelaxed);
    if (e != hipSuccess) {
        vxh_pipeline_destroy(p);
        return gpu_error("pipeline capture", e);
    }
    p->capturing = true;
    vxh::capture_opened();
    // First node: clear the error word, so every launch of the graph starts
    // from zero (a kernel, not a memset node).
    if ((e = vxh::launch_zero(p->d_error, 1, p->stream)) != hipSuccess) {
        vxh_pipeline_destroy(p);
        return gpu_error("pipeline capture", e);
    }
    *out = p;
    return VXH_OK;
}

void *vxh_pipeline_stream(const vxh_pipeline *p) {
    return p && p->capturing ? (void *)p->stream : nullptr;
}

vxh_status vxh_pipeline_retain(vxh_pipeline *p, vxh_texture *tex) {
    if (!p || !tex) return invalid("null argument");
    p->retained.push_back(tex);
    return VXH_OK;
}

vxh_status vxh_pipeline_commit(vxh_pipeline *p) {
    if (!p) return invalid("null pipeline");
    if (p->committed) return VXH_OK;
    vxh::CaptureRelaxed relaxed;
    if (hipSetDevice(p->device) != hipSuccess)
        return vxh::report_error(VXH_ERR_DEVICE_NOT_FOUND, "no HIP device found");
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(p->stream, &g);
    p->capturing = false;
    vxh::capture_closed();
    p->committed = true;
    p->graph = g;
    if (e != hipSuccess || !g) return gpu_error("pipeline capture", e);
    size_t n = 0;
    if ((e = hipGraphGetNodes(g, nullptr, &n)) != hipSuccess)
        return gpu_error("pipeline graph", e);
    p->nodes = n;
    if ((e = hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0)) != hipSuccess) {
        p->exec = nullptr;
        return gpu_error("pipeline graph", e);
    }
    hipEvent_t ev = nullptr;
    if ((e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess)
        return gpu_error("pipeline event", e);
    p->done.store(ev);
    if ((e = hipGraphLaunch(p->exec, p->stream)) != hipSuccess)
        return gpu_error("pipeline launch", e);
    p->launched = true;
    if ((e = hipEventRecord(ev, p->stream)) != hipSuccess)
        return gpu_error("pipeline event", e);
    return VXH_OK;
}

vxh_status vxh_pipeline_wait(vxh_pipeline *p) {
    if (!p) return invalid("null pipeline");
    if (!p->launched) return VXH_OK;
    vxh::CaptureRelaxed relaxed;
    if (hipSetDevice(p->device) != hipSuccess)
        return vxh::report_error(VXH_ERR_DEVICE_NOT_FOUND, "no HIP device found");
    hipError_t e = hipStreamSynchronize(p->stream);
    if (e != hipSuccess) return gpu_error("pipeline wait", e);
    uint32_t err = 0;
    if ((e = hipMemcpyWithStream(&err, p->d_error, 4, hipMemcpyDeviceToHost, p->stream)) !=
        hipSuccess)
        return gpu_error("pipeline error word", e);
    if (err)
        return vxh::report_error(VXH_ERR_GPU, "GPU error: a split scanline's wait timed out");
    return VXH_OK;
}

vxh_status vxh_pipeline_commit_and_wait(vxh_pipeline *p) {
    vxh_status v = vxh_pipeline_commit(p);
    return v == VXH_OK ? vxh_pipeline_wait(p) : v;
}

vxh_status vxh_pipeline_relaunch(vxh_pipeline *p) {
    if (!p) return invalid("null pipeline");
    if (!p->launched) return invalid("pipeline not committed");
    if (hipSetDevice(p->device) != hipSuccess)
        return vxh::report_error(VXH_ERR_DEVICE_NOT_FOUND, "no HIP device found");
    hipError_t e = hipGraphLaunch(p->exec, p->stream);
    if (e == hipSuccess) e = hipEventRecord(p->done.load(), p->stream);
    return e == hipSuccess ? VXH_OK : gpu_error("pipeline launch", e);
}

size_t vxh_pipeline_take_retained(vxh_pipeline *p, vxh_texture **out, size_t cap) {
    if (!p) return 0;
    if (!out) return p->retained.size();
    const size_t n = std::min(cap, p->retained.size());
    std::copy(p->retained.begin(), p->retained.begin() + n, out);
    p->retained.erase(p->retained.begin(), p->retained.begin() + n);
    return n;
}

int vxh_pipeline_is_committed(const vxh_pipeline *p) { return p && p->committed ? 1 : 0; }

size_t vxh_pipeline_node_count(const vxh_pipeline *p) { return p ? p->nodes : 0; }

}  // extern "C"

// ---- TexturePool ----

struct vxh_texture_pool {
    using Key = std::tuple<int, uint32_t, uint32_t, uint32_t>;   // device, w, h, format
    size_t max_per_bucket = 8;
    std::map<Key, std::vector<vxh_texture *>> buckets;
    uint64_t acquires = 0;
    uint64_t hits = 0;
    mutable std::mutex mu;
};

extern "C" {

vxh_status vxh_texture_pool_create(size_t max_per_bucket, vxh_texture_pool **out) {
    if (!out) return invalid("null argument");
    vxh_texture_pool *p = new vxh_texture_pool();
    p->max_per_bucket = max_per_bucket;
    *out = p;
    return VXH_OK;
}

void vxh_texture_pool_clear(vxh_texture_pool *pool) {
    if (!pool) return;
    std::lock_guard<std::mutex> g(pool->mu);
    for (auto &kv : pool->buckets)
        for (vxh_texture *t : kv.second) vxh_texture_destroy(t);
    pool->buckets.clear();
}

void vxh_texture_pool_destroy(vxh_texture_pool *pool) {
    if (!pool) return;
    vxh_texture_pool_clear(pool);
    delete pool;
}

vxh_status vxh_texture_pool_acquire(vxh_texture_pool *pool, vxh_context *ctx, uint32_t w,
                                    uint32_t h, uint32_t format, vxh_texture **out) {
    if (!pool || !ctx || !out) return invalid("null argument");
    *out = nullptr;
    if (format < 1 || format > 3) return invalid("texture format");
    {
        std::lock_guard<std::mutex> g(pool->mu);
        ++pool->acquires;
        auto it = pool->buckets.find(vxh_texture_pool::Key{vxh_context_device(ctx), w, h, format});
        if (it != pool->buckets.end() && !it->second.empty()) {
            ++pool->hits;
            *out = it->second.back();
            it->second.pop_back();
            return VXH_OK;
        }
    }
    switch (format) {
        case 1: return vxh_texture_output_gray8(ctx, w, h, out);
        case 2: return vxh_texture_output_r32float(ctx, w, h, out);
        default: return vxh_texture_output_rgba8(ctx, w, h, out);
    }
}

vxh_status vxh_texture_pool_acquire_gray8(vxh_texture_pool *pool, vxh_context *ctx,
                                          uint32_t w, uint32_t h, vxh_texture **out) {
    return vxh_texture_pool_acquire(pool, ctx, w, h, 1, out);
}

vxh_status vxh_texture_pool_acquire_r32float(vxh_texture_pool *pool, vxh_context *ctx,
                                             uint32_t w, uint32_t h, vxh_texture **out) {
    return vxh_texture_pool_acquire(pool, ctx, w, h, 2, out);
}

vxh_status vxh_texture_pool_acquire_rgba8(vxh_texture_pool *pool, vxh_context *ctx,
                                          uint32_t w, uint32_t h, vxh_texture **out) {
    return vxh_texture_pool_acquire(pool, ctx, w, h, 3, out);
}

void vxh_texture_pool_release(vxh_texture_pool *pool, vxh_texture *tex) {
    if (!tex) return;
    if (!pool) {
        vxh_texture_destroy(tex);
        return;
    }
    {
        std::lock_guard<std::mutex> g(pool->mu);
        std::vector<vxh_texture *> &b =
            pool->buckets[vxh_texture_pool::Key{vxh_texture_device(tex), vxh_texture_width(tex),
                                                vxh_texture_height(tex),
                                                vxh_texture_format(tex)}];
        if (b.size() < pool->max_per_bucket) {
            b.push_back(tex);
            return;
        }
    }
    vxh_texture_destroy(tex);   // the bucket is full: dropped, as the reference does
}

size_t vxh_texture_pool_cached_count(const vxh_texture_pool *pool) {
    if (!pool) return 0;
    std::lock_guard<std::mutex> g(pool->mu);
    size_t n = 0;
    for (const auto &kv : pool->buckets) n += kv.second.size();
    return n;
}

uint64_t vxh_texture_pool_total_acquires(const vxh_texture_pool *pool) {
    if (!pool) return 0;
    std::lock_guard<std::mutex> g(pool->mu);
    return pool->acquires;
}

uint64_t vxh_texture_pool_total_hits(const vxh_texture_pool *pool) {
    if (!pool) return 0;
    std::lock_guard<std::mutex> g(pool->mu);
    return pool->hits;
}

double vxh_texture_pool_hit_rate(const vxh_texture_pool *pool) {
    if (!pool) return 0.0;
    std::lock_guard<std::mutex> g(pool->mu);
    return pool->acquires == 0 ? 0.0 : (double)pool->hits / (double)pool->acquires;
}

}  // extern "C"
