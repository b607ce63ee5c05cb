// Host runtime behind include/vxh/sgm.h: contexts, device textures, the SGM
// handle with its scratch-slot pool, and the stream-ordered pipeline.
//
// Reference counterparts:
//   vxh_context      vx-vision/src/context.rs:17-36, vx-core/src/device.rs:11-56;
//                    also owns the device pool (scratch blocks and streams)
//                    that pipelines draw from
//   vxh_texture      vx-vision/src/texture.rs:29-462 (R8Unorm, R32Float,
//                    RGBA8Unorm)
//   vxh_sgm          vx-vision/src/kernels/sgm.rs:61-288
//   device buffers   vx-core/src/buffer.rs:14-122 (UnifiedBuffer/GpuGuard): here
//                    explicit hipMalloc allocations owned by textures and by
//                    scratch slots, used in stream order; a slot is "in
//                    flight" while a call holds it and is returned only after
//                    that call's stream has drained.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "sgm_internal.h"
#include "sgm_pipeline.h"
#include "vxh/sgm.h"

namespace {

thread_local std::string g_last_error;

// Pipelines between begin and commit, process-wide.  While one is open, a
// blocking copy on the legacy null stream is refused by the runtime (it would
// depend on the capturing stream), so texture copies take a pooled stream.
std::atomic<int> g_open_captures{0};

vxh_status fail(vxh_status st, const std::string &msg) {
    g_last_error = msg;
    return st;
}

// Maps a HIP error to the reference's Error variants (error.rs:6-23).
vxh_status hip_fail(hipError_t e, const char *what, size_t bytes = 0) {
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) {
        char buf[128];
        snprintf(buf, sizeof buf, "buffer allocation failed (%zu bytes)", bytes);
        return fail(VXH_ERR_BUFFER_ALLOC, buf);
    }
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice)
        return fail(VXH_ERR_DEVICE_NOT_FOUND, "no HIP device found");
    return fail(VXH_ERR_GPU, std::string("GPU error: ") + what + ": " +
                                 hipGetErrorString(e));
}

#define VXH_HIP(call, what)                          \
    do {                                             \
        hipError_t e_ = (call);                      \
        if (e_ != hipSuccess) return hip_fail(e_, what); \
    } while (0)


enum TexFormat { kGray8 = 1, kR32F = 2, kRGBA8 = 3 };

size_t texel_bytes(int fmt) { return fmt == kGray8 ? 1 : 4; }

using vxh::WsLayout;
using vxh::workspace_layout;

int auto_mode(const vxh_sgm_config *c) { return vxh::auto_volume_mode(c->p2); }

// Workspace large enough for either format.
size_t workspace_total(uint32_t w, uint32_t h, const vxh_sgm_config *c) {
    return std::max(workspace_layout(w, h, c->num_disparities, (int)c->num_paths,
                                     vxh::kOutU8, c->lr_check != 0).total,
                    workspace_layout(w, h, c->num_disparities, (int)c->num_paths,
                                     vxh::kOutU16, c->lr_check != 0).total);
}

struct Slot {
    hipStream_t stream = nullptr;
    void *ws = nullptr;
    size_t ws_bytes = 0;
};

// Index of (lanes, per_lane) among the layouts of D's class, -1 if none.
int layout_index(int D, uint32_t lanes, uint32_t per_lane) {
    const int c = vxh::layout_class(D);
    for (int i = 0; i < vxh::kClassCount[c]; ++i)
        if ((uint32_t)vxh::kClassLayouts[c][i][0] == lanes &&
            (uint32_t)vxh::kClassLayouts[c][i][1] == per_lane)
            return i;
    return -1;
}

}  // namespace

struct vxh_context {
    struct Block {
        void *ptr;
        size_t bytes;
        bool busy;
    };
    int device = 0;
    std::mutex mu;
    std::vector<Block> blocks;
    std::vector<hipStream_t> idle_streams;
    std::vector<hipStream_t> all_streams;
};

struct vxh_texture {
    vxh_context *ctx;
    uint32_t w, h;
    int format;
    void *dptr;
    size_t bytes;
};

struct vxh_sgm {
    vxh_context *ctx;
    std::mutex mu;
    std::vector<Slot *> free_slots;
    std::vector<Slot *> all_slots;
    vxh::Schedule sched[257] = {};   // per D: layouts used when layout_set[D],
    bool layout_set[257] = {};       // the split level when split_set[D]
    bool split_set[257] = {};
    uint8_t format[257] = {};        // per D: 0 auto, 1 u8, 2 u16
    std::atomic<uint32_t> wait_polls{vxh::kSplitWaitPolls};   // split heads' bounded wait
};

namespace {

// (texture, pipeline) pairs recorded by encodes, process-wide.
std::mutex g_use_mu;
std::vector<std::pair<const vxh_texture *, vxh_pipeline *>> g_uses;

// In-flight state of `t` for a host copy: VXH_OK when no pipeline that used it
// has launched work still running; otherwise, with `wait`, waits for that work
// first, else refuses with INVALID_CONFIG.
std::vector<hipEvent_t> tex_pending(const vxh_texture *t) {
    std::vector<hipEvent_t> pending;
    std::lock_guard<std::mutex> g(g_use_mu);
    for (const auto &u : g_uses) {
        if (u.first != t) continue;
        hipEvent_t ev = vxh::pipeline_done_event(u.second);
        if (ev && hipEventQuery(ev) == hipErrorNotReady) pending.push_back(ev);
    }
    return pending;
}

vxh_status tex_settle(const vxh_texture *t, bool wait) {
    const std::vector<hipEvent_t> pending = tex_pending(t);
    if (pending.empty()) return VXH_OK;
    if (!wait)
        return fail(VXH_ERR_INVALID_CONFIG,
                    "invalid config: texture is in flight on the GPU (a launched pipeline that "
                    "uses it has not finished; wait for it first)");
    for (hipEvent_t ev : pending) {
        const hipError_t e = hipEventSynchronize(ev);
        if (e != hipSuccess) return hip_fail(e, "hipEventSynchronize");
    }
    return VXH_OK;
}

void tex_forget(const vxh_texture *t) {
    std::lock_guard<std::mutex> g(g_use_mu);
    g_uses.erase(std::remove_if(g_uses.begin(), g_uses.end(),
                                [&](const auto &u) { return u.first == t; }),
                 g_uses.end());
}

}  // namespace

namespace {

// Scratch of vxh_sgm_enqueue_static, the entry point for callers that cannot
// allocate: a file-scope device array (every GPU the library runs on holds
// its own copy, at its own address, until the process ends).  Host code never
// passes `scratch` itself: the name is a host-side stand-in, and
// resolve_static_scratch() looks up each device's address once
// (hipGetSymbolAddress with that device current) into g_scratch_ptr.
constexpr size_t kStaticScratchBytes = (size_t)96 << 20;
__attribute__((used)) static __device__ __align__(256) uint8_t scratch[kStaticScratchBytes];
std::mutex g_scratch_mu;
std::vector<void *> g_scratch_ptr;   // per device ordinal; null until resolved

// Looks up `device`'s copy of the array (the caller has made `device`
// current); once per device.
vxh_status resolve_static_scratch(int device) {
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if ((size_t)device >= g_scratch_ptr.size()) g_scratch_ptr.resize((size_t)device + 1, nullptr);
    if (g_scratch_ptr[(size_t)device]) return VXH_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(VXH_ERR_SHADER_MISSING,
                    std::string("shader missing: static scratch symbol: ") + hipGetErrorString(e));
    g_scratch_ptr[(size_t)device] = p;
    return VXH_OK;
}

void *static_scratch(int device) {
    std::lock_guard<std::mutex> g(g_scratch_mu);
    return (size_t)device < g_scratch_ptr.size() ? g_scratch_ptr[(size_t)device] : nullptr;
}

}  // namespace

namespace vxh {

void capture_opened() { g_open_captures.fetch_add(1); }
void capture_closed() { g_open_captures.fetch_sub(1); }

void pipeline_use(vxh_pipeline *p, const vxh_texture *t) {
    std::lock_guard<std::mutex> g(g_use_mu);
    for (const auto &u : g_uses)
        if (u.first == t && u.second == p) return;
    g_uses.emplace_back(t, p);
}

void pipeline_forget(vxh_pipeline *p) {
    std::lock_guard<std::mutex> g(g_use_mu);
    g_uses.erase(std::remove_if(g_uses.begin(), g_uses.end(),
                                [&](const auto &u) { return u.second == p; }),
                 g_uses.end());
}

}  // namespace vxh

// A blocking host<->device copy of texture bytes: hipMemcpy, or, while a
// pipeline is being encoded, hipMemcpyWithStream on a stream of ctx's pool.
// The latter is not ordered after work on the null stream.
static vxh_status copy_blocking(vxh_context *ctx, void *dst, const void *src, size_t n,
                                hipMemcpyKind kind) {
    vxh::CaptureRelaxed relaxed;
    VXH_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    if (g_open_captures.load() == 0) {
        VXH_HIP(hipMemcpy(dst, src, n, kind), "hipMemcpy");
        return VXH_OK;
    }
    hipStream_t st = nullptr;
    vxh_status v = vxh::ctx_take_stream(ctx, &st);
    if (v != VXH_OK) return v;
    const hipError_t e = hipMemcpyWithStream(dst, src, n, kind, st);
    vxh::ctx_give_stream(ctx, st);
    return e == hipSuccess ? VXH_OK : hip_fail(e, "hipMemcpyWithStream");
}

// The schedule of a w x h pair with D disparities and `paths` paths: what the
// handle was set to, else the default for that pair.
static vxh::Schedule schedule_for(vxh_sgm *s, int D, int paths, uint32_t w, uint32_t h) {
    std::lock_guard<std::mutex> g(s->mu);
    vxh::Schedule sc = vxh::plain_schedule(D, paths, w, h);
    if (s->layout_set[D]) {
        sc.horiz = s->sched[D].horiz;
        sc.other = s->sched[D].other;
    }
    if (s->split_set[D]) sc.split = s->sched[D].split;
    sc.wait_polls = s->wait_polls.load();
    return sc;
}

static int mode_for(vxh_sgm *s, const vxh_sgm_config *c) {
    int f;
    {
        std::lock_guard<std::mutex> g(s->mu);
        f = s->format[c->num_disparities];
    }
    return f == 1 ? vxh::kOutU8 : (f == 2 ? vxh::kOutU16 : auto_mode(c));
}

extern "C" {

const char *vxh_last_error(void) { return g_last_error.c_str(); }
const char *vxh_version(void) { return "vxh-sgm 0.1.0 (gfx950)"; }

void vxh_sgm_config_default(vxh_sgm_config *c) {
    c->num_disparities = 128;
    c->p1 = 10.0f;
    c->p2 = 120.0f;
    c->census_radius_x = 4;
    c->census_radius_y = 3;
    c->num_paths = 2;
    c->lr_check = 0;
    c->lr_tolerance = 1;
}

void vxh_sgm_config_new(uint32_t nd, vxh_sgm_config *c) {
    vxh_sgm_config_default(c);
    c->num_disparities = nd;
}

vxh_status vxh_sgm_config_validate(const vxh_sgm_config *c, uint32_t w,
                                   uint32_t h) {
    char buf[160];
    if (!c) return fail(VXH_ERR_INVALID_CONFIG, "invalid config: null config");
    if (w == 0 || h == 0)
        return fail(VXH_ERR_INVALID_CONFIG, "invalid config: empty image");
    if (c->num_disparities < 1 || c->num_disparities > 256) {
        snprintf(buf, sizeof buf,
                 "invalid config: num_disparities %u outside 1..256",
                 c->num_disparities);
        return fail(VXH_ERR_INVALID_CONFIG, buf);
    }
    if (c->num_disparities > w) {
        snprintf(buf, sizeof buf,
                 "invalid config: num_disparities %u exceeds width %u",
                 c->num_disparities, w);
        return fail(VXH_ERR_INVALID_CONFIG, buf);
    }
    uint64_t bits = (2ull * c->census_radius_x + 1) * (2ull * c->census_radius_y + 1) - 1;
    if (c->census_radius_x > 64 || c->census_radius_y > 64 || bits > 64) {
        snprintf(buf, sizeof buf,
                 "invalid config: census window %ux%u needs more than 64 bits",
                 2 * c->census_radius_x + 1, 2 * c->census_radius_y + 1);
        return fail(VXH_ERR_INVALID_CONFIG, buf);
    }
    if (!(c->p1 >= 0.0f && c->p1 < 65536.0f) || !(c->p2 >= 0.0f && c->p2 < 65536.0f))
        return fail(VXH_ERR_INVALID_CONFIG,
                    "invalid config: p1 and p2 must lie in [0, 65536)");
    if (c->num_paths != 2 && c->num_paths != 4 && c->num_paths != 8) {
        snprintf(buf, sizeof buf, "invalid config: num_paths %u not in {2,4,8}",
                 c->num_paths);
        return fail(VXH_ERR_INVALID_CONFIG, buf);
    }
    // Volume entries are addressed with 32-bit indices (as the reference's
    // shader does, SGMStereo.metal:164), so the padded volume must stay below
    // 2^32 entries.
    if ((uint64_t)w * h * (uint64_t)vxh::volume_stride((int)c->num_disparities) >
        0xFFFFFFFFull) {
        snprintf(buf, sizeof buf,
                 "invalid config: %ux%u with %u disparities exceeds 2^32 volume entries", w,
                 h, c->num_disparities);
        return fail(VXH_ERR_INVALID_CONFIG, buf);
    }
    return VXH_OK;
}

int vxh_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

vxh_status vxh_context_create(int device, vxh_context **out) {
    if (!out) return fail(VXH_ERR_INVALID_CONFIG, "invalid config: null out");
    *out = nullptr;
    int n = vxh_device_count();
    if (device < 0 || device >= n)
        return fail(VXH_ERR_DEVICE_NOT_FOUND, "no HIP device found");
    VXH_HIP(hipSetDevice(device), "hipSetDevice");
    vxh_context *c = new vxh_context();
    c->device = device;
    *out = c;
    return VXH_OK;
}

void vxh_context_destroy(vxh_context *ctx) {
    if (!ctx) return;
    vxh::CaptureRelaxed relaxed;
    (void)hipSetDevice(ctx->device);
    for (hipStream_t st : ctx->all_streams) {
        (void)hipStreamSynchronize(st);
        (void)hipStreamDestroy(st);
    }
    for (const vxh_context::Block &b : ctx->blocks) (void)hipFree(b.ptr);
    delete ctx;
}

vxh_status vxh_context_pool_stats(vxh_context *ctx, size_t *out3) {
    if (!ctx || !out3) return fail(VXH_ERR_INVALID_CONFIG, "invalid config: null argument");
    std::lock_guard<std::mutex> g(ctx->mu);
    size_t held = 0, used = 0;
    for (const vxh_context::Block &b : ctx->blocks) {
        held += b.bytes;
        if (b.busy) used += b.bytes;
    }
    out3[0] = ctx->blocks.size();
    out3[1] = held;
    out3[2] = used;
    return VXH_OK;
}

void vxh_context_pool_trim(vxh_context *ctx) {
    if (!ctx) return;
    vxh::CaptureRelaxed relaxed;
    std::lock_guard<std::mutex> g(ctx->mu);
    (void)hipSetDevice(ctx->device);
    std::vector<vxh_context::Block> keep;
    for (const vxh_context::Block &b : ctx->blocks) {
        if (b.busy)
            keep.push_back(b);
        else
            (void)hipFree(b.ptr);
    }
    ctx->blocks.swap(keep);
    for (hipStream_t st : ctx->idle_streams) {
        (void)hipStreamSynchronize(st);
        (void)hipStreamDestroy(st);
        ctx->all_streams.erase(std::find(ctx->all_streams.begin(), ctx->all_streams.end(), st));
    }
    ctx->idle_streams.clear();
}

int vxh_context_device(const vxh_context *ctx) { return ctx ? ctx->device : -1; }

static vxh_status tex_alloc(vxh_context *ctx, uint32_t w, uint32_t h, int fmt,
                            vxh_texture **out) {
    if (!ctx || !out) return fail(VXH_ERR_INVALID_CONFIG, "invalid config: null argument");
    *out = nullptr;
    if (w == 0 || h == 0)
        return fail(VXH_ERR_INVALID_CONFIG, "invalid config: empty texture");
    size_t bytes = (size_t)w * h * texel_bytes(fmt);
    vxh::CaptureRelaxed relaxed;
    VXH_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc", bytes);
    *out = new vxh_texture{ctx, w, h, fmt, p, bytes};
    return VXH_OK;
}

// Texture::from_gray8 / from_rgba8 (texture.rs:38-93, :238-294): a new
// texture of `fmt` holding `len` bytes of host pixels.
static vxh_status tex_upload(vxh_context *ctx, const uint8_t *pixels, size_t len, uint32_t w,
                             uint32_t h, int fmt, const char *name, vxh_texture **out) {
    if (out) *out = nullptr;
    const size_t want = (size_t)w * h * texel_bytes(fmt);
    if (!pixels || len != want) {
        char buf[160];
        snprintf(buf, sizeof buf,
                 "invalid config: expected %zu bytes for %ux%u %s texture, got %zu", want, w,
                 h, name, len);
        return fail(VXH_ERR_INVALID_CONFIG, buf);
    }
    vxh_status st = tex_alloc(ctx, w, h, fmt, out);
    if (st != VXH_OK) return st;
    if ((st = copy_blocking(ctx, (*out)->dptr, pixels, len, hipMemcpyHostToDevice)) != VXH_OK) {
        vxh_texture_destroy(*out);
        *out = nullptr;
    }
    return st;
}

vxh_status vxh_texture_gray8(vxh_context *ctx, const uint8_t *pixels, size_t len,
                             uint32_t w, uint32_t h, vxh_texture **out) {
    return tex_upload(ctx, pixels, len, w, h, kGray8, "R8", out);
}

vxh_status vxh_texture_rgba8(vxh_context *ctx, const uint8_t *pixels, size_t len,
                             uint32_t w, uint32_t h, vxh_texture **out) {
    return tex_upload(ctx, pixels, len, w, h, kRGBA8, "RGBA8", out);
}

vxh_status vxh_texture_output_gray8(vxh_context *ctx, uint32_t w, uint32_t h,
                                    vxh_texture **out) {
    return tex_alloc(ctx, w, h, kGray8, out);
}

vxh_status vxh_texture_output_rgba8(vxh_context *ctx, uint32_t w, uint32_t h,
                                    vxh_texture **out) {
    return tex_alloc(ctx, w, h, kRGBA8, out);
}

vxh_status vxh_texture_output_r32float(vxh_context *ctx, uint32_t w, uint32_t h,
                                       vxh_texture **out) {
    return tex_alloc(ctx, w, h, kR32F, out);
}

vxh_status vxh_texture_r32float(vxh_context *ctx, const float *data, size_t len, uint32_t w,
                                uint32_t h, vxh_texture **out) {
    if (out) *out = nullptr;
    if (!data || len != (size_t)w * h) {
        char buf[160];
        snprintf(buf, sizeof buf,
                 "invalid config: expected %zu floats for %ux%u R32Float texture, got %zu",
                 (size_t)w * h, w, h, len);
        return fail(VXH_ERR_INVALID_CONFIG, buf);
    }
    vxh_status st = tex_alloc(ctx, w, h, kR32F, out);
    if (st != VXH_OK) return st;
    if ((st = copy_blocking(ctx, (*out)->dptr, data, len * 4, hipMemcpyHostToDevice)) !=
        VXH_OK) {
        vxh_texture_destroy(*out);
        *out = nullptr;
    }
    return st;
}

uint32_t vxh_texture_format(const vxh_texture *t) { return t ? (uint32_t)t->format : 0u; }

// UnifiedBuffer::write refuses while in flight (vx-core/src/buffer.rs:59-72):
// so does a write into a texture that a launched, unfinished pipeline uses.
vxh_status vxh_texture_write_gray8(vxh_texture *t, const uint8_t *pixels, size_t len) {
    if (!t || !pixels || t->format != kGray8 || len != (size_t)t->w * t->h)
        return fail(VXH_ERR_INVALID_CONFIG, "invalid config: gray8 write size");
    vxh_status v = tex_settle(t, false);
    if (v != VXH_OK) return v;
    return copy_blocking(t->ctx, t->dptr, pixels, len, hipMemcpyHostToDevice);
}

// Reads wait for the launched pipelines that use the texture, so they return
// what those pipelines wrote.
vxh_status vxh_texture_read_r32float(const vxh_texture *t, float *dst, size_t len) {
    if (!t || t->format != kR32F)
        return fail(VXH_ERR_INVALID_CONFIG, "invalid config: not an R32Float texture");
    if (len != (size_t)t->w * t->h)
        return fail(VXH_ERR_INVALID_CONFIG, "invalid config: read length");
    vxh_status v = tex_settle(t, true);
    if (v != VXH_OK) return v;
    return copy_blocking(t->ctx, dst, t->dptr, len * 4, hipMemcpyDeviceToHost);
}

// Texture::read_gray8 / read_rgba8 (texture.rs:352-402).
static vxh_status tex_read_bytes(const vxh_texture *t, int fmt, uint8_t *dst, size_t len) {
    if (!t || t->format != fmt || !dst)
        return fail(VXH_ERR_INVALID_CONFIG, "invalid config: texture format");
    if (len != t->bytes) return fail(VXH_ERR_INVALID_CONFIG, "invalid config: read length");
    vxh_status v = tex_settle(t, true);
    if (v != VXH_OK) return v;
    return copy_blocking(t->ctx, dst, t->dptr, len, hipMemcpyDeviceToHost);
}

int vxh_texture_in_flight(const vxh_texture *t) { return t && !tex_pending(t).empty() ? 1 : 0; }

vxh_status vxh_texture_read_gray8(const vxh_texture *t, uint8_t *dst, size_t len) {
    return tex_read_bytes(t, kGray8, dst, len);
}

vxh_status vxh_texture_read_rgba8(const vxh_texture *t, uint8_t *dst, size_t len) {
    return tex_read_bytes(t, kRGBA8, dst, len);
}

int vxh_texture_device(const vxh_texture *t) { return t ? t->ctx->device : -1; }
uint32_t vxh_texture_width(const vxh_texture *t) { return t ? t->w : 0; }
uint32_t vxh_texture_height(const vxh_texture *t) { return t ? t->h : 0; }
void *vxh_texture_device_ptr(const vxh_texture *t) { return t ? t->dptr : nullptr; }

void vxh_texture_destroy(vxh_texture *t) {
    if (!t) return;
    tex_forget(t);
    vxh::CaptureRelaxed relaxed;
    (void)hipSetDevice(t->ctx->device);
    (void)hipFree(t->dptr);
    delete t;
}

vxh_status vxh_sgm_create(vxh_context *ctx, vxh_sgm **out) {
    if (!ctx || !out) return fail(VXH_ERR_INVALID_CONFIG, "invalid config: null argument");
    *out = nullptr;
    VXH_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    vxh_status v = resolve_static_scratch(ctx->device);
    if (v != VXH_OK) return v;
    vxh_sgm *s = new vxh_sgm();
    s->ctx = ctx;
    *out = s;
    return VXH_OK;
}

void vxh_sgm_destroy(vxh_sgm *s) {
    if (!s) return;
    (void)hipSetDevice(s->ctx->device);
    for (Slot *sl : s->all_slots) {
        if (sl->stream) {
            (void)hipStreamSynchronize(sl->stream);
            (void)hipStreamDestroy(sl->stream);
        }
        if (sl->ws) (void)hipFree(sl->ws);
        delete sl;
    }
    delete s;
}

size_t vxh_sgm_workspace_bytes(uint32_t w, uint32_t h, const vxh_sgm_config *c) {
    if (!c) return 0;
    return workspace_total(w, h, c);
}

// Aggregates every path of `c` at once, direction i into the i-th volume of
// the workspace (`mode` storage; the u8 flags must already be clear), then
// sums the volumes and takes the winner: into `out` (may be null), `best`
// and `rkey` (the LR check's integer winners and right-view keys, both or
// neither; the keys are cleared here) or, with `sum_out`, the summed u16
// volume.  With `mark`, records it on `st` between the two launches.
static vxh_status paths_and_wta(const uint8_t *cost, uint32_t w, uint32_t h,
                                const vxh_sgm_config *c, char *ws, const WsLayout &L,
                                int mode, vxh::Schedule sc, hipStream_t st, float *out,
                                int16_t *best, uint16_t *sum_out, uint32_t *rkey,
                                hipEvent_t mark = nullptr) {
    const int D = (int)c->num_disparities;
    const int P = (int)c->num_paths;
    vxh::PassParams base{};
    base.cost = cost;
    base.flags = mode == vxh::kOutU8 ? (uint8_t *)(ws + L.flags) : nullptr;
    base.w = (int)w;
    base.h = (int)h;
    base.D = D;
    base.Dp = vxh::volume_stride(D);
    const uint32_t p1 = (uint32_t)c->p1 & 0xFFFFu, p2 = (uint32_t)c->p2 & 0xFFFFu;
    base.p1x2 = p1 | (p1 << 16);
    base.p2x2 = p2 | (p2 << 16);
    const uint32_t xq = p1 - vxh::cost_bias(p1);
    base.xqx2 = xq | (xq << 16);
    base.wait_polls = sc.wait_polls;
    int dirs[vxh::kMaxPasses];
    uint8_t *vols[vxh::kMaxPasses], *his[vxh::kMaxPasses];
    for (int i = 0; i < P; ++i) {
        dirs[i] = i;
        vols[i] = (uint8_t *)(ws + L.vol + (size_t)i * L.vol_bytes);
        his[i] = mode == vxh::kOutU8 ? (uint8_t *)(ws + L.hi + (size_t)i * L.vol_bytes)
                                     : nullptr;
    }
    VXH_HIP(vxh::launch_paths(base, dirs, vols, his, P, sc, mode, (uint32_t *)(ws + L.sync), st),
            "aggregate");
    if (mark) VXH_HIP(hipEventRecord(mark, st), "hipEventRecord");
    const uint8_t *hi0 = mode == vxh::kOutU8 ? (const uint8_t *)(ws + L.hi) : nullptr;
    if (rkey) {
        VXH_HIP(vxh::launch_zero(rkey, (size_t)w * h, st), "zero");
        VXH_HIP(vxh::launch_wta_lr((const uint8_t *)(ws + L.vol), L.vol_bytes, hi0,
                                   (const uint8_t *)base.flags, P, mode, (int)w, (int)h, D,
                                   base.Dp, out, best, rkey, st),
                "wta");
    } else {
        VXH_HIP(vxh::launch_wta_sum((const uint8_t *)(ws + L.vol), L.vol_bytes, hi0,
                                    (const uint8_t *)base.flags, P, mode, (int)(w * h), D,
                                    base.Dp, out, best, sum_out, st),
                "wta");
    }
    return VXH_OK;
}

}  // extern "C"

namespace vxh {

vxh_status report_error(vxh_status st, const std::string &msg) { return fail(st, msg); }

vxh_status ctx_take_block(vxh_context *ctx, size_t bytes, void **ptr, size_t *got) {
    bytes = (bytes + 4095) & ~(size_t)4095;
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        vxh_context::Block *best = nullptr;
        for (vxh_context::Block &b : ctx->blocks)
            if (!b.busy && b.bytes >= bytes && (!best || b.bytes < best->bytes)) best = &b;
        if (best) {
            best->busy = true;
            *ptr = best->ptr;
            if (got) *got = best->bytes;
            return VXH_OK;
        }
    }
    CaptureRelaxed relaxed;
    VXH_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    void *p = nullptr;
    const hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc", bytes);
    std::lock_guard<std::mutex> g(ctx->mu);
    ctx->blocks.push_back(vxh_context::Block{p, bytes, true});
    *ptr = p;
    if (got) *got = bytes;
    return VXH_OK;
}

void ctx_give_block(vxh_context *ctx, void *ptr) {
    std::lock_guard<std::mutex> g(ctx->mu);
    for (vxh_context::Block &b : ctx->blocks)
        if (b.ptr == ptr) b.busy = false;
}

vxh_status ctx_take_stream(vxh_context *ctx, hipStream_t *st) {
    {
        std::lock_guard<std::mutex> g(ctx->mu);
        if (!ctx->idle_streams.empty()) {
            *st = ctx->idle_streams.back();
            ctx->idle_streams.pop_back();
            return VXH_OK;
        }
    }
    CaptureRelaxed relaxed;
    VXH_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    hipStream_t s = nullptr;
    VXH_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate");
    std::lock_guard<std::mutex> g(ctx->mu);
    ctx->all_streams.push_back(s);
    *st = s;
    return VXH_OK;
}

void ctx_give_stream(vxh_context *ctx, hipStream_t st) {
    std::lock_guard<std::mutex> g(ctx->mu);
    ctx->idle_streams.push_back(st);
}

// The whole pipeline of one pair on `st` (inputs validated by the caller;
// `ws` holds workspace_layout(w, h, D, paths, mode, lr) bytes).  With
// `marks`, records marks[0] after the census and cost passes, marks[1] after
// the path aggregation and marks[2] at the end.
vxh_status run_pipeline(const uint8_t *d_left, const uint8_t *d_right, uint32_t w, uint32_t h,
                        const vxh_sgm_config *c, float *d_out, char *base, int mode,
                        Schedule sc, hipStream_t st, hipEvent_t const *marks) {
    const int D = (int)c->num_disparities;
    const bool lr = c->lr_check != 0;
    const WsLayout L = workspace_layout(w, h, D, (int)c->num_paths, mode, lr);
    uint2 *lc = (uint2 *)(base + L.lc);
    uint2 *rc = (uint2 *)(base + L.rc);
    uint2 *flags = mode == kOutU8 ? (uint2 *)(base + L.flags) : nullptr;
    VXH_HIP(launch_census_pair(d_left, d_right, (int)w, (int)h, c->census_radius_x,
                               c->census_radius_y, lc, rc, flags, (uint32_t *)(base + L.sync),
                               sync_words(w, h), st),
            "census");
    uint8_t *cost = (uint8_t *)(base + L.cost);
    VXH_HIP(launch_cost(lc, rc, (int)w, (int)h, D, cost_bias((uint32_t)c->p1 & 0xFFFFu), cost,
                        st),
            "cost");
    if (marks) VXH_HIP(hipEventRecord(marks[0], st), "hipEventRecord");
    // LR check: the integer winners and the right view's keys, in the left
    // census's area (free once the costs are computed).
    int16_t *best = lr ? (int16_t *)(base + L.best) : nullptr;
    uint32_t *rkey = lr ? (uint32_t *)(base + L.lc) : nullptr;
    vxh_status v = paths_and_wta(cost, w, h, c, base, L, mode, sc, st, d_out, best, nullptr,
                                 rkey, marks ? marks[1] : nullptr);
    if (v != VXH_OK) return v;
    if (lr)
        VXH_HIP(launch_lr_apply(best, rkey, (int)w, (int)h, (int)c->lr_tolerance, d_out, st),
                "lr_check");
    if (marks) VXH_HIP(hipEventRecord(marks[2], st), "hipEventRecord");
    return VXH_OK;
}

}  // namespace vxh

extern "C" {

static vxh_status enqueue_impl(vxh_sgm *s, const uint8_t *d_left, const uint8_t *d_right,
                               uint32_t w, uint32_t h, const vxh_sgm_config *c,
                               float *d_out, void *d_ws, size_t ws_bytes,
                               hipStream_t st, hipEvent_t const *marks = nullptr) {
    vxh_status v = vxh_sgm_config_validate(c, w, h);
    if (v != VXH_OK) return v;
    const int D = (int)c->num_disparities;
    if (!d_left || !d_right || !d_out || !d_ws)
        return fail(VXH_ERR_INVALID_CONFIG, "invalid config: null device buffer");
    const int mode = mode_for(s, c);
    const WsLayout L = workspace_layout(w, h, D, (int)c->num_paths, mode, c->lr_check != 0);
    if (ws_bytes < L.total) {
        char buf[128];
        snprintf(buf, sizeof buf, "invalid config: workspace %zu < %zu bytes",
                 ws_bytes, L.total);
        return fail(VXH_ERR_INVALID_CONFIG, buf);
    }
    return vxh::run_pipeline(d_left, d_right, w, h, c, d_out, (char *)d_ws, mode,
                             schedule_for(s, D, (int)c->num_paths, w, h), st, marks);
}

vxh_status vxh_sgm_enqueue(vxh_sgm *s, const uint8_t *d_left, const uint8_t *d_right,
                           uint32_t w, uint32_t h, const vxh_sgm_config *c,
                           float *d_out, void *d_ws, size_t ws_bytes, void *stream) {
    if (!s) return fail(VXH_ERR_INVALID_CONFIG, "invalid config: null handle");
    VXH_HIP(hipSetDevice(s->ctx->device), "hipSetDevice");
    return enqueue_impl(s, d_left, d_right, w, h, c, d_out, d_ws, ws_bytes,
                        (hipStream_t)stream);
}

size_t vxh_sgm_static_scratch_bytes(void) { return kStaticScratchBytes; }

// vxh_sgm_enqueue with the device's static scratch as the workspace: no
// allocation of any kind.  Split scanlines stay off (their bounded waits
// report through the workspace, which no caller holds here); results are the
// same either way.
vxh_status vxh_sgm_enqueue_static(vxh_sgm *s, const uint8_t *d_left, const uint8_t *d_right,
                                  uint32_t w, uint32_t h, const vxh_sgm_config *c, float *d_out,
                                  void *stream) {
    if (!s) return fail(VXH_ERR_INVALID_CONFIG, "invalid config: null handle");
    vxh_status v = vxh_sgm_config_validate(c, w, h);
    if (v != VXH_OK) return v;
    if (!d_left || !d_right || !d_out)
        return fail(VXH_ERR_INVALID_CONFIG, "invalid config: null device buffer");
    const int mode = mode_for(s, c);
    const WsLayout L =
        workspace_layout(w, h, c->num_disparities, (int)c->num_paths, mode, c->lr_check != 0);
    if (L.total > kStaticScratchBytes) {
        char buf[200];
        snprintf(buf, sizeof buf,
                 "invalid config: %ux%u with %u disparities and %u paths needs %zu bytes of "
                 "scratch; the static scratch holds %zu",
                 w, h, c->num_disparities, c->num_paths, L.total, kStaticScratchBytes);
        return fail(VXH_ERR_INVALID_CONFIG, buf);
    }
    VXH_HIP(hipSetDevice(s->ctx->device), "hipSetDevice");
    void *ws = static_scratch(s->ctx->device);
    if (!ws) return fail(VXH_ERR_SHADER_MISSING, "shader missing: static scratch not resolved");
    vxh::Schedule sc = schedule_for(s, (int)c->num_disparities, (int)c->num_paths, w, h);
    sc.split = 0;
    return vxh::run_pipeline(d_left, d_right, w, h, c, d_out, (char *)ws, mode, sc,
                             (hipStream_t)stream);
}

vxh_status vxh_sgm_enqueue_marked(vxh_sgm *s, const uint8_t *d_left, const uint8_t *d_right,
                                  uint32_t w, uint32_t h, const vxh_sgm_config *c,
                                  float *d_out, void *d_ws, size_t ws_bytes, void *stream,
                                  void *const *events3) {
    if (!s) return fail(VXH_ERR_INVALID_CONFIG, "invalid config: null handle");
    if (!events3 || !events3[0] || !events3[1] || !events3[2])
        return fail(VXH_ERR_INVALID_CONFIG, "invalid config: three events needed");
    VXH_HIP(hipSetDevice(s->ctx->device), "hipSetDevice");
    return enqueue_impl(s, d_left, d_right, w, h, c, d_out, d_ws, ws_bytes,
                        (hipStream_t)stream, (hipEvent_t const *)events3);
}

size_t vxh_sgm_workspace_error_offset(void) { return vxh::kErrorWordOffset; }

vxh_status vxh_sgm_workspace_status(vxh_sgm *s, const void *d_ws) {
    if (!s || !d_ws) return fail(VXH_ERR_INVALID_CONFIG, "invalid config: null argument");
    uint32_t err = 0;
    vxh_status v = copy_blocking(s->ctx, &err, (const char *)d_ws + vxh::kErrorWordOffset, 4,
                                 hipMemcpyDeviceToHost);
    if (v != VXH_OK) return v;
    if (err) return fail(VXH_ERR_GPU, "GPU error: a split scanline's wait timed out");
    return VXH_OK;
}

static vxh_status acquire_slot(vxh_sgm *s, size_t bytes, Slot **out) {
    Slot *sl = nullptr;
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (!s->free_slots.empty()) {
            sl = s->free_slots.back();
            s->free_slots.pop_back();
        }
    }
    if (!sl) {
        sl = new Slot();
        hipError_t e = hipStreamCreateWithFlags(&sl->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete sl;
            return hip_fail(e, "hipStreamCreate");
        }
        std::lock_guard<std::mutex> g(s->mu);
        s->all_slots.push_back(sl);
    }
    if (sl->ws_bytes < bytes) {
        if (sl->ws) (void)hipFree(sl->ws);
        sl->ws = nullptr;
        sl->ws_bytes = 0;
        hipError_t e = hipMalloc(&sl->ws, bytes);
        if (e != hipSuccess) {
            std::lock_guard<std::mutex> g(s->mu);
            s->free_slots.push_back(sl);
            return hip_fail(e, "hipMalloc", bytes);
        }
        sl->ws_bytes = bytes;
    }
    *out = sl;
    return VXH_OK;
}

static void release_slot(vxh_sgm *s, Slot *sl) {
    std::lock_guard<std::mutex> g(s->mu);
    s->free_slots.push_back(sl);
}

// Texture checks of the synchronous and the encode form: gray8 inputs and an
// R32Float output of one size, else INVALID_CONFIG / TEXTURE_SIZE_MISMATCH
// (the reference's compute_disparity, sgm.rs:108-119, checks none and takes
// the sizes from the left texture).
static vxh_status check_pair_textures(vxh_sgm *s, const vxh_texture *left,
                                      const vxh_texture *right, const vxh_texture *out) {
    if (!s || !left || !right || !out)
        return fail(VXH_ERR_INVALID_CONFIG, "invalid config: null argument");
    if (left->format != kGray8 || right->format != kGray8 || out->format != kR32F)
        return fail(VXH_ERR_INVALID_CONFIG, "invalid config: texture format");
    if (left->w != right->w || left->h != right->h || left->w != out->w ||
        left->h != out->h)
        return fail(VXH_ERR_TEXTURE_SIZE_MISMATCH, "texture size mismatch");
    return VXH_OK;
}

vxh_status vxh_sgm_encode(vxh_sgm *s, vxh_pipeline *p, const vxh_texture *left,
                          const vxh_texture *right, vxh_texture *out,
                          const vxh_sgm_config *c) {
    if (!p) return fail(VXH_ERR_INVALID_CONFIG, "invalid config: null pipeline");
    vxh_status v = check_pair_textures(s, left, right, out);
    if (v != VXH_OK) return v;
    const uint32_t w = left->w, h = left->h;
    if ((v = vxh_sgm_config_validate(c, w, h)) != VXH_OK) return v;
    hipStream_t st = nullptr;
    if ((v = vxh::pipeline_stream(p, &st)) != VXH_OK) return v;
    const int mode = mode_for(s, c);
    const WsLayout L = workspace_layout(w, h, c->num_disparities, (int)c->num_paths, mode,
                                        c->lr_check != 0);
    void *ws = nullptr;
    if ((v = vxh::pipeline_scratch(p, L.total, &ws)) != VXH_OK) return v;
    v = vxh::run_pipeline((const uint8_t *)left->dptr, (const uint8_t *)right->dptr, w, h, c,
                          (float *)out->dptr, (char *)ws, mode,
                          schedule_for(s, (int)c->num_disparities, (int)c->num_paths, w, h),
                          st);
    if (v != VXH_OK) return v;
    // The split launch's error word (a head whose wait for its tail timed out).
    for (const vxh_texture *t : {left, right, (const vxh_texture *)out}) vxh::pipeline_use(p, t);
    return vxh::pipeline_watch_word(p, (const uint32_t *)((char *)ws + vxh::kErrorWordOffset));
}

vxh_status vxh_sgm_compute_disparity(vxh_sgm *s, vxh_context *ctx,
                                     const vxh_texture *left,
                                     const vxh_texture *right, vxh_texture *out,
                                     const vxh_sgm_config *c) {
    (void)ctx;
    vxh_status v = check_pair_textures(s, left, right, out);
    if (v != VXH_OK) return v;
    const uint32_t w = left->w, h = left->h;
    v = vxh_sgm_config_validate(c, w, h);
    if (v != VXH_OK) return v;
    vxh::CaptureRelaxed relaxed;   // callable while a pipeline is being encoded
    VXH_HIP(hipSetDevice(s->ctx->device), "hipSetDevice");
    Slot *sl = nullptr;
    v = acquire_slot(s, workspace_total(w, h, c), &sl);
    if (v != VXH_OK) return v;
    v = enqueue_impl(s, (const uint8_t *)left->dptr, (const uint8_t *)right->dptr, w, h,
                     c, (float *)out->dptr, sl->ws, sl->ws_bytes, sl->stream);
    hipError_t e = hipStreamSynchronize(sl->stream);
    // The split launch's error word: a head whose wait for its tail timed out.
    uint32_t err = 0;
    if (v == VXH_OK && e == hipSuccess)
        e = hipMemcpyWithStream(&err, (char *)sl->ws + vxh::kErrorWordOffset, 4,
                                hipMemcpyDeviceToHost, sl->stream);
    release_slot(s, sl);
    if (v != VXH_OK) return v;
    if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    if (err) return fail(VXH_ERR_GPU, "GPU error: a split scanline's wait timed out");
    return VXH_OK;
}

vxh_status vxh_sgm_compute(vxh_sgm *s, vxh_context *ctx, const vxh_texture *left,
                           const vxh_texture *right, const vxh_sgm_config *c,
                           const vxh_camera_intrinsics *K, float baseline,
                           vxh_depth_map *out) {
    (void)baseline;
    if (!s || !ctx || !left || !K || !out)
        return fail(VXH_ERR_INVALID_CONFIG, "invalid config: null argument");
    vxh_texture *t = nullptr;
    vxh_status v = vxh_texture_output_r32float(ctx, left->w, left->h, &t);
    if (v != VXH_OK) return v;
    v = vxh_sgm_compute_disparity(s, ctx, left, right, t, c);
    if (v != VXH_OK) {
        vxh_texture_destroy(t);
        return v;
    }
    out->depth = t;
    out->intrinsics = *K;
    out->min_depth = 0.0f;
    out->max_depth = FLT_MAX;
    return VXH_OK;
}

vxh_status vxh_sgm_compute_host(vxh_sgm *s, const uint8_t *left, const uint8_t *right,
                                uint32_t w, uint32_t h, const vxh_sgm_config *c,
                                float *out) {
    if (!s) return fail(VXH_ERR_INVALID_CONFIG, "invalid config: null handle");
    vxh_status v = vxh_sgm_config_validate(c, w, h);
    if (v != VXH_OK) return v;
    vxh_texture *tl = nullptr, *tr = nullptr, *to = nullptr;
    size_t n = (size_t)w * h;
    v = vxh_texture_gray8(s->ctx, left, n, w, h, &tl);
    if (v == VXH_OK) v = vxh_texture_gray8(s->ctx, right, n, w, h, &tr);
    if (v == VXH_OK) v = vxh_texture_output_r32float(s->ctx, w, h, &to);
    if (v == VXH_OK) v = vxh_sgm_compute_disparity(s, s->ctx, tl, tr, to, c);
    if (v == VXH_OK) v = vxh_texture_read_r32float(to, out, n);
    vxh_texture_destroy(tl);
    vxh_texture_destroy(tr);
    vxh_texture_destroy(to);
    return v;
}

vxh_status vxh_sgm_census(const uint8_t *d_img, uint32_t w, uint32_t h, uint32_t rx,
                          uint32_t ry, uint32_t *d_census, void *stream) {
    if (!d_img || !d_census || w == 0 || h == 0)
        return fail(VXH_ERR_INVALID_CONFIG, "invalid config: census arguments");
    if ((2ull * rx + 1) * (2ull * ry + 1) - 1 > 64)
        return fail(VXH_ERR_INVALID_CONFIG, "invalid config: census window over 64 bits");
    VXH_HIP(vxh::launch_census(d_img, w, h, rx, ry, (uint2 *)d_census, nullptr, nullptr, 0,
                               (hipStream_t)stream),
            "census");
    return VXH_OK;
}

vxh_status vxh_sgm_set_schedule(vxh_sgm *s, uint32_t nd, uint32_t lanes_h,
                                uint32_t per_lane_h, uint32_t lanes, uint32_t per_lane) {
    if (!s || nd < 1 || nd > 256)
        return fail(VXH_ERR_INVALID_CONFIG, "invalid config: schedule target");
    if (lanes_h == 0 && per_lane_h == 0 && lanes == 0 && per_lane == 0) {
        std::lock_guard<std::mutex> g(s->mu);
        s->layout_set[nd] = false;
        return VXH_OK;
    }
    const int ih = layout_index((int)nd, lanes_h, per_lane_h);
    const int io = layout_index((int)nd, lanes, per_lane);
    if (ih < 0 || io < 0)
        return fail(VXH_ERR_INVALID_CONFIG, "invalid config: unsupported lane layout");
    std::lock_guard<std::mutex> g(s->mu);
    s->sched[nd].horiz = ih;
    s->sched[nd].other = io;
    s->layout_set[nd] = true;
    return VXH_OK;
}

vxh_status vxh_sgm_set_split(vxh_sgm *s, uint32_t nd, uint32_t split) {
    if (!s || nd < 1 || nd > 256 || (split > 2 && split != 0xFFFFFFFFu))
        return fail(VXH_ERR_INVALID_CONFIG, "invalid config: split setting");
    std::lock_guard<std::mutex> g(s->mu);
    s->split_set[nd] = split != 0xFFFFFFFFu;
    if (s->split_set[nd]) s->sched[nd].split = (int)split;
    return VXH_OK;
}

static void put_schedule(const vxh::Schedule &sc, int D, uint32_t *out5) {
    const int c = vxh::layout_class(D);
    out5[0] = (uint32_t)vxh::kClassLayouts[c][sc.horiz][0];
    out5[1] = (uint32_t)vxh::kClassLayouts[c][sc.horiz][1];
    out5[2] = (uint32_t)vxh::kClassLayouts[c][sc.other][0];
    out5[3] = (uint32_t)vxh::kClassLayouts[c][sc.other][1];
    out5[4] = (uint32_t)sc.split;
}

vxh_status vxh_sgm_get_schedule(vxh_sgm *s, uint32_t nd, uint32_t *out5) {
    if (!s || nd < 1 || nd > 256 || !out5)
        return fail(VXH_ERR_INVALID_CONFIG, "invalid config: schedule target");
    put_schedule(schedule_for(s, (int)nd, 8, 1280, 720), (int)nd, out5);
    return VXH_OK;
}

vxh_status vxh_sgm_schedule_for(vxh_sgm *s, uint32_t w, uint32_t h, const vxh_sgm_config *c,
                                uint32_t *out5) {
    if (!s || !out5) return fail(VXH_ERR_INVALID_CONFIG, "invalid config: schedule target");
    vxh_status v = vxh_sgm_config_validate(c, w, h);
    if (v != VXH_OK) return v;
    const int D = (int)c->num_disparities;
    put_schedule(schedule_for(s, D, (int)c->num_paths, w, h), D, out5);
    return VXH_OK;
}

uint32_t vxh_sgm_volume_format_in_effect(vxh_sgm *s, uint32_t nd, const vxh_sgm_config *c) {
    if (!s || !c || nd < 1 || nd > 256 || c->num_disparities != nd) return 0;
    return mode_for(s, c) == vxh::kOutU8 ? 1u : 2u;
}

vxh_status vxh_sgm_set_split_wait(vxh_sgm *s, uint32_t polls) {
    if (!s) return fail(VXH_ERR_INVALID_CONFIG, "invalid config: null handle");
    s->wait_polls.store(polls == 0xFFFFFFFFu ? vxh::kSplitWaitPolls : polls);
    return VXH_OK;
}

vxh_status vxh_sgm_set_volume_format(vxh_sgm *s, uint32_t nd, uint32_t format) {
    if (!s || nd < 1 || nd > 256 || format > 2)
        return fail(VXH_ERR_INVALID_CONFIG, "invalid config: volume format");
    std::lock_guard<std::mutex> g(s->mu);
    s->format[nd] = (uint8_t)format;
    return VXH_OK;
}

vxh_status vxh_util_copy(void *d_dst, const void *d_src, size_t bytes, void *stream) {
    if (!d_dst || !d_src || bytes % 16 || ((uintptr_t)d_dst | (uintptr_t)d_src) % 16)
        return fail(VXH_ERR_INVALID_CONFIG, "invalid config: copy needs 16-byte multiples");
    VXH_HIP(vxh::launch_copy(d_dst, d_src, bytes, (hipStream_t)stream), "copy");
    return VXH_OK;
}

uint32_t vxh_sgm_volume_stride(uint32_t nd) { return (uint32_t)vxh::volume_stride((int)nd); }

vxh_status vxh_sgm_aggregate_layout(const uint32_t *d_lc, const uint32_t *d_rc,
                                    uint32_t w, uint32_t h, const vxh_sgm_config *c,
                                    uint32_t lanes, uint32_t per_lane,
                                    uint16_t *d_volume, void *stream) {
    vxh_status v = vxh_sgm_config_validate(c, w, h);
    if (v != VXH_OK) return v;
    if (!d_lc || !d_rc || !d_volume)
        return fail(VXH_ERR_INVALID_CONFIG, "invalid config: null device buffer");
    const int D = (int)c->num_disparities;
    vxh::Schedule sc = vxh::default_schedule(D);
    if (lanes != 0 || per_lane != 0) {
        const int i = layout_index(D, lanes, per_lane);
        if (i < 0) return fail(VXH_ERR_INVALID_CONFIG, "invalid config: unsupported lane layout");
        sc = vxh::Schedule{i, i};
    }
    // Stream-ordered scratch (the call stays asynchronous): the cost volume,
    // the flags and the per-direction volumes.
    hipStream_t st = (hipStream_t)stream;
    const size_t n = (size_t)w * h;
    const int mode = auto_mode(c);
    const WsLayout L = workspace_layout(w, h, D, (int)c->num_paths, mode, false);
    const size_t bytes = L.total;
    char *g = nullptr;
    hipError_t e = hipMallocAsync((void **)&g, bytes, st);
    if (e != hipSuccess) return hip_fail(e, "hipMallocAsync", bytes);
    e = vxh::launch_cost((const uint2 *)d_lc, (const uint2 *)d_rc, (int)w, (int)h, D,
                         vxh::cost_bias((uint32_t)c->p1 & 0xFFFFu),
                         (uint8_t *)(g + L.cost), st);
    if (e == hipSuccess && mode == vxh::kOutU8) e = hipMemsetAsync(g + L.flags, 0, n * 8, st);
    if (e == hipSuccess)
        e = vxh::launch_zero((uint32_t *)(g + L.sync), vxh::sync_words(w, h), st);
    vxh_status v2 = e == hipSuccess
                        ? paths_and_wta((const uint8_t *)(g + L.cost), w, h, c, g, L, mode, sc,
                                        st, nullptr, nullptr, d_volume, nullptr)
                        : hip_fail(e, "launch");
    hipError_t e2 = hipFreeAsync(g, st);
    if (v2 != VXH_OK) return v2;
    if (e2 != hipSuccess) return hip_fail(e2, "hipFreeAsync");
    return VXH_OK;
}

vxh_status vxh_sgm_aggregate(const uint32_t *d_lc, const uint32_t *d_rc, uint32_t w,
                             uint32_t h, const vxh_sgm_config *c, uint16_t *d_volume,
                             void *stream) {
    return vxh_sgm_aggregate_layout(d_lc, d_rc, w, h, c, 0, 0, d_volume, stream);
}

vxh_status vxh_sgm_wta(const uint16_t *d_volume, uint32_t w, uint32_t h,
                       const vxh_sgm_config *c, float *d_out, void *stream) {
    vxh_status v = vxh_sgm_config_validate(c, w, h);
    if (v != VXH_OK) return v;
    if (c->lr_check)
        return fail(VXH_ERR_INVALID_CONFIG,
                    "invalid config: the stage-level WTA entry point has no LR check");
    const int D = (int)c->num_disparities;
    VXH_HIP(vxh::launch_wta_sum((const uint8_t *)d_volume, 0, nullptr, nullptr, 1, vxh::kOutU16,
                                (int)(w * h), D, vxh::volume_stride(D), d_out, nullptr,
                                nullptr, (hipStream_t)stream),
            "wta");
    return VXH_OK;
}

}  // extern "C"
