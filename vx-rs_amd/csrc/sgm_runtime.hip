// Host runtime behind include/vxh/sgm.h: contexts, device textures, the SGM
// handle with its scratch-slot pool, and the stream-ordered pipeline.
//
// Reference counterparts:
//   vxh_context      vx-vision/src/context.rs:17-36, vx-core/src/device.rs:11-56
//   vxh_texture      vx-vision/src/texture.rs:29-462 (R8Unorm in, R32Float out)
//   vxh_sgm          vx-vision/src/kernels/sgm.rs:61-288
//   device buffers   vx-core/src/buffer.rs:14-122 (UnifiedBuffer/GpuGuard): here
//                    explicit hipMalloc allocations owned by textures and by
//                    scratch slots, used in stream order; a slot is "in
//                    flight" while a call holds it and is returned only after
//                    that call's stream has drained.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "sgm_internal.h"
#include "vxh/sgm.h"

namespace {

thread_local std::string g_last_error;

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

constexpr size_t kAlign = 256;
inline size_t align_up(size_t v) { return (v + kAlign - 1) & ~(kAlign - 1); }

enum TexFormat { kGray8 = 1, kR32F = 2 };

struct Layout {
    size_t lc, rc, vol, vol_bytes, best, total;
};

// Scratch of one pair: census of both images, `groups` summed volumes (one
// per path group) and the winner index used by the LR check.
Layout workspace_layout(uint32_t w, uint32_t h, uint32_t D, int groups = 1) {
    Layout l;
    size_t n = (size_t)w * h;
    l.lc = 0;
    l.rc = l.lc + vxh::census_span(n);
    l.vol = l.rc + align_up(n * 8);
    l.vol_bytes = align_up(n * (size_t)vxh::volume_stride((int)D) * 2);
    l.best = l.vol + l.vol_bytes * (size_t)groups;
    l.total = l.best + align_up(n * 2);
    return l;
}

struct Slot {
    hipStream_t stream = nullptr;
    void *ws = nullptr;
    size_t ws_bytes = 0;
};

bool groups_valid(uint32_t g) { return g == 1 || g == 2 || g == 4 || g == 8; }

}  // namespace

struct vxh_context {
    int device;
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
    vxh::Layout layouts[257] = {};   // per D; L == 0 selects the default
    uint8_t groups[257] = {};        // per D; 0 selects the default
};

static vxh::Layout layout_for(vxh_sgm *s, int D) {
    std::lock_guard<std::mutex> g(s->mu);
    vxh::Layout l = s->layouts[D];
    return l.L == 0 ? vxh::default_layout(D) : l;
}

static int groups_for(vxh_sgm *s, int D, int P) {
    int g;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        g = s->groups[D] == 0 ? 1 : s->groups[D];
    }
    while (g > P) g /= 2;
    return g;
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
    *out = new vxh_context{device};
    return VXH_OK;
}

void vxh_context_destroy(vxh_context *ctx) { delete ctx; }

static vxh_status tex_alloc(vxh_context *ctx, uint32_t w, uint32_t h, int fmt,
                            vxh_texture **out) {
    if (!ctx || !out) return fail(VXH_ERR_INVALID_CONFIG, "invalid config: null argument");
    *out = nullptr;
    if (w == 0 || h == 0)
        return fail(VXH_ERR_INVALID_CONFIG, "invalid config: empty texture");
    size_t bytes = (size_t)w * h * (fmt == kGray8 ? 1 : 4);
    VXH_HIP(hipSetDevice(ctx->device), "hipSetDevice");
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc", bytes);
    *out = new vxh_texture{ctx, w, h, fmt, p, bytes};
    return VXH_OK;
}

vxh_status vxh_texture_gray8(vxh_context *ctx, const uint8_t *pixels, size_t len,
                             uint32_t w, uint32_t h, vxh_texture **out) {
    if (out) *out = nullptr;
    if (len != (size_t)w * h) {
        char buf[160];
        snprintf(buf, sizeof buf,
                 "invalid config: expected %zu bytes for %ux%u R8 texture, got %zu",
                 (size_t)w * h, w, h, len);
        return fail(VXH_ERR_INVALID_CONFIG, buf);
    }
    vxh_status st = tex_alloc(ctx, w, h, kGray8, out);
    if (st != VXH_OK) return st;
    hipError_t e = hipMemcpy((*out)->dptr, pixels, len, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        vxh_texture_destroy(*out);
        *out = nullptr;
        return hip_fail(e, "hipMemcpy");
    }
    return VXH_OK;
}

vxh_status vxh_texture_output_r32float(vxh_context *ctx, uint32_t w, uint32_t h,
                                       vxh_texture **out) {
    return tex_alloc(ctx, w, h, kR32F, out);
}

vxh_status vxh_texture_write_gray8(vxh_texture *t, const uint8_t *pixels, size_t len) {
    if (!t || t->format != kGray8 || len != (size_t)t->w * t->h)
        return fail(VXH_ERR_INVALID_CONFIG, "invalid config: gray8 write size");
    VXH_HIP(hipSetDevice(t->ctx->device), "hipSetDevice");
    VXH_HIP(hipMemcpy(t->dptr, pixels, len, hipMemcpyHostToDevice), "hipMemcpy");
    return VXH_OK;
}

vxh_status vxh_texture_read_r32float(const vxh_texture *t, float *dst, size_t len) {
    if (!t || t->format != kR32F)
        return fail(VXH_ERR_INVALID_CONFIG, "invalid config: not an R32Float texture");
    if (len != (size_t)t->w * t->h)
        return fail(VXH_ERR_INVALID_CONFIG, "invalid config: read length");
    VXH_HIP(hipSetDevice(t->ctx->device), "hipSetDevice");
    VXH_HIP(hipMemcpy(dst, t->dptr, len * 4, hipMemcpyDeviceToHost), "hipMemcpy");
    return VXH_OK;
}

uint32_t vxh_texture_width(const vxh_texture *t) { return t ? t->w : 0; }
uint32_t vxh_texture_height(const vxh_texture *t) { return t ? t->h : 0; }
void *vxh_texture_device_ptr(const vxh_texture *t) { return t ? t->dptr : nullptr; }

void vxh_texture_destroy(vxh_texture *t) {
    if (!t) return;
    (void)hipSetDevice(t->ctx->device);
    (void)hipFree(t->dptr);
    delete t;
}

vxh_status vxh_sgm_create(vxh_context *ctx, vxh_sgm **out) {
    if (!ctx || !out) return fail(VXH_ERR_INVALID_CONFIG, "invalid config: null argument");
    *out = nullptr;
    VXH_HIP(hipSetDevice(ctx->device), "hipSetDevice");
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
    return workspace_layout(w, h, c->num_disparities).total;
}

size_t vxh_sgm_workspace_bytes_groups(uint32_t w, uint32_t h, const vxh_sgm_config *c,
                                      uint32_t groups) {
    if (!c || !groups_valid(groups)) return 0;
    return workspace_layout(w, h, c->num_disparities, (int)groups).total;
}

// Runs every path into `vol`.  With `wta_out`, the last path also picks the
// winner from its final sums and writes the disparity map instead of storing
// its sums.
static vxh_status aggregate_impl(const uint2 *lc, const uint2 *rc, uint32_t w,
                                 uint32_t h, const vxh_sgm_config *c, uint16_t *vol,
                                 vxh::Layout lay, hipStream_t st,
                                 float *wta_out = nullptr) {
    const int D = (int)c->num_disparities;
    if (!vxh::layout_supported(lay, D))
        return fail(VXH_ERR_INVALID_CONFIG, "invalid config: unsupported lane layout");
    vxh::AggParams p{};
    p.lc = lc;
    p.rc = rc;
    p.vol = vol;
    p.w = (int)w;
    p.h = (int)h;
    p.D = D;
    p.Dp = vxh::volume_stride(D);
    p.p1 = (uint32_t)c->p1 & 0xFFFFu;
    p.p2 = (uint32_t)c->p2 & 0xFFFFu;
    for (uint32_t i = 0; i < c->num_paths; ++i) {
        p.dir = (int)i;
        p.nscan = vxh::num_scanlines((int)i, (int)w, (int)h);
        p.accumulate = i > 0;
        p.wta_out = (i + 1 == c->num_paths) ? wta_out : nullptr;
        VXH_HIP(vxh::launch_aggregate(p, lay, st), "aggregate");
    }
    return VXH_OK;
}

// Runs the P paths as P / groups launches of `groups` concurrent passes:
// launch k runs directions k*groups .. k*groups + groups - 1, direction i
// accumulating into volume i % groups (the slices of `vol_bytes` after
// `vol`), and the summing winner-take-all adds the volumes.  Sums are
// mod 2^16 and therefore independent of the grouping.  With `sum_out` the
// summed volume is stored into the first slice and `best` receives the
// integer winners (for the LR check).
static vxh_status run_path_groups(const uint2 *lc, const uint2 *rc, uint32_t w, uint32_t h,
                                  const vxh_sgm_config *c, uint16_t *vol, size_t vol_bytes,
                                  int groups, vxh::Layout lay, hipStream_t st, float *out,
                                  int16_t *best, bool store_sum) {
    const int D = (int)c->num_disparities;
    if (!vxh::layout_supported(lay, D))
        return fail(VXH_ERR_INVALID_CONFIG, "invalid config: unsupported lane layout");
    const int P = (int)c->num_paths;
    vxh::AggParams base{};
    base.lc = lc;
    base.rc = rc;
    base.w = (int)w;
    base.h = (int)h;
    base.D = D;
    base.Dp = vxh::volume_stride(D);
    base.p1 = (uint32_t)c->p1 & 0xFFFFu;
    base.p2 = (uint32_t)c->p2 & 0xFFFFu;
    for (int k = 0; k < P / groups; ++k) {
        int dirs[vxh::kMaxPasses], acc[vxh::kMaxPasses];
        uint16_t *vols[vxh::kMaxPasses];
        for (int g = 0; g < groups; ++g) {
            dirs[g] = k * groups + g;
            vols[g] = (uint16_t *)((char *)vol + (size_t)g * vol_bytes);
            acc[g] = k > 0;
        }
        VXH_HIP(vxh::launch_aggregate_multi(base, dirs, vols, acc, groups, lay, st),
                "aggregate");
    }
    VXH_HIP(vxh::launch_wta_sum(vol, vol_bytes / 2, groups, (int)(w * h), D, base.Dp, out,
                                best, store_sum ? vol : nullptr, st),
            "wta");
    return VXH_OK;
}

static vxh_status enqueue_impl(vxh_sgm *s, const uint8_t *d_left, const uint8_t *d_right,
                               uint32_t w, uint32_t h, const vxh_sgm_config *c,
                               float *d_out, void *d_ws, size_t ws_bytes,
                               hipStream_t st) {
    vxh_status v = vxh_sgm_config_validate(c, w, h);
    if (v != VXH_OK) return v;
    const int D = (int)c->num_disparities;
    Layout L = workspace_layout(w, h, D);
    if (!d_left || !d_right || !d_out || !d_ws)
        return fail(VXH_ERR_INVALID_CONFIG, "invalid config: null device buffer");
    if (ws_bytes < L.total) {
        char buf[128];
        snprintf(buf, sizeof buf, "invalid config: workspace %zu < %zu bytes",
                 ws_bytes, L.total);
        return fail(VXH_ERR_INVALID_CONFIG, buf);
    }
    // The largest path-group count up to the handle's setting whose volumes
    // fit the workspace given.
    int groups = groups_for(s, D, (int)c->num_paths);
    while (groups > 1 && workspace_layout(w, h, D, groups).total > ws_bytes) groups /= 2;
    L = workspace_layout(w, h, D, groups);
    char *base = (char *)d_ws;
    uint2 *lc = (uint2 *)(base + L.lc);
    uint2 *rc = (uint2 *)(base + L.rc);
    uint16_t *vol = (uint16_t *)(base + L.vol);
    int16_t *best = (int16_t *)(base + L.best);
    VXH_HIP(vxh::launch_census(d_left, w, h, c->census_radius_x, c->census_radius_y, lc, st),
            "census");
    VXH_HIP(vxh::launch_census(d_right, w, h, c->census_radius_x, c->census_radius_y, rc, st),
            "census");
    const vxh::Layout lay = layout_for(s, D);
    const int Dp = vxh::volume_stride(D);
    if (groups > 1) {
        vxh_status v2 = run_path_groups(lc, rc, w, h, c, vol, L.vol_bytes, groups, lay, st,
                                        d_out, c->lr_check ? best : nullptr, c->lr_check);
        if (v2 != VXH_OK || !c->lr_check) return v2;
    } else if (!c->lr_check) {
        return aggregate_impl(lc, rc, w, h, c, vol, lay, st, d_out);
    } else {
        // The LR check reads sums of other pixels, so the volume is completed
        // first and the winner taken by separate passes.
        vxh_status v2 = aggregate_impl(lc, rc, w, h, c, vol, lay, st);
        if (v2 != VXH_OK) return v2;
        VXH_HIP(vxh::launch_wta(vol, w, h, D, Dp, d_out, best, st), "wta");
    }
    VXH_HIP(vxh::launch_lr_check(vol, w, h, D, Dp, (int)c->lr_tolerance, best, d_out, st),
            "lr_check");
    return VXH_OK;
}

vxh_status vxh_sgm_enqueue(vxh_sgm *s, const uint8_t *d_left, const uint8_t *d_right,
                           uint32_t w, uint32_t h, const vxh_sgm_config *c,
                           float *d_out, void *d_ws, size_t ws_bytes, void *stream) {
    if (!s) return fail(VXH_ERR_INVALID_CONFIG, "invalid config: null handle");
    VXH_HIP(hipSetDevice(s->ctx->device), "hipSetDevice");
    return enqueue_impl(s, d_left, d_right, w, h, c, d_out, d_ws, ws_bytes,
                        (hipStream_t)stream);
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

vxh_status vxh_sgm_compute_disparity(vxh_sgm *s, vxh_context *ctx,
                                     const vxh_texture *left,
                                     const vxh_texture *right, vxh_texture *out,
                                     const vxh_sgm_config *c) {
    if (!s || !left || !right || !out)
        return fail(VXH_ERR_INVALID_CONFIG, "invalid config: null argument");
    (void)ctx;
    if (left->format != kGray8 || right->format != kGray8 || out->format != kR32F)
        return fail(VXH_ERR_INVALID_CONFIG, "invalid config: texture format");
    if (left->w != right->w || left->h != right->h || left->w != out->w ||
        left->h != out->h)
        return fail(VXH_ERR_TEXTURE_SIZE_MISMATCH, "texture size mismatch");
    const uint32_t w = left->w, h = left->h;
    vxh_status v = vxh_sgm_config_validate(c, w, h);
    if (v != VXH_OK) return v;
    VXH_HIP(hipSetDevice(s->ctx->device), "hipSetDevice");
    Slot *sl = nullptr;
    const int groups = groups_for(s, (int)c->num_disparities, (int)c->num_paths);
    v = acquire_slot(s, workspace_layout(w, h, c->num_disparities, groups).total, &sl);
    if (v != VXH_OK) return v;
    v = enqueue_impl(s, (const uint8_t *)left->dptr, (const uint8_t *)right->dptr, w, h,
                     c, (float *)out->dptr, sl->ws, sl->ws_bytes, sl->stream);
    hipError_t e = hipStreamSynchronize(sl->stream);
    release_slot(s, sl);
    if (v != VXH_OK) return v;
    if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
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
    VXH_HIP(vxh::launch_census(d_img, w, h, rx, ry, (uint2 *)d_census, (hipStream_t)stream),
            "census");
    return VXH_OK;
}

vxh_status vxh_sgm_set_layout(vxh_sgm *s, uint32_t nd, uint32_t lanes, uint32_t per_lane) {
    if (!s || nd < 1 || nd > 256)
        return fail(VXH_ERR_INVALID_CONFIG, "invalid config: layout target");
    vxh::Layout l{(int)lanes, (int)per_lane};
    if (!(lanes == 0 && per_lane == 0) && !vxh::layout_supported(l, (int)nd))
        return fail(VXH_ERR_INVALID_CONFIG, "invalid config: unsupported lane layout");
    std::lock_guard<std::mutex> g(s->mu);
    s->layouts[nd] = l;
    return VXH_OK;
}

vxh_status vxh_sgm_set_groups(vxh_sgm *s, uint32_t nd, uint32_t groups) {
    if (!s || nd < 1 || nd > 256)
        return fail(VXH_ERR_INVALID_CONFIG, "invalid config: groups target");
    if (groups != 0 && !groups_valid(groups))
        return fail(VXH_ERR_INVALID_CONFIG, "invalid config: path groups not in {1,2,4,8}");
    std::lock_guard<std::mutex> g(s->mu);
    s->groups[nd] = (uint8_t)groups;
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
    vxh::Layout lay = (lanes == 0 && per_lane == 0)
                          ? vxh::default_layout((int)c->num_disparities)
                          : vxh::Layout{(int)lanes, (int)per_lane};
    if (!vxh::layout_supported(lay, (int)c->num_disparities))
        return fail(VXH_ERR_INVALID_CONFIG, "invalid config: unsupported lane layout");
    // The kernels read up to kCensusGuard bytes in front of the right census,
    // which a caller's buffer need not have: copy it behind such a guard
    // (stream-ordered allocation, so the call stays asynchronous).
    hipStream_t st = (hipStream_t)stream;
    const size_t cb = (size_t)w * h * 8;
    char *g = nullptr;
    hipError_t e = hipMallocAsync((void **)&g, vxh::kCensusGuard + cb, st);
    if (e != hipSuccess) return hip_fail(e, "hipMallocAsync", vxh::kCensusGuard + cb);
    e = hipMemcpyAsync(g + vxh::kCensusGuard, d_rc, cb, hipMemcpyDeviceToDevice, st);
    vxh_status v2 = e == hipSuccess
                        ? aggregate_impl((const uint2 *)d_lc,
                                         (const uint2 *)(g + vxh::kCensusGuard), w, h, c,
                                         d_volume, lay, st)
                        : hip_fail(e, "hipMemcpyAsync");
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
    VXH_HIP(vxh::launch_wta(d_volume, w, h, D, vxh::volume_stride(D), d_out, nullptr,
                            (hipStream_t)stream),
            "wta");
    return VXH_OK;
}

}  // extern "C"
