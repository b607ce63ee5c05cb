/*
 * vxh — MI355X-native semi-global matching (SGM) stereo, C ABI.
 *
 * Drop-in boundary for vx-vision's SGMStereo (vx-vision/src/kernels/sgm.rs).
 * Every entry point names the reference interface it replaces.  All calls
 * return a vxh_status; on failure vxh_last_error() returns a thread-local
 * message (the reference's Error Display text, vx-vision/src/error.rs:25-37).
 *
 * Ownership: host buffers passed in are borrowed for the duration of the call;
 * every device buffer is owned by the library (textures by their handle, the
 * per-call scratch by the SGM handle's slot pool).  Handles are thread-safe:
 * concurrent compute calls on one vxh_sgm each take their own scratch slot and
 * HIP stream, as the reference's SGMStereo is Send + Sync
 * (vx-vision/src/kernels/sgm.rs:68-69).
 */
#ifndef VXH_SGM_H
#define VXH_SGM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes, one per variant of the reference's Error enum
 * (vx-vision/src/error.rs:6-23). */
typedef enum {
    VXH_OK = 0,
    VXH_ERR_DEVICE_NOT_FOUND = 1,     /* Error::DeviceNotFound          */
    VXH_ERR_SHADER_MISSING = 2,       /* Error::ShaderMissing(name)     */
    VXH_ERR_PIPELINE_COMPILE = 3,     /* Error::PipelineCompile(msg)    */
    VXH_ERR_BUFFER_ALLOC = 4,         /* Error::BufferAlloc { bytes }   */
    VXH_ERR_TEXTURE_SIZE_MISMATCH = 5,/* Error::TextureSizeMismatch     */
    VXH_ERR_INVALID_CONFIG = 6,       /* Error::InvalidConfig(msg)      */
    VXH_ERR_GPU = 7                   /* Error::Gpu(msg)                */
} vxh_status;

/* Mirrors SGMStereoConfig (vx-vision/src/kernels/sgm.rs:23-37) plus the path
 * count and the optional left-right check, which the reference lacks.
 *   num_paths: 2 = the reference's (+1,0) and (0,+1) (sgm.rs:197-200);
 *              4 adds (-1,0) and (0,-1); 8 adds the four diagonals.
 *   lr_check:  0 off; otherwise left pixels whose match fails
 *              |d - D_R(x - d)| <= lr_tolerance become -1.0f. */
typedef struct {
    uint32_t num_disparities;
    float p1;
    float p2;
    uint32_t census_radius_x;
    uint32_t census_radius_y;
    uint32_t num_paths;
    uint32_t lr_check;
    uint32_t lr_tolerance;
} vxh_sgm_config;

/* SGMStereoConfig::default (sgm.rs:39-49): D=128, P1=10, P2=120, census 4x3;
 * num_paths 2 (the reference's behaviour), lr_check off, tolerance 1. */
void vxh_sgm_config_default(vxh_sgm_config *cfg);
/* SGMStereoConfig::new(num_disparities) (sgm.rs:51-59). */
void vxh_sgm_config_new(uint32_t num_disparities, vxh_sgm_config *cfg);
/* Validates a config for a w x h pair.  VXH_OK or VXH_ERR_INVALID_CONFIG:
 * 1 <= D <= 256, D <= w, (2rx+1)(2ry+1)-1 <= 64 census bits, 0 <= p1, p2 <
 * 65536 (truncated to u16 as the shader's ushort(p) conversion does,
 * SGMStereo.metal:155-157), num_paths in {2,4,8}. */
vxh_status vxh_sgm_config_validate(const vxh_sgm_config *cfg, uint32_t w,
                                   uint32_t h);

/* ---- Context (vx-vision/src/context.rs:17-36, vx-core/src/device.rs) ---- */
typedef struct vxh_context vxh_context;
/* Context::new (context.rs:26-36): device ordinal instead of the system
 * default device; VXH_ERR_DEVICE_NOT_FOUND when no such HIP device exists,
 * VXH_ERR_SHADER_MISSING when the gfx950 code object lacks a kernel. */
vxh_status vxh_context_create(int device, vxh_context **out);
void vxh_context_destroy(vxh_context *ctx);
/* The context's HIP device ordinal (-1 for NULL). */
int vxh_context_device(const vxh_context *ctx);
/* Number of HIP devices visible to the process (0 when none). */
int vxh_device_count(void);

/* ---- Textures (vx-vision/src/texture.rs) ---- */
typedef struct vxh_texture vxh_texture;
/* Context::texture_gray8 -> Texture::from_gray8 (context.rs:39-41,
 * texture.rs:38-93): uploads w*h bytes; len != w*h -> INVALID_CONFIG. */
vxh_status vxh_texture_gray8(vxh_context *ctx, const uint8_t *pixels,
                             size_t len, uint32_t w, uint32_t h,
                             vxh_texture **out);
/* Context::texture_output_r32float (context.rs:64-66, texture.rs:324-349). */
vxh_status vxh_texture_output_r32float(vxh_context *ctx, uint32_t w,
                                       uint32_t h, vxh_texture **out);
/* Context::texture_r32float -> Texture::from_r32float (context.rs:44-46,
 * texture.rs:96-151): uploads w*h floats; len != w*h -> INVALID_CONFIG. */
vxh_status vxh_texture_r32float(vxh_context *ctx, const float *data, size_t len,
                                uint32_t w, uint32_t h, vxh_texture **out);
/* Context::texture_output_gray8 (context.rs:49-51, texture.rs:154-180) and
 * texture_intermediate_gray8 (context.rs:69-71): an uninitialised gray8
 * texture (every texture here is device memory, so both are the same). */
vxh_status vxh_texture_output_gray8(vxh_context *ctx, uint32_t w, uint32_t h,
                                    vxh_texture **out);
/* Context::texture_rgba8 -> Texture::from_rgba8 (context.rs:54-56,
 * texture.rs:238-294): uploads w*h*4 bytes; len != w*h*4 -> INVALID_CONFIG. */
vxh_status vxh_texture_rgba8(vxh_context *ctx, const uint8_t *pixels, size_t len,
                             uint32_t w, uint32_t h, vxh_texture **out);
/* Context::texture_output_rgba8 (context.rs:59-61, texture.rs:296-322). */
vxh_status vxh_texture_output_rgba8(vxh_context *ctx, uint32_t w, uint32_t h,
                                    vxh_texture **out);
/* Texture::read_gray8 / read_rgba8 (texture.rs:352-402): copies w*h (gray8)
 * or w*h*4 (RGBA8) bytes to dst; another length or format -> INVALID_CONFIG. */
vxh_status vxh_texture_read_gray8(const vxh_texture *tex, uint8_t *dst, size_t len);
vxh_status vxh_texture_read_rgba8(const vxh_texture *tex, uint8_t *dst, size_t len);
/* Pixel format of a texture: 1 = gray8 (R8Unorm), 2 = R32Float,
 * 3 = RGBA8Unorm, 0 = none. */
uint32_t vxh_texture_format(const vxh_texture *tex);
/* Texture::read_r32float (texture.rs:404-427): copies w*h floats to dst;
 * len != w*h or a non-R32Float texture -> INVALID_CONFIG. */
vxh_status vxh_texture_read_r32float(const vxh_texture *tex, float *dst,
                                     size_t len);
/* Replaces the bytes of an existing gray8 texture (same size).  Refused with
 * VXH_ERR_INVALID_CONFIG while the texture is in flight: a committed or
 * relaunched pipeline that encoded work on it has not finished (the
 * reference's UnifiedBuffer refuses mutable access while in flight,
 * vx-core/src/buffer.rs:59-72).  Reads of a texture wait for such work. */
vxh_status vxh_texture_write_gray8(vxh_texture *tex, const uint8_t *pixels,
                                   size_t len);
/* UnifiedBuffer::is_in_flight (vx-core/src/buffer.rs:89-92): 1 while a
 * launched pipeline that used the texture has not finished, else 0. */
int vxh_texture_in_flight(const vxh_texture *tex);
/* HIP device ordinal the texture lives on (MTLTexture.device); -1 for NULL. */
int vxh_texture_device(const vxh_texture *tex);
uint32_t vxh_texture_width(const vxh_texture *tex);
uint32_t vxh_texture_height(const vxh_texture *tex);
/* Device address of the texture's pixels (row-major, w*h elements). */
void *vxh_texture_device_ptr(const vxh_texture *tex);
/* Frees the texture; it must be destroyed before its context. */
void vxh_texture_destroy(vxh_texture *tex);

/* ---- SGM kernel struct (vx-vision/src/kernels/sgm.rs:61-105) ---- */
typedef struct vxh_sgm vxh_sgm;
/* SGMStereo::new (sgm.rs:73-105): resolves the kernels once. */
vxh_status vxh_sgm_create(vxh_context *ctx, vxh_sgm **out);
void vxh_sgm_destroy(vxh_sgm *sgm);

/* SGMStereo::compute_disparity (sgm.rs:108-265).  Synchronous, like the
 * reference (commit + waitUntilCompleted, sgm.rs:260-261).  Sizes of left,
 * right and output must agree (TEXTURE_SIZE_MISMATCH otherwise). */
vxh_status vxh_sgm_compute_disparity(vxh_sgm *sgm, vxh_context *ctx,
                                     const vxh_texture *left,
                                     const vxh_texture *right,
                                     vxh_texture *output,
                                     const vxh_sgm_config *cfg);

/* Camera intrinsics and depth map, as CameraIntrinsics / DepthMap
 * (vx-vision/src/types_3d.rs:254-300, :384-430). */
typedef struct {
    float fx, fy, cx, cy;
    uint32_t width, height;
} vxh_camera_intrinsics;
typedef struct {
    vxh_texture *depth;          /* R32Float disparity map, owned by caller */
    vxh_camera_intrinsics intrinsics;
    float min_depth;             /* 0.0      */
    float max_depth;             /* FLT_MAX  */
} vxh_depth_map;
/* SGMStereo::compute (sgm.rs:269-288): allocates the output texture, runs
 * compute_disparity and wraps the disparity map (no depth conversion, as in
 * the reference; baseline is unused there too). */
vxh_status vxh_sgm_compute(vxh_sgm *sgm, vxh_context *ctx,
                           const vxh_texture *left, const vxh_texture *right,
                           const vxh_sgm_config *cfg,
                           const vxh_camera_intrinsics *intrinsics,
                           float baseline, vxh_depth_map *out);

/* Host-buffer convenience: upload, compute_disparity, read back, the
 * sequence of the reference's example caller
 * (examples/depth_to_cloud_demo.rs:46-55). */
vxh_status vxh_sgm_compute_host(vxh_sgm *sgm, const uint8_t *left,
                                const uint8_t *right, uint32_t w, uint32_t h,
                                const vxh_sgm_config *cfg, float *out);

/* ---- Device-resident, stream-ordered form (the encode() pattern the
 * reference's kernels follow, vx-vision/src/pipeline.rs:12-80) ---- */
/* Scratch bytes one device-resident call needs for a w x h pair (enough for
 * either volume format). */
size_t vxh_sgm_workspace_bytes(uint32_t w, uint32_t h,
                               const vxh_sgm_config *cfg);
/* Enqueues the whole pipeline on `stream` (a hipStream_t; NULL = the null
 * stream) with caller-owned device buffers: left/right w*h bytes, out w*h
 * floats, workspace >= vxh_sgm_workspace_bytes.  Does not synchronize. */
vxh_status vxh_sgm_enqueue(vxh_sgm *sgm, const uint8_t *d_left,
                           const uint8_t *d_right, uint32_t w, uint32_t h,
                           const vxh_sgm_config *cfg, float *d_out,
                           void *d_workspace, size_t workspace_bytes,
                           void *stream);

/* vxh_sgm_enqueue that also records three caller-created HIP events
 * (hipEvent_t) on `stream`: events3[0] after the census pass, events3[1]
 * after the path aggregation, events3[2] after the winner-take-all (and LR
 * check), so a caller can time each stage in stream order.  No reference
 * counterpart (the reference has no GPU timestamps). */
vxh_status vxh_sgm_enqueue_marked(vxh_sgm *sgm, const uint8_t *d_left,
                                  const uint8_t *d_right, uint32_t w, uint32_t h,
                                  const vxh_sgm_config *cfg, float *d_out,
                                  void *d_workspace, size_t workspace_bytes,
                                  void *stream, void *const *events3);

/* vxh_sgm_enqueue for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): the same
 * pipeline, with the library's static device scratch of
 * vxh_sgm_static_scratch_bytes() (96 MiB, held by every GPU the library runs
 * on until the process ends) as its workspace instead of a caller-supplied
 * one.  It allocates nothing: no hipMalloc, hipFree or allocating library
 * call.  Pairs whose workspace (vxh_sgm_workspace_bytes) exceeds the static
 * scratch are rejected with VXH_ERR_INVALID_CONFIG (320x240 with D=64 and 8
 * paths fits).  Every call on one GPU shares that GPU's scratch, so calls
 * that overlap on one GPU (other streams, other threads) race; the scratch
 * keeps the last call's contents, and each call clears what must start at
 * zero itself.  Split scanlines stay off.  No counterpart in the reference. */
vxh_status vxh_sgm_enqueue_static(vxh_sgm *sgm, const uint8_t *d_left,
                                  const uint8_t *d_right, uint32_t w, uint32_t h,
                                  const vxh_sgm_config *cfg, float *d_out,
                                  void *stream);
/* Bytes of the static scratch vxh_sgm_enqueue_static works in. */
size_t vxh_sgm_static_scratch_bytes(void);

/* Status of the enqueued call that used `d_workspace`, once the caller has
 * synchronised the stream it was enqueued on: VXH_ERR_GPU when a split
 * scanline's bounded wait expired in that call (its map is then incomplete),
 * else VXH_OK.  Reads the call's error word with a blocking copy.  Every
 * enqueue clears the word first, so it describes the last call on the
 * workspace.  No reference counterpart (the reference's compute_disparity
 * waits with waitUntilCompleted and checks nothing, sgm.rs:260-261). */
vxh_status vxh_sgm_workspace_status(vxh_sgm *sgm, const void *d_workspace);
/* Byte offset of that error word (a uint32, non-zero on failure) in the
 * workspace, for callers that copy it back in stream order themselves. */
size_t vxh_sgm_workspace_error_offset(void);

/* Stage-level entry points on device buffers, for parity tests of the
 * intermediate results.  d_census is 2*w*h uint32 (lo, hi per pixel);
 * d_volume is w*h*vxh_sgm_volume_stride(D) uint16, pixel-major with d
 * innermost (entries d >= D of each pixel are padding). */
uint32_t vxh_sgm_volume_stride(uint32_t num_disparities);
vxh_status vxh_sgm_census(const uint8_t *d_img, uint32_t w, uint32_t h,
                          uint32_t rx, uint32_t ry, uint32_t *d_census,
                          void *stream);
vxh_status vxh_sgm_aggregate(const uint32_t *d_lc, const uint32_t *d_rc,
                             uint32_t w, uint32_t h,
                             const vxh_sgm_config *cfg, uint16_t *d_volume,
                             void *stream);
/* vxh_sgm_aggregate with an explicit lane layout for every direction:
 * `lanes` lanes of a wave own a pair of scanlines and each lane owns
 * `per_lane` consecutive disparities, lanes * per_lane = 64, 128 or 256 for
 * D <= 64, <= 128, <= 256 (per_lane 2..16); 0, 0 selects the default.
 * Unsupported layouts return VXH_ERR_INVALID_CONFIG.  Both aggregate entry
 * points take stream-ordered scratch (hipMallocAsync on `stream`, freed in
 * stream order) for a guarded copy of d_rc and the per-direction volumes. */
vxh_status vxh_sgm_aggregate_layout(const uint32_t *d_lc, const uint32_t *d_rc,
                                    uint32_t w, uint32_t h,
                                    const vxh_sgm_config *cfg, uint32_t lanes,
                                    uint32_t per_lane, uint16_t *d_volume,
                                    void *stream);
vxh_status vxh_sgm_wta(const uint16_t *d_volume, uint32_t w, uint32_t h,
                       const vxh_sgm_config *cfg, float *d_out, void *stream);

/* Selects the aggregation lane layouts this handle uses for pairs with
 * exactly `num_disparities` disparities: (lanes_h, per_lane_h) for the
 * horizontal directions, whose scanlines are the longest, and
 * (lanes, per_lane) for the others, each as for vxh_sgm_aggregate_layout;
 * all zero restores the default, which depends on the pair's size and path
 * count (vxh_sgm_schedule_for).  Results do not depend on the layouts, only
 * the speed does.  No reference counterpart (the reference has one fixed
 * thread per scanline, vx-vision/src/kernels/sgm.rs:346-357). */
vxh_status vxh_sgm_set_schedule(vxh_sgm *sgm, uint32_t num_disparities,
                                uint32_t lanes_h, uint32_t per_lane_h,
                                uint32_t lanes, uint32_t per_lane);
/* Whether this handle splits scanlines into two halves aggregated at once,
 * for pairs with exactly `num_disparities` disparities: 0 never, 1 in the
 * directions whose scanlines are longer than 2/3 of the longest direction's,
 * 2 in every direction; in either case only scanlines of at least 256
 * pixels.  0xFFFFFFFF restores the default, which depends on the pair's size
 * and path count (level 1 for 2 and 4 paths up to D = 128, see
 * vxh_sgm_schedule_for).  The second half
 * starts from the state every path starts from; the first half's wave then
 * continues into it until its state equals the second half's bit for bit
 * and overwrites the values before that point, so the result is identical
 * either way while the longest chain of sequential steps halves.  No
 * reference counterpart. */
vxh_status vxh_sgm_set_split(vxh_sgm *sgm, uint32_t num_disparities, uint32_t split);
/* Polls a split scanline's first half makes of its second half's progress
 * flag before it gives up, raises the call's error word (VXH_ERR_GPU from
 * compute_disparity, vxh_sgm_workspace_status, pipeline waits) and goes on
 * without the flag; 0xFFFFFFFF restores the default (seconds of polling).
 * With 0 every split call reports the error: the way a caller exercises its
 * error path.  No reference counterpart. */
vxh_status vxh_sgm_set_split_wait(vxh_sgm *sgm, uint32_t polls);
/* The schedule in effect for 1280x720 8-path pairs with `num_disparities`
 * disparities: out5 = lanes_h, per_lane_h, lanes, per_lane, split. */
vxh_status vxh_sgm_get_schedule(vxh_sgm *sgm, uint32_t num_disparities,
                                uint32_t *out5);
/* The schedule vxh_sgm_compute_disparity and vxh_sgm_enqueue use for a
 * w x h pair with `cfg`, as out5 above: what the handle was set to for
 * cfg->num_disparities, else the default for that size and path count.
 * vxh_sgm_enqueue_static uses the same layouts with split 0.  No reference
 * counterpart. */
vxh_status vxh_sgm_schedule_for(vxh_sgm *sgm, uint32_t w, uint32_t h,
                                const vxh_sgm_config *cfg, uint32_t *out5);

/* Storage of the per-direction path volumes for pairs with exactly
 * `num_disparities` disparities: 0 = automatic (u8 when 64 + P2 <= 255,
 * else u16), 1 = u8 low bytes plus the high bytes of the rare slices that
 * exceed 255, 2 = u16.  Results are identical for every setting (all sums
 * are mod 2^16).  No reference counterpart (the reference accumulates one
 * u16 volume, SGMStereo.metal:164-165). */
vxh_status vxh_sgm_set_volume_format(vxh_sgm *sgm, uint32_t num_disparities,
                                     uint32_t format);

/* The volume storage calls with `cfg` use on this handle for pairs with
 * `num_disparities` disparities (cfg->num_disparities must equal it): 1 = u8,
 * 2 = u16; 0 for invalid arguments. */
uint32_t vxh_sgm_volume_format_in_effect(vxh_sgm *sgm, uint32_t num_disparities,
                                         const vxh_sgm_config *cfg);

/* Plain device-to-device copy kernel (16-byte vectors, grid-stride), used to
 * calibrate HBM traffic counters against a kernel of known traffic.  bytes
 * must be a multiple of 16 and both pointers 16-byte aligned.  No reference
 * counterpart. */
vxh_status vxh_util_copy(void *d_dst, const void *d_src, size_t bytes,
                         void *stream);

/* Thread-local message describing the last failure on this thread. */
const char *vxh_last_error(void);
/* Library version string. */
const char *vxh_version(void);

#ifdef __cplusplus
}
#endif

#endif
