// Sparse stereo matching behind include/vxh/stereo_match.h
// (vx-vision/src/kernels/stereomatch.rs:17-339,
// vx-vision/shaders/StereoMatch.metal:32-115).
//
// The shader materialises the n_left x n_right u16 distance matrix with one
// thread per pair, then scans each row with one thread per left keypoint and
// appends matches through an atomic counter.  Here the matcher never writes
// the matrix: one wave owns one left keypoint (its descriptor and position in
// scalar registers), the 64 lanes stride over the right keypoints computing
// the same masked distance, and a wave-wide minimum of (distance << 32 | right
// index) gives the first least distance exactly as the sequential scan does.
// Lane 0 triangulates and writes the keypoint's slot; a one-block pass then
// compacts the accepted slots in place, in ascending left index, and writes
// the count (no atomics, a deterministic order).  The matrix kernel, for
// callers that want the matrix, maps lanes to right keypoints so each row is
// written as contiguous 128-byte runs.
#include <hip/hip_runtime.h>

#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "sgm_pipeline.h"
#include "vxh/stereo_match.h"

namespace {

constexpr uint32_t kRejected = 0xFFFFFFFFu;

// Masked Hamming distance of the shader's first kernel (StereoMatch.metal:46-59).
__device__ __forceinline__ uint32_t pair_distance(const uint32_t (&l)[8], float lx, float ly,
                                                  const vxh_corner_point &r, const uint4 *rd,
                                                  const vxh_stereo_config &p) {
    const float dy = fabsf(ly - r.y);
    if (!(dy <= p.max_epipolar)) return 0xFFFFu;
    const float d = lx - r.x;
    if (!(d >= p.min_disparity && d <= p.max_disparity)) return 0xFFFFu;
    const uint4 a = rd[0], b = rd[1];
    return __popc(l[0] ^ a.x) + __popc(l[1] ^ a.y) + __popc(l[2] ^ a.z) + __popc(l[3] ^ a.w) +
           __popc(l[4] ^ b.x) + __popc(l[5] ^ b.y) + __popc(l[6] ^ b.z) + __popc(l[7] ^ b.w);
}

__global__ void __launch_bounds__(256)
stereo_best_kernel(const vxh_corner_point *__restrict__ lk,
                   const vxh_corner_point *__restrict__ rk, const uint32_t *__restrict__ ld,
                   const uint4 *__restrict__ rd, uint32_t nl, uint32_t nr,
                   vxh_stereo_config p, vxh_stereo_match *__restrict__ out) {
    const uint32_t li =
        __builtin_amdgcn_readfirstlane(blockIdx.x * 4u + (threadIdx.x >> 6));
    const uint32_t lane = threadIdx.x & 63u;
    if (li >= nl) return;   // wave-uniform
    uint32_t l[8];
#pragma unroll
    for (int w = 0; w < 8; ++w) l[w] = ld[(size_t)li * 8 + w];
    const vxh_corner_point lp = lk[li];
    uint64_t best = ~0ull;
    for (uint32_t ri = lane; ri < nr; ri += 64) {
        const uint32_t dist = pair_distance(l, lp.x, lp.y, rk[ri], rd + (size_t)ri * 2, p);
        const uint64_t key = ((uint64_t)dist << 32) | ri;
        best = key < best ? key : best;
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const uint64_t o = __shfl_xor(best, off, 64);
        best = o < best ? o : best;
    }
    if (lane != 0) return;
    vxh_stereo_match m;
    m.left_idx = kRejected;
    m.right_idx = 0;
    m.disparity = 0.0f;
    m.point_3d[0] = m.point_3d[1] = m.point_3d[2] = 0.0f;
    m.pad = 0.0f;
    const uint32_t dist = (uint32_t)(best >> 32), ri = (uint32_t)best;
    if (dist <= (p.max_hamming & 0xFFFFu)) {
        // StereoMatch.metal:103-107, each product and quotient in that order.
        const float disp = lp.x - rk[ri].x;
        const float Z = p.fx * p.baseline / disp;
        m.left_idx = li;
        m.right_idx = ri;
        m.disparity = disp;
        m.point_3d[0] = Z * (lp.x - p.cx) / p.fx;
        m.point_3d[1] = Z * (lp.y - p.cy) / p.fy;
        m.point_3d[2] = Z;
    }
    out[li] = m;
}

// One block: keeps the accepted slots of out[0, n) in order, in place, and
// writes their number.  A chunk's reads all precede its writes (barrier), and
// its writes land below the next chunk's reads and above the last chunk's
// writes, so one pass suffices.
__global__ void __launch_bounds__(1024)
stereo_compact_kernel(vxh_stereo_match *out, uint32_t n, uint32_t *count) {
    __shared__ uint32_t wave_sum[16];
    __shared__ uint32_t chunk_total;
    const uint32_t t = threadIdx.x, lane = t & 63u, wave = t >> 6;
    uint32_t base_out = 0;
    for (uint32_t base = 0; base < n; base += 1024) {
        const uint32_t i = base + t;
        vxh_stereo_match m{};
        bool keep = false;
        if (i < n) {
            m = out[i];
            keep = m.left_idx != kRejected;
        }
        const uint64_t ballot = __ballot(keep);
        const uint32_t below = __popcll(ballot & ((1ull << lane) - 1ull));
        if (lane == 0) wave_sum[wave] = __popcll(ballot);
        __syncthreads();
        uint32_t before = 0, total = 0;
        for (uint32_t w = 0; w < 16; ++w) {
            before += w < wave ? wave_sum[w] : 0u;
            total += wave_sum[w];
        }
        if (keep) out[base_out + before + below] = m;
        if (t == 0) chunk_total = total;
        __syncthreads();
        base_out += chunk_total;
        __syncthreads();   // chunk_total and wave_sum are rewritten next chunk
    }
    if (t == 0) *count = base_out;
}

__global__ void __launch_bounds__(256)
stereo_matrix_kernel(const vxh_corner_point *__restrict__ lk,
                     const vxh_corner_point *__restrict__ rk, const uint32_t *__restrict__ ld,
                     const uint4 *__restrict__ rd, uint32_t nl, uint32_t nr,
                     vxh_stereo_config p, uint16_t *__restrict__ mat) {
    const uint32_t li = __builtin_amdgcn_readfirstlane(blockIdx.x * 4u + threadIdx.y);
    const uint32_t ri = blockIdx.y * 64u + threadIdx.x;
    if (li >= nl || ri >= nr) return;
    uint32_t l[8];
#pragma unroll
    for (int w = 0; w < 8; ++w) l[w] = ld[(size_t)li * 8 + w];
    const vxh_corner_point lp = lk[li];
    mat[(size_t)li * nr + ri] =
        (uint16_t)pair_distance(l, lp.x, lp.y, rk[ri], rd + (size_t)ri * 2, p);
}

vxh_status invalid(const std::string &msg) {
    return vxh::report_error(VXH_ERR_INVALID_CONFIG, "invalid config: " + msg);
}

vxh_status launched(const char *what) {
    const hipError_t e = hipGetLastError();
    if (e == hipSuccess) return VXH_OK;
    return vxh::report_error(VXH_ERR_GPU, std::string("GPU error: ") + what + ": " +
                                              hipGetErrorString(e));
}

vxh_status check_buffers(const void *lk, const void *rk, const void *ld, const void *rd,
                         uint32_t nl, uint32_t nr, const vxh_stereo_config *cfg) {
    if (!cfg) return invalid("null config");
    if (nl >= (1u << 31) || nr >= (1u << 31)) return invalid("keypoint count over 2^31");
    if (nl && nr && (!lk || !rk || !ld || !rd)) return invalid("null device buffer");
    if (((uintptr_t)ld | (uintptr_t)rd) % 16) return invalid("descriptors need 16-byte alignment");
    return VXH_OK;
}

}  // namespace

struct vxh_stereo_matcher {
    vxh_context *ctx;
};

extern "C" {

void vxh_stereo_config_default(vxh_stereo_config *c) {
    if (!c) return;
    c->max_epipolar = 2.0f;
    c->min_disparity = 1.0f;
    c->max_disparity = 120.0f;
    c->max_hamming = 50;
    c->ratio_thresh = 0.8f;
    c->fx = 600.0f;
    c->fy = 600.0f;
    c->cx = 320.0f;
    c->cy = 240.0f;
    c->baseline = 0.12f;
}

vxh_status vxh_stereo_match_enqueue(const vxh_corner_point *d_lk, const vxh_corner_point *d_rk,
                                    const uint32_t *d_ld, const uint32_t *d_rd, uint32_t nl,
                                    uint32_t nr, const vxh_stereo_config *cfg,
                                    vxh_stereo_match *d_matches, uint32_t *d_count,
                                    void *stream) {
    vxh_status v = check_buffers(d_lk, d_rk, d_ld, d_rd, nl, nr, cfg);
    if (v != VXH_OK) return v;
    if (!d_count || (nl && !d_matches)) return invalid("null output buffer");
    const hipStream_t s = (hipStream_t)stream;
    if (nl == 0 || nr == 0) {   // stereomatch.rs:137-141: no matches
        if (vxh::launch_zero(d_count, 1, s) != hipSuccess) return launched("stereo count");
        return VXH_OK;
    }
    hipLaunchKernelGGL(stereo_best_kernel, dim3((nl + 3) / 4), dim3(256), 0, s, d_lk, d_rk,
                       d_ld, (const uint4 *)d_rd, nl, nr, *cfg, d_matches);
    if ((v = launched("stereo_best")) != VXH_OK) return v;
    hipLaunchKernelGGL(stereo_compact_kernel, dim3(1), dim3(1024), 0, s, d_matches, nl,
                       d_count);
    return launched("stereo_compact");
}

vxh_status vxh_stereo_hamming_matrix_enqueue(const vxh_corner_point *d_lk,
                                             const vxh_corner_point *d_rk,
                                             const uint32_t *d_ld, const uint32_t *d_rd,
                                             uint32_t nl, uint32_t nr,
                                             const vxh_stereo_config *cfg, uint16_t *d_matrix,
                                             void *stream) {
    vxh_status v = check_buffers(d_lk, d_rk, d_ld, d_rd, nl, nr, cfg);
    if (v != VXH_OK) return v;
    if (nl == 0 || nr == 0) return VXH_OK;
    if (!d_matrix) return invalid("null matrix");
    if (nr > 65535u * 64u) return invalid("matrix rows over 4194240 entries");
    hipLaunchKernelGGL(stereo_matrix_kernel, dim3((nl + 3) / 4, (nr + 63) / 64), dim3(64, 4), 0,
                       (hipStream_t)stream, d_lk, d_rk, d_ld, (const uint4 *)d_rd, nl, nr, *cfg,
                       d_matrix);
    return launched("stereo_matrix");
}

vxh_status vxh_stereo_matcher_create(vxh_context *ctx, vxh_stereo_matcher **out) {
    if (!ctx || !out) return invalid("null argument");
    *out = nullptr;
    if (hipSetDevice(vxh_context_device(ctx)) != hipSuccess)
        return vxh::report_error(VXH_ERR_DEVICE_NOT_FOUND, "no HIP device found");
    *out = new vxh_stereo_matcher{ctx};
    return VXH_OK;
}

void vxh_stereo_matcher_destroy(vxh_stereo_matcher *m) { delete m; }

vxh_status vxh_stereo_matcher_run(vxh_stereo_matcher *m, vxh_context *ctx,
                                  const vxh_corner_point *lk, size_t nl,
                                  const vxh_corner_point *rk, size_t nr,
                                  const vxh_orb_output *ldesc, const vxh_orb_output *rdesc,
                                  const vxh_stereo_config *cfg, vxh_stereo_match *out,
                                  size_t *n_out) {
    if (!m || !ctx || !cfg || !n_out) return invalid("null argument");
    *n_out = 0;
    if (nl == 0 || nr == 0) return VXH_OK;
    if (!lk || !rk || !ldesc || !rdesc || !out) return invalid("null argument");
    if (nl >= (1u << 31) || nr >= (1u << 31)) return invalid("keypoint count over 2^31");
    // One scratch block: keypoints, flattened descriptors, matches, count.
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t o_rk = up(nl * sizeof(vxh_corner_point));
    const size_t o_ld = o_rk + up(nr * sizeof(vxh_corner_point));
    const size_t o_rd = o_ld + up(nl * 32);
    const size_t o_m = o_rd + up(nr * 32);
    const size_t o_c = o_m + up(nl * sizeof(vxh_stereo_match));
    const size_t total = o_c + 256;
    std::vector<uint32_t> words;
    try {
        words.resize((nl + nr) * 8);
    } catch (const std::bad_alloc &) {
        return vxh::report_error(VXH_ERR_BUFFER_ALLOC, "buffer allocation failed (descriptors)");
    }
    for (size_t i = 0; i < nl; ++i) std::memcpy(&words[i * 8], ldesc[i].desc, 32);
    for (size_t i = 0; i < nr; ++i) std::memcpy(&words[(nl + i) * 8], rdesc[i].desc, 32);
    vxh::CaptureRelaxed relaxed;
    if (hipSetDevice(vxh_context_device(ctx)) != hipSuccess)
        return vxh::report_error(VXH_ERR_DEVICE_NOT_FOUND, "no HIP device found");
    void *blk = nullptr;
    vxh_status v = vxh::ctx_take_block(ctx, total, &blk, nullptr);
    if (v != VXH_OK) return v;
    hipStream_t s = nullptr;
    if ((v = vxh::ctx_take_stream(ctx, &s)) != VXH_OK) {
        vxh::ctx_give_block(ctx, blk);
        return v;
    }
    char *g = (char *)blk;
    uint32_t count = 0;
    hipError_t e = hipMemcpyAsync(g, lk, nl * sizeof(vxh_corner_point), hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(g + o_rk, rk, nr * sizeof(vxh_corner_point), hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(g + o_ld, words.data(), nl * 32, hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(g + o_rd, words.data() + nl * 8, nr * 32, hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        v = vxh_stereo_match_enqueue((const vxh_corner_point *)g,
                                     (const vxh_corner_point *)(g + o_rk),
                                     (const uint32_t *)(g + o_ld), (const uint32_t *)(g + o_rd),
                                     (uint32_t)nl, (uint32_t)nr, cfg,
                                     (vxh_stereo_match *)(g + o_m), (uint32_t *)(g + o_c), s);
    if (e == hipSuccess && v == VXH_OK)
        e = hipMemcpyAsync(&count, g + o_c, 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e == hipSuccess && v == VXH_OK && count > 0)
        e = hipMemcpyWithStream(out, g + o_m, (size_t)count * sizeof(vxh_stereo_match),
                                hipMemcpyDeviceToHost, s);
    (void)hipStreamSynchronize(s);
    vxh::ctx_give_stream(ctx, s);
    vxh::ctx_give_block(ctx, blk);
    if (v != VXH_OK) return v;
    if (e != hipSuccess)
        return vxh::report_error(VXH_ERR_GPU, std::string("GPU error: stereo match: ") +
                                                  hipGetErrorString(e));
    *n_out = count < nl ? count : nl;
    return VXH_OK;
}

}  // extern "C"
