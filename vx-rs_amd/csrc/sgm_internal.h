// Internal declarations shared by the SGM kernels and the host runtime.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace vxh {

// Path directions in path-index order.  num_paths 2 uses the first two (the
// reference's scan_dirs, vx-vision/src/kernels/sgm.rs:197-200), 4 the first
// four, 8 all of them.
constexpr int kDirX[8] = {+1, 0, -1, 0, +1, -1, +1, -1};
constexpr int kDirY[8] = {0, +1, 0, -1, +1, +1, -1, -1};

// Number of scanlines of direction `dir` in a w x h image: every pixel whose
// predecessor lies outside the image starts one.
inline int num_scanlines(int dir, int w, int h) {
    int dx = kDirX[dir], dy = kDirY[dir];
    if (dy == 0) return h;
    if (dx == 0) return w;
    return w + h - 1;
}

// Per-pixel stride, in u16 elements, of the summed cost volume: D rounded up
// to 16 so every pixel's disparity slice starts 32-byte aligned and each lane's
// packed segment of it never straddles a pixel.
inline int volume_stride(int D) { return (D + 15) & ~15; }

// Lane layout of the aggregation kernel: L lanes own one scanline, each lane
// K consecutive disparities (L * K >= D).
struct Layout {
    int L, K;
};

// The aggregation reads the right census at x - d for every d of a lane
// through one base kCensusGuard bytes before rc; the reads with x - d < 0
// (whose cost is replaced by 64) then land in the kCensusGuard readable bytes
// that every rc buffer must have in front of it (the left census, or padding).
constexpr size_t kCensusGuard = 2048;

// Bytes reserved for the left census in front of the right one, so that rc
// has kCensusGuard readable bytes before it.
inline size_t census_span(size_t npx) {
    const size_t b = npx * 8 > kCensusGuard ? npx * 8 : kCensusGuard;
    return (b + 255) & ~(size_t)255;
}

struct AggParams {
    const uint2 *lc;       // left census, w*h (lo, hi)
    const uint2 *rc;       // right census, kCensusGuard readable bytes before it
    uint16_t *vol;         // summed volume, (y*w + x)*Dp + d
    int w, h, D, Dp;
    uint32_t p1, p2;       // penalties already truncated to u16
    int dir;
    int nscan;
    int accumulate;        // 0: store (first path), 1: add mod 2^16
    float *wta_out;        // non-null on the last path: winner-take-all is
                           // computed from the final sums instead of storing them
};

// Several independent direction passes in one launch (each into its own
// volume).  Blocks [first_block, first_block + nblocks) belong to a pass;
// nblocks is a multiple of 8 so that the hardware's round-robin placement of
// workgroups over the 8 XCDs can be undone (each XCD gets a contiguous range
// of the pass's scanlines and shares their census rows in its L2).
constexpr int kMaxPasses = 8;
struct PassDesc {
    uint16_t *vol;
    int dir, dx, dy, nscan;
    int accumulate;
    int first_block, nblocks;
};
struct MultiAgg {
    AggParams base;        // shared fields; vol/dir/nscan/accumulate come per pass
    int npass;
    PassDesc pass[kMaxPasses];
};

hipError_t launch_census(const uint8_t *img, int w, int h, int rx, int ry,
                         uint2 *out, hipStream_t s);
// Default lane layout for D disparities.
Layout default_layout(int D);
bool layout_supported(Layout l, int D);
hipError_t launch_aggregate(const AggParams &p, Layout l, hipStream_t s);
// Passes `dirs[0..n)` at once, pass j into `vols[j]` (accumulating when
// `accumulate[j]`); no winner-take-all.  base supplies the shared fields.
hipError_t launch_aggregate_multi(const AggParams &base, const int *dirs,
                                  uint16_t *const *vols, const int *accumulate, int n,
                                  Layout l, hipStream_t s);
// Sums `nvol` volumes (vol0 + k * vol_stride, k < nvol) mod 2^16 and picks
// each pixel's winner with parabola refinement; with `sum_out` the summed
// volume is also stored there (may equal vol0), with `best` the integer winner.
hipError_t launch_wta_sum(const uint16_t *vol0, size_t vol_stride, int nvol, int npx,
                          int D, int Dp, float *out, int16_t *best, uint16_t *sum_out,
                          hipStream_t s);
hipError_t launch_copy(void *dst, const void *src, size_t bytes, hipStream_t s);
hipError_t launch_wta(const uint16_t *vol, int w, int h, int D, int Dp,
                      float *out, int16_t *best, hipStream_t s);
hipError_t launch_lr_check(const uint16_t *vol, int w, int h, int D, int Dp,
                           int tol, const int16_t *best, float *out,
                           hipStream_t s);

}  // namespace vxh
