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

// Longest scanline of direction `dir`: the passes with the longest scanlines
// are placed first in a launch so that their waves start first.
inline int scan_length(int dir, int w, int h) {
    if (kDirY[dir] == 0) return w;
    if (kDirX[dir] == 0) return h;
    return w < h ? w : h;
}

// Per-pixel stride, in u16 elements, of the summed cost volume: D rounded up
// to 16 so every pixel's disparity slice starts 32-byte aligned and each lane's
// packed segment of it never straddles a pixel.
inline int volume_stride(int D) { return (D + 15) & ~15; }

// Storage of the per-direction path volumes (entries pix*Dp + d).
constexpr int kOutU8 = 0;    // low byte, plus high bytes of flagged slices
constexpr int kOutU16 = 1;   // full u16 path cost

// One direction pass of the packed-pair aggregation.
struct PassParams {
    const uint8_t *cost;   // matching costs + bias, entry pix*Dp + d (u8, 64 + bias where
                           // x - d < 0)
    uint8_t *vol;          // this direction's volume
    uint8_t *hi;           // kOutU8: high bytes of flagged slices (1 byte per entry)
    uint8_t *flags;        // kOutU8: 8 bytes per pixel, byte dir = 1 when flagged
    int w, h, D, Dp;
    uint32_t p1x2, p2x2;   // penalties truncated to u16, in both halves
    uint32_t xqx2;         // P1 - the cost volume's bias, in both halves
    int dir, dx, dy, nscan;
    int mode;              // kOutU8 / kOutU16 (read only where the kernel takes it at run time)
    uint32_t wait_polls;   // split heads: polls of a tail's flag before the error word is set
};

// Lane layouts of the aggregation: layout class c covers D <= 64 << c with
// the (L, K) choices kClassLayouts[c][0 .. kClassCount[c]), L lanes owning a
// scanline pair and K disparities each (L * K = 64 << c).
constexpr int kLayoutClasses = 3;
constexpr int kLayoutsPerClass = 3;
constexpr int kClassCount[kLayoutClasses] = {3, 3, 2};
constexpr int kClassLayouts[kLayoutClasses][kLayoutsPerClass][2] = {
    {{32, 2}, {16, 4}, {8, 8}},
    {{64, 2}, {32, 4}, {16, 8}},
    {{64, 4}, {32, 8}, {32, 8}},
};
inline int layout_class(int D) { return D <= 64 ? 0 : (D <= 128 ? 1 : 2); }

// Lane layout of one direction class of a pair: horizontal passes (the
// longest scanlines) and the other directions are laid out separately.
// Polls of a head's bounded wait for its tail's flag before it gives up and
// raises the error word (seconds of polling).
constexpr uint32_t kSplitWaitPolls = 1u << 22;
struct Schedule {
    int horiz;   // index into kClassLayouts[layout_class(D)]
    int other;
    int split;   // long scanlines run as head and tail halves at once: 1 those of
                 // the longest passes, 2 those of every pass; 0 none
    uint32_t wait_polls = kSplitWaitPolls;   // the heads' bounded wait
};

// Words of the split launch's synchronisation area for a w x h pair: an
// unused word, the error flag, one ticket counter per XCD (8), then two flags
// (done, early) per scanline group (at most 8 passes of ceil(scanlines / 2)
// groups).
constexpr size_t kSyncFlags = 10;
inline size_t sync_words(uint32_t w, uint32_t h) {
    return kSyncFlags + 16 * ((size_t)(w + h) / 2 + 1);
}

constexpr size_t kWsAlign = 256;
inline size_t ws_align(size_t v) { return (v + kWsAlign - 1) & ~(kWsAlign - 1); }

// Scratch of one pair, byte offsets from its base: the synchronisation words
// first (word 1 is the pair's error word, so its offset never depends on the
// format: kErrorWordOffset), then the census of both images, the
// matching-cost volume (u8, pix*Dp + d), the exception flags (u8 volumes
// only, 8 bytes per pixel), one volume per path direction (u8 or u16
// entries, vol_bytes apart), the high-byte volumes (u8 only, also vol_bytes
// apart), and for the LR check the winner index (the right view's keys go
// to the left census's area, free by then).
struct WsLayout {
    size_t lc, rc, cost, flags, sync, vol, vol_bytes, hi, best, total;
};
constexpr size_t kErrorWordOffset = 4;
inline WsLayout workspace_layout(uint32_t w, uint32_t h, uint32_t D, int paths, int mode,
                                 bool lr) {
    WsLayout l;
    const size_t n = (size_t)w * h;
    const size_t Dp = (size_t)volume_stride((int)D);
    const bool u8 = mode == kOutU8;
    l.sync = 0;
    l.lc = l.sync + ws_align(sync_words(w, h) * 4);
    l.rc = l.lc + ws_align(n * 8);
    l.cost = l.rc + ws_align(n * 8);
    l.flags = l.cost + ws_align(n * Dp);
    l.vol = l.flags + (u8 ? ws_align(n * 8) : 0);
    l.vol_bytes = ws_align(n * Dp * (u8 ? 1 : 2));
    l.hi = l.vol + (size_t)paths * l.vol_bytes;
    l.best = l.hi + (u8 ? (size_t)paths * l.vol_bytes : 0);
    l.total = l.best + (lr ? ws_align(n * 2) : 0);
    return l;
}

// Volume format a configuration selects by default: u8 entries whenever
// every path cost that has not wrapped fits a byte (costs never exceed
// 64 + P2, SGMStereo.metal:154-160), u16 otherwise.
inline int auto_volume_mode(float p2) {
    return ((uint32_t)p2 & 0xFFFFu) + 64u <= 255u ? kOutU8 : kOutU16;
}

// Several independent direction passes in one launch (each into its own
// volume).  Blocks [first_block, first_block + nblocks) belong to a pass;
// nblocks is a multiple of 8 so that the hardware's round-robin placement of
// workgroups over the 8 XCDs can be undone (each XCD gets a contiguous range
// of the pass's scanlines and shares their cost rows in its L2).
constexpr int kMaxPasses = 8;
struct PassDesc {
    uint8_t *vol, *hi;
    int dir, dx, dy, nscan;
    int layout;
    int first_block, nblocks;   // unsplit launches: blocks of this pass
    int nwaves;                 // scanline groups (waves per role)
    int first_wave;             // split launches: index of the pass's first group
    int split;                  // split launches: this pass runs tail and head roles
};
struct MultiPass {
    PassParams base;       // shared fields; vol/hi/dir/dx/dy/nscan come per pass
    int npass;
    int split;             // some pass runs as tail and head roles (ticketed launch)
    int total_waves;       // split: scanline groups over all passes
    uint32_t *sync;        // split: [1] error, [2 + k] XCD k's ticket counter,
                           // [kSyncFlags + g] done and [kSyncFlags + total_waves + g]
                           // early flag of group g's tail
    PassDesc pass[kMaxPasses];
};


// ---- Multi-pair launches of the batched form ----
// Each launch covers the blocks of all its jobs; jobs are sorted by
// first_block, the first at 0, and block b belongs to the job whose
// [first_block, next job's first_block) holds it.
// Rows of one census tile (64 pixels wide; blocks of 64 x 4 threads).
constexpr int kCensusTileH = 16;
struct CensusJob {   // one image: tiles of 64 x kCensusTileH pixels, nbx per tile row
    const uint8_t *img;
    uint2 *out;
    uint2 *zero_flags;       // as launch_census
    uint32_t *zero_words;
    size_t nwords;
    int w, h, rx, ry, nbx;
    uint32_t first_block;
};
struct CostJob {     // one pair: tiles of 256 pixels of one row, nbx per row
    const uint2 *lc, *rc;
    uint8_t *cost;
    int w, h, D, Dp;
    uint32_t bias;
    int nbx;
    uint32_t first_block;
};
struct WtaJob {      // one pair, as launch_wta_sum (16 pixels per block) or, with
                     // rkey, as launch_wta_lr (wta_job_blocks blocks)
    const uint8_t *vol0;
    size_t vstride;
    const uint8_t *hi0;
    const uint2 *flags;
    float *out;
    int16_t *best;
    uint32_t *rkey;
    int nv, mode, npx, w, D, Dp;
    uint32_t first_block;
};
// Blocks of a winner job, and its kind for launch_wta_multi.
uint32_t wta_job_blocks(const WtaJob &j);
constexpr int kWtaOne = 1, kWtaTwo = 2, kWtaLr = 4;
int wta_job_kind(const WtaJob &j);
hipError_t launch_census_multi(const CensusJob *d_jobs, int njobs, uint32_t blocks,
                               hipStream_t s);
hipError_t launch_cost_multi(const CostJob *d_jobs, int njobs, uint32_t blocks, hipStream_t s);
// kinds: the bitwise or of wta_job_kind over the jobs (one launch per kind);
// every job without right-view keys must have volume format `mode`.
hipError_t launch_wta_multi(const WtaJob *d_jobs, int njobs, uint32_t blocks, int kinds,
                            int mode, hipStream_t s);

// Persistent aggregation over the direction passes of many pairs.  A pass is
// one direction of one pair (p carries its volume, direction and scanline
// count; layout indexes its class's lane layouts).  Item (pass, wave) is the
// scanline group `wave` of a pass.  Waves of block b take tickets from
// counters[b % 8] (the XCD the hardware places the block on) and run item
// items[item_off[k] + ticket] of that XCD's list until it is exhausted; the
// host orders each list pair by pair, so an XCD works through contiguous
// scanlines of one pair's passes at a time.  counters must be zero at launch.
struct BatchPass {
    PassParams p;
    int layout;
};
struct BatchItem {
    uint32_t pass, wave;
};
struct BatchPaths {
    const BatchPass *passes;
    const BatchItem *items;
    uint32_t *counters;
    uint32_t item_off[8], item_cnt[8];
};
hipError_t launch_batch_paths_c0(int mode, const BatchPaths &b, int blocks, hipStream_t s);
hipError_t launch_batch_paths_c1(int mode, const BatchPaths &b, int blocks, hipStream_t s);
hipError_t launch_batch_paths_c2(int mode, const BatchPaths &b, int blocks, hipStream_t s);
// Blocks of 256 threads of the class's persistent kernel one CU holds at once.
int batch_paths_occupancy_c0(int mode);
int batch_paths_occupancy_c1(int mode);
int batch_paths_occupancy_c2(int mode);

// ---- The two-sweep raster form of the batched path ----
// A forward sweep aggregates the four directions whose predecessor lies on the
// current or the previous row (+1,0) (0,+1) (+1,+1) (-1,+1) over the image in
// raster order and stores their sum (u16, pix*Dp + d); a backward sweep runs
// the other four over the image turned by 180 degrees, adds the forward sum
// and takes each pixel's winner, so the summed volume is never stored.  A
// wave owns G = 64 / L groups of L lanes, each group two consecutive rows
// packed into the halves of every register (K = 4 disparities per lane);
// row r runs two steps behind row r - 1.  Row states cross a wave boundary
// through a per-wave ring of slots in global memory (one slot per column,
// kSweepRec bytes per lane) with a progress word, inside one XCD's L2.
constexpr int kSweepK = 4;
// Polls of a sweep wave's bounded wait for the wave above it (tens to a
// hundred milliseconds; a predecessor always runs, so real waits are
// microseconds).
constexpr uint32_t kSweepWaitPolls = 1u << 16;
constexpr int kSweepRec = 48;   // per lane and slot: three directions x (4 u16, pm, pad)
inline int sweep_lanes(int D) { return 16 << layout_class(D); }
inline int sweep_groups(int D) { return 64 / sweep_lanes(D); }
// Waves of one sweep of a w x h pair.
inline int sweep_waves(int h, int D) {
    const int G = sweep_groups(D);
    return ((h + 1) / 2 + G - 1) / G;
}
inline size_t sweep_slot_bytes(int D) {
    return ((size_t)sweep_lanes(D) * kSweepRec + 127) & ~(size_t)127;
}
constexpr size_t sweep_ring_slots(int w) { return (size_t)w + 4; }
struct SweepJob {
    const uint8_t *cost;   // matching costs + bias (as PassParams::cost)
    uint16_t *sf;          // forward sums, pix*Dp + d
    uint8_t *ring;         // nwaves rings of sweep_ring_slots(w) slots
    uint32_t *prog;        // nwaves progress words (cleared before each sweep)
    float *out;            // disparity map
    int16_t *best;         // LR check: the winners (else null)
    uint16_t *sum;         // LR check: the summed volume (else null)
    int w, h, D, Dp;
    uint32_t p1x2, p2x2, xqx2;   // as PassParams
    int nwaves;
    uint32_t slot_bytes;
};
struct SweepItem {
    uint32_t job, wave;
};
// One sweep over every job of one lane-layout class.  Waves take items by
// ticket from the list of the XCD they run on (read from the hardware), each
// list holding whole pairs, ordered by wave index so that a wave's
// predecessor always holds an earlier ticket of the same list.
struct SweepArgs {
    const SweepJob *jobs;
    const SweepItem *items;
    uint32_t *counters;         // 8 ticket counters, zero at launch
    uint32_t *error;            // raised when a bounded wait expires
    const uint8_t *fresh;       // zeroed slot (the row above the image)
    uint32_t item_off[8], item_cnt[8];
    uint32_t wait_polls;
    uint32_t zero;              // always 0 (orders a progress store after a load)
};
hipError_t launch_sweep_c0(bool bwd, const SweepArgs &a, int blocks, hipStream_t s);
hipError_t launch_sweep_c1(bool bwd, const SweepArgs &a, int blocks, hipStream_t s);
hipError_t launch_sweep_c2(bool bwd, const SweepArgs &a, int blocks, hipStream_t s);
// One-wave blocks of the class's sweep kernel one CU holds at once.
int sweep_occupancy_c0(bool bwd);
int sweep_occupancy_c1(bool bwd);
int sweep_occupancy_c2(bool bwd);

// Census of one image; with `zero_flags` also clears the 8 flag bytes of
// every pixel (the kOutU8 exception flags of the pass that follows), with
// `zero_words` the first nwords words there (the split launch's
// synchronisation area, which launch_paths expects cleared).
hipError_t launch_census(const uint8_t *img, int w, int h, int rx, int ry,
                         uint2 *out, uint2 *zero_flags, uint32_t *zero_words, size_t nwords,
                         hipStream_t s);
// launch_census of both images of a pair in one launch (the left one with
// the clearing of launch_census).
hipError_t launch_census_pair(const uint8_t *left, const uint8_t *right, int w, int h, int rx,
                              int ry, uint2 *out_l, uint2 *out_r, uint2 *zero_flags,
                              uint32_t *zero_words, size_t nwords, hipStream_t s);
// Matching-cost volume of a pair from its census (layout of
// PassParams::cost, Dp = volume_stride(D)), every entry plus `bias`
// (cost_bias(p1)).
hipError_t launch_cost(const uint2 *lc, const uint2 *rc, int w, int h, int D, uint32_t bias,
                       uint8_t *cost, hipStream_t s);
// The bias the cost volume carries: P1 itself when cost + P1 fits a byte
// (costs are at most 64), else 0.  The recurrence adds the rest of P1.
inline uint32_t cost_bias(uint32_t p1_u16) { return p1_u16 <= 191u ? p1_u16 : 0u; }
// *dst |= *src with a kernel (stream-ordered, safe inside captured graphs).
hipError_t launch_or_word(uint32_t *dst, const uint32_t *src, hipStream_t s);
// Zeroes n words with a kernel (safe inside captured graphs).
hipError_t launch_zero(uint32_t *p, size_t n, hipStream_t s);
// Default schedule for D disparities: the fastest for a lone 8-path pair,
// and the per-pair schedule of the batch forms.
Schedule default_schedule(int D);
// Default schedule of a lone w x h pair with `paths` paths and D disparities.
Schedule plain_schedule(int D, int paths, uint32_t w, uint32_t h);
bool schedule_valid(Schedule sc, int D);
// The aggregation launchers of layout classes 0, 1 and 2 (one translation
// unit each): paths_split_kernel when `split`, else paths_kernel, in storage
// `mode`, over `blocks` blocks of 256 threads.
hipError_t launch_paths_c0(int mode, bool split, const MultiPass &m, int blocks, hipStream_t s);
hipError_t launch_paths_c1(int mode, bool split, const MultiPass &m, int blocks, hipStream_t s);
hipError_t launch_paths_c2(int mode, bool split, const MultiPass &m, int blocks, hipStream_t s);
// Runs the passes dirs[0..n) at once, pass j into vols[j] (and his[j] for
// kOutU8).  base supplies the shared fields; `sync` (sync_words words,
// cleared beforehand) enables the split launch.
hipError_t launch_paths(const PassParams &base, const int *dirs, uint8_t *const *vols,
                        uint8_t *const *his, int n, Schedule sc, int mode, uint32_t *sync,
                        hipStream_t s);
// Sums `nvol` per-direction volumes (vol0 + k * vol_stride bytes, k < nvol;
// kOutU8 adds 256 * the high bytes of flagged slices, his at the same
// offsets from hi0) mod 2^16 and picks each pixel's winner with parabola
// refinement; with `sum_out` the summed u16 volume is also stored, with
// `best` the integer winners.
hipError_t launch_wta_sum(const uint8_t *vol0, size_t vol_stride, const uint8_t *hi0,
                          const uint8_t *flags, int nvol, int mode, int npx, int D, int Dp,
                          float *out, int16_t *best, uint16_t *sum_out, hipStream_t s);
// launch_wta_sum for a pair with the LR check: the integer winners go to
// `best` (w * h) and the right view's winners, as keys, to `rkey` (w * h
// words, zeroed beforehand); launch_lr_apply then checks `out` against them.
hipError_t launch_wta_lr(const uint8_t *vol0, size_t vol_stride, const uint8_t *hi0,
                         const uint8_t *flags, int nvol, int mode, int w, int h, int D, int Dp,
                         float *out, int16_t *best, uint32_t *rkey, hipStream_t s);
hipError_t launch_lr_apply(const int16_t *best, const uint32_t *rkey, int w, int h, int tol,
                           float *out, hipStream_t s);
hipError_t launch_copy(void *dst, const void *src, size_t bytes, hipStream_t s);
// Left-right check of a pair's map `out` from its summed u16 volume and
// winners; `dr` holds w * h int16 of scratch (the right view's winners; the
// census area of the workspace, free once the costs are computed).
hipError_t launch_lr_check(const uint16_t *vol, int w, int h, int D, int Dp,
                           int tol, const int16_t *best, float *out, int16_t *dr,
                           hipStream_t s);

}  // namespace vxh
