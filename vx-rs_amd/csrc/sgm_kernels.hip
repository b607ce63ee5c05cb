// Per-pair SGM kernels for gfx950: thin grid wrappers around the device
// functions in sgm_device.h / sgm_paths.h, and their host-side launchers.
#include <algorithm>

#include "sgm_paths.h"

namespace vxh {
namespace {

using namespace dev;

// Census of 64 x 4 pixels per block (blockDim (64, 4)).  The default 9 x 7
// window reads its 10 x 72 tile of clamped image rows once into LDS, which
// reproduces clamp-to-edge exactly, so every pixel then takes the same path:
// 63 LDS byte reads at constant offsets, and each bit shifted in with one
// subtraction (n - c has bit 31 set iff n < c) and one alignbit, from the last
// neighbour down so that neighbour k ends at bit k.  Other windows read the
// image directly.
constexpr int kCtW = 64 + 8, kCtH = 4 + 6;

// Tile (bx, by) of the census of one image (block-uniform arguments).
__device__ __forceinline__ void census_tile(const uint8_t *__restrict__ img, int w, int h, int rx,
                                            int ry, uint2 *__restrict__ out,
                                            uint2 *__restrict__ zero_flags,
                                            uint32_t *__restrict__ zero_words, size_t nwords,
                                            int bx, int by) {
    const int x = bx * 64 + threadIdx.x;
    const int y = by * 4 + threadIdx.y;
    uint2 word;
    if (rx == 4 && ry == 3) {   // block-uniform
        __shared__ uint8_t tile[kCtH * kCtW];
        const int x0 = bx * 64 - 4, y0 = by * 4 - 3;
        for (int t = threadIdx.y * 64 + threadIdx.x; t < kCtH * kCtW; t += 256) {
            const int r = t / kCtW, c = t - r * kCtW;
            tile[t] = img[(size_t)min(max(y0 + r, 0), h - 1) * w + min(max(x0 + c, 0), w - 1)];
        }
        __syncthreads();
        if (x >= w || y >= h) return;
        const uint8_t *win = tile + threadIdx.y * kCtW + threadIdx.x;   // window's top-left
        const uint32_t c = win[3 * kCtW + 4];
        uint32_t lo = 0, hi = 0;
#pragma unroll
        for (int k = 61; k >= 0; --k) {
            const int j = k < 31 ? k : k + 1;   // window index, the centre (31) skipped
            const uint32_t d = (uint32_t)win[(j / 9) * kCtW + j % 9] - c;
            if (k >= 32) hi = __builtin_amdgcn_alignbit(hi, d, 31);
            else lo = __builtin_amdgcn_alignbit(lo, d, 31);
        }
        word = make_uint2(lo, hi);
    } else {
        if (x >= w || y >= h) return;
        word = census_at(img, w, h, rx, ry, x, y);
    }
    const size_t i = (size_t)y * w + x;
    out[i] = word;
    if (zero_flags) zero_flags[i] = make_uint2(0, 0);
    for (size_t k = i; k < nwords; k += (size_t)w * h) zero_words[k] = 0;
}

__global__ void __launch_bounds__(256)
census_kernel(const uint8_t *__restrict__ img, int w, int h, int rx, int ry,
              uint2 *__restrict__ out, uint2 *__restrict__ zero_flags,
              uint32_t *__restrict__ zero_words, size_t nwords) {
    census_tile(img, w, h, rx, ry, out, zero_flags, zero_words, nwords, (int)blockIdx.x,
                (int)blockIdx.y);
}

// Index of the table entry whose [first_block, next first_block) holds block
// b (entries sorted by first_block, the first at 0; b wave-uniform).
template <typename E>
__device__ __forceinline__ int find_entry(const E *__restrict__ t, int n, uint32_t b) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (t[mid].first_block <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Census of many images in one launch (blockDim (64, 4)): block b belongs to
// the entry whose block range holds it.
__global__ void __launch_bounds__(256)
census_multi_kernel(const CensusJob *__restrict__ jobs, int njobs) {
    const int j = find_entry(jobs, njobs, blockIdx.x);
    const CensusJob &c = jobs[j];
    const int local = (int)(blockIdx.x - c.first_block);
    census_tile(c.img, c.w, c.h, c.rx, c.ry, c.out, c.zero_flags, c.zero_words, c.nwords,
                local % c.nbx, local / c.nbx);
}

// Matching-cost volume of a pair from its census (SGMStereo.metal:133-142),
// computed once and read by every direction pass.  A block (4 waves) owns
// 256 pixels of one row: it stages the left census words of the tile and the
// right census words the tile reads (the tile and the D - 1 pixels before
// it) in LDS, then each wave computes the cost slices of its 64 pixels, one
// pixel per lane (a lane reads 16 consecutive right-census words per chunk
// at constant offsets), into an LDS tile that it copies out as contiguous
// 16-byte chunks.  Slices wider than 128 entries go in two 128-entry passes,
// each a whole 128-byte line per pixel.
constexpr int kCcTile = 256;                      // pixels per block
constexpr int kCcSeg = 128;                       // slice entries per pass
constexpr int kCcOutRow = kCcSeg + 16;            // LDS row of one pixel (bank spread)
constexpr int kCcSpanR = kCcTile + 256;           // right census words (D - 1 <= 255 before)
constexpr size_t kCcLds = (size_t)(kCcSpanR + kCcTile) * 8 + 4 * 64 * kCcOutRow;

// Tile bx of row y of a pair's cost volume (block-uniform arguments; `lds`
// holds kCcLds bytes).
__device__ __forceinline__ void cost_tile(const uint2 *__restrict__ lc,
                                          const uint2 *__restrict__ rc, int w, int h, int D,
                                          int Dp, uint32_t bias, uint8_t *__restrict__ cost,
                                          int bx, int y, uint8_t *lds) {
    uint2 *cen_r = (uint2 *)lds;                            // kCcSpanR words
    uint2 *cen_l = cen_r + kCcSpanR;                        // kCcTile words
    uint8_t *out_all = (uint8_t *)(cen_l + kCcTile);        // 4 x 64 x kCcOutRow, 16-aligned
    const int t = threadIdx.x;
    const int x0 = bx * kCcTile;
    const int xs0 = max(0, x0 - (D - 1));                   // first right-census pixel
    const int xs1 = min(w, x0 + kCcTile);
    const int nr = xs1 - xs0, nl = xs1 - x0;
    {
        const size_t row = (size_t)y * w;
        for (int i = t; i < nr + nl; i += 256) {
            if (i < nr) cen_r[i] = rc[row + xs0 + i];
            else cen_l[i - nr] = lc[row + x0 + i - nr];
        }
    }
    __syncthreads();
    const int wv = t >> 6, lane = t & 63;
    const int p0 = wv * 64;                                 // first pixel of this wave
    const int npx = min(64, nl - p0);
    if (npx <= 0) return;                                   // wave-uniform
    uint8_t *out = out_all + wv * 64 * kCcOutRow;
    const int x = x0 + p0 + lane;
    const bool live = lane < npx;
    const uint2 lv = live ? cen_l[p0 + lane] : make_uint2(0, 0);
    const int rel = x - xs0;                                // cen_r index of x - d is rel - d
    const bool inside = x0 + p0 >= D - 1;                   // no lane has x - d < 0 for d < D
    uint8_t *dst = cost + ((size_t)y * w + x0 + p0) * Dp;
    for (int seg = 0; seg < Dp; seg += kCcSeg) {
        const int segw = min(kCcSeg, Dp - seg);
        if (live) {
            for (int dc = 0; dc < segw; dc += 16) {
                const int d0 = seg + dc;
                uint32_t wd[4];
                auto chunk = [&](auto checked) {
                    // Words rel - d0 - 15 .. rel - d0, read at constant offsets.
                    const uint2 *rb = cen_r + (decltype(checked)::value ? 0 : rel - d0 - 15);
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        uint32_t v = 0;
#pragma unroll
                        for (int b = 0; b < 4; ++b) {
                            const int i = 4 * q + b, d = d0 + i;
                            uint32_t c;
                            if constexpr (decltype(checked)::value) {
                                const uint2 r = cen_r[max(rel - d, 0)];
                                c = (d < D && d <= x) ? __popc(lv.x ^ r.x) + __popc(lv.y ^ r.y)
                                                      : 64u;
                            } else {
                                const uint2 r = rb[15 - i];
                                c = __popc(lv.x ^ r.x) + __popc(lv.y ^ r.y);
                            }
                            v |= (c + bias) << (8 * b);
                        }
                        wd[q] = v;
                    }
                };
                if (inside && d0 + 16 <= D) chunk(std::false_type{});
                else chunk(std::true_type{});
                *(uint4 *)(out + lane * kCcOutRow + dc) = make_uint4(wd[0], wd[1], wd[2], wd[3]);
            }
        }
        // The tile's rows come from other lanes of this wave.
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // Chunk i is chunk c of pixel px: i / nch by a multiply, exact for
        // nch <= 8 and i < 64 * nch (the error is below 1/128, the gap to the
        // next multiple at least 1/8); offsets from dst fit 32 bits.
        const int nch = segw / 16;
        const uint32_t magic = (65536u + (uint32_t)nch - 1u) / (uint32_t)nch;
        for (int i = lane; i < npx * nch; i += 64) {
            const int px = (int)(((uint32_t)i * magic) >> 16), c = i - px * nch;
            *(uint4 *)(dst + (uint32_t)(px * Dp + seg + c * 16)) =
                *(const uint4 *)(out + px * kCcOutRow + c * 16);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

__global__ void __launch_bounds__(256)
cost_kernel(const uint2 *__restrict__ lc, const uint2 *__restrict__ rc, int w, int h, int D,
            int Dp, uint32_t bias, uint8_t *__restrict__ cost) {
    extern __shared__ __align__(16) uint8_t lds[];
    cost_tile(lc, rc, w, h, D, Dp, bias, cost, (int)blockIdx.x, (int)blockIdx.y, lds);
}

// Cost volumes of many pairs in one launch: block b belongs to the entry
// whose block range holds it (tiles of kCcTile pixels, row-major).
__global__ void __launch_bounds__(256)
cost_multi_kernel(const CostJob *__restrict__ jobs, int njobs) {
    extern __shared__ __align__(16) uint8_t lds[];
    const int j = find_entry(jobs, njobs, blockIdx.x);
    const CostJob &c = jobs[j];
    const int local = (int)(blockIdx.x - c.first_block);
    cost_tile(c.lc, c.rc, c.w, c.h, c.D, c.Dp, c.bias, c.cost, local % c.nbx, local / c.nbx, lds);
}

__global__ void or_word_kernel(uint32_t *dst, const uint32_t *src) {
    if (threadIdx.x == 0) *dst |= *src;
}

__global__ void __launch_bounds__(256) zero_kernel(uint32_t *p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// Sum of NV per-direction volumes (u8 or u16 entries, MODE) and
// winner-take-all, 16 lanes per pixel.
template <int NV, int NC, int MODE>
__global__ void __launch_bounds__(256)
wta_sum_kernel(const uint8_t *__restrict__ vol0, size_t vstride,
               const uint8_t *__restrict__ hi0, const uint2 *__restrict__ flags, int npx,
               int D, int Dp, float *__restrict__ out, int16_t *__restrict__ best,
               uint16_t *__restrict__ sum_out) {
    const int pix = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 4);
    wta_sum_pixel(vol0, vstride, hi0, flags, NV, NC, MODE, pix, npx, D, Dp, out, best,
                  sum_out);
}

// Sum + winner-take-all of many pairs in one launch (16 pixels per block).
__global__ void __launch_bounds__(256)
wta_multi_kernel(const WtaJob *__restrict__ jobs, int njobs) {
    const int j = find_entry(jobs, njobs, blockIdx.x);
    const WtaJob &c = jobs[j];
    const int pix = (int)((blockIdx.x - c.first_block) * 16 + (threadIdx.x >> 4));
    wta_sum_pixel(c.vol0, c.vstride, c.hi0, c.flags, c.nv, c.Dp > 128 ? 2 : 1, c.mode, pix,
                  c.npx, c.D, c.Dp, c.out, c.best, c.sum_out);
}

__global__ void __launch_bounds__(256)
lr_kernel(const uint16_t *__restrict__ vol, int w, int h, int D, int Dp, int tol,
          const int16_t *__restrict__ best, float *__restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)w * h) return;
    lr_pixel(vol, w, D, Dp, tol, i, best, out);
}

__global__ void __launch_bounds__(256)
copy_kernel(uint4 *__restrict__ dst, const uint4 *__restrict__ src, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

}  // namespace

// The fastest schedule of a 1280x720 pair at D = 128, 8 paths, carried over
// to the other classes: the middle layout for the horizontal directions, the
// widest per-lane layout for the others, scanlines not split.
Schedule default_schedule(int D) {
    return layout_class(D) == 2 ? Schedule{0, 1, 0} : Schedule{1, 2, 0};
}

bool schedule_valid(Schedule sc, int D) {
    const int n = kClassCount[layout_class(D)];
    return sc.horiz >= 0 && sc.horiz < n && sc.other >= 0 && sc.other < n &&
           (sc.split >= 0 && sc.split <= 2);
}

hipError_t launch_zero(uint32_t *p, size_t n, hipStream_t s) {
    hipLaunchKernelGGL(zero_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

hipError_t launch_census(const uint8_t *img, int w, int h, int rx, int ry,
                         uint2 *out, uint2 *zero_flags, uint32_t *zero_words, size_t nwords,
                         hipStream_t s) {
    dim3 block(64, 4);
    dim3 grid((w + 63) / 64, (h + 3) / 4);
    hipLaunchKernelGGL(census_kernel, grid, block, 0, s, img, w, h, rx, ry, out, zero_flags,
                       zero_words, zero_words ? nwords : 0);
    return hipGetLastError();
}

hipError_t launch_census_multi(const CensusJob *d_jobs, int njobs, uint32_t blocks,
                               hipStream_t s) {
    if (njobs < 1 || blocks < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(census_multi_kernel, dim3(blocks), dim3(64, 4), 0, s, d_jobs, njobs);
    return hipGetLastError();
}

hipError_t launch_cost_multi(const CostJob *d_jobs, int njobs, uint32_t blocks, hipStream_t s) {
    if (njobs < 1 || blocks < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(cost_multi_kernel, dim3(blocks), dim3(256), kCcLds, s, d_jobs, njobs);
    return hipGetLastError();
}

hipError_t launch_wta_multi(const WtaJob *d_jobs, int njobs, uint32_t blocks, hipStream_t s) {
    if (njobs < 1 || blocks < 1) return hipErrorInvalidValue;
    hipLaunchKernelGGL(wta_multi_kernel, dim3(blocks), dim3(256), 0, s, d_jobs, njobs);
    return hipGetLastError();
}

hipError_t launch_or_word(uint32_t *dst, const uint32_t *src, hipStream_t s) {
    hipLaunchKernelGGL(or_word_kernel, dim3(1), dim3(64), 0, s, dst, src);
    return hipGetLastError();
}

hipError_t launch_cost(const uint2 *lc, const uint2 *rc, int w, int h, int D, uint32_t bias,
                       uint8_t *cost, hipStream_t s) {
    const dim3 grid((unsigned)((w + kCcTile - 1) / kCcTile), (unsigned)h);
    hipLaunchKernelGGL(cost_kernel, grid, dim3(256), kCcLds, s, lc, rc, w, h, D,
                       volume_stride(D), bias, cost);
    return hipGetLastError();
}

hipError_t launch_paths(const PassParams &base, const int *dirs, uint8_t *const *vols,
                        uint8_t *const *his, int n, Schedule sc, int mode, uint32_t *sync,
                        hipStream_t s) {
    if (n < 1 || n > kMaxPasses || base.D < 1 || base.D > 256 || !schedule_valid(sc, base.D) ||
        (mode != kOutU8 && mode != kOutU16))
        return hipErrorInvalidValue;
    const int C = layout_class(base.D);
    MultiPass m{};
    m.base = base;
    m.npass = n;
    int order[kMaxPasses];
    for (int j = 0; j < n; ++j) {
        if (dirs[j] < 0 || dirs[j] > 7) return hipErrorInvalidValue;
        order[j] = j;
    }
    std::stable_sort(order, order + n, [&](int a, int b) {
        return scan_length(dirs[a], base.w, base.h) > scan_length(dirs[b], base.w, base.h);
    });
    int blocks = 0;
    for (int j = 0; j < n; ++j) {
        const int k = order[j];
        PassDesc &d = m.pass[j];
        d.vol = vols[k];
        d.hi = his ? his[k] : nullptr;
        d.dir = dirs[k];
        d.dx = kDirX[d.dir];
        d.dy = kDirY[d.dir];
        d.nscan = num_scanlines(d.dir, base.w, base.h);
        d.layout = d.dy == 0 ? sc.horiz : sc.other;
        const int per_wave = 128 / kClassLayouts[C][d.layout][0];
        const int waves = (d.nscan + per_wave - 1) / per_wave;
        d.first_block = blocks;
        d.nblocks = (((waves + 3) / 4) + 7) & ~7;
        blocks += d.nblocks;
        d.nwaves = waves;
    }
    // Split the passes whose scanlines are long enough: at split level 1
    // those that set the critical path, longer than 2/3 of the longest (at
    // 1280x720 the horizontal ones, whose halves then run no longer than the
    // vertical scanlines); at level 2 every pass of at least kSplitMin.
    int longest = 0;
    for (int j = 0; j < n; ++j) longest = std::max(longest, scan_length(m.pass[j].dir, base.w, base.h));
    int waves_all = 0, nsplit = 0;
    int per_xcd[8] = {};   // tickets of each XCD's counter
    for (int j = 0; j < n; ++j) {
        PassDesc &d = m.pass[j];
        const int len = scan_length(d.dir, base.w, base.h);
        d.split = sc.split && len >= kSplitMin && (sc.split >= 2 || 3 * len > 2 * longest);
        d.first_wave = waves_all;
        waves_all += d.nwaves;
        nsplit += d.split;
        for (int k = 0; k < 8; ++k)
            per_xcd[k] += (d.nwaves * (k + 1) / 8 - d.nwaves * k / 8) * (d.split ? 2 : 1);
    }
    for (int j = n; j < kMaxPasses; ++j) m.pass[j] = m.pass[0];
    m.split = sync != nullptr && nsplit > 0 &&
              kSyncFlags + 2 * (size_t)waves_all <= sync_words(base.w, base.h);
    m.total_waves = waves_all;
    m.sync = sync;
    if (m.split) {
        // Every counter's tickets are taken: 8 equal shares of blocks.
        const int most = *std::max_element(per_xcd, per_xcd + 8);
        blocks = 8 * ((most + 3) / 4);
    }
    if (C == 0) return launch_paths_c0(mode, m.split != 0, m, blocks, s);
    if (C == 1) return launch_paths_c1(mode, m.split != 0, m, blocks, s);
    return launch_paths_c2(mode, m.split != 0, m, blocks, s);
}

hipError_t launch_wta_sum(const uint8_t *vol0, size_t vstride, const uint8_t *hi0,
                          const uint8_t *flags, int nvol, int mode, int npx, int D, int Dp,
                          float *out, int16_t *best, uint16_t *sum_out, hipStream_t s) {
    if (Dp > 256 || Dp % 16 || npx < 1) return hipErrorInvalidValue;
    const unsigned grid = (unsigned)(((size_t)npx * 16 + 255) / 256);
#define VXH_WTA_CASE(NV, NC, MM)                                                             \
    if (mode == MM && nvol == NV && (Dp > 128) == (NC == 2)) {                                \
        hipLaunchKernelGGL((wta_sum_kernel<NV, NC, MM>), dim3(grid), dim3(256), 0, s, vol0,   \
                           vstride, hi0, (const uint2 *)flags, npx, D, Dp, out, best,         \
                           sum_out);                                                          \
        return hipGetLastError();                                                             \
    }
    VXH_WTA_CASE(1, 1, kOutU16) VXH_WTA_CASE(1, 2, kOutU16)
    VXH_WTA_CASE(2, 1, kOutU16) VXH_WTA_CASE(2, 2, kOutU16)
    VXH_WTA_CASE(4, 1, kOutU16) VXH_WTA_CASE(4, 2, kOutU16)
    VXH_WTA_CASE(8, 1, kOutU16) VXH_WTA_CASE(8, 2, kOutU16)
    VXH_WTA_CASE(2, 1, kOutU8) VXH_WTA_CASE(2, 2, kOutU8)
    VXH_WTA_CASE(4, 1, kOutU8) VXH_WTA_CASE(4, 2, kOutU8)
    VXH_WTA_CASE(8, 1, kOutU8) VXH_WTA_CASE(8, 2, kOutU8)
#undef VXH_WTA_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_copy(void *dst, const void *src, size_t bytes, hipStream_t s) {
    const size_t n = bytes / 16;
    const unsigned grid = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(copy_kernel, dim3(grid ? grid : 1), dim3(256), 0, s, (uint4 *)dst,
                       (const uint4 *)src, n);
    return hipGetLastError();
}

hipError_t launch_lr_check(const uint16_t *vol, int w, int h, int D, int Dp, int tol,
                           const int16_t *best, float *out, hipStream_t s) {
    const size_t n = (size_t)w * h;
    hipLaunchKernelGGL(lr_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256),
                       0, s, vol, w, h, D, Dp, tol, best, out);
    return hipGetLastError();
}

}  // namespace vxh
