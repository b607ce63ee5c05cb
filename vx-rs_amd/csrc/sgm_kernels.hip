// Per-pair SGM kernels for gfx950: thin grid wrappers around the device
// functions in sgm_device.h / sgm_paths.h, and their host-side launchers.
#include <algorithm>

#include "sgm_paths.h"

namespace vxh {
namespace {

using namespace dev;

// Census of 64 x kCensusTileH pixels per block (blockDim (64, 4)); each
// thread computes one column of 4 consecutive rows.  The default 9 x 7 window
// first reads the block's (kCensusTileH + 6) x 72 tile of clamped image rows
// into LDS, which reproduces clamp-to-edge exactly, so every pixel then takes
// the same path.  A thread reads each of its 10 window rows as three aligned
// dwords and aligns them to its column with two alignbytes and a shift: the 9
// bytes of a row then sit at constant byte positions, which the byte-select
// operands of the subtractions read directly.  Each bit is shifted in with
// one subtraction (n - c has bit 31 set iff n < c) and one alignbit, from the
// last neighbour down so that neighbour k ends at bit k.  Other windows read
// the image directly.
constexpr int kCtW = 64 + 8, kCtH = kCensusTileH + 6;
constexpr int kCtRows = kCensusTileH / 4;   // rows per thread

// Byte b of a 9-byte window row held as three dwords (b constant).
__device__ __forceinline__ uint32_t row_byte(const uint32_t (&r)[3], int b) {
    return (r[b >> 2] >> (8 * (b & 3))) & 0xFFu;
}

// Tile (bx, by) of the census of one image (block-uniform arguments).
__device__ __forceinline__ void census_tile(const uint8_t *__restrict__ img, int w, int h, int rx,
                                            int ry, uint2 *__restrict__ out,
                                            uint2 *__restrict__ zero_flags,
                                            uint32_t *__restrict__ zero_words, size_t nwords,
                                            int bx, int by) {
    const int x = bx * 64 + threadIdx.x;
    const int yt = by * kCensusTileH + threadIdx.y * kCtRows;   // first row of this thread
    uint2 words[kCtRows];
    if (rx == 4 && ry == 3) {   // block-uniform
        __shared__ __align__(16) uint8_t tile[kCtH * kCtW];
        const int x0 = bx * 64 - 4, y0 = by * kCensusTileH - 3;
        for (int t = threadIdx.y * 64 + threadIdx.x; t < kCtH * kCtW; t += 256) {
            const int r = t / kCtW, c = t - r * kCtW;
            tile[t] = img[(size_t)min(max(y0 + r, 0), h - 1) * w + min(max(x0 + c, 0), w - 1)];
        }
        __syncthreads();
        if (x >= w || yt >= h) return;
        // Window rows of this thread: tile rows threadIdx.y * kCtRows + 0 .. kCtRows + 5,
        // columns threadIdx.x .. threadIdx.x + 8.
        const uint32_t *src = (const uint32_t *)(tile + threadIdx.y * kCtRows * kCtW) +
                              (threadIdx.x >> 2);
        const uint32_t sh = threadIdx.x & 3;
        uint32_t rows[kCtRows + 6][3];
#pragma unroll
        for (int r = 0; r < kCtRows + 6; ++r) {
            const uint32_t w0 = src[r * (kCtW / 4)], w1 = src[r * (kCtW / 4) + 1],
                           w2 = src[r * (kCtW / 4) + 2];
            rows[r][0] = __builtin_amdgcn_alignbyte(w1, w0, sh);
            rows[r][1] = __builtin_amdgcn_alignbyte(w2, w1, sh);
            rows[r][2] = w2 >> (8 * sh);
        }
#pragma unroll
        for (int p = 0; p < kCtRows; ++p) {
            const uint32_t c = row_byte(rows[p + 3], 4);
            uint32_t lo = 0, hi = 0;
#pragma unroll
            for (int k = 61; k >= 0; --k) {
                const int j = k < 31 ? k : k + 1;   // window index, the centre (31) skipped
                const uint32_t d = row_byte(rows[p + j / 9], j % 9) - c;
                if (k >= 32) hi = __builtin_amdgcn_alignbit(hi, d, 31);
                else lo = __builtin_amdgcn_alignbit(lo, d, 31);
            }
            words[p] = make_uint2(lo, hi);
        }
    } else {
        if (x >= w || yt >= h) return;
#pragma unroll
        for (int p = 0; p < kCtRows; ++p)
            words[p] = yt + p < h ? census_at(img, w, h, rx, ry, x, yt + p) : make_uint2(0, 0);
    }
#pragma unroll
    for (int p = 0; p < kCtRows; ++p) {
        if (yt + p >= h) break;
        const size_t i = (size_t)(yt + p) * w + x;
        out[i] = words[p];
        if (zero_flags) zero_flags[i] = make_uint2(0, 0);
        for (size_t k = i; k < nwords; k += (size_t)w * h) zero_words[k] = 0;
    }
}

__global__ void __launch_bounds__(256)
census_kernel(const uint8_t *__restrict__ img, int w, int h, int rx, int ry,
              uint2 *__restrict__ out, uint2 *__restrict__ zero_flags,
              uint32_t *__restrict__ zero_words, size_t nwords) {
    census_tile(img, w, h, rx, ry, out, zero_flags, zero_words, nwords, (int)blockIdx.x,
                (int)blockIdx.y);
}

// Census of both images of a pair in one launch (blockIdx.z: 0 left, 1
// right); only the left one clears the flags and words.
__global__ void __launch_bounds__(256)
census_pair_kernel(const uint8_t *__restrict__ left, const uint8_t *__restrict__ right, int w,
                   int h, int rx, int ry, uint2 *__restrict__ out_l, uint2 *__restrict__ out_r,
                   uint2 *__restrict__ zero_flags, uint32_t *__restrict__ zero_words,
                   size_t nwords) {
    const bool r = blockIdx.z != 0;
    census_tile(r ? right : left, w, h, rx, ry, r ? out_r : out_l, r ? nullptr : zero_flags,
                r ? nullptr : zero_words, r ? 0 : nwords, (int)blockIdx.x, (int)blockIdx.y);
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
// kCcTile pixels of one row and stages in LDS the left census words of the
// tile and the right census words it reads (the tile and the 255 pixels
// before it).  Each wave then walks 64 consecutive pixels, NG pixels per
// step: a group of 64 / NG lanes covers one pixel's slice, lane s of the
// group the four entries 4s .. 4s + 3, stored as one dword, so one step
// stores NG whole slices contiguously.
//
// A lane's right words form one ascending stream: entry j of the group's
// pixel x is the word of x - 4s - j, and the group's next pixel is NG
// further.  The lane reads the stream in aligned 16-byte chunks of four words
// (one per dword array: the low and the high halves are staged apart), three
// chunks ahead of use, so the LDS reads 8 / 4 / 2 bytes per entry for slices
// of <= 64 / 128 / 256 entries without waiting on them.  The lanes of a
// group read adjacent chunks (no bank conflicts); each group has its own copy
// of the arrays, shifted so that its chunks are 16-byte aligned.
constexpr int kCcTile = 256;                      // pixels per block
constexpr int kCcSpanR = kCcTile + 256 + 4;       // right census words: 255 before, 4 after
constexpr int kCcArr = kCcSpanR + 4;              // dwords per array (room for the shift)
constexpr size_t kCcLds = (size_t)4 * 2 * kCcArr * 4 + kCcTile * 8;   // up to 4 copies + left

// popcount(x) + acc in one instruction (the compiler otherwise adds the bias
// and the two halves' counts separately).
__device__ __forceinline__ uint32_t bcnt_acc(uint32_t x, uint32_t acc) {
    uint32_t r;
    asm("v_bcnt_u32_b32 %0, %1, %2" : "=v"(r) : "v"(x), "v"(acc));
    return r;
}

__device__ __forceinline__ uint32_t elem(const uint4 &v, int i) {
    return i == 0 ? v.x : (i == 1 ? v.y : (i == 2 ? v.z : v.w));
}

// Packed costs of four entries: byte j is popcount(l ^ r_j) + bias.
__device__ __forceinline__ uint32_t cost_word(uint2 l, const uint32_t (&lo)[4],
                                              const uint32_t (&hi)[4], uint32_t bias) {
    uint32_t c[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) c[j] = bcnt_acc(l.y ^ hi[j], bcnt_acc(l.x ^ lo[j], bias));
    // Three byte permutes: (c0, c1) and (c2, c3) into halves, then the halves.
    const uint32_t v01 = __builtin_amdgcn_perm(c[1], c[0], 0x0c0c0400u);
    const uint32_t v23 = __builtin_amdgcn_perm(c[3], c[2], 0x0c0c0400u);
    return __builtin_amdgcn_perm(v23, v01, 0x05040100u);
}

// One wave's run of 64 pixels from xw.  `slo` / `shi` point at the lane's
// stream (16-byte aligned): element m is the word of pixel x - 4s - 3 + m.
// cen_l[k] is the left word of pixel x0 + k.  FAST: every pixel of the run
// is inside the row, every lane stores, and no entry reads a pixel left of
// the row or lies past D; otherwise those entries get 64 + bias and the
// stores are predicated.
template <int NG, bool FAST>
__device__ __forceinline__ void cost_run(const uint4 *slo, const uint4 *shi, const uint2 *cen_l,
                                         int x0, int xw, int w, int D, int Dp, uint32_t bias,
                                         uint8_t *__restrict__ row_out) {
    constexpr int G = 64 / NG;                    // lanes per pixel
    constexpr int NC = (NG * (64 / NG - 1) + 3) / 4 + 1;   // chunks of the stream used
    const int lane = (int)(threadIdx.x & 63);
    const int g = lane / G, d0 = 4 * (lane % G);
    const int x = xw + g;
    const uint32_t fill = (64u + bias) * 0x01010101u;
    // Entries d >= D of this lane (a slice's padding up to Dp).
    const uint32_t mask_d = d0 >= D ? ~0u : (uint32_t)(~0ull << (8 * min(D - d0, 4)));
    const bool stores = d0 < Dp;
    // A fast run's slices are exactly the group's 4G entries: the stride is
    // then a constant and the stores take immediate offsets.
    const int Ds = FAST ? 4 * G : Dp;
    uint32_t *dst = (uint32_t *)(row_out + (size_t)x * Ds + d0);
    const int step = NG * Ds / 4;                 // dwords from one pixel of the group to its next
    const uint2 *lp = cen_l + (x - x0);
    // Four slots: chunk q in slot q % 4, fetched three chunks ahead.  The
    // steps run as 4 rounds of 4 chunks each, so slots are compile-time.
    constexpr int SR = 16 / NG;                   // steps per round
    uint4 bl[4], bh[4];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        bl[q] = slo[q];
        bh[q] = shi[q];
    }
#pragma unroll 1
    for (int r = 0; r < 4; ++r) {
#pragma unroll
        for (int ii = 0; ii < SR; ++ii) {
            const int m0 = NG * ii, qq = m0 / 4;          // chunk 4r + qq
            if (m0 % 4 < NG && 4 * r + qq + 3 < NC) {     // first step in the chunk
                bl[(qq + 3) % 4] = slo[4 * r + qq + 3];
                bh[(qq + 3) % 4] = shi[4 * r + qq + 3];
            }
            uint32_t lo[4], hi[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {                 // entry j: element m0 + 3 - j
                const int m = m0 + 3 - j;
                lo[j] = elem(bl[(m / 4) % 4], m % 4);
                hi[j] = elem(bh[(m / 4) % 4], m % 4);
            }
            const int i = SR * r + ii;
            uint32_t v = cost_word(lp[NG * i], lo, hi, bias);
            if (FAST) {
                dst[i * step] = v;
            } else {
                // Entries with x - d < 0 (j > x - d0) and d >= D cost 64.
                const int t = min(max(x + NG * i - d0 + 1, 0), 4);
                const uint32_t mk = (uint32_t)(~0ull << (8 * t)) | mask_d;
                v = (mk & fill) | (~mk & v);
                if (stores && x + NG * i < w) dst[i * step] = v;
            }
        }
    }
}

// Tile bx of row y of a pair's cost volume (block-uniform arguments; `lds`
// holds kCcLds bytes).
__device__ __forceinline__ void cost_tile(const uint2 *__restrict__ lc,
                                          const uint2 *__restrict__ rc, int w, int h, int D,
                                          int Dp, uint32_t bias, uint8_t *__restrict__ cost,
                                          int bx, int y, uint8_t *lds) {
    const int ng = Dp <= 64 ? 4 : (Dp <= 128 ? 2 : 1);   // groups (pixels) per wave step
    uint32_t *arr = (uint32_t *)lds;                      // copy c: low at arr + c * 2 * kCcArr
    uint2 *cen_l = (uint2 *)(arr + 4 * 2 * kCcArr);       // + shift(c), high kCcArr further
    const int t = threadIdx.x;
    const int x0 = bx * kCcTile;
    const int xr0 = x0 - 255;                             // pixel of stream index 0
    const size_t row = (size_t)y * w;
    // Copy c holds index k at dword shift(c) + k, shift(c) = (4 - c) % 4: group
    // c's stream starts at an index = c (mod 4), so its chunks are aligned.
    // Words outside the row are clamped: every entry that reads one is
    // replaced by 64 + bias (x - d < 0) or lies past the row's end.
    for (int i = t; i < kCcSpanR; i += 256) {
        const uint2 v = rc[row + min(max(xr0 + i, 0), w - 1)];
        for (int c = 0; c < ng; ++c) {
            uint32_t *lo = arr + c * 2 * kCcArr + ((4 - c) & 3);
            lo[i] = v.x;
            lo[i + kCcArr] = v.y;
        }
    }
    cen_l[t] = lc[row + min(x0 + t, w - 1)];
    __syncthreads();
    const int xw = x0 + (t >> 6) * 64;                    // first pixel of this wave
    if (xw >= w) return;                                  // wave-uniform
    uint8_t *row_out = cost + (size_t)y * w * Dp;
    const int G = 64 / ng, lane = t & 63;
    const int g = lane / G, d0 = 4 * (lane % G);
    // The lane's stream starts at the word of pixel xw + g - d0 - 3.
    const uint32_t *lo = arr + g * 2 * kCcArr + ((4 - g) & 3) + (xw + g - d0 - 3 - xr0);
    const uint4 *slo = (const uint4 *)lo, *shi = (const uint4 *)(lo + kCcArr);
    // Fast runs: inside the row, no entry left of it (x - d >= 0 for every
    // d < Dp), no padding entries, and every lane of a group stores.
    const bool fast = xw >= Dp - 1 && xw + 64 <= w && D == Dp && Dp == 4 * G;
    if (ng == 4) {
        if (fast) cost_run<4, true>(slo, shi, cen_l, x0, xw, w, D, Dp, bias, row_out);
        else cost_run<4, false>(slo, shi, cen_l, x0, xw, w, D, Dp, bias, row_out);
    } else if (ng == 2) {
        if (fast) cost_run<2, true>(slo, shi, cen_l, x0, xw, w, D, Dp, bias, row_out);
        else cost_run<2, false>(slo, shi, cen_l, x0, xw, w, D, Dp, bias, row_out);
    } else {
        if (fast) cost_run<1, true>(slo, shi, cen_l, x0, xw, w, D, Dp, bias, row_out);
        else cost_run<1, false>(slo, shi, cen_l, x0, xw, w, D, Dp, bias, row_out);
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
// winner-take-all, 16 lanes per pixel.  With two chunks (D > 128) each group
// takes two pixels, 16 apart, with both pixels' loads issued before either
// is summed (1920x1080, D = 256: 697 against 726 us); with one chunk that
// is slower (1280x720, D = 128: 162 against 156 us).
template <int NV, int NC, int MODE>
__global__ void __launch_bounds__(256)
wta_sum_kernel(const uint8_t *__restrict__ vol0, size_t vstride,
               const uint8_t *__restrict__ hi0, const uint2 *__restrict__ flags, int npx,
               int D, int Dp, float *__restrict__ out, int16_t *__restrict__ best,
               uint16_t *__restrict__ sum_out) {
    if constexpr (NC == 2) {
        const int p0 = (int)(blockIdx.x * 32 + (threadIdx.x >> 4));
        WtaIn<MODE> a, b;
        wta_load(a, vol0, vstride, flags, NV, NC, p0, npx, Dp);
        wta_load(b, vol0, vstride, flags, NV, NC, p0 + 16, npx, Dp);
        wta_finish(a, hi0, vstride, NV, NC, p0, npx, D, Dp, out, best, sum_out);
        wta_finish(b, hi0, vstride, NV, NC, p0 + 16, npx, D, Dp, out, best, sum_out);
    } else {
        const int pix = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 4);
        wta_sum_pixel(vol0, vstride, hi0, flags, NV, NC, MODE, pix, npx, D, Dp, out, best,
                      sum_out);
    }
}

// Sum + winner with the right view's keys for the left-right check: block
// (bx, y) takes kWtaLrP pixels of row y from bx * kWtaLrP (wta_lr_tile).
// 16 pixels and 256 threads per block: at 32 or 64 (512 or 1024 threads) the
// fewer global atomics bought nothing and 720p, D = 128 ran up to 20 us
// slower.
constexpr int kWtaLrP = 16;
template <int NV, int NC, int MODE>
__global__ void __launch_bounds__(16 * kWtaLrP)
wta_lr_kernel(const uint8_t *__restrict__ vol0, size_t vstride, const uint8_t *__restrict__ hi0,
              const uint2 *__restrict__ flags, int w, int h, int D, int Dp,
              float *__restrict__ out, int16_t *__restrict__ best, uint32_t *__restrict__ rkey) {
    __shared__ uint32_t lkey[kWtaLrP + 255];
    wta_lr_tile<MODE, kWtaLrP>(vol0, vstride, hi0, flags, NV, NC, w, h, (int)blockIdx.y,
                               (int)blockIdx.x * kWtaLrP, D, Dp, out, best, rkey, lkey);
}

// Sum + winner-take-all of many pairs in one launch: 16 pixels per block, of
// one row for a pair with right-view keys (the LR check, as wta_lr_kernel),
// and for the other pairs with two chunks (D > 128; TWO) 32 pixels per block,
// two per group as wta_sum_kernel takes them.  Instantiated per kind and, for
// the plain kinds, per volume format (a launch's pairs share one: the batch
// groups them by format), so that each keeps its own registers; each
// instantiation skips the blocks of the other kinds' pairs (wta_job_blocks
// gives every pair's block count).
template <bool LR, bool TWO, int MODE>
__global__ void __launch_bounds__(256)
wta_multi_kernel(const WtaJob *__restrict__ jobs, int njobs) {
    const int j = find_entry(jobs, njobs, blockIdx.x);
    const WtaJob &c = jobs[j];
    if ((c.rkey != nullptr) != LR) return;
    const int local = (int)(blockIdx.x - c.first_block);
    const int nc = c.Dp > 128 ? 2 : 1;
    if constexpr (LR) {
        __shared__ uint32_t lkey[kWtaLrP + 255];
        const int nbx = (c.w + kWtaLrP - 1) / kWtaLrP;
        const int h = c.npx / c.w;
        if (c.mode == kOutU16)
            wta_lr_tile<kOutU16, kWtaLrP>(c.vol0, c.vstride, c.hi0, c.flags, c.nv, nc, c.w, h,
                                          local / nbx, (local % nbx) * kWtaLrP, c.D, c.Dp, c.out,
                                          c.best, c.rkey, lkey);
        else
            wta_lr_tile<kOutU8, kWtaLrP>(c.vol0, c.vstride, c.hi0, c.flags, c.nv, nc, c.w, h,
                                         local / nbx, (local % nbx) * kWtaLrP, c.D, c.Dp, c.out,
                                         c.best, c.rkey, lkey);
    } else {
        // Not this instantiation's pair: another launch of launch_wta_multi takes it.
        if (nc != (TWO ? 2 : 1) || c.mode != MODE) return;
        const int p0 = local * (TWO ? 32 : 16) + (int)(threadIdx.x >> 4);
        WtaIn<MODE> a, b;
        wta_load(a, c.vol0, c.vstride, c.flags, c.nv, nc, p0, c.npx, c.Dp);
        if constexpr (TWO) wta_load(b, c.vol0, c.vstride, c.flags, c.nv, nc, p0 + 16, c.npx, c.Dp);
        wta_finish(a, c.hi0, c.vstride, c.nv, nc, p0, c.npx, c.D, c.Dp, c.out, c.best, nullptr);
        if constexpr (TWO)
            wta_finish(b, c.hi0, c.vstride, c.nv, nc, p0 + 16, c.npx, c.D, c.Dp, c.out, c.best,
                       nullptr);
    }
}

// Left-right check in two passes.  lr_right_kernel finds D_R(xr) for every
// right pixel xr of a row: the lowest k < D with xr + k < w minimising
// S(xr + k, y, k), scanned upwards from 0xFFFF with d = 0 kept when nothing
// is smaller (SGMStereo has no right view; this is the check as the build
// defines it).  A block owns kLrT right pixels of one row and walks d in
// chunks of kLrC: it stages the chunk's entries of the kLrT + kLrC - 1
// pixels the anti-diagonals cross (each a contiguous run of the pixel's
// slice) in LDS, rows padded to kLrRow entries so that the lanes' reads of
// one anti-diagonal fall in different banks, and each lane scans its own.
// lr_apply_kernel then keeps pixel x iff x - best >= 0 and
// |best - D_R(x - best)| <= tol, else writes -1.0f.  Reading the volume once
// per right pixel along anti-diagonals straight from memory would touch a
// different cache line per entry.
constexpr int kLrT = 256, kLrC = 64, kLrRow = kLrC + 2;

__global__ void __launch_bounds__(256)
lr_right_kernel(const uint16_t *__restrict__ vol, int w, int D, int Dp,
                int16_t *__restrict__ dr) {
    __shared__ __align__(16) uint16_t tile[(kLrT + kLrC - 1) * kLrRow];
    const int t = (int)threadIdx.x, y = (int)blockIdx.y, x0 = (int)blockIdx.x * kLrT;
    const int xr = x0 + t;
    const uint16_t *__restrict__ row = vol + (size_t)y * w * Dp;
    uint32_t bc = 0xFFFFu;
    int bd = 0;
    for (int c0 = 0; c0 < D; c0 += kLrC) {
        const int ne = min(kLrC, D - c0);                 // entries of this chunk
        const int nw = (ne + 1) >> 1;                      // dwords (c0 + 2 nw <= Dp)
        const int p0 = x0 + c0;                            // pixel of tile row 0
        const int np = min(kLrT + kLrC - 1, w - p0);       // pixels inside the row
        __syncthreads();                                   // the last chunk's reads are done
        if (ne == kLrC) {
            // Whole chunks: 16-byte loads, 8 per pixel (aligned: Dp and c0 are
            // multiples of 16 entries), written as four dwords.
            for (int i = t; i < np * (kLrC / 8); i += 256) {
                const int p = i >> 3, q = i & 7;
                const uint4 v = *(const uint4 *)(row + (size_t)(p0 + p) * Dp + c0 + 8 * q);
                uint32_t *dst = (uint32_t *)(tile + p * kLrRow + 8 * q);
                dst[0] = v.x;
                dst[1] = v.y;
                dst[2] = v.z;
                dst[3] = v.w;
            }
        } else {
            for (int i = t; i < np * nw; i += 256) {
                const int p = i / nw, q = i - p * nw;
                const uint32_t v = *(const uint32_t *)(row + (size_t)(p0 + p) * Dp + c0 + 2 * q);
                *(uint32_t *)(tile + p * kLrRow + 2 * q) = v;
            }
        }
        __syncthreads();
        const int kmax = min(ne, w - xr - c0);             // k = c0 + dd valid for dd < kmax
        const uint16_t *diag = tile + t * kLrRow;          // entry dd of pixel xr + c0 + dd
        if (kmax == kLrC) {
#pragma unroll 16
            for (int dd = 0; dd < kLrC; ++dd) {
                const uint32_t c = diag[dd * (kLrRow + 1)];
                if (c < bc) {
                    bc = c;
                    bd = c0 + dd;
                }
            }
        } else {
            for (int dd = 0; dd < kmax; ++dd) {
                const uint32_t c = diag[dd * (kLrRow + 1)];
                if (c < bc) {
                    bc = c;
                    bd = c0 + dd;
                }
            }
        }
    }
    if (xr < w) dr[(size_t)y * w + xr] = (int16_t)bd;
}

// The check from the right view's keys (lr_right_of_key): keeps pixel x iff
// x - best >= 0 and |best - D_R(x - best)| <= tol, else writes -1.0f.
__global__ void __launch_bounds__(256)
lr_apply_keys_kernel(const int16_t *__restrict__ best, const uint32_t *__restrict__ rkey, int w,
                     int h, int tol, float *__restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)w * h) return;
    const int x = (int)(i % w);
    const int d = best[i];
    const int xr = x - d;
    const bool ok = xr >= 0 && abs(d - (int)lr_right_of_key(rkey[i - d])) <= tol;
    if (!ok) out[i] = -1.0f;
}

__global__ void __launch_bounds__(256)
lr_apply_kernel(const int16_t *__restrict__ best, const int16_t *__restrict__ dr, int w, int h,
                int tol, float *__restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)w * h) return;
    const int x = (int)(i % w);
    const int d = best[i];
    const int xr = x - d;
    const bool ok = xr >= 0 && abs(d - (int)dr[i - d]) <= tol;
    if (!ok) out[i] = -1.0f;
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

// A lone pair's schedule by path count, from per-pair timings of every
// layout pair and split level at 320x240 to 1920x1080 for each class:
//  - 2 paths: one horizontal and one vertical pass leave most of the chip
//    idle, and halving the longest chain of steps wins everywhere: the
//    middle layout, split (10-30 % faster).
//  - 4 paths: split again; the widest per-lane layout for D in (64, 128]
//    (it leads at 1920x1080 and stays within 10 % elsewhere); D > 128
//    keeps the 8-path schedule, which no split beat.
//  - 8 paths: the 8-path schedule, except that D <= 64 uses the middle
//    layout everywhere and splits when level 1 would split only the
//    longest passes (320x240 and 1280x720: 4-10 % faster); where it would
//    split the diagonals and the shorter axis too (480x360, 640x480) the
//    doubled wave count costs up to 40 %.
Schedule plain_schedule(int D, int paths, uint32_t w, uint32_t h) {
    const int c = layout_class(D);
    if (paths == 2) return Schedule{1, 1, 1};
    if (paths == 4) return c == 2 ? Schedule{0, 1, 0} : Schedule{c == 1 ? 2 : 1, c == 1 ? 2 : 1, 1};
    if (c == 0) {
        const uint32_t lo = std::min(w, h), hi = std::max(w, h);
        const bool longest_only = lo < (uint32_t)kSplitMin || 3 * lo <= 2 * hi;
        return Schedule{1, 1, longest_only ? 1 : 0};
    }
    return default_schedule(D);
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
    dim3 grid((w + 63) / 64, (h + kCensusTileH - 1) / kCensusTileH);
    hipLaunchKernelGGL(census_kernel, grid, block, 0, s, img, w, h, rx, ry, out, zero_flags,
                       zero_words, zero_words ? nwords : 0);
    return hipGetLastError();
}

hipError_t launch_census_pair(const uint8_t *left, const uint8_t *right, int w, int h, int rx,
                              int ry, uint2 *out_l, uint2 *out_r, uint2 *zero_flags,
                              uint32_t *zero_words, size_t nwords, hipStream_t s) {
    dim3 block(64, 4);
    dim3 grid((w + 63) / 64, (h + kCensusTileH - 1) / kCensusTileH, 2);
    hipLaunchKernelGGL(census_pair_kernel, grid, block, 0, s, left, right, w, h, rx, ry, out_l,
                       out_r, zero_flags, zero_words, zero_words ? nwords : 0);
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

hipError_t launch_wta_multi(const WtaJob *d_jobs, int njobs, uint32_t blocks, int kinds,
                            int mode, hipStream_t s) {
    if (njobs < 1 || blocks < 1 || kinds < 1 || kinds > 7 ||
        (mode != kOutU8 && mode != kOutU16))
        return hipErrorInvalidValue;
    const dim3 g(blocks), b(256);
    if ((kinds & kWtaOne) && mode == kOutU8)
        hipLaunchKernelGGL((wta_multi_kernel<false, false, kOutU8>), g, b, 0, s, d_jobs, njobs);
    if ((kinds & kWtaOne) && mode == kOutU16)
        hipLaunchKernelGGL((wta_multi_kernel<false, false, kOutU16>), g, b, 0, s, d_jobs, njobs);
    if ((kinds & kWtaTwo) && mode == kOutU8)
        hipLaunchKernelGGL((wta_multi_kernel<false, true, kOutU8>), g, b, 0, s, d_jobs, njobs);
    if ((kinds & kWtaTwo) && mode == kOutU16)
        hipLaunchKernelGGL((wta_multi_kernel<false, true, kOutU16>), g, b, 0, s, d_jobs, njobs);
    if (kinds & kWtaLr)
        hipLaunchKernelGGL((wta_multi_kernel<true, false, kOutU8>), g, b, 0, s, d_jobs, njobs);
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
    // 16 pixels per block, 32 with two chunks (wta_sum_kernel).
    const unsigned grid1 = (unsigned)(((size_t)npx + 15) / 16);
    const unsigned grid2 = (unsigned)(((size_t)npx + 31) / 32);
#define VXH_WTA_CASE(NV, NC, MM)                                                             \
    if (mode == MM && nvol == NV && (Dp > 128) == (NC == 2)) {                                \
        hipLaunchKernelGGL((wta_sum_kernel<NV, NC, MM>), dim3(NC == 2 ? grid2 : grid1),       \
                           dim3(256), 0, s, vol0,                                             \
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

uint32_t wta_job_blocks(const WtaJob &j) {
    if (j.rkey) return (uint32_t)(((j.w + kWtaLrP - 1) / kWtaLrP) * (j.npx / j.w));
    if (j.Dp > 128) return (uint32_t)((j.npx + 31) / 32);
    return (uint32_t)((j.npx + 15) / 16);
}

int wta_job_kind(const WtaJob &j) {
    return j.rkey ? kWtaLr : (j.Dp > 128 ? kWtaTwo : kWtaOne);
}

hipError_t launch_wta_lr(const uint8_t *vol0, size_t vstride, const uint8_t *hi0,
                         const uint8_t *flags, int nvol, int mode, int w, int h, int D, int Dp,
                         float *out, int16_t *best, uint32_t *rkey, hipStream_t s) {
    if (Dp > 256 || Dp % 16 || w < 1 || h < 1 || D < 1 || D > Dp || !best || !rkey)
        return hipErrorInvalidValue;
    const dim3 grid((unsigned)((w + kWtaLrP - 1) / kWtaLrP), (unsigned)h);
#define VXH_WTA_LR_CASE(NV, NC, MM)                                                          \
    if (mode == MM && nvol == NV && (Dp > 128) == (NC == 2)) {                                \
        hipLaunchKernelGGL((wta_lr_kernel<NV, NC, MM>), grid, dim3(16 * kWtaLrP), 0, s,      \
                           vol0, vstride, hi0, (const uint2 *)flags, w, h, D, Dp, out, best,  \
                           rkey);                                                             \
        return hipGetLastError();                                                             \
    }
    VXH_WTA_LR_CASE(1, 1, kOutU16) VXH_WTA_LR_CASE(1, 2, kOutU16)
    VXH_WTA_LR_CASE(2, 1, kOutU16) VXH_WTA_LR_CASE(2, 2, kOutU16)
    VXH_WTA_LR_CASE(4, 1, kOutU16) VXH_WTA_LR_CASE(4, 2, kOutU16)
    VXH_WTA_LR_CASE(8, 1, kOutU16) VXH_WTA_LR_CASE(8, 2, kOutU16)
    VXH_WTA_LR_CASE(2, 1, kOutU8) VXH_WTA_LR_CASE(2, 2, kOutU8)
    VXH_WTA_LR_CASE(4, 1, kOutU8) VXH_WTA_LR_CASE(4, 2, kOutU8)
    VXH_WTA_LR_CASE(8, 1, kOutU8) VXH_WTA_LR_CASE(8, 2, kOutU8)
#undef VXH_WTA_LR_CASE
    return hipErrorInvalidValue;
}

hipError_t launch_lr_apply(const int16_t *best, const uint32_t *rkey, int w, int h, int tol,
                           float *out, hipStream_t s) {
    if (w < 1 || h < 1) return hipErrorInvalidValue;
    const size_t n = (size_t)w * h;
    hipLaunchKernelGGL(lr_apply_keys_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                       best, rkey, w, h, tol, out);
    return hipGetLastError();
}

hipError_t launch_copy(void *dst, const void *src, size_t bytes, hipStream_t s) {
    const size_t n = bytes / 16;
    const unsigned grid = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(copy_kernel, dim3(grid ? grid : 1), dim3(256), 0, s, (uint4 *)dst,
                       (const uint4 *)src, n);
    return hipGetLastError();
}

hipError_t launch_lr_check(const uint16_t *vol, int w, int h, int D, int Dp, int tol,
                           const int16_t *best, float *out, int16_t *dr, hipStream_t s) {
    if (w < 1 || h < 1 || D < 1 || D > Dp) return hipErrorInvalidValue;
    hipLaunchKernelGGL(lr_right_kernel, dim3((unsigned)((w + kLrT - 1) / kLrT), (unsigned)h),
                       dim3(256), 0, s, vol, w, D, Dp, dr);
    const size_t n = (size_t)w * h;
    hipLaunchKernelGGL(lr_apply_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, best,
                       dr, w, h, tol, out);
    return hipGetLastError();
}

}  // namespace vxh
