/*
 * Scalar CPU oracle for vx-vision's SGMStereo path (TEST INFRASTRUCTURE ONLY:
 * the checker the HIP path is compared against, never part of the product).
 *
 * Restates, function by function:
 *   census      vx-vision/shaders/SGMStereo.metal:25-66   (sgm_census_transform)
 *   cost        vx-vision/shaders/SGMStereo.metal:133-142 (inline Hamming cost)
 *   aggregate   vx-vision/shaders/SGMStereo.metal:72-173  (sgm_cost_aggregate)
 *   wta         vx-vision/shaders/SGMStereo.metal:179-214 (sgm_wta_disparity)
 *   pipeline    vx-vision/src/kernels/sgm.rs:108-265      (compute_disparity)
 *
 * Semantics kept exactly:
 *   - census: dy outer, dx inner, centre skipped, clamp-to-edge, bit b -> lo
 *     for b < 32 else hi; comparison neighbour < centre on the u8 values
 *     (R8Unorm reads k/255 are strictly monotone in k, so the float compare of
 *     the shader equals the integer compare);
 *   - cost 64 when x - d < 0, otherwise popcount(lo^lo') + popcount(hi^hi');
 *   - recurrence in place over ascending d on one array, so the d-1 term reads
 *     the value already updated in this step while the d and d+1 terms read
 *     the previous step; every sum is reduced mod 2^16 (ushort), and min is the
 *     unsigned ushort minimum;
 *   - the summed volume starts at zero and each path adds mod 2^16;
 *   - WTA keeps the first (lowest d) strict minimum starting from 0xFFFF, and
 *     refines with a 3-point parabola in IEEE fp32.
 *
 * Defined here because the reference leaves them undefined (it returns early
 * for diagonals, SGMStereo.metal:100-104, and has no left-right check):
 *   - path set: index 0..7 = (+1,0) (0,+1) (-1,0) (0,-1) (+1,+1) (-1,+1)
 *     (+1,-1) (-1,-1); num_paths 2 uses the reference's two (sgm.rs:197-200),
 *     4 adds the reverse cardinals the shader already handles
 *     (SGMStereo.metal:88-99), 8 adds the diagonals;
 *   - a path starts at every pixel whose predecessor p - r lies outside the
 *     image and runs while inside; its first pixel copies the cost, later
 *     pixels use the recurrence above unchanged;
 *   - left-right check: D_R(x') = lowest d minimising S(x'+d, y, d) over
 *     x'+d < w; a left pixel with integer winner d is kept iff x - d >= 0 and
 *     |d - D_R(x - d)| <= tol, otherwise it becomes -1.0f.
 */
#include "sgm_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <thread>
#include <vector>

namespace {

inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* Directions in path-index order. */
const int kDirX[8] = {+1, 0, -1, 0, +1, -1, +1, -1};
const int kDirY[8] = {0, +1, 0, -1, +1, +1, -1, -1};

/* Start pixels of direction (dx, dy): every pixel whose predecessor lies
 * outside the image.  Horizontal and vertical directions reduce to the
 * reference's start column/row (SGMStereo.metal:88-99). */
void start_pixels(uint32_t w, uint32_t h, int dx, int dy,
                  std::vector<std::pair<int, int>> &out) {
    out.clear();
    for (int y = 0; y < (int)h; ++y)
        for (int x = 0; x < (int)w; ++x) {
            int px = x - dx, py = y - dy;
            if (px < 0 || px >= (int)w || py < 0 || py >= (int)h)
                out.emplace_back(x, y);
        }
}

/* One scanline of SGMStereo.metal:106-172 starting at (sx, sy). */
void scanline(const uint32_t *lc, const uint32_t *rc, uint32_t w, uint32_t h,
              uint32_t nd, uint32_t P1, uint32_t P2, int dx, int dy, int sx,
              int sy, uint16_t *S) {
    uint16_t prev[256];
    for (uint32_t d = 0; d < nd && d < 256; ++d) prev[d] = 0;  /* :112-114 */
    uint16_t pm = 0;                                            /* :115 */
    for (int step = 0;; ++step) {
        int x = sx + step * dx, y = sy + step * dy;             /* :121-122 */
        if (x < 0 || x >= (int)w || y < 0 || y >= (int)h) break;  /* :124 */
        uint16_t cur = 0xFFFF;                                  /* :129 */
        for (uint32_t d = 0; d < nd && d < 256; ++d) {          /* :131 */
            uint16_t c = (uint16_t)oracle_cost(lc, rc, w, x, y, d);
            uint16_t agg;
            if (step == 0) {
                agg = c;                                        /* :152 */
            } else {
                uint16_t o0 = prev[d];                          /* :154 */
                uint16_t o1 = d > 0 ? (uint16_t)(prev[d - 1] + P1) : 0xFFFF;
                uint16_t o2 = d + 1 < nd ? (uint16_t)(prev[d + 1] + P1) : 0xFFFF;
                uint16_t o3 = (uint16_t)(pm + P2);              /* :157 */
                uint16_t best = std::min(std::min(o0, o1), std::min(o2, o3));
                agg = (uint16_t)(c + best - pm);                /* :160 */
            }
            size_t idx = ((size_t)y * w + (size_t)x) * nd + d;  /* :164 */
            S[idx] = (uint16_t)(S[idx] + agg);                  /* :165 */
            prev[d] = agg;                                      /* :167 */
            cur = std::min(cur, agg);                           /* :168 */
        }
        pm = cur;                                               /* :171 */
    }
}

void run_parallel(uint32_t nthreads, size_t n,
                  const std::function<void(size_t, size_t)> &fn) {
    if (nthreads <= 1 || n < 2) {
        fn(0, n);
        return;
    }
    uint32_t t = std::min<size_t>(nthreads, n);
    std::vector<std::thread> th;
    for (uint32_t i = 0; i < t; ++i) {
        size_t b = n * i / t, e = n * (i + 1) / t;
        th.emplace_back([=, &fn] { fn(b, e); });
    }
    for (auto &x : th) x.join();
}

void census_rows(const uint8_t *img, uint32_t w, uint32_t h, uint32_t rx,
                 uint32_t ry, uint32_t *out, size_t y0, size_t y1) {
    /* SGMStereo.metal:31-65 */
    int W = (int)w - 1, H = (int)h - 1;
    for (size_t cy = y0; cy < y1; ++cy)
        for (int cx = 0; cx <= W; ++cx) {
            uint8_t centre = img[cy * w + cx];
            uint32_t lo = 0, hi = 0;
            int bit = 0;
            for (int dy = -(int)ry; dy <= (int)ry; ++dy)
                for (int dx = -(int)rx; dx <= (int)rx; ++dx) {
                    if (dx == 0 && dy == 0) continue;
                    int sx = clampi(cx + dx, 0, W);
                    int sy = clampi((int)cy + dy, 0, H);
                    if (img[(size_t)sy * w + sx] < centre) {
                        if (bit < 32) lo |= 1u << bit;
                        else hi |= 1u << (bit - 32);
                    }
                    ++bit;
                }
            out[2 * (cy * w + cx)] = lo;
            out[2 * (cy * w + cx) + 1] = hi;
        }
}

void aggregate_dir_mt(const uint32_t *lc, const uint32_t *rc, uint32_t w,
                      uint32_t h, uint32_t nd, uint32_t P1, uint32_t P2, int dx,
                      int dy, uint16_t *S, uint32_t nthreads) {
    std::vector<std::pair<int, int>> starts;
    start_pixels(w, h, dx, dy, starts);
    /* Scanlines of one direction touch disjoint pixels, so they may run on
     * separate threads without changing any bit. */
    run_parallel(nthreads, starts.size(), [&](size_t b, size_t e) {
        for (size_t i = b; i < e; ++i)
            scanline(lc, rc, w, h, nd, P1, P2, dx, dy, starts[i].first,
                     starts[i].second, S);
    });
}

/* Validation of a user config, returning the product's status codes:
 * 0 ok, 5 size mismatch (not reachable here), 6 invalid config. */
int validate(uint32_t w, uint32_t h, uint32_t nd, float p1, float p2,
             uint32_t rx, uint32_t ry, uint32_t num_paths) {
    if (w == 0 || h == 0) return 6;
    if (nd < 1 || nd > 256 || nd > w) return 6;
    if ((2 * (uint64_t)rx + 1) * (2 * (uint64_t)ry + 1) - 1 > 64) return 6;
    if (!(p1 >= 0.0f && p1 < 65536.0f)) return 6;
    if (!(p2 >= 0.0f && p2 < 65536.0f)) return 6;
    if (num_paths != 2 && num_paths != 4 && num_paths != 8) return 6;
    return 0;
}

}  // namespace

extern "C" {

uint32_t oracle_cost(const uint32_t *lc, const uint32_t *rc, uint32_t w,
                     uint32_t x, uint32_t y, uint32_t d) {
    /* SGMStereo.metal:133-142 */
    int rx_coord = (int)x - (int)d;
    if (rx_coord < 0) return 64;
    size_t li = (size_t)y * w + x, ri = (size_t)y * w + (size_t)rx_coord;
    return (uint32_t)__builtin_popcount(lc[2 * li] ^ rc[2 * ri]) +
           (uint32_t)__builtin_popcount(lc[2 * li + 1] ^ rc[2 * ri + 1]);
}

void oracle_census(const uint8_t *img, uint32_t w, uint32_t h, uint32_t rx,
                   uint32_t ry, uint32_t *out) {
    census_rows(img, w, h, rx, ry, out, 0, h);
}

void oracle_aggregate_dir(const uint32_t *lc, const uint32_t *rc, uint32_t w,
                          uint32_t h, uint32_t nd, uint32_t p1, uint32_t p2,
                          int dx, int dy, uint16_t *S) {
    aggregate_dir_mt(lc, rc, w, h, nd, p1, p2, dx, dy, S, 1);
}

void oracle_wta_index(const uint16_t *S, uint32_t w, uint32_t h, uint32_t nd,
                      int32_t *out) {
    /* SGMStereo.metal:190-199 */
    for (size_t i = 0; i < (size_t)w * h; ++i) {
        const uint16_t *v = S + i * nd;
        uint16_t best = 0xFFFF;
        int bd = 0;
        for (uint32_t d = 0; d < nd; ++d)
            if (v[d] < best) { best = v[d]; bd = (int)d; }
        out[i] = bd;
    }
}

void oracle_wta(const uint16_t *S, uint32_t w, uint32_t h, uint32_t nd,
                float *out) {
    /* SGMStereo.metal:185-213 */
    for (size_t i = 0; i < (size_t)w * h; ++i) {
        const uint16_t *v = S + i * nd;
        uint16_t best = 0xFFFF;
        int bd = 0;
        for (uint32_t d = 0; d < nd; ++d)
            if (v[d] < best) { best = v[d]; bd = (int)d; }
        float disp = (float)bd;
        if (bd > 0 && bd < (int)nd - 1) {
            float c0 = (float)v[bd - 1];
            float c1 = (float)best;
            float c2 = (float)v[bd + 1];
            volatile float denom = 2.0f * (c0 + c2 - 2.0f * c1);
            if (std::fabs(denom) > 1e-6f) {
                volatile float q = (c0 - c2) / denom;
                disp = disp + q;
            }
        }
        out[i] = disp;
    }
}

void oracle_lr_check(const uint16_t *S, uint32_t w, uint32_t h, uint32_t nd,
                     uint32_t tol, float *disp) {
    std::vector<int32_t> dl((size_t)w * h);
    oracle_wta_index(S, w, h, nd, dl.data());
    std::vector<int32_t> dr(w);
    for (uint32_t y = 0; y < h; ++y) {
        for (uint32_t xr = 0; xr < w; ++xr) {
            uint16_t best = 0xFFFF;
            int bd = 0;
            for (uint32_t d = 0; d < nd && xr + d < w; ++d) {
                uint16_t v = S[((size_t)y * w + xr + d) * nd + d];
                if (v < best) { best = v; bd = (int)d; }
            }
            dr[xr] = bd;
        }
        for (uint32_t x = 0; x < w; ++x) {
            int d = dl[(size_t)y * w + x];
            int xr = (int)x - d;
            bool ok = xr >= 0 && std::abs(d - dr[xr]) <= (int)tol;
            if (!ok) disp[(size_t)y * w + x] = -1.0f;
        }
    }
}

int oracle_direction(uint32_t index, int *dx, int *dy) {
    if (index >= 8) return -1;
    *dx = kDirX[index];
    *dy = kDirY[index];
    return 0;
}

int oracle_sgm_volume(const uint8_t *left, const uint8_t *right, uint32_t w,
                      uint32_t h, uint32_t nd, float p1, float p2, uint32_t rx,
                      uint32_t ry, uint32_t num_paths, uint32_t nthreads,
                      uint16_t *S) {
    int st = validate(w, h, nd, p1, p2, rx, ry, num_paths);
    if (st) return st;
    size_t npx = (size_t)w * h;
    std::vector<uint32_t> lc(2 * npx), rc(2 * npx);
    run_parallel(nthreads, h, [&](size_t b, size_t e) {
        census_rows(left, w, h, rx, ry, lc.data(), b, e);
        census_rows(right, w, h, rx, ry, rc.data(), b, e);
    });
    std::memset(S, 0, npx * nd * sizeof(uint16_t));   /* sgm.rs:128-129 */
    /* Metal float -> ushort conversion truncates (SGMStereo.metal:155-157). */
    uint32_t P1 = (uint32_t)p1, P2 = (uint32_t)p2;
    for (uint32_t i = 0; i < num_paths; ++i)
        aggregate_dir_mt(lc.data(), rc.data(), w, h, nd, P1 & 0xFFFF,
                         P2 & 0xFFFF, kDirX[i], kDirY[i], S, nthreads);
    return 0;
}

int oracle_sgm(const uint8_t *left, const uint8_t *right, uint32_t w,
               uint32_t h, uint32_t nd, float p1, float p2, uint32_t rx,
               uint32_t ry, uint32_t num_paths, uint32_t lr_check,
               uint32_t lr_tol, uint32_t nthreads, float *out) {
    int st = validate(w, h, nd, p1, p2, rx, ry, num_paths);
    if (st) return st;
    std::vector<uint16_t> S((size_t)w * h * nd);
    oracle_sgm_volume(left, right, w, h, nd, p1, p2, rx, ry, num_paths,
                      nthreads, S.data());
    run_parallel(nthreads, h, [&](size_t b, size_t e) {
        oracle_wta(S.data() + b * w * nd, w, (uint32_t)(e - b), nd,
                   out + b * w);
    });
    if (lr_check) oracle_lr_check(S.data(), w, h, nd, lr_tol, out);
    return 0;
}

void oracle_scanline(const uint8_t *costs, uint32_t steps, uint32_t nd,
                     uint32_t p1, uint32_t p2, uint16_t *out) {
    /* The recurrence of SGMStereo.metal:106-172 on explicit costs. */
    uint16_t prev[256] = {0};
    uint16_t pm = 0;
    for (uint32_t s = 0; s < steps; ++s) {
        uint16_t cur = 0xFFFF;
        for (uint32_t d = 0; d < nd; ++d) {
            uint16_t c = costs[s * nd + d], agg;
            if (s == 0) {
                agg = c;
            } else {
                uint16_t o0 = prev[d];
                uint16_t o1 = d > 0 ? (uint16_t)(prev[d - 1] + p1) : 0xFFFF;
                uint16_t o2 = d + 1 < nd ? (uint16_t)(prev[d + 1] + p1) : 0xFFFF;
                uint16_t o3 = (uint16_t)(pm + p2);
                uint16_t best = std::min(std::min(o0, o1), std::min(o2, o3));
                agg = (uint16_t)(c + best - pm);
            }
            out[s * nd + d] = agg;
            prev[d] = agg;
            cur = std::min(cur, agg);
        }
        pm = cur;
    }
}

}  // extern "C"
