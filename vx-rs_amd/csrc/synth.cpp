// Seeded, integer-only synthetic stereo pairs; the algorithm is stated in
// include/vxh/synth.h.
#include "vxh/synth.h"

#include <algorithm>
#include <vector>

namespace {

struct SplitMix64 {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    uint32_t u(uint32_t n) { return (uint32_t)(next() % n); }
};

int64_t floor_div(int64_t a, int64_t b) {
    int64_t q = a / b;
    if ((a % b != 0) && ((a < 0) != (b < 0))) --q;
    return q;
}

}  // namespace

extern "C" int vxh_synth_pair(uint32_t w, uint32_t h, uint32_t nd,
                              uint32_t n_rects, uint32_t n_planes,
                              uint64_t seed, uint8_t *left, uint8_t *right,
                              int32_t *truth) {
    if (w == 0 || h == 0 || nd == 0) return 6;
    SplitMix64 rng{seed};
    const size_t n = (size_t)w * h;
    std::vector<uint8_t> raw(n);
    for (size_t i = 0; i < n; ++i) raw[i] = (uint8_t)(rng.next() & 0xFF);
    for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x) {
            uint32_t s = 0;
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    int sx = std::clamp((int)x + dx, 0, (int)w - 1);
                    int sy = std::clamp((int)y + dy, 0, (int)h - 1);
                    s += raw[(size_t)sy * w + sx];
                }
            left[(size_t)y * w + x] = (uint8_t)((s + 4) / 9);
        }

    std::vector<int32_t> disp(n, (int32_t)std::min<uint32_t>(8, nd - 1));
    for (uint32_t r = 0; r < n_rects; ++r) {
        uint32_t x0 = rng.u(w), y0 = rng.u(h);
        uint32_t rw = 1 + rng.u(std::max<uint32_t>(1, w / 4));
        uint32_t rh = 1 + rng.u(std::max<uint32_t>(1, h / 4));
        int32_t d = (int32_t)rng.u(nd);
        for (uint32_t y = y0; y < std::min(h, y0 + rh); ++y)
            for (uint32_t x = x0; x < std::min(w, x0 + rw); ++x)
                disp[(size_t)y * w + x] = d;
    }
    for (uint32_t p = 0; p < n_planes; ++p) {
        uint32_t x0 = rng.u(w), y0 = rng.u(h);
        uint32_t rw = 1 + rng.u(std::max<uint32_t>(1, w / 3));
        uint32_t rh = 1 + rng.u(std::max<uint32_t>(1, h / 3));
        int64_t a = rng.u(nd);
        int64_t bx = (int64_t)rng.u(65) - 32, by = (int64_t)rng.u(65) - 32;
        for (uint32_t y = y0; y < std::min(h, y0 + rh); ++y)
            for (uint32_t x = x0; x < std::min(w, x0 + rw); ++x) {
                int64_t num = 1024 * a + bx * (int64_t)(x - x0) +
                              by * (int64_t)(y - y0) + 512;
                int64_t d = std::clamp<int64_t>(floor_div(num, 1024), 0, nd - 1);
                disp[(size_t)y * w + x] = (int32_t)d;
            }
    }

    std::vector<int32_t> zbuf(n, -1);
    for (uint32_t y = 0; y < h; ++y)
        for (uint32_t x = 0; x < w; ++x) {
            int32_t d = disp[(size_t)y * w + x];
            int32_t xr = (int32_t)x - d;
            if (xr < 0) continue;
            size_t j = (size_t)y * w + (size_t)xr;
            if (d > zbuf[j]) {
                zbuf[j] = d;
                right[j] = left[(size_t)y * w + x];
            }
        }
    for (size_t i = 0; i < n; ++i)
        if (zbuf[i] < 0) right[i] = (uint8_t)(rng.next() & 0xFF);
    for (size_t i = 0; i < n; ++i) {
        int64_t s = 0;
        for (int k = 0; k < 12; ++k) s += (int64_t)(rng.next() & 0xFFFF);
        int64_t noise = floor_div(s - 393216 + 16384, 32768);
        right[i] = (uint8_t)std::clamp<int64_t>((int64_t)right[i] + noise, 0, 255);
    }
    if (truth)
        for (size_t i = 0; i < n; ++i) truth[i] = disp[i];
    return 0;
}
