/*
 * Scalar CPU oracle for vx-vision's sparse stereo matcher (TEST
 * INFRASTRUCTURE ONLY: the checker the HIP kernels are compared against,
 * never part of the product).
 *
 * Restates vx-vision/shaders/StereoMatch.metal:
 *   matrix  stereo_hamming (:32-62): 0xFFFF unless |ly - ry| <= max_epipolar
 *           and min_disparity <= lx - rx <= max_disparity, else the popcount
 *           of the xor of the eight descriptor words;
 *   match   stereo_extract (:68-115): scan right indices in order keeping the
 *           first strictly smaller distance; reject when best >
 *           ushort(max_hamming); disparity lx - rx, Z = fx * baseline /
 *           disparity, X = Z * (lx - cx) / fx, Y = Z * (ly - cy) / fy in fp32
 *           (IEEE division), appended in ascending left index.
 * Keypoints are 4 floats/words each (x, y, response, level), descriptors 8
 * words each; params are max_epipolar, min_disparity, max_disparity,
 * ratio_thresh (unused), fx, fy, cx, cy, baseline, and max_hamming apart.
 *
 * Parity pins (tests/test_oracle_stereo_match.py): Python's int.bit_count on
 * every pair, the band/range masks evaluated independently, identical (0) and
 * complementary (256) descriptors, the first-of-ties rule, the 16-bit
 * truncation of max_hamming, and triangulation against float64.
 */
#include <cmath>
#include <cstdint>
#include <cstring>

#include "sgm_oracle.h"

namespace {

struct Kp {
    float x, y, response;
    uint32_t level;
};

uint16_t distance(const Kp &l, const Kp &r, const uint32_t *ld, const uint32_t *rd,
                  const float *p) {
    const float dy = std::fabs(l.y - r.y);
    if (dy <= p[0]) {
        const float d = l.x - r.x;
        if (d >= p[1] && d <= p[2]) {
            uint32_t dist = 0;
            for (int w = 0; w < 8; ++w) dist += (uint32_t)__builtin_popcount(ld[w] ^ rd[w]);
            return (uint16_t)dist;
        }
    }
    return 0xFFFF;
}

}  // namespace

extern "C" {

void oracle_stereo_matrix(const void *lk, const void *rk, const uint32_t *ld,
                          const uint32_t *rd, uint32_t nl, uint32_t nr, const float *p,
                          uint16_t *out) {
    const Kp *L = (const Kp *)lk, *R = (const Kp *)rk;
    for (uint32_t i = 0; i < nl; ++i)
        for (uint32_t j = 0; j < nr; ++j)
            out[(size_t)i * nr + j] = distance(L[i], R[j], ld + (size_t)i * 8,
                                               rd + (size_t)j * 8, p);
}

/* out: 7 words per match (left_idx, right_idx, disparity, X, Y, Z, pad 0).
 * Returns the number of matches. */
uint32_t oracle_stereo_match(const void *lk, const void *rk, const uint32_t *ld,
                             const uint32_t *rd, uint32_t nl, uint32_t nr, const float *p,
                             uint32_t max_hamming, uint32_t *out) {
    const Kp *L = (const Kp *)lk, *R = (const Kp *)rk;
    const float fx = p[4], fy = p[5], cx = p[6], cy = p[7], baseline = p[8];
    uint32_t n = 0;
    if (nl == 0 || nr == 0) return 0;
    for (uint32_t i = 0; i < nl; ++i) {
        uint16_t best = 0xFFFF, second = 0xFFFF;
        uint32_t best_ri = 0;
        for (uint32_t j = 0; j < nr; ++j) {
            const uint16_t d = distance(L[i], R[j], ld + (size_t)i * 8, rd + (size_t)j * 8, p);
            if (d < best) {
                second = best;
                best = d;
                best_ri = j;
            } else if (d < second) {
                second = d;
            }
        }
        (void)second;   // the ratio test that would use it is disabled (:98-101)
        if (best > (uint16_t)max_hamming) continue;
        const float disp = L[i].x - R[best_ri].x;
        const float Z = fx * baseline / disp;
        const float X = Z * (L[i].x - cx) / fx;
        const float Y = Z * (L[i].y - cy) / fy;
        uint32_t *o = out + (size_t)n * 7;
        o[0] = i;
        o[1] = best_ri;
        std::memcpy(&o[2], &disp, 4);
        std::memcpy(&o[3], &X, 4);
        std::memcpy(&o[4], &Y, 4);
        std::memcpy(&o[5], &Z, 4);
        o[6] = 0;
        ++n;
    }
    return n;
}

}  // extern "C"
