// Two-sweep raster aggregation (the batched form's sweep kernels).
//
// Same recurrence as sgm_paths.h (vx-vision/shaders/SGMStereo.metal:150-171,
// in place over ascending d, mod 2^16), scheduled differently: instead of one
// direction per scanline pass, the four directions whose predecessor lies on
// the current or the previous row advance together pixel by pixel in raster
// order, and their sum is stored once (forward sweep); the other four run the
// same way over the image turned by 180 degrees, add that sum and pick each
// pixel's winner (backward sweep, SGMStereo.metal:185-211).  Under the turn
// (x, y) -> (w-1-x, h-1-y) the backward directions (-1,0) (0,-1) (-1,-1)
// (+1,-1) become (+1,0) (0,+1) (+1,+1) (-1,+1), so one kernel body serves
// both sweeps; path sums are mod 2^16, so their order does not matter.
//
// Lanes.  A group of L lanes owns two consecutive rows, the first in the low
// half of every register and the second in the high half; lane `sub` owns
// disparities 4 sub .. 4 sub + 3.  Row r' of a wave (r' = 2 g or 2 g + 1)
// processes column x = t - 2 r' at step t, so the entries its vertical and
// diagonal terms need from the row above (columns x, x - 1, x + 1) were
// produced 2, 3 and 1 steps earlier.  Each lane keeps those outputs of its own
// rows (vertical 2 steps back, x-1 diagonal 3 back, x+1 diagonal 1 back): the
// second row of a group takes its operands from the low halves, the first row
// of a group from the previous group's high halves (a lane shift by L), and
// the wave's first group from the previous wave's last row, through that
// wave's ring of column slots in global memory.  The last group writes the
// same bundle it shifts, one slot per step; a wave publishes how many of its
// steps are in memory and its successor polls that word before loading
// slots.  Columns outside the image and rows below it hold the fresh state
// (zero entries, zero minimum), which is the start rule of every path, so
// the first and last column and the first row need no special case.
//
// Each wave's pair lives in one XCD's list (read from the hardware), so the
// ring traffic stays in that XCD's L2.
#pragma once

#include "sgm_paths.h"

namespace vxh {
namespace dev {

constexpr int kSwU = 6;     // steps per unrolled round (two halves of 3)
constexpr int kSwPF = 2;    // cost (and forward-sum) loads run this many steps ahead
constexpr int kSwRPF = 1;   // ring loads run this many steps ahead
constexpr int kSwNR = kSwRPF + 1;
static_assert(kSwU % (kSwPF + 1) == 0 && kSwU % kSwNR == 0 && kSwU % 2 == 0 && kSwU % 3 == 0,
              "the unrolled round must hold whole prefetch rings and whole row rings");
static_assert(kSweepRec == 48, "ring records are three 16-byte fields per lane");

#define VXH_GLOBAL __attribute__((address_space(1)))
typedef VXH_GLOBAL uint8_t gbyte;
typedef uint32_t nv2 __attribute__((ext_vector_type(2)));
typedef uint32_t nv4 __attribute__((ext_vector_type(4)));

// Global loads and stores at base + off (off aligned to the type).
template <typename T>
__device__ __forceinline__ T gld(const void *base, uint32_t off) {
    return *(const VXH_GLOBAL T *)((const gbyte *)base + off);
}
template <typename T>
__device__ __forceinline__ void gst(void *base, uint32_t off, T v) {
    *(VXH_GLOBAL T *)((gbyte *)base + off) = v;
}

// One direction of one step for both rows, in two parts around the carry
// check.  sw_spec: from the predecessor entries `prev` (prevq = prev + P1) and
// their minimum pm, with this pixel's stored costs C (cost + bias), the terms
// that do not depend on d-1 (A), the cost terms (B) and the chain
// q = agg + P1 assuming no carry into the lane; `take` masks the halves
// (rows) that take a carry from the lane below.  sw_rounds: the fixed-point
// rounds over the carry (only when some lane's carry changed).  sw_finish:
// agg = q - P1 and its minimum over d.
struct SwDir {
    uint32_t A[4], B[4], q[4], eu, e;
};
template <int L, bool ALIGNED>
__device__ __forceinline__ void sw_spec(SwDir &s, const uint32_t (&C)[4],
                                        const uint32_t (&prev)[4], const uint32_t (&prevq)[4],
                                        uint32_t pm, uint32_t take, uint32_t p2x2,
                                        uint32_t xqx2, const uint32_t (&om)[4]) {
    constexpr int K = 4;
    const uint32_t o3 = pk_add(pm, p2x2);
    const uint32_t pmx = pk_sub(pm, xqx2);
    const uint32_t nbq = from_upper_lane<L>(prevq[0]);
#pragma unroll
    for (int i = 0; i < K; ++i) {
        uint32_t o2 = (i + 1 < K) ? prevq[(i + 1) % K] : nbq;
        if (!ALIGNED || i == K - 1) o2 |= om[i];
        s.A[i] = pk_min(pk_min(prev[i], o2), o3);
        s.B[i] = pk_sub(C[i], pmx);
    }
    s.q[0] = pk_add(s.B[0], s.A[0]);
#pragma unroll
    for (int i = 1; i < K; ++i) s.q[i] = pk_add(s.B[i], pk_min(s.A[i], s.q[i - 1]));
    s.eu = s.A[0];
    const uint32_t cand = pk_min(s.A[0], from_lower_lane<L>(s.q[K - 1]));
    s.e = (cand & take) | (s.eu & ~take);
}
template <int L>
__device__ __forceinline__ void sw_rounds(SwDir &s, uint32_t take) {
    constexpr int K = 4;
    for (int r = 0; r < L - 1; ++r) {
        if (s.e != s.eu) {
            s.eu = s.e;
            s.q[0] = pk_add(s.B[0], s.e);
#pragma unroll
            for (int i = 1; i < K; ++i) s.q[i] = pk_add(s.B[i], pk_min(s.A[i], s.q[i - 1]));
        }
        const uint32_t cand = pk_min(s.A[0], from_lower_lane<L>(s.q[K - 1]));
        s.e = (cand & take) | (s.eu & ~take);
        if (lanes_ne(s.e, s.eu) == 0) break;
    }
}
template <int L, bool ALIGNED>
__device__ __forceinline__ void sw_finish(const SwDir &s, uint32_t p1x2,
                                          const uint32_t (&pad)[4], uint32_t lane_pad,
                                          uint32_t (&agg)[4], uint32_t &pmo) {
    uint32_t m = 0xFFFFFFFFu;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        agg[i] = pk_sub(s.q[i], p1x2);
        m = pk_min(m, ALIGNED ? agg[i] : (agg[i] | pad[i]));
    }
    if (ALIGNED) m |= lane_pad;
    pmo = group_min_pk<L>(m);
}

// The source word of a group's first row: the ring word `rw` in the first
// group, the previous group's word `hw` (lane - L) in the others.  For L = 32
// one permlane32 swap does both: its first result holds rw in lanes 0-31 and
// hw's lanes 0-31 in lanes 32-63; hw keeps its lanes 32-63 (the second
// result), which is all the last group (lanes 32-63) still needs of it.
template <int L>
__device__ __forceinline__ uint32_t first_row_src(uint32_t rw, uint32_t &hw, int g) {
    if constexpr (L == 32) {
        const auto p = __builtin_amdgcn_permlane32_swap(rw, hw, false, false);
        hw = (uint32_t)p[1];
        return (uint32_t)p[0];
    } else if constexpr (L == 16) {
        const int src = (int)(((threadIdx.x & 63) + 48) & 63);
        const uint32_t sh = (uint32_t)__builtin_amdgcn_ds_bpermute(src * 4, (int)hw);
        return g == 0 ? rw : sh;
    } else {
        return rw;
    }
}

// Per-pixel inputs of one step, loaded kSwPF steps ahead: the cost words of
// both rows and (backward sweep) their forward sums.
struct SwCost {
    uint32_t lo, hi;
    nv2 sl, sh;
};
// A ring slot (per lane): per direction (vertical, x-1 diagonal, x+1
// diagonal) the two words of the row's u16 entries, its minimum and a pad
// word, 16 bytes each.
struct SwRing {
    nv4 tb, dr, dl;
};

// Bounded wait until *word >= need (wave-uniform).  On expiry raises *error
// and returns a value that satisfies every later wait, so the wave finishes.
__device__ __forceinline__ uint32_t sw_wait(const uint32_t *word, uint32_t need,
                                            uint32_t *error, uint32_t polls) {
    for (uint32_t n = 0; n < polls; ++n) {
        const uint32_t v = (uint32_t)__builtin_amdgcn_readfirstlane(
            (int)__hip_atomic_load(word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        if (v >= need) return v;
        __builtin_amdgcn_s_sleep(1);
    }
    __hip_atomic_store(error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return 0xFFFFFFFFu;
}

// Wave-uniform copies (scalar registers) of a pointer and an integer read
// from memory with vector loads.
template <typename P>
__device__ __forceinline__ P *uni(P *p) {
    const uint64_t x = (uint64_t)p;
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)x);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(x >> 32));
    return (P *)(((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ uint32_t uni(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)x);
}

// One wave of a sweep: rows 2 (v G + g) and 2 (v G + g) + 1 of the (turned,
// for BWD) image for every group g.
template <int L, bool BWD, bool ALIGNED>
__device__ __forceinline__ void sweep_wave(const SweepJob &jb, int v, const uint8_t *fresh,
                                           uint32_t *error, uint32_t polls, uint32_t zero,
                                           uint32_t *lds) {
    constexpr int K = 4, G = 64 / L, U = kSwU, PF = kSwPF, RPF = kSwRPF, NR = kSwNR;
    const int lane = (int)(threadIdx.x & 63);
    const int g = lane / L, sub = lane % L, d0 = sub * K;
    const int W = uni(jb.w), H = uni(jb.h), D = uni(jb.D), Dp = uni(jb.Dp);
    const int nwaves = uni(jb.nwaves);
    const bool lane_real = d0 < D;
    const uint32_t dsafe = d0 < Dp ? (uint32_t)d0 : 0u;
    const bool lane_store = d0 < Dp;
    const uint32_t p1x2 = uni(jb.p1x2), p2x2 = uni(jb.p2x2), xqx2 = uni(jb.xqx2);
    // Rows of this group (logical, in sweep order) and their memory rows.
    const int rlo = 2 * (v * G + g);
    const bool vlo = rlo < H, vhi = rlo + 1 < H;
    const int mlo = BWD ? H - 1 - rlo : rlo, mhi = BWD ? H - 2 - rlo : rlo + 1;
    // Memory pixel of logical column x = pb + x * xs (rows outside the image
    // read row 0; their results are discarded).
    const int xs = BWD ? -1 : 1;
    const int pbl = (vlo ? mlo : 0) * W + (BWD ? W - 1 : 0);
    const int pbh = (vhi ? mhi : 0) * W + (BWD ? W - 1 : 0);
    // Every row of the wave is in the image and every lane stores (wave-uniform).
    const bool full_ok = 2 * (v * G + G - 1) + 1 < H && L * K <= Dp;
    // Steps of every wave: the last group's second row passes column W, and
    // its bundle for column W - 1 reaches the ring one step later.
    const int T = W + 4 * G;
    const int nslot = (int)sweep_ring_slots(W);
    const uint32_t slot = uni(jb.slot_bytes);

    uint32_t om[K], pad[K];
#pragma unroll
    for (int i = 0; i < K; ++i) {
        om[i] = (d0 + i + 1 < D) ? 0u : 0xFFFFFFFFu;
        pad[i] = (d0 + i < D) ? 0u : 0xFFFFFFFFu;
    }
    const uint32_t lane_pad = lane_real ? 0u : 0xFFFFFFFFu;
    const uint32_t take_full = (sub > 0 && lane_real) ? 0xFFFFFFFFu : 0u;

    // This wave's ring (written by its last group) and the previous wave's
    // (read by its first group; the zeroed slot above the first wave).
    // Addresses are a wave-uniform base plus a 32-bit lane offset.
    uint8_t *const ring_out = uni(jb.ring) + (size_t)v * nslot * slot;
    const uint8_t *const ring_in = v == 0 ? fresh : ring_out - (size_t)nslot * slot;
    const uint32_t in_stride = v == 0 ? 0u : slot;
    const uint32_t rec = (uint32_t)sub * kSweepRec;
    uint32_t *const prog = uni(jb.prog);
    uint32_t *const prog_out = prog + v;
    const uint32_t *const prog_in = v == 0 ? prog : prog + v - 1;
    uint32_t seen = v == 0 ? 0xFFFFFFFFu : 0u;   // the predecessor's published progress

    const uint8_t *const cost = uni(jb.cost);
    uint16_t *const sf = uni(jb.sf);
    float *const outp = uni(jb.out);
    int16_t *const bestp = uni(jb.best);
    uint16_t *const sump = uni(jb.sum);
    uint8_t *const sink = store_sink<uint8_t>();

    // Element offsets (pix * Dp + d0) of both rows at step t (columns clamped
    // into the image when !FULL).
    auto offsets = [&](int t, bool full, uint32_t &ol, uint32_t &oh, int &pl, int &ph) {
        int xl = t - 4 * g, xh = xl - 2;
        if (!full) {
            xl = min(max(xl, 0), W - 1);
            xh = min(max(xh, 0), W - 1);
        }
        pl = pbl + xl * xs;
        ph = pbh + xh * xs;
        ol = (uint32_t)pl * (uint32_t)Dp + dsafe;
        oh = (uint32_t)ph * (uint32_t)Dp + dsafe;
    };
    auto issue_cost = [&](SwCost &b, int t, auto fullc) {
        constexpr bool FULL = decltype(fullc)::value;
        uint32_t ol, oh;
        int pl, ph;
        offsets(t, FULL, ol, oh, pl, ph);
        b.lo = gld<uint32_t>(cost, ol);
        b.hi = gld<uint32_t>(cost, oh);
        if constexpr (BWD) {
            b.sl = gld<nv2>(sf, 2 * ol);
            b.sh = gld<nv2>(sf, 2 * oh);
        }
    };
    // The first group's first row at step t reads the slot of its own
    // column t (slot index t + 1).
    auto issue_ring = [&](SwRing &b, int t, auto fullc) {
        constexpr bool FULL = decltype(fullc)::value;
        int idx = t + 1;
        if (!FULL) idx = min(idx, nslot - 1);
        const uint32_t o = (uint32_t)idx * in_stride + rec;
        b.tb = gld<nv4>(ring_in, o);
        b.dr = gld<nv4>(ring_in, o + 16);
        b.dl = gld<nv4>(ring_in, o + 32);
    };
    // Progress a wave needs from its predecessor before loading the slot of
    // step s: the predecessor writes it at its step s + 4 G.
    auto wait_pred = [&](int s) {
        const uint32_t need = (uint32_t)min(s + 4 * G + 1, T);
        if (seen < need) seen = sw_wait(prog_in, need, error, polls);
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
    };

    // Running state: the rows' own (+1,0) path, and the last outputs of the
    // three directions the row below needs (vertical 2 steps back, the x-1
    // diagonal 3 back, the x+1 diagonal 1 back): agg entries and minimum.
    uint32_t lrp[K], lrq[K], lrpm = 0;
    uint32_t tb[2][K], tbm[2], dr[3][K], drm[3], dl[K], dlm = 0;
#pragma unroll
    for (int i = 0; i < K; ++i) {
        lrp[i] = 0;
        lrq[i] = p1x2;
        dl[i] = 0;
        tb[0][i] = tb[1][i] = 0;
        dr[0][i] = dr[1][i] = dr[2][i] = 0;
    }
    tbm[0] = tbm[1] = 0;
    drm[0] = drm[1] = drm[2] = 0;

    SwCost cb[PF + 1];
    SwRing rb[NR];
#pragma unroll
    for (int j = 0; j < PF; ++j) issue_cost(cb[j], j, std::false_type{});
    wait_pred(RPF - 1);
#pragma unroll
    for (int j = 0; j < RPF; ++j) issue_ring(rb[j], j, std::false_type{});

    auto step = [&](auto uc, int t, auto fullc) {
        constexpr int u = decltype(uc)::value;
        constexpr bool FULL = decltype(fullc)::value;
        issue_cost(cb[(u + PF) % (PF + 1)], t + PF, fullc);
        issue_ring(rb[(u + RPF) % NR], t + RPF, fullc);
        const SwCost &c = cb[u % (PF + 1)];
        const SwRing &rr = rb[u % NR];
        const int xl = t - 4 * g, xh = xl - 2;
        const bool alo = FULL || (vlo && xl >= 0 && xl < W);
        const bool ahi = FULL || (vhi && xh >= 0 && xh < W);
        const uint32_t keep = (alo ? 0x0000FFFFu : 0u) | (ahi ? 0xFFFF0000u : 0u);
        const uint32_t take = FULL ? take_full : (take_full & keep);

        uint32_t C[K];
#pragma unroll
        for (int i = 0; i < K; ++i)
            C[i] = perm(c.hi, c.lo, 0x0C000C00u | ((4u + i) << 16) | (uint32_t)i);

        // Operands of the vertical and diagonal directions: the second row's
        // from this lane's low halves, the first row's from the ring or the
        // previous group.  hw: this lane's high halves (its second row's
        // entries and minimum), which the next group and the ring take.
        uint32_t (&tbs)[K] = tb[u % 2];
        uint32_t (&drs)[K] = dr[u % 3];
        auto assemble = [&](const uint32_t (&own)[K], uint32_t ownm, const nv4 &ring, nv4 &hw,
                            uint32_t (&pv)[K], uint32_t (&pq)[K], uint32_t &pmv) {
            uint32_t hx = perm(own[1], own[0], 0x07060302u);
            uint32_t hy = perm(own[3], own[2], 0x07060302u);
            uint32_t hz = ownm >> 16;
            const uint32_t sx = first_row_src<L>(ring.x, hx, g);
            const uint32_t sy = first_row_src<L>(ring.y, hy, g);
            const uint32_t sz = first_row_src<L>(ring.z, hz, g);
            hw = nv4{hx, hy, hz, 0u};
            pv[0] = perm(own[0], sx, 0x05040100u);
            pv[1] = perm(own[1], sx, 0x05040302u);
            pv[2] = perm(own[2], sy, 0x05040100u);
            pv[3] = perm(own[3], sy, 0x05040302u);
            pmv = perm(ownm, sz, 0x05040100u);
#pragma unroll
            for (int i = 0; i < K; ++i) pq[i] = pk_add(pv[i], p1x2);
        };
        uint32_t tp[K], tq[K], dp[K], dq[K], lp[K], lq[K], tpm, dpm, lpm;
        nv4 htb, hdr, hdl;
        assemble(tbs, tbm[u % 2], rr.tb, htb, tp, tq, tpm);
        assemble(drs, drm[u % 3], rr.dr, hdr, dp, dq, dpm);
        assemble(dl, dlm, rr.dl, hdl, lp, lq, lpm);

        // The last group's second row feeds the next wave: the bundle just
        // assembled (vertical entries of column x - 2, the x-1 diagonal's of
        // column x - 3, the x+1 diagonal's of column x - 1, x = its column)
        // is what that wave's first row reads at column x - 2.
        if (g == G - 1) {
            int idx = t - 4 * G + 1;
            if (!FULL) idx = min(max(idx, 0), nslot - 1);
            const uint32_t o = (uint32_t)idx * slot + rec;
            gst<nv4>(ring_out, o, htb);
            gst<nv4>(ring_out, o + 16, hdr);
            gst<nv4>(ring_out, o + 32, hdl);
        }

        // The four directions, with one carry check for all of them.
        SwDir s0, s1, s2, s3;
        sw_spec<L, ALIGNED>(s0, C, lrp, lrq, lrpm, take, p2x2, xqx2, om);
        sw_spec<L, ALIGNED>(s1, C, tp, tq, tpm, take, p2x2, xqx2, om);
        sw_spec<L, ALIGNED>(s2, C, dp, dq, dpm, take, p2x2, xqx2, om);
        sw_spec<L, ALIGNED>(s3, C, lp, lq, lpm, take, p2x2, xqx2, om);
        if (__builtin_expect((lanes_ne(s0.e, s0.eu) | lanes_ne(s1.e, s1.eu) |
                              lanes_ne(s2.e, s2.eu) | lanes_ne(s3.e, s3.eu)) != 0, 0)) {
            sw_rounds<L>(s0, take);
            sw_rounds<L>(s1, take);
            sw_rounds<L>(s2, take);
            sw_rounds<L>(s3, take);
        }
        uint32_t a0[K], m0, a1[K], m1, a2[K], m2, a3[K], m3;
        sw_finish<L, ALIGNED>(s0, p1x2, pad, lane_pad, a0, m0);
        sw_finish<L, ALIGNED>(s1, p1x2, pad, lane_pad, a1, m1);
        sw_finish<L, ALIGNED>(s2, p1x2, pad, lane_pad, a2, m2);
        sw_finish<L, ALIGNED>(s3, p1x2, pad, lane_pad, a3, m3);
        uint32_t q0[K];
#pragma unroll
        for (int i = 0; i < K; ++i) q0[i] = s0.q[i];
        if (!FULL) {
            // Columns outside the image and rows below it hold the fresh
            // state: what a path's first pixel starts from.
#pragma unroll
            for (int i = 0; i < K; ++i) {
                a0[i] &= keep;
                q0[i] = (q0[i] & keep) | (p1x2 & ~keep);
                a1[i] &= keep;
                a2[i] &= keep;
                a3[i] &= keep;
            }
            m0 &= keep;
            m1 &= keep;
            m2 &= keep;
            m3 &= keep;
        }
        uint32_t S[K];
#pragma unroll
        for (int i = 0; i < K; ++i) {
            S[i] = pk_add(pk_add(a0[i], a1[i]), pk_add(a2[i], a3[i]));
            lrp[i] = a0[i];
            lrq[i] = q0[i];
            tbs[i] = a1[i];
            drs[i] = a2[i];
            dl[i] = a3[i];
        }
        lrpm = m0;
        tbm[u % 2] = m1;
        drm[u % 3] = m2;
        dlm = m3;

        uint32_t ol, oh;
        int pl, ph;
        offsets(t, FULL, ol, oh, pl, ph);
        // A row's entries in disparity order: low halves, high halves.
        const nv2 wl = nv2{perm(S[1], S[0], 0x05040100u), perm(S[3], S[2], 0x05040100u)};
        const nv2 wh = nv2{perm(S[1], S[0], 0x07060302u), perm(S[3], S[2], 0x07060302u)};
        auto store_rows = [&](uint16_t *base, nv2 xl2, nv2 xh2) {
            if constexpr (FULL) {
                gst<nv2>(base, 2 * ol, xl2);
                gst<nv2>(base, 2 * oh, xh2);
            } else {
                const bool sl = alo && lane_store, sh = ahi && lane_store;
                gst<nv2>(sl ? (void *)base : (void *)sink, sl ? 2 * ol : 0u, xl2);
                gst<nv2>(sh ? (void *)base : (void *)sink, sh ? 2 * oh : 8u, xh2);
            }
        };
        if constexpr (!BWD) {
            store_rows(sf, wl, wh);   // the forward sum
        } else {
            // Full sum, then each row's first minimum and its neighbours.
#pragma unroll
            for (int i = 0; i < K; ++i) {
                const uint32_t fl = i < 2 ? c.sl.x : c.sl.y, fh = i < 2 ? c.sh.x : c.sh.y;
                S[i] = pk_add(S[i], perm(fh, fl, (i & 1) ? 0x07060302u : 0x05040100u));
            }
            if (sump != nullptr) {
                const nv2 sl2 = nv2{perm(S[1], S[0], 0x05040100u), perm(S[3], S[2], 0x05040100u)};
                const nv2 sh2 = nv2{perm(S[1], S[0], 0x07060302u), perm(S[3], S[2], 0x07060302u)};
                store_rows(sump, sl2, sh2);
            }
            uint32_t mn = 0xFFFFFFFFu;
#pragma unroll
            for (int i = 0; i < K; ++i) mn = pk_min(mn, S[i] | pad[i]);
            mn = group_min_pk<L>(mn);
            // First d with S[d] == min, both rows at once: d where equal,
            // 0xFFFF elsewhere, then the minimum.
            uint32_t key = 0xFFFFFFFFu;
#pragma unroll
            for (int i = 0; i < K; ++i) {
                const uint32_t ne = pk_min(pk_sub(S[i], mn), 0x00010001u);   // 1 where S > min
                const uint32_t dd = (uint32_t)(d0 + i) * 0x00010001u;
                key = pk_min(key, dd | pk_sub(0u, ne) | pad[i]);
            }
            key = group_min_pk<L>(key);
            // Neighbours S[bd - 1], S[bd + 1] through LDS (entry e of the
            // group at word g * 4L + e).
            const uint32_t *gl = lds + g * 4 * L;
            *(uint4 *)(lds + lane * 4) = make_uint4(S[0], S[1], S[2], S[3]);
            const int bl = (int)(key & 0xFFFFu), bh = (int)(key >> 16);
            const int cap = 4 * L - 1;
            const uint32_t lm = gl[min(max(bl - 1, 0), cap)], lp2 = gl[min(bl + 1, cap)];
            const uint32_t hm = gl[min(max(bh - 1, 0), cap)], hp2 = gl[min(bh + 1, cap)];
            if (sub == 0) {
                auto finish = [&](int bd, uint32_t bc, uint32_t c0u, uint32_t c2u, bool act,
                                  int pix) {
                    const int bdw = bc == 0xFFFFu ? 0 : bd;
                    float disp = (float)bdw;
                    if (bdw > 0 && bdw < D - 1) {
                        const float c0 = (float)c0u;
                        const float c1 = (float)bc;
                        const float c2 = (float)c2u;
                        const float den = 2.0f * (c0 + c2 - 2.0f * c1);
                        if (fabsf(den) > 1e-6f) disp = disp + (c0 - c2) / den;
                    }
                    if (act) {
                        gst<float>(outp, (uint32_t)pix * 4u, disp);
                        if (bestp) gst<int16_t>(bestp, (uint32_t)pix * 2u, (int16_t)bdw);
                    }
                };
                finish(bl, mn & 0xFFFFu, lm & 0xFFFFu, lp2 & 0xFFFFu, alo, pl);
                finish(bh, mn >> 16, hm >> 16, hp2 >> 16, ahi, ph);
            }
        }
    };

    // After step t the cost loads it used were issued before the stores of
    // step t - PF, and memory operations complete in issue order, so every
    // store of the steps before t - PF is in memory: publish t - PF (ordered
    // after that use by a data dependence on the loaded word).
    auto publish = [&](int t, const SwCost &used) {
        if (v + 1 < nwaves) {
            const uint32_t dep = (uint32_t)__builtin_amdgcn_readfirstlane((int)used.lo) & zero;
            if (lane == 0)
                __hip_atomic_store(prog_out, (uint32_t)(t - PF) | dep, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        }
    };
    const int rounds = (T + U - 1) / U;
    // Rounds whose steps and prefetches all lie inside every row of the wave
    // (and whose ring slots need no clamp).
    const int full_from = (4 * G - 1 + U - 1) / U;
    const int full_to = full_ok ? (W - PF) / U : 0;   // r < full_to: r U + U - 1 + PF <= W - 1
    // Half rounds of 3 steps: wait for the predecessor before each (its ring
    // loads reach step t0 + 2 + RPF), publish after each.  Steps keep their
    // memory operations in program order (the published progress relies on
    // it).
    auto round = [&](int t0, auto fullc) {
        static_for<U>([&](auto uc) {
            constexpr int u = decltype(uc)::value;
            if constexpr (u % 3 == 0) wait_pred(t0 + u + 2 + RPF);
            step(uc, t0 + u, fullc);
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            if constexpr (u % 3 == 2) publish(t0 + u, cb[u % (PF + 1)]);
        });
    };
    for (int r = 0; r < rounds; ++r) {
        if (r >= full_from && r < full_to)
            round(r * U, std::true_type{});
        else
            round(r * U, std::false_type{});
    }
    // Every store of the wave is in memory before its last progress word.
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    if (lane == 0 && v + 1 < nwaves)
        __hip_atomic_store(prog_out, (uint32_t)T, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace dev

namespace {

// The hardware XCD a wave runs on.
__device__ __forceinline__ int xcd_id() {
    return (int)(__builtin_amdgcn_s_getreg((31 << 11) | 20) & 7);   // HW_REG_XCC_ID
}

template <int L, bool BWD>
__global__ void __launch_bounds__(64, 2)
sweep_kernel(SweepArgs a) {
    __shared__ uint32_t lds[256];
    const int k = xcd_id();
    const uint32_t n = a.item_cnt[k];
    const SweepItem *__restrict__ items = a.items + a.item_off[k];
    const int lane = (int)(threadIdx.x & 63);
    for (;;) {
        uint32_t tk = 0;
        if (lane == 0)
            tk = __hip_atomic_fetch_add(a.counters + k, 1u, __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t ticket = (uint32_t)__builtin_amdgcn_readfirstlane((int)tk);
        if (ticket >= n) break;   // wave-uniform
        const SweepItem it = items[ticket];
        const uint32_t job = dev::uni(it.job), wave = dev::uni(it.wave);
        const SweepJob &jb = a.jobs[job];
        if (dev::uni(jb.D) % 4 == 0)
            dev::sweep_wave<L, BWD, true>(jb, (int)wave, a.fresh, a.error, a.wait_polls,
                                          a.zero, lds);
        else
            dev::sweep_wave<L, BWD, false>(jb, (int)wave, a.fresh, a.error, a.wait_polls,
                                           a.zero, lds);
    }
}

}  // namespace
}  // namespace vxh

// Defines launch_sweep_c<CC>() and sweep_occupancy_c<CC>() for lane count
// 16 << CC.
#define VXH_DEFINE_SWEEP_LAUNCHER(CC)                                                        \
    namespace vxh {                                                                          \
    hipError_t launch_sweep_c##CC(bool bwd, const SweepArgs &a, int blocks, hipStream_t s) { \
        if (bwd)                                                                             \
            hipLaunchKernelGGL((sweep_kernel<(16 << CC), true>), dim3(blocks), dim3(64), 0, \
                               s, a);                                                        \
        else                                                                                 \
            hipLaunchKernelGGL((sweep_kernel<(16 << CC), false>), dim3(blocks), dim3(64), 0, \
                               s, a);                                                        \
        return hipGetLastError();                                                            \
    }                                                                                        \
    int sweep_occupancy_c##CC(bool bwd) {                                                    \
        int n = 0;                                                                           \
        const hipError_t e =                                                                 \
            bwd ? hipOccupancyMaxActiveBlocksPerMultiprocessor(                              \
                      &n, sweep_kernel<(16 << CC), true>, 64, 0)                             \
                : hipOccupancyMaxActiveBlocksPerMultiprocessor(                              \
                      &n, sweep_kernel<(16 << CC), false>, 64, 0);                           \
        return e == hipSuccess ? n : 0;                                                      \
    }                                                                                        \
    }
