// Path aggregation over scanline pairs, packed two per register (gfx950).
//
// Reproduces vx-vision/shaders/SGMStereo.metal:72-173 (sgm_cost_aggregate) for
// one direction, bit for bit, and extends it to the diagonals (a path starts
// at every pixel whose predecessor lies outside the image).
//
// Lane layout.  A group of L lanes of one wave owns TWO scanlines of the same
// direction (scanlines 2g and 2g+1 of the wave's range): every register holds
// a u16 of the first scanline in its low half and of the second in its high
// half, so the whole recurrence runs on packed u16 instructions
// (v_pk_add_u16 / v_pk_sub_u16 / v_pk_min_u16), which wrap mod 2^16 per half
// exactly as the shader's ushort arithmetic does.  Lane `sub` of the group owns
// disparities d0 = sub*K .. d0+K-1 of both scanlines; cross-lane steps (the
// d+1 neighbour, the carry from d-1, the minimum over d) are DPP / permlane
// moves inside the wave.
//
// The reference updates prev[] in place over ascending d, so the d-1 term
// reads the value already produced in the same step (SGMStereo.metal:155,167).
// The d, d+1 and P2 terms do not depend on d-1 and are evaluated for all d at
// once (A[d]); the d-1 chain runs sequentially inside each lane and is carried
// across lanes by fixed-point rounds: every lane first assumes no carry-in,
// then re-evaluates whenever the carry-in from the lane below changes
// min(A[first], o1).  Lane 0's carry-in is fixed (none), so after r rounds
// lanes 0..r are final; the loop stops as soon as no lane changes, which is
// exactly the sequential result including every wrap.
//
// Output.  Each direction writes its own volume (entry pix*Dp + d):
//   kOutU16  the u16 path cost;
//   kOutU8   its low byte, plus, for the rare slices in which some entry
//            exceeds 255 (only after a mod-2^16 wrap when 64 + P2 <= 255),
//            the high bytes of the whole slice into a second volume and a
//            flag byte (pixel, direction) that tells the summing pass to add
//            them.  Sums are mod 2^16, so low + 256 * high reproduces the
//            u16 path cost exactly.
#pragma once

#include <type_traits>

#include "sgm_device.h"

namespace vxh {
namespace dev {

__device__ __forceinline__ uint32_t pk_add(uint32_t a, uint32_t b) {
    return __builtin_bit_cast(uint32_t,
                              (u16x2)(__builtin_bit_cast(u16x2, a) + __builtin_bit_cast(u16x2, b)));
}
__device__ __forceinline__ uint32_t pk_sub(uint32_t a, uint32_t b) {
    return __builtin_bit_cast(uint32_t,
                              (u16x2)(__builtin_bit_cast(u16x2, a) - __builtin_bit_cast(u16x2, b)));
}
__device__ __forceinline__ uint32_t pk_min(uint32_t a, uint32_t b) {
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_min(__builtin_bit_cast(u16x2, a),
                                                                  __builtin_bit_cast(u16x2, b)));
}
__device__ __forceinline__ uint32_t pk_max(uint32_t a, uint32_t b) {
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(u16x2, a),
                                                                  __builtin_bit_cast(u16x2, b)));
}
// Byte permute: result byte k = byte sel[k] of the 8-byte value (hi:lo), with
// selector 0x0c giving zero.
__device__ __forceinline__ uint32_t perm(uint32_t hi, uint32_t lo, uint32_t sel) {
    return __builtin_amdgcn_perm(hi, lo, sel);
}

template <int L>
__device__ __forceinline__ uint32_t group_min_pk(uint32_t v) {
    return group_reduce<L>(v, [](uint32_t a, uint32_t b) { return pk_min(a, b); });
}

// Stores the first BYTES bytes of w (BYTES in {2, 4, 8, 16, 32}; p aligned
// to min(BYTES, 16)).  Plain stores: non-temporal ones for the direction
// volumes were measured 1.5x slower for the whole launch.
template <int BYTES>
__device__ __forceinline__ void store_seg(uint8_t *p, const uint32_t *w) {
    if constexpr (BYTES == 2) *(uint16_t *)p = (uint16_t)w[0];
    else if constexpr (BYTES == 4) *(uint32_t *)p = w[0];
    else if constexpr (BYTES == 8) *(uint2 *)p = make_uint2(w[0], w[1]);
    else {
#pragma unroll
        for (int j = 0; j < BYTES / 16; ++j)
            ((uint4 *)p)[j] = make_uint4(w[4 * j], w[4 * j + 1], w[4 * j + 2], w[4 * j + 3]);
    }
}

// One byte of every half of v[0..K) (byte 0 of each half for sel 0x06020400,
// byte 1 for 0x07030501), regrouped per scanline in disparity order:
// wa = bytes of the low halves, wb = of the high halves, 4 per word (for
// K = 2, the two bytes in the low 16 bits).
template <int K>
__device__ __forceinline__ void bytes_of_pairs(const uint32_t *v, uint32_t sel, uint32_t *wa,
                                               uint32_t *wb) {
    uint32_t t[K / 2];   // [a_2j a_2j+1 b_2j b_2j+1]
#pragma unroll
    for (int j = 0; j < K / 2; ++j) t[j] = perm(v[2 * j + 1], v[2 * j], sel);
    if constexpr (K == 2) {
        wa[0] = t[0] & 0xFFFFu;
        wb[0] = t[0] >> 16;
    } else {
#pragma unroll
        for (int k = 0; k < K / 4; ++k) {
            wa[k] = perm(t[2 * k + 1], t[2 * k], 0x05040100u);
            wb[k] = perm(t[2 * k + 1], t[2 * k], 0x07060302u);
        }
    }
}

// Cost loads of one step for both scanlines of a lane's group: the lane's K
// entries of each pixel's cost slice (K bytes).
template <int K>
struct PairBuf {
    uint32_t c[2][(K + 3) / 4];
};

// Split scanlines (head role): the tail's stored values of one step, both
// scanlines (this lane's entries: u8 K bytes or u16 2K bytes) and their
// exception flag bytes.
template <int K>
struct TailBuf {
    uint32_t v[2][(2 * K + 3) / 4];
    uint32_t flag[2];
};

// A scanline of length len >= kSplitMin is split at m = split_point(len).
// Its tail role starts kSplitWarm steps before m from the same fresh state as
// any path start, runs those warm-up steps without storing (the state forgets
// its start within a few dozen steps) and stores [m, len).  Its head role runs
// [0, m) and then continues into the tail's range until its state equals the
// tail's stored one bit for bit (every later step then agrees, the recurrence
// being deterministic), overwriting the tail's values until then; after the
// warm-up that is almost always the first step, m.  m balances the two:
// m = len - m + kSplitWarm.
constexpr int kSplitWarm = 96;
constexpr int kSplitMin = 256;
__device__ __forceinline__ int split_point(int len) {
    return len >= kSplitMin ? (len + kSplitWarm) / 2 : len;
}

// Roles of a wave in a split launch.
constexpr int kRoleWhole = 0, kRoleHead = 1, kRoleTail = 2;

// Bounded wait for one of a tail's flags (`polls` polls, seconds of polling
// by default): a head that times out raises *error and treats the flag as
// set.
__device__ __forceinline__ void wait_flag(const uint32_t *flag, uint32_t *error,
                                          uint32_t polls) {
    for (uint32_t n = 0; n < polls; ++n) {
        const uint32_t v = __hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if ((uint32_t)__builtin_amdgcn_readfirstlane((int)v) != 0) {
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            return;
        }
        __builtin_amdgcn_s_sleep(4);
    }
    __hip_atomic_store(error, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// One wave aggregates scanlines s_first .. s_first + 2 * (64 / L) - 1 of
// direction p.dir (those below p.nscan) into the direction's volume, in
// storage MODE (kOutU8 or kOutU16; -1 = p.mode at run time).  ALIGNED: the
// caller guarantees D % K == 0, so every lane is either wholly inside D or
// wholly padding and only the element at D - 1 needs its d+1 term masked.
// SPLIT: `role` is kRoleWhole, kRoleHead or kRoleTail of split scanlines
// (see split_point).  A tail sets *tail_early (release) once its values of
// the first kSplitRange steps from m are stored; its caller sets *tail_done
// once all of them are.  A head reads the tail's values of [m, m + range)
// after *tail_early and any later ones after *tail_done, which also ends the
// tail's stores, so the head may then overwrite them.
// Wave-level: every lane of the wave must call it with the same arguments.
template <int L, int K, int MODE, bool ALIGNED, bool SPLIT = false>
__device__ __forceinline__ void agg_pair_wave(const PassParams &p, int s_first,
                                              int role = kRoleWhole,
                                              const uint32_t *tail_done = nullptr,
                                              uint32_t *tail_early = nullptr,
                                              uint32_t *error = nullptr) {
    static_assert(K >= 2 && K <= 16 && (K & (K - 1)) == 0, "K must be 2, 4, 8 or 16");
    const int lane = threadIdx.x & 63;
    const int g = lane / L, sub = lane % L;
    const int w = p.w, h = p.h, D = p.D, Dp = p.Dp;
    const int dx = p.dx, dy = p.dy;
    const int d0 = sub * K;
    const bool lane_real = d0 < D;
    const bool lane_store = d0 < Dp;
    const uint32_t p1x2 = p.p1x2, p2x2 = p.p2x2, xqx2 = p.xqx2;
    // Lanes past the volume stride read their pixel's first entries (unused).
    const uint8_t *__restrict__ cost = p.cost + (lane_store ? d0 : 0);

    int sx[2] = {0, 0}, len[2] = {0, 0}, pix0[2] = {0, 0};
    // Head role: steps [vfrom, len) of a scanline check the tail's values.
    int vfrom[2] = {0x7FFFFFFF, 0x7FFFFFFF};
    // Tail role: the steps before sfrom are the unstored warm-up.
    int sfrom[2] = {0, 0};
#pragma unroll
    for (int hh = 0; hh < 2; ++hh) {
        const int s = s_first + 2 * g + hh;
        int sy = 0;
        if (s < p.nscan) scan_start(s, dx, dy, w, h, sx[hh], sy, len[hh]);
        if (SPLIT && role == kRoleTail) {
            const int m = split_point(len[hh]);
            const int start = m < len[hh] ? m - kSplitWarm : len[hh];
            sx[hh] += start * dx;
            sy += start * dy;
            len[hh] -= start;
            sfrom[hh] = m - start;
        } else if (SPLIT && role == kRoleHead) {
            const int m = split_point(len[hh]);
            if (m < len[hh]) vfrom[hh] = m;
        }
        pix0[hh] = sy * w + sx[hh];
    }
    int maxlen = max(len[0], len[1]), minlen = min(len[0], len[1]);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        maxlen = max(maxlen, __shfl_xor(maxlen, o));
        minlen = min(minlen, __shfl_xor(minlen, o));
    }
    if (maxlen == 0) return;
    const int pstep = dy * w + dx;
    const int mode = MODE >= 0 ? MODE : p.mode;
    const size_t ebytes = mode == kOutU16 ? 2 : 1;
    // Volume pointers of this lane's segment for both scanlines, advanced by
    // one pixel step per step (64-bit: a volume may exceed 4 GB).
    uint8_t *vp[2];
#pragma unroll
    for (int hh = 0; hh < 2; ++hh)
        vp[hh] = p.vol + ((size_t)pix0[hh] * Dp + d0) * ebytes;
    const ptrdiff_t vstep = (ptrdiff_t)pstep * Dp * (ptrdiff_t)ebytes;
    int tcur = 0;   // the step vp points at

    // o2 is 0xFFFF where d + 1 >= D (SGMStereo.metal:156); entries d >= D
    // are padding and never enter the minimum.
    uint32_t om[K], pad[K];
#pragma unroll
    for (int i = 0; i < K; ++i) {
        om[i] = (d0 + i + 1 < D) ? 0u : 0xFFFFFFFFu;
        pad[i] = (d0 + i < D) ? 0u : 0xFFFFFFFFu;
    }
    // Aligned: a padding lane's minimum is masked once per step.
    const uint32_t lane_pad = lane_real ? 0u : 0xFFFFFFFFu;
    // prevq[i] = prev[i] + P1, kept from the chain of the step that produced
    // prev (the d+1 term of the next step reads it, as the d-1 term of the
    // chain did), so no step adds P1 to prev again.
    uint32_t prev[K], prevq[K];
#pragma unroll
    for (int i = 0; i < K; ++i) {
        prev[i] = 0;
        prevq[i] = p1x2;
    }
    uint32_t pm = 0;

    // Head role: the first step at which any scanline of the wave reads the
    // tail's values (the wait for the tail comes before its loads), and the
    // scanlines whose state has met the tail's (from then on the tail's
    // values stand and this wave stores nothing more for them).
    int vfirst = min(vfrom[0], vfrom[1]);
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) vfirst = min(vfirst, __shfl_xor(vfirst, o));
    vfirst = __builtin_amdgcn_readfirstlane(vfirst);
    bool met[2] = {false, false};
    // Bytes of this lane's stored segment that hold real entries (d < D).
    uint32_t realmask[(2 * K + 3) / 4];
#pragma unroll
    for (int q = 0; q < (2 * K + 3) / 4; ++q) {
        uint32_t m = 0;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int e = (q * 4 + b) / (int)ebytes;   // entry index of byte b
            if (d0 + e < D) m |= 0xFFu << (8 * b);
        }
        realmask[q] = m;
    }

    // Loads run PF steps ahead through a ring of NB = PF + 1 buffers (see the
    // main loop below).
    constexpr int PF = K <= 4 ? 6 : (K <= 8 ? 4 : 3);
    constexpr int NB = PF + 1;
    // Steps from m that a tail publishes early (two rounds).
    constexpr int kRange = 2 * NB;

    // FULL: step t lies inside every scanline of the wave (no clamp needed);
    // the loads then come from fptr, which the full rounds advance by one
    // pixel step per issue instead of recomputing pix * Dp.
    const uint8_t *fptr[2];
    const ptrdiff_t cstep = (ptrdiff_t)pstep * Dp;
    auto issue = [&](PairBuf<K> &b, int t, auto full) {
        constexpr bool FULL = decltype(full)::value;
#pragma unroll
        for (int hh = 0; hh < 2; ++hh) {
            // Inactive scanlines read pixel 0.
            const uint8_t *src;
            if constexpr (FULL) {
                src = fptr[hh];
                fptr[hh] += cstep;
            } else {
                const uint32_t pix = t < len[hh] ? (uint32_t)(pix0[hh] + t * pstep) : 0u;
                src = cost + (size_t)pix * (uint32_t)Dp;
            }
            if constexpr (K == 2) b.c[hh][0] = *(const uint16_t *)src;
            else if constexpr (K == 4) b.c[hh][0] = *(const uint32_t *)src;
            else if constexpr (K == 8) {
                const uint2 q = *(const uint2 *)src;
                b.c[hh][0] = q.x;
                b.c[hh][1] = q.y;
            } else {
                const uint4 q = *(const uint4 *)src;
                b.c[hh][0] = q.x;
                b.c[hh][1] = q.y;
                b.c[hh][2] = q.z;
                b.c[hh][3] = q.w;
            }
        }
    };
    // Head role: the tail's stored values of step t (loaded only after the
    // tail has finished; other lanes and steps read the sink).
    auto issue_tail = [&](TailBuf<K> &tb, int t) {
        if (t < vfirst) return;   // wave-uniform: only the checking steps load
        if (t == vfirst) wait_flag(tail_early, error, p.wait_polls);
        if (t == vfirst + kRange) wait_flag(tail_done, error, p.wait_polls);
#pragma unroll
        for (int hh = 0; hh < 2; ++hh) {
            const bool v = t >= vfrom[hh] && t < len[hh] && lane_store;
            const uint8_t *src = v ? vp[hh] + (ptrdiff_t)(t - tcur) * vstep
                                   : (const uint8_t *)store_sink<uint8_t>();
            if (mode == kOutU16) {
                if constexpr (K == 2) tb.v[hh][0] = *(const uint32_t *)src;
                else if constexpr (K == 4) { uint2 q = *(const uint2 *)src; tb.v[hh][0] = q.x; tb.v[hh][1] = q.y; }
                else {
#pragma unroll
                    for (int j = 0; j < K / 8; ++j) {
                        uint4 q = ((const uint4 *)src)[j];
                        tb.v[hh][4 * j] = q.x; tb.v[hh][4 * j + 1] = q.y;
                        tb.v[hh][4 * j + 2] = q.z; tb.v[hh][4 * j + 3] = q.w;
                    }
                }
                tb.flag[hh] = 0;
            } else {
                if constexpr (K == 2) tb.v[hh][0] = *(const uint16_t *)src;
                else if constexpr (K == 4) tb.v[hh][0] = *(const uint32_t *)src;
                else if constexpr (K == 8) { uint2 q = *(const uint2 *)src; tb.v[hh][0] = q.x; tb.v[hh][1] = q.y; }
                else { uint4 q = *(const uint4 *)src; tb.v[hh][0] = q.x; tb.v[hh][1] = q.y; tb.v[hh][2] = q.z; tb.v[hh][3] = q.w; }
                const uint8_t *fp = v ? p.flags + (size_t)(pix0[hh] + t * pstep) * 8 + p.dir
                                      : (const uint8_t *)store_sink<uint8_t>();
                tb.flag[hh] = *fp;
            }
        }
    };

    // Slices of step tp with an entry above 255 are rare (only after a wrap).
    // `high` holds that step's high bytes (OR over the lane's entries); the
    // wave tests them one step later, when the test no longer waits on the
    // recurrence, and only then works out which scanline of which group is
    // affected: their high bytes (still in prev) go to the high-byte volume
    // and the pixel's flag is set.
    uint32_t high = 0;
    uint64_t high_lanes = 0;   // lanes with high != 0, taken when high is set
    auto flush_high = [&](int tp) {
        if (high_lanes == 0) return;
        // `high` was masked to the scanlines that stored at step tp.
        const uint32_t exc = lane_real ? (((high & 0x0000FF00u) ? 1u : 0u) |
                                          ((high & 0xFF000000u) ? 2u : 0u))
                                       : 0u;
        const uint32_t f = group_or<L>(exc);
        uint32_t ha[(K + 3) / 4], hb[(K + 3) / 4];
        bytes_of_pairs<K>(prev, 0x07030501u, ha, hb);
        const ptrdiff_t hoff = p.hi - p.vol;
        if ((f & 1u) && lane_store) store_seg<K>(vp[0] - vstep + hoff, ha);
        if ((f & 2u) && lane_store) store_seg<K>(vp[1] - vstep + hoff, hb);
        if (sub == 0) {
#pragma unroll
            for (int hh = 0; hh < 2; ++hh)
                if (f & (1u << hh)) p.flags[(size_t)(pix0[hh] + tp * pstep) * 8 + p.dir] = 1;
        }
        high = 0;
        high_lanes = 0;
    };

    // VERIFY: a head's step at or after vfirst (compares with the tail).
    // FULL: every scanline of the wave is active at step t and every lane
    // stores (no tail warm-up), so no predicate depends on t.
    auto step = [&](PairBuf<K> &b, TailBuf<K> &tb, int t, auto verify, auto full) {
        constexpr bool VERIFY = decltype(verify)::value;
        constexpr bool FULL = decltype(full)::value;
        const bool act0 = FULL || t < len[0], act1 = FULL || t < len[1];
        // Costs of both scanlines, packed: byte i of each slice segment into
        // the low (first scanline) and high (second) half.
        uint32_t C[K];
#pragma unroll
        for (int i = 0; i < K; ++i)
            C[i] = perm(b.c[1][i / 4], b.c[0][i / 4],
                        0x0C000C00u | ((4u + (i & 3)) << 16) | (uint32_t)(i & 3));

        // With prev = 0 and pm = 0 before the first step, the recurrence
        // below gives agg = C there (best = min(0, ...) = 0), the first
        // pixel's copy of the cost (SGMStereo.metal:152), so no step is
        // special-cased.
        // The chain runs on q[i] = agg[i] + P1 (the d-1 term of d + 1):
        // q[i] = C[i] + P1 - pm + min(A[i], q[i-1]).  The cost volume holds
        // C + bias, so B[i] = that cost - (pm - (P1 - bias)) is C - pm + P1
        // with one subtraction per entry; agg = q - P1 leaves the chain.
        uint32_t agg[K], q[K];
        {
            const uint32_t o3 = pk_add(pm, p2x2);
            const uint32_t pmx = pk_sub(pm, xqx2);
            const uint32_t nbq = from_upper_lane<L>(prevq[0]);
            uint32_t A[K], B[K];
#pragma unroll
            for (int i = 0; i < K; ++i) {
                uint32_t o2 = (i + 1 < K) ? prevq[(i + 1) % K] : nbq;
                if (!ALIGNED || i == K - 1) o2 |= om[i];
                A[i] = pk_min(pk_min(prev[i], o2), o3);
                B[i] = pk_sub(C[i], pmx);
            }
            // Speculative pass: no carry-in from the lane below.
            q[0] = pk_add(B[0], A[0]);
#pragma unroll
            for (int i = 1; i < K; ++i) q[i] = pk_add(B[i], pk_min(A[i], q[i - 1]));
            // Fixed-point rounds over the carry-in.  Halves whose scanline has
            // ended, the group's first lane and padding lanes take no carry.
            const uint32_t take = (sub > 0 && lane_real)
                                      ? ((act0 ? 0x0000FFFFu : 0u) | (act1 ? 0xFFFF0000u : 0u))
                                      : 0u;
            uint32_t e_used = A[0];
            auto carry = [&]() {
                const uint32_t cand = pk_min(A[0], from_lower_lane<L>(q[K - 1]));
                return (cand & take) | (e_used & ~take);
            };
            uint32_t e = carry();
            // Usually no lane's carry changes: that path falls straight through,
            // and the rounds (at most L - 1 recomputations) sit out of line.
            if (__builtin_expect(lanes_ne(e, e_used) != 0, 0)) {
                for (int r = 0; r < L - 1; ++r) {
                    if (e != e_used) {
                        e_used = e;
                        q[0] = pk_add(B[0], e);
#pragma unroll
                        for (int i = 1; i < K; ++i) q[i] = pk_add(B[i], pk_min(A[i], q[i - 1]));
                    }
                    e = carry();
                    if (lanes_ne(e, e_used) == 0) break;
                }
            }
#pragma unroll
            for (int i = 0; i < K; ++i) agg[i] = pk_sub(q[i], p1x2);
        }

        uint32_t m = 0xFFFFFFFFu;
#pragma unroll
        for (int i = 0; i < K; ++i) m = pk_min(m, ALIGNED ? agg[i] : (agg[i] | pad[i]));
        if (ALIGNED) m |= lane_pad;
        pm = group_min_pk<L>(m);
        // The previous step's high-byte slices, if any (see flush_high),
        // before prev is overwritten.
        if (mode == kOutU8 && __builtin_expect(high_lanes != 0, 0)) flush_high(t - 1);
#pragma unroll
        for (int i = 0; i < K; ++i) {
            prev[i] = agg[i];
            prevq[i] = q[i];
        }

        bool st0 = FULL || (act0 && lane_store), st1 = FULL || (act1 && lane_store);
        if constexpr (SPLIT && !FULL) {
            // Tail warm-up steps store nothing (sfrom is 0 for other roles).
            st0 = st0 && t >= sfrom[0];
            st1 = st1 && t >= sfrom[1];
        }
        uint32_t clear_flag = 0;   // head role: u8 flags this step rewrites (bit per scanline)
        if (VERIFY && t >= vfirst) {
            // Compare this step's state with the tail's stored one, scanline
            // by scanline over the whole group (entries d < D): from the
            // first step where they agree, the tail's values stand.
#pragma unroll
            for (int hh = 0; hh < 2; ++hh) {
                const bool ver = t >= vfrom[hh] && t < len[hh] && !met[hh];
                uint32_t diff = 0;
                if (mode == kOutU16) {
#pragma unroll
                    for (int i = 0; i < K / 2; ++i) {
                        const uint32_t mine = perm(agg[2 * i + 1], agg[2 * i],
                                                   hh == 0 ? 0x05040100u : 0x07060302u);
                        diff |= (mine ^ tb.v[hh][i]) & realmask[i];
                    }
                } else {
                    uint32_t lo[(K + 3) / 4], hi[(K + 3) / 4], lo2[(K + 3) / 4], hi2[(K + 3) / 4];
                    bytes_of_pairs<K>(agg, 0x06020400u, lo, lo2);
                    bytes_of_pairs<K>(agg, 0x07030501u, hi, hi2);
#pragma unroll
                    for (int i = 0; i < (K + 3) / 4; ++i) {
                        const uint32_t l = hh == 0 ? lo[i] : lo2[i];
                        const uint32_t hb = hh == 0 ? hi[i] : hi2[i];
                        diff |= ((l ^ tb.v[hh][i]) | hb) & realmask[i];
                    }
                    diff |= tb.flag[hh];
                }
                const uint32_t g_diff = group_or<L>(lane_real && ver && diff != 0 ? 1u : 0u);
                if (ver && g_diff == 0) met[hh] = true;
            }
            // Scanlines that have met store nothing more; the others
            // overwrite the tail's values, and (u8) their flag byte, which the
            // high-byte pass of the next step sets again if needed.
            st0 = st0 && !met[0];
            st1 = st1 && !met[1];
            clear_flag = (st0 && t >= vfrom[0] ? 1u : 0u) | (st1 && t >= vfrom[1] ? 2u : 0u);
        }
        if (mode == kOutU16) {
            // Halves to pixel order: [a_i | a_i+1] and [b_i | b_i+1].
            uint32_t wa[K / 2], wb[K / 2];
#pragma unroll
            for (int i = 0; i < K / 2; ++i) {
                wa[i] = perm(agg[2 * i + 1], agg[2 * i], 0x05040100u);
                wb[i] = perm(agg[2 * i + 1], agg[2 * i], 0x07060302u);
            }
            store_seg<K * 2>(st0 ? vp[0] : store_sink<uint8_t>(), wa);
            store_seg<K * 2>(st1 ? vp[1] : store_sink<uint8_t>(), wb);
        } else {
            uint32_t wa[(K + 3) / 4], wb[(K + 3) / 4];
            bytes_of_pairs<K>(agg, 0x06020400u, wa, wb);
            store_seg<K>(st0 ? vp[0] : store_sink<uint8_t>(), wa);
            store_seg<K>(st1 ? vp[1] : store_sink<uint8_t>(), wb);
            if (SPLIT && clear_flag != 0 && sub == 0) {
#pragma unroll
                for (int hh = 0; hh < 2; ++hh)
                    if (clear_flag & (1u << hh))
                        p.flags[(size_t)(pix0[hh] + t * pstep) * 8 + p.dir] = 0;
            }
            // High bytes of this step, tested (and stored) one step later.
            uint32_t o = 0;
#pragma unroll
            for (int i = 0; i < K; ++i) o |= agg[i];
            high = o & ((st0 ? 0x0000FF00u : 0u) | (st1 ? 0xFF000000u : 0u));
            high_lanes = lanes_ne(high, 0u);
        }
        vp[0] += vstep;
        vp[1] += vstep;
        ++tcur;
    };

    // Loads run PF steps ahead through a ring of PF + 1 buffers; the loop is
    // unrolled by the ring size so every buffer stays in fixed registers, and
    // runs a whole number of rounds (the steps past maxlen are inactive on
    // every lane: their loads read in-range words, their stores go to the
    // sink).  With no branch around any load or store, every step issues the
    // same memory operations and waits only for its own PF-old loads.
    // Split roles run the same loop: a tail publishes its early range once
    // (at round publish_round), a head leaves it at the round holding vfirst
    // and continues in a second loop that also loads and checks the tail's
    // values; it ends as soon as every scanline of the wave has met its tail
    // or ended (checked once per round).
    auto publish = [&]() {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_store(tail_early, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    };
    const int rounds = __builtin_amdgcn_readfirstlane((maxlen + NB - 1) / NB);
    // A tail's values of step kSplitWarm + kRange - 1 are flushed (high bytes
    // included) once step kSplitWarm + kRange has run, at the end of this
    // round: every step it publishes early is then in memory before the flag.
    // A tail too short for that publishes after its last flush instead.
    const int publish_round =
        SPLIT && role == kRoleTail && (kSplitWarm + kRange) / NB < rounds - 1
            ? (kSplitWarm + kRange) / NB
            : -1;
    const int rounds_main =
        SPLIT && role == kRoleHead ? min(rounds, __builtin_amdgcn_readfirstlane(vfirst) / NB)
                                   : rounds;
    // Rounds whose steps and prefetches all lie inside every scanline of the
    // wave run the FULL step (every lane of the wave stores when L * K <= Dp).
    // A tail's warm-up stores nothing, so its FULL rounds start after the
    // warm-up (every split scanline of a tail wave has the same warm-up, and
    // a wave holding a scanline too short to split has no FULL rounds).
    const bool full_ok = L * K <= Dp;
    const int full_from = SPLIT && role == kRoleTail ? (kSplitWarm + NB - 1) / NB : 0;
    const int full_rounds =
        full_ok ? min(rounds_main, max(0, (__builtin_amdgcn_readfirstlane(minlen) - PF) / NB)) : 0;
    PairBuf<K> buf[NB];
    TailBuf<K> tbuf[SPLIT ? NB : 1];
#pragma unroll
    for (int j = 0; j < PF; ++j) {
        issue(buf[j], j, std::false_type{});
        __builtin_amdgcn_sched_barrier(0);
    }
    int r = 0;
    if constexpr (SPLIT) {
        // A tail's warm-up rounds (predicated stores).
        const int warm = min(full_from, rounds_main);
        for (; r < warm; ++r) {
            const int t = r * NB;
#pragma unroll
            for (int u = 0; u < NB; ++u) {
                issue(buf[(u + PF) % NB], t + u + PF, std::false_type{});
                step(buf[u], tbuf[0], t + u, std::false_type{}, std::false_type{});
            }
            if (r == publish_round) publish();
        }
    }
#pragma unroll
    for (int hh = 0; hh < 2; ++hh)
        fptr[hh] = cost + (ptrdiff_t)(pix0[hh] + (r * NB + PF) * pstep) * Dp;
    for (; r < full_rounds; ++r) {
        const int t = r * NB;
#pragma unroll
        for (int u = 0; u < NB; ++u) {
            issue(buf[(u + PF) % NB], t + u + PF, std::true_type{});
            step(buf[u], tbuf[0], t + u, std::false_type{}, std::true_type{});
        }
        if (SPLIT && r == publish_round) publish();
    }
    for (; r < rounds_main; ++r) {
        const int t = r * NB;
#pragma unroll
        for (int u = 0; u < NB; ++u) {
            issue(buf[(u + PF) % NB], t + u + PF, std::false_type{});
            step(buf[u], tbuf[0], t + u, std::false_type{}, std::false_type{});
        }
        if (SPLIT && r == publish_round) publish();
    }
    if constexpr (SPLIT) {
        if (role == kRoleHead && r < rounds) {
            // The tail's values of the steps whose costs the ring holds.
#pragma unroll
            for (int j = 0; j < PF; ++j) issue_tail(tbuf[j], r * NB + j);
            for (; r < rounds; ++r) {
                const int t = r * NB;
#pragma unroll
                for (int u = 0; u < NB; ++u) {
                    issue(buf[(u + PF) % NB], t + u + PF, std::false_type{});
                    issue_tail(tbuf[(u + PF) % NB], t + u + PF);
                    step(buf[u], tbuf[u], t + u, std::true_type{}, std::false_type{});
                }
                const int tn = t + NB;
                const bool live = (tn < len[0] && !met[0]) || (tn < len[1] && !met[1]);
                if (tn > vfirst && __builtin_amdgcn_uicmp((uint32_t)live, 0u, 33) == 0) {
                    ++r;
                    break;
                }
            }
        }
    }
    if (mode == kOutU8) flush_high(r * NB - 1);
    if (SPLIT && role == kRoleTail && publish_round < 0) publish();
    // Loads still in flight (the ring runs PF steps ahead) complete before
    // the wave ends.
    __builtin_amdgcn_s_waitcnt(0);
}

// Winner-take-all with parabola refinement of one pixel from its summed
// slice, 16 lanes per pixel: in chunk c lane j holds entries
// 128c + 8j .. 128c + 8j + 7 in v[c][0..3] (pairs in disparity order).  The
// first strict minimum from 0xFFFF (SGMStereo.metal:185-199) is the smallest
// key (value << 16 | d) over d < D; a slice of all 0xFFFF keeps d = 0.
__device__ __forceinline__ void wta_slice(const u16x2 (&v)[2][4], int nc, int j, int D,
                                          bool valid, int pix, float *__restrict__ out,
                                          int16_t *__restrict__ best) {
    uint32_t key = 0xFFFFFFFFu;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        if (c >= nc) break;
        const int d0 = 128 * c + 8 * j;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const uint32_t val = v[c][e / 2][e & 1];
            if (d0 + e < D) key = min(key, (val << 16) | (uint32_t)(d0 + e));
        }
    }
    key = group_min<16>(key);
    const int bd = (int)(key & 0xFFFFu);
    const uint32_t bc = key >> 16;
    uint32_t nb2 = 0;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        if (c >= nc) break;
        const int d0 = 128 * c + 8 * j;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const uint32_t val = v[c][e / 2][e & 1];
            if (d0 + e == bd - 1) nb2 |= val << 16;
            if (d0 + e == bd + 1) nb2 |= val;
        }
    }
    nb2 = group_or<16>(nb2);
    if (valid && j == 0) {
        // Parabola through the neighbours (SGMStereo.metal:201-211), IEEE fp32.
        const int bdw = bc == 0xFFFFu ? 0 : bd;
        float disp = (float)bdw;
        if (bdw > 0 && bdw < D - 1) {
            const float c0 = (float)(nb2 >> 16);
            const float c1 = (float)bc;
            const float c2 = (float)(nb2 & 0xFFFFu);
            const float den = 2.0f * (c0 + c2 - 2.0f * c1);
            if (fabsf(den) > 1e-6f) disp = disp + (c0 - c2) / den;
        }
        if (out) out[pix] = disp;
        if (best) best[pix] = (int16_t)bdw;
    }
}

// Raw loads of one pixel's direction slices for the summing winner-take-all:
// q[c][k] holds volume k's entries 128c + 8j .. +7 of the pixel (E = uint4,
// 16 bytes, for u16 entries; E = uint2, 8 bytes, for u8), fl the pixel's
// exception flags (u8 only).  Split from the sums so that a caller can keep
// the loads of the next pixels in flight while it sums the current ones.
typedef uint32_t v2u __attribute__((ext_vector_type(2)));
typedef uint32_t v4u __attribute__((ext_vector_type(4)));

template <int MODE>
struct WtaIn {
    using E = typename std::conditional<MODE == kOutU16, uint4, uint2>::type;
    uint2 fl;
    E q[2][8];
};

// Loads of pixel `pix` (16 lanes per pixel, j = lane & 15; nc = 1 for
// Dp <= 128, else 2; volume k at vol0 + k * vstride bytes, entries
// pix*Dp + d).
template <int MODE>
__device__ __forceinline__ void wta_load(WtaIn<MODE> &in, const uint8_t *__restrict__ vol0,
                                         size_t vstride, const uint2 *__restrict__ flags,
                                         int nv, int nc, int pix, int npx, int Dp) {
    using E = typename WtaIn<MODE>::E;
    const int j = (int)(threadIdx.x & 15);
    const bool valid = pix < npx;
    const size_t base = (size_t)(valid ? pix : 0) * (size_t)Dp;
    in.fl = MODE == kOutU8 ? flags[valid ? pix : 0] : make_uint2(0, 0);
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        if (c >= nc) break;
        const int d0 = 128 * c + 8 * j;
        if (d0 >= Dp) continue;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            if (k >= nv) break;
            // Read once: non-temporal loads (measured 25% faster here).
            const E *src = (const E *)(vol0 + k * vstride + (base + d0) * sizeof(E) / 8);
            if constexpr (sizeof(E) == 8) {
                const v2u v = __builtin_nontemporal_load((const v2u *)src);
                in.q[c][k] = E{v.x, v.y};
            } else {
                const v4u v = __builtin_nontemporal_load((const v4u *)src);
                in.q[c][k] = E{v.x, v.y, v.z, v.w};
            }
        }
    }
}

// Sums the loaded slices mod 2^16.
//   kOutU16: u16 entries added as u16 pairs.
//   kOutU8:  the byte sums of at most 8 volumes stay below 2^16, so even and
//            odd bytes are summed as two u16 pairs without carries; slices
//            flagged in the pixel's 8 flag bytes (byte k by direction k) add
//            256 * their high bytes (hi0 + k * vstride, same entry offsets).
// v[c][q] receives entries 128c + 8j + 2q and + 1 (pairs in disparity order).
template <int MODE>
__device__ __forceinline__ void wta_sums(const WtaIn<MODE> &in, const uint8_t *__restrict__ hi0,
                                         size_t vstride, int nv, int nc, int pix, int npx,
                                         int Dp, u16x2 (&v)[2][4]) {
    const int j = (int)(threadIdx.x & 15);
    const bool valid = pix < npx;
    const size_t base = (size_t)(valid ? pix : 0) * (size_t)Dp;
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int q = 0; q < 4; ++q) v[c][q] = (u16x2){0, 0};
    if constexpr (MODE == kOutU16) {
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            if (c >= nc) break;
            if (128 * c + 8 * j >= Dp) continue;
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                if (k >= nv) break;
                const uint4 x = in.q[c][k];
                v[c][0] += __builtin_bit_cast(u16x2, x.x);
                v[c][1] += __builtin_bit_cast(u16x2, x.y);
                v[c][2] += __builtin_bit_cast(u16x2, x.z);
                v[c][3] += __builtin_bit_cast(u16x2, x.w);
            }
        }
    } else {
        const uint2 fl = in.fl;
        const bool flagged = (fl.x | fl.y) != 0;
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            if (c >= nc) break;
            const int d0 = 128 * c + 8 * j;
            uint32_t e0 = 0, o0 = 0, e1 = 0, o1 = 0;
            if (d0 < Dp) {
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    if (k >= nv) break;
                    const uint2 x = in.q[c][k];
                    e0 += x.x & 0x00FF00FFu;
                    o0 += perm(0u, x.x, 0x0c030c01u);
                    e1 += x.y & 0x00FF00FFu;
                    o1 += perm(0u, x.y, 0x0c030c01u);
                }
                if (flagged) {
#pragma unroll
                    for (int k = 0; k < 8; ++k) {
                        if (k >= nv) break;
                        const uint32_t fb = k < 4 ? (fl.x >> (8 * k)) : (fl.y >> (8 * (k - 4)));
                        if (fb & 0xFFu) {
                            const uint2 x = *(const uint2 *)(hi0 + k * vstride + base + d0);
                            e0 = pk_add(e0, (x.x & 0x00FF00FFu) << 8);
                            o0 = pk_add(o0, perm(0u, x.x, 0x030c010cu));
                            e1 = pk_add(e1, (x.y & 0x00FF00FFu) << 8);
                            o1 = pk_add(o1, perm(0u, x.y, 0x030c010cu));
                        }
                    }
                }
            }
            // Entries d0..d0+7 in disparity order: [e0.lo o0.lo] [e0.hi o0.hi] ...
            v[c][0] = __builtin_bit_cast(u16x2, perm(o0, e0, 0x05040100u));
            v[c][1] = __builtin_bit_cast(u16x2, perm(o0, e0, 0x07060302u));
            v[c][2] = __builtin_bit_cast(u16x2, perm(o1, e1, 0x05040100u));
            v[c][3] = __builtin_bit_cast(u16x2, perm(o1, e1, 0x07060302u));
        }
    }
}

// wta_sums, then the pixel's winner; with sum_out the summed u16 slice is
// also stored (pixel-major, stride Dp).  All lanes of the wave must call it
// (the winner uses cross-lane moves).
template <int MODE>
__device__ __forceinline__ void wta_finish(const WtaIn<MODE> &in,
                                           const uint8_t *__restrict__ hi0, size_t vstride,
                                           int nv, int nc, int pix, int npx, int D, int Dp,
                                           float *__restrict__ out, int16_t *__restrict__ best,
                                           uint16_t *__restrict__ sum_out) {
    const int j = (int)(threadIdx.x & 15);
    const bool valid = pix < npx;
    const size_t base = (size_t)(valid ? pix : 0) * (size_t)Dp;
    u16x2 v[2][4];
    wta_sums(in, hi0, vstride, nv, nc, pix, npx, Dp, v);
    if (sum_out != nullptr && valid) {
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            if (c >= nc) break;
            const int d0 = 128 * c + 8 * j;
            if (d0 < Dp)
                *(uint4 *)(sum_out + base + d0) =
                    make_uint4(__builtin_bit_cast(uint32_t, v[c][0]),
                               __builtin_bit_cast(uint32_t, v[c][1]),
                               __builtin_bit_cast(uint32_t, v[c][2]),
                               __builtin_bit_cast(uint32_t, v[c][3]));
        }
    }
    wta_slice(v, nc, j, D, valid, pix, out, best);
}

// Sum of the nv direction slices of pixel `pix` and its winner (wta_load
// then wta_finish), in the volume storage `mode`.
__device__ __forceinline__ void wta_sum_pixel(const uint8_t *__restrict__ vol0, size_t vstride,
                                              const uint8_t *__restrict__ hi0,
                                              const uint2 *__restrict__ flags, int nv, int nc,
                                              int mode, int pix, int npx, int D, int Dp,
                                              float *__restrict__ out,
                                              int16_t *__restrict__ best,
                                              uint16_t *__restrict__ sum_out) {
    if (mode == kOutU16) {
        WtaIn<kOutU16> in;
        wta_load(in, vol0, vstride, flags, nv, nc, pix, npx, Dp);
        wta_finish(in, hi0, vstride, nv, nc, pix, npx, D, Dp, out, best, sum_out);
    } else {
        WtaIn<kOutU8> in;
        wta_load(in, vol0, vstride, flags, nv, nc, pix, npx, Dp);
        wta_finish(in, hi0, vstride, nv, nc, pix, npx, D, Dp, out, best, sum_out);
    }
}

// Left-right check, right view (the check as the build defines it; the
// reference has no right view): D_R(xr) is the lowest k < D with xr + k < w
// minimising S(xr + k, y, k), taken strictly below 0xFFFF, and 0 when no
// entry is below it.  Entry (x, d) of a summed slice is entry k = d of right
// pixel x - d, so every pixel offers each of its entries with S < 0xFFFF and
// d <= x as the inverted key ~(S << 16 | d) to right pixel x - d; the largest
// inverted key is the smallest S and, among equal S, the smallest d.  Keys
// start at 0 (no offer).
__device__ __forceinline__ uint32_t lr_right_of_key(uint32_t key) {
    return key ? (~key) & 0xFFFFu : 0u;
}

// The offers of one pixel x (its slice in v, 16 lanes, as wta_slice) into
// lkey[x - d - xb] (LDS words; atomic max).
__device__ __forceinline__ void wta_offer_right(const u16x2 (&v)[2][4], int nc, int j, int D,
                                                int x, int xb, uint32_t *lkey) {
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        if (c >= nc) break;
        const int d0 = 128 * c + 8 * j;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int d = d0 + e;
            const uint32_t val = v[c][e / 2][e & 1];
            if (d < D && d <= x && val != 0xFFFFu)
                atomicMax(&lkey[x - d - xb], ~((val << 16) | (uint32_t)d));
        }
    }
}

// Sum + winner of the TP pixels of row y from x0 (16 lanes each, as
// wta_sum_pixel; the block has 16 TP threads), with their right-view offers
// gathered in lkey (TP + 255 LDS words: right pixels x0 - D + 1 .. x0 + TP - 1)
// and folded into the row's keys rkey[y * w + xr] with one atomic max per
// right pixel offered to, (TP + D - 1) / TP per pixel.  Taking a wider tile
// 16 pixels at a time instead keeps fewer loads in flight per CU, and the
// pass is bound by its loads.  Block-level: every thread must call it.
template <int MODE, int TP>
__device__ __forceinline__ void wta_lr_tile(const uint8_t *__restrict__ vol0, size_t vstride,
                                            const uint8_t *__restrict__ hi0,
                                            const uint2 *__restrict__ flags, int nv, int nc,
                                            int w, int h, int y, int x0, int D, int Dp,
                                            float *__restrict__ out, int16_t *__restrict__ best,
                                            uint32_t *__restrict__ rkey, uint32_t *lkey) {
    constexpr int kThreads = 16 * TP;
    const int t = (int)threadIdx.x, j = t & 15;
    const int xb = x0 - (D - 1);
    const int nk = TP + D - 1;
    const int npx = w * h;
    const int x = x0 + (t >> 4);
    const int pix = x < w ? y * w + x : npx;
    WtaIn<MODE> in;
    wta_load(in, vol0, vstride, flags, nv, nc, pix, npx, Dp);
    for (int i = t; i < nk; i += kThreads) lkey[i] = 0;
    __syncthreads();
    u16x2 v[2][4];
    wta_sums(in, hi0, vstride, nv, nc, pix, npx, Dp, v);
    if (x < w) wta_offer_right(v, nc, j, D, x, xb, lkey);
    wta_slice(v, nc, j, D, x < w, pix, out, best);
    __syncthreads();
    uint32_t *row = rkey + (size_t)y * w;
    for (int i = t; i < nk; i += kThreads) {
        const int xr = xb + i;
        const uint32_t k = lkey[i];
        if (k != 0 && xr >= 0 && xr < w) atomicMax(&row[xr], k);
    }
}

}  // namespace dev
}  // namespace vxh
