// check_core.h -- the check-node update on one word (4 frames) in packed-16
// byte lanes, shared by the per-batch and the persistent decode kernels.
//
// What is computed (per frame, the reference's serial table order;
// src/decoder_oms/cuda/CUDA_OMS_SIMD.cu:148-191):
//     t_j  = sat8(post[v_j] - msg_j)            (first iteration: post[v_j])
//     min1, min2 = two smallest |t_j| (unsigned, |-128| = 128, both start at 127)
//     msg_j = +-(|t_j| == min1 ? f(min2) : g(min1)),  sign = parity of the others' t > 0
//     post[v_j] = sat8(t_j + msg_j)
#pragma once
#include "byte_lanes.h"
#include "kernels.h"

namespace ldpc {

// learned parameter of one frame in the form magnitudes() takes
__device__ __forceinline__ uint32_t learned_half(int variant, int v) {
    return variant == VAR_LNMS ? ((uint32_t)v << 5) : ((uint32_t)v << 8);
}

// ---------------------------------------------------------------------------
// per-check magnitudes  (cste_1 from min2, cste_2 from min1)
// ---------------------------------------------------------------------------
// lp: learned parameter of each frame of the half (LOMS: beta << 8, LNMS:
// alpha << 5), so frames of one word may sit at different iterations.
__device__ __forceinline__ void magnitudes(int variant, u16x2 min1, u16x2 min2, bool nocap, u16x2 lp,
                                           u16x2 &k1, u16x2 &k2) {
    const u16x2 cap31 = splat(0x1F00u);
    switch (variant) {
    case VAR_MS:  // CUDA_MS_SIMD.cu:173-174
        k1 = umin(min2, cap31);
        k2 = umin(min1, cap31);
        break;
    case VAR_OMS:      // CUDA_OMS_SIMD.cu:173-174; no cap: :115-116
    case VAR_OMS_V2: {  // CUDA_OMS_SIMD_v2.cu:145-146
        const u16x2 cap = nocap ? splat(0x7F00u) : cap31;
        const u16x2 one = splat(0x0100u);
        k1 = umin(sub_sat_u(min2, one), cap);
        k2 = umin(sub_sat_u(min1, one), cap);
        break;
    }
    case VAR_LOMS:  // learned offset beta (LSB), cap 31
        k1 = umin(sub_sat_u(min2, lp), cap31);
        k2 = umin(sub_sat_u(min1, lp), cap31);
        break;
    case VAR_NMS: {  // (char)(0.75f*m) = floor(3m/4): m<<8 - m<<6 keeps 3m/4 exactly
        const u16x2 hi = splat(0xFF00u);
        k1 = (min2 - (min2 >> (u16x2)(2))) & hi;
        k2 = (min1 - (min1 >> (u16x2)(2))) & hi;
        break;
    }
    case VAR_2NMS: {  // 0.875 on min2, 0.75 on min1 (CUDA_2NMS_SIMD.cu:73-85)
        const u16x2 hi = splat(0xFF00u);
        k1 = (min2 - (min2 >> (u16x2)(3))) & hi;
        k2 = (min1 - (min1 >> (u16x2)(2))) & hi;
        break;
    }
    default: {  // VAR_LNMS: min(floor(m * alpha / 8), 31), alpha in [0, 8]
        const u16x2 hi = splat(0xFF00u);
        k1 = umin(((min2 >> (u16x2)(8)) * lp) & hi, cap31);
        k2 = umin(((min1 >> (u16x2)(8)) * lp) & hi, cap31);
        break;
    }
    }
}

// ---------------------------------------------------------------------------
// one check of degree D on one packed-16 half (2 frames, each in the high
// byte of a 16-bit lane; low bytes zero on input).
// x: posterior halves, m: message halves (ignored when first).
// Outputs: pn = sat(t + msg) and mn = msg, both with clean low bytes where
// consumers need them (repack reads only high bytes).
// EQ = false: r = k1 - min(a - min1, k1 - k2).  Valid when k1 - k2 <= min2 - min1
//   (every 1-Lipschitz magnitude map: MS, OMS, NMS, learned tables): a lane
//   that is not the minimum has a >= min2.  The residue 0x7FFF (a saturated
//   +127, or |-128|) only occurs with a >= 127 >= min2, so it lands on k2 as
//   the reference's lane does, or on k1 == k2 when min1 == min2 == 127.
// EQ = true (2NMS, k1 - k2 can exceed min2 - min1): t is cleaned to its high
//   byte and the test a == min1 is made explicit.
// Returns the parity-of-positives bits at 0x80008000 (bit 15: the half's
// first frame, bit 31: its second).
// ---------------------------------------------------------------------------
template <int D, bool EQ>
__device__ __forceinline__ uint32_t check_half(const u16x2 (&x)[D], const u16x2 (&m)[D], bool first, int variant,
                                               bool nocap, u16x2 lp, u16x2 (&pn)[D], u16x2 (&mn)[D]) {
    // KEEP: |t| and the sign of t stay live across the two passes (D <= 12);
    // larger checks recompute them in the second pass to bound registers.
    // With KEEP the sign is kept as a mask and the message's sign applied by
    // a multiply by +-1; without it, as the top bit and a conditional negate
    // (the mask form costs one op more per edge when recomputed).
    constexpr bool KEEP = D <= 12;
    u16x2 t[D], ng[KEEP ? D : 1], ab[KEEP ? D : 1];
    u16x2 m1 = splat(0x7F00u), m2 = m1;
    u16x2 sg = splat(0u), yp = splat(0u);
    const u16x2 zero = splat(0u);
#pragma unroll
    for (int j = 0; j < D; j++) {
        u16x2 v = x[j];
        if (!first) v = sub_sat_s(v, m[j]);
        if (EQ) v &= splat(0xFF00u);
        t[j] = v;
        u16x2 y = sub_sat_s(zero, v);  // -t, top bit = (t > 0)
        const u16x2 a = umin(v, y);    // |t| as uint
        if constexpr (KEEP) y = sign_mask(y);  // 0xFFFF where t > 0
        // parity of the positives (top bit; with KEEP every bit): two edges
        // per 3-input XOR (v_bitop3 0x96)
        if (j & 1)
            sg = u2(__builtin_amdgcn_bitop3_b32(w32(sg), w32(yp), w32(y), 0x96));
        else if (j == D - 1)
            sg ^= y;
        yp = y;
        if (j == 0) {
            // min2 = min(127, max(a, 127)) = 127 after the first edge
            m1 = umin(m1, a);
        } else {
            m2 = umin(m2, umax(a, m1));
            m1 = umin(m1, a);
        }
        if constexpr (KEEP) {
            ng[j] = y;
            ab[j] = a;
        }
    }
    const uint32_t parity = w32(sg) & 0x80008000u;

    u16x2 k1, k2;
    magnitudes(variant, m1, m2, nocap, lp, k1, k2);
    const u16x2 K = k1 - k2;
    // k1 - min(a - min1, k1 - k2) == max(sat_u(k1 + min1 - a), k2): one op
    // less per edge (k1 + min1 <= 0xFE00 cannot wrap)
    const u16x2 C1 = k1 + m1;
    if constexpr (!KEEP) {
        sg ^= splat(0x8000u);  // top bit now: parity of the others is even -> negate
        // opaque, or the compiler re-associates (sg ^ c) ^ y into (sg ^ y) ^ c:
        // one XOR more per edge
        uint32_t sw = w32(sg);
        __asm__("" : "+v"(sw));
        sg = u2(sw);
    }

#pragma unroll
    for (int j = 0; j < D; j++) {
        u16x2 y, a;
        if constexpr (KEEP) {
            y = ng[j];
            a = ab[j];
        } else {
            y = sub_sat_s(zero, t[j]);
            a = umin(t[j], y);
        }
        u16x2 r;
        if (EQ) {
            const u16x2 d = a - m1;  // >= 0, zero exactly on the minimum
            r = k2 + sub_sat_u(splat(0x0001u), d) * K;
        } else {
            r = umax(sub_sat_u(C1, a), k2);
        }
        // r has a clean low byte (so has -r)
        if constexpr (KEEP) {
            // negative where the parity of the others' positives is even:
            // s = ~(sg ^ y) | 1 is -1 there and +1 elsewhere (one v_bitop3,
            // truth table 0xEB), msg = r * s (one v_pk_mul_lo_u16)
            const u16x2 s = u2(__builtin_amdgcn_bitop3_b32(w32(sg), w32(y), 0x00010001u, 0xEB));
            mn[j] = r * s;
        } else {
            const u16x2 sj = sign_mask(sg ^ y);
            mn[j] = (r ^ sj) - sj;
        }
        pn[j] = add_sat_s(t[j], mn[j]);
    }
    return parity;
}

// How a word's message bytes are unpacked: the even half by v_perm with
// `sel_even`, the odd half by AND with `mask_odd`.  A byte lane whose frame
// reads zero messages (a frame that starts now in the frame-queue decoder)
// gets the zero selector 0x0C in the even half and a zero mask byte in the
// odd half, so clearing it costs nothing beyond the unpack itself.
struct MsgUnpack {
    uint32_t sel_even = 0x020C000Cu, mask_odd = 0xFF00FF00u;
};
// keep: 0xFF for each byte lane whose messages are read, 0x00 for a lane
// whose messages count as zero
__device__ __forceinline__ MsgUnpack msg_unpack(uint32_t keep) {
    MsgUnpack u;
    if (!(keep & 0x000000FFu)) u.sel_even = (u.sel_even & 0xFFFF00FFu) | 0x00000C00u;
    if (!(keep & 0x00FF0000u)) u.sel_even = (u.sel_even & 0x00FFFFFFu) | 0x0C000000u;
    u.mask_odd = keep & 0xFF00FF00u;
    return u;
}

// one check of degree D on one word (4 frames = even + odd half).
// Returns the parity bits of the even (x) and odd (y) halves at 0x80008000:
// frame 0 = x bit 15, 1 = y bit 15, 2 = x bit 31, 3 = y bit 31.
template <int D, bool EQ>
__device__ __forceinline__ uint2 check_core(const uint32_t (&pw)[D], const uint32_t (&mw)[D], bool first,
                                            int variant, bool nocap, u16x2 lpe, u16x2 lpo, uint32_t (&pn)[D],
                                            uint32_t (&mn)[D], MsgUnpack mu = MsgUnpack()) {
    u16x2 xe[D], xo[D], me[D], mo[D], pe[D], po[D], ne[D], no[D];
#pragma unroll
    for (int j = 0; j < D; j++) {
        xe[j] = unpack_even(pw[j]);
        xo[j] = unpack_odd(pw[j]);
        if (!first) {
            me[j] = u2(__builtin_amdgcn_perm(mw[j], mw[j], mu.sel_even));
            mo[j] = u2(mw[j] & mu.mask_odd);
        }
    }
    const uint32_t pe_bits = check_half<D, EQ>(xe, me, first, variant, nocap, lpe, pe, ne);
    const uint32_t po_bits = check_half<D, EQ>(xo, mo, first, variant, nocap, lpo, po, no);
#pragma unroll
    for (int j = 0; j < D; j++) {
        mn[j] = repack(ne[j], no[j]);
        pn[j] = repack(pe[j], po[j]);
    }
    return make_uint2(pe_bits, po_bits);
}

// Load / store the D message words of one lane (8-byte aligned runs when D is even).
template <int D>
__device__ __forceinline__ void load_msgs(const uint32_t *__restrict__ mp, uint32_t (&mw)[D]) {
    if constexpr (D % 2 == 0) {
        const uint2 *q = reinterpret_cast<const uint2 *>(mp);
#pragma unroll
        for (int h = 0; h < D / 2; h++) {
            const uint2 x = q[h];
            mw[2 * h] = x.x;
            mw[2 * h + 1] = x.y;
        }
    } else {
#pragma unroll
        for (int j = 0; j < D; j++) mw[j] = mp[j];
    }
}
template <int D>
__device__ __forceinline__ void store_msgs(uint32_t *__restrict__ mp, const uint32_t (&mn)[D]) {
    if constexpr (D % 2 == 0) {
        uint2 *q = reinterpret_cast<uint2 *>(mp);
#pragma unroll
        for (int h = 0; h < D / 2; h++) q[h] = make_uint2(mn[2 * h], mn[2 * h + 1]);
    } else {
#pragma unroll
        for (int j = 0; j < D; j++) mp[j] = mn[j];
    }
}

// ---------------------------------------------------------------------------
// One level's items of one thread for a code whose checks all have degree D:
// items c_first, c_first + cstride, ... < le of word w.  The check's edges
// start at c*D, its LDS byte offsets (v * P * 4) come packed two per dword
// (edge_off16), and the index and message loads of G items are issued
// together before any of them is computed: full batches of G while G items
// remain, then single items, so no load or update is predicated inside a
// batch.  FIRST: no message read (the decode's first iteration).  keep: byte
// mask of the message lanes read (a lane whose frame starts now reads zero
// messages; folded into the unpack).  VK >= 0: the variant known at compile
// time (its magnitude map is then straight-line code with no per-item
// branch), -1: read from `variant`.  Returns the OR of the items' parity
// bits (check_core's layout).
// Offsets are 32-bit byte offsets from the (uniform) table and message
// bases, so the loads and stores take the scalar-base + 32-bit vector-offset
// form and no 64-bit multiply is formed per item.
// ---------------------------------------------------------------------------
// One batch of uniform_items: G items, all valid, at table byte offset ob and
// message byte offset mb (the next item ostep / mstep further).
template <int D, int G, bool EQ, bool FIRST, int VK>
__device__ __forceinline__ uint2 uniform_batch(const char *obase, char *mbytes, char *pbase, uint32_t ob,
                                               uint32_t mb, uint32_t ostep, uint32_t mstep, int variant,
                                               const MsgUnpack &mu, u16x2 lpe, u16x2 lpo) {
    constexpr int H = (D + 1) / 2;
    uint32_t off[G][H];
    uint32_t mw[G][D];
#pragma unroll
    for (int g = 0; g < G; g++) {
        const uint32_t *op = reinterpret_cast<const uint32_t *>(obase + (ob + g * ostep));
#pragma unroll
        for (int h = 0; h < H; h++) off[g][h] = op[h];
        if (!FIRST) load_msgs<D>(reinterpret_cast<const uint32_t *>(mbytes + (mb + g * mstep)), mw[g]);
    }
    uint2 ov = make_uint2(0u, 0u);
#pragma unroll
    for (int g = 0; g < G; g++) {
        uint32_t *vp[D];
        uint32_t pw[D], pn[D], mn[D];
#pragma unroll
        for (int j = 0; j < D; j++)
            vp[j] = reinterpret_cast<uint32_t *>(pbase + ((j & 1) ? (off[g][j >> 1] >> 16) : (off[g][j >> 1] & 0xFFFFu)));
#pragma unroll
        for (int j = 0; j < D; j++) pw[j] = *vp[j];
        const uint2 pr = check_core<D, EQ>(pw, mw[g], FIRST, VK >= 0 ? VK : variant, false, lpe, lpo, pn, mn, mu);
        ov.x |= pr.x;
        ov.y |= pr.y;
        store_msgs<D>(reinterpret_cast<uint32_t *>(mbytes + (mb + g * mstep)), mn);
#pragma unroll
        for (int j = 0; j < D; j++) *vp[j] = pn[j];
    }
    return ov;
}

template <int D, int G, bool EQ, bool FIRST, int VK = -1>
__device__ __forceinline__ uint2 uniform_items(const uint32_t *__restrict__ edge_off16, int P, int variant,
                                               uint32_t *post, uint32_t *mbase, int c_first, int le, int cstride,
                                               int w, uint32_t keep, u16x2 lpe, u16x2 lpo) {
    constexpr int H = (D + 1) / 2;
    char *const pbase = reinterpret_cast<char *>(post) + 4 * w;
    const char *const obase = reinterpret_cast<const char *>(edge_off16);
    char *const mbytes = reinterpret_cast<char *>(mbase);
    const MsgUnpack mu = msg_unpack(keep);
    // check indices and P fit 24 bits: v_mul_u32_u24, not v_mul_lo_u32
    const uint32_t ostep = __umul24(cstride, 4 * H), mstep = __umul24(__umul24(cstride, P), 4 * D);
    uint32_t ob = __umul24(c_first, 4 * H), mb = __umul24(__umul24(c_first, P) + w, 4 * D);
    uint2 ov = make_uint2(0u, 0u);
    int c = c_first;
    for (; c + (G - 1) * cstride < le; c += G * cstride, ob += G * ostep, mb += G * mstep) {
        const uint2 pr = uniform_batch<D, G, EQ, FIRST, VK>(obase, mbytes, pbase, ob, mb, ostep, mstep, variant, mu,
                                                            lpe, lpo);
        ov.x |= pr.x;
        ov.y |= pr.y;
    }
    if constexpr (G > 1) {
        for (; c < le; c += cstride, ob += ostep, mb += mstep) {
            const uint2 pr = uniform_batch<D, 1, EQ, FIRST, VK>(obase, mbytes, pbase, ob, mb, ostep, mstep, variant,
                                                                mu, lpe, lpo);
            ov.x |= pr.x;
            ov.y |= pr.y;
        }
    }
    return ov;
}

}  // namespace ldpc
