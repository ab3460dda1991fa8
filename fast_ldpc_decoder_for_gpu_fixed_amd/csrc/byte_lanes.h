// byte_lanes.h -- saturating int8 lane arithmetic for gfx950, two frames per
// packed-16 VALU op.
//
// The reference computes with NVIDIA per-byte SIMD intrinsics on 4 frames per
// 32-bit word (src/decoder_ms/cuda/simd_functions.h:89-131: vsubss4, vaddss4,
// vabs4, vminu4, vmaxu4, vcmpgts4, vneg4, vsubus4).  gfx950 has no per-byte
// ALU, so a 32-bit word of 4 int8 lanes [b0 b1 b2 b3] is split into two
// packed-16 registers holding each byte in the HIGH byte of a 16-bit half:
//
//     even = { b0 << 8, b2 << 8 }   (one v_perm_b32)
//     odd  = { b1 << 8, b3 << 8 }   (one v_and_b32)
//
// With the low byte zero, v_pk_sub_i16/v_pk_add_i16 with clamp saturate at
// exactly the int8 limits (0x7FFF has high byte 0x7F, 0x8000 is -128), the
// u16 min/max order equals the uint8 order and v_pk_ashrrev_i16 by 15 gives a
// per-frame sign mask.  A saturated positive result keeps 0xFF in its low
// byte; every consumer below only reads high bytes or compares against values
// where that residue cannot change the outcome (argued at each use in
// decode_kernels.hip).  One repack (v_perm_b32) returns to 4 lanes per word.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ldpc {

typedef short s16x2 __attribute__((ext_vector_type(2)));
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ u16x2 as_u(s16x2 x) { return __builtin_bit_cast(u16x2, x); }
__device__ __forceinline__ s16x2 as_s(u16x2 x) { return __builtin_bit_cast(s16x2, x); }
__device__ __forceinline__ u16x2 u2(uint32_t x) { return __builtin_bit_cast(u16x2, x); }
__device__ __forceinline__ uint32_t w32(u16x2 x) { return __builtin_bit_cast(uint32_t, x); }
__device__ __forceinline__ uint32_t w32(s16x2 x) { return __builtin_bit_cast(uint32_t, x); }

// bytes 0 and 2 into the high bytes of the two halves
__device__ __forceinline__ u16x2 unpack_even(uint32_t x) {
    return u2(__builtin_amdgcn_perm(x, x, 0x020C000Cu));
}
// bytes 1 and 3 already sit in the high bytes
__device__ __forceinline__ u16x2 unpack_odd(uint32_t x) { return u2(x & 0xFF00FF00u); }
// inverse: word = [even.h0, odd.h0, even.h1, odd.h1] (high bytes)
__device__ __forceinline__ uint32_t repack(u16x2 even, u16x2 odd) {
    return __builtin_amdgcn_perm(w32(odd), w32(even), 0x07030501u);
}

__device__ __forceinline__ u16x2 sub_sat_s(u16x2 a, u16x2 b) {
    return as_u(__builtin_elementwise_sub_sat(as_s(a), as_s(b)));
}
__device__ __forceinline__ u16x2 add_sat_s(u16x2 a, u16x2 b) {
    return as_u(__builtin_elementwise_add_sat(as_s(a), as_s(b)));
}
__device__ __forceinline__ u16x2 sub_sat_u(u16x2 a, u16x2 b) {
    return __builtin_elementwise_sub_sat(a, b);
}
__device__ __forceinline__ u16x2 umin(u16x2 a, u16x2 b) { return __builtin_elementwise_min(a, b); }
__device__ __forceinline__ u16x2 umax(u16x2 a, u16x2 b) { return __builtin_elementwise_max(a, b); }
// arithmetic shift of each half by 15: 0xFFFF where the top bit is set (the
// sign mask the check update turns into a +-1 factor, or uses in the
// conditional negate (r ^ m) - m).  Written as one VALU instruction in asm:
// left visible, the compiler rewrites the shift and its use into a per-half
// compare/select sequence (about three times longer).
__device__ __forceinline__ u16x2 sign_mask(u16x2 x) {
    uint32_t r;
    asm("v_pk_ashrrev_i16 %0, 15, %1 op_sel_hi:[0,1]" : "=v"(r) : "v"(w32(x)));
    return u2(r);
}
__device__ __forceinline__ u16x2 splat(uint32_t h) { return u2(h * 0x00010001u); }

}  // namespace ldpc
