// narrow_float.hpp -- a 16-bit float with a narrow, fixed range for scaled
// population storage.  Host and device code share it (the host build of the
// exhaustive codec test compiles it with g++).
//
// Code layout: bit 0 is the sign, bits 1..15 the magnitude code m, split into
// EB exponent bits e (high) and M = 15 - EB mantissa bits f (low).  With
// EMIN = 9 - 2^EB the magnitude is
//     |v| = 2^(EMIN - 1 + e) (1 + f / 2^M)   for e >= 1 (binades [2^EMIN, 2^8)),
//     |v| = f 2^(EMIN - M)                   for e = 0  (evenly spaced below),
// so EB = 2 holds [0, 256) with 14 significant bits over [32, 256) and a step
// of 2^-8 below 32.  No infinities or NaNs: magnitudes beyond the largest
// code, infinities and NaNs saturate to it, keeping their sign bit.  A
// per-pair scale spends no bits on range a population never reaches.
//
// Both directions are a few FP32 and integer operations on
//     G = max(|v|, |v| / 2 + 2^(EMIN-1)),
// which is |v| itself in the normal binades and |v|/2 + 2^(EMIN-1) in
// [2^(EMIN-1), 2^EMIN) below them: the bit pattern of G, less that of
// 2^(EMIN-1), is the magnitude code shifted left by 23 - M, with no branch.
// Decoding inverts it exactly as min(2 G - 2^EMIN, G).  Encoding rounds G to
// M mantissa bits, ties to even; below 2^EMIN, G itself is first rounded in
// FP32 (at 2^-10 of a step), so there an input within that distance of a tie
// may take either neighbour.  Elsewhere the result is the nearest code.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define M3LB_HD __host__ __device__ __forceinline__
#else
#define M3LB_HD inline
#endif

namespace m3lb {

template <int EB>
struct NarrowFloat16 {
  static_assert(EB >= 1 && EB <= 4, "exponent bits");
  static constexpr int M = 15 - EB;
  static constexpr int EMIN = 9 - (1 << EB);
  static constexpr int SHIFT = 23 - M;
  static constexpr uint32_t BASE = uint32_t(127 + EMIN - 1) << 23;  // bits of 2^(EMIN-1)
  static constexpr float HALF_MIN = __builtin_bit_cast(float, BASE);
  static constexpr float P_MIN = 2.0f * HALF_MIN;                   // 2^EMIN
  static constexpr uint32_t MAXM = 0x7fffu;

  M3LB_HD static uint16_t enc(float x) {
    const float a = __builtin_fabsf(x);
    const float g = __builtin_fmaxf(a, __builtin_fmaf(a, 0.5f, HALF_MIN));
    const uint32_t t = __builtin_bit_cast(uint32_t, g) - BASE;
    uint32_t m = (t + ((1u << (SHIFT - 1)) - 1u) + ((t >> SHIFT) & 1u)) >> SHIFT;
    m = m < MAXM ? m : MAXM;
    return uint16_t((m << 1) | (__builtin_bit_cast(uint32_t, x) >> 31));
  }
  M3LB_HD static float dec(uint16_t c) {
    // the sign bit lands below the mantissa's last kept bit and is cleared
    // after the add (a shift-add is one instruction, a masked shift is not)
    const uint32_t gb = ((uint32_t(c) << (SHIFT - 1)) + BASE) & ~(1u << (SHIFT - 1));
    // both operands are non-negative (g >= 2^(EMIN-1)), so the smaller float
    // has the smaller bit pattern: an integer min needs no quieting of NaNs
    const uint32_t lo = __builtin_bit_cast(uint32_t, __builtin_fmaf(__builtin_bit_cast(float, gb), 2.0f, -P_MIN));
    const uint32_t mag = lo < gb ? lo : gb;
#if defined(__HIP_DEVICE_COMPILE__)
    // one shift-or: written as two operations, the compiler splits most of
    // them into a shift and an or (one more vector instruction per decode)
    uint32_t r;
    asm("v_lshl_or_b32 %0, %1, 31, %2" : "=v"(r) : "v"(uint32_t(c)), "v"(mag));
    return __builtin_bit_cast(float, r);
#else
    return __builtin_bit_cast(float, mag | (uint32_t(c) << 31));
#endif
  }

  // enc(x * s) and dec(c) * inv for a power-of-two scale s = 1 / inv, with
  // the scale folded into the constants: multiplying by 2^k adds k to the
  // exponent field, so G / s = max(|x|, |x| / 2 + 2^(EMIN-1) / s) carries
  // the bits of G less k << 23 (exactly, while G / s stays a normal float:
  // |k| <= 60 here).  The constants depend only on s and are uniform over a
  // wavefront, so they cost scalar integer work, not vector operations.
  M3LB_HD static uint16_t enc_scaled(float x, float s) {
    const uint32_t kb = __builtin_bit_cast(uint32_t, s) - 0x3f800000u;  // k << 23
    const float h = __builtin_bit_cast(float, BASE - kb);               // 2^(EMIN-1) / s
    const float a = __builtin_fabsf(x);
    const uint32_t gb = __builtin_bit_cast(uint32_t, __builtin_fmaxf(a, __builtin_fmaf(a, 0.5f, h)));
    // gb + kb - BASE is the rebased pattern of G; bit SHIFT of it is bit SHIFT
    // of gb, since kb and BASE are multiples of 2^23
    uint32_t m = (gb + (kb - BASE) + ((1u << (SHIFT - 1)) - 1u) + ((gb >> SHIFT) & 1u)) >> SHIFT;
    m = m < MAXM ? m : MAXM;
    return uint16_t((m << 1) | (__builtin_bit_cast(uint32_t, x) >> 31));
  }
  M3LB_HD static float dec_scaled(uint16_t c, float inv) {
    const uint32_t kb = __builtin_bit_cast(uint32_t, inv) - 0x3f800000u;  // -k << 23
    // the sign bit lands below the mantissa's last kept bit and is cleared
    // after the add (shift-add is one instruction, the masked shift is not)
    const uint32_t gb = ((uint32_t(c) << (SHIFT - 1)) + (BASE + kb)) & ~(1u << (SHIFT - 1));
    const float g = __builtin_bit_cast(float, gb);
    const float p = __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, P_MIN) + kb);  // 2^EMIN inv
    const uint32_t lo = __builtin_bit_cast(uint32_t, __builtin_fmaf(g, 2.0f, -p));
    const uint32_t mag = lo < gb ? lo : gb;  // non-negative: as in dec()
    return __builtin_bit_cast(float, mag | (uint32_t(c) << 31));
  }
};

}  // namespace m3lb
