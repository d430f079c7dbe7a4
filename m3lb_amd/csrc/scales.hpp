// scales.hpp -- choice of the power-of-two scales of 16-bit storage, shared
// by the single-box solver and the batched server (host code only), so that a
// box gets the same scales, and therefore the same bits, in either.
//
// A scale is a power of two 2^k per opposite pair.
//   F16, BF16: the pair's largest |g| maps into (2^7, 2^8]: FP16 then has 2^8
//     of headroom under its largest finite code (65504) for the flow to grow
//     into, and keeps full precision for values down to 2^-22 of that maximum
//     (its smallest normal is 2^-14).  A floating format's relative precision
//     is the same at every scale, so nothing is lost by mapping low.
//   I16: scales proportional to the weights (see pair_scales), the largest
//     |g| mapped into (2^11, 2^12]: 8x to 16x of headroom under the
//     saturation at 32767, and a uniform error of half a code.
//   E2M13: scales proportional to the weights like I16, the largest |g|
//     mapped into (32, 64]: 4x to 8x of headroom under the saturation at 256,
//     and one step, 2^-8 (2^-13 to 2^-14 of the largest |g|), for every value
//     up to 64.  A format whose step grows within the range the values occupy
//     lets the slow viscous decay of the largest values stall below half a
//     step: such an offset float (|v| + 32 stored with 14 significant bits)
//     ran the 512^3 Taylor-Green case with a kinetic energy 60x to 100x
//     further from FP64 than this one.
// A pair whose largest |g| is below a floor is scaled as if it were at the
// floor: 1e-3 w_q for the floating formats, whose range is wide, and 0.1 w_q
// for I16 and E2M13, whose narrow headroom would otherwise be set by a field
// at rest: a cavity started at rest and driven to u ~ 0.05 (|g| ~ 0.15 w)
// saturated I16 and lost the flow (velocity error 0.7), while 0.1 w_q leaves
// room for |g| up to about 0.8 w_q at a resolution of 2.4e-5 w_q.
#pragma once

#include <algorithm>
#include <cmath>

#include "../../include/m3lb.h"

namespace m3lb {

inline bool is_16bit(int dtype) { return dtype >= M3LB_F16; }

inline bool narrow_range(int dtype) { return dtype == M3LB_I16 || dtype == M3LB_E2M13; }

inline double scale_floor(int dtype) { return narrow_range(dtype) ? 0.1 : 1e-3; }

inline double scale_target(int dtype) {
  return dtype == M3LB_I16 ? 4096.0 : dtype == M3LB_E2M13 ? 64.0 : 256.0;
}

inline double weight_of(int Q, int q) {
  static const int wn9[9] = {16, 4, 4, 4, 4, 1, 1, 1, 1};
  static const int wn27[27] = {64, 16, 16, 16, 16, 16, 16, 4, 4, 4, 4, 4, 4, 4,
                               4,  4,  4,  4,  4,  1,  1,  1, 1, 1, 1, 1, 1};
  return Q == 9 ? wn9[q] / 36.0 : wn27[q] / 216.0;
}

// floor(log2(x)) exactly, for finite x > 0
inline int floor_log2(double x) {
  int e = 0;
  std::frexp(x, &e);  // x = m 2^e, m in [0.5, 1)
  return e - 1;
}

inline int scale_exponent(int dtype, int Q, int q, double max_abs) {
  const double m = std::max(max_abs, scale_floor(dtype) * weight_of(Q, q));
  return std::max(-60, std::min(60, floor_log2(scale_target(dtype) / m)));
}

// Provisional scales before the first seeding: |g_q| <= w_q/2 assumed.
inline void provisional_scales(int dtype, int Q, float* scale, float* inv_scale) {
  for (int q = 0; q < 27; ++q) { scale[q] = 1.0f; inv_scale[q] = 1.0f; }
  if (!is_16bit(dtype)) return;
  for (int q = 0; q < Q; ++q) {
    const int k = scale_exponent(dtype, Q, q, 0.5 * weight_of(Q, q));
    scale[q] = std::ldexp(1.0f, k);
    inv_scale[q] = std::ldexp(1.0f, -k);
  }
}

// The scales a field calls for, from its largest |g_q| per direction (mx).
// I16 and E2M13 scale every pair in proportion to its weight, from the
// largest |g_q| / w_q over all directions: a direction that starts near zero
// (e.g. along an axis the initial flow has no component on) would otherwise
// get a scale it saturates at once the flow turns into it.
inline void pair_scales(int dtype, int Q, const double* mx, double* want) {
  double ratio = 0;
  for (int q = 0; q < Q; ++q) ratio = std::max(ratio, mx[q] / weight_of(Q, q));
  for (int q = 0; q < Q; ++q) {
    const int o = q == 0 ? 0 : (q & 1 ? q + 1 : q - 1);
    const double m = narrow_range(dtype) ? ratio * weight_of(Q, q) : std::max(mx[q], mx[o]);
    want[q] = std::ldexp(1.0, scale_exponent(dtype, Q, q, m));
  }
}

}  // namespace m3lb
