// Exhaustive host check of the narrow 16-bit float codec (NarrowFloat16<EB>)
// against its definition: every code decodes to the value its fields define,
// every code re-encodes to itself, encoding a float gives the nearest code
// with ties to the even code (below the normal binades, within 2^-10 of a
// step of a tie, either neighbour), magnitudes beyond the largest code,
// infinities and NaNs saturate with their sign, and the forms with a
// power-of-two scale folded in agree with scaling first.  Prints "ok" when
// all hold.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <limits>
#include <random>
#include <vector>

#include "../../m3lb_amd/csrc/narrow_float.hpp"

namespace {
int fails = 0;
void check(bool ok, const char* what, double a, double b) {
  if (!ok && fails++ < 10) std::printf("FAIL %s: %.17g vs %.17g\n", what, a, b);
}

template <int EB>
void run() {
  using F = m3lb::NarrowFloat16<EB>;
  const int M = 15 - EB, emin = 9 - (1 << EB);
  // magnitude of each magnitude code m from its fields
  std::vector<double> val(0x8000);
  for (uint32_t m = 0; m < 0x8000; ++m) {
    const uint32_t e = m >> M, f = m & ((1u << M) - 1);
    val[m] = e == 0 ? std::ldexp(double(f), emin - M)
                    : std::ldexp(1.0 + std::ldexp(double(f), -M), emin - 1 + int(e));
  }
  check(val[1u << M] == std::ldexp(1.0, emin), "first normal", val[1u << M], emin);
  check(std::ldexp(1.0, 8) - val[0x7fff] == std::ldexp(1.0, 7 - M), "largest code", val[0x7fff], 256);
  for (uint32_t c = 0; c < 0x10000; ++c) {
    const double want = (c & 1) ? -val[c >> 1] : val[c >> 1];
    const float got = F::dec(uint16_t(c));
    check(double(got) == want, "decode", got, want);
    check(std::signbit(got) == bool(c & 1), "decode sign", got, want);
    check(F::enc(got) == c, "re-encode", double(F::enc(got)), double(c));
  }
  for (uint32_t m = 1; m < 0x8000; ++m) check(val[m] > val[m - 1], "increasing", val[m], val[m - 1]);

  auto code = [](uint32_t m, bool neg) { return (m << 1) | (neg ? 1u : 0u); };
  std::mt19937_64 rng(12345);
  std::uniform_real_distribution<double> lg(-40.0, 9.0), un(-300.0, 300.0);
  for (int k = 0; k < 4000000; ++k) {
    const float x = (k & 1) ? float(std::exp2(lg(rng)) * ((k & 2) ? -1 : 1)) : float(un(rng));
    const double a = std::fabs(double(x));
    const uint32_t got = F::enc(x);
    check((got & 1) == (std::signbit(x) ? 1u : 0u), "encode sign", double(got), x);
    if (a >= val[0x7fff]) {
      check((got >> 1) == 0x7fff, "saturate", double(got), x);
      continue;
    }
    uint32_t lo = 0, hi = 0x7fff;  // val[lo] <= a < val[hi]
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) / 2;
      (val[mid] <= a ? lo : hi) = mid;
    }
    const double tie = 0.5 * (val[lo] + val[hi]);
    const uint32_t near = a < tie ? lo : a > tie ? hi : ((lo & 1) ? hi : lo);
    const bool band = a < std::ldexp(1.0, emin) && std::fabs(a - tie) <= std::ldexp(val[hi] - val[lo], -10);
    if (!band) check((got >> 1) == near, "encode nearest", double(got >> 1), near);
    else check((got >> 1) == lo || (got >> 1) == hi, "encode near a tie", double(got >> 1), near);
  }
  // every tie (to the even code), and inputs just outside the band around it
  for (uint32_t m = 0; m + 1 < 0x8000; ++m) {
    const double tie = 0.5 * (val[m] + val[m + 1]), band = std::ldexp(val[m + 1] - val[m], -9);
    const uint32_t even = (m & 1) ? m + 1 : m;
    check(double(float(tie)) == tie, "tie exact in FP32", tie, 0);
    check(F::enc(float(tie)) == code(even, false), "tie to even", double(F::enc(float(tie))), even);
    check(F::enc(-float(tie)) == code(even, true), "negative tie", double(F::enc(-float(tie))), even);
    check(F::enc(float(tie - band)) == code(m, false), "below tie", double(F::enc(float(tie - band))), m);
    check(F::enc(float(tie + band)) == code(m + 1, false), "above tie", double(F::enc(float(tie + band))), m + 1);
  }
  // the scaled forms equal the plain ones at every power-of-two scale used
  for (int k = -60; k <= 60; k += 7) {
    const float sc = std::ldexp(1.0f, k), inv = std::ldexp(1.0f, -k);
    for (uint32_t c = 0; c < 0x10000; c += 7)
      check(F::dec_scaled(uint16_t(c), inv) == F::dec(uint16_t(c)) * inv, "scaled decode",
            F::dec_scaled(uint16_t(c), inv), k);
    for (int j = 0; j < 200000; ++j) {
      const float x = float(std::exp2(lg(rng)) * ((j & 1) ? -1 : 1)) * inv;
      check(F::enc_scaled(x, sc) == F::enc(x * sc), "scaled encode", double(F::enc_scaled(x, sc)), k);
    }
    check(F::enc_scaled(std::nanf(""), sc) == code(0x7fff, false), "scaled NaN", k, 0);
  }
  const float inf = std::numeric_limits<float>::infinity();
  check(F::enc(inf) == code(0x7fff, false) && F::enc(-inf) == code(0x7fff, true), "infinity saturates", 0, 0);
  check(F::enc(1e30f) == code(0x7fff, false) && F::enc(std::nanf("")) == code(0x7fff, false), "saturation", 0, 0);
  check(F::enc(-0.0f) == 1u && F::enc(0.0f) == 0u, "signed zero", double(F::enc(-0.0f)), 1);
}
}  // namespace

int main() {
  run<2>();
  run<3>();
  if (fails) {
    std::printf("%d failures\n", fails);
    return 1;
  }
  std::printf("ok\n");
  return 0;
}
