// fastdiv.hpp -- exact unsigned 32-bit division by a run-time constant with
// one high multiply (Granlund & Montgomery, "Division by invariant integers
// using multiplication", PLDI 1994, Fig. 4.1): for d >= 1 and l = ceil(log2 d),
//   m  = floor(2^32 (2^l - d) / d) + 1
//   t  = mulhi(m, n)
//   q  = (t + ((n - t) >> min(l, 1))) >> max(l - 1, 0)
// equals floor(n / d) for every 32-bit n.  Plain C++ on the host (so the CPU
// tests can check it exhaustively over ranges), HIP on the device.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define M3LB_HD __host__ __device__ __forceinline__
#else
#define M3LB_HD inline
#endif

namespace m3lb {

struct FastDiv {
  uint32_t d, m;
  uint32_t s1, s2;
};

inline FastDiv make_fastdiv(uint32_t d) {
  FastDiv f;
  f.d = d;
  uint32_t l = 0;
  while (l < 32 && (uint64_t(1) << l) < d) ++l;
  f.m = uint32_t(((uint64_t(1) << 32) * ((uint64_t(1) << l) - d)) / d + 1);
  f.s1 = l < 1 ? l : 1;
  f.s2 = l > 1 ? l - 1 : 0;
  return f;
}

M3LB_HD uint32_t fd_div(const FastDiv& f, uint32_t n) {
  const uint32_t t = uint32_t((uint64_t(f.m) * n) >> 32);
  return (t + ((n - t) >> f.s1)) >> f.s2;
}

}  // namespace m3lb
