// lbm_device.hpp -- device-side building blocks of the gfx950 esoteric-pull step:
// velocity sets, population codecs, and the collision operators, all with
// compile-time population and moment indices so that every array below lives
// in VGPRs (no scratch, no indexed table loads).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>
#include <utility>

#include "../../include/m3lb.h"
#include "narrow_float.hpp"

namespace m3lb {

// ---------------------------------------------------------------------------
// Compile-time loop: f(std::integral_constant<int, I>) for I = 0..N-1.
// ---------------------------------------------------------------------------
template <class F, int... I>
__host__ __device__ __forceinline__ void sfor_impl(F&& f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int N, class F>
__host__ __device__ __forceinline__ void sfor(F&& f) {
  sfor_impl(f, std::make_integer_sequence<int, N>{});
}

// ---------------------------------------------------------------------------
// Velocity sets in esoteric-pull order: rest at 0, opposite pairs (i, i+1) for
// odd i (src/lattice/Lattices.hpp:72-96, 125-159).
// ---------------------------------------------------------------------------
template <int Q>
struct Lat;

template <>
struct Lat<9> {
  static constexpr int Q = 9, D = 2, wden = 36;
  static constexpr int cx[9] = {0, 1, -1, 0, 0, 1, -1, 1, -1};
  static constexpr int cy[9] = {0, 0, 0, 1, -1, 1, -1, -1, 1};
  static constexpr int cz[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  static constexpr int wn[9] = {16, 4, 4, 4, 4, 1, 1, 1, 1};
};

template <>
struct Lat<27> {
  static constexpr int Q = 27, D = 3, wden = 216;
  static constexpr int cx[27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1,
                                 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1};
  static constexpr int cy[27] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0,
                                 1, -1, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1};
  static constexpr int cz[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1,
                                 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1};
  static constexpr int wn[27] = {64, 16, 16, 16, 16, 16, 16, 4, 4, 4, 4, 4, 4, 4,
                                 4,  4,  4,  4,  4,  1,  1,  1, 1, 1, 1, 1, 1};
};

template <int Q>
__host__ __device__ constexpr int cvel(int i, int a) {
  return a == 0 ? Lat<Q>::cx[i] : (a == 1 ? Lat<Q>::cy[i] : Lat<Q>::cz[i]);
}
template <int Q>
__host__ __device__ constexpr int opp(int i) {
  return i == 0 ? 0 : (i & 1 ? i + 1 : i - 1);
}
// Population index with velocity (a-1, b-1, c-1); the product-basis slot map.
template <int Q>
__host__ __device__ constexpr int dir_of(int a, int b, int c) {
  for (int i = 0; i < Q; ++i)
    if (Lat<Q>::cx[i] == a - 1 && Lat<Q>::cy[i] == b - 1 &&
        Lat<Q>::cz[i] == (Q == 27 ? c - 1 : 0))
      return i;
  return -1;
}

// ---------------------------------------------------------------------------
// Population codecs.  Compute type R is double for F64 and float otherwise.
// 16-bit codes hold s_q * g_q; s_q is a power of two shared by each opposite
// pair, so the slot a value sits in never changes its scale.
// ---------------------------------------------------------------------------
template <int DT>
struct Codec;

template <>
struct Codec<M3LB_F64> {
  using T = double;
  using R = double;
  static constexpr bool scaled = false;
  __device__ __forceinline__ static R dec(T v, float) { return v; }
  __device__ __forceinline__ static T enc(R v, float) { return v; }
};
template <>
struct Codec<M3LB_F32> {
  using T = float;
  using R = float;
  static constexpr bool scaled = false;
  __device__ __forceinline__ static R dec(T v, float) { return v; }
  __device__ __forceinline__ static T enc(R v, float) { return v; }
};
template <>
struct Codec<M3LB_F16> {
  using T = _Float16;
  using R = float;
  static constexpr bool scaled = true;
  __device__ __forceinline__ static R dec(T v, float inv) { return float(v) * inv; }
  __device__ __forceinline__ static T enc(R v, float s) { return T(v * s); }
};
template <>
struct Codec<M3LB_BF16> {
  using T = __bf16;
  using R = float;
  static constexpr bool scaled = true;
  __device__ __forceinline__ static R dec(T v, float inv) { return float(v) * inv; }
  __device__ __forceinline__ static T enc(R v, float s) { return T(v * s); }
};
// 16-bit fixed point: code = round-to-nearest-even(s_q * g_q), saturated to
// [-32767, 32767].  The error of a value is uniform (half a code), where
// FP16's grows with the value itself (2^-11 relative).
template <>
struct Codec<M3LB_I16> {
  using T = int16_t;
  using R = float;
  static constexpr bool scaled = true;
  __device__ __forceinline__ static R dec(T v, float inv) { return float(v) * inv; }
  __device__ __forceinline__ static T enc(R v, float s) {
    const float r = __builtin_rintf(v * s);
    return T(__builtin_fminf(__builtin_fmaxf(r, -32767.0f), 32767.0f));
  }
};

// ---------------------------------------------------------------------------
// E2M13: the narrow 16-bit float of narrow_float.hpp with 2 exponent and 13
// mantissa bits: magnitudes [0, 256), 14 significant bits over [32, 256), a
// step of 2^-8 below.
template <>
struct Codec<M3LB_E2M13> {
  using T = uint16_t;
  using R = float;
  using F = NarrowFloat16<2>;
  static constexpr bool scaled = true;
  // decoding multiplies by the scale's inverse, which the compiler folds into
  // the moment sums; the folded decoder (dec_scaled) measured more vector
  // instructions in the step kernel
  __device__ __forceinline__ static R dec(T v, float inv) { return F::dec(v) * inv; }
  __device__ __forceinline__ static T enc(R v, float s) { return F::enc_scaled(v, s); }
};

// ---------------------------------------------------------------------------
// Kernel configuration: everything the reference makes a template parameter
// of FluidSolver / the collision policy (FluidSolver.hpp:34-35, BGK.hpp:31-32,
// MomentCollision.hpp:65-66).
// ---------------------------------------------------------------------------
template <int Q_, int OP_, int ST_, int DT_, int FORCE_>
struct Cfg {
  static constexpr int Q = Q_, D = Lat<Q_>::D, OP = OP_, DT = DT_;
  static constexpr bool SHIFTED = (OP_ != M3LB_OP_NULL) && ST_ == M3LB_SHIFTED;
  // FORCE_: M3LB_FORCE_NONE / _GUO (uniform F) / _FIELD (F read per node)
  static constexpr bool FORCE = FORCE_ != M3LB_FORCE_NONE && OP_ != M3LB_OP_NULL;
  static constexpr bool FIELD = FORCE_ == M3LB_FORCE_FIELD && OP_ != M3LB_OP_NULL;
  using C = Codec<DT_>;
  using T = typename C::T;
  using R = typename C::R;
};

template <class R>
struct Params {
  R omega, omega_m, omega_b;
  R F[3];
};

template <class R>
struct Macro {
  R dens, ux, uy, uz;  // dens = rho (raw) or rho - 1 (shifted), BGK.hpp:26-29
};

template <class K>
__device__ __forceinline__ typename K::R weight(int i) {
  using R = typename K::R;
  return R(Lat<K::Q>::wn[i]) / R(Lat<K::Q>::wden);
}

// Collision::density, BGK.hpp:53-56
template <class K>
__device__ __forceinline__ typename K::R density(const Macro<typename K::R>& m) {
  using R = typename K::R;
  if constexpr (K::SHIFTED) return R(1) + m.dens;
  else return m.dens;
}

// Raw monomial moments S_pqr = sum_i c_ix^p c_iy^q c_iz^r f_i of order <= 2,
// from one factorised pass: sums and differences along z, then y, then x
// (the first-order part of ProductBasis::to_moments at u_b = 0,
// ProductBasis.hpp:159-181).  Unused moments are dead code after inlining.
template <class R>
struct RawMoments {
  R S0, Sx, Sy, Sz, Sxx, Syy, Szz, Sxy, Sxz, Syz;
};

template <class K>
__device__ __forceinline__ RawMoments<typename K::R> raw_moments(const typename K::R* f) {
  using R = typename K::R;
  constexpr int Q = K::Q;
  RawMoments<R> S;
  if constexpr (K::D == 3) {
    R z0[3][3], z1[3][3], z2[3][3];
    sfor<9>([&](auto AB) {
      constexpr int a = AB / 3, b = AB % 3;
      const R vm = f[dir_of<Q>(a, b, 0)], v0 = f[dir_of<Q>(a, b, 1)], vp = f[dir_of<Q>(a, b, 2)];
      z2[a][b] = vp + vm;
      z1[a][b] = vp - vm;
      z0[a][b] = z2[a][b] + v0;
    });
    R Y00[3], Y10[3], Y20[3], Y01[3], Y11[3], Y02[3];
    for (int a = 0; a < 3; ++a) {
      Y20[a] = z0[a][2] + z0[a][0];
      Y10[a] = z0[a][2] - z0[a][0];
      Y00[a] = Y20[a] + z0[a][1];
      Y11[a] = z1[a][2] - z1[a][0];
      Y01[a] = (z1[a][2] + z1[a][0]) + z1[a][1];
      Y02[a] = (z2[a][2] + z2[a][0]) + z2[a][1];
    }
    S.Sxx = Y00[2] + Y00[0]; S.Sx = Y00[2] - Y00[0]; S.S0 = S.Sxx + Y00[1];
    S.Sxy = Y10[2] - Y10[0]; S.Sy = (Y10[2] + Y10[0]) + Y10[1];
    S.Syy = (Y20[2] + Y20[0]) + Y20[1];
    S.Sxz = Y01[2] - Y01[0]; S.Sz = (Y01[2] + Y01[0]) + Y01[1];
    S.Syz = (Y11[2] + Y11[0]) + Y11[1];
    S.Szz = (Y02[2] + Y02[0]) + Y02[1];
  } else {
    R Y0[3], Y1[3], Y2[3];
    sfor<3>([&](auto A) {
      constexpr int a = A;
      const R vm = f[dir_of<Q>(a, 0, 0)], v0 = f[dir_of<Q>(a, 1, 0)], vp = f[dir_of<Q>(a, 2, 0)];
      Y2[a] = vp + vm;
      Y1[a] = vp - vm;
      Y0[a] = Y2[a] + v0;
    });
    S.Sxx = Y0[2] + Y0[0]; S.Sx = Y0[2] - Y0[0]; S.S0 = S.Sxx + Y0[1];
    S.Sxy = Y1[2] - Y1[0]; S.Sy = (Y1[2] + Y1[0]) + Y1[1];
    S.Syy = (Y2[2] + Y2[0]) + Y2[1];
    S.Sz = R(0); S.Szz = R(0); S.Sxz = R(0); S.Syz = R(0);
  }
  return S;
}

// rho and u from S0 and S_a (BGK.hpp:63-78) with Guo::shift_velocity
// (Forcing.hpp:65-69).
template <class K>
__device__ __forceinline__ Macro<typename K::R> macro_of(const RawMoments<typename K::R>& S,
                                                         const Params<typename K::R>& p);

// macroscopic(): sum f, sum c f, u = m / rho, Guo half shift
// (BGK.hpp:63-78, Forcing.hpp:65-69); NullCollision gives {1,0,0,0}.
template <class K>
__device__ __forceinline__ Macro<typename K::R> macroscopic(const typename K::R* f,
                                                            const Params<typename K::R>& p) {
  using R = typename K::R;
  constexpr int Q = K::Q;
  Macro<R> m;
  if constexpr (K::OP == M3LB_OP_NULL) {
    m.dens = R(1); m.ux = R(0); m.uy = R(0); m.uz = R(0);
    return m;
  } else {
    return macro_of<K>(raw_moments<K>(f), p);
  }
}

template <class K>
__device__ __forceinline__ Macro<typename K::R> macro_of(const RawMoments<typename K::R>& S,
                                                         const Params<typename K::R>& p) {
  using R = typename K::R;
  Macro<R> m;
  m.dens = S.S0;
  const R rho = density<K>(m);
  // 16-bit storage rounds every population to 14 bits or fewer: the hardware
  // reciprocal (1 ulp) replaces the correctly rounded division there
  R ir;
  if constexpr (K::C::scaled) ir = __builtin_amdgcn_rcpf(rho);
  else ir = R(1) / rho;
  m.ux = S.Sx * ir; m.uy = S.Sy * ir; m.uz = S.Sz * ir;
  if constexpr (K::FORCE) {
    const R h = R(0.5) * ir;
    m.ux += h * p.F[0]; m.uy += h * p.F[1]; m.uz += h * p.F[2];
  }
  return m;
}

// ---------------------------------------------------------------------------
// Equilibria (src/equilibrium/Equilibrium.hpp:28-136).
// ---------------------------------------------------------------------------
template <class R, int Q, int i>
__device__ __forceinline__ R phi_second(R ux, R uy, R uz) {
  const R cu = R(cvel<Q>(i, 0)) * ux + R(cvel<Q>(i, 1)) * uy + R(cvel<Q>(i, 2)) * uz;
  const R uu = ux * ux + uy * uy + uz * uz;
  return R(3) * cu + R(4.5) * cu * cu - R(1.5) * uu;
}

template <class R, int c>
__device__ __forceinline__ R chi(R u) {
  if constexpr (c == 0) return R(1) - R(1.5) * u * u;
  else return R(1) + R(3) * u * (R(c) + u);
}

template <class R, int Q, int i>
__device__ __forceinline__ R phi_product(R ux, R uy, R uz) {
  if constexpr (Q == 27) {
    return chi<R, Lat<Q>::cx[i]>(ux) * chi<R, Lat<Q>::cy[i]>(uy) * chi<R, Lat<Q>::cz[i]>(uz) -
           R(1);
  } else {
    constexpr R cs2v = R(1) / R(3);
    const R cx = R(Lat<Q>::cx[i]);
    const R cy = R(Lat<Q>::cy[i]);
    const R ax = cx * cx - cs2v;
    const R ay = cy * cy - cs2v;
    const R o1 = (cx * ux + cy * uy) * R(3);
    const R o2 = (ax * ux * ux + ay * uy * uy + R(2) * cx * cy * ux * uy) * R(4.5);
    const R o3 = (ax * cy * ux * ux * uy + ay * cx * ux * uy * uy) * R(13.5);
    const R o4 = (ax * ay * ux * ux * uy * uy) * R(20.25);
    return o1 + o2 + o3 + o4;
  }
}

template <class K, bool PRODUCT, int i>
__device__ __forceinline__ typename K::R equilibrium(const Macro<typename K::R>& m) {
  using R = typename K::R;
  const R phi = PRODUCT ? phi_product<R, K::Q, i>(m.ux, m.uy, m.uz)
                        : phi_second<R, K::Q, i>(m.ux, m.uy, m.uz);
  const R w = weight<K>(i);
  if constexpr (K::SHIFTED) return w * (m.dens + (R(1) + m.dens) * phi);  // eq_shifted
  else return w * m.dens * (R(1) + phi);                                   // eq
}

// guo_source_raw, Forcing.hpp:26-38
template <class K, int i>
__device__ __forceinline__ typename K::R guo_raw(const Macro<typename K::R>& m,
                                                 const Params<typename K::R>& p) {
  using R = typename K::R;
  constexpr int Q = K::Q;
  const R cx = R(cvel<Q>(i, 0)), cy = R(cvel<Q>(i, 1)), cz = R(cvel<Q>(i, 2));
  const R cu = cx * m.ux + cy * m.uy + cz * m.uz;
  const R bx = (cx - m.ux) * R(3) + cu * R(9) * cx;
  const R by = (cy - m.uy) * R(3) + cu * R(9) * cy;
  const R bz = (cz - m.uz) * R(3) + cu * R(9) * cz;
  return weight<K>(i) * (bx * p.F[0] + by * p.F[1] + bz * p.F[2]);
}

// ---------------------------------------------------------------------------
// BGK (BGK.hpp:81-92) and TRT (TRT.hpp:92-127).
// ---------------------------------------------------------------------------
template <class K, bool PRODUCT>
__device__ __forceinline__ void collide_bgk(typename K::R* f, const Macro<typename K::R>& m,
                                            const Params<typename K::R>& p) {
  using R = typename K::R;
  const R w = p.omega;
  const R cf = R(1) - R(0.5) * w;
  sfor<K::Q>([&](auto I) {
    constexpr int i = I;
    const R feq = equilibrium<K, PRODUCT, i>(m);
    f[i] += w * (feq - f[i]);
    if constexpr (K::FORCE) f[i] += cf * guo_raw<K, i>(m, p);
  });
}

// TRT<L, Eq, F, S> with either equilibrium (TRT.hpp:32-33, 86-90): PRODUCT
// selects the product form, as for BGK.
template <class K, bool PRODUCT>
__device__ __forceinline__ void collide_trt(typename K::R* f, const Macro<typename K::R>& m,
                                            const Params<typename K::R>& p) {
  using R = typename K::R;
  const R wp = p.omega, wm = p.omega_m;
  const R cp = R(0.5) * (R(1) - R(0.5) * wp);
  const R cm = R(0.5) * (R(1) - R(0.5) * wm);
  {
    const R e0 = equilibrium<K, PRODUCT, 0>(m);
    f[0] += wp * (e0 - f[0]);
    if constexpr (K::FORCE) f[0] += R(2) * cp * guo_raw<K, 0>(m, p);
  }
  sfor<(K::Q - 1) / 2>([&](auto P) {
    constexpr int i = 2 * P + 1, j = i + 1;
    const R ei = equilibrium<K, PRODUCT, i>(m), ej = equilibrium<K, PRODUCT, j>(m);
    R dp = R(0.5) * wp * ((ei + ej) - (f[i] + f[j]));
    R dm = R(0.5) * wm * ((ei - ej) - (f[i] - f[j]));
    if constexpr (K::FORCE) {
      const R si = guo_raw<K, i>(m, p), sj = guo_raw<K, j>(m, p);
      dp += cp * (si + sj);
      dm += cm * (si - sj);
    }
    f[i] += dp + dm;
    f[j] += dp - dm;
  });
}

// ---------------------------------------------------------------------------
// Factorised product basis {1, C, C^2 - cs2}, C = c - u_b
// (src/collision/ProductBasis.hpp:137-205).  Moment slot of (p, q, r) is
// 9p + 3q + r on D3Q27 and 3p + q on D2Q9.
// ---------------------------------------------------------------------------
template <int D>
__host__ __device__ constexpr int mslot(int p, int q, int r) {
  return D == 2 ? p * 3 + q : (p * 3 + q) * 3 + r;
}

// One forward pass on the values at c = -1, 0, +1 with precomputed 2u and
// u^2 - cs2.
template <class R>
__device__ __forceinline__ void fwd1d(R& a, R& b, R& c, R u, R u2m) {
  const R s0 = a + b + c;
  const R s1 = c - a;
  const R s2 = c + a;
  a = s0;
  b = s1 - u * s0;
  c = s2 - R(2) * u * s1 + u2m * s0;
}
template <class R>
__device__ __forceinline__ void inv1d(R& a, R& b, R& c, R u, R u2m) {
  const R m0 = a, m1 = b, m2 = c;
  const R s1 = m1 + u * m0;
  const R s2 = m2 + R(2) * u * s1 - u2m * m0;
  a = R(0.5) * (s2 - s1);
  b = m0 - s2;
  c = R(0.5) * (s2 + s1);
}
// The same inverse in seven operations (inv1d as written takes eight or
// nine: the halvings of s2 - s1 and s2 + s1 become one product and two fused
// multiply-adds).  Equal to inv1d up to rounding; the 16-bit storage forms,
// whose step is bound by arithmetic, use it.  u2 = 2 u.
template <class R>
__device__ __forceinline__ void inv1d_fast(R& a, R& b, R& c, R u, R u2, R u2m) {
  const R m0 = a;
  const R s1 = __builtin_fmaf(u, m0, b);
  const R s2 = __builtin_fmaf(u2, s1, __builtin_fmaf(-u2m, m0, c));
  const R hs1 = R(0.5) * s1;
  a = __builtin_fmaf(R(0.5), s2, -hs1);
  b = m0 - s2;
  c = __builtin_fmaf(R(0.5), s2, hs1);
}

// MomentCollision::collide, MomentCollision.hpp:234-314; CENTRAL selects u_b = u.
template <class K, bool CENTRAL>
__device__ __forceinline__ void collide_moment(typename K::R* f, const Macro<typename K::R>& m,
                                               const Params<typename K::R>& p) {
  using R = typename K::R;
  constexpr int Q = K::Q, D = K::D;
  constexpr R cs2 = R(1) / R(3);
  const R rho = density<K>(m);
  const R u[3] = {m.ux, m.uy, m.uz};
  R ub[3], du[3], u2m[3];
  for (int a = 0; a < 3; ++a) {
    ub[a] = CENTRAL ? u[a] : R(0);
    du[a] = u[a] - ub[a];
    u2m[a] = ub[a] * ub[a] - cs2;
  }
  // 1D equilibrium factors: Qf = {1, du, du^2}, Aw = {1, -ub, ub^2}
  R Qf[3][3], Aw[3][3];
  for (int a = 0; a < 3; ++a) {
    Qf[a][0] = R(1); Qf[a][1] = du[a]; Qf[a][2] = du[a] * du[a];
    Aw[a][0] = R(1); Aw[a][1] = -ub[a]; Aw[a][2] = ub[a] * ub[a];
  }
  auto eqm = [&](auto N) -> R {
    constexpr int n = N;
    constexpr int pp = D == 2 ? n / 3 : n / 9;
    constexpr int qq = D == 2 ? n % 3 : (n / 3) % 3;
    constexpr int rr = D == 2 ? 0 : n % 3;
    R e = rho * Qf[0][pp] * Qf[1][qq];
    R w = Aw[0][pp] * Aw[1][qq];
    if constexpr (D == 3) { e *= Qf[2][rr]; w *= Aw[2][rr]; }
    if constexpr (K::SHIFTED) return e - w;
    else return e;
  };

  R k[27];
  // to_moments
  sfor<3>([&](auto A) {
    sfor<3>([&](auto B) {
      sfor<(D == 3 ? 3 : 1)>([&](auto C) {
        k[mslot<D>(A, B, C)] = f[dir_of<Q>(A, B, C)];
      });
    });
  });
  if constexpr (D == 3) {
    sfor<9>([&](auto AB) {
      constexpr int a = AB / 3, b = AB % 3;
      fwd1d(k[mslot<3>(a, b, 0)], k[mslot<3>(a, b, 1)], k[mslot<3>(a, b, 2)], ub[2], u2m[2]);
    });
    sfor<9>([&](auto AR) {
      constexpr int a = AR / 3, r = AR % 3;
      fwd1d(k[mslot<3>(a, 0, r)], k[mslot<3>(a, 1, r)], k[mslot<3>(a, 2, r)], ub[1], u2m[1]);
    });
    sfor<9>([&](auto QR) {
      constexpr int q = QR / 3, r = QR % 3;
      fwd1d(k[mslot<3>(0, q, r)], k[mslot<3>(1, q, r)], k[mslot<3>(2, q, r)], ub[0], u2m[0]);
    });
  } else {
    sfor<3>([&](auto A) {
      constexpr int a = A;
      fwd1d(k[a * 3 + 0], k[a * 3 + 1], k[a * 3 + 2], ub[1], u2m[1]);
    });
    sfor<3>([&](auto Qd) {
      constexpr int q = Qd;
      fwd1d(k[q], k[3 + q], k[6 + q], ub[0], u2m[0]);
    });
  }

  const R omega = p.omega;
  const R wb = p.omega_b < R(0) ? omega : p.omega_b;
  const R w2 = R(1) - R(0.5) * omega;
  // order 1: the force enters here
  if constexpr (K::FORCE) {
    k[mslot<D>(1, 0, 0)] += p.F[0];
    k[mslot<D>(0, 1, 0)] += p.F[1];
    if constexpr (D == 3) k[mslot<D>(0, 0, 1)] += p.F[2];
  }
  // order 2: trace at omega_bulk, deviator and shear at omega
  {
    constexpr int i2d[3] = {mslot<D>(2, 0, 0), mslot<D>(0, 2, 0), mslot<D>(0, 0, 2)};
    R d[3], e[3], g[3];
    R tr = R(0), tre = R(0), trg = R(0);
    sfor<D>([&](auto A) {
      constexpr int a = A;
      d[a] = k[i2d[a]];
      e[a] = eqm(std::integral_constant<int, i2d[a]>{});
      g[a] = K::FORCE ? R(2) * p.F[a] * du[a] : R(0);
      tr += d[a]; tre += e[a]; trg += g[a];
    });
    constexpr R invD = R(1) / R(D);
    const R tr_post = (R(1) - wb) * tr + wb * tre + (R(1) - R(0.5) * wb) * trg;
    sfor<D>([&](auto A) {
      constexpr int a = A;
      k[i2d[a]] = (R(1) - omega) * (d[a] - tr * invD) + omega * (e[a] - tre * invD) +
                  w2 * (g[a] - trg * invD) + tr_post * invD;
    });
    sfor<D>([&](auto A) {
      sfor<D>([&](auto B) {
        constexpr int a = A, b = B;
        if constexpr (b > a) {
          constexpr int id = mslot<D>((a == 0 || b == 0), (a == 1 || b == 1),
                                      D == 3 && (a == 2 || b == 2));
          const R ge = K::FORCE ? (p.F[b] * du[a] + p.F[a] * du[b]) : R(0);
          k[id] = (R(1) - omega) * k[id] + omega * eqm(std::integral_constant<int, id>{}) +
                  w2 * ge;
        }
      });
    });
  }
  // order >= 3: straight to equilibrium
  sfor<(D == 3 ? 27 : 9)>([&](auto N) {
    constexpr int n = N;
    constexpr int ord = D == 2 ? (n / 3 + n % 3) : (n / 9 + (n / 3) % 3 + n % 3);
    if constexpr (ord >= 3) k[n] = eqm(N);
  });
  // to_populations
  if constexpr (D == 3) {
    sfor<9>([&](auto QR) {
      constexpr int q = QR / 3, r = QR % 3;
      inv1d(k[mslot<3>(0, q, r)], k[mslot<3>(1, q, r)], k[mslot<3>(2, q, r)], ub[0], u2m[0]);
    });
    sfor<9>([&](auto AR) {
      constexpr int a = AR / 3, r = AR % 3;
      inv1d(k[mslot<3>(a, 0, r)], k[mslot<3>(a, 1, r)], k[mslot<3>(a, 2, r)], ub[1], u2m[1]);
    });
    sfor<9>([&](auto AB) {
      constexpr int a = AB / 3, b = AB % 3;
      inv1d(k[mslot<3>(a, b, 0)], k[mslot<3>(a, b, 1)], k[mslot<3>(a, b, 2)], ub[2], u2m[2]);
    });
  } else {
    sfor<3>([&](auto Qd) {
      constexpr int q = Qd;
      inv1d(k[q], k[3 + q], k[6 + q], ub[0], u2m[0]);
    });
    sfor<3>([&](auto A) {
      constexpr int a = A;
      inv1d(k[a * 3 + 0], k[a * 3 + 1], k[a * 3 + 2], ub[1], u2m[1]);
    });
  }
  sfor<3>([&](auto A) {
    sfor<3>([&](auto B) {
      sfor<(D == 3 ? 3 : 1)>([&](auto C) {
        f[dir_of<Q>(A, B, C)] = k[mslot<D>(A, B, C)];
      });
    });
  });
}

// The step's moment operators, fused with macroscopic(): the ten raw monomial
// moments S_pqr = sum c_x^p c_y^q c_z^r f (p+q+r <= 2) come out of one
// factorised pass (sums and differences along z, then y, then x), give rho and
// u, and are shifted to the Hermite basis {1, C, C^2 - cs2} at u_b.  Exactly
// the first ten moments of ProductBasis::to_moments (ProductBasis.hpp:159-181);
// the moments of order >= 3 it would also produce are overwritten by their
// equilibria (MomentCollision.hpp:287-288), so they are never formed.
template <class K, bool CENTRAL>
__device__ __forceinline__ Macro<typename K::R> collide_moment_fused(typename K::R* f,
                                                                      const Params<typename K::R>& p) {
  using R = typename K::R;
  constexpr int Q = K::Q, D = K::D;
  constexpr R cs2 = R(1) / R(3);
  const RawMoments<R> RM = raw_moments<K>(f);
  const Macro<R> m = macro_of<K>(RM, p);
  const R rho = density<K>(m);
  const R S0 = RM.S0, Sx = RM.Sx, Sy = RM.Sy, Sz = RM.Sz, Sxx = RM.Sxx, Syy = RM.Syy,
          Szz = RM.Szz, Sxy = RM.Sxy, Sxz = RM.Sxz, Syz = RM.Syz;
  const R u[3] = {m.ux, m.uy, m.uz};
  R ub[3], du[3], u2m[3];
  for (int a = 0; a < 3; ++a) {
    ub[a] = CENTRAL ? u[a] : R(0);
    // u - u_b, written as the constant 0 for central moments: the same value
    // for every finite u, and the terms it multiplies then fold away
    du[a] = CENTRAL ? R(0) : u[a];
    u2m[a] = ub[a] * ub[a] - cs2;
  }
  R Qf[3][3], Aw[3][3];
  for (int a = 0; a < 3; ++a) {
    Qf[a][0] = R(1); Qf[a][1] = du[a]; Qf[a][2] = du[a] * du[a];
    Aw[a][0] = R(1); Aw[a][1] = -ub[a]; Aw[a][2] = ub[a] * ub[a];
  }
  auto eqm = [&](auto N) -> R {
    constexpr int n = N;
    constexpr int pp = D == 2 ? n / 3 : n / 9;
    constexpr int qq = D == 2 ? n % 3 : (n / 3) % 3;
    constexpr int rr = D == 2 ? 0 : n % 3;
    R e = rho * Qf[0][pp] * Qf[1][qq];
    R w = Aw[0][pp] * Aw[1][qq];
    if constexpr (D == 3) { e *= Qf[2][rr]; w *= Aw[2][rr]; }
    if constexpr (K::SHIFTED) return e - w;
    else return e;
  };
  // basis moments of order <= 2 at u_b
  R k[27];
  k[mslot<D>(0, 0, 0)] = S0;
  k[mslot<D>(1, 0, 0)] = Sx - ub[0] * S0;
  k[mslot<D>(0, 1, 0)] = Sy - ub[1] * S0;
  k[mslot<D>(2, 0, 0)] = Sxx - R(2) * ub[0] * Sx + u2m[0] * S0;
  k[mslot<D>(0, 2, 0)] = Syy - R(2) * ub[1] * Sy + u2m[1] * S0;
  k[mslot<D>(1, 1, 0)] = Sxy - ub[0] * Sy - ub[1] * Sx + ub[0] * ub[1] * S0;
  if constexpr (D == 3) {
    k[mslot<D>(0, 0, 1)] = Sz - ub[2] * S0;
    k[mslot<D>(0, 0, 2)] = Szz - R(2) * ub[2] * Sz + u2m[2] * S0;
    k[mslot<D>(1, 0, 1)] = Sxz - ub[0] * Sz - ub[2] * Sx + ub[0] * ub[2] * S0;
    k[mslot<D>(0, 1, 1)] = Syz - ub[1] * Sz - ub[2] * Sy + ub[1] * ub[2] * S0;
  }

  const R omega = p.omega;
  const R wb = p.omega_b < R(0) ? omega : p.omega_b;
  const R w2 = R(1) - R(0.5) * omega;
  if constexpr (K::FORCE) {
    k[mslot<D>(1, 0, 0)] += p.F[0];
    k[mslot<D>(0, 1, 0)] += p.F[1];
    if constexpr (D == 3) k[mslot<D>(0, 0, 1)] += p.F[2];
  }
  {
    constexpr int i2d[3] = {mslot<D>(2, 0, 0), mslot<D>(0, 2, 0), mslot<D>(0, 0, 2)};
    R d[3], e[3], g[3];
    R tr = R(0), tre = R(0), trg = R(0);
    sfor<D>([&](auto A) {
      constexpr int a = A;
      d[a] = k[i2d[a]];
      e[a] = eqm(std::integral_constant<int, i2d[a]>{});
      g[a] = K::FORCE ? R(2) * p.F[a] * du[a] : R(0);
      tr += d[a]; tre += e[a]; trg += g[a];
    });
    constexpr R invD = R(1) / R(D);
    const R tr_post = (R(1) - wb) * tr + wb * tre + (R(1) - R(0.5) * wb) * trg;
    sfor<D>([&](auto A) {
      constexpr int a = A;
      k[i2d[a]] = (R(1) - omega) * (d[a] - tr * invD) + omega * (e[a] - tre * invD) +
                  w2 * (g[a] - trg * invD) + tr_post * invD;
    });
    sfor<D>([&](auto A) {
      sfor<D>([&](auto B) {
        constexpr int a = A, b = B;
        if constexpr (b > a) {
          constexpr int id = mslot<D>((a == 0 || b == 0), (a == 1 || b == 1),
                                      D == 3 && (a == 2 || b == 2));
          const R ge = K::FORCE ? (p.F[b] * du[a] + p.F[a] * du[b]) : R(0);
          k[id] = (R(1) - omega) * k[id] + omega * eqm(std::integral_constant<int, id>{}) +
                  w2 * ge;
        }
      });
    });
  }
  sfor<(D == 3 ? 27 : 9)>([&](auto N) {
    constexpr int n = N;
    constexpr int ord = D == 2 ? (n / 3 + n % 3) : (n / 9 + (n / 3) % 3 + n % 3);
    if constexpr (ord >= 3) k[n] = eqm(N);
  });
  if constexpr (D == 3 && K::C::scaled && std::is_same_v<R, float>) {
    const R u2[3] = {R(2) * ub[0], R(2) * ub[1], R(2) * ub[2]};
    sfor<9>([&](auto QR) {
      constexpr int q = QR / 3, r = QR % 3;
      inv1d_fast(k[mslot<3>(0, q, r)], k[mslot<3>(1, q, r)], k[mslot<3>(2, q, r)], ub[0], u2[0],
                 u2m[0]);
    });
    sfor<9>([&](auto AR) {
      constexpr int a = AR / 3, r = AR % 3;
      inv1d_fast(k[mslot<3>(a, 0, r)], k[mslot<3>(a, 1, r)], k[mslot<3>(a, 2, r)], ub[1], u2[1],
                 u2m[1]);
    });
    sfor<9>([&](auto AB) {
      constexpr int a = AB / 3, b = AB % 3;
      inv1d_fast(k[mslot<3>(a, b, 0)], k[mslot<3>(a, b, 1)], k[mslot<3>(a, b, 2)], ub[2], u2[2],
                 u2m[2]);
    });
  } else if constexpr (D == 3) {
    sfor<9>([&](auto QR) {
      constexpr int q = QR / 3, r = QR % 3;
      inv1d(k[mslot<3>(0, q, r)], k[mslot<3>(1, q, r)], k[mslot<3>(2, q, r)], ub[0], u2m[0]);
    });
    sfor<9>([&](auto AR) {
      constexpr int a = AR / 3, r = AR % 3;
      inv1d(k[mslot<3>(a, 0, r)], k[mslot<3>(a, 1, r)], k[mslot<3>(a, 2, r)], ub[1], u2m[1]);
    });
    sfor<9>([&](auto AB) {
      constexpr int a = AB / 3, b = AB % 3;
      inv1d(k[mslot<3>(a, b, 0)], k[mslot<3>(a, b, 1)], k[mslot<3>(a, b, 2)], ub[2], u2m[2]);
    });
  } else {
    sfor<3>([&](auto Qd) {
      constexpr int q = Qd;
      inv1d(k[q], k[3 + q], k[6 + q], ub[0], u2m[0]);
    });
    sfor<3>([&](auto A) {
      constexpr int a = A;
      inv1d(k[a * 3 + 0], k[a * 3 + 1], k[a * 3 + 2], ub[1], u2m[1]);
    });
  }
  sfor<3>([&](auto A) {
    sfor<3>([&](auto B) {
      sfor<(D == 3 ? 3 : 1)>([&](auto C) {
        f[dir_of<Q>(A, B, C)] = k[mslot<D>(A, B, C)];
      });
    });
  });
  return m;
}

// macroscopic() + collide() as the step kernel runs them.
template <class K>
__device__ __forceinline__ Macro<typename K::R> collide_fused(typename K::R* f,
                                                             const Params<typename K::R>& p) {
  if constexpr (K::OP == M3LB_OP_MRT) {
    return collide_moment_fused<K, false>(f, p);
  } else if constexpr (K::OP == M3LB_OP_CM) {
    return collide_moment_fused<K, true>(f, p);
  } else {
    const Macro<typename K::R> m = macroscopic<K>(f, p);
    collide<K>(f, m, p);
    return m;
  }
}

template <class K>
__device__ __forceinline__ void collide(typename K::R* f, const Macro<typename K::R>& m,
                                        const Params<typename K::R>& p) {
  if constexpr (K::OP == M3LB_OP_BGK) collide_bgk<K, false>(f, m, p);
  else if constexpr (K::OP == M3LB_OP_BGK_PRODUCT) collide_bgk<K, true>(f, m, p);
  else if constexpr (K::OP == M3LB_OP_TRT) collide_trt<K, false>(f, m, p);
  else if constexpr (K::OP == M3LB_OP_TRT_PRODUCT) collide_trt<K, true>(f, m, p);
  else if constexpr (K::OP == M3LB_OP_MRT) collide_moment<K, false>(f, m, p);
  else if constexpr (K::OP == M3LB_OP_CM) collide_moment<K, true>(f, m, p);
}

// seed_value for all Q directions: BGK.hpp:95-99, TRT.hpp:129-133,
// MomentCollision.hpp:320-330 (inverse transform of the raw-basis equilibrium
// moments), NullCollision tests/test_streaming.cpp:37.
template <class K>
__device__ __forceinline__ void seed(typename K::R* out, typename K::R rho, typename K::R ux,
                                     typename K::R uy, typename K::R uz) {
  using R = typename K::R;
  constexpr int Q = K::Q, D = K::D;
  if constexpr (K::OP == M3LB_OP_NULL) {
    sfor<Q>([&](auto I) { out[int(I)] = rho; });
  } else if constexpr (K::OP == M3LB_OP_MRT || K::OP == M3LB_OP_CM) {
    constexpr R cs2 = R(1) / R(3);
    const R u[3] = {ux, uy, uz};
    R k[27];
    sfor<(D == 3 ? 27 : 9)>([&](auto N) {
      constexpr int n = N;
      constexpr int pp = D == 2 ? n / 3 : n / 9;
      constexpr int qq = D == 2 ? n % 3 : (n / 3) % 3;
      constexpr int rr = D == 2 ? 0 : n % 3;
      R qx = pp == 0 ? R(1) : (pp == 1 ? u[0] : u[0] * u[0]);
      R qy = qq == 0 ? R(1) : (qq == 1 ? u[1] : u[1] * u[1]);
      R qz = rr == 0 ? R(1) : (rr == 1 ? u[2] : u[2] * u[2]);
      R e = rho * qx * qy;
      if constexpr (D == 3) e *= qz;
      if constexpr (K::SHIFTED) e -= (n == 0 ? R(1) : R(0));
      k[n] = e;
    });
    const R z = R(0), zm = -cs2;
    if constexpr (D == 3) {
      sfor<9>([&](auto QR) {
        constexpr int q = QR / 3, r = QR % 3;
        inv1d(k[mslot<3>(0, q, r)], k[mslot<3>(1, q, r)], k[mslot<3>(2, q, r)], z, zm);
      });
      sfor<9>([&](auto AR) {
        constexpr int a = AR / 3, r = AR % 3;
        inv1d(k[mslot<3>(a, 0, r)], k[mslot<3>(a, 1, r)], k[mslot<3>(a, 2, r)], z, zm);
      });
      sfor<9>([&](auto AB) {
        constexpr int a = AB / 3, b = AB % 3;
        inv1d(k[mslot<3>(a, b, 0)], k[mslot<3>(a, b, 1)], k[mslot<3>(a, b, 2)], z, zm);
      });
    } else {
      sfor<3>([&](auto Qd) {
        constexpr int q = Qd;
        inv1d(k[q], k[3 + q], k[6 + q], z, zm);
      });
      sfor<3>([&](auto A) {
        constexpr int a = A;
        inv1d(k[a * 3 + 0], k[a * 3 + 1], k[a * 3 + 2], z, zm);
      });
    }
    sfor<3>([&](auto A) {
      sfor<3>([&](auto B) {
        sfor<(D == 3 ? 3 : 1)>([&](auto C) {
          out[dir_of<Q>(A, B, C)] = k[mslot<D>(A, B, C)];
        });
      });
    });
  } else {
    Macro<R> m;
    m.dens = K::SHIFTED ? rho - R(1) : rho;
    m.ux = ux; m.uy = uy; m.uz = uz;
    constexpr bool PRODUCT = K::OP == M3LB_OP_BGK_PRODUCT || K::OP == M3LB_OP_TRT_PRODUCT;
    sfor<Q>([&](auto I) { out[int(I)] = equilibrium<K, PRODUCT, I>(m); });
  }
}

}  // namespace m3lb
