/*
 * oracle_core.c -- precision-generic body of the CPU restatement.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Compiled twice by oracle/Makefile:
 * -DORC_PREC=64 (REAL = double) and -DORC_PREC=32 (REAL = float), mirroring the
 * reference's configure-time typedef (src/core/Types.hpp:16-20).  Every routine
 * restates the reference routine cited beside it, in the reference's operation
 * order, so an FP64 run reproduces the reference's arithmetic term for term.
 * Node order is ascending, which is Kokkos Serial's order, the order the
 * reference's published results were made in.  The esoteric-pull step is
 * order-independent (src/memory/EsotericPull.hpp:19-21) except where a wall
 * node reads a neighbour's slots (outlets, corner stencils), and the seeding
 * scatter is not at halo layers; ascending order is what both leave.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "oracle.h"
#include "oracle_core.h"

#if ORC_PREC == 64
typedef double REAL;
#define CORE(name) core64_##name
#else
typedef float REAL;
#define CORE(name) core32_##name
#endif

#define R(x) ((REAL)(x))

/* ---------------------------------------------------------------------------
 * Lattices, src/lattice/Lattices.hpp:72-96, 125-159 (D2Q9, D3Q27).
 * ------------------------------------------------------------------------- */
static const int CX9[9] = {0, 1, -1, 0, 0, 1, -1, 1, -1};
static const int CY9[9] = {0, 0, 0, 1, -1, 1, -1, -1, 1};
static const int W9[9] = {16, 4, 4, 4, 4, 1, 1, 1, 1};
static const int CX27[27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1,
                             0, 0, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1};
static const int CY27[27] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0,
                             1, -1, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1};
static const int CZ27[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1,
                             1, -1, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1};
static const int W27[27] = {64, 16, 16, 16, 16, 16, 16, 4, 4, 4, 4, 4, 4, 4,
                            4,  4,  4,  4,  4,  1,  1,  1, 1, 1, 1, 1, 1};

typedef struct {
  int Q, D;
  const int *cx, *cy, *cz, *wn;
  int wden;
} Lat;

static const int ZERO27[27] = {0};

static Lat lat_of(int lattice) {
  Lat l;
  if (lattice == M3LB_D2Q9) {
    l.Q = 9; l.D = 2; l.cx = CX9; l.cy = CY9; l.cz = ZERO27; l.wn = W9; l.wden = 36;
  } else {
    l.Q = 27; l.D = 3; l.cx = CX27; l.cy = CY27; l.cz = CZ27; l.wn = W27; l.wden = 216;
  }
  return l;
}

/* cvel<L>(i, a), Lattices.hpp:178-185 */
static inline int cvel(const Lat* l, int i, int a) {
  return a == 0 ? l->cx[i] : (a == 1 ? l->cy[i] : l->cz[i]);
}
/* weight<L,T>(i), Lattices.hpp:166-169 */
static inline REAL weight(const Lat* l, int i) { return R(l->wn[i]) / R(l->wden); }
/* cs2 = 1/3 and inv_cs2 = 3 on both lattices, Lattices.hpp:170-177 */
#define CS2 (R(1) / R(3))
#define ICS2 (R(3) / R(1))

/* ---------------------------------------------------------------------------
 * Collision policy state (BGK.hpp:43-44, TRT.hpp:46-48,
 * MomentCollision.hpp:89-91, Forcing.hpp:58-61).
 * ------------------------------------------------------------------------- */
typedef struct {
  Lat l;
  int op, shifted, forcing;
  REAL omega, omega_m, omega_b;
  REAL F[3];
} Coll;

static Coll coll_of(const m3lb_config* c) {
  Coll k;
  k.l = lat_of(c->lattice);
  k.op = c->op;
  k.shifted = c->storage == M3LB_SHIFTED;
  k.forcing = c->forcing == M3LB_FORCE_GUO || c->forcing == M3LB_FORCE_FIELD;
  k.omega = R(c->omega);
  k.omega_m = R(c->omega_minus);
  k.omega_b = R(c->omega_bulk);
  k.F[0] = R(c->force[0]);
  k.F[1] = R(c->force[1]);
  k.F[2] = R(c->force[2]);
  return k;
}

/* SecondOrderPhi::phi, Equilibrium.hpp:49-60 */
static inline REAL phi2(const Lat* l, int i, REAL ux, REAL uy, REAL uz) {
  const REAL ics2 = ICS2;
  const REAL cu = R(cvel(l, i, 0)) * ux + R(cvel(l, i, 1)) * uy + R(cvel(l, i, 2)) * uz;
  const REAL uu = ux * ux + uy * uy + uz * uz;
  return ics2 * cu + R(0.5) * ics2 * ics2 * cu * cu - R(0.5) * ics2 * uu;
}

/* ProductFormPhi<D3Q27>::chi, Equilibrium.hpp:114-118 */
static inline REAL chi(int c, REAL u) {
  return (c == 0) ? R(1) - R(1.5) * u * u : R(1) + R(3) * u * (R(c) + u);
}

/* ProductFormPhi<D2Q9>::phi (Equilibrium.hpp:71-90) and <D3Q27>
 * (Equilibrium.hpp:110-124) */
static inline REAL phip(const Lat* l, int i, REAL ux, REAL uy, REAL uz) {
  if (l->D == 3)
    return chi(l->cx[i], ux) * chi(l->cy[i], uy) * chi(l->cz[i], uz) - R(1);
  const REAL cs2v = CS2;
  const REAL cx = R(l->cx[i]);
  const REAL cy = R(l->cy[i]);
  const REAL ax = cx * cx - cs2v;
  const REAL ay = cy * cy - cs2v;
  const REAL o1 = (cx * ux + cy * uy) / cs2v;
  const REAL o2 = (ax * ux * ux + ay * uy * uy + R(2) * cx * cy * ux * uy) /
                  (R(2) * cs2v * cs2v);
  const REAL o3 = (ax * cy * ux * ux * uy + ay * cx * ux * uy * uy) /
                  (R(2) * cs2v * cs2v * cs2v);
  const REAL o4 = (ax * ay * ux * ux * uy * uy) / (R(4) * cs2v * cs2v * cs2v * cs2v);
  return o1 + o2 + o3 + o4;
}

static inline REAL phi_of(const Lat* l, int product, int i, REAL ux, REAL uy, REAL uz) {
  return product ? phip(l, i, ux, uy, uz) : phi2(l, i, ux, uy, uz);
}

/* Equilibrium::eq / eq_shifted, Equilibrium.hpp:33-42 */
static inline REAL eq(const Lat* l, int product, int i, REAL rho, REAL ux, REAL uy,
                      REAL uz) {
  return weight(l, i) * rho * (R(1) + phi_of(l, product, i, ux, uy, uz));
}
static inline REAL eq_shifted(const Lat* l, int product, int i, REAL drho, REAL ux,
                              REAL uy, REAL uz) {
  return weight(l, i) * (drho + (R(1) + drho) * phi_of(l, product, i, ux, uy, uz));
}

/* guo_source_raw, Forcing.hpp:26-38 */
static inline REAL guo_source_raw(const Lat* l, int i, const REAL F[3], REAL ux, REAL uy,
                                  REAL uz) {
  const REAL ics2 = ICS2;
  const REAL cx = R(cvel(l, i, 0));
  const REAL cy = R(cvel(l, i, 1));
  const REAL cz = R(cvel(l, i, 2));
  const REAL cu = cx * ux + cy * uy + cz * uz;
  const REAL bx = (cx - ux) * ics2 + cu * ics2 * ics2 * cx;
  const REAL by = (cy - uy) * ics2 + cu * ics2 * ics2 * cy;
  const REAL bz = (cz - uz) * ics2 + cu * ics2 * ics2 * cz;
  return weight(l, i) * (bx * F[0] + by * F[1] + bz * F[2]);
}

typedef struct {
  REAL dens, ux, uy, uz;
} Macro; /* BGK.hpp:26-29 */

/* Collision::density, BGK.hpp:53-56 (NullCollision: m.dens,
 * tests/test_streaming.cpp:36) */
static inline REAL density(const Coll* c, const Macro* m) {
  if (c->op == M3LB_OP_NULL) return m->dens;
  return c->shifted ? R(1) + m->dens : m->dens;
}

/* BGK/TRT/MomentCollision::macroscopic (BGK.hpp:63-78, TRT.hpp:69-84,
 * MomentCollision.hpp:108-123, identical bodies) with Guo::shift_velocity
 * (Forcing.hpp:65-69).  NullCollision returns {1,0,0,0}
 * (tests/test_streaming.cpp:32-34). */
static inline Macro macroscopic(const Coll* c, const REAL* f) {
  Macro m;
  if (c->op == M3LB_OP_NULL) {
    m.dens = R(1); m.ux = R(0); m.uy = R(0); m.uz = R(0);
    return m;
  }
  const Lat* l = &c->l;
  REAL s = R(0), mx = R(0), my = R(0), mz = R(0);
  for (int i = 0; i < l->Q; ++i) {
    s += f[i];
    mx += f[i] * R(cvel(l, i, 0));
    my += f[i] * R(cvel(l, i, 1));
    mz += f[i] * R(cvel(l, i, 2));
  }
  m.dens = s;
  m.ux = R(0); m.uy = R(0); m.uz = R(0);
  const REAL rho = density(c, &m);
  const REAL ir = R(1) / rho;
  m.ux = mx * ir; m.uy = my * ir; m.uz = mz * ir;
  if (c->forcing) {
    const REAL h = R(0.5) / rho;
    m.ux += h * c->F[0]; m.uy += h * c->F[1]; m.uz += h * c->F[2];
  }
  return m;
}

/* BGK::collide, BGK.hpp:81-92 with Guo::source (Forcing.hpp:76-80) */
static void collide_bgk(const Coll* c, int product, REAL* f, const Macro* m) {
  const Lat* l = &c->l;
  for (int i = 0; i < l->Q; ++i) {
    const REAL feq = c->shifted ? eq_shifted(l, product, i, m->dens, m->ux, m->uy, m->uz)
                                : eq(l, product, i, m->dens, m->ux, m->uy, m->uz);
    f[i] += c->omega * (feq - f[i]);
    if (c->forcing)
      f[i] += (R(1) - R(0.5) * c->omega) * guo_source_raw(l, i, c->F, m->ux, m->uy, m->uz);
  }
}

/* TRT::collide, TRT.hpp:92-127, with the second-order (product = 0) or the
 * product-form (product = 1) equilibrium: TRT<L, Eq, F, S> takes either
 * (TRT.hpp:32-33, eq_of :86-90). */
static void collide_trt(const Coll* c, int product, REAL* f, const Macro* m) {
  const Lat* l = &c->l;
  REAL e[27], sr[27];
  for (int i = 0; i < l->Q; ++i)
    e[i] = c->shifted ? eq_shifted(l, product, i, m->dens, m->ux, m->uy, m->uz)
                      : eq(l, product, i, m->dens, m->ux, m->uy, m->uz);
  if (c->forcing)
    for (int i = 0; i < l->Q; ++i) sr[i] = guo_source_raw(l, i, c->F, m->ux, m->uy, m->uz);
  const REAL op = c->omega, om = c->omega_m;
  const REAL cp = R(1) - R(0.5) * op;
  const REAL cm = R(1) - R(0.5) * om;
  f[0] += op * (e[0] - f[0]);
  if (c->forcing) f[0] += cp * sr[0];
  for (int i = 1; i < l->Q; i += 2) {
    const int j = i + 1;
    const REAL fp = R(0.5) * (f[i] + f[j]);
    const REAL fm = R(0.5) * (f[i] - f[j]);
    const REAL ep = R(0.5) * (e[i] + e[j]);
    const REAL em = R(0.5) * (e[i] - e[j]);
    REAL dp = op * (ep - fp);
    REAL dm = om * (em - fm);
    if (c->forcing) {
      dp += cp * R(0.5) * (sr[i] + sr[j]);
      dm += cm * R(0.5) * (sr[i] - sr[j]);
    }
    f[i] += dp + dm;
    f[j] += dp - dm;
  }
}

/* ---------------------------------------------------------------------------
 * ProductBasis<L>, src/collision/ProductBasis.hpp:55-206.
 * ------------------------------------------------------------------------- */
/* dir_index<L>(a,b,c), ProductBasis.hpp:34-40 */
static int dir_index(const Lat* l, int a, int b, int c) {
  for (int i = 0; i < l->Q; ++i)
    if (cvel(l, i, 0) == a && cvel(l, i, 1) == b && cvel(l, i, 2) == c) return i;
  return -1;
}
/* pi(a,b,c) table, ProductBasis.hpp:118-131 */
static int pi_of(const Lat* l, int a, int b, int c) {
  return dir_index(l, a - 1, b - 1, l->D == 3 ? c - 1 : 0);
}
/* mi(p,q,r), ProductBasis.hpp:61-64 */
static inline int mi(int D, int p, int q, int r) {
  return D == 2 ? (p * 3 + q) : ((p * 3 + q) * 3 + r);
}

/* fwd1d / inv1d, ProductBasis.hpp:137-156 */
static inline void fwd1d(REAL* a, REAL* b, REAL* c, REAL u) {
  const REAL cs2v = CS2;
  const REAL s0 = *a + *b + *c;
  const REAL s1 = *c - *a;
  const REAL s2 = *c + *a;
  *a = s0;
  *b = s1 - u * s0;
  *c = s2 - R(2) * u * s1 + (u * u - cs2v) * s0;
}
static inline void inv1d(REAL* a, REAL* b, REAL* c, REAL u) {
  const REAL cs2v = CS2;
  const REAL m0 = *a, m1 = *b, m2 = *c;
  const REAL s1 = m1 + u * m0;
  const REAL s2 = m2 + R(2) * u * s1 - (u * u - cs2v) * m0;
  *a = R(0.5) * (s2 - s1);
  *b = m0 - s2;
  *c = R(0.5) * (s2 + s1);
}

typedef struct {
  int pi[27];
} PiTab;

static PiTab pitab_of(const Lat* l) {
  PiTab t;
  memset(&t, 0, sizeof t);
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      for (int c = 0; c < (l->D == 3 ? 3 : 1); ++c)
        t.pi[mi(l->D, a, b, c)] = pi_of(l, a, b, c);
  return t;
}

/* to_moments, ProductBasis.hpp:159-181 */
static void to_moments(const Lat* l, const PiTab* t, const REAL* f, const REAL ub[3],
                       REAL* k) {
  const int D = l->D;
  if (D == 2) {
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) k[a * 3 + b] = f[t->pi[a * 3 + b]];
    for (int a = 0; a < 3; ++a) fwd1d(&k[a * 3 + 0], &k[a * 3 + 1], &k[a * 3 + 2], ub[1]);
    for (int q = 0; q < 3; ++q) fwd1d(&k[0 * 3 + q], &k[1 * 3 + q], &k[2 * 3 + q], ub[0]);
  } else {
    for (int n = 0; n < 27; ++n) k[n] = f[t->pi[n]];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        fwd1d(&k[mi(3, a, b, 0)], &k[mi(3, a, b, 1)], &k[mi(3, a, b, 2)], ub[2]);
    for (int a = 0; a < 3; ++a)
      for (int r = 0; r < 3; ++r)
        fwd1d(&k[mi(3, a, 0, r)], &k[mi(3, a, 1, r)], &k[mi(3, a, 2, r)], ub[1]);
    for (int q = 0; q < 3; ++q)
      for (int r = 0; r < 3; ++r)
        fwd1d(&k[mi(3, 0, q, r)], &k[mi(3, 1, q, r)], &k[mi(3, 2, q, r)], ub[0]);
  }
}

/* to_populations, ProductBasis.hpp:183-205 */
static void to_populations(const Lat* l, const PiTab* t, REAL* k, const REAL ub[3],
                           REAL* f) {
  const int D = l->D;
  if (D == 2) {
    for (int q = 0; q < 3; ++q) inv1d(&k[0 * 3 + q], &k[1 * 3 + q], &k[2 * 3 + q], ub[0]);
    for (int a = 0; a < 3; ++a) inv1d(&k[a * 3 + 0], &k[a * 3 + 1], &k[a * 3 + 2], ub[1]);
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) f[t->pi[a * 3 + b]] = k[a * 3 + b];
  } else {
    for (int q = 0; q < 3; ++q)
      for (int r = 0; r < 3; ++r)
        inv1d(&k[mi(3, 0, q, r)], &k[mi(3, 1, q, r)], &k[mi(3, 2, q, r)], ub[0]);
    for (int a = 0; a < 3; ++a)
      for (int r = 0; r < 3; ++r)
        inv1d(&k[mi(3, a, 0, r)], &k[mi(3, a, 1, r)], &k[mi(3, a, 2, r)], ub[1]);
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        inv1d(&k[mi(3, a, b, 0)], &k[mi(3, a, b, 1)], &k[mi(3, a, b, 2)], ub[2]);
    for (int n = 0; n < 27; ++n) f[t->pi[n]] = k[n];
  }
}

/* ---------------------------------------------------------------------------
 * MomentCollision<L,F,S,Central>, src/collision/MomentCollision.hpp:65-330.
 * ------------------------------------------------------------------------- */
typedef struct {
  REAL Qf[3][3], Aw[3][3];
} EqF;

/* eq_factors / ProductBasis::eq_1d, MomentCollision.hpp:128-132,
 * ProductBasis.hpp:107-111 */
static EqF eq_factors(const REAL du[3], const REAL ub[3]) {
  EqF e;
  for (int a = 0; a < 3; ++a) {
    e.Qf[a][0] = R(1); e.Qf[a][1] = du[a]; e.Qf[a][2] = du[a] * du[a];
    e.Aw[a][0] = R(1); e.Aw[a][1] = -ub[a]; e.Aw[a][2] = ub[a] * ub[a];
  }
  return e;
}

/* eq_moment<N>, MomentCollision.hpp:192-201 */
static inline REAL eq_moment(int D, int shifted, int N, REAL rho, const EqF* e) {
  const int p = D == 2 ? N / 3 : N / 9;
  const int q = D == 2 ? N % 3 : (N / 3) % 3;
  const int r = D == 2 ? 0 : N % 3;
  REAL ev = rho * e->Qf[0][p] * e->Qf[1][q];
  REAL wv = e->Aw[0][p] * e->Aw[1][q];
  if (D == 3) { ev *= e->Qf[2][r]; wv *= e->Aw[2][r]; }
  return shifted ? ev - wv : ev;
}
static inline int order_of(int D, int N) {
  return D == 2 ? (N / 3 + N % 3) : (N / 9 + (N / 3) % 3 + N % 3);
}
/* i1 / i2d / i2s, MomentCollision.hpp:335-344 */
static inline int i1(int D, int a) { return mi(D, a == 0, a == 1, D == 3 && a == 2); }
static inline int i2d(int D, int a) {
  return mi(D, 2 * (a == 0), 2 * (a == 1), D == 3 ? 2 * (a == 2) : 0);
}
static inline int i2s(int D, int a, int b) {
  return mi(D, (a == 0 || b == 0), (a == 1 || b == 1), D == 3 && (a == 2 || b == 2));
}

/* MomentCollision::collide, MomentCollision.hpp:234-314 */
static void collide_moment(const Coll* c, const PiTab* t, int central, REAL* f,
                           const Macro* m) {
  const Lat* l = &c->l;
  const int D = l->D, NM = D == 2 ? 9 : 27;
  const REAL rho = density(c, m);
  const REAL u[3] = {m->ux, m->uy, m->uz};
  REAL ub[3], du[3];
  for (int a = 0; a < 3; ++a) {
    ub[a] = central ? u[a] : R(0);
    du[a] = u[a] - ub[a];
  }
  const EqF e = eq_factors(du, ub);
  REAL k[27];
  to_moments(l, t, f, ub, k);

  REAL F[3] = {R(0), R(0), R(0)};
  if (c->forcing) { F[0] = c->F[0]; F[1] = c->F[1]; F[2] = c->F[2]; }
  const REAL w2 = R(1) - R(0.5) * c->omega;
  const REAL wb = c->omega_b < R(0) ? c->omega : c->omega_b; /* bulk(), :99-101 */
  const REAL omega = c->omega;

  if (c->forcing)
    for (int a = 0; a < D; ++a) k[i1(D, a)] += F[a];

  {
    REAL d[3], ee[3], g[3];
    REAL tr = R(0), tre = R(0), trg = R(0);
    for (int a = 0; a < D; ++a) {
      d[a] = k[i2d(D, a)];
      ee[a] = eq_moment(D, c->shifted, i2d(D, a), rho, &e);
      g[a] = c->forcing ? R(2) * F[a] * du[a] : R(0);
      tr += d[a]; tre += ee[a]; trg += g[a];
    }
    const REAL invD = R(1) / R(D);
    const REAL tr_post = (R(1) - wb) * tr + wb * tre + (R(1) - R(0.5) * wb) * trg;
    for (int a = 0; a < D; ++a)
      k[i2d(D, a)] = (R(1) - omega) * (d[a] - tr * invD) + omega * (ee[a] - tre * invD) +
                     w2 * (g[a] - trg * invD) + tr_post * invD;
    for (int a = 0; a < D; ++a)
      for (int b = a + 1; b < D; ++b) {
        const int id = i2s(D, a, b);
        const REAL ge = c->forcing ? (F[b] * du[a] + F[a] * du[b]) : R(0);
        k[id] = (R(1) - omega) * k[id] + omega * eq_moment(D, c->shifted, id, rho, &e) +
                w2 * ge;
      }
  }
  for (int N = 0; N < NM; ++N)
    if (order_of(D, N) >= 3) k[N] = eq_moment(D, c->shifted, N, rho, &e);
  to_populations(l, t, k, ub, f);
}

static void collide(const Coll* c, const PiTab* t, REAL* f, const Macro* m) {
  switch (c->op) {
    case M3LB_OP_NULL: return;
    case M3LB_OP_BGK: collide_bgk(c, 0, f, m); return;
    case M3LB_OP_BGK_PRODUCT: collide_bgk(c, 1, f, m); return;
    case M3LB_OP_TRT: collide_trt(c, 0, f, m); return;
    case M3LB_OP_TRT_PRODUCT: collide_trt(c, 1, f, m); return;
    case M3LB_OP_MRT: collide_moment(c, t, 0, f, m); return;
    default: collide_moment(c, t, 1, f, m); return;
  }
}

/* seed_value for every operator: BGK.hpp:95-99, TRT.hpp:129-133,
 * MomentCollision.hpp:320-330 (all Q values at once),
 * NullCollision tests/test_streaming.cpp:37. */
static void seed_all(const Coll* c, const PiTab* t, REAL rho, REAL ux, REAL uy, REAL uz,
                     REAL* out) {
  const Lat* l = &c->l;
  if (c->op == M3LB_OP_NULL) {
    for (int i = 0; i < l->Q; ++i) out[i] = rho;
    return;
  }
  if (c->op == M3LB_OP_MRT || c->op == M3LB_OP_CM) {
    const int D = l->D, NM = D == 2 ? 9 : 27;
    const REAL ub[3] = {R(0), R(0), R(0)};
    const REAL du[3] = {ux, uy, uz};
    const EqF e = eq_factors(du, ub);
    REAL k[27];
    for (int N = 0; N < NM; ++N) k[N] = eq_moment(D, c->shifted, N, rho, &e);
    to_populations(l, t, k, ub, out);
    return;
  }
  const int product = c->op == M3LB_OP_BGK_PRODUCT || c->op == M3LB_OP_TRT_PRODUCT;
  for (int i = 0; i < l->Q; ++i)
    out[i] = c->shifted ? eq_shifted(l, product, i, rho - R(1), ux, uy, uz)
                        : eq(l, product, i, rho, ux, uy, uz);
}

/* ---------------------------------------------------------------------------
 * Domain, src/grid/Domain.hpp:37-124.
 * ------------------------------------------------------------------------- */
typedef struct {
  int64_t nx, ny, nz, hx, hy, hz, sx, sy, sz, np;
  int per[3];
} Dom;

static int64_t halo_for(int64_t n, int per) { /* Domain.hpp:104-108, mpi_halo = 0 */
  if (n <= 1) return 0;
  if (per) return 0;
  return 1;
}

static Dom dom_of(const m3lb_config* c) {
  Dom d;
  d.nx = c->nx; d.ny = c->ny; d.nz = c->nz;
  d.per[0] = c->periodic[0] != 0; d.per[1] = c->periodic[1] != 0; d.per[2] = c->periodic[2] != 0;
  d.hx = halo_for(d.nx, d.per[0]);
  d.hy = halo_for(d.ny, d.per[1]);
  d.hz = halo_for(d.nz, d.per[2]);
  d.sx = d.nx + 2 * d.hx; d.sy = d.ny + 2 * d.hy; d.sz = d.nz + 2 * d.hz;
  d.np = d.sx * d.sy * d.sz;
  return d;
}

static inline int64_t idp(const Dom* d, int64_t px, int64_t py, int64_t pz) {
  return (pz * d->sy + py) * d->sx + px;
}
static inline void coords(const Dom* d, int64_t n, int64_t* px, int64_t* py, int64_t* pz) {
  *px = n % d->sx;
  *py = (n / d->sx) % d->sy;
  *pz = n / (d->sx * d->sy);
}
static inline int64_t shift(int64_t v, int64_t h, int64_t n, int64_t s, int per) {
  if (per) {
    if (v < h) return v + n;
    if (v >= h + n) return v - n;
    return v;
  }
  if (v < 0) return 0;
  if (v >= s) return s - 1;
  return v;
}

/* fill_neighbours<L, First, Stride>, Domain.hpp:82-101 */
static void fill_neighbours(const Dom* d, const Lat* l, int first, int stride, int64_t n,
                            int64_t* j) {
  int64_t x, y, z;
  coords(d, n, &x, &y, &z);
  int core = (x > 0 && x < d->sx - 1) && (y > 0 && y < d->sy - 1);
  if (l->D == 3) core = core && (z > 0 && z < d->sz - 1);
  if (core) {
    for (int i = first; i < l->Q; i += stride)
      j[i] = n + cvel(l, i, 0) + cvel(l, i, 1) * d->sx + cvel(l, i, 2) * d->sx * d->sy;
  } else {
    for (int i = first; i < l->Q; i += stride)
      j[i] = idp(d, shift(x + cvel(l, i, 0), d->hx, d->nx, d->sx, d->per[0]),
                 shift(y + cvel(l, i, 1), d->hy, d->ny, d->sy, d->per[1]),
                 shift(z + cvel(l, i, 2), d->hz, d->nz, d->sz, d->per[2]));
  }
}

/* ---------------------------------------------------------------------------
 * Solver state (FluidSolver.hpp:42-69, 1255-1353).
 * ------------------------------------------------------------------------- */
typedef struct {
  m3lb_config cfg;
  Coll c;
  PiTab pt;
  Dom d;
  int streaming; /* 0 esoteric pull, 1 two-lattice */
  int Q;
  uint8_t* flags;
  REAL* f;       /* (n, q), q fastest: the host default layout, Types.hpp:35-43 */
  REAL* f2;      /* second lattice (TwoLattice only) */
  int64_t* active;
  int64_t n_active;
  uint64_t t;
  int nthreads;
  REAL* ff; /* FORCE_FIELD: fx, fy, fz planes of n_padded values, or NULL */
  /* regularised walls (FluidSolver.hpp:1259-1290), per padded node */
  uint8_t* bc_nrm;
  uint16_t* bc_tag;
  uint32_t* bc_unk;
  uint8_t* bc_ext;
  REAL* bc_rho;
  double* wtab; /* (ux, uy, uz, rho) per table entry */
  int ntab;
  int has_corners, fd_corners;
  int64_t* walls;
  int64_t n_walls;
  /* outflow codes: donor node and its axis face (NrmOutFree), the table as
   * given (set_wall_velocity_scale), the per-node wall velocity field */
  int64_t* bc_don;
  uint8_t* bc_onrm;
  double* wtab0;
  REAL* wall_un; /* ux, uy, uz planes of n_padded values, or NULL */
  REAL un_scale;
  int force_bc, out_order;
  REAL outflow_rho;
  /* specular walls (FluidSolver.hpp:1262-1270): ghost normal, on-node faces */
  uint8_t* spec_nrm;
  uint8_t* spec_faces;
  int mirror[3][27]; /* MirrorTable<L>::m (Specular.hpp:104-121) */
} S;

/* The collision policy at node n: for a per-node force (a forcing policy
 * with at(n, F), e.g. validation/forcing_3d.cpp:64-91) a copy carrying that
 * node's F; Guo's terms then use it as they use the uniform one
 * (Forcing.hpp:489-496). */
static inline const Coll* coll_at(const S* s, int64_t n, Coll* tmp) {
  if (!s->ff) return &s->c;
  *tmp = s->c;
  tmp->F[0] = s->ff[n];
  tmp->F[1] = s->ff[s->d.np + n];
  tmp->F[2] = s->ff[2 * s->d.np + n];
  return tmp;
}

#define F(s, n, q) ((s)->f[(n) * (int64_t)(s)->Q + (q)])

/* EsotericPull::Access<P>::load, EsotericPull.hpp:65-72 */
static inline REAL ep_load(const S* s, int P, const int64_t* j, int64_t n, int i) {
  if (i == 0) return F(s, n, 0);
  if (P == 0) return (i & 1) ? F(s, n, i) : F(s, j[i - 1], i);
  return (i & 1) ? F(s, n, i + 1) : F(s, j[i - 1], i - 1);
}
/* EsotericPull::Access<P>::scatter, EsotericPull.hpp:104-111 */
static inline void ep_scatter(S* s, int P, const int64_t* j, int64_t n, int i, REAL v) {
  if (i == 0) { F(s, n, 0) = v; return; }
  if (P == 0) {
    if (i & 1) F(s, n, i) = v; else F(s, j[i - 1], i) = v;
  } else {
    if (i & 1) F(s, n, i + 1) = v; else F(s, j[i - 1], i - 1) = v;
  }
}

/* TwoLattice::Access::load / scatter, TwoLattice.hpp:39-58 (src = f) */
static inline int opp(int i) { return i == 0 ? 0 : (i & 1 ? i + 1 : i - 1); }
static inline REAL tl_load(const S* s, const int64_t* j, int64_t n, int i) {
  return F(s, i == 0 ? n : j[opp(i)], i);
}

/* rebuild_lists, FluidSolver.hpp:157-177.  The RegWall nodes go to their own
 * list (they are updated before the rest of the step, see run_step). */
static void rebuild_lists(S* s) {
  int64_t k = 0, w = 0;
  for (int64_t n = 0; n < s->d.np; ++n) {
    const uint8_t fl = s->flags[n];
    if (fl == M3LB_REGWALL) s->walls[w++] = n;
    if (fl == M3LB_EXCLUDED || fl == M3LB_REGWALL) continue;
    if (fl == M3LB_SOLID && s->streaming == 0) continue;
    s->active[k++] = n;
  }
  s->n_active = k;
  s->n_walls = w;
}

/* A node that holds a fluid state (FluidSolver.hpp:557, 575, 1054, 1215). */
static inline int fluid_like(uint8_t fl) {
  return fl == M3LB_FLUID || fl == M3LB_REGWALL || fl == M3LB_SPECNODE;
}

/* MirrorTable<L> (Specular.hpp:104-121): m[a][i] is the direction of c_i with
 * component a negated. */
static void build_mirror(S* s) {
  const Lat* l = &s->c.l;
  for (int a = 0; a < 3; ++a)
    for (int i = 0; i < l->Q; ++i) {
      int c[3] = {cvel(l, i, 0), cvel(l, i, 1), cvel(l, i, 2)};
      c[a] = -c[a];
      s->mirror[a][i] = i;
      for (int j = 0; j < l->Q; ++j)
        if (cvel(l, j, 0) == c[0] && cvel(l, j, 1) == c[1] && cvel(l, j, 2) == c[2]) {
          s->mirror[a][i] = j;
          break;
        }
    }
}

/* face_sign (Specular.hpp:253-258): outward sign of the face mask on axis a */
static inline int face_sign(uint8_t faces, int a) {
  return (faces & (1u << (2 * a))) ? 1 : ((faces & (1u << (2 * a + 1))) ? -1 : 0);
}

/* mirror_unknowns_faces (Specular.hpp:274-289): every direction pointing into
 * the domain on a masked axis takes the value of its image through the
 * composition of those axis mirrors. */
static void mirror_unknowns_faces(const S* s, REAL* f, uint8_t faces) {
  REAL h[27];
  for (int i = 0; i < s->Q; ++i) h[i] = f[i];
  for (int i = 0; i < s->Q; ++i) {
    int j = i, unknown = 0;
    for (int a = 0; a < 3; ++a) {
      const int sg = face_sign(faces, a);
      if (sg != 0 && cvel(&s->c.l, i, a) * sg < 0) {
        j = s->mirror[a][j];
        unknown = 1;
      }
    }
    if (unknown) f[i] = h[j];
  }
}

static int is_interior(const Dom* d, int64_t px, int64_t py, int64_t pz) {
  return px >= d->hx && px < d->hx + d->nx && py >= d->hy && py < d->hy + d->ny &&
         pz >= d->hz && pz < d->hz + d->nz;
}

void* CORE(create)(const m3lb_config* cfg, int streaming) {
  S* s = (S*)calloc(1, sizeof(S));
  s->cfg = *cfg;
  s->c = coll_of(cfg);
  s->pt = pitab_of(&s->c.l);
  s->d = dom_of(cfg);
  s->streaming = streaming;
  s->Q = s->c.l.Q;
  s->nthreads = 1;
  s->flags = (uint8_t*)malloc((size_t)s->d.np);
  /* Views are zero-initialised (Kokkos default), so are these. */
  s->f = (REAL*)calloc((size_t)(s->d.np * s->Q), sizeof(REAL));
  if (streaming) s->f2 = (REAL*)calloc((size_t)(s->d.np * s->Q), sizeof(REAL));
  s->active = (int64_t*)malloc(sizeof(int64_t) * (size_t)(s->d.np > 0 ? s->d.np : 1));
  if (cfg->forcing == M3LB_FORCE_FIELD) s->ff = (REAL*)calloc((size_t)(3 * s->d.np), sizeof(REAL));
  s->fd_corners = 1; /* fd_corners_ = has_shear_omega<Collision> (FluidSolver.hpp:1287) */
  s->force_bc = 1;   /* force_bc_ (:1288) */
  s->out_order = 1;  /* out_order_ (:1291) */
  s->outflow_rho = R(1);
  s->un_scale = R(1);
  s->walls = (int64_t*)malloc(sizeof(int64_t) * (size_t)(s->d.np > 0 ? s->d.np : 1));
  build_mirror(s);
  for (int64_t n = 0; n < s->d.np; ++n) {
    int64_t px, py, pz;
    coords(&s->d, n, &px, &py, &pz);
    s->flags[n] = is_interior(&s->d, px, py, pz) ? M3LB_FLUID : M3LB_EXCLUDED;
  }
  rebuild_lists(s);
  return s;
}

void CORE(destroy)(void* p) {
  S* s = (S*)p;
  if (!s) return;
  free(s->flags); free(s->f); free(s->f2); free(s->active); free(s->ff);
  free(s->bc_nrm); free(s->bc_tag); free(s->bc_unk); free(s->bc_ext); free(s->bc_rho);
  free(s->wtab); free(s->walls);
  free(s->bc_don); free(s->bc_onrm); free(s->wtab0); free(s->wall_un);
  free(s->spec_nrm); free(s->spec_faces);
  free(s);
}

void CORE(set_threads)(void* p, int n) { ((S*)p)->nthreads = n > 0 ? n : 1; }

void CORE(domain)(const void* p, int64_t out[10]) {
  const S* s = (const S*)p;
  const Dom* d = &s->d;
  out[0] = d->nx; out[1] = d->ny; out[2] = d->nz; out[3] = d->hx; out[4] = d->hy;
  out[5] = d->hz; out[6] = d->sx; out[7] = d->sy; out[8] = d->sz; out[9] = d->np;
}

/* set_geometry, FluidSolver.hpp:74-82 */
void CORE(set_geometry)(void* p, const uint8_t* fl) {
  S* s = (S*)p;
  const Dom* d = &s->d;
  int64_t k = 0;
  for (int64_t z = 0; z < d->nz; ++z)
    for (int64_t y = 0; y < d->ny; ++y)
      for (int64_t x = 0; x < d->nx; ++x)
        s->flags[idp(d, x + d->hx, y + d->hy, z + d->hz)] = fl[k++];
  rebuild_lists(s);
}

/* scatter<P>(fn), FluidSolver.hpp:1231-1241: ascending n over the padded box */
static void scatter_fn(S* s, const REAL* vals /* np*Q logical */) {
  const int P = (int)(s->t % 2);
  int64_t j[27];
  if (s->streaming == 0) {
    for (int64_t n = 0; n < s->d.np; ++n) {
      fill_neighbours(&s->d, &s->c.l, 1, 2, n, j);
      for (int i = 0; i < s->Q; ++i) ep_scatter(s, P, j, n, i, vals[n * s->Q + i]);
    }
  } else {
    for (int64_t n = 0; n < s->d.np; ++n) {
      fill_neighbours(&s->d, &s->c.l, 1, 1, n, j);
      for (int i = 0; i < s->Q; ++i) F(s, i == 0 ? n : j[opp(i)], i) = vals[n * s->Q + i];
    }
  }
}

/* initialize / initialize_field, FluidSolver.hpp:553-579 */
static void init_common(S* s, int field, REAL rho0, REAL ux0, REAL uy0, REAL uz0,
                        const double* rho, const double* ux, const double* uy,
                        const double* uz) {
  s->t = 0;
  const int Q = s->Q;
  REAL* vals = (REAL*)malloc(sizeof(REAL) * (size_t)(s->d.np * Q));
  REAL rest[27];
  if (field)
    seed_all(&s->c, &s->pt, R(1), R(0), R(0), R(0), rest); /* FlowState{} */
  else
    seed_all(&s->c, &s->pt, rho0, R(0), R(0), R(0), rest);
  for (int64_t n = 0; n < s->d.np; ++n) {
    if (fluid_like(s->flags[n])) {
      if (field)
        seed_all(&s->c, &s->pt, R(rho[n]), R(ux[n]), R(uy[n]), R(uz[n]), &vals[n * Q]);
      else
        seed_all(&s->c, &s->pt, rho0, ux0, uy0, uz0, &vals[n * Q]);
    } else {
      memcpy(&vals[n * Q], rest, sizeof(REAL) * (size_t)Q);
    }
  }
  scatter_fn(s, vals);
  free(vals);
}

void CORE(initialize)(void* p, double rho0, double ux, double uy, double uz) {
  init_common((S*)p, 0, R(rho0), R(ux), R(uy), R(uz), NULL, NULL, NULL, NULL);
}
void CORE(initialize_field)(void* p, const double* rho, const double* ux, const double* uy,
                            const double* uz) {
  init_common((S*)p, 1, R(0), R(0), R(0), R(0), rho, ux, uy, uz);
}

void CORE(seed_populations)(void* p, const double* in) {
  S* s = (S*)p;
  REAL* vals = (REAL*)malloc(sizeof(REAL) * (size_t)(s->d.np * s->Q));
  for (int64_t k = 0; k < s->d.np * s->Q; ++k) vals[k] = R(in[k]);
  scatter_fn(s, vals);
  free(vals);
}

/* gather<P>, FluidSolver.hpp:1243-1253 */
void CORE(gather_populations)(void* p, double* out) {
  S* s = (S*)p;
  const int P = (int)(s->t % 2);
  int64_t j[27];
  for (int64_t n = 0; n < s->d.np; ++n) {
    if (s->streaming == 0) {
      fill_neighbours(&s->d, &s->c.l, 1, 2, n, j);
      for (int i = 0; i < s->Q; ++i) out[n * s->Q + i] = (double)ep_load(s, P, j, n, i);
    } else {
      fill_neighbours(&s->d, &s->c.l, 1, 1, n, j);
      for (int i = 0; i < s->Q; ++i) out[n * s->Q + i] = (double)tl_load(s, j, n, i);
    }
  }
}

void CORE(storage)(void* p, double* out) {
  S* s = (S*)p;
  for (int64_t k = 0; k < s->d.np * s->Q; ++k) out[k] = (double)s->f[k];
}

/* ---------------------------------------------------------------------------
 * Regularised velocity walls (Latt et al. 2008), src/boundary/Regularized.hpp
 * with FluidSolver.hpp:372-460 (setup), 663-710 (corner density),
 * 793-1020 (the wall path of run_step) and 1076-1185 (macro_kernel):
 * velocity walls on straight faces and corners, and the three outflow codes.
 * ------------------------------------------------------------------------- */
#define NRM_CORNER 7
#define NRM_OUT_XP 8
#define NRM_OUT_FREE 9
#define NRM_OUT_EQ 10

static void normal_of(int code, int n[3]) {
  n[0] = n[1] = n[2] = 0;
  switch (code) {
    case 1: n[0] = 1; break;
    case 2: n[0] = -1; break;
    case 3: n[1] = 1; break;
    case 4: n[1] = -1; break;
    case 5: n[2] = 1; break;
    case 6: n[2] = -1; break;
    case 8: n[0] = 1; break; /* NrmOutXp: outward normal, for the unknown set */
    default: break;
  }
}

/* Regularized::normal_velocity: the density closure inverted, rho known */
static REAL bc_normal_velocity(const Lat* l, const REAL* f, const int nrm[3], REAL rho) {
  REAL out = R(0), tang = R(0);
  for (int i = 0; i < l->Q; ++i) {
    const int cn = cvel(l, i, 0) * nrm[0] + cvel(l, i, 1) * nrm[1] + cvel(l, i, 2) * nrm[2];
    if (cn > 0) out += f[i];
    else if (cn == 0) tang += f[i];
  }
  return (R(2) * out + tang) / rho - R(1);
}

/* The wall velocity of node n: the per-node field when one is set
 * (set_wall_velocity_field), else its table entry (FluidSolver.hpp:686-689). */
static void wall_vel(const S* s, int64_t n, REAL uw[3]) {
  if (s->wall_un) {
    for (int a = 0; a < 3; ++a) uw[a] = s->un_scale * s->wall_un[a * s->d.np + n];
    return;
  }
  const double* w = &s->wtab[4 * (int64_t)s->bc_tag[n]];
  uw[0] = R(w[0]); uw[1] = R(w[1]); uw[2] = R(w[2]);
}

/* EquilibriumOf<Collision>: the operator's own equilibrium for BGK and TRT,
 * the product form for the moment operators (Regularized.hpp:14-19,
 * Equilibrium.hpp:130-136). */
static int wall_product(const Coll* c) { return c->op != M3LB_OP_BGK && c->op != M3LB_OP_TRT; }

/* Regularized::density */
static REAL bc_density(const Lat* l, const REAL* f, const int nrm[3], const REAL u[3]) {
  REAL out = R(0), tang = R(0);
  for (int i = 0; i < l->Q; ++i) {
    const int cn = cvel(l, i, 0) * nrm[0] + cvel(l, i, 1) * nrm[1] + cvel(l, i, 2) * nrm[2];
    if (cn > 0) out += f[i];
    else if (cn == 0) tang += f[i];
  }
  const REAL uperp = u[0] * R(nrm[0]) + u[1] * R(nrm[1]) + u[2] * R(nrm[2]);
  return (R(2) * out + tang) / (R(1) + uperp);
}

static inline REAL qpi(const Lat* l, int i, const REAL Pi[6]) {
  const REAL cs2v = CS2;
  const REAL cx = R(cvel(l, i, 0)), cy = R(cvel(l, i, 1)), cz = R(cvel(l, i, 2));
  return (cx * cx - cs2v) * Pi[0] + (cy * cy - cs2v) * Pi[1] +
         ((l->D == 3) ? (cz * cz - cs2v) * Pi[2] : R(0)) +
         R(2) * (cx * cy * Pi[3] + ((l->D == 3) ? (cx * cz * Pi[4] + cy * cz * Pi[5]) : R(0)));
}

/* Regularized::apply_with_stress */
static void bc_apply_with_stress(const Coll* c, REAL* f, REAL rho, const REAL u[3],
                                 const REAL Pi[6]) {
  const Lat* l = &c->l;
  const REAL cs2v = CS2, cs4v = cs2v * cs2v;
  const int pr = wall_product(c);
  for (int i = 0; i < l->Q; ++i)
    f[i] = eq(l, pr, i, rho, u[0], u[1], u[2]) + weight(l, i) / (R(2) * cs4v) * qpi(l, i, Pi);
}

/* Regularized::stress_from_gradient */
static void bc_stress_from_gradient(REAL rho, REAL omega, const REAL g[3][3], REAL Pi[6]) {
  const REAL k = -R(2) * CS2 * rho / omega;
  Pi[0] = k * g[0][0];
  Pi[1] = k * g[1][1];
  Pi[2] = k * g[2][2];
  Pi[3] = k * R(0.5) * (g[0][1] + g[1][0]);
  Pi[4] = k * R(0.5) * (g[0][2] + g[2][0]);
  Pi[5] = k * R(0.5) * (g[1][2] + g[2][1]);
}

/* Regularized::apply: unknown directions from the bounced non-equilibrium
 * part (with the Guo odd part when a force is given), then every direction
 * rebuilt from (rho, u, Pi^(1)). */
static void bc_apply(const Coll* c, REAL* f, REAL rho, const REAL u[3], uint32_t unk,
                     const REAL* Fv) {
  const Lat* l = &c->l;
  const REAL cs2v = CS2, cs4v = cs2v * cs2v;
  const int pr = wall_product(c);
  REAL feq[27];
  for (int i = 0; i < l->Q; ++i) feq[i] = eq(l, pr, i, rho, u[0], u[1], u[2]);
  const REAL hodd = R(0.5) / cs2v;
  for (int i = 0; i < l->Q; ++i) {
    if (!(unk & (1u << i))) continue;
    const REAL cF = Fv ? (R(cvel(l, i, 0)) * Fv[0] + R(cvel(l, i, 1)) * Fv[1] +
                          R(cvel(l, i, 2)) * Fv[2])
                       : R(0);
    const REAL odd = weight(l, i) * hodd * cF;
    f[i] = (unk & (1u << opp(i))) ? feq[i] - odd
                                  : feq[i] + (f[opp(i)] - feq[opp(i)]) - R(2) * odd;
  }
  REAL Pi[6] = {0, 0, 0, 0, 0, 0};
  for (int i = 0; i < l->Q; ++i) {
    const REAL d = f[i] - feq[i];
    const REAL cx = R(cvel(l, i, 0)), cy = R(cvel(l, i, 1)), cz = R(cvel(l, i, 2));
    Pi[0] += d * cx * cx; Pi[1] += d * cy * cy; Pi[2] += d * cz * cz;
    Pi[3] += d * cx * cy; Pi[4] += d * cx * cz; Pi[5] += d * cy * cz;
  }
  for (int i = 0; i < l->Q; ++i) {
    const REAL cx = R(cvel(l, i, 0)), cy = R(cvel(l, i, 1)), cz = R(cvel(l, i, 2));
    const REAL cF = Fv ? (cx * Fv[0] + cy * Fv[1] + cz * Fv[2]) : R(0);
    f[i] = feq[i] - weight(l, i) * hodd * cF + weight(l, i) / (R(2) * cs4v) * qpi(l, i, Pi);
  }
}

/* Incoming set of node n at parity P, unshifted to f. */
static void load_raw(const S* s, int P, int64_t n, REAL* f) {
  int64_t j[27];
  fill_neighbours(&s->d, &s->c.l, 1, 2, n, j);
  for (int i = 0; i < s->Q; ++i) f[i] = ep_load(s, P, j, n, i);
  if (s->c.shifted)
    for (int i = 0; i < s->Q; ++i) f[i] += weight(&s->c.l, i);
}

static inline int64_t interior_id(const Dom* d, int64_t x, int64_t y, int64_t z) {
  return ((z + d->hz) * d->sy + (y + d->hy)) * d->sx + x + d->hx;
}

/* corner_density<P> (FluidSolver.hpp:663-710): rho at every straight-wall
 * node from its incoming set, then each corner extrapolated along the axis
 * chosen at setup (2 rho(n1) - rho(n2)), or 1 without a stencil. */
static void corner_density(S* s, int P) {
  REAL f[27];
  for (int64_t k = 0; k < s->n_walls; ++k) {
    const int64_t n = s->walls[k];
    const int code = s->bc_nrm[n];
    if (code == NRM_CORNER || code == 0) continue;
    load_raw(s, P, n, f);
    int nrm[3];
    normal_of(code, nrm);
    REAL uw[3];
    wall_vel(s, n, uw);
    const REAL trho = R(s->wtab[4 * (int64_t)s->bc_tag[n] + 3]);
    s->bc_rho[n] = (code == NRM_OUT_XP || code == NRM_OUT_EQ) ? trho
                 : (code == NRM_OUT_FREE) ? s->outflow_rho
                                          : bc_density(&s->c.l, f, nrm, uw);
  }
  static const int dirs[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
  for (int64_t k = 0; k < s->n_walls; ++k) {
    const int64_t n = s->walls[k];
    if (s->bc_nrm[n] != NRM_CORNER) continue;
    const int e = s->bc_ext[n];
    if (e == 0) { s->bc_rho[n] = R(1); continue; }
    int64_t px, py, pz;
    coords(&s->d, n, &px, &py, &pz);
    const int64_t x = px - s->d.hx, y = py - s->d.hy, z = pz - s->d.hz;
    const int* v = dirs[e - 1];
    const int64_t n1 = interior_id(&s->d, x + v[0], y + v[1], z + v[2]);
    const int64_t n2 = interior_id(&s->d, x + 2 * v[0], y + 2 * v[1], z + 2 * v[2]);
    s->bc_rho[n] = R(2) * s->bc_rho[n1] - s->bc_rho[n2];
  }
}

/* velocity at node m for the finite-difference corner stress: the wall
 * velocity of a wall node, else from m's incoming set. */
static void fd_vel(const S* s, int P, int64_t m, REAL out[3]) {
  if (s->flags[m] == M3LB_REGWALL) {
    const double* w = &s->wtab[4 * (int64_t)s->bc_tag[m]];
    out[0] = R(w[0]); out[1] = R(w[1]); out[2] = R(w[2]);
    return;
  }
  REAL g[27];
  load_raw(s, P, m, g);
  REAL sm = 0, mx = 0, my = 0, mz = 0;
  for (int i = 0; i < s->Q; ++i) {
    sm += g[i];
    mx += g[i] * R(cvel(&s->c.l, i, 0));
    my += g[i] * R(cvel(&s->c.l, i, 1));
    mz += g[i] * R(cvel(&s->c.l, i, 2));
  }
  const REAL ir = R(1) / sm;
  out[0] = mx * ir; out[1] = my * ir; out[2] = mz * ir;
}

/* The mean velocity of node m's incoming set (the upstream and donor
 * velocities of the outflow codes, FluidSolver.hpp:935-950).  guard: return 0
 * for a non-positive sum instead of dividing. */
static void mean_velocity(const S* s, int P, int64_t m, REAL* sm_out, REAL v[3], int guard) {
  REAL g[27];
  load_raw(s, P, m, g);
  REAL sm = R(0);
  v[0] = v[1] = v[2] = R(0);
  for (int i = 0; i < s->Q; ++i) {
    sm += g[i];
    for (int a = 0; a < 3; ++a) v[a] += g[i] * R(cvel(&s->c.l, i, a));
  }
  const REAL ir = guard ? ((sm > R(0)) ? R(1) / sm : R(0)) : R(1) / sm;
  for (int a = 0; a < 3; ++a) v[a] *= ir;
  if (sm_out) *sm_out = sm;
}

/* NrmOutXp's velocity (FluidSolver.hpp:912-951, 1143-1180): rho imposed, the
 * normal component from the inverted closure on the node's own set f, the
 * tangential ones zero-gradient (order 1) or linearly extrapolated (order 2)
 * from the upstream neighbours along -x. */
static void outflow_xp_velocity(const S* s, int P, int64_t n, const REAL* f, REAL rw, REAL ur[3]) {
  int64_t px, py, pz;
  coords(&s->d, n, &px, &py, &pz);
  const int64_t bx = px - s->d.hx, by = py - s->d.hy, bz = pz - s->d.hz;
  REAL ut[3];
  mean_velocity(s, P, interior_id(&s->d, bx - 1, by, bz), NULL, ut, 0);
  if (s->out_order >= 2) {
    REAL v2[3];
    mean_velocity(s, P, interior_id(&s->d, bx - 2, by, bz), NULL, v2, 0);
    for (int a = 0; a < 3; ++a) ut[a] = R(2) * ut[a] - v2[a];
  }
  const int nrm[3] = {1, 0, 0};
  const REAL un = bc_normal_velocity(&s->c.l, f, nrm, rw);
  REAL utn = R(0);
  for (int a = 0; a < 3; ++a) utn += ut[a] * R(nrm[a]);
  for (int a = 0; a < 3; ++a) ur[a] = ut[a] + (un - utn) * R(nrm[a]);
}

/* The RegWall node update of run_step (FluidSolver.hpp:793-1033): velocity
 * walls, corners and the outflow codes; rebuild, collide, store (the
 * population-extrapolation and equilibrium outlets store without colliding). */
static void wall_update(S* s, int P, int64_t n) {
  const Lat* l = &s->c.l;
  const int Q = s->Q;
  int64_t j[27];
  REAL f[27];
  fill_neighbours(&s->d, l, 1, 2, n, j);
  f[0] = F(s, n, 0);
  for (int i = 1; i < Q; i += 2) {
    if (P == 0) { f[i] = F(s, n, i); f[i + 1] = F(s, j[i], i + 1); }
    else        { f[i] = F(s, n, i + 1); f[i + 1] = F(s, j[i], i); }
  }
  if (s->c.shifted)
    for (int i = 0; i < Q; ++i) f[i] += weight(l, i);
  const int code = s->bc_nrm[n];
  int nrm[3];
  normal_of(code, nrm);
  const int tg = s->bc_tag[n];
  REAL uw[3];
  wall_vel(s, n, uw);
  Coll cn;
  const Coll* cp = coll_at(s, n, &cn);
  REAL Fv[3] = {R(0), R(0), R(0)};
  if (cp->forcing && s->force_bc) { Fv[0] = cp->F[0]; Fv[1] = cp->F[1]; Fv[2] = cp->F[2]; }
  REAL ur[3] = {uw[0], uw[1], uw[2]};
  REAL rw;
  int collide_it = 1;
  if (code == NRM_OUT_FREE) {
    /* the donor's distribution rescaled to the imposed density (:843-880) */
    REAL g[27];
    load_raw(s, P, s->bc_don[n], g);
    REAL sm = R(0);
    for (int i = 0; i < Q; ++i) sm += g[i];
    if (sm > R(0)) {
      const REAL sc = s->outflow_rho / sm;
      for (int i = 0; i < Q; ++i) f[i] = g[i] * sc;
    }
    collide_it = 0;
  } else if (code == NRM_OUT_EQ) {
    /* equilibrium at the imposed rho and the upstream velocity (:881-911) */
    const REAL rw2 = R(s->wtab[4 * (int64_t)tg + 3]);
    int64_t px, py, pz;
    coords(&s->d, n, &px, &py, &pz);
    REAL mu[3];
    mean_velocity(s, P, interior_id(&s->d, px - s->d.hx - 1, py - s->d.hy, pz - s->d.hz), NULL,
                  mu, 1);
    for (int i = 0; i < Q; ++i) f[i] = eq(l, wall_product(&s->c), i, rw2, mu[0], mu[1], mu[2]);
    collide_it = 0;
  } else if (code == NRM_OUT_XP) {
    rw = R(s->wtab[4 * (int64_t)tg + 3]);
    outflow_xp_velocity(s, P, n, f, rw, ur);
  } else {
    rw = (code == NRM_CORNER) ? s->bc_rho[n] : bc_density(l, f, nrm, ur);
  }
  if (collide_it) {
    if (s->fd_corners && code == NRM_CORNER) {
      int64_t px, py, pz;
      coords(&s->d, n, &px, &py, &pz);
      const int64_t ic[3] = {px - s->d.hx, py - s->d.hy, pz - s->d.hz};
      const int64_t nn[3] = {s->d.nx, s->d.ny, s->d.nz};
      REAL grad[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
      for (int a = 0; a < l->D; ++a) {
        if (nn[a] < 3) continue;
        int dir = 0;
        if (!s->d.per[a]) {
          if (ic[a] == 0) dir = 1;
          else if (ic[a] == nn[a] - 1) dir = -1;
        }
        int64_t c1[3] = {ic[0], ic[1], ic[2]}, c2[3] = {ic[0], ic[1], ic[2]};
        REAL u1[3], u2[3];
        if (dir != 0) {
          c1[a] += dir; c2[a] += 2 * dir;
          fd_vel(s, P, interior_id(&s->d, c1[0], c1[1], c1[2]), u1);
          fd_vel(s, P, interior_id(&s->d, c2[0], c2[1], c2[2]), u2);
          for (int b = 0; b < l->D; ++b)
            grad[a][b] = R(dir) * (R(-1.5) * uw[b] + R(2) * u1[b] - R(0.5) * u2[b]);
        } else {
          c1[a] = (ic[a] + 1) % nn[a];
          c2[a] = (ic[a] - 1 + nn[a]) % nn[a];
          fd_vel(s, P, interior_id(&s->d, c1[0], c1[1], c1[2]), u1);
          fd_vel(s, P, interior_id(&s->d, c2[0], c2[1], c2[2]), u2);
          for (int b = 0; b < l->D; ++b) grad[a][b] = R(0.5) * (u1[b] - u2[b]);
        }
      }
      REAL Pi[6];
      bc_stress_from_gradient(rw, s->c.omega, grad, Pi);
      bc_apply_with_stress(&s->c, f, rw, ur, Pi);
    } else {
      bc_apply(&s->c, f, rw, ur, s->bc_unk[n], Fv);
    }
  }
  if (s->c.shifted)
    for (int i = 0; i < Q; ++i) f[i] -= weight(l, i);
  if (collide_it) {
    const Macro m = macroscopic(cp, f);
    collide(cp, &s->pt, f, &m);
  }
  F(s, n, 0) = f[0];
  for (int i = 1; i < Q; i += 2) {
    if (P == 0) { F(s, j[i], i + 1) = f[i]; F(s, n, i) = f[i + 1]; }
    else        { F(s, j[i], i) = f[i]; F(s, n, i + 1) = f[i + 1]; }
  }
}

/* run_step<P>, FluidSolver.hpp:712-1038, Fluid and Solid paths only; one
 * contiguous range of the active list. */
typedef struct {
  S* s;
  int P;
  int64_t lo, hi;
} Range;

static void* step_range(void* arg) {
  Range* r = (Range*)arg;
  S* s = r->s;
  const int P = r->P;
  const int Q = s->Q;
  int64_t j[27];
  REAL f[27];
  for (int64_t ai = r->lo; ai < r->hi; ++ai) {
    const int64_t n = s->active[ai];
    const uint8_t flag = s->flags[n];
    if (flag == M3LB_EXCLUDED) continue;
    if (s->streaming == 0) {
      if (flag == M3LB_SPECWALL) {
        /* specular ghost (:745-756): the incoming set permuted about the
         * wall normal's axis and stored back, no collision */
        int nv[3];
        normal_of(s->spec_nrm[n], nv);
        const int ax = nv[0] ? 0 : (nv[1] ? 1 : 2);
        fill_neighbours(&s->d, &s->c.l, 1, 2, n, j);
        REAL fin[27];
        fin[0] = F(s, n, 0);
        for (int i = 1; i < Q; i += 2) {
          if (P == 0) { fin[i] = F(s, n, i); fin[i + 1] = F(s, j[i], i + 1); }
          else        { fin[i] = F(s, n, i + 1); fin[i + 1] = F(s, j[i], i); }
        }
        for (int i = 0; i < Q; ++i) f[i] = fin[s->mirror[ax][i]];
        F(s, n, 0) = f[0];
        for (int i = 1; i < Q; i += 2) {
          if (P == 0) { F(s, j[i], i + 1) = f[i]; F(s, n, i) = f[i + 1]; }
          else        { F(s, j[i], i) = f[i]; F(s, n, i + 1) = f[i + 1]; }
        }
        continue;
      }
      if (flag == M3LB_SOLID) continue; /* implicit bounce-back, :759-761 */
      fill_neighbours(&s->d, &s->c.l, 1, 2, n, j);
      /* load_rest / load_pair, EsotericPull.hpp:87-97 */
      f[0] = F(s, n, 0);
      for (int i = 1; i < Q; i += 2) {
        if (P == 0) { f[i] = F(s, n, i); f[i + 1] = F(s, j[i], i + 1); }
        else        { f[i] = F(s, n, i + 1); f[i + 1] = F(s, j[i], i); }
      }
      /* on-node specular (:782-793): mirror the unknowns, then collide */
      if (flag == M3LB_SPECNODE) mirror_unknowns_faces(s, f, s->spec_faces[n]);
      Coll cn;
      const Coll* cp = coll_at(s, n, &cn);
      const Macro m = macroscopic(cp, f);
      collide(cp, &s->pt, f, &m);
      F(s, n, 0) = f[0];
      for (int i = 1; i < Q; i += 2) {
        if (P == 0) { F(s, j[i], i + 1) = f[i]; F(s, n, i) = f[i + 1]; }
        else        { F(s, j[i], i) = f[i]; F(s, n, i + 1) = f[i + 1]; }
      }
    } else {
      fill_neighbours(&s->d, &s->c.l, 1, 1, n, j);
      /* TwoLattice load_rest / load_pair, TwoLattice.hpp:49-57 */
      f[0] = F(s, n, 0);
      for (int i = 1; i < Q; i += 2) {
        f[i] = F(s, j[i + 1], i);
        f[i + 1] = F(s, j[i], i + 1);
      }
      REAL* dst = s->f2;
      if (flag == M3LB_SOLID) { /* explicit halfway bounce-back, :771-778 */
        for (int i = 0; i < Q; ++i) dst[n * Q + i] = f[opp(i)];
        continue;
      }
      Coll cn;
      const Coll* cp = coll_at(s, n, &cn);
      const Macro m = macroscopic(cp, f);
      collide(cp, &s->pt, f, &m);
      for (int i = 0; i < Q; ++i) dst[n * Q + i] = f[i];
    }
  }
  return NULL;
}

static void run_step(S* s) {
  const int P = (int)(s->t % 2);
  const int nt = s->nthreads;
  if (s->n_walls && s->streaming == 0) {
    /* FluidSolver::step: the corner-density pass (:583), then one sweep of
     * the active list in ascending node order, the order of the reference's
     * serial execution space, in which its published results were made.
     * Every node reads and writes only its own slots, so the order matters
     * only where a wall node reads a NEIGHBOUR's slots (an outlet's upstream
     * nodes, an outflow donor, a corner's finite-difference stencil): a
     * neighbour earlier in the order has already been updated and its slots
     * hold its post-collision set (pair-swapped), a later one has not. */
    if (s->has_corners) corner_density(s, P);
    int64_t wi = 0;
    for (int64_t ai = 0; ai <= s->n_active; ++ai) {
      const int64_t nn = ai < s->n_active ? s->active[ai] : INT64_MAX;
      while (wi < s->n_walls && s->walls[wi] < nn) wall_update(s, P, s->walls[wi++]);
      if (ai < s->n_active) {
        Range r = {s, P, ai, ai + 1};
        step_range(&r);
      }
    }
  } else if (nt <= 1 || s->n_active < 4096) {
    Range r = {s, P, 0, s->n_active};
    step_range(&r);
  } else {
    pthread_t th[256];
    Range rs[256];
    const int n = nt > 256 ? 256 : nt;
    for (int k = 0; k < n; ++k) {
      rs[k].s = s; rs[k].P = P;
      rs[k].lo = s->n_active * k / n;
      rs[k].hi = s->n_active * (k + 1) / n;
      pthread_create(&th[k], NULL, step_range, &rs[k]);
    }
    for (int k = 0; k < n; ++k) pthread_join(th[k], NULL);
  }
  if (s->streaming) { /* TwoLattice::end_of_step swap, TwoLattice.hpp:78 */
    REAL* tmp = s->f; s->f = s->f2; s->f2 = tmp;
  }
  ++s->t;
}

void CORE(step)(void* p, int nsteps) {
  for (int k = 0; k < nsteps; ++k) run_step((S*)p);
}

double CORE(time_steps)(void* p, int nsteps) {
  S* s = (S*)p;
  run_step(s); /* warm-up, src/app/main.cpp:42-43 */
  struct timespec a, b;
  clock_gettime(CLOCK_MONOTONIC, &a);
  for (int k = 0; k < nsteps; ++k) run_step(s);
  clock_gettime(CLOCK_MONOTONIC, &b);
  return (double)(b.tv_sec - a.tv_sec) + 1e-9 * (double)(b.tv_nsec - a.tv_nsec);
}

uint64_t CORE(timestep)(const void* p) { return ((const S*)p)->t; }

/* set_regularized_walls (FluidSolver.hpp:372-530): normals (0..10) and
 * tags per INTERIOR node (x fastest), the deduplicated wall-state table
 * (ux, uy, uz, rho per entry).  Marks the nodes RegWall, picks the corner
 * extrapolation axis, the unknown-direction masks and the outflow donors
 * exactly as the reference does, from the flags as they are now.  Returns 0,
 * or -1 for a bad argument (then nothing changes). */
int CORE(set_regularized_walls)(void* p, const uint8_t* normals, const uint16_t* tags,
                                const double* table, int ntab) {
  S* s = (S*)p;
  const Dom* d = &s->d;
  const int64_t np = d->np;
  const int64_t ni = d->nx * d->ny * d->nz;
  if (s->streaming != 0 || ntab < 1) return -1;
  for (int64_t k = 0; k < ni; ++k)
    if (normals[k] > NRM_OUT_EQ || (normals[k] && tags[k] >= ntab)) return -1;
  if (!s->bc_nrm) {
    s->bc_nrm = (uint8_t*)calloc((size_t)np, 1);
    s->bc_tag = (uint16_t*)calloc((size_t)np, sizeof(uint16_t));
    s->bc_unk = (uint32_t*)calloc((size_t)np, sizeof(uint32_t));
    s->bc_ext = (uint8_t*)calloc((size_t)np, 1);
    s->bc_rho = (REAL*)calloc((size_t)np, sizeof(REAL));
    s->bc_don = (int64_t*)calloc((size_t)np, sizeof(int64_t));
    s->bc_onrm = (uint8_t*)calloc((size_t)np, 1);
  }
  free(s->wtab);
  free(s->wtab0);
  s->wtab = (double*)malloc(sizeof(double) * 4 * (size_t)ntab);
  s->wtab0 = (double*)malloc(sizeof(double) * 4 * (size_t)ntab);
  /* the table holds Reals (View2D<Real> wall_u_) */
  for (int64_t k = 0; k < 4 * (int64_t)ntab; ++k) s->wtab[k] = s->wtab0[k] = (double)R(table[k]);
  s->ntab = ntab;
  for (int64_t n = 0; n < np; ++n) { s->bc_nrm[n] = 0; s->bc_tag[n] = 0; s->bc_ext[n] = 0; s->bc_unk[n] = 0; }
  int64_t k = 0;
  for (int64_t z = 0; z < d->nz; ++z)
    for (int64_t y = 0; y < d->ny; ++y)
      for (int64_t x = 0; x < d->nx; ++x, ++k) {
        if (normals[k] == 0) continue;
        const int64_t n = interior_id(d, x, y, z);
        s->bc_nrm[n] = normals[k];
        s->bc_tag[n] = tags[k];
        s->flags[n] = M3LB_REGWALL;
      }
  static const int dirs[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
  s->has_corners = 0;
  for (int64_t z = 0; z < d->nz; ++z)
    for (int64_t y = 0; y < d->ny; ++y)
      for (int64_t x = 0; x < d->nx; ++x) {
        const int64_t n = interior_id(d, x, y, z);
        if (s->bc_nrm[n] != NRM_CORNER) continue;
        s->has_corners = 1;
        for (int q = 0; q < (d->nz > 1 ? 6 : 4); ++q) {
          const int64_t x1 = x + dirs[q][0], y1 = y + dirs[q][1], z1 = z + dirs[q][2];
          const int64_t x2 = x + 2 * dirs[q][0], y2 = y + 2 * dirs[q][1], z2 = z + 2 * dirs[q][2];
          if (x1 < 0 || y1 < 0 || z1 < 0 || x2 < 0 || y2 < 0 || z2 < 0) continue;
          if (x1 >= d->nx || y1 >= d->ny || z1 >= d->nz) continue;
          if (x2 >= d->nx || y2 >= d->ny || z2 >= d->nz) continue;
          const int a = s->bc_nrm[interior_id(d, x1, y1, z1)];
          const int b = s->bc_nrm[interior_id(d, x2, y2, z2)];
          if (a != 0 && a != NRM_CORNER && b != 0 && b != NRM_CORNER) {
            s->bc_ext[n] = (uint8_t)(q + 1);
            break;
          }
        }
      }
  /* unknown directions: the source p - c_i outside a non-periodic extent, or
   * an Excluded or Solid node */
  const int64_t next[3] = {d->nx, d->ny, d->nz};
  for (int64_t z = 0; z < d->nz; ++z)
    for (int64_t y = 0; y < d->ny; ++y)
      for (int64_t x = 0; x < d->nx; ++x) {
        const int64_t n = interior_id(d, x, y, z);
        if (s->bc_nrm[n] == 0) continue;
        uint32_t m = 0;
        for (int i = 0; i < s->Q; ++i) {
          int64_t src[3] = {x - cvel(&s->c.l, i, 0), y - cvel(&s->c.l, i, 1), z - cvel(&s->c.l, i, 2)};
          int outside = 0;
          for (int a = 0; a < 3; ++a) {
            if (d->per[a]) src[a] = (src[a] + next[a]) % next[a];
            else if (src[a] < 0 || src[a] >= next[a]) outside = 1;
          }
          if (!outside) {
            const uint8_t sf = s->flags[interior_id(d, src[0], src[1], src[2])];
            if (sf == M3LB_EXCLUDED || sf == M3LB_SOLID) outside = 1;
          }
          if (outside) m |= (1u << i);
        }
        s->bc_unk[n] = m;
      }
  /* NrmOutFree donors (:475-520): the first axis step that leaves the fluid
   * (outside or Solid) whose opposite neighbour is Fluid; a node without one
   * points at itself */
  {
    static const int dd[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
    static const uint8_t codes[6] = {2, 1, 4, 3, 6, 5};
    for (int64_t n = 0; n < np; ++n) { s->bc_don[n] = n; s->bc_onrm[n] = 0; }
    for (int64_t z = 0; z < d->nz; ++z)
      for (int64_t y = 0; y < d->ny; ++y)
        for (int64_t x = 0; x < d->nx; ++x) {
          const int64_t n = interior_id(d, x, y, z);
          if (s->bc_nrm[n] != NRM_OUT_FREE) continue;
          for (int q = 0; q < 6; ++q) {
            const int64_t qx = x + dd[q][0], qy = y + dd[q][1], qz = z + dd[q][2];
            const int leaves = (qx < 0 || qx >= d->nx || qy < 0 || qy >= d->ny || qz < 0 ||
                                qz >= d->nz) ||
                               s->flags[interior_id(d, qx, qy, qz)] == M3LB_SOLID;
            if (!leaves) continue;
            const int o = q ^ 1;
            const int64_t rx = x + dd[o][0], ry = y + dd[o][1], rz = z + dd[o][2];
            if (rx < 0 || rx >= d->nx || ry < 0 || ry >= d->ny || rz < 0 || rz >= d->nz) continue;
            const int64_t m = interior_id(d, rx, ry, rz);
            if (s->flags[m] != M3LB_FLUID) continue;
            s->bc_don[n] = m;
            s->bc_onrm[n] = codes[q];
            break;
          }
        }
  }
  rebuild_lists(s);
  return 0;
}

/* set_fd_corners (FluidSolver.hpp:272) */
void CORE(set_fd_corners)(void* p, int on) { ((S*)p)->fd_corners = on ? 1 : 0; }
/* set_force_correction (:1311), set_outflow_order (:1324), set_outflow_density (:1330) */
void CORE(set_force_correction)(void* p, int on) { ((S*)p)->force_bc = on ? 1 : 0; }
void CORE(set_outflow_order)(void* p, int k) { ((S*)p)->out_order = k; }
void CORE(set_outflow_density)(void* p, double r) { ((S*)p)->outflow_rho = R(r); }

/* set_wall_velocity_scale (:303-312): the velocities of the table as given,
 * times sc (in Real) */
void CORE(set_wall_velocity_scale)(void* p, double sc) {
  S* s = (S*)p;
  for (int k = 0; k < s->ntab; ++k)
    for (int a = 0; a < 3; ++a) s->wtab[4 * k + a] = (double)(R(s->wtab0[4 * k + a]) * R(sc));
}
/* set_wall_density (:330-335): slot 3 of every state, and of the table as given */
void CORE(set_wall_density)(void* p, double r) {
  S* s = (S*)p;
  for (int k = 0; k < s->ntab; ++k) s->wtab[4 * k + 3] = s->wtab0[4 * k + 3] = (double)R(r);
}
/* set_wall_velocity_field (:353-367): one velocity per interior node (x
 * fastest), read instead of the table from now on; its scale restarts at 1 */
void CORE(set_wall_velocity_field)(void* p, const double* ux, const double* uy, const double* uz) {
  S* s = (S*)p;
  const Dom* d = &s->d;
  if (!s->wall_un) s->wall_un = (REAL*)malloc(sizeof(REAL) * 3 * (size_t)d->np);
  for (int64_t n = 0; n < 3 * d->np; ++n) s->wall_un[n] = R(0);
  int64_t k = 0;
  for (int64_t z = 0; z < d->nz; ++z)
    for (int64_t y = 0; y < d->ny; ++y)
      for (int64_t x = 0; x < d->nx; ++x, ++k) {
        const int64_t n = interior_id(d, x, y, z);
        s->wall_un[n] = R(ux[k]);
        s->wall_un[d->np + n] = R(uy[k]);
        s->wall_un[2 * d->np + n] = R(uz[k]);
      }
  s->un_scale = R(1);
}
void CORE(set_wall_velocity_field_scale)(void* p, double sc) { ((S*)p)->un_scale = R(sc); }

/* set_specular_walls (:193-216): axis normals 1..6 per interior node mark
 * ghost specular cells; every other cell's normal is cleared.  Returns -1
 * (and changes nothing) for a code above 6. */
int CORE(set_specular_walls)(void* p, const uint8_t* normals) {
  S* s = (S*)p;
  const Dom* d = &s->d;
  const int64_t ni = d->nx * d->ny * d->nz;
  if (s->streaming != 0) return -1;
  for (int64_t k = 0; k < ni; ++k)
    if (normals[k] > 6) return -1;
  if (!s->spec_nrm) s->spec_nrm = (uint8_t*)calloc((size_t)d->np, 1);
  for (int64_t n = 0; n < d->np; ++n) s->spec_nrm[n] = 0;
  int64_t k = 0;
  for (int64_t z = 0; z < d->nz; ++z)
    for (int64_t y = 0; y < d->ny; ++y)
      for (int64_t x = 0; x < d->nx; ++x, ++k) {
        if (!normals[k]) continue;
        const int64_t n = interior_id(d, x, y, z);
        s->spec_nrm[n] = normals[k];
        s->flags[n] = M3LB_SPECWALL;
      }
  rebuild_lists(s);
  return 0;
}

/* set_specular_nodes (:236-270): a face mask per interior node marks on-node
 * specular nodes.  Returns -1 (and changes nothing) for bits above 0x3F or
 * both faces of one axis. */
int CORE(set_specular_nodes)(void* p, const uint8_t* faces) {
  S* s = (S*)p;
  const Dom* d = &s->d;
  const int64_t ni = d->nx * d->ny * d->nz;
  if (s->streaming != 0) return -1;
  for (int64_t k = 0; k < ni; ++k) {
    if (faces[k] > 0x3F) return -1;
    for (int a = 0; a < 3; ++a)
      if ((faces[k] & (1u << (2 * a))) && (faces[k] & (1u << (2 * a + 1)))) return -1;
  }
  if (!s->spec_faces) s->spec_faces = (uint8_t*)calloc((size_t)d->np, 1);
  for (int64_t n = 0; n < d->np; ++n) s->spec_faces[n] = 0;
  int64_t k = 0;
  for (int64_t z = 0; z < d->nz; ++z)
    for (int64_t y = 0; y < d->ny; ++y)
      for (int64_t x = 0; x < d->nx; ++x, ++k) {
        if (!faces[k]) continue;
        const int64_t n = interior_id(d, x, y, z);
        s->spec_faces[n] = faces[k];
        s->flags[n] = M3LB_SPECNODE;
      }
  rebuild_lists(s);
  return 0;
}

void CORE(set_force_field)(void* p, const double* fx, const double* fy, const double* fz) {
  S* s = (S*)p;
  const int64_t np = s->d.np;
  if (!s->ff) s->ff = (REAL*)calloc((size_t)(3 * np), sizeof(REAL));
  for (int64_t n = 0; n < np; ++n) {
    s->ff[n] = R(fx[n]);
    s->ff[np + n] = R(fy[n]);
    s->ff[2 * np + n] = R(fz[n]);
  }
}

void CORE(set_force)(void* p, double fx, double fy, double fz) {
  S* s = (S*)p;
  s->c.F[0] = R(fx); s->c.F[1] = R(fy); s->c.F[2] = R(fz);
  s->cfg.force[0] = fx; s->cfg.force[1] = fy; s->cfg.force[2] = fz;
}

/* Gather a node's incoming set at the current parity (load_rest + load_pair) */
static void load_node(const S* s, int64_t n, REAL* f) {
  const int P = (int)(s->t % 2);
  int64_t j[27];
  if (s->streaming == 0) {
    fill_neighbours(&s->d, &s->c.l, 1, 2, n, j);
    for (int i = 0; i < s->Q; ++i) f[i] = ep_load(s, P, j, n, i);
  } else {
    fill_neighbours(&s->d, &s->c.l, 1, 1, n, j);
    for (int i = 0; i < s->Q; ++i) f[i] = tl_load(s, j, n, i);
  }
}

/* macro_kernel<P> fluid path, FluidSolver.hpp:1042-1056, 1192-1199 */
void CORE(compute_macroscopic)(void* p, double* rho, double* ux, double* uy, double* uz) {
  S* s = (S*)p;
  REAL f[27];
  for (int64_t n = 0; n < s->d.np; ++n) {
    REAL r = R(0), a = R(0), b = R(0), c = R(0);
    if (s->flags[n] == M3LB_FLUID) {
      load_node(s, n, f);
      Coll cn;
      const Macro m = macroscopic(coll_at(s, n, &cn), f);
      r = density(&s->c, &m); a = m.ux; b = m.uy; c = m.uz;
    } else if (s->flags[n] == M3LB_SPECNODE) {
      /* macro_kernel's SpecNode path (:1057-1077): raw moments of the
       * mirrored incoming set */
      load_node(s, n, f);
      mirror_unknowns_faces(s, f, s->spec_faces[n]);
      if (s->c.shifted)
        for (int i = 0; i < s->Q; ++i) f[i] += weight(&s->c.l, i);
      REAL sm = R(0), mx = R(0), my = R(0), mz = R(0);
      for (int i = 0; i < s->Q; ++i) {
        sm += f[i];
        mx += f[i] * R(cvel(&s->c.l, i, 0));
        my += f[i] * R(cvel(&s->c.l, i, 1));
        mz += f[i] * R(cvel(&s->c.l, i, 2));
      }
      const REAL ir = R(1) / sm;
      r = sm; a = mx * ir; b = my * ir; c = mz * ir;
    } else if (s->flags[n] == M3LB_REGWALL) {
      /* macro_kernel's RegWall path (:1078-1185): the imposed (or outflow)
       * velocity and the closure (or imposed) density; a corner reports the
       * last pass's rho */
      const int P = (int)(s->t % 2);
      const int code = s->bc_nrm[n];
      const int tg = s->bc_tag[n];
      REAL uw[3];
      wall_vel(s, n, uw);
      load_raw(s, P, n, f);
      int nrm[3];
      normal_of(code, nrm);
      if (code == NRM_OUT_FREE) {
        r = s->outflow_rho;
        int onr[3];
        normal_of(s->bc_onrm[n], onr);
        REAL sm, m4[3];
        mean_velocity(s, P, s->bc_don[n], &sm, m4, 1);
        const REAL un = bc_normal_velocity(&s->c.l, f, onr, r);
        REAL utn = R(0);
        for (int k = 0; k < 3; ++k) utn += m4[k] * R(onr[k]);
        for (int k = 0; k < 3; ++k) uw[k] = m4[k] + (un - utn) * R(onr[k]);
      } else if (code == NRM_OUT_EQ) {
        r = R(s->wtab[4 * (int64_t)tg + 3]);
        int64_t px, py, pz;
        coords(&s->d, n, &px, &py, &pz);
        mean_velocity(s, P, interior_id(&s->d, px - s->d.hx - 1, py - s->d.hy, pz - s->d.hz),
                      NULL, uw, 1);
      } else if (code == NRM_OUT_XP) {
        r = R(s->wtab[4 * (int64_t)tg + 3]);
        REAL ur[3];
        outflow_xp_velocity(s, P, n, f, r, ur);
        uw[0] = ur[0]; uw[1] = ur[1]; uw[2] = ur[2];
      } else {
        r = (code == NRM_CORNER) ? s->bc_rho[n] : bc_density(&s->c.l, f, nrm, uw);
      }
      a = uw[0]; b = uw[1]; c = uw[2];
    }
    if (rho) rho[n] = (double)r;
    if (ux) ux[n] = (double)a;
    if (uy) uy[n] = (double)b;
    if (uz) uz[n] = (double)c;
  }
}

/* reduce<P>(axis), FluidSolver.hpp:1205-1229 (accumulated in double) */
static double reduce(S* s, int axis) {
  REAL f[27];
  double sum = 0;
  for (int64_t n = 0; n < s->d.np; ++n) {
    if (!fluid_like(s->flags[n])) continue;
    load_node(s, n, f);
    /* the streamed unknowns of a SpecNode hold nothing until mirrored */
    if (s->flags[n] == M3LB_SPECNODE) mirror_unknowns_faces(s, f, s->spec_faces[n]);
    Coll cn;
    const Macro m = macroscopic(coll_at(s, n, &cn), f);
    const REAL r = density(&s->c, &m);
    sum += (double)((axis < 0) ? r : r * (axis == 0 ? m.ux : (axis == 1 ? m.uy : m.uz)));
  }
  return sum;
}
double CORE(total_mass)(void* p) { return reduce((S*)p, -1); }
double CORE(total_momentum)(void* p, int axis) { return reduce((S*)p, axis); }

void CORE(storage_sum_exact)(void* p, int64_t out[2]) {
  S* s = (S*)p;
  unsigned __int128 acc = 0;
  for (int64_t k = 0; k < s->d.np * s->Q; ++k) {
    int64_t v[2];
    orc_fixed100((double)s->f[k], v);
    const unsigned __int128 x =
        ((unsigned __int128)(uint64_t)v[0] << 64) | (unsigned __int128)(uint64_t)v[1];
    acc += x;
  }
  out[0] = (int64_t)(uint64_t)(acc >> 64);
  out[1] = (int64_t)(uint64_t)acc;
}

/* campaign::diagnostics, validation/Campaign.hpp:99-130 */
void CORE(energy_enstrophy)(void* p, double* e_out, double* w_out, int* finite) {
  S* s = (S*)p;
  const Dom* d = &s->d;
  const int64_t np = d->np;
  double* hx = (double*)malloc(sizeof(double) * (size_t)np);
  double* hy = (double*)malloc(sizeof(double) * (size_t)np);
  double* hz = (double*)malloc(sizeof(double) * (size_t)np);
  CORE(compute_macroscopic)(p, NULL, hx, hy, hz);
  const int64_t nx = d->nx, ny = d->ny, nz = d->nz;
#define U(x, y, z, c)                                                                  \
  ((c) == 0 ? hx[idp(d, ((x) + nx) % nx + d->hx, ((y) + ny) % ny + d->hy,              \
                     ((z) + nz) % nz + d->hz)]                                         \
            : ((c) == 1 ? hy[idp(d, ((x) + nx) % nx + d->hx, ((y) + ny) % ny + d->hy,  \
                                 ((z) + nz) % nz + d->hz)]                             \
                        : hz[idp(d, ((x) + nx) % nx + d->hx, ((y) + ny) % ny + d->hy,  \
                                 ((z) + nz) % nz + d->hz)]))
  double e = 0, w = 0;
  int ok = 1;
  for (int64_t z = 0; z < nz; ++z)
    for (int64_t y = 0; y < ny; ++y)
      for (int64_t x = 0; x < nx; ++x) {
        const int64_t n = idp(d, x + d->hx, y + d->hy, z + d->hz);
        const double a = hx[n], b = hy[n], c = hz[n];
        if (!isfinite(a) || !isfinite(b) || !isfinite(c)) { ok = 0; continue; }
        e += 0.5 * (a * a + b * b + c * c);
        const double wz = 0.5 * (U(x + 1, y, z, 1) - U(x - 1, y, z, 1)) -
                          0.5 * (U(x, y + 1, z, 0) - U(x, y - 1, z, 0));
        double wx = 0, wy = 0;
        if (nz > 1) {
          wx = 0.5 * (U(x, y + 1, z, 2) - U(x, y - 1, z, 2)) -
               0.5 * (U(x, y, z + 1, 1) - U(x, y, z - 1, 1));
          wy = 0.5 * (U(x, y, z + 1, 0) - U(x, y, z - 1, 0)) -
               0.5 * (U(x + 1, y, z, 2) - U(x - 1, y, z, 2));
        }
        w += 0.5 * (wx * wx + wy * wy + wz * wz);
      }
#undef U
  free(hx); free(hy); free(hz);
  *e_out = e; *w_out = w; *finite = ok;
}

/* ---- single-node entry points -------------------------------------------- */
void CORE(collide_node)(const m3lb_config* cfg, const double* fin, double* fout,
                        double mo[4]) {
  const Coll c = coll_of(cfg);
  const PiTab t = pitab_of(&c.l);
  REAL f[27];
  for (int i = 0; i < c.l.Q; ++i) f[i] = R(fin[i]);
  const Macro m = macroscopic(&c, f);
  collide(&c, &t, f, &m);
  for (int i = 0; i < c.l.Q; ++i) fout[i] = (double)f[i];
  if (mo) { mo[0] = (double)m.dens; mo[1] = (double)m.ux; mo[2] = (double)m.uy; mo[3] = (double)m.uz; }
}

void CORE(macroscopic_node)(const m3lb_config* cfg, const double* fin, double mo[4]) {
  const Coll c = coll_of(cfg);
  REAL f[27];
  for (int i = 0; i < c.l.Q; ++i) f[i] = R(fin[i]);
  const Macro m = macroscopic(&c, f);
  mo[0] = (double)m.dens; mo[1] = (double)m.ux; mo[2] = (double)m.uy; mo[3] = (double)m.uz;
}

void CORE(seed_node)(const m3lb_config* cfg, double rho, double ux, double uy, double uz,
                     double* out) {
  const Coll c = coll_of(cfg);
  const PiTab t = pitab_of(&c.l);
  REAL f[27];
  seed_all(&c, &t, R(rho), R(ux), R(uy), R(uz), f);
  for (int i = 0; i < c.l.Q; ++i) out[i] = (double)f[i];
}

void CORE(equilibrium)(const m3lb_config* cfg, int product, int shifted, double rho,
                       double ux, double uy, double uz, double* out) {
  const Lat l = lat_of(cfg->lattice);
  for (int i = 0; i < l.Q; ++i)
    out[i] = shifted ? (double)eq_shifted(&l, product, i, R(rho) - R(1), R(ux), R(uy), R(uz))
                     : (double)eq(&l, product, i, R(rho), R(ux), R(uy), R(uz));
}

void CORE(to_moments)(const m3lb_config* cfg, const double* fin, const double ub[3],
                      double* kout) {
  const Lat l = lat_of(cfg->lattice);
  const PiTab t = pitab_of(&l);
  REAL f[27], k[27];
  const REAL u[3] = {R(ub[0]), R(ub[1]), R(ub[2])};
  for (int i = 0; i < l.Q; ++i) f[i] = R(fin[i]);
  to_moments(&l, &t, f, u, k);
  for (int i = 0; i < l.Q; ++i) kout[i] = (double)k[i];
}

void CORE(to_populations)(const m3lb_config* cfg, const double* kin, const double ub[3],
                          double* fout) {
  const Lat l = lat_of(cfg->lattice);
  const PiTab t = pitab_of(&l);
  REAL f[27], k[27];
  const REAL u[3] = {R(ub[0]), R(ub[1]), R(ub[2])};
  for (int i = 0; i < l.Q; ++i) k[i] = R(kin[i]);
  to_populations(&l, &t, k, u, f);
  for (int i = 0; i < l.Q; ++i) fout[i] = (double)f[i];
}
