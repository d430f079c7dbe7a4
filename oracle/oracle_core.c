/*
 * oracle_core.c -- precision-generic body of the CPU restatement.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Compiled twice by oracle/Makefile:
 * -DORC_PREC=64 (REAL = double) and -DORC_PREC=32 (REAL = float), mirroring the
 * reference's configure-time typedef (src/core/Types.hpp:16-20).  Every routine
 * restates the reference routine cited beside it, in the reference's operation
 * order, so an FP64 run reproduces the reference's arithmetic term for term.
 * Node order is ascending, which is Kokkos Serial's order; the esoteric-pull
 * step is order-independent (src/memory/EsotericPull.hpp:19-21), the seeding
 * scatter is not at halo layers and ascending order is what it leaves.
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
 * Lattices, src/lattice/Lattices.hpp:130-214 (D2Q9, D3Q27).
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

/* cvel<L>(i, a), Lattices.hpp:234-240 */
static inline int cvel(const Lat* l, int i, int a) {
  return a == 0 ? l->cx[i] : (a == 1 ? l->cy[i] : l->cz[i]);
}
/* weight<L,T>(i), Lattices.hpp:221-224 */
static inline REAL weight(const Lat* l, int i) { return R(l->wn[i]) / R(l->wden); }
/* cs2 = 1/3 and inv_cs2 = 3 on both lattices, Lattices.hpp:225-232 */
#define CS2 (R(1) / R(3))
#define ICS2 (R(3) / R(1))

/* ---------------------------------------------------------------------------
 * Collision policy state (BGK.hpp:181-182, TRT.hpp:46-48,
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
} Macro; /* BGK.hpp:164-167 */

/* Collision::density, BGK.hpp:191-194 (NullCollision: m.dens,
 * tests/test_streaming.cpp:36) */
static inline REAL density(const Coll* c, const Macro* m) {
  if (c->op == M3LB_OP_NULL) return m->dens;
  return c->shifted ? R(1) + m->dens : m->dens;
}

/* BGK/TRT/MomentCollision::macroscopic (BGK.hpp:201-216, TRT.hpp:69-84,
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

/* BGK::collide, BGK.hpp:219-230 with Guo::source (Forcing.hpp:76-80) */
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

/* TRT::collide, TRT.hpp:92-127 (SecondOrderEquilibrium) */
static void collide_trt(const Coll* c, REAL* f, const Macro* m) {
  const Lat* l = &c->l;
  REAL e[27], sr[27];
  for (int i = 0; i < l->Q; ++i)
    e[i] = c->shifted ? eq_shifted(l, 0, i, m->dens, m->ux, m->uy, m->uz)
                      : eq(l, 0, i, m->dens, m->ux, m->uy, m->uz);
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
 * ProductBasis<L>, src/collision/ProductBasis.hpp:191-342.
 * ------------------------------------------------------------------------- */
/* dir_index<L>(a,b,c), ProductBasis.hpp:171-176 */
static int dir_index(const Lat* l, int a, int b, int c) {
  for (int i = 0; i < l->Q; ++i)
    if (cvel(l, i, 0) == a && cvel(l, i, 1) == b && cvel(l, i, 2) == c) return i;
  return -1;
}
/* pi(a,b,c) table, ProductBasis.hpp:254-267 */
static int pi_of(const Lat* l, int a, int b, int c) {
  return dir_index(l, a - 1, b - 1, l->D == 3 ? c - 1 : 0);
}
/* mi(p,q,r), ProductBasis.hpp:198-200 */
static inline int mi(int D, int p, int q, int r) {
  return D == 2 ? (p * 3 + q) : ((p * 3 + q) * 3 + r);
}

/* fwd1d / inv1d, ProductBasis.hpp:273-292 */
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

/* to_moments, ProductBasis.hpp:295-317 */
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

/* to_populations, ProductBasis.hpp:319-341 */
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
 * ProductBasis.hpp:243-247 */
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
    case M3LB_OP_TRT: collide_trt(c, f, m); return;
    case M3LB_OP_MRT: collide_moment(c, t, 0, f, m); return;
    default: collide_moment(c, t, 1, f, m); return;
  }
}

/* seed_value for every operator: BGK.hpp:233-237, TRT.hpp:129-133,
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
  const int product = c->op == M3LB_OP_BGK_PRODUCT;
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

/* EsotericPull::Access<P>::load, EsotericPull.hpp:191-198 */
static inline REAL ep_load(const S* s, int P, const int64_t* j, int64_t n, int i) {
  if (i == 0) return F(s, n, 0);
  if (P == 0) return (i & 1) ? F(s, n, i) : F(s, j[i - 1], i);
  return (i & 1) ? F(s, n, i + 1) : F(s, j[i - 1], i - 1);
}
/* EsotericPull::Access<P>::scatter, EsotericPull.hpp:230-237 */
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

/* rebuild_lists, FluidSolver.hpp:157-177 */
static void rebuild_lists(S* s) {
  int64_t k = 0;
  for (int64_t n = 0; n < s->d.np; ++n) {
    const uint8_t fl = s->flags[n];
    if (fl == M3LB_EXCLUDED) continue;
    if (fl == M3LB_SOLID && s->streaming == 0) continue;
    s->active[k++] = n;
  }
  s->n_active = k;
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
    if (s->flags[n] == M3LB_FLUID) {
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
      if (flag == M3LB_SOLID) continue; /* implicit bounce-back, :759-761 */
      fill_neighbours(&s->d, &s->c.l, 1, 2, n, j);
      /* load_rest / load_pair, EsotericPull.hpp:213-223 */
      f[0] = F(s, n, 0);
      for (int i = 1; i < Q; i += 2) {
        if (P == 0) { f[i] = F(s, n, i); f[i + 1] = F(s, j[i], i + 1); }
        else        { f[i] = F(s, n, i + 1); f[i + 1] = F(s, j[i], i); }
      }
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
  if (nt <= 1 || s->n_active < 4096) {
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
    if (s->flags[n] != M3LB_FLUID) continue;
    load_node(s, n, f);
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
