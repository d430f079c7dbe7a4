/*
 * oracle.c -- precision dispatch for the CPU restatement.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h): the checker the HIP path is
 * compared against, never part of the product.
 */
#include <math.h>
#include <stdlib.h>

#include "oracle.h"
#include "oracle_core.h"

struct orc_solver {
  int p64;
  void* impl;
};

#define DISPATCH(s, name, ...) \
  ((s)->p64 ? core64_##name((s)->impl, __VA_ARGS__) : core32_##name((s)->impl, __VA_ARGS__))

orc_solver* orc_create(const m3lb_config* cfg, int streaming) {
  orc_solver* s = (orc_solver*)calloc(1, sizeof(orc_solver));
  s->p64 = cfg->dtype == M3LB_F64;
  s->impl = s->p64 ? core64_create(cfg, streaming) : core32_create(cfg, streaming);
  return s;
}
void orc_destroy(orc_solver* s) {
  if (!s) return;
  if (s->p64) core64_destroy(s->impl); else core32_destroy(s->impl);
  free(s);
}
void orc_set_threads(orc_solver* s, int n) { DISPATCH(s, set_threads, n); }
void orc_domain(const orc_solver* s, int64_t out[10]) { DISPATCH(s, domain, out); }
void orc_set_geometry(orc_solver* s, const uint8_t* f) { DISPATCH(s, set_geometry, f); }
void orc_initialize(orc_solver* s, double rho0, double ux, double uy, double uz) {
  DISPATCH(s, initialize, rho0, ux, uy, uz);
}
void orc_initialize_field(orc_solver* s, const double* rho, const double* ux,
                          const double* uy, const double* uz) {
  DISPATCH(s, initialize_field, rho, ux, uy, uz);
}
void orc_step(orc_solver* s, int nsteps) { DISPATCH(s, step, nsteps); }
void orc_compute_macroscopic(orc_solver* s, double* rho, double* ux, double* uy,
                             double* uz) {
  DISPATCH(s, compute_macroscopic, rho, ux, uy, uz);
}
double orc_total_mass(orc_solver* s) {
  return s->p64 ? core64_total_mass(s->impl) : core32_total_mass(s->impl);
}
double orc_total_momentum(orc_solver* s, int axis) { return DISPATCH(s, total_momentum, axis); }
void orc_storage_sum_exact(orc_solver* s, int64_t out[2]) { DISPATCH(s, storage_sum_exact, out); }
void orc_energy_enstrophy(orc_solver* s, double* e, double* w, int* finite) {
  DISPATCH(s, energy_enstrophy, e, w, finite);
}
void orc_gather_populations(orc_solver* s, double* out) { DISPATCH(s, gather_populations, out); }
void orc_seed_populations(orc_solver* s, const double* in) { DISPATCH(s, seed_populations, in); }
void orc_storage(orc_solver* s, double* out) { DISPATCH(s, storage, out); }
uint64_t orc_timestep(const orc_solver* s) {
  return s->p64 ? core64_timestep(s->impl) : core32_timestep(s->impl);
}
void orc_set_force(orc_solver* s, double fx, double fy, double fz) {
  DISPATCH(s, set_force, fx, fy, fz);
}
void orc_set_force_field(orc_solver* s, const double* fx, const double* fy, const double* fz) {
  DISPATCH(s, set_force_field, fx, fy, fz);
}
int orc_set_regularized_walls(orc_solver* s, const uint8_t* normals, const uint16_t* tags,
                              const double* table, int ntab) {
  return DISPATCH(s, set_regularized_walls, normals, tags, table, ntab);
}
void orc_set_fd_corners(orc_solver* s, int on) { DISPATCH(s, set_fd_corners, on); }
void orc_set_force_correction(orc_solver* s, int on) { DISPATCH(s, set_force_correction, on); }
void orc_set_outflow_order(orc_solver* s, int k) { DISPATCH(s, set_outflow_order, k); }
void orc_set_outflow_density(orc_solver* s, double r) { DISPATCH(s, set_outflow_density, r); }
void orc_set_wall_velocity_scale(orc_solver* s, double sc) { DISPATCH(s, set_wall_velocity_scale, sc); }
void orc_set_wall_density(orc_solver* s, double r) { DISPATCH(s, set_wall_density, r); }
void orc_set_wall_velocity_field(orc_solver* s, const double* ux, const double* uy,
                                 const double* uz) {
  DISPATCH(s, set_wall_velocity_field, ux, uy, uz);
}
void orc_set_wall_velocity_field_scale(orc_solver* s, double sc) {
  DISPATCH(s, set_wall_velocity_field_scale, sc);
}
int orc_set_specular_walls(orc_solver* s, const uint8_t* normals) {
  return DISPATCH(s, set_specular_walls, normals);
}
int orc_set_specular_nodes(orc_solver* s, const uint8_t* faces) {
  return DISPATCH(s, set_specular_nodes, faces);
}
double orc_time_steps(orc_solver* s, int nsteps) { return DISPATCH(s, time_steps, nsteps); }

#define NODE(name, cfg, ...) \
  ((cfg)->dtype == M3LB_F64 ? core64_##name(cfg, __VA_ARGS__) : core32_##name(cfg, __VA_ARGS__))

void orc_collide_node(const m3lb_config* cfg, const double* fin, double* fout, double mo[4]) {
  NODE(collide_node, cfg, fin, fout, mo);
}
void orc_macroscopic_node(const m3lb_config* cfg, const double* f, double mo[4]) {
  NODE(macroscopic_node, cfg, f, mo);
}
void orc_seed_node(const m3lb_config* cfg, double rho, double ux, double uy, double uz,
                   double* out) {
  NODE(seed_node, cfg, rho, ux, uy, uz, out);
}
void orc_equilibrium(const m3lb_config* cfg, int product, int shifted, double rho, double ux,
                     double uy, double uz, double* out) {
  NODE(equilibrium, cfg, product, shifted, rho, ux, uy, uz, out);
}
void orc_to_moments(const m3lb_config* cfg, const double* f, const double ub[3], double* k) {
  NODE(to_moments, cfg, f, ub, k);
}
void orc_to_populations(const m3lb_config* cfg, const double* k, const double ub[3],
                        double* f) {
  NODE(to_populations, cfg, k, ub, f);
}

/* value * 2^100, truncated toward zero, as a two's-complement 128-bit integer.
 * |value| < 2^26 is assumed (population values are O(1)); larger values
 * saturate.  Exact for every double: the 53-bit mantissa is shifted, never
 * rounded. */
void orc_fixed100(double v, int64_t out[2]) {
  unsigned __int128 x = 0;
  if (v != 0.0 && isfinite(v)) {
    int e;
    const double fr = frexp(fabs(v), &e);           /* |v| = fr * 2^e, fr in [0.5,1) */
    const uint64_t mant = (uint64_t)ldexp(fr, 53);  /* exact 53-bit integer */
    const int sh = e - 53 + 100;                    /* |v| * 2^100 = mant * 2^sh */
    if (sh >= 0) {
      x = sh >= 74 ? ((unsigned __int128)1 << 126) : ((unsigned __int128)mant << sh);
    } else if (sh > -64) {
      x = (unsigned __int128)(mant >> (-sh));
    }
    if (v < 0) x = ~x + 1;
  }
  out[0] = (int64_t)(uint64_t)(x >> 64);
  out[1] = (int64_t)(uint64_t)x;
}
