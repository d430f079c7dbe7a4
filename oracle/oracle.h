/*
 * oracle.h -- CPU restatement of the reference's fused esoteric-pull step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path is
 * compared against; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product (m3lb_amd/, libm3lb_hip.so) never
 * links, loads or calls it.
 *
 * Every function follows the reference file:line named at its definition in
 * oracle_core.c.  The configuration struct is the product's public m3lb_config
 * so that a test drives both sides with one value.  F64 runs in double, every
 * other dtype in float (the reference's two typedef choices,
 * src/core/Types.hpp:16-20); 16-bit storage is not restated (no counterpart in
 * the reference).
 */
#ifndef M3LB_ORACLE_H
#define M3LB_ORACLE_H

#include <stdint.h>
#include "../include/m3lb.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_solver orc_solver;

/* streaming: 0 = EsotericPull (src/memory/EsotericPull.hpp),
 *            1 = TwoLattice   (src/memory/TwoLattice.hpp), the reference's
 *                permanent bisection reference. */
orc_solver* orc_create(const m3lb_config* cfg, int streaming);
void orc_destroy(orc_solver* s);
void orc_set_threads(orc_solver* s, int nthreads);
void orc_domain(const orc_solver* s, int64_t out[10]);
void orc_set_geometry(orc_solver* s, const uint8_t* interior_flags);
void orc_initialize(orc_solver* s, double rho0, double ux, double uy, double uz);
void orc_initialize_field(orc_solver* s, const double* rho, const double* ux,
                          const double* uy, const double* uz);
void orc_step(orc_solver* s, int nsteps);
void orc_compute_macroscopic(orc_solver* s, double* rho, double* ux, double* uy,
                             double* uz);
double orc_total_mass(orc_solver* s);
double orc_total_momentum(orc_solver* s, int axis);
void orc_storage_sum_exact(orc_solver* s, int64_t out[2]);
void orc_energy_enstrophy(orc_solver* s, double* e, double* w, int* finite);
void orc_gather_populations(orc_solver* s, double* out);
void orc_seed_populations(orc_solver* s, const double* in);
void orc_storage(orc_solver* s, double* out);
uint64_t orc_timestep(const orc_solver* s);
void orc_set_force(orc_solver* s, double fx, double fy, double fz);
/* Per-node Guo force for M3LB_FORCE_FIELD configs: three padded planes. */
void orc_set_force_field(orc_solver* s, const double* fx, const double* fy, const double* fz);
/* Regularised walls: normal code per INTERIOR node (0 none, 1..6 =
 * +x,-x,+y,-y,+z,-z outward normal, 7 corner, 8..10 the outflow codes),
 * table index per node, and the table of (ux, uy, uz, rho).  Returns 0 or -1. */
int orc_set_regularized_walls(orc_solver* s, const uint8_t* normals, const uint16_t* tags,
                              const double* table, int ntab);
void orc_set_fd_corners(orc_solver* s, int on);
void orc_set_force_correction(orc_solver* s, int on);
void orc_set_outflow_order(orc_solver* s, int k);
void orc_set_outflow_density(orc_solver* s, double r);
void orc_set_wall_velocity_scale(orc_solver* s, double sc);
void orc_set_wall_density(orc_solver* s, double r);
/* one wall velocity per INTERIOR node (x fastest) */
void orc_set_wall_velocity_field(orc_solver* s, const double* ux, const double* uy,
                                 const double* uz);
void orc_set_wall_velocity_field_scale(orc_solver* s, double sc);
/* Specular ghost cells: axis normal 1..6 per interior node.  Returns 0 or -1. */
int orc_set_specular_walls(orc_solver* s, const uint8_t* normals);
/* On-node specular nodes: face mask per interior node.  Returns 0 or -1. */
int orc_set_specular_nodes(orc_solver* s, const uint8_t* faces);
/* Wall-clock seconds for nsteps steps after one untimed warm-up step,
 * timed as src/app/main.cpp:41-47 does. */
double orc_time_steps(orc_solver* s, int nsteps);

/* Single-node entry points for the operator-level tests
 * (tests/test_collision.cpp, tests/test_moments.cpp, tests/test_equilibrium.cpp
 * of the reference).  Arrays are double; arithmetic is in the config's
 * precision. */
void orc_collide_node(const m3lb_config* cfg, const double* fin, double* fout,
                      double macro_out[4]);
void orc_macroscopic_node(const m3lb_config* cfg, const double* f, double macro_out[4]);
void orc_seed_node(const m3lb_config* cfg, double rho, double ux, double uy, double uz,
                   double* out);
/* shifted = 0: Eq::eq(i, rho, u); 1: Eq::eq_shifted(i, rho - 1, u).
 * product = 0: SecondOrderPhi, 1: ProductFormPhi. */
void orc_equilibrium(const m3lb_config* cfg, int product, int shifted, double rho,
                     double ux, double uy, double uz, double* out);
void orc_to_moments(const m3lb_config* cfg, const double* f, const double ub[3],
                    double* k);
void orc_to_populations(const m3lb_config* cfg, const double* k, const double ub[3],
                        double* f);
/* Fixed-point conversion used by the exact storage sum: value * 2^100
 * truncated toward zero, as {hi, lo}. */
void orc_fixed100(double v, int64_t out[2]);

#ifdef __cplusplus
}
#endif
#endif
