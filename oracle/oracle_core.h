/* oracle_core.h -- declarations of the two precision instantiations of
 * oracle_core.c (TEST INFRASTRUCTURE ONLY, see oracle.h). Generated list,
 * kept by hand. */
#ifndef M3LB_ORACLE_CORE_H
#define M3LB_ORACLE_CORE_H
#include <stdint.h>
#include "../include/m3lb.h"

void* core64_create(const m3lb_config* cfg, int streaming);
void core64_destroy(void* p);
void core64_set_threads(void* p, int n);
void core64_domain(const void* p, int64_t out[10]);
void core64_set_geometry(void* p, const uint8_t* fl);
void core64_initialize(void* p, double rho0, double ux, double uy, double uz);
void core64_initialize_field(void* p, const double* rho, const double* ux, const double* uy, const double* uz);
void core64_seed_populations(void* p, const double* in);
void core64_gather_populations(void* p, double* out);
void core64_storage(void* p, double* out);
void core64_step(void* p, int nsteps);
double core64_time_steps(void* p, int nsteps);
uint64_t core64_timestep(const void* p);
void core64_set_force(void* p, double fx, double fy, double fz);
void core64_set_force_field(void* p, const double* fx, const double* fy, const double* fz);
int core64_set_regularized_walls(void* p, const uint8_t* normals, const uint16_t* tags,
                                 const double* table, int ntab);
void core64_set_fd_corners(void* p, int on);
void core64_set_force_correction(void* p, int on);
void core64_set_outflow_order(void* p, int k);
void core64_set_outflow_density(void* p, double r);
void core64_set_wall_velocity_scale(void* p, double sc);
void core64_set_wall_density(void* p, double r);
void core64_set_wall_velocity_field(void* p, const double* ux, const double* uy, const double* uz);
void core64_set_wall_velocity_field_scale(void* p, double sc);
int core64_set_specular_walls(void* p, const uint8_t* normals);
int core64_set_specular_nodes(void* p, const uint8_t* faces);
void core64_compute_macroscopic(void* p, double* rho, double* ux, double* uy, double* uz);
double core64_total_mass(void* p);
double core64_total_momentum(void* p, int axis);
void core64_storage_sum_exact(void* p, int64_t out[2]);
void core64_energy_enstrophy(void* p, double* e_out, double* w_out, int* finite);
void core64_collide_node(const m3lb_config* cfg, const double* fin, double* fout, double mo[4]);
void core64_macroscopic_node(const m3lb_config* cfg, const double* fin, double mo[4]);
void core64_seed_node(const m3lb_config* cfg, double rho, double ux, double uy, double uz, double* out);
void core64_equilibrium(const m3lb_config* cfg, int product, int shifted, double rho, double ux, double uy, double uz, double* out);
void core64_to_moments(const m3lb_config* cfg, const double* fin, const double ub[3], double* kout);
void core64_to_populations(const m3lb_config* cfg, const double* kin, const double ub[3], double* fout);

void* core32_create(const m3lb_config* cfg, int streaming);
void core32_destroy(void* p);
void core32_set_threads(void* p, int n);
void core32_domain(const void* p, int64_t out[10]);
void core32_set_geometry(void* p, const uint8_t* fl);
void core32_initialize(void* p, double rho0, double ux, double uy, double uz);
void core32_initialize_field(void* p, const double* rho, const double* ux, const double* uy, const double* uz);
void core32_seed_populations(void* p, const double* in);
void core32_gather_populations(void* p, double* out);
void core32_storage(void* p, double* out);
void core32_step(void* p, int nsteps);
double core32_time_steps(void* p, int nsteps);
uint64_t core32_timestep(const void* p);
void core32_set_force(void* p, double fx, double fy, double fz);
void core32_set_force_field(void* p, const double* fx, const double* fy, const double* fz);
int core32_set_regularized_walls(void* p, const uint8_t* normals, const uint16_t* tags,
                                 const double* table, int ntab);
void core32_set_fd_corners(void* p, int on);
void core32_set_force_correction(void* p, int on);
void core32_set_outflow_order(void* p, int k);
void core32_set_outflow_density(void* p, double r);
void core32_set_wall_velocity_scale(void* p, double sc);
void core32_set_wall_density(void* p, double r);
void core32_set_wall_velocity_field(void* p, const double* ux, const double* uy, const double* uz);
void core32_set_wall_velocity_field_scale(void* p, double sc);
int core32_set_specular_walls(void* p, const uint8_t* normals);
int core32_set_specular_nodes(void* p, const uint8_t* faces);
void core32_compute_macroscopic(void* p, double* rho, double* ux, double* uy, double* uz);
double core32_total_mass(void* p);
double core32_total_momentum(void* p, int axis);
void core32_storage_sum_exact(void* p, int64_t out[2]);
void core32_energy_enstrophy(void* p, double* e_out, double* w_out, int* finite);
void core32_collide_node(const m3lb_config* cfg, const double* fin, double* fout, double mo[4]);
void core32_macroscopic_node(const m3lb_config* cfg, const double* fin, double mo[4]);
void core32_seed_node(const m3lb_config* cfg, double rho, double ux, double uy, double uz, double* out);
void core32_equilibrium(const m3lb_config* cfg, int product, int shifted, double rho, double ux, double uy, double uz, double* out);
void core32_to_moments(const m3lb_config* cfg, const double* fin, const double ub[3], double* kout);
void core32_to_populations(const m3lb_config* cfg, const double* kin, const double ub[3], double* fout);

#endif
