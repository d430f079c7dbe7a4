/*
 * m3lb.h -- C-ABI of the MI355X (gfx950) fused esoteric-pull stream+collide
 * step.  Plain C types only: opaque handle, pointers, sizes and int status.
 *
 * The reference has no FFI for this path: its "operator API" is the set of
 * compile-time policies consumed by
 *     FluidSolver<L, EsotericPull<L>, Collision>   (src/solver/FluidSolver.hpp:34-35)
 * so every entry point below names the FluidSolver member (or policy) it stands
 * in for.  Per-node arrays use the reference's padded node index
 *     n = (pz*sy + py)*sx + px                     (src/grid/Domain.hpp:60-62)
 * with halo layers only on non-periodic axes (src/grid/Domain.hpp:104-108),
 * so a View1D/View2D of the reference and an array here hold the same value at
 * the same index.  Population arrays are (n, q) ordered, q fastest, as the
 * reference's host mirrors are.
 *
 * Status: 0 on success, a negative M3LB_E* code otherwise; the message of the
 * last failure on the calling thread is returned by m3lb_last_error().  The
 * reference aborts or throws instead (FluidSolver.hpp:204-209, 524-526).
 *
 * Threading: one host thread per handle.  Every call on a handle is ordered on
 * that handle's own HIP stream.  m3lb_step is asynchronous, as the Kokkos
 * device step is (no fence in FluidSolver::step, FluidSolver.hpp:582-591);
 * every call that returns host data synchronises the stream first.
 */
#ifndef M3LB_H
#define M3LB_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Velocity sets, esoteric-pull ordering (src/lattice/Lattices.hpp:130-214). */
enum { M3LB_D2Q9 = 9, M3LB_D3Q27 = 27 };

/* Collision operators.
 *   NULL        identity collision, the reference's NullCollision
 *               (tests/test_streaming.cpp:28-38): isolates streaming
 *   BGK         BGK<L, SecondOrderEquilibrium<L>, F, S>   (src/collision/BGK.hpp:169-238)
 *   BGK_PRODUCT BGK<L, HighOrderEquilibrium<L>, F, S>     (Equilibrium.hpp:110-136;
 *               the validation campaign's "bgk", validation/Campaign.hpp:36)
 *   TRT         TRT<L, SecondOrderEquilibrium<L>, F, S>   (src/collision/TRT.hpp:33-134)
 *   MRT         MomentCollision<L, F, S, false>           (src/collision/MomentCollision.hpp:380-381)
 *   CM          MomentCollision<L, F, S, true>            (src/collision/MomentCollision.hpp:378-379)
 */
enum {
  M3LB_OP_NULL = 0,
  M3LB_OP_BGK = 1,
  M3LB_OP_BGK_PRODUCT = 2,
  M3LB_OP_TRT = 3,
  M3LB_OP_MRT = 4,
  M3LB_OP_CM = 5
};

/* Stored variable: f (RawPopulations) or g = f - w (ShiftedPopulations),
 * src/memory/Storage.hpp:27-35. */
enum { M3LB_RAW = 0, M3LB_SHIFTED = 1 };

/* Storage dtype of the population array.  F64 computes in FP64; F32, F16 and
 * BF16 compute in FP32.  F16/BF16 hold s_q * g_q with one scale per direction
 * and require SHIFTED storage.  The reference selects FP32/FP64 with a
 * configure-time typedef (src/core/Types.hpp:16-20) and has no 16-bit form. */
enum { M3LB_F64 = 0, M3LB_F32 = 1, M3LB_F16 = 2, M3LB_BF16 = 3 };

/* Forcing policy: NoForcing or uniform Guo (src/forcing/Forcing.hpp:41-82). */
enum { M3LB_FORCE_NONE = 0, M3LB_FORCE_GUO = 1 };

/* Cell types (src/boundary/Flags.hpp:20-23).  Solid is halfway bounce-back,
 * the identity on esoteric-pull storage (src/memory/EsotericPull.hpp:27-33). */
enum { M3LB_FLUID = 0, M3LB_SOLID = 1, M3LB_EXCLUDED = 2 };

/* Status codes. */
enum {
  M3LB_OK = 0,
  M3LB_EINVAL = -1,     /* bad argument or unsupported combination */
  M3LB_ENOMEM = -2,     /* device allocation failed */
  M3LB_EDEVICE = -3,    /* HIP runtime error */
  M3LB_ESTATE = -4      /* call not valid in the handle's current state */
};

/* Everything FluidSolver's constructor and the collision policy's members
 * carry (FluidSolver.hpp:42-69; BGK.hpp:181-182; TRT.hpp:46-48;
 * MomentCollision.hpp:89-91; Domain.hpp:47-56). */
typedef struct m3lb_config {
  int32_t lattice;       /* M3LB_D2Q9 | M3LB_D3Q27 */
  int32_t op;            /* M3LB_OP_* */
  int32_t storage;       /* M3LB_RAW | M3LB_SHIFTED */
  int32_t dtype;         /* M3LB_F64 | M3LB_F32 | M3LB_F16 | M3LB_BF16 */
  int32_t nx, ny, nz;    /* interior extents; nz = 1 for a 2D run */
  int32_t periodic[3];   /* 1 = periodic along x, y, z */
  double omega;          /* BGK omega, TRT omega_plus, moment shear omega */
  double omega_minus;    /* TRT omega_minus (ignored otherwise) */
  double omega_bulk;     /* moment ops: trace rate, < 0 follows omega */
  int32_t forcing;       /* M3LB_FORCE_* */
  int32_t reserved;      /* must be 0 */
  double force[3];       /* Guo force density (fx, fy, fz) */
} m3lb_config;

typedef struct m3lb_solver m3lb_solver;

/* Domain extents as the reference computes them (Domain.hpp:47-56):
 * out = {nx, ny, nz, hx, hy, hz, sx, sy, sz, n_padded}. */
int m3lb_domain(const m3lb_config* cfg, int64_t out[10]);

/* FluidSolver(Domain, Collision) (FluidSolver.hpp:42-69): allocates the one
 * in-place population array, the flag array (halo Excluded, interior Fluid)
 * and the macroscopic arrays on the current HIP device. */
int m3lb_create(const m3lb_config* cfg, m3lb_solver** out);
int m3lb_destroy(m3lb_solver* s);
const char* m3lb_last_error(void);

/* set_geometry(fn(x,y,z) -> CellType) (FluidSolver.hpp:74-82): one flag per
 * INTERIOR node, x fastest (nx*ny*nz bytes). */
int m3lb_set_geometry(m3lb_solver* s, const uint8_t* interior_flags);

/* initialize(rho0, u0) (FluidSolver.hpp:553-561): t = 0, fluid seeded with
 * the collision's seed_value at (rho0, u0), Solid and halo at rest. */
int m3lb_initialize(m3lb_solver* s, double rho0, double ux, double uy, double uz);

/* initialize_field(fn(n) -> FlowState) (FluidSolver.hpp:568-579): padded
 * arrays of n_padded values; only Fluid entries are read, the rest are seeded
 * at rest with rho = 1. */
int m3lb_initialize_field(m3lb_solver* s, const double* rho, const double* ux,
                          const double* uy, const double* uz);

/* step(store_macroscopic), called nsteps times (FluidSolver.hpp:582-591). */
int m3lb_step(m3lb_solver* s, int32_t nsteps, int32_t store_macroscopic);

/* compute_macroscopic() + rho()/ux()/uy()/uz() (FluidSolver.hpp:597-599,
 * 611-614, 1042-1203): padded arrays, 0 on non-fluid nodes.  Any pointer may
 * be NULL.  After a step with store_macroscopic = 1, m3lb_macroscopic_stored
 * returns what that step wrote without recomputing. */
int m3lb_compute_macroscopic(m3lb_solver* s, double* rho, double* ux, double* uy,
                             double* uz);
int m3lb_macroscopic_stored(m3lb_solver* s, double* rho, double* ux, double* uy,
                            double* uz);

/* total_mass() / total_momentum(a) (FluidSolver.hpp:601-606, 1205-1229):
 * sum of rho (rho*u_a) over fluid nodes, accumulated in FP64. */
int m3lb_total_mass(m3lb_solver* s, double* out);
int m3lb_total_momentum(m3lb_solver* s, int32_t axis, double* out);

/* Sum of every stored slot (n_padded * Q values, solid- and halo-owned slots
 * included) as a two's-complement 128-bit fixed-point integer with 100
 * fraction bits, each value truncated toward zero before an exact integer
 * sum: out = {high 64 bits, low 64 bits}.  Independent of summation order,
 * so any permutation of the slots leaves it bit-identical.  For 16-bit
 * storage the summed value is the stored code divided by its direction's
 * scale.  No counterpart in the reference. */
int m3lb_storage_sum_exact(m3lb_solver* s, int64_t out[2]);

/* validation/Campaign.hpp:99-130 on the device: E = sum 1/2|u|^2 and
 * Omega = sum 1/2|curl u|^2 over interior nodes, central differences with
 * periodic wrap, FP64 accumulation; *finite = 0 if any u is not finite. */
int m3lb_energy_enstrophy(m3lb_solver* s, double* energy, double* enstrophy,
                          int32_t* finite);

/* gather_populations() / seed_populations(fn) (FluidSolver.hpp:625-637,
 * 1231-1253): the logical population travelling in direction q that arrives
 * at node n, (n, q) ordered, n_padded * Q doubles.  Seeding writes each slot
 * from the highest-indexed node whose load reads it, which is what the
 * reference's ascending scatter leaves behind. */
int m3lb_gather_populations(m3lb_solver* s, double* out);
int m3lb_seed_populations(m3lb_solver* s, const double* in);

/* EsotericPull::storage() (src/memory/EsotericPull.hpp:123): the raw slot
 * image, (n, q) ordered, each stored value converted exactly to double (for
 * 16-bit storage: the code divided by its scale). */
int m3lb_storage(m3lb_solver* s, double* out);

/* timestep() (FluidSolver.hpp:616). */
int m3lb_timestep(const m3lb_solver* s, uint64_t* t);

/* collision().forcing = Guo{fx,fy,fz} and collision().omega* at run time. */
int m3lb_set_force(m3lb_solver* s, double fx, double fy, double fz);
int m3lb_set_omega(m3lb_solver* s, double omega, double omega_minus, double omega_bulk);

/* Per-direction scales of 16-bit storage: the stored code is s_q * g_q.
 * Setting them re-encodes the array in place.  Ignored for F32/F64. */
int m3lb_get_scales(m3lb_solver* s, double* scales_q);
int m3lb_set_scales(m3lb_solver* s, const double* scales_q);
/* Re-chooses the 16-bit scales from the current field: per opposite pair the
 * largest power of two keeping max |g_q| * s_q below 2^14.  Runs after every
 * initialize; callers may repeat it on a schedule.  No-op for F32/F64. */
int m3lb_calibrate_scales(m3lb_solver* s);

/* Runs nsteps steps bracketed by HIP events recorded on the handle's stream
 * and returns the elapsed milliseconds.  Used by the benchmark.  No
 * counterpart in the reference. */
int m3lb_time_steps(m3lb_solver* s, int32_t nsteps, double* ms);

/* Order in which the step kernel's tiles are assigned to the GPU's 8 XCDs:
 * 0 = hardware order, -1 = each XCD takes one contiguous eighth of the box,
 * run > 0 = XCDs take turns with runs of `run` consecutive tiles (default 64
 * for 4- and 8-byte populations, 2 for 16-bit storage).
 * A tuning knob only: results are bit-identical for every value.  No
 * counterpart in the reference. */
int m3lb_set_tile_order(m3lb_solver* s, int32_t run);

/* The block -> tile map the step kernel uses for `grid` blocks and a tile
 * order `run` (host-side copy of the device function): out[b] = tile of block
 * b.  A permutation of [0, grid) for every value; exposed so that this can be
 * checked without a GPU. */
int m3lb_tile_order_map(uint32_t grid, int32_t run, uint32_t* out);

/* Streams `bytes` of device memory in place (read + write, one float4 per
 * thread, nontemporal or not), `reps` times, and returns the fastest time in
 * milliseconds: the streaming ceiling of the device, which the step kernel is
 * measured against, and a copy kernel of known size for calibrating HBM
 * traffic counters.  Allocates and frees its own buffer.  No counterpart in
 * the reference. */
int m3lb_stream_probe(uint64_t bytes, int32_t nontemporal, int32_t reps, double* best_ms);

/* Device bytes held by the handle (populations + flags + macroscopic). */
int m3lb_device_bytes(const m3lb_solver* s, uint64_t* bytes);

#ifdef __cplusplus
}
#endif

#endif /* M3LB_H */
