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

/* Velocity sets, esoteric-pull ordering (src/lattice/Lattices.hpp:72-96, 125-159). */
enum { M3LB_D2Q9 = 9, M3LB_D3Q27 = 27 };

/* Collision operators.
 *   NULL        identity collision, the reference's NullCollision
 *               (tests/test_streaming.cpp:28-38): isolates streaming
 *   BGK         BGK<L, SecondOrderEquilibrium<L>, F, S>   (src/collision/BGK.hpp:31-100)
 *   BGK_PRODUCT BGK<L, HighOrderEquilibrium<L>, F, S>     (Equilibrium.hpp:110-136;
 *               the validation campaign's "bgk", validation/Campaign.hpp:36)
 *   TRT         TRT<L, SecondOrderEquilibrium<L>, F, S>   (src/collision/TRT.hpp:33-134)
 *   TRT_PRODUCT TRT<L, HighOrderEquilibrium<L>, F, S>     (the same operator with the
 *               product-form equilibrium; TRT takes any Eq, TRT.hpp:32-33, 86-90)
 *   MRT         MomentCollision<L, F, S, false>           (src/collision/MomentCollision.hpp:380-381)
 *   CM          MomentCollision<L, F, S, true>            (src/collision/MomentCollision.hpp:378-379)
 */
enum {
  M3LB_OP_NULL = 0,
  M3LB_OP_BGK = 1,
  M3LB_OP_BGK_PRODUCT = 2,
  M3LB_OP_TRT = 3,
  M3LB_OP_MRT = 4,
  M3LB_OP_CM = 5,
  M3LB_OP_TRT_PRODUCT = 6
};

/* Stored variable: f (RawPopulations) or g = f - w (ShiftedPopulations),
 * src/memory/Storage.hpp:27-35. */
enum { M3LB_RAW = 0, M3LB_SHIFTED = 1 };

/* Storage dtype of the population array.  F64 computes in FP64; F32 and the
 * 16-bit forms compute in FP32.  The 16-bit forms hold s_q * g_q with one
 * power-of-two scale per opposite pair and require SHIFTED storage: F16 and
 * BF16 as floating point, I16 as fixed point (round to nearest, saturating at
 * +-32767), E2M13 as a narrow 16-bit float: bit 0 the sign, bits 15..14 an
 * exponent e, bits 13..1 a mantissa f, magnitude 2^(4+e) (1 + f/2^13) for
 * e >= 1 and f 2^-8 for e = 0 (so [0, 256): 14 significant bits over
 * [32, 256), a step of 2^-8 below), rounded to nearest even, saturating at
 * the largest code, no infinities or NaNs.
 * The
 * reference selects FP32/FP64 with a configure-time typedef
 * (src/core/Types.hpp:16-20) and has no 16-bit form. */
enum { M3LB_F64 = 0, M3LB_F32 = 1, M3LB_F16 = 2, M3LB_BF16 = 3, M3LB_I16 = 4, M3LB_E2M13 = 5 };

/* Forcing policy: NoForcing, uniform Guo (src/forcing/Forcing.hpp:41-82), or
 * Guo with a per-node force F(n) set by m3lb_set_force_field -- the form of
 * any forcing policy with at(n, F) (e.g. SinForce, validation/forcing_3d.cpp:
 * 64-91), whose F enters the velocity shift and the Guo terms exactly as the
 * uniform force does (Forcing.hpp:489-496). */
enum { M3LB_FORCE_NONE = 0, M3LB_FORCE_GUO = 1, M3LB_FORCE_FIELD = 2 };

/* Cell types (src/boundary/Flags.hpp:20-23).  Solid is halfway bounce-back,
 * the identity on esoteric-pull storage (src/memory/EsotericPull.hpp:27-33). */
enum { M3LB_FLUID = 0, M3LB_SOLID = 1, M3LB_EXCLUDED = 2 };
/* Boundary node types set only by their own calls, never by
 * m3lb_set_geometry (src/boundary/Flags.hpp:24-33): a regularised wall node
 * (m3lb_set_regularized_walls), a ghost specular cell (m3lb_set_specular_walls:
 * permutes its incoming set about the wall normal, no collision, no
 * macroscopic state) and an on-node specular node (m3lb_set_specular_nodes:
 * mirrors its unknown directions, then collides as fluid). */
enum { M3LB_REGWALL = 3, M3LB_SPECWALL = 4, M3LB_SPECNODE = 5 };

/* Wall normal codes of m3lb_set_regularized_walls and m3lb_set_specular_walls
 * (boundary/Regularized.hpp:97-142): outward normals, a corner, and the three
 * outflow closures (rho imposed from the table, or for OUT_FREE from
 * m3lb_set_outflow_density). */
enum {
  M3LB_NRM_NONE = 0, M3LB_NRM_XP = 1, M3LB_NRM_XM = 2, M3LB_NRM_YP = 3, M3LB_NRM_YM = 4,
  M3LB_NRM_ZP = 5, M3LB_NRM_ZM = 6, M3LB_NRM_CORNER = 7, M3LB_NRM_OUT_XP = 8,
  M3LB_NRM_OUT_FREE = 9, M3LB_NRM_OUT_EQ = 10
};
/* Face mask bits of m3lb_set_specular_nodes (Specular.hpp:241-249). */
enum {
  M3LB_SPEC_XP = 1, M3LB_SPEC_XM = 2, M3LB_SPEC_YP = 4, M3LB_SPEC_YM = 8,
  M3LB_SPEC_ZP = 16, M3LB_SPEC_ZM = 32
};

/* Status codes. */
enum {
  M3LB_OK = 0,
  M3LB_EINVAL = -1,     /* bad argument or unsupported combination */
  M3LB_ENOMEM = -2,     /* device allocation failed */
  M3LB_EDEVICE = -3,    /* HIP runtime error */
  M3LB_ESTATE = -4      /* call not valid in the handle's current state */
};

/* Everything FluidSolver's constructor and the collision policy's members
 * carry (FluidSolver.hpp:42-69; BGK.hpp:43-44; TRT.hpp:46-48;
 * MomentCollision.hpp:89-91; Domain.hpp:47-56). */
typedef struct m3lb_config {
  int32_t lattice;       /* M3LB_D2Q9 | M3LB_D3Q27 */
  int32_t op;            /* M3LB_OP_* */
  int32_t storage;       /* M3LB_RAW | M3LB_SHIFTED */
  int32_t dtype;         /* M3LB_F64 | M3LB_F32 | M3LB_F16 | M3LB_BF16 | M3LB_I16 | M3LB_E2M13 */
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

/* Momentum exchange with a body (validation/square_cylinder.cpp:27-50,
 * 340-350): over every link from a Fluid node n to a neighbour n + c_i in the
 * body, F += 2 c_i f(n, opp(i)), f(n, opp(i)) being the population arriving
 * at n along opp(i) (gather_populations, unshifted).  body: one byte per
 * INTERIOR node, x fastest, nonzero = in the body; NULL = every Solid node.
 * FP64 accumulation; force = {Fx, Fy, Fz}.  The reference evaluates the same
 * sum on the host from gather_populations. */
int m3lb_momentum_exchange(m3lb_solver* s, const uint8_t* body, double force[3]);
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
/* set_regularized_walls(fn -> WallSpec) (FluidSolver.hpp:84-103, 372-530):
 * regularised walls (Latt et al. 2008) on the nodes themselves.
 * normals: one M3LB_NRM_* code per INTERIOR node, x fastest -- 0 none, 1..6
 * the OUTWARD normal +x, -x, +y, -y, +z, -z, 7 a corner (density
 * extrapolated along a wall, stress from finite-difference gradients unless
 * m3lb_set_fd_corners(s, 0)), 8 a back-pressure outlet on a +x face (rho from
 * the table, normal velocity from the inverted closure, tangential velocity
 * from upstream), 9 an outlet on any face (the donor neighbour's distribution
 * rescaled to m3lb_set_outflow_density), 10 an equilibrium outlet (rho from
 * the table, velocity from upstream); tags: index into table, the
 * deduplicated (ux, uy, uz, rho) states.  The nodes become M3LB_REGWALL;
 * unknown directions and outflow donors come from the flags as they are at
 * this call (call after m3lb_set_geometry, which removes the walls).  A node
 * that reads a neighbour (the outlets, a corner's stencil) sees an earlier
 * node (lower padded index) after its update in the same step and a later
 * one before it, the reference's serial order.  Codes 8 and 10 need two and
 * one nodes upstream along -x.  A rejected call changes nothing. */
typedef struct m3lb_wall {
  double ux, uy, uz, rho;
} m3lb_wall;
int m3lb_set_regularized_walls(m3lb_solver* s, const uint8_t* normals, const uint16_t* tags,
                                const m3lb_wall* table, int32_t ntable);
/* set_fd_corners(on) (FluidSolver.hpp:272); on by default. */
int m3lb_set_fd_corners(m3lb_solver* s, int32_t on);
/* set_force_correction(on) (:1311): the Guo odd part in the wall
 * reconstruction; on by default. */
int m3lb_set_force_correction(m3lb_solver* s, int32_t on);
/* set_outflow_order(k) (:1324): tangential outlet velocity zero-gradient (1,
 * the default) or linearly extrapolated (2). */
int m3lb_set_outflow_order(m3lb_solver* s, int32_t order);
/* set_outflow_density(rho) (:1330): the density code 9 imposes; 1 by default. */
int m3lb_set_outflow_density(m3lb_solver* s, double rho);
/* set_wall_velocity_scale(sc) (:303-312): every state's velocity is the table
 * as given times sc (absolute, not compounding). */
int m3lb_set_wall_velocity_scale(m3lb_solver* s, double scale);
/* set_wall_density(rho) (:330-335): every state's density slot. */
int m3lb_set_wall_density(m3lb_solver* s, double rho);
/* set_wall_velocity_field(fn) (:353-367): one wall velocity per INTERIOR
 * node, x fastest, read instead of the table (by the density closure, the
 * reconstruction and the macroscopic readout; a corner's stencil keeps the
 * table) until the solver is destroyed; its scale restarts at 1. */
int m3lb_set_wall_velocity_field(m3lb_solver* s, const double* ux, const double* uy,
                                 const double* uz);
int m3lb_set_wall_velocity_field_scale(m3lb_solver* s, double scale);
/* set_specular_walls(fn -> NormalCode) (:193-216): free-slip ghost cells;
 * an axis normal 1..6 per INTERIOR node marks the cell M3LB_SPECWALL, 0 leaves
 * it.  Call after m3lb_set_geometry.  Codes above 6 are rejected. */
int m3lb_set_specular_walls(m3lb_solver* s, const uint8_t* normals);
/* set_specular_nodes(fn -> SpecFace mask) (:236-270): on-node free-slip
 * nodes (M3LB_SPECNODE), M3LB_SPEC_* bits per INTERIOR node; edges and
 * corners carry several faces, never both faces of one axis. */
int m3lb_set_specular_nodes(m3lb_solver* s, const uint8_t* faces);
/* A forcing policy with a per-node F(n) (at(n, F), e.g. SinForce,
 * validation/forcing_3d.cpp:64-91): for handles created with
 * M3LB_FORCE_FIELD, three padded arrays of n_padded values (only Fluid nodes
 * read them), converted to the compute type.  Until set, F = 0.  M3LB_ESTATE
 * for other handles. */
int m3lb_set_force_field(m3lb_solver* s, const double* fx, const double* fy, const double* fz);
int m3lb_set_omega(m3lb_solver* s, double omega, double omega_minus, double omega_bulk);

/* Per-direction scales of 16-bit storage: the stored code is s_q * g_q.
 * Setting them re-encodes the array in place.  Each scale is a power of two
 * in 2^-60 .. 2^60, shared by opposite directions (else M3LB_EINVAL).
 * Ignored for F32/F64. */
int m3lb_get_scales(m3lb_solver* s, double* scales_q);
int m3lb_set_scales(m3lb_solver* s, const double* scales_q);
/* Re-chooses the 16-bit scales from the current field: per opposite pair the
 * power of two that maps max |g_q| into (2^7, 2^8] for F16/BF16 (256x
 * headroom under FP16's largest finite value; a pair below 1e-3 w_q counts
 * as 1e-3 w_q), and for I16 and E2M13 the power of two that maps r w_q into
 * (2^11, 2^12] (I16, 8x headroom under saturation) or (32, 64] (E2M13, 4x),
 * with r the largest max |g_q| / w_q over all directions, at least 0.1.
 * Runs after every initialize; callers may repeat it on a schedule.  No-op
 * for F32/F64. */
int m3lb_calibrate_scales(m3lb_solver* s);

/* Runs nsteps steps bracketed by HIP events recorded on the handle's stream
 * and returns the elapsed milliseconds.  Used by the benchmark.  No
 * counterpart in the reference. */
int m3lb_time_steps(m3lb_solver* s, int32_t nsteps, double* ms);

/* Order in which the step kernel's tiles are assigned to the GPU's 8 XCDs:
 * 0 = hardware order, -1 = each XCD takes one contiguous eighth of the box,
 * run > 0 = XCDs take turns with runs of `run` consecutive tiles (default 64
 * for 4- and 8-byte populations, 8 for E2M13, 2 for the other 16-bit forms).
 * A tuning knob only: results are bit-identical for every value.  No
 * counterpart in the reference. */
int m3lb_set_tile_order(m3lb_solver* s, int32_t run);

/* How the step kernel accesses the slots a node owns at its own index:
 * nontemporal (STREAM) or plain (PLAIN).  AUTO, the default, takes PLAIN when
 * the population array fits the 256 MB last-level cache (there it streams
 * about 25% faster) and STREAM above.  A tuning knob only: results are
 * bit-identical for every value.  No counterpart in the reference. */
enum { M3LB_CACHE_AUTO = 0, M3LB_CACHE_STREAM = 1, M3LB_CACHE_PLAIN = 2 };
int m3lb_set_cache_policy(m3lb_solver* s, int32_t policy);

/* Two steps in one pass over the population array.  Step t+1 follows step t
 * a few z-planes behind inside one persistent launch, so that it reads step
 * t's output from the caches, and a short second launch finishes the rows
 * whose neighbours another stripe of the box updates: a pair of steps then
 * moves about the HBM bytes of one.  The result is bit-identical to one launch
 * per step.  No counterpart in the reference.
 *   mode  M3LB_FUSION_AUTO (default): pair the steps of FP32 and FP64 boxes
 *         whose population array exceeds the 256 MB last-level cache (the
 *         16-bit forms run faster one launch per step);
 *         M3LB_FUSION_OFF: never; M3LB_FUSION_ON: whenever the box allows it.
 * A box allows it when it is D3Q27 without a per-node force, has no
 * regularised walls or specular cells, and nz >= lag + 3.  The tuning values
 * take 0 for their default: stripes (y-slabs of the box, 1..64, at least 3
 * rows each; default 8, one per XCD), lag (planes step t+1 trails step t by;
 * default from the box: 2 at 512^3, 5 at 256^3), blocks_per_cu (1..8;
 * default 4) and group (256-node tiles per ticket, 1..64; default 1).  A pair
 * never covers the step that stores the macroscopic fields.  A pass whose
 * bounded waits give up makes the next synchronising call on the handle fail
 * with M3LB_EDEVICE. */
enum { M3LB_FUSION_AUTO = 0, M3LB_FUSION_OFF = 1, M3LB_FUSION_ON = 2 };
int m3lb_set_step_fusion(m3lb_solver* s, int32_t mode, int32_t stripes, int32_t lag,
                         int32_t blocks_per_cu, int32_t group);
/* total_mass and total_momentum(0..2) at once, out = {mass, Mx, My, Mz},
 * bit-identical to those two calls, without allocating: the per-block
 * partials live in a static device array of 4096 * 4 + 64 doubles (131,584
 * bytes), one per GPU, kept until the process ends.  No counterpart in the
 * reference; for callers that cannot allocate device memory (tight serving
 * loops, code that forbids hipMalloc).  Calls that overlap on one GPU share
 * that array and race; one call at a time per GPU. */
int m3lb_totals_static(m3lb_solver* s, double out[4]);

/* Pairs of steps the handle has run as one pass since creation. */
int m3lb_fused_passes(const m3lb_solver* s, uint64_t* passes);
/* Failure-path test hook: with on != 0 every wait of a two-step pass is made
 * unsatisfiable and gives up after a few polls, so the pass drains at once
 * and the next synchronising call reports M3LB_EDEVICE.  Never set outside
 * tests of that error path; off by default. */
int m3lb_test_fail_fused_waits(m3lb_solver* s, int32_t on);

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

/* Identity of the kernels in this library: a hash of every device source
 * file, this header and the compile flags, fixed at build time.  A stored
 * measurement (a traffic count, a profile) records the id it was taken on,
 * and is only reused while the loaded library reports the same id.  No
 * counterpart in the reference. */
const char* m3lb_kernel_build_id(void);

/* Device bytes held by the handle (populations + flags + macroscopic). */
int m3lb_device_bytes(const m3lb_solver* s, uint64_t* bytes);
/* initialize_field with the Taylor-Green vortex of
 * validation/tgv3d_bench.cpp:79-90 evaluated on the device: w = 0,
 *   u = u0 sin X cos Y cos Z,  v = -u0 cos X sin Y cos Z,
 *   rho = 1 + 3 u0^2/16 (cos 2X + cos 2Y)(cos 2Z + 2),
 * X = 2 pi (x - hx) / nx and likewise Y, Z with their own extents (for a cube,
 * the reference's field).  Non-fluid nodes at rest with rho = 1. */
int m3lb_initialize_taylor_green(m3lb_solver* s, double u0);

/* ------------------------------------------------------------------------
 * Batched serving.  No counterpart in the reference.  A server keeps many
 * independent boxes (requests) resident in one device arena of a fixed byte
 * budget and advances all of them together: each step is one persistent
 * kernel over the packed tile list of every resident box; a chunk of
 * steps_per_chunk steps is one hipGraph launch.  Requests are admitted first
 * come, first served as slots and memory free up, and complete
 * independently.  Behind a queue head that does not fit yet, a later request
 * is admitted only when its last step comes no later than the step after
 * which the head fits (every resident box advances one step per step), so
 * passing never delays the head; at most max_bypass pass one head.  A box computes bit-identical results alone, in any mix, and
 * in the single-box solver with the same configuration and forcing policy.
 * Every box of a server shares one kernel configuration (lattice, operator,
 * storage, dtype, forcing policy).  16-bit storage (shifted only) chooses
 * each box's scales at admission as m3lb_initialize does (seed, measure,
 * seed again), which synchronises the server's stream once per admitted box.
 * One host thread per server; calls are ordered on the server's own stream.
 * ------------------------------------------------------------------------ */
enum { M3LB_SHAPE_ROWS = 0, M3LB_SHAPE_LINEAR = 1, M3LB_SHAPE_AUTO = 2 };
enum { M3LB_TILE_BLOCK_RANGES = -2 };
typedef struct m3lb_server_config {
  int32_t lattice, op, storage, dtype; /* as in m3lb_config */
  int32_t forcing;         /* M3LB_FORCE_NONE | M3LB_FORCE_GUO (force per request) */
  int32_t max_boxes;       /* resident slots, 1 .. 65536 */
  int32_t steps_per_chunk; /* steps per launch sequence, 1 .. 1024 */
  int32_t use_graph;       /* 1: a full chunk is one hipGraph launch */
  int32_t blocks_per_cu;   /* persistent grid = CUs * blocks_per_cu, 1 .. 64 */
  int32_t tile_run;        /* order of the packed tiles: >= -1 as in
                              m3lb_set_tile_order (one box lookup per tile), or
                              M3LB_TILE_BLOCK_RANGES (one contiguous range of
                              the packed list per block) */
  int32_t tile_shape;      /* M3LB_SHAPE_*: ROWS = whole x-rows per tile (as the
                              single-box step), LINEAR = 256 consecutive
                              interior nodes, AUTO = ROWS when nx % 64 == 0 */
  int32_t max_bypass;      /* admissions allowed past one waiting queue head
                              (each also bounded as above) */
  uint64_t budget_bytes;   /* device arena for resident boxes */
  uint64_t max_resident_nodes; /* admission stops while the resident boxes hold
                              this many interior nodes (one box is always
                              admitted into an empty server); 0 = no limit.
                              Bounds the work shared per step, so a request's
                              time to completion stays near its own run time
                              once the device is saturated. */
} m3lb_server_config;
enum { M3LB_INIT_UNIFORM = 0, M3LB_INIT_TAYLOR_GREEN = 1 };
enum { M3LB_WALLS_NONE = 0, M3LB_WALLS_Y = 1 };
typedef struct m3lb_request {
  int32_t nx, ny, nz;
  int32_t periodic[3];
  int32_t init;            /* M3LB_INIT_UNIFORM (rho0, u) or M3LB_INIT_TAYLOR_GREEN
                              (amplitude u[0], as m3lb_initialize_taylor_green) */
  int32_t walls;           /* M3LB_WALLS_Y: interior rows y = 0, ny-1 Solid;
                              ignored when flags are passed to submit */
  int32_t steps;           /* >= 1 */
  int32_t keep;            /* 1: the box stays resident after completion for
                              the fetch calls, until m3lb_srv_release */
  double omega, omega_minus, omega_bulk;
  double force[3];         /* Guo force (servers with M3LB_FORCE_GUO) */
  double rho0, u[3];
  double arrival;          /* server-clock seconds; < 0: the time of submit */
} m3lb_request;
enum { M3LB_REQ_QUEUED = 0, M3LB_REQ_RESIDENT = 1, M3LB_REQ_DONE = 2, M3LB_REQ_CANCELLED = 3 };
typedef struct m3lb_result {
  int64_t id;
  int32_t state;           /* M3LB_REQ_* */
  int32_t steps;           /* steps completed */
  double arrival, admitted, completed; /* server-clock seconds */
  double mass, momentum[3]; /* at completion: sums over fluid nodes, FP64 */
} m3lb_result;
typedef struct m3lb_server_stats {
  int64_t submitted, rejected, queued, resident, done, cancelled;
  int64_t chunks, steps;   /* chunks and batched steps issued */
  int64_t lattice_updates; /* interior nodes x steps of retired chunks */
  uint64_t bytes_used, bytes_peak, bytes_budget;
  double device_seconds;   /* retired chunks, event-timed on the stream */
  int32_t grid, tile_run, tile_shape, steps_per_chunk;
  uint64_t resident_nodes; /* interior nodes of the resident boxes */
} m3lb_server_stats;
typedef struct m3lb_tune_result {
  int32_t tile_shape, tile_run, blocks_per_cu, candidates;
  double best_glups;       /* chosen setting on the tuning mix */
  double initial_glups;    /* the server's setting before tuning */
} m3lb_tune_result;
typedef struct m3lb_server m3lb_server;
/* Allocates the arena (budget_bytes) on the current device. */
int m3lb_srv_create(const m3lb_server_config* cfg, m3lb_server** out);
int m3lb_srv_destroy(m3lb_server* srv);
/* Arena bytes one request needs (populations + flags); no device call. */
int m3lb_srv_footprint(const m3lb_server_config* cfg, const m3lb_request* req,
                       uint64_t* bytes);
/* Queues a request.  interior_flags (nx*ny*nz bytes, x fastest, values as
 * m3lb_set_geometry) may be NULL.  A request larger than the whole budget is
 * refused with M3LB_EINVAL; one that does not fit now waits in the queue. */
int m3lb_srv_submit(m3lb_server* srv, const m3lb_request* req,
                    const uint8_t* interior_flags, int64_t* id);
/* Retires finished chunks, admits queued requests, issues the next chunks
 * (at most two in flight).  wait_us > 0 waits up to that long for a chunk to
 * retire when none has, < 0 waits for one, 0 returns at once.
 * *n_completed (may be NULL) = requests completed by this call. */
int m3lb_srv_progress(m3lb_server* srv, int32_t wait_us, int32_t* n_completed);
/* Runs progress until no request is queued or resident-and-unfinished. */
int m3lb_srv_drain(m3lb_server* srv);
/* Pops up to max records of requests completed or cancelled since the last
 * poll.  A popped record of a request without keep is forgotten. */
int m3lb_srv_poll(m3lb_server* srv, m3lb_result* out, int32_t max, int32_t* n);
int m3lb_srv_result(m3lb_server* srv, int64_t id, m3lb_result* out);
/* Fields of a completed request submitted with keep = 1, as
 * m3lb_compute_macroscopic / m3lb_gather_populations of its box. */
int m3lb_srv_fetch_macroscopic(m3lb_server* srv, int64_t id, double* rho, double* ux,
                               double* uy, double* uz);
int m3lb_srv_fetch_populations(m3lb_server* srv, int64_t id, double* out);
/* Frees a kept box and forgets its request. */
int m3lb_srv_release(m3lb_server* srv, int64_t id);
/* Removes a queued or resident request; its memory returns to the arena. */
int m3lb_srv_cancel(m3lb_server* srv, int64_t id);
int m3lb_srv_stats(m3lb_server* srv, m3lb_server_stats* out);
int m3lb_srv_clock(m3lb_server* srv, double* seconds);
/* Times every tile shape x tile order x blocks-per-CU candidate on a fixed mix
 * of Taylor-Green boxes (edges log-uniform on [32, 192]) for `steps` steps
 * each and keeps the fastest.  Only on an idle server (no request queued or
 * resident).  Results of later requests do not depend on the choice. */
int m3lb_srv_autotune(m3lb_server* srv, int32_t steps, m3lb_tune_result* out);

#ifdef __cplusplus
}
#endif

#endif /* M3LB_H */
