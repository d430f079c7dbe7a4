// launch.hpp -- host launchers of one kernel configuration, gathered into a
// table the solver selects at creation from the runtime config.  This table is
// the only runtime dispatch: every kernel below is specialised at compile time
// on lattice, operator, storage, dtype, forcing, parity and geometry, as the
// reference's FluidSolver / collision templates are (FluidSolver.hpp:34-35,
// 582-591).
#pragma once

#include "kernels.hpp"
#include "pair.hpp"

namespace m3lb {

struct KernelSet {
  void (*step)(const StepArgs&, int parity, bool geom, unsigned grid, hipStream_t);
  void (*seed)(const SeedArgs&, hipStream_t);
  void (*gather)(const ReadArgs&, hipStream_t);
  void (*storage)(const ReadArgs&, hipStream_t);
  void (*macro)(const ReadArgs&, hipStream_t);
  void (*reduce_mass)(const ReadArgs&, unsigned grid, hipStream_t);
  void (*sum_exact)(const ReadArgs&, unsigned grid, hipStream_t);
  void (*absmax)(const ReadArgs&, unsigned grid, hipStream_t);
  void (*rescale)(void* f, uint64_t stride, int np, const Ratio&, unsigned grid, hipStream_t);
  void (*exchange)(const ReadArgs&, const uint8_t* body, unsigned grid, hipStream_t);
  // regularised walls: corner-density pass (both halves) and the wall update
  void (*wall_density)(const WallArgs&, int parity, hipStream_t);
  void (*walls)(const WallArgs&, int parity, hipStream_t);
  // the neighbour sets wall nodes read, taken before the step
  void (*snapshot)(const SnapArgs&, int parity, hipStream_t);
  // specular ghosts and on-node specular nodes
  void (*specular)(const SpecArgs&, int parity, hipStream_t);
  // batched server step; null for the identity op and a per-node force
  void (*batched)(const BatchArgs&, unsigned grid, hipStream_t);
  // two steps in one pass (pair.hpp): launch 1 on a persistent grid, then the
  // deferred rows; null where not instantiated (D2Q9, per-node force)
  void (*pair)(const PairArgs&, int parity, bool geom, unsigned grid, unsigned tail_grid,
               unsigned nclear, unsigned* host_err, hipStream_t);
  int elem_bytes;    // bytes per stored population
  int real_bytes;    // bytes of the compute type
};

inline unsigned blocks_for(long long n) { return unsigned((n + kBlock - 1) / kBlock); }

template <class K>
struct Launch {
  static constexpr bool kBatched = K::OP != M3LB_OP_NULL && !K::FIELD;
  // Occupancy cap of the streaming step: the kernel uses no LDS, but a launch
  // that reserves a.lds_cap bytes of it per block (a quarter of the CU's LDS
  // plus one byte, set up by the solver from the device's attributes) fits at
  // most 3 blocks (12 waves) per CU.  Measured for D3Q27 CM FP32 at 512^3,
  // interleaved runs: 25.0 GLUPS at 3 or 4 blocks per CU, 24.8 at 2,
  // 24.4-24.55 uncapped (8 waves per SIMD).  The 16-bit forms lose at low
  // occupancy (FP16 47.5 -> 41.1, E2M13 44.6 -> 34.0 at 3 blocks), so only
  // the measured case is capped: D3Q27 CM with 4-byte storage, no geometry
  // and no per-node force, on the nontemporal path.
  static constexpr bool kCapped = sizeof(typename K::T) == 4 && K::Q == 27 &&
                                  K::OP == M3LB_OP_CM && !K::FIELD;
  template <int P, bool G>
  static void step_p(const StepArgs& a, unsigned grid, hipStream_t s) {
    const size_t lds = (kCapped && !G) ? a.lds_cap : 0;
    if (a.cached) k_step<K, P, G, false><<<grid, kBlock, 0, s>>>(a);
    else k_step<K, P, G, true><<<grid, kBlock, lds, s>>>(a);
  }
  static void step(const StepArgs& a, int parity, bool geom, unsigned grid, hipStream_t s) {
    if (parity == 0) {
      if (geom) step_p<0, true>(a, grid, s);
      else step_p<0, false>(a, grid, s);
    } else {
      if (geom) step_p<1, true>(a, grid, s);
      else step_p<1, false>(a, grid, s);
    }
  }
  static void seed(const SeedArgs& a, hipStream_t s) {
    k_seed<K><<<blocks_for(a.d.np), kBlock, 0, s>>>(a);
  }
  static void gather(const ReadArgs& a, hipStream_t s) {
    k_gather<K><<<blocks_for(a.d.np), kBlock, 0, s>>>(a);
  }
  static void storage(const ReadArgs& a, hipStream_t s) {
    k_storage<K><<<blocks_for(a.d.np), kBlock, 0, s>>>(a);
  }
  static void macro(const ReadArgs& a, hipStream_t s) {
    k_macro<K><<<blocks_for(a.d.np), kBlock, 0, s>>>(a);
  }
  static void reduce_mass(const ReadArgs& a, unsigned grid, hipStream_t s) {
    k_reduce_mass<K><<<grid, kBlock, 0, s>>>(a);
  }
  static void sum_exact(const ReadArgs& a, unsigned grid, hipStream_t s) {
    k_sum_exact<K><<<grid, kBlock, 0, s>>>(a);
  }
  static void absmax(const ReadArgs& a, unsigned grid, hipStream_t s) {
    k_absmax<K><<<grid, kBlock, 0, s>>>(a);
  }
  static void rescale(void* f, uint64_t stride, int np, const Ratio& r, unsigned grid,
                      hipStream_t s) {
    if constexpr (K::C::scaled) k_rescale<K><<<grid, kBlock, 0, s>>>(f, stride, np, r);
  }
  static void wall_density(const WallArgs& a, int parity, hipStream_t s) {
    if constexpr (K::OP != M3LB_OP_NULL) {
      const unsigned g = blocks_for(a.nwall);
      if (parity == 0) k_wall_density<K, 0><<<g, kBlock, 0, s>>>(a);
      else k_wall_density<K, 1><<<g, kBlock, 0, s>>>(a);
      k_corner_density<K><<<g, kBlock, 0, s>>>(a);
    }
  }
  static void walls(const WallArgs& a, int parity, hipStream_t s) {
    if constexpr (K::OP != M3LB_OP_NULL) {
      const unsigned g = blocks_for(a.nwall);
      if (parity == 0) k_walls<K, 0><<<g, kBlock, 0, s>>>(a);
      else k_walls<K, 1><<<g, kBlock, 0, s>>>(a);
    }
  }
  static void snapshot(const SnapArgs& a, int parity, hipStream_t s) {
    const unsigned g = blocks_for(a.count);
    if (parity == 0) k_snapshot<K, 0><<<g, kBlock, 0, s>>>(a);
    else k_snapshot<K, 1><<<g, kBlock, 0, s>>>(a);
  }
  static void specular(const SpecArgs& a, int parity, hipStream_t s) {
    if constexpr (K::OP != M3LB_OP_NULL) {
      const unsigned g = blocks_for(a.count);
      if (parity == 0) k_specular<K, 0><<<g, kBlock, 0, s>>>(a);
      else k_specular<K, 1><<<g, kBlock, 0, s>>>(a);
    }
  }
  static void exchange(const ReadArgs& a, const uint8_t* body, unsigned grid, hipStream_t s) {
    k_momentum_exchange<K><<<grid, kBlock, 0, s>>>(a, body);
  }
  static constexpr bool kPair = K::Q == 27 && !K::FIELD;
  template <int P, bool G>
  static void pair_p(const PairArgs& a, unsigned grid, unsigned tail_grid, unsigned nclear,
                     unsigned* host_err, hipStream_t s) {
    k_pair<K, P, G><<<grid, kBlock, 0, s>>>(a);
    k_pair_tail<K, P, G><<<tail_grid, kBlock, 0, s>>>(a, nclear, host_err);
  }
  static void pair(const PairArgs& a, int parity, bool geom, unsigned grid, unsigned tail_grid,
                   unsigned nclear, unsigned* host_err, hipStream_t s) {
    if constexpr (kPair) {
      if (parity == 0) {
        if (geom) pair_p<0, true>(a, grid, tail_grid, nclear, host_err, s);
        else pair_p<0, false>(a, grid, tail_grid, nclear, host_err, s);
      } else {
        if (geom) pair_p<1, true>(a, grid, tail_grid, nclear, host_err, s);
        else pair_p<1, false>(a, grid, tail_grid, nclear, host_err, s);
      }
    }
  }
  static void batched(const BatchArgs& a, unsigned grid, hipStream_t s) {
    if constexpr (kBatched) {
      if (a.run == kRunBlockRanges) {
        if (a.cached) k_step_batched<K, true, false><<<grid, kBlock, 0, s>>>(a);
        else k_step_batched<K, true, true><<<grid, kBlock, 0, s>>>(a);
      } else {
        if (a.cached) k_step_batched<K, false, false><<<grid, kBlock, 0, s>>>(a);
        else k_step_batched<K, false, true><<<grid, kBlock, 0, s>>>(a);
      }
    }
  }
  static const KernelSet* table() {
    static const KernelSet t = {step,     seed,      gather, storage,
                                macro,    reduce_mass, sum_exact, absmax,
                                rescale,  exchange, wall_density, walls,
                                snapshot, specular,
                                kBatched ? batched : nullptr,
                                kPair ? pair : nullptr,
                                int(sizeof(typename K::T)), int(sizeof(typename K::R))};
    return &t;
  }
};

// Kernel table of a runtime configuration, or null (defined with the C-ABI of
// the single-box solver).
const KernelSet* kernel_set(int q, int op, int storage, int dtype, int force);
// Records msg as the calling thread's m3lb_last_error() and returns code.
int set_error(int code, const char* msg);

}  // namespace m3lb

// One instantiation unit per (lattice, operator); each exports this lookup.
#define M3LB_LOOKUP_NAME(Q, OP) m3lb_kernels_q##Q##_op##OP
#define M3LB_DECLARE_LOOKUP(Q, OP) \
  const m3lb::KernelSet* M3LB_LOOKUP_NAME(Q, OP)(int storage, int dtype, int force)
