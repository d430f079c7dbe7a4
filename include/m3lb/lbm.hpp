// lbm.hpp -- header-only C++ facade over the C-ABI in m3lb.h, with the
// reference's names and policy structure, so that code written against
//     FluidSolver<L, EsotericPull<L>, Collision>      (src/solver/FluidSolver.hpp:34-35)
// compiles against this header and runs the gfx950 kernels.  Each policy type
// maps to a C-ABI enum at compile time; a combination the device path does not
// implement fails with a static_assert, as the reference's own policies do
// (src/collision/MomentCollision.hpp:75-87).
//
// Differences from the reference, all at the host boundary:
//   * rho()/ux()/uy()/uz()/flags() return host std::vector<double>/<uint8_t>
//     copies instead of device View handles;
//   * initialize_field / seed_populations take host callables, evaluated on
//     the host over the padded node index, instead of device lambdas;
//   * the storage precision is a constructor argument (FP64, FP32, FP16 or
//     BF16 storage) instead of the configure-time typedef of
//     src/core/Types.hpp:16-20; its default follows LBM_SINGLE_PRECISION.
#pragma once

#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../m3lb.h"

namespace m3lb {

using Index = std::int32_t;  // src/core/Types.hpp:28

// ---- lattices (src/lattice/Lattices.hpp:72-96, 125-159) ----------------------------
struct D2Q9 {
  static constexpr const char* name = "D2Q9";
  static constexpr int D = 2, Q = 9, id = M3LB_D2Q9;
};
struct D3Q27 {
  static constexpr const char* name = "D3Q27";
  static constexpr int D = 3, Q = 27, id = M3LB_D3Q27;
};

constexpr int opp(int i) noexcept { return i == 0 ? 0 : (i & 1 ? i + 1 : i - 1); }

// ---- storage and forcing tags (Storage.hpp:27-35, Forcing.hpp:41-82) --------
struct RawPopulations {
  static constexpr const char* name = "raw";
  static constexpr bool shifted = false;
};
struct ShiftedPopulations {
  static constexpr const char* name = "shifted";
  static constexpr bool shifted = true;
};
struct NoForcing {
  static constexpr const char* name = "None";
  static constexpr bool active = false;
};
struct Guo {
  static constexpr const char* name = "Guo";
  static constexpr bool active = true;
  double fx = 0, fy = 0, fz = 0;
};
// A forcing policy with a per-node force, as the reference's SinForce
// (validation/forcing_3d.cpp:64-91): any type with
//     static constexpr bool active = true, per_node = true;
//     void at(Index n, double F[3]) const;      // n: padded node index
// evaluated on the host for every node and uploaded (M3LB_FORCE_FIELD).  After
// changing the policy's parameters, call FluidSolver::update_force_field().
template <class T, class = void>
struct is_per_node_force : std::false_type {};
template <class T>
struct is_per_node_force<T, std::void_t<decltype(T::per_node)>>
    : std::integral_constant<bool, T::per_node> {};

// ---- equilibria (Equilibrium.hpp:126-136) -----------------------------------
template <class L>
struct SecondOrderEquilibrium {
  static constexpr bool product = false;
};
template <class L>
struct ProductFormEquilibrium {
  static constexpr bool product = true;
};
template <class L>
using HighOrderEquilibrium = ProductFormEquilibrium<L>;

// ---- streaming (EsotericPull.hpp:44-129) ------------------------------------
template <class L>
struct EsotericPull {
  using Lattice = L;
  static constexpr const char* name = "EsotericPull";
  static constexpr bool implicit_bounce_back = true;
  static constexpr std::size_t bytes_per_node_fp64() { return L::Q * 8; }
};
// The reference's two-array bisection scheme has no device counterpart here.
template <class L>
struct TwoLattice {
  using Lattice = L;
};

inline double omega_from_viscosity(double nu) { return 1.0 / (nu * 3.0 + 0.5); }
inline double viscosity_from_omega(double w) { return (1.0 / w - 0.5) / 3.0; }

// ---- collision policies -----------------------------------------------------
template <class L, class Eq = SecondOrderEquilibrium<L>, class Forcing = NoForcing,
          class Store = RawPopulations>
struct BGK {  // src/collision/BGK.hpp:31-100
  using Lattice = L;
  using Storage = Store;
  using ForcingPolicy = Forcing;
  static constexpr const char* name = "BGK";
  static constexpr int op = Eq::product ? M3LB_OP_BGK_PRODUCT : M3LB_OP_BGK;
  double omega = 1.0;
  Forcing forcing{};
  static double omega_from_viscosity(double nu) { return m3lb::omega_from_viscosity(nu); }
  static double viscosity_from_omega(double w) { return m3lb::viscosity_from_omega(w); }
  double rate(int k) const { return k == 0 ? omega : (k == 1 ? omega : -1.0); }
};

template <class L, class Eq = SecondOrderEquilibrium<L>, class Forcing = NoForcing,
          class Store = RawPopulations>
struct TRT {  // src/collision/TRT.hpp:33-134
  using Lattice = L;
  using Storage = Store;
  using ForcingPolicy = Forcing;
  static constexpr const char* name = "TRT";
  // TRT<L, Eq, F, S> takes either equilibrium (TRT.hpp:32-33, 86-90)
  static constexpr int op = Eq::product ? M3LB_OP_TRT_PRODUCT : M3LB_OP_TRT;
  static constexpr double magic_3_16 = 3.0 / 16.0;
  double omega_p = 1.0, omega_m = 1.0;
  Forcing forcing{};
  static double omega_from_viscosity(double nu) { return m3lb::omega_from_viscosity(nu); }
  static double omega_minus_for(double omega_plus, double lambda) {
    return 1.0 / (lambda / (1.0 / omega_plus - 0.5) + 0.5);
  }
  static double magic_parameter(double wp, double wm) { return (1.0 / wp - 0.5) * (1.0 / wm - 0.5); }
  double rate(int k) const { return k == 0 ? omega_p : (k == 1 ? omega_m : -1.0); }
};

template <class L, class Forcing = NoForcing, class Store = RawPopulations, bool Central = true>
struct MomentCollision {  // src/collision/MomentCollision.hpp:65-376
  using Lattice = L;
  using Storage = Store;
  using ForcingPolicy = Forcing;
  static constexpr const char* name = Central ? "CentralMoments" : "MRT";
  static constexpr int op = Central ? M3LB_OP_CM : M3LB_OP_MRT;
  double omega = 1.0;
  double omega_bulk = -1.0;  // < 0 follows omega
  Forcing forcing{};
  static double omega_from_viscosity(double nu) { return m3lb::omega_from_viscosity(nu); }
  static double viscosity_from_omega(double w) { return m3lb::viscosity_from_omega(w); }
  double rate(int k) const { return k == 0 ? omega : (k == 1 ? omega : omega_bulk); }
};
template <class L, class F = NoForcing, class S = RawPopulations>
using CentralMoments = MomentCollision<L, F, S, true>;
template <class L, class F = NoForcing, class S = RawPopulations>
using MRT = MomentCollision<L, F, S, false>;

// ---- cells and domain (Flags.hpp:20-23, Domain.hpp:37-124) ------------------
enum CellType : std::uint8_t {
  Fluid = 0, Solid = 1, Excluded = 2, RegWall = 3, SpecWall = 4, SpecNode = 5
};

// wall normal codes, pointing out of the fluid, and the outflow codes
// (boundary/Regularized.hpp:97-142)
enum WallNormal : std::uint8_t {
  NrmNone = 0, NrmXp = 1, NrmXm = 2, NrmYp = 3, NrmYm = 4, NrmZp = 5, NrmZm = 6, NrmCorner = 7,
  NrmOutXp = 8, NrmOutFree = 9, NrmOutEq = 10
};

// on-node specular face mask bits (boundary/Specular.hpp:241-249)
enum SpecFace : std::uint8_t {
  SpecNone = 0, SpecXp = 1, SpecXm = 2, SpecYp = 4, SpecYm = 8, SpecZp = 16, SpecZm = 32
};

struct FlowState {  // src/solver/FluidSolver.hpp:30-32
  double rho = 1.0, ux = 0.0, uy = 0.0, uz = 0.0;
};

class Domain {
 public:
  Index nx = 1, ny = 1, nz = 1;
  Index hx = 0, hy = 0, hz = 0;
  Index sx = 1, sy = 1, sz = 1;
  Index n_padded = 1;
  bool periodic[3] = {true, true, true};

  Domain() = default;
  Domain(Index nx_, Index ny_, Index nz_ = 1, bool px = true, bool py = true, bool pz = true)
      : nx(nx_), ny(ny_), nz(nz_) {
    periodic[0] = px; periodic[1] = py; periodic[2] = pz;
    hx = halo(nx, px); hy = halo(ny, py); hz = halo(nz, pz);
    sx = nx + 2 * hx; sy = ny + 2 * hy; sz = nz + 2 * hz;
    n_padded = sx * sy * sz;
  }
  Index n_interior() const { return nx * ny * nz; }
  Index idp(Index px, Index py, Index pz) const { return (pz * sy + py) * sx + px; }
  Index id(Index x, Index y, Index z = 0) const { return idp(x + hx, y + hy, z + hz); }
  void coords(Index n, Index& px, Index& py, Index& pz) const {
    px = n % sx; py = (n / sx) % sy; pz = n / (sx * sy);
  }
  bool is_interior(Index px, Index py, Index pz) const {
    return px >= hx && px < hx + nx && py >= hy && py < hy + ny && pz >= hz && pz < hz + nz;
  }

 private:
  static Index halo(Index n, bool per) { return (n <= 1 || per) ? 0 : 1; }
};

// ---- error handling ---------------------------------------------------------
inline void check(int rc, const char* what) {
  if (rc != M3LB_OK)
    throw std::runtime_error(std::string(what) + ": " + m3lb_last_error());
}

enum class Precision : int {
  FP64 = M3LB_F64,
  FP32 = M3LB_F32,
  FP16 = M3LB_F16,
  BF16 = M3LB_BF16,
  I16 = M3LB_I16,
  E2M13 = M3LB_E2M13
};

#if defined(LBM_SINGLE_PRECISION)
constexpr Precision kDefaultPrecision = Precision::FP32;
#else
constexpr Precision kDefaultPrecision = Precision::FP64;
#endif

// ---- the solver -------------------------------------------------------------
template <class L, class Streaming, class Collision>
class FluidSolver {
  static_assert(std::is_same<Streaming, EsotericPull<L>>::value,
                "the device path implements EsotericPull streaming only");
  static_assert(std::is_same<typename Collision::Lattice, L>::value,
                "collision policy built for another lattice");

 public:
  using Lattice = L;
  static constexpr int Q = L::Q;

  FluidSolver(const Domain& d, const Collision& coll, Precision prec = kDefaultPrecision)
      : dom_(d), coll_(coll) {
    m3lb_config c{};
    c.lattice = L::id;
    c.op = Collision::op;
    c.storage = Collision::Storage::shifted ? M3LB_SHIFTED : M3LB_RAW;
    c.dtype = static_cast<int>(prec);
    c.nx = d.nx; c.ny = d.ny; c.nz = d.nz;
    for (int a = 0; a < 3; ++a) c.periodic[a] = d.periodic[a] ? 1 : 0;
    push_rates(c);
    c.forcing = kPerNode ? M3LB_FORCE_FIELD
                         : (Collision::ForcingPolicy::active ? M3LB_FORCE_GUO : M3LB_FORCE_NONE);
    push_force(c);
    cfg_ = c;
    check(m3lb_create(&cfg_, &h_), "m3lb_create");
    if constexpr (kPerNode) update_force_field();
  }

  // Re-evaluates a per-node forcing policy's at(n, F) on every padded node.
  void update_force_field() {
    static_assert(kPerNode, "the forcing policy has no per-node force");
    const std::size_t np = std::size_t(dom_.n_padded);
    std::vector<double> fx(np, 0.0), fy(np, 0.0), fz(np, 0.0);
    for (Index n = 0; n < dom_.n_padded; ++n) {
      double F[3] = {0, 0, 0};
      coll_.forcing.at(n, F);
      fx[std::size_t(n)] = F[0]; fy[std::size_t(n)] = F[1]; fz[std::size_t(n)] = F[2];
    }
    check(m3lb_set_force_field(h_, fx.data(), fy.data(), fz.data()), "set_force_field");
  }
  ~FluidSolver() { m3lb_destroy(h_); }
  FluidSolver(const FluidSolver&) = delete;
  FluidSolver& operator=(const FluidSolver&) = delete;

  template <class Fn>
  std::vector<std::uint8_t> interior_bytes(Fn fn) const {
    std::vector<std::uint8_t> b(std::size_t(dom_.n_interior()));
    std::size_t k = 0;
    for (Index z = 0; z < dom_.nz; ++z)
      for (Index y = 0; y < dom_.ny; ++y)
        for (Index x = 0; x < dom_.nx; ++x) b[k++] = static_cast<std::uint8_t>(fn(x, y, z));
    return b;
  }

  // set_geometry(fn(x, y, z) -> CellType), FluidSolver.hpp:74-82
  template <class Fn>
  void set_geometry(Fn fn) {
    std::vector<std::uint8_t> fl(std::size_t(dom_.n_interior()));
    std::size_t k = 0;
    for (Index z = 0; z < dom_.nz; ++z)
      for (Index y = 0; y < dom_.ny; ++y)
        for (Index x = 0; x < dom_.nx; ++x) fl[k++] = static_cast<std::uint8_t>(fn(x, y, z));
    check(m3lb_set_geometry(h_, fl.data()), "set_geometry");
  }

  // set_regularized_walls(fn(x, y, z) -> WallSpec), FluidSolver.hpp:84-103,
  // 372-530: walls and outlets on the nodes themselves; the states are
  // deduplicated into a table.  Call after set_geometry.
  struct WallSpec {
    std::uint8_t normal = NrmNone;
    double ux = 0, uy = 0, uz = 0;
    double rho = 1.0;  // read by the outflow codes only
  };
  template <class Fn>
  void set_regularized_walls(Fn fn) {
    const std::size_t ni = std::size_t(dom_.n_interior());
    std::vector<std::uint8_t> nrm(ni, 0);
    std::vector<std::uint16_t> tag(ni, 0);
    std::vector<m3lb_wall> table;
    std::size_t k = 0;
    for (Index z = 0; z < dom_.nz; ++z)
      for (Index y = 0; y < dom_.ny; ++y)
        for (Index x = 0; x < dom_.nx; ++x, ++k) {
          const WallSpec w = fn(x, y, z);
          if (w.normal == NrmNone) continue;
          std::size_t t = 0;
          while (t < table.size() && !(table[t].ux == w.ux && table[t].uy == w.uy &&
                                       table[t].uz == w.uz && table[t].rho == w.rho))
            ++t;
          if (t == table.size()) {
            if (t == 65536) throw std::runtime_error("set_regularized_walls: over 65536 wall states");
            table.push_back(m3lb_wall{w.ux, w.uy, w.uz, w.rho});
          }
          nrm[k] = w.normal;
          tag[k] = std::uint16_t(t);
        }
    if (table.empty()) table.push_back(m3lb_wall{0, 0, 0, 1});
    check(m3lb_set_regularized_walls(h_, nrm.data(), tag.data(), table.data(),
                                     std::int32_t(table.size())),
          "set_regularized_walls");
  }
  // set_fd_corners(on), FluidSolver.hpp:272
  void set_fd_corners(bool on) { check(m3lb_set_fd_corners(h_, on ? 1 : 0), "set_fd_corners"); }
  // FluidSolver.hpp:303-367, 1311-1330
  void set_force_correction(bool on) {
    check(m3lb_set_force_correction(h_, on ? 1 : 0), "set_force_correction");
  }
  void set_outflow_order(int k) { check(m3lb_set_outflow_order(h_, k), "set_outflow_order"); }
  void set_outflow_density(double r) { check(m3lb_set_outflow_density(h_, r), "set_outflow_density"); }
  void set_wall_velocity_scale(double sc) {
    check(m3lb_set_wall_velocity_scale(h_, sc), "set_wall_velocity_scale");
  }
  void set_wall_density(double r) { check(m3lb_set_wall_density(h_, r), "set_wall_density"); }
  // fn(x, y, z) -> std::array<double, 3>, interior coordinates
  template <class Fn>
  void set_wall_velocity_field(Fn fn) {
    const std::size_t ni = std::size_t(dom_.n_interior());
    std::vector<double> u(ni), v(ni), w(ni);
    std::size_t k = 0;
    for (Index z = 0; z < dom_.nz; ++z)
      for (Index y = 0; y < dom_.ny; ++y)
        for (Index x = 0; x < dom_.nx; ++x, ++k) {
          const auto c = fn(x, y, z);
          u[k] = c[0]; v[k] = c[1]; w[k] = c[2];
        }
    check(m3lb_set_wall_velocity_field(h_, u.data(), v.data(), w.data()), "set_wall_velocity_field");
  }
  void set_wall_velocity_field_scale(double sc) {
    check(m3lb_set_wall_velocity_field_scale(h_, sc), "set_wall_velocity_field_scale");
  }
  // set_specular_walls(fn(x, y, z) -> NrmXp..NrmZm or NrmNone), FluidSolver.hpp:193-216
  template <class Fn>
  void set_specular_walls(Fn fn) {
    check(m3lb_set_specular_walls(h_, interior_bytes(fn).data()), "set_specular_walls");
  }
  // set_specular_nodes(fn(x, y, z) -> SpecFace mask), FluidSolver.hpp:236-270
  template <class Fn>
  void set_specular_nodes(Fn fn) {
    check(m3lb_set_specular_nodes(h_, interior_bytes(fn).data()), "set_specular_nodes");
  }

  void initialize(double rho0, double ux0 = 0, double uy0 = 0, double uz0 = 0) {
    sync_policy();
    check(m3lb_initialize(h_, rho0, ux0, uy0, uz0), "initialize");
  }

  // initialize_field(fn(Index n) -> FlowState), n the padded node index
  template <class Fn>
  void initialize_field(Fn fn) {
    sync_policy();
    const std::size_t np = std::size_t(dom_.n_padded);
    std::vector<double> r(np, 1.0), u(np, 0.0), v(np, 0.0), w(np, 0.0);
    for (Index n = 0; n < dom_.n_padded; ++n) {
      Index px, py, pz;
      dom_.coords(n, px, py, pz);
      if (!dom_.is_interior(px, py, pz)) continue;
      const FlowState s = fn(n);
      r[n] = s.rho; u[n] = s.ux; v[n] = s.uy; w[n] = s.uz;
    }
    check(m3lb_initialize_field(h_, r.data(), u.data(), v.data(), w.data()), "initialize_field");
  }

  void step(bool store_macroscopic = false) {
    sync_policy();
    check(m3lb_step(h_, 1, store_macroscopic ? 1 : 0), "step");
  }
  void steps(int n, bool store_macroscopic_last = false) {
    sync_policy();
    check(m3lb_step(h_, n, store_macroscopic_last ? 1 : 0), "step");
  }

  void compute_macroscopic() {
    const std::size_t np = std::size_t(dom_.n_padded);
    rho_.assign(np, 0.0); ux_.assign(np, 0.0); uy_.assign(np, 0.0); uz_.assign(np, 0.0);
    check(m3lb_compute_macroscopic(h_, rho_.data(), ux_.data(), uy_.data(), uz_.data()),
          "compute_macroscopic");
  }
  double total_mass() {
    double m = 0;
    check(m3lb_total_mass(h_, &m), "total_mass");
    return m;
  }
  double total_momentum(int a) {
    double m = 0;
    check(m3lb_total_momentum(h_, a, &m), "total_momentum");
    return m;
  }

  const Domain& domain() const { return dom_; }
  Collision& collision() { return coll_; }
  const std::vector<double>& rho() const { return rho_; }
  const std::vector<double>& ux() const { return ux_; }
  const std::vector<double>& uy() const { return uy_; }
  const std::vector<double>& uz() const { return uz_; }
  std::size_t timestep() const {
    std::uint64_t t = 0;
    check(m3lb_timestep(h_, &t), "timestep");
    return std::size_t(t);
  }

  // gather_populations(): (n, q), n_padded * Q values
  std::vector<double> gather_populations() const {
    std::vector<double> out(std::size_t(dom_.n_padded) * Q);
    check(m3lb_gather_populations(h_, out.data()), "gather_populations");
    return out;
  }
  // seed_populations(fn(n, i) -> value)
  template <class Fn>
  void seed_populations(Fn fn) {
    std::vector<double> in(std::size_t(dom_.n_padded) * Q);
    for (Index n = 0; n < dom_.n_padded; ++n)
      for (int i = 0; i < Q; ++i) in[std::size_t(n) * Q + i] = fn(n, i);
    check(m3lb_seed_populations(h_, in.data()), "seed_populations");
  }

  m3lb_solver* handle() { return h_; }

 private:
  void push_rates(m3lb_config& c) const {
    c.omega = coll_.rate(0);
    c.omega_minus = coll_.rate(1);
    c.omega_bulk = coll_.rate(2);
  }
  static constexpr bool kPerNode = is_per_node_force<typename Collision::ForcingPolicy>::value;
  void push_force(m3lb_config& c) const {
    if constexpr (Collision::ForcingPolicy::active && !kPerNode) {
      c.force[0] = coll_.forcing.fx; c.force[1] = coll_.forcing.fy; c.force[2] = coll_.forcing.fz;
    } else {
      c.force[0] = c.force[1] = c.force[2] = 0.0;
    }
  }
  // collision() hands out the policy by reference, as the reference does, so a
  // caller may change omega or the force between steps; push what changed.
  void sync_policy() {
    m3lb_config c = cfg_;
    push_rates(c);
    push_force(c);
    if (c.omega != cfg_.omega || c.omega_minus != cfg_.omega_minus || c.omega_bulk != cfg_.omega_bulk)
      check(m3lb_set_omega(h_, c.omega, c.omega_minus, c.omega_bulk), "set_omega");
    if constexpr (Collision::ForcingPolicy::active && !kPerNode) {
      if (c.force[0] != cfg_.force[0] || c.force[1] != cfg_.force[1] || c.force[2] != cfg_.force[2])
        check(m3lb_set_force(h_, c.force[0], c.force[1], c.force[2]), "set_force");
    }
    cfg_ = c;
  }

  Domain dom_;
  Collision coll_;
  m3lb_config cfg_{};
  m3lb_solver* h_ = nullptr;
  std::vector<double> rho_, ux_, uy_, uz_;
};

}  // namespace m3lb
