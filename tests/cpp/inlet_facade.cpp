// The inlet-driven Poiseuille channel of the reference's
// validation/poiseuille_inlet.cpp through the C++ facade, with the walls
// given the way that file gives them: a lambda returning a WallSpec per node
// -- walls at rest, a parabolic velocity inlet, and a constant back-pressure
// outlet whose WallSpec carries the imposed density.  Same stopping rule.
// Usage: inlet_facade Ly [order].  Prints "<Ly> <steps> <rel L2 error>".
#include <m3lb/lbm.hpp>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

using namespace m3lb;
using L = D2Q9;
using CMS = CentralMoments<L, NoForcing, ShiftedPopulations>;
using Solver = FluidSolver<L, EsotericPull<L>, CMS>;

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  const Index Ly = std::atoi(argv[1]);
  const int order = argc > 2 ? std::atoi(argv[2]) : 1;
  const Index Lx = 21;
  const double U0 = 0.005, Re = 10.0, nu = U0 * double(Ly) / Re;
  Domain d(Lx, Ly, 1, false, false, true);
  CMS coll;
  coll.omega = CMS::omega_from_viscosity(nu);
  Solver s(d, coll, Precision::FP64);
  s.set_geometry([&](Index, Index, Index) { return Fluid; });
  using WS = Solver::WallSpec;
  s.set_regularized_walls([&](Index x, Index y, Index) -> WS {
    if (y == 0) return WS{NrmYm, 0, 0, 0};
    if (y == Ly - 1) return WS{NrmYp, 0, 0, 0};
    if (x == 0) {
      const double yy = double(y) / double(Ly - 1);
      return WS{NrmXm, 4.0 * U0 * yy * (1.0 - yy), 0, 0};
    }
    if (x == Lx - 1) return WS{NrmOutXp, 0, 0, 0, 1.0};
    return WS{};
  });
  s.set_outflow_order(order);
  s.initialize(1.0);
  const double tau_d = Re * double(Ly) / U0;
  const std::size_t probe = std::max<std::size_t>(500, std::size_t(tau_d / 200.0));
  const std::size_t tmin = std::size_t(0.5 * tau_d), cap = std::size_t(30.0 * tau_d);
  std::vector<double> prev(std::size_t(Lx) * std::size_t(Ly), 0.0);
  std::size_t taken = 0;
  for (std::size_t t = 0; t < cap; t += probe) {
    s.steps(int(probe));
    taken += probe;
    s.compute_macroscopic();
    double num = 0, den = 0;
    for (Index y = 0; y < Ly; ++y)
      for (Index x = 0; x < Lx; ++x) {
        const std::size_t k = std::size_t(y) * std::size_t(Lx) + std::size_t(x);
        const double v = s.ux()[std::size_t(d.id(x, y))];
        num += (v - prev[k]) * (v - prev[k]);
        den += v * v;
        prev[k] = v;
      }
    if (taken >= tmin && std::sqrt(num / std::max(den, 1e-300)) < 1e-10) break;
  }
  s.compute_macroscopic();
  double num = 0, den = 0;
  for (Index y = 0; y < Ly; ++y) {
    const double yy = double(y) / double(Ly - 1), ana = 4.0 * U0 * yy * (1.0 - yy);
    const double dv = s.ux()[std::size_t(d.id(Lx / 2, y))] - ana;
    num += dv * dv;
    den += ana * ana;
  }
  std::printf("%d %zu %.10e\n", int(Ly), taken, std::sqrt(num / den));
  return 0;
}
