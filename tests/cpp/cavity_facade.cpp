// The lid-driven cavity of the reference's validation/cavity_cm.cpp through
// the C++ facade, with the walls given the way that file gives them: a
// lambda returning a WallSpec per node (lid moving, corners at rest).
// Usage: cavity_facade N Re steps.  Prints u/U along x = N/2 (y = 0..N-1)
// after `steps` steps, one value per line.
#include <m3lb/lbm.hpp>

#include <cstdio>
#include <cstdlib>

using namespace m3lb;
using L = D2Q9;
using CMS = CentralMoments<L, NoForcing, ShiftedPopulations>;
using Solver = FluidSolver<L, EsotericPull<L>, CMS>;

int main(int argc, char** argv) {
  if (argc < 4) return 2;
  const Index N = std::atoi(argv[1]);
  const double Re = std::atof(argv[2]);
  const int steps = std::atoi(argv[3]);
  const double U = 0.02, nu = U * double(N - 1) / Re;
  Domain d(N, N, 1, false, false, true);
  CMS coll;
  coll.omega = 1.0 / (3.0 * nu + 0.5);
  Solver s(d, coll, Precision::FP64);
  s.set_geometry([&](Index, Index, Index) { return Fluid; });
  using WS = Solver::WallSpec;
  s.set_regularized_walls([&](Index x, Index y, Index) -> WS {
    const bool xl = x == 0, xr = x == N - 1, yb = y == 0, yt = y == N - 1;
    if ((xl || xr) && (yb || yt)) return WS{NrmCorner, 0, 0, 0};
    if (yt) return WS{NrmYp, U, 0, 0};
    if (yb) return WS{NrmYm, 0, 0, 0};
    if (xl) return WS{NrmXm, 0, 0, 0};
    if (xr) return WS{NrmXp, 0, 0, 0};
    return WS{};
  });
  s.initialize(1.0);
  s.steps(steps);
  s.compute_macroscopic();
  for (Index y = 0; y < N; ++y) std::printf("%.17e\n", s.ux()[std::size_t(d.id(N / 2, y))] / U);
  return 0;
}
