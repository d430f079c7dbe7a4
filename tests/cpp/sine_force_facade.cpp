// The fully three-dimensional forced flow of the reference's
// validation/forcing_3d.cpp (PART 2) through the C++ facade, with the forcing
// policy written the way that file writes SinForce: a struct with at(n, F).
// Periodic cube, Fx = F0 sin(k y) sin(k z), exact steady state
// ux = umax sin(k y) sin(k z).  Prints "<N> <tau> <L2> <steps>" per run.
#include <m3lb/lbm.hpp>

#include <cmath>
#include <cstdio>

using namespace m3lb;
using L = D3Q27;

struct SinForce {
  static constexpr const char* name = "SinForce";
  static constexpr bool active = true, per_node = true;
  double F0 = 0, k = 0;
  Domain dom{1, 1, 1, true, true, true};
  void at(Index n, double F[3]) const {
    Index px, py, pz;
    dom.coords(n, px, py, pz);
    const double y = double(py - dom.hy), z = double(pz - dom.hz);
    F[0] = F0 * std::sin(k * y) * std::sin(k * z);
    F[1] = 0;
    F[2] = 0;
  }
};
using CMS = CentralMoments<L, SinForce, ShiftedPopulations>;

static void run(Index N, double tau, double umax) {
  const double nu = (tau - 0.5) / 3.0, k = 2.0 * M_PI / double(N);
  Domain d(N, N, N, true, true, true);
  CMS coll;
  coll.omega = 1.0 / tau;
  coll.forcing.F0 = umax * 2.0 * nu * k * k;
  coll.forcing.k = k;
  coll.forcing.dom = d;
  FluidSolver<L, EsotericPull<L>, CMS> s(d, coll, Precision::FP64);
  s.initialize(1.0);
  const Index probe = d.id(0, N / 4, N / 4);
  double prev = 0;
  std::size_t taken = 0;
  for (std::size_t t = 0; t < 4000000; t += 100) {
    s.steps(100);
    taken += 100;
    s.compute_macroscopic();
    const double cur = s.ux()[std::size_t(probe)];
    if (t > 0 && std::abs(cur - prev) < 1e-14 * (std::abs(cur) + 1e-30)) break;
    prev = cur;
  }
  s.compute_macroscopic();
  double num = 0, den = 0;
  for (Index z = 0; z < N; ++z)
    for (Index y = 0; y < N; ++y)
      for (Index x = 0; x < N; ++x) {
        const std::size_t id = std::size_t(d.id(x, y, z));
        const double ana = umax * std::sin(k * double(y)) * std::sin(k * double(z));
        const double ex = s.ux()[id] - ana;
        num += ex * ex + s.uy()[id] * s.uy()[id] + s.uz()[id] * s.uz()[id];
        den += ana * ana;
      }
  std::printf("%d %.4f %.10e %zu\n", int(N), tau, std::sqrt(num / den), taken);
}

int main() {
  for (double tau : {0.8, 1.2})
    for (Index N : {Index(16), Index(32)}) run(N, tau, 0.02);
  return 0;
}
