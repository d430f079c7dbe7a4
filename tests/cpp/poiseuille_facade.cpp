// Force-driven Poiseuille through the C++ facade, written the way the
// reference's validation/poiseuille.cpp drives FluidSolver: geometry lambda,
// initialize, step loop to steady state, compute_macroscopic, measure.
// Prints "<operator> <tau> <L2> <slip> <steps>" per run.
#include <m3lb/lbm.hpp>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <vector>

using namespace m3lb;
using L = D2Q9;

static void parabola_roots(const std::vector<double>& y, const std::vector<double>& u, double& lo,
                           double& hi) {
  double S[3][4] = {};
  for (std::size_t k = 0; k < y.size(); ++k) {
    const double p[3] = {1.0, y[k], y[k] * y[k]};
    for (int i = 0; i < 3; ++i) {
      for (int j = 0; j < 3; ++j) S[i][j] += p[i] * p[j];
      S[i][3] += p[i] * u[k];
    }
  }
  for (int c = 0; c < 3; ++c) {
    int piv = c;
    for (int r = c; r < 3; ++r)
      if (std::abs(S[r][c]) > std::abs(S[piv][c])) piv = r;
    for (int j = 0; j < 4; ++j) std::swap(S[c][j], S[piv][j]);
    const double d = S[c][c];
    for (int j = 0; j < 4; ++j) S[c][j] /= d;
    for (int r = 0; r < 3; ++r)
      if (r != c) {
        const double f = S[r][c];
        for (int j = 0; j < 4; ++j) S[r][j] -= f * S[c][j];
      }
  }
  const double a = S[0][3], b = S[1][3], c2 = S[2][3];
  const double disc = std::sqrt(b * b - 4 * c2 * a);
  lo = (-b + disc) / (2 * c2);
  hi = (-b - disc) / (2 * c2);
  if (lo > hi) std::swap(lo, hi);
}

template <class Coll, class Setup>
static void run(const char* nm, double tau, Index H, Setup setup) {
  const Index nx = 8, ny = H + 2;
  const double nu = (tau - 0.5) / 3.0, omega = 1.0 / tau;
  const double G = 8.0 * nu * 0.01 / (double(H) * double(H));
  Domain d(nx, ny, 1, true, false, true);
  Coll coll;
  setup(coll, omega, G);
  FluidSolver<L, EsotericPull<L>, Coll> s(d, coll, Precision::FP64);
  s.set_geometry([&](Index, Index y, Index) { return (y == 0 || y == ny - 1) ? Solid : Fluid; });
  s.initialize(1.0);
  double prev = 0;
  std::size_t taken = 0;
  for (std::size_t t = 0; t < 600000; t += 200) {
    s.steps(200);
    taken += 200;
    s.compute_macroscopic();
    const double cur = s.ux()[d.id(nx / 2, ny / 2)];
    if (t > 0 && std::abs(cur - prev) < 1e-15 * std::abs(cur)) break;
    prev = cur;
  }
  s.compute_macroscopic();
  std::vector<double> yy, uu;
  double rsum = 0;
  for (Index j = 1; j <= H; ++j) {
    yy.push_back(double(j));
    uu.push_back(s.ux()[d.id(nx / 2, j)]);
    rsum += s.rho()[d.id(nx / 2, j)];
  }
  const double A = G / (2.0 * (rsum / H) * nu);
  double num = 0, den = 0;
  for (Index j = 1; j <= H; ++j) {
    const double ana = A * (j - 0.5) * (H + 0.5 - j);
    num += (uu[j - 1] - ana) * (uu[j - 1] - ana);
    den += ana * ana;
  }
  double lo, hi;
  parabola_roots(yy, uu, lo, hi);
  std::printf("%s %.7f %.6e %+.6e %zu\n", nm, tau, std::sqrt(num / den), 0.5 - lo, taken);
}

int main() {
  auto plain = [](auto& c, double w, double G) { c.omega = w; c.forcing = Guo{G, 0, 0}; };
  auto trt = [](auto& c, double w, double G) {
    using T = std::decay_t<decltype(c)>;
    c.omega_p = w;
    c.omega_m = T::omega_minus_for(w, T::magic_3_16);
    c.forcing = Guo{G, 0, 0};
  };
  using BGKf = BGK<L, SecondOrderEquilibrium<L>, Guo, ShiftedPopulations>;
  using TRTf = TRT<L, SecondOrderEquilibrium<L>, Guo, ShiftedPopulations>;
  // TRT takes any equilibrium (TRT.hpp:32-33); with u_y = 0 the D2Q9 product
  // form reduces to the second-order one, so it must give TRT's slip row.
  using TRTp = TRT<L, HighOrderEquilibrium<L>, Guo, ShiftedPopulations>;
  using MRTf = MRT<L, Guo, ShiftedPopulations>;
  using CMf = CentralMoments<L, Guo, ShiftedPopulations>;
  const double tb = 0.5 + std::sqrt(3.0) / 4.0;
  for (double tau : {0.6, 0.875, tb}) {
    run<BGKf>("BGK", tau, 32, plain);
    run<TRTf>("TRT", tau, 32, trt);
    run<TRTp>("TRT_PRODUCT", tau, 32, trt);
    run<MRTf>("MRT", tau, 32, plain);
    run<CMf>("CM", tau, 32, plain);
  }
  return 0;
}
