// Must NOT compile: the device path implements esoteric pull only.
#include <m3lb/lbm.hpp>
int main() {
  using namespace m3lb;
  Domain d(8, 8, 1);
  FluidSolver<D2Q9, TwoLattice<D2Q9>, BGK<D2Q9>> s(d, BGK<D2Q9>{});
  return 0;
}
