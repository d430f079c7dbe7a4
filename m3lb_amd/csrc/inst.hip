// inst.hip -- instantiates every kernel of one (lattice, operator) pair.
// Compiled once per pair with -DM3LB_Q=<9|27> -DM3LB_OP=<0..6>, so the
// configurations build in parallel and each object stays small.
#include "launch.hpp"

#ifndef M3LB_Q
#error "M3LB_Q must be defined (9 or 27)"
#endif
#ifndef M3LB_OP
#error "M3LB_OP must be defined (0..6)"
#endif

namespace {
template <int ST, int DT, int F>
const m3lb::KernelSet* pick() {
  return m3lb::Launch<m3lb::Cfg<M3LB_Q, M3LB_OP, ST, DT, F>>::table();
}
}  // namespace

#define M3LB_JOIN2(a, b) a##b
#define M3LB_NAME(Q, OP) M3LB_LOOKUP_NAME(Q, OP)
#define M3LB_EXPAND_LOOKUP(Q, OP) M3LB_DECLARE_LOOKUP(Q, OP)

M3LB_EXPAND_LOOKUP(M3LB_Q, M3LB_OP);
M3LB_EXPAND_LOOKUP(M3LB_Q, M3LB_OP) {
#if M3LB_OP == 0
  // identity collision: raw storage, no force (tests/test_streaming.cpp:28-38)
  (void)storage;
  (void)force;
  if (dtype == M3LB_F64) return pick<M3LB_RAW, M3LB_F64, M3LB_FORCE_NONE>();
  if (dtype == M3LB_F32) return pick<M3LB_RAW, M3LB_F32, M3LB_FORCE_NONE>();
  return nullptr;
#else
  // every forcing policy for F64/F32; uniform Guo or none for 16-bit storage
#define M3LB_PICK3(ST, DT)                                                  \
  (force == M3LB_FORCE_FIELD ? pick<ST, DT, M3LB_FORCE_FIELD>()             \
   : force == M3LB_FORCE_GUO ? pick<ST, DT, M3LB_FORCE_GUO>()               \
                             : pick<ST, DT, M3LB_FORCE_NONE>())
#define M3LB_PICK2(ST, DT)                                                  \
  (force == M3LB_FORCE_FIELD ? nullptr                                      \
   : force == M3LB_FORCE_GUO ? pick<ST, DT, M3LB_FORCE_GUO>()               \
                             : pick<ST, DT, M3LB_FORCE_NONE>())
  if (storage == M3LB_RAW) {
    if (dtype == M3LB_F64) return M3LB_PICK3(M3LB_RAW, M3LB_F64);
    if (dtype == M3LB_F32) return M3LB_PICK3(M3LB_RAW, M3LB_F32);
    return nullptr;
  }
  if (dtype == M3LB_F64) return M3LB_PICK3(M3LB_SHIFTED, M3LB_F64);
  if (dtype == M3LB_F32) return M3LB_PICK3(M3LB_SHIFTED, M3LB_F32);
  if (dtype == M3LB_F16) return M3LB_PICK2(M3LB_SHIFTED, M3LB_F16);
  if (dtype == M3LB_BF16) return M3LB_PICK2(M3LB_SHIFTED, M3LB_BF16);
  if (dtype == M3LB_I16) return M3LB_PICK2(M3LB_SHIFTED, M3LB_I16);
  if (dtype == M3LB_E2M13) return M3LB_PICK2(M3LB_SHIFTED, M3LB_E2M13);
  return nullptr;
#endif
}
