// solver.hip -- the C-ABI of include/m3lb.h: a handle owning one in-place
// population array on the current HIP device, its flags and diagnostics
// buffers, and one HIP stream that orders every call on the handle.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "launch.hpp"
#include "scales.hpp"

using m3lb::KernelSet;

#define M3LB_X(Q, OP) M3LB_DECLARE_LOOKUP(Q, OP);
M3LB_X(9, 0) M3LB_X(9, 1) M3LB_X(9, 2) M3LB_X(9, 3) M3LB_X(9, 4) M3LB_X(9, 5) M3LB_X(9, 6)
M3LB_X(27, 0) M3LB_X(27, 1) M3LB_X(27, 2) M3LB_X(27, 3) M3LB_X(27, 4) M3LB_X(27, 5) M3LB_X(27, 6)
#undef M3LB_X

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIP_OK(expr)                                                                  \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess)                                                             \
      return fail(M3LB_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

const KernelSet* lookup(int q, int op, int st, int dt, int force) {
  typedef const KernelSet* (*Fn)(int, int, int);
  static const Fn q9[7] = {m3lb_kernels_q9_op0, m3lb_kernels_q9_op1, m3lb_kernels_q9_op2,
                           m3lb_kernels_q9_op3, m3lb_kernels_q9_op4, m3lb_kernels_q9_op5,
                           m3lb_kernels_q9_op6};
  static const Fn q27[7] = {m3lb_kernels_q27_op0, m3lb_kernels_q27_op1, m3lb_kernels_q27_op2,
                            m3lb_kernels_q27_op3, m3lb_kernels_q27_op4, m3lb_kernels_q27_op5,
                            m3lb_kernels_q27_op6};
  if (op < 0 || op > 6) return nullptr;
  return (q == 9 ? q9 : q27)[op](st, dt, force);
}

void domain_of(const m3lb_config& c, int64_t out[10]) {
  const int64_t n[3] = {c.nx, c.ny, c.nz};
  int64_t h[3];
  for (int a = 0; a < 3; ++a) h[a] = (n[a] <= 1 || c.periodic[a]) ? 0 : 1;  // Domain.hpp:104-108
  for (int a = 0; a < 3; ++a) { out[a] = n[a]; out[3 + a] = h[a]; out[6 + a] = n[a] + 2 * h[a]; }
  out[9] = out[6] * out[7] * out[8];
}

int validate(const m3lb_config* c) {
  if (!c) return fail(M3LB_EINVAL, "config is null");
  if (c->lattice != M3LB_D2Q9 && c->lattice != M3LB_D3Q27)
    return fail(M3LB_EINVAL, "lattice must be M3LB_D2Q9 or M3LB_D3Q27");
  if (c->op < M3LB_OP_NULL || c->op > M3LB_OP_TRT_PRODUCT) return fail(M3LB_EINVAL, "unknown operator");
  if (c->storage != M3LB_RAW && c->storage != M3LB_SHIFTED)
    return fail(M3LB_EINVAL, "storage must be M3LB_RAW or M3LB_SHIFTED");
  if (c->dtype < M3LB_F64 || c->dtype > M3LB_E2M13) return fail(M3LB_EINVAL, "unknown dtype");
  if (c->dtype >= M3LB_F16 &&
      (c->storage != M3LB_SHIFTED || c->op == M3LB_OP_NULL))
    return fail(M3LB_EINVAL, "16-bit storage holds shifted populations (g = f - w) only");
  if (c->forcing != M3LB_FORCE_NONE && c->forcing != M3LB_FORCE_GUO && c->forcing != M3LB_FORCE_FIELD)
    return fail(M3LB_EINVAL, "forcing must be M3LB_FORCE_NONE, M3LB_FORCE_GUO or M3LB_FORCE_FIELD");
  if (c->reserved != 0) return fail(M3LB_EINVAL, "reserved must be 0");
  if (c->nx < 1 || c->ny < 1 || c->nz < 1) return fail(M3LB_EINVAL, "extents must be >= 1");
  if (c->lattice == M3LB_D2Q9 && c->nz != 1) return fail(M3LB_EINVAL, "D2Q9 needs nz == 1");
  if (c->op == M3LB_OP_NULL && c->forcing != M3LB_FORCE_NONE)
    return fail(M3LB_EINVAL, "the identity collision takes no force");
  if (!(std::isfinite(c->omega) && std::isfinite(c->omega_minus) && std::isfinite(c->omega_bulk)))
    return fail(M3LB_EINVAL, "relaxation rates must be finite");
  return M3LB_OK;
}

}  // namespace

// Boundary nodes set by their own calls (FluidSolver.hpp:193-530,
// 1259-1300): regularised walls (per padded node: normal code, table tag,
// unknown mask, corner axis, outflow donor and its face, closure density) and
// specular cells (ghost normal, on-node face mask).  Host copies are kept to
// build the per-step plan.
struct Boundaries {
  // regularised walls, host
  std::vector<uint8_t> nrm, ext, onrm;
  std::vector<uint16_t> tag;
  std::vector<uint32_t> unk;
  std::vector<unsigned> don;
  std::vector<double> wtab, wtab0;  // (ux, uy, uz, rho) per state, as the compute type holds them
  bool walls_set = false, has_corners = false, un_on = false;
  int fd_corners = 1;  // fd_corners_ = has_shear_omega<Collision> (FluidSolver.hpp:1287)
  int force_bc = 1;    // force_bc_ (:1288)
  int out_order = 1;   // out_order_ (:1291)
  double outflow_rho = 1.0, un_scale = 1.0;
  // regularised walls, device
  uint8_t *nrm_d = nullptr, *ext_d = nullptr, *onrm_d = nullptr;
  uint16_t* tag_d = nullptr;
  uint32_t* unk_d = nullptr;
  unsigned* don_d = nullptr;
  void* rho_d = nullptr;
  double* wtab_d = nullptr;
  void* un_d = nullptr;
  // specular cells
  std::vector<uint8_t> spec_nrm, spec_faces;
  uint8_t *spec_nrm_d = nullptr, *spec_faces_d = nullptr;
  // the per-step plan, rebuilt from the flags when they or the reads change
  bool dirty = true;
  int nwall = 0, nspec = 0, nsnap = 0;
  std::vector<int> level_off;  // wall list offsets per level, then the end
  unsigned *wlist_d = nullptr, *slist_d = nullptr, *snap_nodes_d = nullptr;
  int* rd_d = nullptr;
  void* snap_d = nullptr;
};

struct m3lb_solver {
  m3lb_config cfg;
  int Q = 0;
  int64_t dom[10];
  m3lb::Dom d;
  const KernelSet* ks = nullptr;
  int device = 0;
  hipStream_t stream = nullptr;
  void* f = nullptr;
  uint64_t stride = 0;  // elements per direction
  uint8_t* flags_d = nullptr;
  std::vector<uint8_t> flags_h;
  bool geometry = false;  // any interior node not Fluid
  void* macro_d[4] = {nullptr, nullptr, nullptr, nullptr};  // compute type, StoreMacro
  bool macro_valid = false;
  double* partial_d = nullptr;  // per-block partials
  size_t partial_n = 0;
  uint64_t t = 0;
  float scale[27];
  float inv_scale[27];
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int remap = 64;  // step-kernel tile order (kernels.hpp tile_of)
  unsigned lds_cap = 0;  // LDS a capped step launch reserves: 3 blocks per CU
  int cache_policy = M3LB_CACHE_AUTO;  // step-kernel accesses to the local slots
  void* ff_d = nullptr;  // M3LB_FORCE_FIELD: fx, fy, fz planes of `stride` (compute type)
  Boundaries b;  // regularised walls and specular cells
  // two steps per pass (pair.hpp): mode, stripes, lag, blocks per CU, tiles per
  // ticket (0: default)
  int fusion = M3LB_FUSION_AUTO, pair_stripes = 0, pair_lag = 0, pair_bpc = 0, pair_group = 0;
  int cus = 0;                    // compute units of the device
  unsigned* pair_sync = nullptr;  // counters of launch 1
  unsigned pair_sync_n = 0;       // words allocated
  unsigned* pair_err_h = nullptr; // pinned: a failed wait in a pass, set by the tail launch
  uint64_t pair_passes = 0;       // pairs of steps run as one pass
  bool pair_fail_test = false;    // every wait of a pass gives up (failure-path test)
};

namespace {

// Synchronises the handle's stream and reports a two-step pass whose bounded
// waits gave up (its steps' results are then not valid).
int stream_sync(m3lb_solver* s) {
  HIP_OK(hipStreamSynchronize(s->stream));
  if (s->pair_err_h && *s->pair_err_h) {
    *s->pair_err_h = 0;
    return fail(M3LB_EDEVICE, "two-step pass: a wait for neighbouring tiles timed out; "
                              "the populations are not valid");
  }
  return M3LB_OK;
}
#define SYNC_OK(s)                          \
  do {                                      \
    const int rc_ = stream_sync(s);         \
    if (rc_) return rc_;                    \
  } while (0)

unsigned reduce_grid(const m3lb_solver* s) {
  return std::max(1u, std::min(4096u, m3lb::blocks_for(s->d.np)));
}

int ensure_partials(m3lb_solver* s, size_t n_doubles) {
  if (s->partial_n >= n_doubles) return M3LB_OK;
  if (s->partial_d) HIP_OK(hipFree(s->partial_d));
  s->partial_d = nullptr;
  HIP_OK(hipMalloc(&s->partial_d, n_doubles * sizeof(double)));
  s->partial_n = n_doubles;
  return M3LB_OK;
}

m3lb::StepArgs step_args(const m3lb_solver* s) {
  m3lb::StepArgs a;
  std::memset(&a, 0, sizeof a);
  a.f = s->f;
  a.stride = s->stride;
  a.flags = s->flags_d;
  a.d = s->d;
  const int nx = s->d.nx;
  a.bx_log2 = nx > 128 ? 8 : (nx > 64 ? 7 : 6);
  a.chunks = (nx + m3lb::kBlock - 1) / m3lb::kBlock;
  a.rows = s->d.ny * s->d.nz;
  a.div_ny = m3lb::make_fastdiv(uint32_t(s->d.ny));
  a.remap = s->remap;
  a.lds_cap = s->lds_cap;
  const uint64_t bytes = s->stride * uint64_t(s->Q) * uint64_t(s->ks->elem_bytes);
  a.cached = s->cache_policy == M3LB_CACHE_PLAIN ||
             (s->cache_policy == M3LB_CACHE_AUTO && bytes <= m3lb::kCacheResidentBytes);
  a.omega = s->cfg.omega;
  a.omega_m = s->cfg.omega_minus;
  a.omega_b = s->cfg.omega_bulk;
  for (int k = 0; k < 3; ++k) a.F[k] = s->cfg.force[k];
  a.ff = s->ff_d;
  for (int q = 0; q < 27; ++q) { a.scale[q] = s->scale[q]; a.inv_scale[q] = s->inv_scale[q]; }
  return a;
}

m3lb::ReadArgs read_args(const m3lb_solver* s) {
  m3lb::ReadArgs a;
  std::memset(&a, 0, sizeof a);
  a.f = s->f;
  a.stride = s->stride;
  a.flags = s->flags_d;
  a.d = s->d;
  a.parity = int(s->t % 2);
  a.omega = s->cfg.omega;
  a.omega_m = s->cfg.omega_minus;
  a.omega_b = s->cfg.omega_bulk;
  for (int k = 0; k < 3; ++k) a.F[k] = s->cfg.force[k];
  a.ff = s->ff_d;
  for (int q = 0; q < 27; ++q) a.inv_scale[q] = s->inv_scale[q];
  const Boundaries& b = s->b;
  if (b.walls_set) {
    a.bc_nrm = b.nrm_d;
    a.bc_tag = b.tag_d;
    a.bc_rho = b.rho_d;
    a.wtab = b.wtab_d;
    a.bc_don = b.don_d;
    a.bc_onrm = b.onrm_d;
    a.wall_un = b.un_on ? b.un_d : nullptr;
    a.un_scale = b.un_scale;
    a.outflow_rho = b.outflow_rho;
    a.out_order = b.out_order;
  }
  a.spec_faces = b.spec_faces_d;
  return a;
}

m3lb::WallArgs wall_args(const m3lb_solver* s) {
  const Boundaries& b = s->b;
  m3lb::WallArgs a;
  std::memset(&a, 0, sizeof a);
  a.f = s->f;
  a.stride = s->stride;
  a.flags = s->flags_d;
  a.d = s->d;
  a.list = b.wlist_d;
  a.nwall = b.nwall;
  a.fd_corners = b.fd_corners;
  a.bc_nrm = b.nrm_d;
  a.bc_tag = b.tag_d;
  a.bc_unk = b.unk_d;
  a.bc_ext = b.ext_d;
  a.bc_rho = b.rho_d;
  a.wtab = b.wtab_d;
  a.bc_don = b.don_d;
  a.wall_un = b.un_on ? b.un_d : nullptr;
  a.un_scale = b.un_scale;
  a.outflow_rho = b.outflow_rho;
  a.out_order = b.out_order;
  a.force_bc = b.force_bc;
  a.rd = b.rd_d;
  a.snap = b.snap_d;
  a.omega = s->cfg.omega;
  a.omega_m = s->cfg.omega_minus;
  a.omega_b = s->cfg.omega_bulk;
  for (int k = 0; k < 3; ++k) a.F[k] = s->cfg.force[k];
  a.ff = s->ff_d;
  for (int q = 0; q < 27; ++q) { a.scale[q] = s->scale[q]; a.inv_scale[q] = s->inv_scale[q]; }
  return a;
}

m3lb::SpecArgs spec_args(const m3lb_solver* s) {
  const Boundaries& b = s->b;
  m3lb::SpecArgs a;
  std::memset(&a, 0, sizeof a);
  a.f = s->f;
  a.stride = s->stride;
  a.flags = s->flags_d;
  a.d = s->d;
  a.list = b.slist_d;
  a.count = b.nspec;
  a.spec_nrm = b.spec_nrm_d;
  a.spec_faces = b.spec_faces_d;
  a.omega = s->cfg.omega;
  a.omega_m = s->cfg.omega_minus;
  a.omega_b = s->cfg.omega_bulk;
  for (int k = 0; k < 3; ++k) a.F[k] = s->cfg.force[k];
  a.ff = s->ff_d;
  for (int q = 0; q < 27; ++q) { a.scale[q] = s->scale[q]; a.inv_scale[q] = s->inv_scale[q]; }
  return a;
}

template <class P>
void free_dev(P*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

void free_boundaries(m3lb_solver* s) {
  Boundaries& b = s->b;
  free_dev(b.nrm_d); free_dev(b.ext_d); free_dev(b.onrm_d); free_dev(b.tag_d);
  free_dev(b.unk_d); free_dev(b.don_d); free_dev(b.rho_d); free_dev(b.wtab_d);
  free_dev(b.un_d); free_dev(b.spec_nrm_d); free_dev(b.spec_faces_d); free_dev(b.wlist_d);
  free_dev(b.slist_d); free_dev(b.snap_nodes_d); free_dev(b.rd_d); free_dev(b.snap_d);
}

// A value as the compute type holds it (the reference's tables are Real).
double as_real(const m3lb_solver* s, double v) {
  return s->ks->real_bytes == 8 ? v : double(float(v));
}
double real_mul(const m3lb_solver* s, double a, double c) {
  return s->ks->real_bytes == 8 ? a * c : double(float(a) * float(c));
}

// Re-uploads a small host vector into a device array of its size.
template <class T>
int upload(T*& dst, const std::vector<T>& src, hipStream_t st) {
  free_dev(dst);
  HIP_OK(hipMalloc(&dst, std::max<size_t>(1, src.size()) * sizeof(T)));
  if (!src.empty())
    HIP_OK(hipMemcpyAsync(dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, st));
  return M3LB_OK;
}

// The per-step plan.  The reference's results were made with its serial
// execution space, which sweeps the nodes in ascending order in one pass; a
// wall node that reads a NEIGHBOUR's slots (an outlet's upstream nodes, an
// outflow donor, a corner's finite-difference stencil) then sees an earlier
// neighbour after that neighbour's update (its post-collision set, pair-swapped
// in the slots) and a later one before.  Here every wall node runs after the
// fluid and specular kernels, in levels so that a wall read by a later wall
// runs first, and the sets of later neighbours are copied before the step
// (k_snapshot); rd says which source each read uses.  Nodes that read no
// neighbour are unaffected by order.
int rebuild_plan(m3lb_solver* s) {
  Boundaries& b = s->b;
  const m3lb::Dom& d = s->d;
  const size_t np = size_t(d.np);
  std::vector<unsigned> walls, spec;
  for (size_t n = 0; n < np; ++n) {
    const uint8_t f = s->flags_h[n];
    if (f == M3LB_REGWALL) walls.push_back(unsigned(n));
    else if (f == M3LB_SPECWALL || f == M3LB_SPECNODE) spec.push_back(unsigned(n));
  }
  auto coords = [&](unsigned n, int& x, int& y, int& z) {
    x = int(n % unsigned(d.sx)) - d.hx;
    y = int((n / unsigned(d.sx)) % unsigned(d.sy)) - d.hy;
    z = int(n / unsigned(d.sx * d.sy)) - d.hz;
  };
  auto id = [&](int x, int y, int z) { return unsigned(((z + d.hz) * d.sy + (y + d.hy)) * d.sx + x + d.hx); };
  const int D = s->Q == 9 ? 2 : 3;
  // level of each wall, indexed like `walls` (ascending node index)
  std::vector<int> level(walls.size(), 0);
  auto wall_pos = [&](int64_t m) {
    return size_t(std::lower_bound(walls.begin(), walls.end(), unsigned(m)) - walls.begin());
  };
  std::vector<std::array<int64_t, m3lb::kWallReads>> reads(walls.size());
  int maxlev = 0;
  for (size_t w = 0; w < walls.size(); ++w) {
    const unsigned n = walls[w];
    auto& r = reads[w];
    r.fill(-1);  // -1: no read
    const int code = b.nrm[n];
    int x, y, z;
    coords(n, x, y, z);
    if (code == m3lb::kNrmOutXp || code == m3lb::kNrmOutEq) {
      r[0] = id(x - 1, y, z);
      if (code == m3lb::kNrmOutXp && b.out_order >= 2) r[1] = id(x - 2, y, z);
    } else if (code == m3lb::kNrmOutFree) {
      r[0] = b.don[n];
    } else if (code == m3lb::kNrmCorner && b.fd_corners) {
      const int ic[3] = {x, y, z}, nn[3] = {d.nx, d.ny, d.nz}, per[3] = {d.px, d.py, d.pz};
      for (int ax = 0; ax < D; ++ax) {
        if (nn[ax] < 3) continue;
        int dir = 0;
        if (!per[ax]) {
          if (ic[ax] == 0) dir = 1;
          else if (ic[ax] == nn[ax] - 1) dir = -1;
        }
        int c1[3] = {ic[0], ic[1], ic[2]}, c2[3] = {ic[0], ic[1], ic[2]};
        if (dir != 0) {
          c1[ax] += dir; c2[ax] += 2 * dir;
        } else {
          c1[ax] = (ic[ax] + 1) % nn[ax];
          c2[ax] = (ic[ax] - 1 + nn[ax]) % nn[ax];
        }
        const unsigned m1 = id(c1[0], c1[1], c1[2]), m2 = id(c2[0], c2[1], c2[2]);
        if (s->flags_h[m1] != M3LB_REGWALL) r[2 * ax] = m1;  // walls: table velocity
        if (s->flags_h[m2] != M3LB_REGWALL) r[2 * ax + 1] = m2;
      }
    }
    for (int k = 0; k < m3lb::kWallReads; ++k)
      if (r[k] >= 0 && unsigned(r[k]) < n && s->flags_h[size_t(r[k])] == M3LB_REGWALL)
        level[w] = std::max(level[w], level[wall_pos(r[k])] + 1);
    maxlev = std::max(maxlev, level[w]);
  }
  // wall list by (level, node); rd rows in the same order
  std::vector<size_t> order(walls.size());
  for (size_t w = 0; w < walls.size(); ++w) order[w] = w;
  std::stable_sort(order.begin(), order.end(),
                   [&](size_t a, size_t c) { return level[a] < level[c]; });
  std::vector<unsigned> wlist(walls.size()), snap_nodes;
  std::vector<int> rd(walls.size() * m3lb::kWallReads, -1);
  b.level_off.assign(size_t(maxlev) + 2, int(walls.size()));
  for (size_t k = 0; k < order.size(); ++k) {
    const size_t w = order[k];
    const unsigned n = walls[w];
    wlist[k] = n;
    const int lv = level[w];
    if (b.level_off[size_t(lv)] > int(k)) b.level_off[size_t(lv)] = int(k);
    for (int r = 0; r < m3lb::kWallReads; ++r) {
      const int64_t m = reads[w][size_t(r)];
      if (m > int64_t(n)) {  // a later node: its set from before the step
        rd[k * m3lb::kWallReads + size_t(r)] = int(snap_nodes.size());
        snap_nodes.push_back(unsigned(m));
      }
    }
  }
  for (int lv = maxlev; lv >= 0; --lv)  // levels are contiguous; fill any gap
    b.level_off[size_t(lv)] = std::min(b.level_off[size_t(lv)], b.level_off[size_t(lv) + 1]);
  b.nwall = int(walls.size());
  b.nspec = int(spec.size());
  b.nsnap = int(snap_nodes.size());
  int rc = upload(b.wlist_d, wlist, s->stream);
  if (!rc) rc = upload(b.slist_d, spec, s->stream);
  if (!rc) rc = upload(b.snap_nodes_d, snap_nodes, s->stream);
  if (!rc) rc = upload(b.rd_d, rd, s->stream);
  if (rc) return rc;
  free_dev(b.snap_d);
  HIP_OK(hipMalloc(&b.snap_d, std::max<size_t>(1, snap_nodes.size()) * size_t(s->Q) *
                                  size_t(s->ks->real_bytes)));
  SYNC_OK(s);
  b.dirty = false;
  return M3LB_OK;
}

void refresh_geometry(m3lb_solver* s) {
  const m3lb::Dom& d = s->d;
  bool geom = false;
  for (int z = 0; z < d.nz && !geom; ++z)
    for (int y = 0; y < d.ny && !geom; ++y)
      for (int x = 0; x < d.nx && !geom; ++x)
        geom = s->flags_h[size_t(((z + d.hz) * d.sy + (y + d.hy)) * d.sx + x + d.hx)] != M3LB_FLUID;
  s->geometry = geom;
}

m3lb::SeedArgs seed_args(const m3lb_solver* s) {
  m3lb::SeedArgs a;
  std::memset(&a, 0, sizeof a);
  a.f = s->f;
  a.stride = s->stride;
  a.flags = s->flags_d;
  a.d = s->d;
  a.parity = int(s->t % 2);
  for (int q = 0; q < 27; ++q) a.scale[q] = s->scale[q];
  return a;
}

bool is16(const m3lb_solver* s) {
  return s->cfg.dtype >= M3LB_F16;
}

// Provisional 16-bit scales before the first seeding (scales.hpp).
void provisional_scales(m3lb_solver* s) {
  m3lb::provisional_scales(s->cfg.dtype, s->Q, s->scale, s->inv_scale);
}

int launch_ok() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(M3LB_EDEVICE, std::string("kernel launch: ") + hipGetErrorString(e));
  return M3LB_OK;
}

int sync(m3lb_solver* s) {
  SYNC_OK(s);
  return M3LB_OK;
}

int calibrate(m3lb_solver* s);
int seed_calibrated(m3lb_solver* s, m3lb::SeedArgs a);

}  // namespace

namespace m3lb {
const KernelSet* kernel_set(int q, int op, int storage, int dtype, int force) {
  if (q != M3LB_D2Q9 && q != M3LB_D3Q27) return nullptr;
  return lookup(q, op, storage, dtype, force);
}
int set_error(int code, const char* msg) { return fail(code, msg); }
}  // namespace m3lb

extern "C" {

const char* m3lb_last_error(void) { return g_err.c_str(); }

int m3lb_domain(const m3lb_config* cfg, int64_t out[10]) {
  int rc = validate(cfg);
  if (rc) return rc;
  domain_of(*cfg, out);
  return M3LB_OK;
}

int m3lb_create(const m3lb_config* cfg, m3lb_solver** out) {
  if (!out) return fail(M3LB_EINVAL, "out is null");
  *out = nullptr;
  int rc = validate(cfg);
  if (rc) return rc;
  const KernelSet* ks = lookup(cfg->lattice, cfg->op, cfg->storage, cfg->dtype, cfg->forcing);
  if (!ks) return fail(M3LB_EINVAL, "unsupported lattice/operator/storage/dtype combination");
  int64_t dom[10];
  domain_of(*cfg, dom);
  const int64_t np = dom[9];
  if (np >= (int64_t(1) << 31))
    return fail(M3LB_EINVAL, "n_padded exceeds the reference's int32 node index");
  if (np * int64_t(ks->elem_bytes) >= (int64_t(1) << 32))
    return fail(M3LB_EINVAL, "one direction of the population array exceeds 4 GiB");

  m3lb_solver* s = new m3lb_solver();
  s->cfg = *cfg;
  s->Q = cfg->lattice;
  std::memcpy(s->dom, dom, sizeof dom);
  s->d.nx = int(dom[0]); s->d.ny = int(dom[1]); s->d.nz = int(dom[2]);
  s->d.hx = int(dom[3]); s->d.hy = int(dom[4]); s->d.hz = int(dom[5]);
  s->d.sx = int(dom[6]); s->d.sy = int(dom[7]); s->d.sz = int(dom[8]);
  s->d.px = cfg->periodic[0] ? 1 : 0;
  s->d.py = cfg->periodic[1] ? 1 : 0;
  s->d.pz = cfg->periodic[2] ? 1 : 0;
  s->d.np = int(np);
  s->ks = ks;
  s->stride = (uint64_t(np) + 63) / 64 * 64;
  provisional_scales(s);
  // Tile order measured fastest at 512^3 on MI355X: runs of 64 tiles per XCD
  // for 4- and 8-byte populations, runs of 2 (one row) for 2-byte ones except
  // E2M13, whose step (more instructions per node) ran fastest with runs of
  // 8 at 256^3, 384^3 and 512^3 (2.3-3.4% faster than runs of 2; FP16 was
  // 0.8% slower with 8 at 512^3).
  s->remap = ks->elem_bytes == 2 ? (cfg->dtype == M3LB_E2M13 ? 8 : 2) : 64;

  auto bail = [&](int code) {
    m3lb_destroy(s);
    return code;
  };
  if (hipGetDevice(&s->device) != hipSuccess) return bail(fail(M3LB_EDEVICE, "no HIP device"));
  if (hipStreamCreate(&s->stream) != hipSuccess) return bail(fail(M3LB_EDEVICE, "hipStreamCreate failed"));
  if (hipDeviceGetAttribute(&s->cus, hipDeviceAttributeMultiprocessorCount, s->device) != hipSuccess ||
      s->cus < 1)
    s->cus = 1;
  {
    // a quarter of the CU's LDS plus one byte: at most 3 blocks of the capped
    // step kernel fit a CU (no cap when the attributes are not available)
    int per_cu = 0, per_block = 0;
    if (hipDeviceGetAttribute(&per_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor,
                              s->device) == hipSuccess &&
        hipDeviceGetAttribute(&per_block, hipDeviceAttributeMaxSharedMemoryPerBlock, s->device) ==
            hipSuccess &&
        per_cu > 0 && per_cu / 4 + 1 <= per_block)
      s->lds_cap = unsigned(per_cu / 4 + 1);
  }
  if (hipEventCreate(&s->ev0) != hipSuccess || hipEventCreate(&s->ev1) != hipSuccess)
    return bail(fail(M3LB_EDEVICE, "hipEventCreate failed"));
  const size_t fbytes = s->stride * size_t(s->Q) * size_t(ks->elem_bytes);
  if (hipMalloc(&s->f, fbytes) != hipSuccess)
    return bail(fail(M3LB_ENOMEM, "population array: hipMalloc of " + std::to_string(fbytes) + " bytes failed"));
  // Kokkos Views are zero-initialised; so is this.
  if (hipMemsetAsync(s->f, 0, fbytes, s->stream) != hipSuccess)
    return bail(fail(M3LB_EDEVICE, "hipMemsetAsync failed"));
  if (hipMalloc(&s->flags_d, size_t(np)) != hipSuccess)
    return bail(fail(M3LB_ENOMEM, "flags: hipMalloc failed"));
  if (cfg->forcing == M3LB_FORCE_FIELD) {
    // the per-node force starts at zero everywhere
    const size_t ffb = 3 * s->stride * size_t(ks->real_bytes);
    if (hipMalloc(&s->ff_d, ffb) != hipSuccess)
      return bail(fail(M3LB_ENOMEM, "force field: hipMalloc failed"));
    if (hipMemsetAsync(s->ff_d, 0, ffb, s->stream) != hipSuccess)
      return bail(fail(M3LB_EDEVICE, "hipMemsetAsync failed"));
  }
  // Interior Fluid, halo Excluded (FluidSolver.hpp:59-67).
  s->flags_h.assign(size_t(np), uint8_t(M3LB_EXCLUDED));
  for (int z = 0; z < s->d.nz; ++z)
    for (int y = 0; y < s->d.ny; ++y)
      for (int x = 0; x < s->d.nx; ++x)
        s->flags_h[size_t(((z + s->d.hz) * s->d.sy + (y + s->d.hy)) * s->d.sx + x + s->d.hx)] = M3LB_FLUID;
  if (hipMemcpyAsync(s->flags_d, s->flags_h.data(), size_t(np), hipMemcpyHostToDevice, s->stream) != hipSuccess)
    return bail(fail(M3LB_EDEVICE, "flags upload failed"));
  if (hipStreamSynchronize(s->stream) != hipSuccess) return bail(fail(M3LB_EDEVICE, "sync failed"));
  *out = s;
  return M3LB_OK;
}

int m3lb_destroy(m3lb_solver* s) {
  if (!s) return M3LB_OK;
  if (s->stream) hipStreamSynchronize(s->stream);
  if (s->f) hipFree(s->f);
  if (s->flags_d) hipFree(s->flags_d);
  if (s->ff_d) hipFree(s->ff_d);
  free_boundaries(s);
  for (void* p : s->macro_d)
    if (p) hipFree(p);
  if (s->partial_d) hipFree(s->partial_d);
  if (s->pair_sync) hipFree(s->pair_sync);
  if (s->pair_err_h) hipHostFree(s->pair_err_h);
  if (s->ev0) hipEventDestroy(s->ev0);
  if (s->ev1) hipEventDestroy(s->ev1);
  if (s->stream) hipStreamDestroy(s->stream);
  delete s;
  return M3LB_OK;
}

int m3lb_set_geometry(m3lb_solver* s, const uint8_t* fl) {
  if (!s || !fl) return fail(M3LB_EINVAL, "null argument");
  const m3lb::Dom& d = s->d;
  const size_t ni = size_t(d.nx) * size_t(d.ny) * size_t(d.nz);
  for (size_t k = 0; k < ni; ++k)
    if (fl[k] > M3LB_EXCLUDED) return fail(M3LB_EINVAL, "flag values are Fluid(0), Solid(1), Excluded(2)");
  // every interior flag is rewritten, so wall and specular nodes set earlier
  // are gone (their per-node data stays, unread)
  SYNC_OK(s);
  size_t k = 0;
  for (int z = 0; z < d.nz; ++z)
    for (int y = 0; y < d.ny; ++y)
      for (int x = 0; x < d.nx; ++x, ++k)
        s->flags_h[size_t(((z + d.hz) * d.sy + (y + d.hy)) * d.sx + x + d.hx)] = fl[k];
  refresh_geometry(s);
  s->b.dirty = true;
  HIP_OK(hipMemcpyAsync(s->flags_d, s->flags_h.data(), s->flags_h.size(), hipMemcpyHostToDevice, s->stream));
  return sync(s);
}

int m3lb_initialize(m3lb_solver* s, double rho0, double ux, double uy, double uz) {
  if (!s) return fail(M3LB_EINVAL, "null handle");
  s->t = 0;
  s->macro_valid = false;
  provisional_scales(s);
  m3lb::SeedArgs a = seed_args(s);
  a.mode = m3lb::kSeedUniform;
  a.rho0 = rho0;
  a.u0[0] = ux; a.u0[1] = uy; a.u0[2] = uz;
  return seed_calibrated(s, a);
}

int m3lb_initialize_taylor_green(m3lb_solver* s, double u0) {
  if (!s) return fail(M3LB_EINVAL, "null handle");
  if (!std::isfinite(u0)) return fail(M3LB_EINVAL, "amplitude must be finite");
  s->t = 0;
  s->macro_valid = false;
  provisional_scales(s);
  m3lb::SeedArgs a = seed_args(s);
  a.mode = m3lb::kSeedTaylorGreen;
  a.u0[0] = u0;
  return seed_calibrated(s, a);
}

int m3lb_initialize_field(m3lb_solver* s, const double* rho, const double* ux, const double* uy,
                          const double* uz) {
  if (!s || !rho || !ux || !uy || !uz) return fail(M3LB_EINVAL, "null argument");
  s->t = 0;
  s->macro_valid = false;
  provisional_scales(s);
  const size_t np = size_t(s->d.np);
  double* buf = nullptr;
  HIP_OK(hipMalloc(&buf, 4 * np * sizeof(double)));
  const double* src[4] = {rho, ux, uy, uz};
  for (int k = 0; k < 4; ++k)
    HIP_OK(hipMemcpyAsync(buf + k * np, src[k], np * sizeof(double), hipMemcpyHostToDevice, s->stream));
  m3lb::SeedArgs a = seed_args(s);
  a.mode = m3lb::kSeedField;
  for (int k = 0; k < 4; ++k) a.field[k] = buf + k * np;
  const int rc = seed_calibrated(s, a);  // synchronises before returning
  (void)hipFree(buf);
  return rc;
}

int m3lb_seed_populations(m3lb_solver* s, const double* in) {
  if (!s || !in) return fail(M3LB_EINVAL, "null argument");
  s->macro_valid = false;
  const size_t n = size_t(s->d.np) * size_t(s->Q);
  double* buf = nullptr;
  HIP_OK(hipMalloc(&buf, n * sizeof(double)));
  HIP_OK(hipMemcpyAsync(buf, in, n * sizeof(double), hipMemcpyHostToDevice, s->stream));
  m3lb::SeedArgs a = seed_args(s);
  a.mode = m3lb::kSeedArray;
  a.arr = buf;
  s->ks->seed(a, s->stream);
  int rc = launch_ok();
  hipError_t e = hipStreamSynchronize(s->stream);
  hipFree(buf);
  if (rc) return rc;
  if (e != hipSuccess) return fail(M3LB_EDEVICE, hipGetErrorString(e));
  SYNC_OK(s);  // reports a failed two-step pass
  return M3LB_OK;
}

}  // extern "C"

namespace {

int pair_stripes(const m3lb_solver* s) {
  const int want = s->pair_stripes > 0 ? s->pair_stripes : 8;
  return std::min(want, s->d.ny / 3);
}
int pair_bpc(const m3lb_solver* s) { return s->pair_bpc > 0 ? s->pair_bpc : 4; }
int pair_group(const m3lb_solver* s) { return s->pair_group > 0 ? s->pair_group : 1; }
// Planes step t+1 trails step t by.  A B ticket should find the step-t tiles
// it waits for finished: those were handed out about lag - 1 planes earlier,
// so that distance should exceed the tickets one XCD holds at a time (the
// blocks on it).  Measured D3Q27 CM FP32, 4 blocks per CU, 8 stripes: 512^3
// (128 tickets per plane) 32.0 GLUPS at lag 2, 31.8 at 3, 17.7 at 1; 256^3
// (32 tickets per plane) 29.9 GLUPS at lag 4 against 21.5 at lag 2.
int pair_lag(const m3lb_solver* s) {
  if (s->pair_lag > 0) return s->pair_lag;
  const int S = std::max(1, pair_stripes(s));
  const int rows = s->d.ny / S;
  const unsigned ta = m3lb::pair_per_ticket(
      m3lb::pair_tiles_a(rows, (s->d.nx + m3lb::kBlock - 1) / m3lb::kBlock), pair_group(s));
  const unsigned per_xcd = unsigned(s->cus * pair_bpc(s) + 7) / 8u;
  const int want = int(std::min(64u, std::max(2u, (per_xcd + ta - 1) / std::max(1u, ta) + 1u)));
  return std::max(1, std::min(want, s->d.nz - 3));  // any lag in 1 .. nz-3 is exact
}

// The two-step pass applies when its kernels exist for this configuration,
// no boundary kernel has to run between the steps, the box has room for the
// stripes and the lag, and (in AUTO) the populations exceed the last-level
// cache, where the pass saves HBM traffic, in 4- or 8-byte storage.  Measured
// at 512^3, D3Q27 CM: FP32 31.7 against 24.5 GLUPS, FP64 14.6 against 12.3;
// the 16-bit forms lose at every setting of stripes (8, 16), blocks per CU
// (4, 6, 7) and tiles per ticket (1, 2, 4): at the best, FP16 2.92 against
// 2.80 ms per step and E2M13 2.98 against 2.78 (at the defaults 3.59 and
// 3.94), so AUTO leaves them to one launch per step.
bool pair_eligible(const m3lb_solver* s, const m3lb::StepArgs& a) {
  if (!s->ks->pair || s->fusion == M3LB_FUSION_OFF) return false;
  if (s->b.nwall || s->b.nspec || s->b.nsnap) return false;
  if (s->fusion == M3LB_FUSION_AUTO && (a.cached || s->ks->elem_bytes < 4)) return false;
  const int S = pair_stripes(s);
  return S >= 1 && s->d.nz >= 4 && s->d.nz >= pair_lag(s) + 3;
}

int launch_pair(m3lb_solver* s, const m3lb::StepArgs& a, int parity) {
  const m3lb::Dom& d = s->d;
  m3lb::PairArgs pa;
  std::memset(&pa, 0, sizeof pa);
  pa.stripes = pair_stripes(s);
  const unsigned words = m3lb::pair_sync_words(d.nz, pa.stripes);
  if (s->pair_sync_n < words) {
    // a zeroed area; the tail launch of every pass clears what it used
    if (s->pair_sync) {
      SYNC_OK(s);
      HIP_OK(hipFree(s->pair_sync));
      s->pair_sync = nullptr;
      s->pair_sync_n = 0;
    }
    HIP_OK(hipMalloc(&s->pair_sync, size_t(words) * 4));
    HIP_OK(hipMemsetAsync(s->pair_sync, 0, size_t(words) * 4, s->stream));
    s->pair_sync_n = words;
  }
  if (!s->pair_err_h) {
    HIP_OK(hipHostMalloc(&s->pair_err_h, sizeof(unsigned)));
    *s->pair_err_h = 0;
  }
  pa.a = a;
  for (int j = 0; j < 4; ++j) pa.a.macro[j] = nullptr;
  pa.lag = pair_lag(s);
  pa.group = pair_group(s);
  for (int k = 0; k <= pa.stripes; ++k) pa.y0[k] = int(int64_t(k) * d.ny / pa.stripes);
  pa.sync = s->pair_sync;
  pa.spin_limit = s->pair_fail_test ? 256u : m3lb::kPairSpinLimit;
  pa.wait_extra = s->pair_fail_test ? 1u : 0u;
  const unsigned bpc = unsigned(pair_bpc(s));
  const unsigned tail = std::max(m3lb::pair_tail_rows(d.ny, d.nz, pa.stripes) * unsigned(a.chunks),
                                 (words + unsigned(m3lb::kBlock) - 1) / unsigned(m3lb::kBlock));
  s->ks->pair(pa, parity, s->geometry, unsigned(s->cus) * bpc, tail, words, s->pair_err_h,
              s->stream);
  ++s->pair_passes;
  return launch_ok();
}

}  // namespace

extern "C" {

int m3lb_set_step_fusion(m3lb_solver* s, int32_t mode, int32_t stripes, int32_t lag,
                         int32_t blocks_per_cu, int32_t group) {
  if (!s) return fail(M3LB_EINVAL, "null handle");
  if (mode < M3LB_FUSION_AUTO || mode > M3LB_FUSION_ON)
    return fail(M3LB_EINVAL, "mode must be M3LB_FUSION_AUTO, _OFF or _ON");
  if (stripes < 0 || stripes > m3lb::kPairMaxStripes)
    return fail(M3LB_EINVAL, "stripes must be 0 (default) or 1..64");
  if (lag < 0 || blocks_per_cu < 0 || blocks_per_cu > 8 || group < 0 || group > 64)
    return fail(M3LB_EINVAL, "lag must be >= 0, blocks_per_cu 0..8 and group 0..64 (0: default)");
  SYNC_OK(s);
  s->fusion = mode;
  s->pair_stripes = stripes;
  s->pair_lag = lag;
  s->pair_bpc = blocks_per_cu;
  s->pair_group = group;
  return M3LB_OK;
}

int m3lb_fused_passes(const m3lb_solver* s, uint64_t* passes) {
  if (!s || !passes) return fail(M3LB_EINVAL, "null argument");
  *passes = s->pair_passes;
  return M3LB_OK;
}

int m3lb_test_fail_fused_waits(m3lb_solver* s, int32_t on) {
  if (!s) return fail(M3LB_EINVAL, "null handle");
  SYNC_OK(s);
  s->pair_fail_test = on != 0;
  return M3LB_OK;
}

int m3lb_step(m3lb_solver* s, int32_t nsteps, int32_t store_macroscopic) {
  if (!s) return fail(M3LB_EINVAL, "null handle");
  if (nsteps < 0) return fail(M3LB_EINVAL, "nsteps must be >= 0");
  if (nsteps == 0) return M3LB_OK;
  m3lb::StepArgs a = step_args(s);
  if (store_macroscopic) {
    for (int k = 0; k < 4; ++k) {
      if (!s->macro_d[k]) HIP_OK(hipMalloc(&s->macro_d[k], size_t(s->d.np) * size_t(s->ks->real_bytes)));
    }
  }
  const unsigned rows_per_block = unsigned(m3lb::kBlock >> a.bx_log2);
  const unsigned grid = a.chunks > 1 ? unsigned(a.rows) * unsigned(a.chunks)
                                     : unsigned((a.rows + rows_per_block - 1) / rows_per_block);
  Boundaries& b = s->b;
  if (b.dirty) {
    const int rc = rebuild_plan(s);
    if (rc) return rc;
  }
  m3lb::WallArgs wa = wall_args(s);
  m3lb::SpecArgs sa = spec_args(s);
  m3lb::SnapArgs na;
  std::memset(&na, 0, sizeof na);
  na.f = s->f;
  na.stride = s->stride;
  na.d = s->d;
  na.nodes = b.snap_nodes_d;
  na.count = b.nsnap;
  na.out = b.snap_d;
  for (int q = 0; q < 27; ++q) na.inv_scale[q] = s->inv_scale[q];
  const bool pairs = pair_eligible(s, a);
  for (int k = 0; k < nsteps; ++k) {
    const bool last = k == nsteps - 1;
    // two steps in one pass while two are left (never the step that stores
    // the macroscopic fields)
    if (pairs && k + 1 < nsteps && !(store_macroscopic && k + 2 == nsteps)) {
      const int rc = launch_pair(s, a, int(s->t % 2));
      if (rc) return rc;
      s->t += 2;
      ++k;
      continue;
    }
    for (int j = 0; j < 4; ++j) {
      a.macro[j] = (store_macroscopic && last) ? s->macro_d[j] : nullptr;
      wa.macro[j] = sa.macro[j] = a.macro[j];
    }
    const int parity = int(s->t % 2);
    // FluidSolver::step (:582-591): the corner-density pass, then one sweep
    // in the reference's node order (see rebuild_plan)
    if (b.nwall && b.has_corners) {
      wa.list = b.wlist_d;
      wa.nwall = b.nwall;
      s->ks->wall_density(wa, parity, s->stream);
    }
    if (b.nsnap) s->ks->snapshot(na, parity, s->stream);
    if (b.nspec) s->ks->specular(sa, parity, s->stream);
    s->ks->step(a, parity, s->geometry, grid, s->stream);
    for (size_t lv = 0; lv + 1 < b.level_off.size(); ++lv) {
      const int lo = b.level_off[lv], hi = b.level_off[lv + 1];
      if (hi <= lo) continue;
      wa.list = b.wlist_d + lo;
      wa.rd = b.rd_d + size_t(lo) * m3lb::kWallReads;
      wa.nwall = hi - lo;
      s->ks->walls(wa, parity, s->stream);
    }
    ++s->t;
  }
  s->macro_valid = store_macroscopic != 0;
  return launch_ok();
}

int m3lb_time_steps(m3lb_solver* s, int32_t nsteps, double* ms) {
  if (!s || !ms) return fail(M3LB_EINVAL, "null argument");
  if (s->b.dirty) {  // the boundary plan is host work: keep it out of the timing
    const int rc = rebuild_plan(s);
    if (rc) return rc;
  }
  HIP_OK(hipEventRecord(s->ev0, s->stream));
  int rc = m3lb_step(s, nsteps, 0);
  if (rc) return rc;
  HIP_OK(hipEventRecord(s->ev1, s->stream));
  HIP_OK(hipEventSynchronize(s->ev1));
  SYNC_OK(s);  // reports a failed two-step pass
  float t = 0;
  HIP_OK(hipEventElapsedTime(&t, s->ev0, s->ev1));
  *ms = double(t);
  return M3LB_OK;
}

int m3lb_compute_macroscopic(m3lb_solver* s, double* rho, double* ux, double* uy, double* uz) {
  if (!s) return fail(M3LB_EINVAL, "null handle");
  const size_t np = size_t(s->d.np);
  double* buf = nullptr;
  HIP_OK(hipMalloc(&buf, 4 * np * sizeof(double)));
  m3lb::ReadArgs a = read_args(s);
  for (int k = 0; k < 4; ++k) a.out2[k] = buf + k * np;
  s->ks->macro(a, s->stream);
  int rc = launch_ok();
  double* dst[4] = {rho, ux, uy, uz};
  for (int k = 0; k < 4 && !rc; ++k)
    if (dst[k] && hipMemcpyAsync(dst[k], buf + k * np, np * sizeof(double), hipMemcpyDeviceToHost, s->stream) != hipSuccess)
      rc = fail(M3LB_EDEVICE, "macroscopic download failed");
  hipError_t e = hipStreamSynchronize(s->stream);
  hipFree(buf);
  if (rc) return rc;
  if (e != hipSuccess) return fail(M3LB_EDEVICE, hipGetErrorString(e));
  SYNC_OK(s);  // reports a failed two-step pass
  return M3LB_OK;
}

int m3lb_macroscopic_stored(m3lb_solver* s, double* rho, double* ux, double* uy, double* uz) {
  if (!s) return fail(M3LB_EINVAL, "null handle");
  if (!s->macro_valid) return fail(M3LB_ESTATE, "no step with store_macroscopic = 1 since the last change");
  const size_t np = size_t(s->d.np);
  const size_t rb = size_t(s->ks->real_bytes);
  std::vector<char> tmp(np * rb);
  double* dst[4] = {rho, ux, uy, uz};
  for (int k = 0; k < 4; ++k) {
    if (!dst[k]) continue;
    HIP_OK(hipMemcpyAsync(tmp.data(), s->macro_d[k], np * rb, hipMemcpyDeviceToHost, s->stream));
    SYNC_OK(s);
    for (size_t n = 0; n < np; ++n)
      dst[k][n] = rb == 8 ? reinterpret_cast<const double*>(tmp.data())[n]
                          : double(reinterpret_cast<const float*>(tmp.data())[n]);
    // nodes the step does not collide (Solid, Excluded, specular ghosts and
    // the population-extrapolation and equilibrium outlets): report 0
    for (size_t n = 0; n < np; ++n) {
      const uint8_t f = s->flags_h[n];
      bool stored = f == M3LB_FLUID || f == M3LB_SPECNODE;
      if (f == M3LB_REGWALL)
        stored = s->b.nrm[n] != m3lb::kNrmOutFree && s->b.nrm[n] != m3lb::kNrmOutEq;
      if (!stored) dst[k][n] = 0.0;
    }
  }
  return M3LB_OK;
}

static int reduce4(m3lb_solver* s, double out[4]) {
  const unsigned grid = reduce_grid(s);
  int rc = ensure_partials(s, size_t(grid) * 4);
  if (rc) return rc;
  m3lb::ReadArgs a = read_args(s);
  a.out = s->partial_d;
  s->ks->reduce_mass(a, grid, s->stream);
  rc = launch_ok();
  if (rc) return rc;
  std::vector<double> h(size_t(grid) * 4);
  HIP_OK(hipMemcpyAsync(h.data(), s->partial_d, h.size() * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  SYNC_OK(s);
  for (int k = 0; k < 4; ++k) out[k] = 0;
  for (unsigned b = 0; b < grid; ++b)
    for (int k = 0; k < 4; ++k) out[k] += h[size_t(b) * 4 + k];
  return M3LB_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Totals without allocation (m3lb_totals_static).  The per-block partials and
// the folded result live in a file-scope device array of fixed size, one copy
// per GPU, addressed through the pointer hipGetSymbolAddress returns for the
// device (the array's own name is a host-side stand-in, not a device address).
// Every call on a GPU uses that GPU's copy: calls that overlap on one GPU race.
// ---------------------------------------------------------------------------
constexpr unsigned kScratchBlocks = 4096;  // = the largest reduce_grid
constexpr size_t kScratchDoubles = size_t(kScratchBlocks) * 4 + 64;
static __device__ __align__(256) double scratch[kScratchDoubles];

// Sums each of the 4 columns of `grid` rows of partials in row order into
// out[0..3]: the order, and so the bits, of reduce4's host loop.
__global__ void k_fold_rows4(const double* partial, unsigned grid, double* out) {
  const unsigned k = threadIdx.x;
  if (k >= 4) return;
  double s = 0;
  for (unsigned b = 0; b < grid; ++b) s += partial[size_t(b) * 4 + k];
  out[k] = s;
}

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// This device's copy of `scratch`, looked up once per device with it current.
int scratch_of(int device, double** out) {
  if (device < 0 || device >= kMaxDevices) return fail(M3LB_EDEVICE, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    int cur = -1;
    HIP_OK(hipGetDevice(&cur));
    if (cur != device) HIP_OK(hipSetDevice(device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != device) HIP_OK(hipSetDevice(cur));
    if (e != hipSuccess || !p)
      return fail(M3LB_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = static_cast<double*>(p);
  }
  *out = g_scratch_ptr[device];
  return M3LB_OK;
}
}  // namespace

extern "C" {

int m3lb_totals_static(m3lb_solver* s, double out[4]) {
  if (!s || !out) return fail(M3LB_EINVAL, "null argument");
  const unsigned grid = reduce_grid(s);
  if (size_t(grid) * 4 + 4 > kScratchDoubles)
    return fail(M3LB_EINVAL, "totals_static: " + std::to_string(grid) +
                                 " reduction blocks exceed the static scratch of " +
                                 std::to_string(kScratchBlocks));
  double* sc = nullptr;
  int rc = scratch_of(s->device, &sc);
  if (rc) return rc;
  double* result = sc + size_t(kScratchBlocks) * 4;
  m3lb::ReadArgs a = read_args(s);
  a.out = sc;  // every block writes its 4 partials: nothing to clear
  s->ks->reduce_mass(a, grid, s->stream);
  rc = launch_ok();
  if (rc) return rc;
  k_fold_rows4<<<1, 64, 0, s->stream>>>(sc, grid, result);
  rc = launch_ok();
  if (rc) return rc;
  HIP_OK(hipMemcpyAsync(out, result, 4 * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  SYNC_OK(s);
  return M3LB_OK;
}

int m3lb_total_mass(m3lb_solver* s, double* out) {
  if (!s || !out) return fail(M3LB_EINVAL, "null argument");
  double r[4];
  int rc = reduce4(s, r);
  if (rc) return rc;
  *out = r[0];
  return M3LB_OK;
}

int m3lb_total_momentum(m3lb_solver* s, int32_t axis, double* out) {
  if (!s || !out) return fail(M3LB_EINVAL, "null argument");
  if (axis < 0 || axis > 2) return fail(M3LB_EINVAL, "axis must be 0, 1 or 2");
  double r[4];
  int rc = reduce4(s, r);
  if (rc) return rc;
  *out = r[1 + axis];
  return M3LB_OK;
}

int m3lb_momentum_exchange(m3lb_solver* s, const uint8_t* body, double force[3]) {
  if (!s || !force) return fail(M3LB_EINVAL, "null argument");
  const m3lb::Dom& d = s->d;
  std::vector<uint8_t> pad(size_t(d.np), 0);
  if (body) {
    size_t k = 0;
    for (int z = 0; z < d.nz; ++z)
      for (int y = 0; y < d.ny; ++y)
        for (int x = 0; x < d.nx; ++x, ++k)
          pad[size_t(((z + d.hz) * d.sy + (y + d.hy)) * d.sx + x + d.hx)] = body[k] ? 1 : 0;
  } else {
    for (size_t n = 0; n < pad.size(); ++n) pad[n] = s->flags_h[n] == M3LB_SOLID ? 1 : 0;
  }
  const unsigned grid = reduce_grid(s);
  int rc = ensure_partials(s, size_t(grid) * 3);
  if (rc) return rc;
  uint8_t* body_d = nullptr;
  HIP_OK(hipMalloc(&body_d, pad.size()));
  rc = M3LB_OK;
  if (hipMemcpyAsync(body_d, pad.data(), pad.size(), hipMemcpyHostToDevice, s->stream) != hipSuccess)
    rc = fail(M3LB_EDEVICE, "body upload failed");
  std::vector<double> h(size_t(grid) * 3);
  if (!rc) {
    m3lb::ReadArgs a = read_args(s);
    a.out = s->partial_d;
    s->ks->exchange(a, body_d, grid, s->stream);
    rc = launch_ok();
  }
  if (!rc && hipMemcpyAsync(h.data(), s->partial_d, h.size() * sizeof(double), hipMemcpyDeviceToHost, s->stream) != hipSuccess)
    rc = fail(M3LB_EDEVICE, "partials download failed");
  const hipError_t e = hipStreamSynchronize(s->stream);
  (void)hipFree(body_d);
  if (rc) return rc;
  if (e != hipSuccess) return fail(M3LB_EDEVICE, hipGetErrorString(e));
  SYNC_OK(s);  // reports a failed two-step pass
  for (int k = 0; k < 3; ++k) force[k] = 0;
  for (unsigned b = 0; b < grid; ++b)
    for (int k = 0; k < 3; ++k) force[k] += h[size_t(b) * 3 + k];
  return M3LB_OK;
}

int m3lb_storage_sum_exact(m3lb_solver* s, int64_t out[2]) {
  if (!s || !out) return fail(M3LB_EINVAL, "null argument");
  const unsigned grid = reduce_grid(s);
  int rc = ensure_partials(s, size_t(grid) * 2);
  if (rc) return rc;
  m3lb::ReadArgs a = read_args(s);
  a.out = s->partial_d;
  s->ks->sum_exact(a, grid, s->stream);
  rc = launch_ok();
  if (rc) return rc;
  std::vector<uint64_t> h(size_t(grid) * 2);
  HIP_OK(hipMemcpyAsync(h.data(), s->partial_d, h.size() * sizeof(uint64_t), hipMemcpyDeviceToHost, s->stream));
  SYNC_OK(s);
  unsigned __int128 acc = 0;
  for (unsigned b = 0; b < grid; ++b)
    acc += ((unsigned __int128)h[2 * b] << 64) | (unsigned __int128)h[2 * b + 1];
  out[0] = int64_t(uint64_t(acc >> 64));
  out[1] = int64_t(uint64_t(acc));
  return M3LB_OK;
}

}  // extern "C"

// E = sum 1/2|u|^2 and Omega = sum 1/2|curl u|^2 over interior nodes with
// periodic-wrap central differences (validation/Campaign.hpp:99-130).
namespace {
__global__ void __launch_bounds__(m3lb::kBlock) k_energy(m3lb::Dom d, const double* ux,
                                                         const double* uy, const double* uz,
                                                         double* out, int* finite) {
  const long long N = (long long)d.nx * d.ny * d.nz;
  double acc[2] = {0, 0};
  bool ok = true;
  for (long long i = blockIdx.x * (long long)m3lb::kBlock + threadIdx.x; i < N;
       i += (long long)gridDim.x * m3lb::kBlock) {
    const int x = int(i % d.nx), y = int((i / d.nx) % d.ny), z = int(i / ((long long)d.nx * d.ny));
    auto id = [&](int xx, int yy, int zz) {
      xx = (xx + d.nx) % d.nx; yy = (yy + d.ny) % d.ny; zz = (zz + d.nz) % d.nz;
      return ((zz + d.hz) * d.sy + (yy + d.hy)) * d.sx + xx + d.hx;
    };
    const int n = id(x, y, z);
    const double a = ux[n], b = uy[n], c = uz[n];
    if (!isfinite(a) || !isfinite(b) || !isfinite(c)) { ok = false; continue; }
    acc[0] += 0.5 * (a * a + b * b + c * c);
    const double wz = 0.5 * (uy[id(x + 1, y, z)] - uy[id(x - 1, y, z)]) -
                      0.5 * (ux[id(x, y + 1, z)] - ux[id(x, y - 1, z)]);
    double wx = 0, wy = 0;
    if (d.nz > 1) {
      wx = 0.5 * (uz[id(x, y + 1, z)] - uz[id(x, y - 1, z)]) -
           0.5 * (uy[id(x, y, z + 1)] - uy[id(x, y, z - 1)]);
      wy = 0.5 * (ux[id(x, y, z + 1)] - ux[id(x, y, z - 1)]) -
           0.5 * (uz[id(x + 1, y, z)] - uz[id(x - 1, y, z)]);
    }
    acc[1] += 0.5 * (wx * wx + wy * wy + wz * wz);
  }
  if (!ok) *finite = 0;
  m3lb::block_sum<2>(acc, out);
}

// Plain in-place stream, one float4 per thread: the best a memory-bound
// kernel with no reuse reaches on this device.
typedef float f4v __attribute__((ext_vector_type(4)));
template <bool NT>
__global__ void __launch_bounds__(m3lb::kBlock) k_stream(f4v* a, uint64_t n4) {
  const uint64_t i = uint64_t(blockIdx.x) * m3lb::kBlock + threadIdx.x;
  if (i >= n4) return;
  f4v v;
  if constexpr (NT) v = __builtin_nontemporal_load(&a[i]);
  else v = a[i];
  v = v * 1.0000001f;
  if constexpr (NT) __builtin_nontemporal_store(v, &a[i]);
  else a[i] = v;
}
}  // namespace

extern "C" int m3lb_stream_probe(uint64_t bytes, int32_t nontemporal, int32_t reps,
                                 double* best_ms) {
  if (!best_ms || bytes < 16 || reps < 1) return fail(M3LB_EINVAL, "bad argument");
  const uint64_t n4 = bytes / 16;
  f4v* a = nullptr;
  if (hipMalloc(&a, n4 * 16) != hipSuccess) return fail(M3LB_ENOMEM, "stream probe: hipMalloc failed");
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = M3LB_OK;
  if (hipStreamCreate(&st) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess || hipMemsetAsync(a, 0, n4 * 16, st) != hipSuccess)
    rc = fail(M3LB_EDEVICE, "stream probe: setup failed");
  const unsigned grid = unsigned((n4 + m3lb::kBlock - 1) / m3lb::kBlock);
  double best = 1e30;
  for (int r = 0; r < reps && !rc; ++r) {
    hipEventRecord(e0, st);
    if (nontemporal) k_stream<true><<<grid, m3lb::kBlock, 0, st>>>(a, n4);
    else k_stream<false><<<grid, m3lb::kBlock, 0, st>>>(a, n4);
    hipEventRecord(e1, st);
    rc = launch_ok();
    if (!rc && hipEventSynchronize(e1) != hipSuccess) rc = fail(M3LB_EDEVICE, "stream probe: sync failed");
    float ms = 0;
    if (!rc) hipEventElapsedTime(&ms, e0, e1);
    if (!rc && double(ms) < best) best = double(ms);
  }
  if (st) hipStreamSynchronize(st);
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  if (st) hipStreamDestroy(st);
  hipFree(a);
  if (rc) return rc;
  *best_ms = best;
  return M3LB_OK;
}

extern "C" int m3lb_energy_enstrophy(m3lb_solver* s, double* energy, double* enstrophy,
                                     int32_t* finite) {
  if (!s || !energy || !enstrophy) return fail(M3LB_EINVAL, "null argument");
  const size_t np = size_t(s->d.np);
  double* buf = nullptr;
  int* fl = nullptr;
  HIP_OK(hipMalloc(&buf, 3 * np * sizeof(double)));
  HIP_OK(hipMalloc(&fl, sizeof(int)));
  const int one = 1;
  HIP_OK(hipMemcpyAsync(fl, &one, sizeof(int), hipMemcpyHostToDevice, s->stream));
  m3lb::ReadArgs a = read_args(s);
  a.out2[0] = nullptr;
  for (int k = 0; k < 3; ++k) a.out2[1 + k] = buf + k * np;
  s->ks->macro(a, s->stream);
  const long long N = (long long)s->d.nx * s->d.ny * s->d.nz;
  const unsigned grid = std::max(1u, std::min(4096u, unsigned((N + m3lb::kBlock - 1) / m3lb::kBlock)));
  int rc = ensure_partials(s, size_t(grid) * 2);
  if (!rc) {
    k_energy<<<grid, m3lb::kBlock, 0, s->stream>>>(s->d, buf, buf + np, buf + 2 * np, s->partial_d, fl);
    rc = launch_ok();
  }
  std::vector<double> h(size_t(grid) * 2);
  int okv = 1;
  if (!rc) {
    if (hipMemcpyAsync(h.data(), s->partial_d, h.size() * sizeof(double), hipMemcpyDeviceToHost, s->stream) != hipSuccess ||
        hipMemcpyAsync(&okv, fl, sizeof(int), hipMemcpyDeviceToHost, s->stream) != hipSuccess)
      rc = fail(M3LB_EDEVICE, "energy download failed");
  }
  hipError_t e = hipStreamSynchronize(s->stream);
  hipFree(buf);
  hipFree(fl);
  if (rc) return rc;
  if (e != hipSuccess) return fail(M3LB_EDEVICE, hipGetErrorString(e));
  SYNC_OK(s);  // reports a failed two-step pass
  double E = 0, W = 0;
  for (unsigned b = 0; b < grid; ++b) { E += h[2 * b]; W += h[2 * b + 1]; }
  *energy = E;
  *enstrophy = W;
  if (finite) *finite = okv;
  return M3LB_OK;
}

extern "C" {

static int read_pops(m3lb_solver* s, double* out, bool logical) {
  if (!s || !out) return fail(M3LB_EINVAL, "null argument");
  const size_t n = size_t(s->d.np) * size_t(s->Q);
  double* buf = nullptr;
  HIP_OK(hipMalloc(&buf, n * sizeof(double)));
  m3lb::ReadArgs a = read_args(s);
  a.out = buf;
  if (logical) s->ks->gather(a, s->stream);
  else s->ks->storage(a, s->stream);
  int rc = launch_ok();
  if (!rc && hipMemcpyAsync(out, buf, n * sizeof(double), hipMemcpyDeviceToHost, s->stream) != hipSuccess)
    rc = fail(M3LB_EDEVICE, "population download failed");
  hipError_t e = hipStreamSynchronize(s->stream);
  hipFree(buf);
  if (rc) return rc;
  if (e != hipSuccess) return fail(M3LB_EDEVICE, hipGetErrorString(e));
  SYNC_OK(s);  // reports a failed two-step pass
  return M3LB_OK;
}

int m3lb_gather_populations(m3lb_solver* s, double* out) { return read_pops(s, out, true); }
int m3lb_storage(m3lb_solver* s, double* out) { return read_pops(s, out, false); }

int m3lb_timestep(const m3lb_solver* s, uint64_t* t) {
  if (!s || !t) return fail(M3LB_EINVAL, "null argument");
  *t = s->t;
  return M3LB_OK;
}

int m3lb_set_regularized_walls(m3lb_solver* s, const uint8_t* normals, const uint16_t* tags,
                                const m3lb_wall* table, int32_t ntab) {
  if (!s || !normals || !tags || !table) return fail(M3LB_EINVAL, "null argument");
  if (ntab < 1 || ntab > 65536) return fail(M3LB_EINVAL, "the wall table holds 1..65536 entries");
  if (s->cfg.op == M3LB_OP_NULL) return fail(M3LB_EINVAL, "the identity collision has no walls");
  const m3lb::Dom& d = s->d;
  const size_t np = size_t(d.np);
  auto id = [&](int x, int y, int z) { return size_t(((z + d.hz) * d.sy + (y + d.hy)) * d.sx + x + d.hx); };
  // validate everything before changing anything
  {
    size_t k = 0;
    for (int z = 0; z < d.nz; ++z)
      for (int y = 0; y < d.ny; ++y)
        for (int x = 0; x < d.nx; ++x, ++k) {
          const int c = normals[k];
          if (c > m3lb::kNrmOutEq) return fail(M3LB_EINVAL, "normal codes are 0..10");
          if (c && tags[k] >= ntab) return fail(M3LB_EINVAL, "wall tag outside the table");
          // the outlets read their upstream nodes along -x (two for order 2)
          if ((c == m3lb::kNrmOutXp && x < 2) || (c == m3lb::kNrmOutEq && x < 1))
            return fail(M3LB_EINVAL, "an outflow node needs upstream nodes along -x");
        }
  }
  SYNC_OK(s);
  Boundaries& b = s->b;
  b.nrm.assign(np, 0);
  b.tag.assign(np, 0);
  b.ext.assign(np, 0);
  b.unk.assign(np, 0);
  b.onrm.assign(np, 0);
  b.don.resize(np);
  for (size_t n = 0; n < np; ++n) b.don[n] = unsigned(n);
  size_t k = 0;
  for (int z = 0; z < d.nz; ++z)
    for (int y = 0; y < d.ny; ++y)
      for (int x = 0; x < d.nx; ++x, ++k) {
        if (normals[k] == 0) continue;
        b.nrm[id(x, y, z)] = normals[k];
        b.tag[id(x, y, z)] = tags[k];
        s->flags_h[id(x, y, z)] = M3LB_REGWALL;
      }
  // corners: an axis direction whose next two nodes are straight-wall nodes
  static const int dirs[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
  bool corners = false;
  for (int z = 0; z < d.nz; ++z)
    for (int y = 0; y < d.ny; ++y)
      for (int x = 0; x < d.nx; ++x) {
        if (b.nrm[id(x, y, z)] != m3lb::kNrmCorner) continue;
        corners = true;
        for (int q = 0; q < (d.nz > 1 ? 6 : 4); ++q) {
          const int x1 = x + dirs[q][0], y1 = y + dirs[q][1], z1 = z + dirs[q][2];
          const int x2 = x + 2 * dirs[q][0], y2 = y + 2 * dirs[q][1], z2 = z + 2 * dirs[q][2];
          if (x1 < 0 || y1 < 0 || z1 < 0 || x2 < 0 || y2 < 0 || z2 < 0) continue;
          if (x1 >= d.nx || y1 >= d.ny || z1 >= d.nz || x2 >= d.nx || y2 >= d.ny || z2 >= d.nz) continue;
          const int a = b.nrm[id(x1, y1, z1)], c = b.nrm[id(x2, y2, z2)];
          if (a != 0 && a != m3lb::kNrmCorner && c != 0 && c != m3lb::kNrmCorner) {
            b.ext[id(x, y, z)] = uint8_t(q + 1);
            break;
          }
        }
      }
  // unknown directions: source p - c_i outside a non-periodic extent, or an
  // Excluded or Solid node (FluidSolver.hpp:420-455)
  const int next[3] = {d.nx, d.ny, d.nz}, per[3] = {d.px, d.py, d.pz};
  for (int z = 0; z < d.nz; ++z)
    for (int y = 0; y < d.ny; ++y)
      for (int x = 0; x < d.nx; ++x) {
        const size_t n = id(x, y, z);
        if (!b.nrm[n]) continue;
        uint32_t m = 0;
        for (int i = 0; i < s->Q; ++i) {
          const int c[3] = {s->Q == 9 ? m3lb::Lat<9>::cx[i] : m3lb::Lat<27>::cx[i],
                            s->Q == 9 ? m3lb::Lat<9>::cy[i] : m3lb::Lat<27>::cy[i],
                            s->Q == 9 ? m3lb::Lat<9>::cz[i] : m3lb::Lat<27>::cz[i]};
          int src[3] = {x - c[0], y - c[1], z - c[2]};
          bool outside = false;
          for (int a = 0; a < 3; ++a) {
            if (per[a]) src[a] = (src[a] + next[a]) % next[a];
            else if (src[a] < 0 || src[a] >= next[a]) outside = true;
          }
          if (!outside) {
            const uint8_t sf = s->flags_h[id(src[0], src[1], src[2])];
            if (sf == M3LB_EXCLUDED || sf == M3LB_SOLID) outside = true;
          }
          if (outside) m |= 1u << i;
        }
        b.unk[n] = m;
      }
  // NrmOutFree donors (:475-520): the first axis step that leaves the fluid
  // (outside, or Solid) whose opposite neighbour is Fluid; none: the node itself
  {
    static const int dd[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
    static const uint8_t codes[6] = {2, 1, 4, 3, 6, 5};
    for (int z = 0; z < d.nz; ++z)
      for (int y = 0; y < d.ny; ++y)
        for (int x = 0; x < d.nx; ++x) {
          const size_t n = id(x, y, z);
          if (b.nrm[n] != m3lb::kNrmOutFree) continue;
          for (int q = 0; q < 6; ++q) {
            const int qx = x + dd[q][0], qy = y + dd[q][1], qz = z + dd[q][2];
            const bool leaves = qx < 0 || qx >= d.nx || qy < 0 || qy >= d.ny || qz < 0 ||
                                qz >= d.nz || s->flags_h[id(qx, qy, qz)] == M3LB_SOLID;
            if (!leaves) continue;
            const int o = q ^ 1;
            const int rx = x + dd[o][0], ry = y + dd[o][1], rz = z + dd[o][2];
            if (rx < 0 || rx >= d.nx || ry < 0 || ry >= d.ny || rz < 0 || rz >= d.nz) continue;
            const size_t m = id(rx, ry, rz);
            if (s->flags_h[m] != M3LB_FLUID) continue;
            b.don[n] = unsigned(m);
            b.onrm[n] = codes[q];
            break;
          }
        }
  }
  b.wtab.resize(size_t(ntab) * 4);
  for (int t = 0; t < ntab; ++t) {
    const double v[4] = {table[t].ux, table[t].uy, table[t].uz, table[t].rho};
    for (int a = 0; a < 4; ++a) b.wtab[size_t(t) * 4 + size_t(a)] = as_real(s, v[a]);
  }
  b.wtab0 = b.wtab;
  b.has_corners = corners;
  int rc = upload(b.nrm_d, b.nrm, s->stream);
  if (!rc) rc = upload(b.tag_d, b.tag, s->stream);
  if (!rc) rc = upload(b.unk_d, b.unk, s->stream);
  if (!rc) rc = upload(b.ext_d, b.ext, s->stream);
  if (!rc) rc = upload(b.onrm_d, b.onrm, s->stream);
  if (!rc) rc = upload(b.don_d, b.don, s->stream);
  if (!rc) rc = upload(b.wtab_d, b.wtab, s->stream);
  if (rc) return rc;
  if (!b.rho_d) {
    HIP_OK(hipMalloc(&b.rho_d, np * size_t(s->ks->real_bytes)));
    HIP_OK(hipMemsetAsync(b.rho_d, 0, np * size_t(s->ks->real_bytes), s->stream));
  }
  HIP_OK(hipMemcpyAsync(s->flags_d, s->flags_h.data(), np, hipMemcpyHostToDevice, s->stream));
  b.walls_set = true;
  b.dirty = true;
  refresh_geometry(s);
  s->macro_valid = false;
  return sync(s);
}

int m3lb_set_fd_corners(m3lb_solver* s, int32_t on) {
  if (!s) return fail(M3LB_EINVAL, "null handle");
  s->b.fd_corners = on ? 1 : 0;
  s->b.dirty = true;
  return M3LB_OK;
}

int m3lb_set_force_correction(m3lb_solver* s, int32_t on) {
  if (!s) return fail(M3LB_EINVAL, "null handle");
  s->b.force_bc = on ? 1 : 0;
  return M3LB_OK;
}

int m3lb_set_outflow_order(m3lb_solver* s, int32_t order) {
  if (!s) return fail(M3LB_EINVAL, "null handle");
  s->b.out_order = order;
  s->b.dirty = true;
  return M3LB_OK;
}

int m3lb_set_outflow_density(m3lb_solver* s, double rho) {
  if (!s) return fail(M3LB_EINVAL, "null handle");
  if (!std::isfinite(rho)) return fail(M3LB_EINVAL, "density must be finite");
  s->b.outflow_rho = as_real(s, rho);
  return M3LB_OK;
}

int m3lb_set_wall_velocity_scale(m3lb_solver* s, double scale) {
  if (!s) return fail(M3LB_EINVAL, "null handle");
  Boundaries& b = s->b;
  if (!b.walls_set) return fail(M3LB_ESTATE, "no regularised walls are set");
  SYNC_OK(s);
  for (size_t t = 0; t + 3 < b.wtab.size(); t += 4)
    for (size_t a = 0; a < 3; ++a) b.wtab[t + a] = real_mul(s, b.wtab0[t + a], scale);
  HIP_OK(hipMemcpy(b.wtab_d, b.wtab.data(), b.wtab.size() * sizeof(double), hipMemcpyHostToDevice));
  return M3LB_OK;
}

int m3lb_set_wall_density(m3lb_solver* s, double rho) {
  if (!s) return fail(M3LB_EINVAL, "null handle");
  Boundaries& b = s->b;
  if (!b.walls_set) return fail(M3LB_ESTATE, "no regularised walls are set");
  SYNC_OK(s);
  for (size_t t = 0; t + 3 < b.wtab.size(); t += 4) b.wtab[t + 3] = b.wtab0[t + 3] = as_real(s, rho);
  HIP_OK(hipMemcpy(b.wtab_d, b.wtab.data(), b.wtab.size() * sizeof(double), hipMemcpyHostToDevice));
  return M3LB_OK;
}

int m3lb_set_wall_velocity_field(m3lb_solver* s, const double* ux, const double* uy, const double* uz) {
  if (!s || !ux || !uy || !uz) return fail(M3LB_EINVAL, "null argument");
  const m3lb::Dom& d = s->d;
  const size_t np = size_t(d.np), rb = size_t(s->ks->real_bytes);
  std::vector<char> h(3 * np * rb, 0);
  size_t k = 0;
  for (int z = 0; z < d.nz; ++z)
    for (int y = 0; y < d.ny; ++y)
      for (int x = 0; x < d.nx; ++x, ++k) {
        const size_t n = size_t(((z + d.hz) * d.sy + (y + d.hy)) * d.sx + x + d.hx);
        const double v[3] = {ux[k], uy[k], uz[k]};
        for (int a = 0; a < 3; ++a) {
          if (rb == 8) reinterpret_cast<double*>(h.data())[size_t(a) * np + n] = v[a];
          else reinterpret_cast<float*>(h.data())[size_t(a) * np + n] = float(v[a]);
        }
      }
  SYNC_OK(s);
  Boundaries& b = s->b;
  if (!b.un_d) HIP_OK(hipMalloc(&b.un_d, h.size()));
  HIP_OK(hipMemcpy(b.un_d, h.data(), h.size(), hipMemcpyHostToDevice));
  b.un_on = true;
  b.un_scale = 1.0;
  return M3LB_OK;
}

int m3lb_set_wall_velocity_field_scale(m3lb_solver* s, double scale) {
  if (!s) return fail(M3LB_EINVAL, "null handle");
  s->b.un_scale = as_real(s, scale);
  return M3LB_OK;
}

int m3lb_set_specular_walls(m3lb_solver* s, const uint8_t* normals) {
  if (!s || !normals) return fail(M3LB_EINVAL, "null argument");
  if (s->cfg.op == M3LB_OP_NULL) return fail(M3LB_EINVAL, "the identity collision has no walls");
  const m3lb::Dom& d = s->d;
  const size_t np = size_t(d.np), ni = size_t(d.nx) * size_t(d.ny) * size_t(d.nz);
  for (size_t k = 0; k < ni; ++k)
    if (normals[k] > M3LB_NRM_ZM)
      return fail(M3LB_EINVAL, "specular reflection is defined for the axis normals 1..6 only");
  SYNC_OK(s);
  Boundaries& b = s->b;
  b.spec_nrm.assign(np, 0);
  size_t k = 0;
  for (int z = 0; z < d.nz; ++z)
    for (int y = 0; y < d.ny; ++y)
      for (int x = 0; x < d.nx; ++x, ++k) {
        if (!normals[k]) continue;
        const size_t n = size_t(((z + d.hz) * d.sy + (y + d.hy)) * d.sx + x + d.hx);
        b.spec_nrm[n] = normals[k];
        s->flags_h[n] = M3LB_SPECWALL;
      }
  int rc = upload(b.spec_nrm_d, b.spec_nrm, s->stream);
  if (rc) return rc;
  if (!b.spec_faces_d) {
    b.spec_faces.assign(np, 0);
    rc = upload(b.spec_faces_d, b.spec_faces, s->stream);
    if (rc) return rc;
  }
  HIP_OK(hipMemcpyAsync(s->flags_d, s->flags_h.data(), np, hipMemcpyHostToDevice, s->stream));
  b.dirty = true;
  refresh_geometry(s);
  s->macro_valid = false;
  return sync(s);
}

int m3lb_set_specular_nodes(m3lb_solver* s, const uint8_t* faces) {
  if (!s || !faces) return fail(M3LB_EINVAL, "null argument");
  if (s->cfg.op == M3LB_OP_NULL) return fail(M3LB_EINVAL, "the identity collision has no walls");
  const m3lb::Dom& d = s->d;
  const size_t np = size_t(d.np), ni = size_t(d.nx) * size_t(d.ny) * size_t(d.nz);
  for (size_t k = 0; k < ni; ++k) {
    if (faces[k] > 0x3F) return fail(M3LB_EINVAL, "face masks use bits 0..5 only");
    for (int a = 0; a < 3; ++a)
      if ((faces[k] & (1u << (2 * a))) && (faces[k] & (1u << (2 * a + 1))))
        return fail(M3LB_EINVAL, "a face mask carries both faces of one axis");
  }
  SYNC_OK(s);
  Boundaries& b = s->b;
  b.spec_faces.assign(np, 0);
  size_t k = 0;
  for (int z = 0; z < d.nz; ++z)
    for (int y = 0; y < d.ny; ++y)
      for (int x = 0; x < d.nx; ++x, ++k) {
        if (!faces[k]) continue;
        const size_t n = size_t(((z + d.hz) * d.sy + (y + d.hy)) * d.sx + x + d.hx);
        b.spec_faces[n] = faces[k];
        s->flags_h[n] = M3LB_SPECNODE;
      }
  int rc = upload(b.spec_faces_d, b.spec_faces, s->stream);
  if (rc) return rc;
  if (!b.spec_nrm_d) {
    b.spec_nrm.assign(np, 0);
    rc = upload(b.spec_nrm_d, b.spec_nrm, s->stream);
    if (rc) return rc;
  }
  HIP_OK(hipMemcpyAsync(s->flags_d, s->flags_h.data(), np, hipMemcpyHostToDevice, s->stream));
  b.dirty = true;
  refresh_geometry(s);
  s->macro_valid = false;
  return sync(s);
}

int m3lb_set_force_field(m3lb_solver* s, const double* fx, const double* fy, const double* fz) {
  if (!s || !fx || !fy || !fz) return fail(M3LB_EINVAL, "null argument");
  if (s->cfg.forcing != M3LB_FORCE_FIELD || !s->ff_d)
    return fail(M3LB_ESTATE, "the handle was not created with M3LB_FORCE_FIELD");
  const size_t np = size_t(s->d.np), st = size_t(s->stride);
  const double* src[3] = {fx, fy, fz};
  SYNC_OK(s);  // the previous field may still be read
  if (s->ks->real_bytes == 8) {
    std::vector<double> h(3 * st, 0.0);
    for (int k = 0; k < 3; ++k) std::copy(src[k], src[k] + np, h.begin() + k * st);
    HIP_OK(hipMemcpy(s->ff_d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
  } else {
    std::vector<float> h(3 * st, 0.0f);
    for (int k = 0; k < 3; ++k)
      for (size_t n = 0; n < np; ++n) h[k * st + n] = float(src[k][n]);
    HIP_OK(hipMemcpy(s->ff_d, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
  }
  s->macro_valid = false;
  return M3LB_OK;
}

int m3lb_set_force(m3lb_solver* s, double fx, double fy, double fz) {
  if (!s) return fail(M3LB_EINVAL, "null handle");
  if (s->cfg.forcing != M3LB_FORCE_GUO) return fail(M3LB_ESTATE, "the handle was created without Guo forcing");
  s->cfg.force[0] = fx; s->cfg.force[1] = fy; s->cfg.force[2] = fz;
  return M3LB_OK;
}

int m3lb_set_omega(m3lb_solver* s, double omega, double omega_minus, double omega_bulk) {
  if (!s) return fail(M3LB_EINVAL, "null handle");
  if (!(std::isfinite(omega) && std::isfinite(omega_minus) && std::isfinite(omega_bulk)))
    return fail(M3LB_EINVAL, "relaxation rates must be finite");
  s->cfg.omega = omega; s->cfg.omega_minus = omega_minus; s->cfg.omega_bulk = omega_bulk;
  return M3LB_OK;
}

int m3lb_get_scales(m3lb_solver* s, double* sc) {
  if (!s || !sc) return fail(M3LB_EINVAL, "null argument");
  for (int q = 0; q < s->Q; ++q) sc[q] = s->scale[q];
  return M3LB_OK;
}

static int apply_scales(m3lb_solver* s, const double* want) {
  m3lb::Ratio r;
  for (int q = 0; q < 27; ++q) r.r[q] = 1.0f;
  for (int q = 0; q < s->Q; ++q) r.r[q] = float(want[q] / double(s->scale[q]));
  const unsigned grid = reduce_grid(s);
  s->ks->rescale(s->f, s->stride, s->d.np, r, grid, s->stream);
  int rc = launch_ok();
  if (rc) return rc;
  for (int q = 0; q < s->Q; ++q) {
    s->scale[q] = float(want[q]);
    s->inv_scale[q] = float(1.0 / want[q]);
  }
  return sync(s);
}

int m3lb_set_scales(m3lb_solver* s, const double* sc) {
  if (!s || !sc) return fail(M3LB_EINVAL, "null argument");
  if (!is16(s)) return M3LB_OK;
  for (int q = 0; q < s->Q; ++q) {
    int e;
    if (!(sc[q] > 0) || std::frexp(sc[q], &e) != 0.5)
      return fail(M3LB_EINVAL, "scales must be positive powers of two");
    // the codecs' folded scale constants are exact for 2^-60 .. 2^60, the
    // range calibration itself keeps to (scales.hpp, scale_exponent)
    if (e - 1 < -60 || e - 1 > 60)
      return fail(M3LB_EINVAL, "scales must lie in 2^-60 .. 2^60");
    const int o = q == 0 ? 0 : (q & 1 ? q + 1 : q - 1);
    if (sc[q] != sc[o]) return fail(M3LB_EINVAL, "opposite directions must share a scale");
  }
  return apply_scales(s, sc);
}

int m3lb_set_tile_order(m3lb_solver* s, int32_t run) {
  if (!s) return fail(M3LB_EINVAL, "null handle");
  s->remap = run;
  return M3LB_OK;
}

int m3lb_set_cache_policy(m3lb_solver* s, int32_t policy) {
  if (!s) return fail(M3LB_EINVAL, "null handle");
  if (policy < M3LB_CACHE_AUTO || policy > M3LB_CACHE_PLAIN)
    return fail(M3LB_EINVAL, "cache policy is AUTO (0), STREAM (1) or PLAIN (2)");
  s->cache_policy = policy;
  return M3LB_OK;
}

int m3lb_tile_order_map(uint32_t grid, int32_t run, uint32_t* out) {
  if (!out) return fail(M3LB_EINVAL, "null argument");
  for (uint32_t b = 0; b < grid; ++b) out[b] = m3lb::tile_of(b, grid, run);
  return M3LB_OK;
}

#ifndef M3LB_BUILD_ID
#define M3LB_BUILD_ID "unknown"
#endif
const char* m3lb_kernel_build_id(void) { return M3LB_BUILD_ID; }

int m3lb_device_bytes(const m3lb_solver* s, uint64_t* bytes) {
  if (!s || !bytes) return fail(M3LB_EINVAL, "null argument");
  uint64_t b = s->stride * uint64_t(s->Q) * uint64_t(s->ks->elem_bytes) + uint64_t(s->d.np);
  for (void* p : s->macro_d)
    if (p) b += uint64_t(s->d.np) * uint64_t(s->ks->real_bytes);
  b += s->partial_n * sizeof(double);
  *bytes = b;
  return M3LB_OK;
}

int m3lb_calibrate_scales(m3lb_solver* s) {
  if (!s) return fail(M3LB_EINVAL, "null handle");
  return is16(s) ? calibrate(s) : M3LB_OK;
}

}  // extern "C"

namespace {
// Per-pair scales the current field calls for (scales.hpp, pair_scales).
int measure_scales(m3lb_solver* s, std::vector<double>& want) {
  const unsigned grid = reduce_grid(s);
  int rc = ensure_partials(s, size_t(grid) * size_t(s->Q));
  if (rc) return rc;
  m3lb::ReadArgs a = read_args(s);
  a.out = s->partial_d;
  s->ks->absmax(a, grid, s->stream);
  rc = launch_ok();
  if (rc) return rc;
  std::vector<double> h(size_t(grid) * size_t(s->Q));
  HIP_OK(hipMemcpyAsync(h.data(), s->partial_d, h.size() * sizeof(double), hipMemcpyDeviceToHost, s->stream));
  SYNC_OK(s);
  std::vector<double> mx(size_t(s->Q), 0.0);
  want.assign(size_t(s->Q), 1.0);
  for (unsigned b = 0; b < grid; ++b)
    for (int q = 0; q < s->Q; ++q) mx[size_t(q)] = std::max(mx[size_t(q)], h[size_t(b) * s->Q + q]);
  m3lb::pair_scales(s->cfg.dtype, s->Q, mx.data(), want.data());
  return M3LB_OK;
}

// Scales from the current field; the codes are re-encoded in place.
int calibrate(m3lb_solver* s) {
  std::vector<double> want;
  const int rc = measure_scales(s, want);
  return rc ? rc : apply_scales(s, want.data());
}

// Seeds, and for 16-bit storage seeds a second time from the same source at
// the scales the first seeding calls for, so no value keeps the rounding of
// the provisional scale (which a rescale of fixed-point codes cannot undo).
int seed_calibrated(m3lb_solver* s, m3lb::SeedArgs a) {
  s->ks->seed(a, s->stream);
  int rc = launch_ok();
  if (!rc) rc = sync(s);
  if (rc || !is16(s)) return rc;
  std::vector<double> want;
  rc = measure_scales(s, want);
  if (rc) return rc;
  for (int q = 0; q < s->Q; ++q) {
    s->scale[q] = float(want[size_t(q)]);
    s->inv_scale[q] = float(1.0 / want[size_t(q)]);
    a.scale[q] = s->scale[q];
  }
  s->ks->seed(a, s->stream);
  rc = launch_ok();
  return rc ? rc : sync(s);
}
}  // namespace
