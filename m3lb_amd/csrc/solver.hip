// solver.hip -- the C-ABI of include/m3lb.h: a handle owning one in-place
// population array on the current HIP device, its flags and diagnostics
// buffers, and one HIP stream that orders every call on the handle.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "launch.hpp"

using m3lb::KernelSet;

#define M3LB_X(Q, OP) M3LB_DECLARE_LOOKUP(Q, OP);
M3LB_X(9, 0) M3LB_X(9, 1) M3LB_X(9, 2) M3LB_X(9, 3) M3LB_X(9, 4) M3LB_X(9, 5)
M3LB_X(27, 0) M3LB_X(27, 1) M3LB_X(27, 2) M3LB_X(27, 3) M3LB_X(27, 4) M3LB_X(27, 5)
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
  static const Fn q9[6] = {m3lb_kernels_q9_op0, m3lb_kernels_q9_op1, m3lb_kernels_q9_op2,
                           m3lb_kernels_q9_op3, m3lb_kernels_q9_op4, m3lb_kernels_q9_op5};
  static const Fn q27[6] = {m3lb_kernels_q27_op0, m3lb_kernels_q27_op1, m3lb_kernels_q27_op2,
                            m3lb_kernels_q27_op3, m3lb_kernels_q27_op4, m3lb_kernels_q27_op5};
  if (op < 0 || op > 5) return nullptr;
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
  if (c->op < M3LB_OP_NULL || c->op > M3LB_OP_CM) return fail(M3LB_EINVAL, "unknown operator");
  if (c->storage != M3LB_RAW && c->storage != M3LB_SHIFTED)
    return fail(M3LB_EINVAL, "storage must be M3LB_RAW or M3LB_SHIFTED");
  if (c->dtype < M3LB_F64 || c->dtype > M3LB_I16) return fail(M3LB_EINVAL, "unknown dtype");
  if ((c->dtype == M3LB_F16 || c->dtype == M3LB_BF16 || c->dtype == M3LB_I16) &&
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
  void* ff_d = nullptr;  // M3LB_FORCE_FIELD: fx, fy, fz planes of `stride` (compute type)
  // regularised walls (m3lb_set_regularized_walls)
  int nwall = 0;
  bool has_corners = false;
  int fd_corners = 1;  // fd_corners_ = has_shear_omega<Collision> (FluidSolver.hpp:1287)
  uint8_t* nrm_d = nullptr;
  uint16_t* tag_d = nullptr;
  uint32_t* unk_d = nullptr;
  uint8_t* ext_d = nullptr;
  void* wrho_d = nullptr;
  double* wtab_d = nullptr;
  unsigned* wlist_d = nullptr;
};

namespace {

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
  a.remap = s->remap;
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
  if (s->nwall) {
    a.bc_nrm = s->nrm_d;
    a.bc_tag = s->tag_d;
    a.bc_rho = s->wrho_d;
    a.wtab = s->wtab_d;
  }
  return a;
}

m3lb::WallArgs wall_args(const m3lb_solver* s) {
  m3lb::WallArgs a;
  std::memset(&a, 0, sizeof a);
  a.f = s->f;
  a.stride = s->stride;
  a.flags = s->flags_d;
  a.d = s->d;
  a.list = s->wlist_d;
  a.nwall = s->nwall;
  a.fd_corners = s->fd_corners;
  a.bc_nrm = s->nrm_d;
  a.bc_tag = s->tag_d;
  a.bc_unk = s->unk_d;
  a.bc_ext = s->ext_d;
  a.bc_rho = s->wrho_d;
  a.wtab = s->wtab_d;
  a.omega = s->cfg.omega;
  a.omega_m = s->cfg.omega_minus;
  a.omega_b = s->cfg.omega_bulk;
  for (int k = 0; k < 3; ++k) a.F[k] = s->cfg.force[k];
  a.ff = s->ff_d;
  for (int q = 0; q < 27; ++q) { a.scale[q] = s->scale[q]; a.inv_scale[q] = s->inv_scale[q]; }
  return a;
}

void free_walls(m3lb_solver* s) {
  void* ps[] = {s->nrm_d, s->tag_d, s->unk_d, s->ext_d, s->wrho_d, s->wtab_d, s->wlist_d};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  s->nrm_d = nullptr; s->tag_d = nullptr; s->unk_d = nullptr; s->ext_d = nullptr;
  s->wrho_d = nullptr; s->wtab_d = nullptr; s->wlist_d = nullptr;
  s->nwall = 0;
  s->has_corners = false;
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
  return s->cfg.dtype == M3LB_F16 || s->cfg.dtype == M3LB_BF16 || s->cfg.dtype == M3LB_I16;
}

// 16-bit scales.  A scale is a power of two 2^k per opposite pair.
//   F16, BF16: the pair's largest |g| maps into (2^7, 2^8]: FP16 then has 2^8
//     of headroom under its largest finite code (65504) for the flow to grow
//     into, and keeps full precision for values down to 2^-22 of that maximum
//     (its smallest normal is 2^-14).  A floating format's relative precision
//     is the same at every scale, so nothing is lost by mapping low.
//   I16: scales proportional to the weights (see measure_scales), the
//     largest |g| mapped into (2^11, 2^12]: 8x to 16x of headroom under the
//     saturation at 32767, and a uniform error of half a code.
// A pair whose largest |g| is below 1e-3 w_q (a field at rest) is scaled as
// if it were 1e-3 w_q.
constexpr double kScaleFloor = 1e-3;

double scale_target(int dtype) { return dtype == M3LB_I16 ? 4096.0 : 256.0; }

double weight_of(int Q, int q) {
  static const int wn9[9] = {16, 4, 4, 4, 4, 1, 1, 1, 1};
  static const int wn27[27] = {64, 16, 16, 16, 16, 16, 16, 4, 4, 4, 4, 4, 4, 4,
                               4,  4,  4,  4,  4,  1,  1,  1, 1, 1, 1, 1, 1};
  return Q == 9 ? wn9[q] / 36.0 : wn27[q] / 216.0;
}

// floor(log2(x)) exactly, for finite x > 0
int floor_log2(double x) {
  int e = 0;
  std::frexp(x, &e);  // x = m 2^e, m in [0.5, 1)
  return e - 1;
}

int scale_exponent(int dtype, int Q, int q, double max_abs) {
  const double m = std::max(max_abs, kScaleFloor * weight_of(Q, q));
  return std::max(-60, std::min(60, floor_log2(scale_target(dtype) / m)));
}

// Provisional 16-bit scales before the first seeding: |g_q| <= w_q/2 assumed.
void provisional_scales(m3lb_solver* s) {
  for (int q = 0; q < 27; ++q) { s->scale[q] = 1.0f; s->inv_scale[q] = 1.0f; }
  if (!is16(s)) return;
  for (int q = 0; q < s->Q; ++q) {
    const int k = scale_exponent(s->cfg.dtype, s->Q, q, 0.5 * weight_of(s->Q, q));
    s->scale[q] = std::ldexp(1.0f, k);
    s->inv_scale[q] = std::ldexp(1.0f, -k);
  }
}

int launch_ok() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(M3LB_EDEVICE, std::string("kernel launch: ") + hipGetErrorString(e));
  return M3LB_OK;
}

int sync(m3lb_solver* s) {
  HIP_OK(hipStreamSynchronize(s->stream));
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
  // for 4- and 8-byte populations, runs of 2 (one row) for 2-byte ones.
  s->remap = ks->elem_bytes == 2 ? 2 : 64;

  auto bail = [&](int code) {
    m3lb_destroy(s);
    return code;
  };
  if (hipGetDevice(&s->device) != hipSuccess) return bail(fail(M3LB_EDEVICE, "no HIP device"));
  if (hipStreamCreate(&s->stream) != hipSuccess) return bail(fail(M3LB_EDEVICE, "hipStreamCreate failed"));
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
  free_walls(s);
  for (void* p : s->macro_d)
    if (p) hipFree(p);
  if (s->partial_d) hipFree(s->partial_d);
  if (s->ev0) hipEventDestroy(s->ev0);
  if (s->ev1) hipEventDestroy(s->ev1);
  if (s->stream) hipStreamDestroy(s->stream);
  delete s;
  return M3LB_OK;
}

int m3lb_set_geometry(m3lb_solver* s, const uint8_t* fl) {
  if (!s || !fl) return fail(M3LB_EINVAL, "null argument");
  // every interior flag is rewritten, so regularised walls set earlier are gone
  if (s->nwall) {
    HIP_OK(hipStreamSynchronize(s->stream));
    free_walls(s);
  }
  const m3lb::Dom& d = s->d;
  bool geom = false;
  size_t k = 0;
  for (int z = 0; z < d.nz; ++z)
    for (int y = 0; y < d.ny; ++y)
      for (int x = 0; x < d.nx; ++x, ++k) {
        const uint8_t v = fl[k];
        if (v > M3LB_EXCLUDED) return fail(M3LB_EINVAL, "flag values are Fluid(0), Solid(1), Excluded(2)");
        s->flags_h[size_t(((z + d.hz) * d.sy + (y + d.hy)) * d.sx + x + d.hx)] = v;
        geom = geom || v != M3LB_FLUID;
      }
  s->geometry = geom;
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
  const m3lb::WallArgs wa = s->nwall ? wall_args(s) : m3lb::WallArgs{};
  for (int k = 0; k < nsteps; ++k) {
    const bool last = k == nsteps - 1;
    for (int j = 0; j < 4; ++j) a.macro[j] = (store_macroscopic && last) ? s->macro_d[j] : nullptr;
    const int parity = int(s->t % 2);
    if (s->nwall) {
      // FluidSolver::step (:583): the corner-density pass, then the step; the
      // wall nodes' share of the step runs first (see kernels.hpp, k_walls)
      if (s->has_corners) s->ks->wall_density(wa, parity, s->stream);
      s->ks->walls(wa, parity, s->stream);
    }
    s->ks->step(a, parity, s->geometry, grid, s->stream);
    ++s->t;
  }
  s->macro_valid = store_macroscopic != 0;
  return launch_ok();
}

int m3lb_time_steps(m3lb_solver* s, int32_t nsteps, double* ms) {
  if (!s || !ms) return fail(M3LB_EINVAL, "null argument");
  HIP_OK(hipEventRecord(s->ev0, s->stream));
  int rc = m3lb_step(s, nsteps, 0);
  if (rc) return rc;
  HIP_OK(hipEventRecord(s->ev1, s->stream));
  HIP_OK(hipEventSynchronize(s->ev1));
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
    HIP_OK(hipStreamSynchronize(s->stream));
    for (size_t n = 0; n < np; ++n)
      dst[k][n] = rb == 8 ? reinterpret_cast<const double*>(tmp.data())[n]
                          : double(reinterpret_cast<const float*>(tmp.data())[n]);
    // non-fluid nodes: the step never writes them; report 0 as compute_macroscopic does
    for (size_t n = 0; n < np; ++n)
      if (s->flags_h[n] != M3LB_FLUID) dst[k][n] = 0.0;
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
  HIP_OK(hipStreamSynchronize(s->stream));
  for (int k = 0; k < 4; ++k) out[k] = 0;
  for (unsigned b = 0; b < grid; ++b)
    for (int k = 0; k < 4; ++k) out[k] += h[size_t(b) * 4 + k];
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
  HIP_OK(hipStreamSynchronize(s->stream));
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
  std::vector<uint8_t> nrm(np, 0), ext(np, 0);
  std::vector<uint16_t> tag(np, 0);
  std::vector<uint32_t> unk(np, 0);
  size_t k = 0;
  for (int z = 0; z < d.nz; ++z)
    for (int y = 0; y < d.ny; ++y)
      for (int x = 0; x < d.nx; ++x, ++k) {
        if (normals[k] > m3lb::kNrmCorner) return fail(M3LB_EINVAL, "normal codes are 0..7");
        if (normals[k] == 0) continue;
        if (tags[k] >= ntab) return fail(M3LB_EINVAL, "wall tag outside the table");
        nrm[id(x, y, z)] = normals[k];
        tag[id(x, y, z)] = tags[k];
      }
  // corners: an axis direction whose next two nodes are straight-wall nodes
  static const int dirs[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
  bool corners = false;
  for (int z = 0; z < d.nz; ++z)
    for (int y = 0; y < d.ny; ++y)
      for (int x = 0; x < d.nx; ++x) {
        if (nrm[id(x, y, z)] != m3lb::kNrmCorner) continue;
        corners = true;
        for (int q = 0; q < (d.nz > 1 ? 6 : 4); ++q) {
          const int x1 = x + dirs[q][0], y1 = y + dirs[q][1], z1 = z + dirs[q][2];
          const int x2 = x + 2 * dirs[q][0], y2 = y + 2 * dirs[q][1], z2 = z + 2 * dirs[q][2];
          if (x1 < 0 || y1 < 0 || z1 < 0 || x2 < 0 || y2 < 0 || z2 < 0) continue;
          if (x1 >= d.nx || y1 >= d.ny || z1 >= d.nz || x2 >= d.nx || y2 >= d.ny || z2 >= d.nz) continue;
          const int a = nrm[id(x1, y1, z1)], b = nrm[id(x2, y2, z2)];
          if (a != 0 && a != m3lb::kNrmCorner && b != 0 && b != m3lb::kNrmCorner) {
            ext[id(x, y, z)] = uint8_t(q + 1);
            break;
          }
        }
      }
  // flags: the wall nodes become RegWall
  for (size_t n = 0; n < np; ++n)
    if (nrm[n]) s->flags_h[n] = M3LB_REGWALL;
  // unknown directions: source p - c_i outside a non-periodic extent, or an
  // Excluded or Solid node (FluidSolver.hpp:420-455)
  const int next[3] = {d.nx, d.ny, d.nz}, per[3] = {d.px, d.py, d.pz};
  for (int z = 0; z < d.nz; ++z)
    for (int y = 0; y < d.ny; ++y)
      for (int x = 0; x < d.nx; ++x) {
        const size_t n = id(x, y, z);
        if (!nrm[n]) continue;
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
        unk[n] = m;
      }
  std::vector<unsigned> list;
  for (size_t n = 0; n < np; ++n)
    if (nrm[n]) list.push_back(unsigned(n));
  std::vector<double> tab(size_t(ntab) * 4);
  for (int t = 0; t < ntab; ++t) {
    tab[size_t(t) * 4 + 0] = table[t].ux;
    tab[size_t(t) * 4 + 1] = table[t].uy;
    tab[size_t(t) * 4 + 2] = table[t].uz;
    tab[size_t(t) * 4 + 3] = table[t].rho;
  }
  HIP_OK(hipStreamSynchronize(s->stream));
  free_walls(s);
  HIP_OK(hipMalloc(&s->nrm_d, np));
  HIP_OK(hipMalloc(&s->tag_d, np * sizeof(uint16_t)));
  HIP_OK(hipMalloc(&s->unk_d, np * sizeof(uint32_t)));
  HIP_OK(hipMalloc(&s->ext_d, np));
  HIP_OK(hipMalloc(&s->wrho_d, np * size_t(s->ks->real_bytes)));
  HIP_OK(hipMalloc(&s->wtab_d, tab.size() * sizeof(double)));
  HIP_OK(hipMalloc(&s->wlist_d, std::max<size_t>(1, list.size()) * sizeof(unsigned)));
  HIP_OK(hipMemcpy(s->nrm_d, nrm.data(), np, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(s->tag_d, tag.data(), np * sizeof(uint16_t), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(s->unk_d, unk.data(), np * sizeof(uint32_t), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(s->ext_d, ext.data(), np, hipMemcpyHostToDevice));
  HIP_OK(hipMemset(s->wrho_d, 0, np * size_t(s->ks->real_bytes)));
  HIP_OK(hipMemcpy(s->wtab_d, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
  if (!list.empty())
    HIP_OK(hipMemcpy(s->wlist_d, list.data(), list.size() * sizeof(unsigned), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(s->flags_d, s->flags_h.data(), np, hipMemcpyHostToDevice));
  s->nwall = int(list.size());
  s->has_corners = corners;
  s->geometry = s->geometry || s->nwall > 0;
  s->macro_valid = false;
  return M3LB_OK;
}

int m3lb_set_fd_corners(m3lb_solver* s, int32_t on) {
  if (!s) return fail(M3LB_EINVAL, "null handle");
  s->fd_corners = on ? 1 : 0;
  return M3LB_OK;
}

int m3lb_set_force_field(m3lb_solver* s, const double* fx, const double* fy, const double* fz) {
  if (!s || !fx || !fy || !fz) return fail(M3LB_EINVAL, "null argument");
  if (s->cfg.forcing != M3LB_FORCE_FIELD || !s->ff_d)
    return fail(M3LB_ESTATE, "the handle was not created with M3LB_FORCE_FIELD");
  const size_t np = size_t(s->d.np), st = size_t(s->stride);
  const double* src[3] = {fx, fy, fz};
  HIP_OK(hipStreamSynchronize(s->stream));  // the previous field may still be read
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

int m3lb_tile_order_map(uint32_t grid, int32_t run, uint32_t* out) {
  if (!out) return fail(M3LB_EINVAL, "null argument");
  for (uint32_t b = 0; b < grid; ++b) out[b] = m3lb::tile_of(b, grid, run);
  return M3LB_OK;
}

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
// Per-pair scales the current field calls for (see scale_exponent).
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
  HIP_OK(hipStreamSynchronize(s->stream));
  std::vector<double> mx(size_t(s->Q), 0.0);
  want.assign(size_t(s->Q), 1.0);
  for (unsigned b = 0; b < grid; ++b)
    for (int q = 0; q < s->Q; ++q) mx[size_t(q)] = std::max(mx[size_t(q)], h[size_t(b) * s->Q + q]);
  // I16 scales every pair in proportion to its weight, from the largest
  // |g_q| / w_q over all directions: a direction that starts near zero (e.g.
  // along an axis the initial flow has no component on) would otherwise get a
  // scale it saturates at once the flow turns into it.
  double ratio = 0;
  for (int q = 0; q < s->Q; ++q) ratio = std::max(ratio, mx[size_t(q)] / weight_of(s->Q, q));
  for (int q = 0; q < s->Q; ++q) {
    const int o = q == 0 ? 0 : (q & 1 ? q + 1 : q - 1);
    const double m = s->cfg.dtype == M3LB_I16 ? ratio * weight_of(s->Q, q)
                                              : std::max(mx[size_t(q)], mx[size_t(o)]);
    want[size_t(q)] = std::ldexp(1.0, scale_exponent(s->cfg.dtype, s->Q, q, m));
  }
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
