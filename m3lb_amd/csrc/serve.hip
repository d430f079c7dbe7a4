// serve.hip -- continuous-batching server of include/m3lb.h (m3lb_srv_*).
//
// Boxes live in one device arena of the budget's size.  Each chunk of up to K
// steps is prepared on the host (box descriptors + packed tile prefix sums in
// pinned memory), copied to the device and run as K launches of the batched
// step kernel -- one hipGraph launch for a full chunk -- followed by the
// completion summaries of the boxes whose last step is in the chunk.  Two
// chunks are in flight at most: while one runs, the host retires the previous
// one, admits queued requests into freed slots and memory, and issues the
// next.  Every device operation is ordered on the server's one stream, so an
// arena range freed by the host can be handed to a new box at once: the new
// box's first write is queued behind every launch that used the old one.
// The reference has no serving layer.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "launch.hpp"
#include "serve_plan.hpp"
#include "scales.hpp"

using m3lb::BoxDesc;
using m3lb::KernelSet;

namespace {

using Clock = std::chrono::steady_clock;
constexpr int kDepth = 2;          // chunks in flight
constexpr int kMaxChunk = 1024;    // steps per chunk and per autotune measurement
constexpr unsigned kSumGrid = 1024;  // blocks of a completion reduction

int err(int code, const std::string& m) { return m3lb::set_error(code, m.c_str()); }

#define SRV_HIP(expr)                                                               \
  do {                                                                              \
    hipError_t e_ = (expr);                                                         \
    if (e_ != hipSuccess)                                                           \
      return err(M3LB_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Folds per-block partials (grid x 4) into 4 values in a fixed order.
__global__ void __launch_bounds__(256) k_fold4(const double* part, unsigned grid, double* out) {
  __shared__ double lds[4][256];
  double acc[4] = {0, 0, 0, 0};
  for (unsigned b = threadIdx.x; b < grid; b += 256)
    for (int k = 0; k < 4; ++k) acc[k] += part[b * 4 + k];
  for (int k = 0; k < 4; ++k) lds[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (unsigned w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      for (int k = 0; k < 4; ++k) lds[k][threadIdx.x] += lds[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 4) out[threadIdx.x] = lds[threadIdx.x][0];
}

// Flags of a box from a preset: halo Excluded (FluidSolver.hpp:59-67),
// interior Fluid, and with walls = 1 the interior rows y = 0 and y = ny - 1
// Solid (the channel of validation/poiseuille.cpp:86-90, validation/channel3d.cpp).
__global__ void __launch_bounds__(m3lb::kBlock) k_box_flags(uint8_t* flags, m3lb::Dom d, int walls) {
  const int n = blockIdx.x * m3lb::kBlock + threadIdx.x;
  if (n >= d.np) return;
  const int x = n % d.sx - d.hx, y = (n / d.sx) % d.sy - d.hy, z = n / (d.sx * d.sy) - d.hz;
  uint8_t v = M3LB_EXCLUDED;
  if (x >= 0 && x < d.nx && y >= 0 && y < d.ny && z >= 0 && z < d.nz)
    v = (walls && (y == 0 || y == d.ny - 1)) ? M3LB_SOLID : M3LB_FLUID;
  flags[n] = v;
}

struct Box {
  int64_t id = 0;
  m3lb_request req;
  int state = M3LB_REQ_QUEUED;
  m3lb::Dom d;
  uint64_t stride = 0, fbytes = 0, bytes = 0, off = m3lb::Arena::kNone;
  std::vector<uint8_t> flags;  // padded flags when the caller passed a geometry
  int steps_left = 0;          // steps not yet issued
  int steps_done = 0;          // steps of retired chunks
  int parity = 0;              // parity of the next issued step
  int slot = -1;
  double t_arrival = 0, t_admit = -1, t_done = -1;
  double summary[4] = {0, 0, 0, 0};
  float scale[27] = {}, inv_scale[27] = {};  // 16-bit storage: chosen at seeding (scales.hpp)
  BoxDesc desc;
  unsigned tiles = 0;
};

struct Chunk {
  hipEvent_t begin = nullptr, end = nullptr;
  std::vector<int64_t> boxes;      // boxes stepped, for steps_done
  std::vector<int> steps;          // steps of each
  std::vector<int64_t> finishing;  // boxes whose last step is in this chunk
  uint64_t lups = 0;
};

}  // namespace

struct m3lb_server {
  m3lb_server_config cfg;
  const KernelSet* ks = nullptr;
  int Q = 0, device = 0, cus = 0;
  unsigned grid = 0;
  hipStream_t stream = nullptr;
  char* arena_d = nullptr;
  m3lb::Arena arena;
  m3lb::Admission adm;
  std::unordered_map<int64_t, Box> boxes;
  std::vector<int64_t> resident;  // admission order
  std::vector<int> free_slots;
  std::vector<int64_t> completed;  // not yet polled
  uint64_t resident_nodes = 0;
  int64_t next_id = 1;
  // per buffer: descriptors and {boxes, tiles, tile_start[]} (pinned + device)
  BoxDesc* desc_h[kDepth] = {nullptr, nullptr};
  BoxDesc* desc_d[kDepth] = {nullptr, nullptr};
  unsigned* hdr_h[kDepth] = {nullptr, nullptr};
  unsigned* hdr_d[kDepth] = {nullptr, nullptr};
  hipGraph_t graph[kDepth] = {nullptr, nullptr};
  hipGraphExec_t exec[kDepth] = {nullptr, nullptr};
  Chunk chunk[kDepth];
  int64_t issued = 0, retired = 0;
  double* partial_d = nullptr;
  double* summary_d = nullptr;  // max_boxes x 4
  double* summary_h = nullptr;  // pinned mirror
  Clock::time_point t0;
  m3lb_server_stats st;
};

namespace {

double now_s(const m3lb_server* s) {
  return std::chrono::duration<double>(Clock::now() - s->t0).count();
}

int validate(const m3lb_server_config* c) {
  if (!c) return err(M3LB_EINVAL, "config is null");
  if (c->lattice != M3LB_D2Q9 && c->lattice != M3LB_D3Q27)
    return err(M3LB_EINVAL, "lattice must be M3LB_D2Q9 or M3LB_D3Q27");
  if (c->op < M3LB_OP_BGK || c->op > M3LB_OP_TRT_PRODUCT)
    return err(M3LB_EINVAL, "the server runs the collision operators BGK .. CM");
  if (c->storage != M3LB_RAW && c->storage != M3LB_SHIFTED)
    return err(M3LB_EINVAL, "storage must be M3LB_RAW or M3LB_SHIFTED");
  if (c->dtype < M3LB_F64 || c->dtype > M3LB_E2M13) return err(M3LB_EINVAL, "unknown dtype");
  if (m3lb::is_16bit(c->dtype) && c->storage != M3LB_SHIFTED)
    return err(M3LB_EINVAL, "16-bit storage holds shifted populations (g = f - w) only");
  if (c->forcing != M3LB_FORCE_NONE && c->forcing != M3LB_FORCE_GUO)
    return err(M3LB_EINVAL, "forcing must be M3LB_FORCE_NONE or M3LB_FORCE_GUO");
  if (c->max_boxes < 1 || c->max_boxes > 65536) return err(M3LB_EINVAL, "max_boxes must be in 1..65536");
  if (c->steps_per_chunk < 1 || c->steps_per_chunk > kMaxChunk)
    return err(M3LB_EINVAL, "steps_per_chunk must be in 1..1024");
  if (c->blocks_per_cu < 1 || c->blocks_per_cu > 64) return err(M3LB_EINVAL, "blocks_per_cu must be in 1..64");
  if (c->tile_shape < M3LB_SHAPE_ROWS || c->tile_shape > M3LB_SHAPE_AUTO)
    return err(M3LB_EINVAL, "unknown tile_shape");
  if (c->max_bypass < 0) return err(M3LB_EINVAL, "max_bypass must be >= 0");
  if (c->tile_run < M3LB_TILE_BLOCK_RANGES) return err(M3LB_EINVAL, "unknown tile_run");
  if (c->budget_bytes < m3lb::Arena::kAlign) return err(M3LB_EINVAL, "budget_bytes too small");
  return M3LB_OK;
}

// Domain of a request as the single-box solver computes it (Domain.hpp:47-56).
int request_domain(const m3lb_server_config& c, const m3lb_request* r, m3lb::Dom& d) {
  if (!r) return err(M3LB_EINVAL, "request is null");
  if (r->nx < 1 || r->ny < 1 || r->nz < 1) return err(M3LB_EINVAL, "extents must be >= 1");
  if (c.lattice == M3LB_D2Q9 && r->nz != 1) return err(M3LB_EINVAL, "D2Q9 needs nz == 1");
  const int n[3] = {r->nx, r->ny, r->nz};
  int h[3], sdim[3];
  for (int a = 0; a < 3; ++a) {
    h[a] = (n[a] <= 1 || r->periodic[a]) ? 0 : 1;
    sdim[a] = n[a] + 2 * h[a];
  }
  const int64_t np = int64_t(sdim[0]) * sdim[1] * sdim[2];
  if (np >= (int64_t(1) << 31)) return err(M3LB_EINVAL, "n_padded exceeds the int32 node index");
  d.nx = n[0]; d.ny = n[1]; d.nz = n[2];
  d.hx = h[0]; d.hy = h[1]; d.hz = h[2];
  d.sx = sdim[0]; d.sy = sdim[1]; d.sz = sdim[2];
  d.px = r->periodic[0] ? 1 : 0;
  d.py = r->periodic[1] ? 1 : 0;
  d.pz = r->periodic[2] ? 1 : 0;
  d.np = int(np);
  return M3LB_OK;
}

int elem_bytes(const m3lb_server_config& c) {
  return c.dtype == M3LB_F64 ? 8 : (m3lb::is_16bit(c.dtype) ? 2 : 4);
}

void footprint_of(const m3lb_server_config& c, const m3lb::Dom& d, uint64_t& stride,
                  uint64_t& fbytes, uint64_t& total) {
  stride = (uint64_t(d.np) + 63) / 64 * 64;
  fbytes = m3lb::Arena::round(stride * uint64_t(c.lattice) * uint64_t(elem_bytes(c)));
  total = fbytes + m3lb::Arena::round(uint64_t(d.np));
}

int validate_request(const m3lb_server_config& c, const m3lb_request* r) {
  if (r->steps < 1) return err(M3LB_EINVAL, "steps must be >= 1");
  if (r->init != M3LB_INIT_UNIFORM && r->init != M3LB_INIT_TAYLOR_GREEN)
    return err(M3LB_EINVAL, "unknown init");
  if (r->walls != M3LB_WALLS_NONE && r->walls != M3LB_WALLS_Y) return err(M3LB_EINVAL, "unknown walls");
  const double v[] = {r->omega, r->omega_minus, r->omega_bulk, r->force[0], r->force[1],
                      r->force[2], r->rho0, r->u[0], r->u[1], r->u[2]};
  for (double x : v)
    if (!std::isfinite(x)) return err(M3LB_EINVAL, "request parameters must be finite");
  if (c.forcing == M3LB_FORCE_NONE && (r->force[0] != 0 || r->force[1] != 0 || r->force[2] != 0))
    return err(M3LB_EINVAL, "a force needs a server with M3LB_FORCE_GUO");
  return M3LB_OK;
}

int shape_for(const m3lb_server* s, int nx) {
  const int p = s->cfg.tile_shape;
  if (p == M3LB_SHAPE_AUTO) return nx % 64 == 0 ? m3lb::kShapeRows : m3lb::kShapeLinear;
  return p == M3LB_SHAPE_ROWS ? m3lb::kShapeRows : m3lb::kShapeLinear;
}

// Static part of a box's descriptor and its tile count under the current shape.
void describe(const m3lb_server* s, Box& b) {
  BoxDesc& x = b.desc;
  std::memset(&x, 0, sizeof x);
  x.f = s->arena_d + b.off;
  x.stride = b.stride;
  x.flags = reinterpret_cast<const uint8_t*>(s->arena_d + b.off + b.fbytes);
  x.d = b.d;
  x.shape = shape_for(s, b.d.nx);
  x.bx_log2 = b.d.nx > 128 ? 8 : (b.d.nx > 64 ? 7 : 6);
  x.chunks = (b.d.nx + m3lb::kBlock - 1) / m3lb::kBlock;
  x.rows = b.d.ny * b.d.nz;
  x.nodes = unsigned(b.d.nx) * unsigned(b.d.ny) * unsigned(b.d.nz);
  x.div_x = m3lb::make_fastdiv(unsigned(b.d.nx));
  x.div_y = m3lb::make_fastdiv(unsigned(b.d.ny));
  x.div_xy = m3lb::make_fastdiv(unsigned(b.d.nx) * unsigned(b.d.ny));
  x.omega = b.req.omega;
  x.omega_m = b.req.omega_minus;
  x.omega_b = b.req.omega_bulk;
  for (int k = 0; k < 3; ++k) x.F[k] = s->cfg.forcing == M3LB_FORCE_GUO ? b.req.force[k] : 0.0;
  const bool scaled = m3lb::is_16bit(s->cfg.dtype);  // the scales are set at seeding
  for (int q = 0; q < 27; ++q) {
    x.scale[q] = scaled ? b.scale[q] : 1.0f;
    x.inv_scale[q] = scaled ? b.inv_scale[q] : 1.0f;
  }
  if (x.shape == m3lb::kShapeLinear) {
    b.tiles = (x.nodes + m3lb::kBlock - 1) / m3lb::kBlock;
  } else {
    const unsigned rpb = unsigned(m3lb::kBlock >> x.bx_log2);
    b.tiles = x.chunks > 1 ? unsigned(x.rows) * unsigned(x.chunks) : (unsigned(x.rows) + rpb - 1) / rpb;
  }
}

m3lb::ReadArgs read_args(const m3lb_server* s, const Box& b, double* out) {
  m3lb::ReadArgs a;
  std::memset(&a, 0, sizeof a);
  a.f = s->arena_d + b.off;
  a.stride = b.stride;
  a.flags = b.desc.flags;
  a.d = b.d;
  a.parity = b.parity;
  a.out = out;
  a.omega = b.req.omega;
  a.omega_m = b.req.omega_minus;
  a.omega_b = b.req.omega_bulk;
  for (int k = 0; k < 3; ++k) a.F[k] = b.desc.F[k];
  const bool scaled = m3lb::is_16bit(s->cfg.dtype);
  for (int q = 0; q < 27; ++q) a.inv_scale[q] = scaled ? b.inv_scale[q] : 1.0f;
  return a;
}

int launch_ok() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return err(M3LB_EDEVICE, std::string("kernel launch: ") + hipGetErrorString(e));
  return M3LB_OK;
}

// Zeroes the box's populations (as the single-box solver's allocation is),
// writes its flags and seeds it at t = 0.
int seed_box(m3lb_server* s, Box& b) {
  char* base = s->arena_d + b.off;
  SRV_HIP(hipMemsetAsync(base, 0, b.fbytes, s->stream));
  uint8_t* fl = reinterpret_cast<uint8_t*>(base + b.fbytes);
  if (!b.flags.empty()) {
    // pageable source: the copy is complete when the call returns
    SRV_HIP(hipMemcpyAsync(fl, b.flags.data(), b.flags.size(), hipMemcpyHostToDevice, s->stream));
    SRV_HIP(hipStreamSynchronize(s->stream));
    std::vector<uint8_t>().swap(b.flags);
  } else {
    k_box_flags<<<m3lb::blocks_for(b.d.np), m3lb::kBlock, 0, s->stream>>>(fl, b.d, b.req.walls);
  }
  m3lb::SeedArgs a;
  std::memset(&a, 0, sizeof a);
  a.f = base;
  a.stride = b.stride;
  a.flags = fl;
  a.d = b.d;
  a.parity = 0;
  if (b.req.init == M3LB_INIT_TAYLOR_GREEN) {
    a.mode = m3lb::kSeedTaylorGreen;
    a.u0[0] = b.req.u[0];
  } else {
    a.mode = m3lb::kSeedUniform;
    a.rho0 = b.req.rho0;
    for (int k = 0; k < 3; ++k) a.u0[k] = b.req.u[k];
  }
  // 16-bit storage: as the single-box solver does, seed at provisional
  // scales, measure the largest |g_q| of the field, and seed again at the
  // scales that calls for (scales.hpp), so a box gets the solver's bits
  m3lb::provisional_scales(s->cfg.dtype, s->Q, b.scale, b.inv_scale);
  for (int q = 0; q < 27; ++q) a.scale[q] = b.scale[q];
  s->ks->seed(a, s->stream);
  int rc = launch_ok();
  if (rc || !m3lb::is_16bit(s->cfg.dtype)) return rc;
  const unsigned g = std::max(1u, std::min(kSumGrid, m3lb::blocks_for(b.d.np)));
  m3lb::ReadArgs r = read_args(s, b, s->partial_d);
  s->ks->absmax(r, g, s->stream);
  rc = launch_ok();
  if (rc) return rc;
  std::vector<double> h(size_t(g) * size_t(s->Q));
  SRV_HIP(hipMemcpyAsync(h.data(), s->partial_d, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                         s->stream));
  SRV_HIP(hipStreamSynchronize(s->stream));
  std::vector<double> mx(size_t(s->Q), 0.0), want(size_t(s->Q), 1.0);
  for (unsigned k = 0; k < g; ++k)
    for (int q = 0; q < s->Q; ++q) mx[size_t(q)] = std::max(mx[size_t(q)], h[size_t(k) * s->Q + q]);
  m3lb::pair_scales(s->cfg.dtype, s->Q, mx.data(), want.data());
  for (int q = 0; q < s->Q; ++q) {
    b.scale[q] = float(want[size_t(q)]);
    b.inv_scale[q] = float(1.0 / want[size_t(q)]);
    a.scale[q] = b.scale[q];
    b.desc.scale[q] = b.scale[q];
    b.desc.inv_scale[q] = b.inv_scale[q];
  }
  s->ks->seed(a, s->stream);
  return launch_ok();
}

void free_box_memory(m3lb_server* s, Box& b) {
  if (b.off != m3lb::Arena::kNone) s->arena.release(b.off, b.bytes);
  b.off = m3lb::Arena::kNone;
}

void drop_resident(m3lb_server* s, int64_t id) {
  auto it = std::find(s->resident.begin(), s->resident.end(), id);
  if (it == s->resident.end()) return;
  s->resident.erase(it);
  const m3lb::Dom& d = s->boxes.at(id).d;
  s->resident_nodes -= uint64_t(d.nx) * uint64_t(d.ny) * uint64_t(d.nz);
}

// A resident box leaves: its slot and memory are free for the next admission
// and it is reported cancelled.  Launches already issued may still step it;
// they are ordered before anything a later owner of its memory or slot issues.
void cancel_resident(m3lb_server* s, int64_t id) {
  Box& b = s->boxes.at(id);
  drop_resident(s, id);
  s->free_slots.push_back(b.slot);
  b.slot = -1;
  free_box_memory(s, b);
  b.steps_left = 0;
  b.state = M3LB_REQ_CANCELLED;
  b.t_done = now_s(s);
  s->completed.push_back(id);
  ++s->st.cancelled;
}

int admit(m3lb_server* s) {
  if (s->adm.size() == 0 || s->free_slots.empty()) return M3LB_OK;
  int rc = M3LB_OK;
  const uint64_t cap = s->cfg.max_resident_nodes;
  uint64_t nodes = s->resident_nodes;
  // the resident boxes, and those admitted in this pass, as admission sees them
  std::vector<m3lb::Holding> hold;
  uint64_t resident_bytes = 0;
  for (int64_t id : s->resident) {
    const Box& b = s->boxes.at(id);
    hold.push_back({b.steps_left, uint64_t(b.d.nx) * uint64_t(b.d.ny) * uint64_t(b.d.nz), b.bytes});
    resident_bytes += m3lb::Arena::round(b.bytes);
  }
  // arena bytes no resident box holds: kept results, which stepping never frees
  const uint64_t fixed = s->arena.used() - resident_bytes;
  int64_t head = -1, horizon = 0;
  std::vector<int64_t> ids = s->adm.run(
      s->free_slots.size(),
      [&](int64_t id) {
        Box& b = s->boxes.at(id);
        const uint64_t bn = uint64_t(b.d.nx) * uint64_t(b.d.ny) * uint64_t(b.d.nz);
        if (cap && nodes > 0 && nodes + bn > cap) return false;
        const uint64_t off = s->arena.alloc(b.bytes);
        if (off == m3lb::Arena::kNone) return false;
        b.off = off;
        nodes += bn;
        hold.push_back({b.req.steps, bn, b.bytes});
        return true;
      },
      // A request passes the waiting head only if its last step is issued no
      // later than the step after which the head fits: every resident box
      // advances one step per step, so it has left by then and the head is
      // admitted no later than without it.
      [&](int64_t waiting, int64_t id) {
        if (waiting != head) {
          head = waiting;
          const Box& h = s->boxes.at(waiting);
          horizon = m3lb::fit_horizon(hold, fixed, cap, s->arena.capacity(),
                                      uint64_t(h.d.nx) * uint64_t(h.d.ny) * uint64_t(h.d.nz),
                                      h.bytes);
        }
        return int64_t(s->boxes.at(id).req.steps) <= horizon;
      });
  s->resident_nodes = nodes;
  const double t = now_s(s);
  for (int64_t id : ids) {
    Box& b = s->boxes.at(id);
    b.slot = s->free_slots.back();
    s->free_slots.pop_back();
    b.state = M3LB_REQ_RESIDENT;
    b.t_admit = t;
    b.steps_left = b.req.steps;
    b.parity = 0;
    describe(s, b);
    s->resident.push_back(id);
  }
  // A box whose seeding fails, and every box admitted after it, is cancelled:
  // its slot and memory go back, and nothing steps it.
  rc = m3lb::seed_admitted(
      ids, [&](int64_t id) { return seed_box(s, s->boxes.at(id)); },
      [&](int64_t id) { cancel_resident(s, id); });
  return rc;
}

m3lb::BatchArgs batch_args(const m3lb_server* s, int buf, int j) {
  m3lb::BatchArgs a;
  a.desc = s->desc_d[buf];
  a.hdr = s->hdr_d[buf];
  a.j = j;
  a.run = s->cfg.tile_run;
  // a node cap whose populations fit the last-level cache keeps the whole
  // resident set there: plain accesses then stream faster
  a.cached = s->cfg.max_resident_nodes > 0 &&
             s->cfg.max_resident_nodes * uint64_t(s->Q) * uint64_t(s->ks->elem_bytes) <=
                 m3lb::kCacheResidentBytes;
  return a;
}

void destroy_graphs(m3lb_server* s) {
  for (int k = 0; k < kDepth; ++k) {
    if (s->exec[k]) (void)hipGraphExecDestroy(s->exec[k]);
    if (s->graph[k]) (void)hipGraphDestroy(s->graph[k]);
    s->exec[k] = nullptr;
    s->graph[k] = nullptr;
  }
}

// One graph per buffer: steps_per_chunk launches of the batched kernel.
int capture_graphs(m3lb_server* s) {
  destroy_graphs(s);
  if (!s->cfg.use_graph) return M3LB_OK;
  for (int k = 0; k < kDepth; ++k) {
    SRV_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
    for (int j = 0; j < s->cfg.steps_per_chunk; ++j) s->ks->batched(batch_args(s, k, j), s->grid, s->stream);
    SRV_HIP(hipStreamEndCapture(s->stream, &s->graph[k]));
    SRV_HIP(hipGraphInstantiate(&s->exec[k], s->graph[k], nullptr, nullptr, 0));
  }
  return M3LB_OK;
}

// Completion summary: mass and momentum of the box after its last step.
int summarize(m3lb_server* s, const Box& b) {
  const unsigned g = std::max(1u, std::min(kSumGrid, m3lb::blocks_for(b.d.np)));
  s->ks->reduce_mass(read_args(s, b, s->partial_d), g, s->stream);
  double* out = s->summary_d + size_t(b.slot) * 4;
  k_fold4<<<1, 256, 0, s->stream>>>(s->partial_d, g, out);
  int rc = launch_ok();
  if (rc) return rc;
  SRV_HIP(hipMemcpyAsync(s->summary_h + size_t(b.slot) * 4, out, 4 * sizeof(double),
                         hipMemcpyDeviceToHost, s->stream));
  return M3LB_OK;
}

// Issues the next chunk; *issued = false when no resident box has steps left.
int issue(m3lb_server* s, bool* issued) {
  *issued = false;
  const int buf = int(s->issued % kDepth);
  Chunk& c = s->chunk[buf];
  c.boxes.clear();
  c.steps.clear();
  c.finishing.clear();
  c.lups = 0;
  const int K = s->cfg.steps_per_chunk;
  int jmax = 0;
  std::vector<unsigned> tiles;
  for (int64_t id : s->resident) {
    Box& b = s->boxes.at(id);
    if (b.steps_left <= 0) continue;
    const int k = std::min(K, b.steps_left);
    c.boxes.push_back(id);
    c.steps.push_back(k);
    tiles.push_back(b.tiles);
  }
  const size_t n = c.boxes.size();
  if (n == 0) return M3LB_OK;
  // boxes that stop within the chunk go last, so that launch j walks only
  // the tiles of the boxes that step at j
  std::vector<size_t> order;
  m3lb::pack_chunk(tiles, c.steps, K, order, s->hdr_h[buf]);
  for (size_t k = 0; k < n; ++k) {
    Box& b = s->boxes.at(c.boxes[order[k]]);
    const int steps = c.steps[order[k]];
    BoxDesc& x = s->desc_h[buf][k];
    x = b.desc;
    x.parity0 = b.parity;
    x.steps = steps;
    b.steps_left -= steps;
    b.parity = (b.parity + steps) & 1;
    c.lups += uint64_t(b.desc.nodes) * uint64_t(steps);
    jmax = std::max(jmax, steps);
    if (b.steps_left == 0) c.finishing.push_back(b.id);
  }
  SRV_HIP(hipEventRecord(c.begin, s->stream));
  SRV_HIP(hipMemcpyAsync(s->desc_d[buf], s->desc_h[buf], n * sizeof(BoxDesc),
                         hipMemcpyHostToDevice, s->stream));
  SRV_HIP(hipMemcpyAsync(s->hdr_d[buf], s->hdr_h[buf], m3lb::pack_words(n, K) * sizeof(unsigned),
                         hipMemcpyHostToDevice, s->stream));
  if (s->exec[buf] && jmax == K) {
    SRV_HIP(hipGraphLaunch(s->exec[buf], s->stream));
  } else {
    for (int j = 0; j < jmax; ++j) s->ks->batched(batch_args(s, buf, j), s->grid, s->stream);
    int rc = launch_ok();
    if (rc) return rc;
  }
  for (int64_t id : c.finishing) {
    int rc = summarize(s, s->boxes.at(id));
    if (rc) return rc;
  }
  SRV_HIP(hipEventRecord(c.end, s->stream));
  ++s->issued;
  ++s->st.chunks;
  s->st.steps += jmax;
  *issued = true;
  return M3LB_OK;
}


// Retires the oldest chunk if it has finished (or wait: once it finishes).
int retire_one(m3lb_server* s, bool wait, bool* done) {
  *done = false;
  if (s->retired == s->issued) return M3LB_OK;
  Chunk& c = s->chunk[int(s->retired % kDepth)];
  if (wait) {
    SRV_HIP(hipEventSynchronize(c.end));
  } else {
    const hipError_t e = hipEventQuery(c.end);
    if (e == hipErrorNotReady) return M3LB_OK;
    if (e != hipSuccess) return err(M3LB_EDEVICE, std::string("chunk: ") + hipGetErrorString(e));
  }
  float ms = 0;
  SRV_HIP(hipEventElapsedTime(&ms, c.begin, c.end));
  s->st.device_seconds += double(ms) * 1e-3;
  s->st.lattice_updates += int64_t(c.lups);
  const double t = now_s(s);
  for (size_t k = 0; k < c.boxes.size(); ++k) {
    auto it = s->boxes.find(c.boxes[k]);
    if (it != s->boxes.end() && it->second.state == M3LB_REQ_RESIDENT) it->second.steps_done += c.steps[k];
  }
  for (int64_t id : c.finishing) {
    auto it = s->boxes.find(id);
    if (it == s->boxes.end() || it->second.state != M3LB_REQ_RESIDENT) continue;  // cancelled
    Box& b = it->second;
    b.state = M3LB_REQ_DONE;
    b.t_done = t;
    std::memcpy(b.summary, s->summary_h + size_t(b.slot) * 4, sizeof b.summary);
    s->free_slots.push_back(b.slot);
    b.slot = -1;
    drop_resident(s, id);
    if (!b.req.keep) free_box_memory(s, b);
    s->completed.push_back(id);
    ++s->st.done;
  }
  ++s->retired;
  *done = true;
  return M3LB_OK;
}

int fill_and_issue(m3lb_server* s) {
  while (s->issued - s->retired < kDepth) {
    int rc = admit(s);
    if (rc) return rc;
    bool issued = false;
    rc = issue(s, &issued);
    if (rc) return rc;
    if (!issued) break;
  }
  return M3LB_OK;
}

void fill_result(const m3lb_server* s, const Box& b, m3lb_result* r) {
  r->id = b.id;
  r->state = b.state;
  r->steps = b.steps_done;
  r->arrival = b.t_arrival;
  r->admitted = b.t_admit;
  r->completed = b.t_done;
  r->mass = b.summary[0];
  for (int k = 0; k < 3; ++k) r->momentum[k] = b.summary[1 + k];
  (void)s;
}

int find_box(m3lb_server* s, int64_t id, Box** out) {
  auto it = s->boxes.find(id);
  if (it == s->boxes.end()) return err(M3LB_EINVAL, "unknown request id " + std::to_string(id));
  *out = &it->second;
  return M3LB_OK;
}

int kept_done(m3lb_server* s, int64_t id, Box** out) {
  int rc = find_box(s, id, out);
  if (rc) return rc;
  if ((*out)->state != M3LB_REQ_DONE || !(*out)->req.keep)
    return err(M3LB_ESTATE, "fields are kept only for completed requests submitted with keep = 1");
  return M3LB_OK;
}

bool idle(const m3lb_server* s) {
  return s->adm.size() == 0 && s->resident.empty() && s->issued == s->retired;
}

}  // namespace

extern "C" {

int m3lb_srv_footprint(const m3lb_server_config* cfg, const m3lb_request* req, uint64_t* bytes) {
  int rc = validate(cfg);
  if (rc) return rc;
  if (!bytes) return err(M3LB_EINVAL, "bytes is null");
  m3lb::Dom d;
  rc = request_domain(*cfg, req, d);
  if (rc) return rc;
  uint64_t stride, fbytes;
  footprint_of(*cfg, d, stride, fbytes, *bytes);
  return M3LB_OK;
}

int m3lb_srv_destroy(m3lb_server* s) {
  if (!s) return M3LB_OK;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  destroy_graphs(s);
  for (int k = 0; k < kDepth; ++k) {
    if (s->chunk[k].begin) (void)hipEventDestroy(s->chunk[k].begin);
    if (s->chunk[k].end) (void)hipEventDestroy(s->chunk[k].end);
    if (s->desc_h[k]) (void)hipHostFree(s->desc_h[k]);
    if (s->hdr_h[k]) (void)hipHostFree(s->hdr_h[k]);
    if (s->desc_d[k]) (void)hipFree(s->desc_d[k]);
    if (s->hdr_d[k]) (void)hipFree(s->hdr_d[k]);
  }
  if (s->partial_d) (void)hipFree(s->partial_d);
  if (s->summary_d) (void)hipFree(s->summary_d);
  if (s->summary_h) (void)hipHostFree(s->summary_h);
  if (s->arena_d) (void)hipFree(s->arena_d);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
  return M3LB_OK;
}

int m3lb_srv_create(const m3lb_server_config* cfg, m3lb_server** out) {
  if (!out) return err(M3LB_EINVAL, "out is null");
  *out = nullptr;
  int rc = validate(cfg);
  if (rc) return rc;
  const KernelSet* ks = m3lb::kernel_set(cfg->lattice, cfg->op, cfg->storage, cfg->dtype, cfg->forcing);
  if (!ks || !ks->batched) return err(M3LB_EINVAL, "no batched kernel for this configuration");
  m3lb_server* s = new m3lb_server();
  s->cfg = *cfg;
  s->ks = ks;
  s->Q = cfg->lattice;
  s->adm.set_max_bypass(cfg->max_bypass);
  s->t0 = Clock::now();
  std::memset(&s->st, 0, sizeof s->st);
  auto bail = [&](int code) {
    m3lb_srv_destroy(s);
    return code;
  };
  if (hipGetDevice(&s->device) != hipSuccess) return bail(err(M3LB_EDEVICE, "no HIP device"));
  if (hipDeviceGetAttribute(&s->cus, hipDeviceAttributeMultiprocessorCount, s->device) != hipSuccess)
    return bail(err(M3LB_EDEVICE, "device attribute query failed"));
  s->grid = unsigned(s->cus * cfg->blocks_per_cu + 7) / 8 * 8;
  if (hipStreamCreate(&s->stream) != hipSuccess) return bail(err(M3LB_EDEVICE, "hipStreamCreate failed"));
  if (hipMalloc(&s->arena_d, cfg->budget_bytes) != hipSuccess)
    return bail(err(M3LB_ENOMEM, "arena: hipMalloc of " + std::to_string(cfg->budget_bytes) + " bytes failed"));
  s->arena.reset(cfg->budget_bytes);
  const size_t nb = size_t(cfg->max_boxes);
  for (int k = 0; k < kDepth; ++k) {
    if (hipEventCreate(&s->chunk[k].begin) != hipSuccess || hipEventCreate(&s->chunk[k].end) != hipSuccess)
      return bail(err(M3LB_EDEVICE, "hipEventCreate failed"));
    if (hipHostMalloc(reinterpret_cast<void**>(&s->desc_h[k]), nb * sizeof(BoxDesc), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void**>(&s->hdr_h[k]), m3lb::pack_words(nb, kMaxChunk) * sizeof(unsigned), hipHostMallocDefault) != hipSuccess)
      return bail(err(M3LB_ENOMEM, "pinned descriptor buffers"));
    if (hipMalloc(&s->desc_d[k], nb * sizeof(BoxDesc)) != hipSuccess ||
        hipMalloc(&s->hdr_d[k], m3lb::pack_words(nb, kMaxChunk) * sizeof(unsigned)) != hipSuccess)
      return bail(err(M3LB_ENOMEM, "device descriptor buffers"));
  }
  // completion sums (4 per block) and, for 16-bit storage, the largest
  // |g_q| per block at seeding (Q per block)
  if (hipMalloc(&s->partial_d, kSumGrid * 27 * sizeof(double)) != hipSuccess ||
      hipMalloc(&s->summary_d, nb * 4 * sizeof(double)) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&s->summary_h), nb * 4 * sizeof(double), hipHostMallocDefault) != hipSuccess)
    return bail(err(M3LB_ENOMEM, "summary buffers"));
  s->free_slots.reserve(nb);
  for (int k = cfg->max_boxes - 1; k >= 0; --k) s->free_slots.push_back(k);
  s->st.bytes_budget = s->arena.capacity();
  rc = capture_graphs(s);
  if (rc) return bail(rc);
  *out = s;
  return M3LB_OK;
}

int m3lb_srv_submit(m3lb_server* s, const m3lb_request* req, const uint8_t* fl, int64_t* id) {
  if (!s || !req || !id) return err(M3LB_EINVAL, "null argument");
  *id = 0;
  Box b;
  int rc = request_domain(s->cfg, req, b.d);
  if (!rc) rc = validate_request(s->cfg, req);
  if (rc) {
    ++s->st.rejected;
    return rc;
  }
  footprint_of(s->cfg, b.d, b.stride, b.fbytes, b.bytes);
  if (uint64_t(b.d.np) * uint64_t(elem_bytes(s->cfg)) >= (uint64_t(1) << 32)) {
    ++s->st.rejected;
    return err(M3LB_EINVAL, "one direction of the population array exceeds 4 GiB");
  }
  if (b.bytes > s->arena.capacity()) {
    ++s->st.rejected;
    return err(M3LB_EINVAL, "request needs " + std::to_string(b.bytes) + " bytes, more than the budget of " +
                                std::to_string(s->arena.capacity()));
  }
  b.req = *req;
  if (fl) {
    const m3lb::Dom& d = b.d;
    b.flags.assign(size_t(d.np), uint8_t(M3LB_EXCLUDED));
    size_t k = 0;
    for (int z = 0; z < d.nz; ++z)
      for (int y = 0; y < d.ny; ++y)
        for (int x = 0; x < d.nx; ++x, ++k) {
          const uint8_t v = fl[k];
          if (v > M3LB_EXCLUDED) {
            ++s->st.rejected;
            return err(M3LB_EINVAL, "flag values are Fluid(0), Solid(1), Excluded(2)");
          }
          b.flags[size_t(((z + d.hz) * d.sy + (y + d.hy)) * d.sx + x + d.hx)] = v;
        }
  }
  b.id = s->next_id++;
  b.t_arrival = req->arrival >= 0 ? req->arrival : now_s(s);
  s->adm.push(b.id);
  s->boxes.emplace(b.id, std::move(b));
  ++s->st.submitted;
  *id = s->next_id - 1;
  return M3LB_OK;
}

int m3lb_srv_progress(m3lb_server* s, int32_t wait_us, int32_t* n_completed) {
  if (!s) return err(M3LB_EINVAL, "null handle");
  const int64_t before = s->st.done;
  bool done = true;
  while (done) {
    int rc = retire_one(s, false, &done);
    if (rc) return rc;
  }
  int rc = fill_and_issue(s);
  if (rc) return rc;
  if (s->st.done == before && wait_us != 0 && s->issued > s->retired) {
    if (wait_us < 0) {
      rc = retire_one(s, true, &done);
    } else {
      const auto until = Clock::now() + std::chrono::microseconds(wait_us);
      done = false;
      while (!done && Clock::now() < until) {
        rc = retire_one(s, false, &done);
        if (rc) return rc;
        if (!done) std::this_thread::sleep_for(std::chrono::microseconds(20));
      }
    }
    if (rc) return rc;
    while (done) {
      rc = retire_one(s, false, &done);
      if (rc) return rc;
    }
    rc = fill_and_issue(s);
    if (rc) return rc;
  }
  if (n_completed) *n_completed = int32_t(s->st.done - before);
  return M3LB_OK;
}

int m3lb_srv_drain(m3lb_server* s) {
  if (!s) return err(M3LB_EINVAL, "null handle");
  while (!idle(s)) {
    const bool stuck = s->issued == s->retired && s->resident.empty();
    int rc = m3lb_srv_progress(s, -1, nullptr);
    if (rc) return rc;
    if (stuck && s->issued == s->retired && s->resident.empty() && s->adm.size() > 0)
      return err(M3LB_ESTATE, "queued requests cannot be admitted (slots or memory)");
  }
  return M3LB_OK;
}

int m3lb_srv_poll(m3lb_server* s, m3lb_result* out, int32_t max, int32_t* n) {
  if (!s || !n || (max > 0 && !out)) return err(M3LB_EINVAL, "null argument");
  int32_t k = 0;
  size_t used = 0;
  for (; used < s->completed.size() && k < max; ++used) {
    auto it = s->boxes.find(s->completed[used]);
    if (it == s->boxes.end()) continue;
    fill_result(s, it->second, &out[k++]);
    if (!it->second.req.keep || it->second.state == M3LB_REQ_CANCELLED) s->boxes.erase(it);
  }
  s->completed.erase(s->completed.begin(), s->completed.begin() + std::ptrdiff_t(used));
  *n = k;
  return M3LB_OK;
}

int m3lb_srv_result(m3lb_server* s, int64_t id, m3lb_result* out) {
  if (!s || !out) return err(M3LB_EINVAL, "null argument");
  Box* b;
  int rc = find_box(s, id, &b);
  if (rc) return rc;
  fill_result(s, *b, out);
  return M3LB_OK;
}

int m3lb_srv_fetch_macroscopic(m3lb_server* s, int64_t id, double* rho, double* ux, double* uy,
                               double* uz) {
  if (!s) return err(M3LB_EINVAL, "null handle");
  Box* b;
  int rc = kept_done(s, id, &b);
  if (rc) return rc;
  const size_t np = size_t(b->d.np);
  double* buf = nullptr;
  SRV_HIP(hipMalloc(&buf, 4 * np * sizeof(double)));
  m3lb::ReadArgs a = read_args(s, *b, nullptr);
  for (int k = 0; k < 4; ++k) a.out2[k] = buf + k * np;
  s->ks->macro(a, s->stream);
  rc = launch_ok();
  double* dst[4] = {rho, ux, uy, uz};
  for (int k = 0; k < 4 && !rc; ++k)
    if (dst[k] && hipMemcpyAsync(dst[k], buf + k * np, np * sizeof(double), hipMemcpyDeviceToHost, s->stream) != hipSuccess)
      rc = err(M3LB_EDEVICE, "macroscopic download failed");
  const hipError_t e = hipStreamSynchronize(s->stream);
  (void)hipFree(buf);
  if (rc) return rc;
  if (e != hipSuccess) return err(M3LB_EDEVICE, hipGetErrorString(e));
  return M3LB_OK;
}

int m3lb_srv_fetch_populations(m3lb_server* s, int64_t id, double* out) {
  if (!s || !out) return err(M3LB_EINVAL, "null argument");
  Box* b;
  int rc = kept_done(s, id, &b);
  if (rc) return rc;
  const size_t n = size_t(b->d.np) * size_t(s->Q);
  double* buf = nullptr;
  SRV_HIP(hipMalloc(&buf, n * sizeof(double)));
  s->ks->gather(read_args(s, *b, buf), s->stream);
  rc = launch_ok();
  if (!rc && hipMemcpyAsync(out, buf, n * sizeof(double), hipMemcpyDeviceToHost, s->stream) != hipSuccess)
    rc = err(M3LB_EDEVICE, "population download failed");
  const hipError_t e = hipStreamSynchronize(s->stream);
  (void)hipFree(buf);
  if (rc) return rc;
  if (e != hipSuccess) return err(M3LB_EDEVICE, hipGetErrorString(e));
  return M3LB_OK;
}

int m3lb_srv_release(m3lb_server* s, int64_t id) {
  if (!s) return err(M3LB_EINVAL, "null handle");
  Box* b;
  int rc = find_box(s, id, &b);
  if (rc) return rc;
  if (b->state != M3LB_REQ_DONE) return err(M3LB_ESTATE, "only completed requests are released");
  free_box_memory(s, *b);
  auto pos = std::find(s->completed.begin(), s->completed.end(), id);
  if (pos != s->completed.end()) s->completed.erase(pos);
  s->boxes.erase(id);
  return M3LB_OK;
}

int m3lb_srv_cancel(m3lb_server* s, int64_t id) {
  if (!s) return err(M3LB_EINVAL, "null handle");
  Box* b;
  int rc = find_box(s, id, &b);
  if (rc) return rc;
  if (b->state == M3LB_REQ_RESIDENT) {
    cancel_resident(s, id);
    return M3LB_OK;
  }
  if (b->state == M3LB_REQ_QUEUED) {
    s->adm.remove(id);
  } else {
    return err(M3LB_ESTATE, "request already completed or cancelled");
  }
  b->state = M3LB_REQ_CANCELLED;
  b->t_done = now_s(s);
  s->completed.push_back(id);
  ++s->st.cancelled;
  return M3LB_OK;
}

int m3lb_srv_stats(m3lb_server* s, m3lb_server_stats* out) {
  if (!s || !out) return err(M3LB_EINVAL, "null argument");
  s->st.queued = int64_t(s->adm.size());
  s->st.resident = int64_t(s->resident.size());
  s->st.bytes_used = s->arena.used();
  s->st.bytes_peak = s->arena.peak();
  s->st.bytes_budget = s->arena.capacity();
  s->st.grid = int32_t(s->grid);
  s->st.tile_run = s->cfg.tile_run;
  s->st.tile_shape = s->cfg.tile_shape;
  s->st.steps_per_chunk = s->cfg.steps_per_chunk;
  s->st.resident_nodes = s->resident_nodes;
  *out = s->st;
  return M3LB_OK;
}

int m3lb_srv_clock(m3lb_server* s, double* seconds) {
  if (!s || !seconds) return err(M3LB_EINVAL, "null argument");
  *seconds = now_s(s);
  return M3LB_OK;
}

int m3lb_srv_autotune(m3lb_server* s, int32_t steps, m3lb_tune_result* out) {
  if (!s || !out) return err(M3LB_EINVAL, "null argument");
  if (steps < 1 || steps > kMaxChunk) return err(M3LB_EINVAL, "steps must be in 1..1024");
  if (!idle(s)) return err(M3LB_ESTATE, "autotune needs an idle server");
  // The mix: Taylor-Green boxes, edges log-uniform on [32, 192] from a fixed
  // sequence, up to max_boxes / 256 boxes, half the budget and 8 GiB, and
  // within the node cap (the resident set the server will run), so that the
  // tuned regime -- cache-resident or not -- is the one served.
  const uint64_t cap = std::min<uint64_t>(s->arena.capacity() / 2, uint64_t(8) << 30);
  const uint64_t node_cap = s->cfg.max_resident_nodes;
  std::vector<Box> mix;
  uint64_t lcg = 0x9E3779B97F4A7C15ull, used = 0, used_nodes = 0;
  auto uni = [&]() {
    lcg = lcg * 6364136223846793005ull + 1442695040888963407ull;
    return double(lcg >> 11) * (1.0 / 9007199254740992.0);
  };
  const bool is3d = s->Q == M3LB_D3Q27;
  while (int(mix.size()) < std::min(s->cfg.max_boxes, 256)) {
    m3lb_request r;
    std::memset(&r, 0, sizeof r);
    r.nx = int(std::floor(32.0 * std::pow(6.0, uni())));
    r.ny = int(std::floor(32.0 * std::pow(6.0, uni())));
    r.nz = is3d ? int(std::floor(32.0 * std::pow(6.0, uni()))) : 1;
    r.periodic[0] = r.periodic[1] = r.periodic[2] = 1;
    r.init = M3LB_INIT_TAYLOR_GREEN;
    r.steps = 1;
    r.omega = r.omega_minus = 1.6;
    r.omega_bulk = -1.0;
    r.u[0] = 0.04;
    Box b;
    if (request_domain(s->cfg, &r, b.d) != M3LB_OK) break;
    footprint_of(s->cfg, b.d, b.stride, b.fbytes, b.bytes);
    if (used + b.bytes > cap) break;
    const uint64_t bn = uint64_t(b.d.nx) * uint64_t(b.d.ny) * uint64_t(b.d.nz);
    if (node_cap && !mix.empty() && used_nodes + bn > node_cap) break;
    used_nodes += bn;
    b.off = s->arena.alloc(b.bytes);
    if (b.off == m3lb::Arena::kNone) break;
    used += b.bytes;
    b.req = r;
    mix.push_back(std::move(b));
  }
  if (mix.empty()) return err(M3LB_ESTATE, "the budget holds no tuning box");
  uint64_t nodes = 0;
  for (const Box& b : mix) nodes += uint64_t(b.d.nx) * uint64_t(b.d.ny) * uint64_t(b.d.nz);

  const m3lb_server_config saved = s->cfg;
  int rc = M3LB_OK;
  for (Box& b : mix) {
    describe(s, b);
    if (!rc) rc = seed_box(s, b);
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
    rc = err(M3LB_EDEVICE, "hipEventCreate failed");

  // `steps` steps of the whole mix as one chunk; returns GLUPS of the second
  // of two runs (the first warms up).
  auto measure = [&](double* glups) -> int {
    s->grid = unsigned(s->cus * s->cfg.blocks_per_cu + 7) / 8 * 8;
    float ms = 0;
    for (int rep = 0; rep < 2; ++rep) {
      const size_t n = mix.size();
      std::vector<unsigned> tiles;
      for (Box& b : mix) {
        describe(s, b);
        tiles.push_back(b.tiles);
      }
      std::vector<size_t> order;
      m3lb::pack_chunk(tiles, std::vector<int>(n, steps), steps, order, s->hdr_h[0]);
      for (size_t k = 0; k < n; ++k) {
        Box& b = mix[order[k]];
        BoxDesc& x = s->desc_h[0][k];
        x = b.desc;
        x.parity0 = b.parity;
        x.steps = steps;
        b.parity = (b.parity + steps) & 1;
      }
      SRV_HIP(hipMemcpyAsync(s->desc_d[0], s->desc_h[0], n * sizeof(BoxDesc), hipMemcpyHostToDevice, s->stream));
      SRV_HIP(hipMemcpyAsync(s->hdr_d[0], s->hdr_h[0], m3lb::pack_words(n, steps) * sizeof(unsigned),
                             hipMemcpyHostToDevice, s->stream));
      SRV_HIP(hipEventRecord(e0, s->stream));
      for (int j = 0; j < steps; ++j) s->ks->batched(batch_args(s, 0, j), s->grid, s->stream);
      int lrc = launch_ok();
      if (lrc) return lrc;
      SRV_HIP(hipEventRecord(e1, s->stream));
      SRV_HIP(hipEventSynchronize(e1));
      SRV_HIP(hipEventElapsedTime(&ms, e0, e1));
    }
    *glups = ms > 0 ? double(nodes) * double(steps) / (double(ms) * 1e-3) * 1e-9 : 0.0;
    return M3LB_OK;
  };

  double initial = 0, best = -1;
  m3lb_server_config best_cfg = saved;
  int candidates = 0;
  if (!rc) rc = measure(&initial);
  static const int shapes[] = {M3LB_SHAPE_ROWS, M3LB_SHAPE_LINEAR, M3LB_SHAPE_AUTO};
  static const int runs[] = {0, 8, 64, M3LB_TILE_BLOCK_RANGES};
  static const int bpcu[] = {4, 8, 16};
  for (int sh : shapes)
    for (int run : runs)
      for (int bp : bpcu) {
        if (rc) break;
        s->cfg.tile_shape = sh;
        s->cfg.tile_run = run;
        s->cfg.blocks_per_cu = bp;
        double g = 0;
        rc = measure(&g);
        ++candidates;
        if (!rc && g > best) {
          best = g;
          best_cfg = s->cfg;
        }
      }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(s->stream);
  for (Box& b : mix) s->arena.release(b.off, b.bytes);
  s->cfg = rc ? saved : best_cfg;
  s->grid = unsigned(s->cus * s->cfg.blocks_per_cu + 7) / 8 * 8;
  const int grc = capture_graphs(s);
  if (rc) return rc;
  if (grc) return grc;
  out->tile_shape = s->cfg.tile_shape;
  out->tile_run = s->cfg.tile_run;
  out->blocks_per_cu = s->cfg.blocks_per_cu;
  out->candidates = candidates;
  out->best_glups = best;
  out->initial_glups = initial;
  return M3LB_OK;
}

}  // extern "C"
