// kernels.hpp -- gfx950 kernels of the esoteric-pull step and its diagnostics.
//
// Population layout in HBM: one array, SoA, A[q][n] at base + q*stride + n with
// n the reference's padded x-fastest node index (src/grid/Domain.hpp:60-62) and
// stride = n_padded rounded up to 64 so every direction starts 256-B aligned.
// One thread updates one node; a block covers whole x-rows, so the x-shifted
// neighbour accesses of a wavefront share cache lines only with wavefronts of
// the same block.
#pragma once

#include "fastdiv.hpp"
#include "lbm_device.hpp"

namespace m3lb {

constexpr int kBlock = 256;

// Domain as the kernels see it: interior extents, halos, padded extents.
struct Dom {
  int nx, ny, nz, hx, hy, hz, sx, sy, sz;
  int px, py, pz;  // periodic flags
  int np;          // n_padded
};

// A node that holds a fluid state: Fluid, a regularised-wall node and an
// on-node specular node (FluidSolver.hpp:557, 575, 1054, 1215).
__host__ __device__ __forceinline__ bool fluid_like(uint8_t fl) {
  return fl == M3LB_FLUID || fl == M3LB_REGWALL || fl == M3LB_SPECNODE;
}

// ---------------------------------------------------------------------------
// Specular reflection (src/boundary/Specular.hpp).  The mirror of direction i
// about a set of axes is the direction with those components of c_i negated
// (MirrorTable<L>, :104-121); every index is a compile-time constant.
// ---------------------------------------------------------------------------
template <int Q, int i, int fx, int fy, int fz>
__host__ __device__ constexpr int mirror_of() {
  return dir_of<Q>((fx ? -cvel<Q>(i, 0) : cvel<Q>(i, 0)) + 1,
                   (fy ? -cvel<Q>(i, 1) : cvel<Q>(i, 1)) + 1,
                   (fz ? -cvel<Q>(i, 2) : cvel<Q>(i, 2)) + 1);
}

// face_sign (Specular.hpp:253-258): outward sign of a face mask on axis a
__device__ __forceinline__ int face_sign(uint8_t faces, int a) {
  return (faces & (1u << (2 * a))) ? 1 : ((faces & (1u << (2 * a + 1))) ? -1 : 0);
}

// mirror_unknowns_faces (Specular.hpp:274-289): each direction that points
// into the domain on a masked axis takes the value of its image through the
// composition of those axis mirrors.
template <class K>
__device__ __forceinline__ void mirror_unknowns_faces(typename K::R* f, uint8_t faces) {
  using R = typename K::R;
  constexpr int Q = K::Q;
  const int sx = face_sign(faces, 0), sy = face_sign(faces, 1), sz = face_sign(faces, 2);
  R h[Q];
  sfor<Q>([&](auto I) { h[int(I)] = f[int(I)]; });
  sfor<Q>([&](auto I) {
    constexpr int i = I;
    constexpr int cx = cvel<Q>(i, 0), cy = cvel<Q>(i, 1), cz = cvel<Q>(i, 2);
    const bool fx = cx * sx < 0, fy = cy * sy < 0, fz = cz * sz < 0;
    f[i] = fx ? (fy ? (fz ? h[mirror_of<Q, i, 1, 1, 1>()] : h[mirror_of<Q, i, 1, 1, 0>()])
                    : (fz ? h[mirror_of<Q, i, 1, 0, 1>()] : h[mirror_of<Q, i, 1, 0, 0>()]))
              : (fy ? (fz ? h[mirror_of<Q, i, 0, 1, 1>()] : h[mirror_of<Q, i, 0, 1, 0>()])
                    : (fz ? h[mirror_of<Q, i, 0, 0, 1>()] : h[i]));
  });
}

struct StepArgs {
  void* f;
  uint64_t stride;
  const uint8_t* flags;  // used when the geometry template flag is set
  Dom d;
  int bx_log2;     // threads per x-chunk = 1 << bx_log2 (64, 128 or 256)
  int chunks;      // x-chunks per row (1 unless nx > 256)
  int rows;        // ny * nz
  int remap;       // tile order, see tile_of()
  int cached;      // 1: plain accesses everywhere (a cache-resident working set)
  unsigned lds_cap;  // bytes of LDS a capped step launch reserves (0: none)
  FastDiv div_ny;  // row -> (y, z) without an integer division
  double omega, omega_m, omega_b;
  double F[3];
  void* macro[4];  // rho, ux, uy, uz (compute type), or null
  const void* ff;  // per-node force, 3 planes of `stride` (compute type), FIELD only
  float scale[27], inv_scale[27];
};

// The force a node sees: the uniform one, or for a per-node force policy the
// node's own F from three planes of `stride` values.
template <class K>
__device__ __forceinline__ void force_at(Params<typename K::R>& p, const void* ff,
                                         uint64_t stride, unsigned n) {
  if constexpr (K::FIELD) {
    const typename K::R* F = static_cast<const typename K::R*>(ff);
    p.F[0] = F[n];
    p.F[1] = F[stride + n];
    p.F[2] = F[2 * stride + n];
  }
}

template <class K>
__device__ __forceinline__ Params<typename K::R> params_of(const StepArgs& a) {
  using R = typename K::R;
  Params<R> p;
  p.omega = R(a.omega); p.omega_m = R(a.omega_m); p.omega_b = R(a.omega_b);
  p.F[0] = R(a.F[0]); p.F[1] = R(a.F[1]); p.F[2] = R(a.F[2]);
  return p;
}

// Neighbour offsets of an INTERIOR node along each axis (-1, 0, +1 -> index
// delta), with periodic wrap inside the interior; non-periodic axes have a halo
// so interior +- 1 is always inside the padded box (Domain.hpp:104-123).
struct Offs {
  int xm, xp, ym, yp, zm, zp;
};

__device__ __forceinline__ Offs offsets(const Dom& d, int x, int y, int z) {
  Offs o;
  const int sxy = d.sx * d.sy;
  o.xm = (d.px && x == 0) ? d.nx - 1 : -1;
  o.xp = (d.px && x == d.nx - 1) ? -(d.nx - 1) : 1;
  o.ym = (d.py && y == 0) ? (d.ny - 1) * d.sx : -d.sx;
  o.yp = (d.py && y == d.ny - 1) ? -(d.ny - 1) * d.sx : d.sx;
  o.zm = (d.pz && z == 0) ? (d.nz - 1) * sxy : -sxy;
  o.zp = (d.pz && z == d.nz - 1) ? -(d.nz - 1) * sxy : sxy;
  return o;
}

template <int Q, int i>
__device__ __forceinline__ int nb_off(const Offs& o) {
  constexpr int cx = Lat<Q>::cx[i], cy = Lat<Q>::cy[i], cz = Lat<Q>::cz[i];
  int v = 0;
  if constexpr (cx < 0) v += o.xm;
  if constexpr (cx > 0) v += o.xp;
  if constexpr (cy < 0) v += o.ym;
  if constexpr (cy > 0) v += o.yp;
  if constexpr (cz < 0) v += o.zm;
  if constexpr (cz > 0) v += o.zp;
  return v;
}

// Element (q, m) of the population array.  The per-direction base is uniform
// (SGPRs) and the node part a 32-bit byte offset, which maps onto gfx950's
// global_load/store "saddr + voffset" form: one VGPR addresses every local slot
// of a node.  Requires n_padded * sizeof(T) < 2^32 (checked at creation).
template <class T>
__device__ __forceinline__ const T* qbase(const T* A, uint64_t stride, int q) {
  return A + uint64_t(q) * stride;
}
// Population accesses.  The slots a node owns at its own index ("local",
// aligned with the wavefront) are read and written nontemporally when the
// array is far larger than the caches (the step kernels' NTL template flag,
// chosen at run time from the working set); the slots at n + c_i ("remote",
// shifted by one element along x so that neighbouring wavefronts share their
// boundary cache lines) always use plain accesses: nontemporal remote
// accesses lose up to 15% because a boundary line then reaches HBM once per
// wavefront that touches it.  A working set that fits the 256 MB of last-level
// cache streams about 25% faster with plain accesses everywhere (measured
// D3Q27 CM FP32: 64^3..128^3 at 28-30 against 24 GLUPS).  M3LB_NT sets the
// default the other kernels use: 0 plain everywhere, 1 nontemporal local.
#ifndef M3LB_NT
#define M3LB_NT 1
#endif
template <class T>
struct Word {
  using type = T;
};
template <>
struct Word<__bf16> {
  using type = uint16_t;  // the nontemporal builtins take integer/float words
};

template <bool NT, class T>
__device__ __forceinline__ T ld(const T* base, unsigned node) {
  using W = typename Word<T>::type;
  const W* p = reinterpret_cast<const W*>(reinterpret_cast<const char*>(base) +
                                          unsigned(node * unsigned(sizeof(T))));
  W w;
  if constexpr (NT) w = __builtin_nontemporal_load(p);
  else w = *p;
  return __builtin_bit_cast(T, w);
}
template <bool NT, class T>
__device__ __forceinline__ void st(T* base, unsigned node, T v) {
  using W = typename Word<T>::type;
  W* p = reinterpret_cast<W*>(reinterpret_cast<char*>(base) + unsigned(node * unsigned(sizeof(T))));
  if constexpr (NT) __builtin_nontemporal_store(__builtin_bit_cast(W, v), p);
  else *p = __builtin_bit_cast(W, v);
}
constexpr bool kNtLocal = M3LB_NT >= 1;
constexpr bool kNtRemote = false;
// Population arrays up to this size run the step kernels with plain
// accesses (the MI355X last-level cache is 256 MB).
constexpr uint64_t kCacheResidentBytes = 256ull << 20;

// load_rest / load_pair of EsotericPull::Access<P> (EsotericPull.hpp:87-97)
template <class K, int P, bool NTL = kNtLocal>
__device__ __forceinline__ void load_node(const typename K::T* __restrict__ A, uint64_t stride,
                                          unsigned n, const Offs& o, const float* inv,
                                          typename K::R* f) {
  using C = typename K::C;
  constexpr int Q = K::Q;
  f[0] = C::dec(ld<NTL>(qbase(A, stride, 0), n), inv[0]);
  sfor<(Q - 1) / 2>([&](auto Pr) {
    constexpr int i = 2 * Pr + 1;
    const unsigned nb = unsigned(int(n) + nb_off<Q, i>(o));
    if constexpr (P == 0) {
      f[i] = C::dec(ld<NTL>(qbase(A, stride, i), n), inv[i]);
      f[i + 1] = C::dec(ld<kNtRemote>(qbase(A, stride, i + 1), nb), inv[i + 1]);
    } else {
      f[i] = C::dec(ld<NTL>(qbase(A, stride, i + 1), n), inv[i + 1]);
      f[i + 1] = C::dec(ld<kNtRemote>(qbase(A, stride, i), nb), inv[i]);
    }
  });
}

// store_rest / store_pair (EsotericPull.hpp:92-98): writes crossed into the
// very slots the load read.
template <class K, int P, bool NTL = kNtLocal>
__device__ __forceinline__ void store_node(typename K::T* __restrict__ A, uint64_t stride,
                                           unsigned n, const Offs& o, const float* s,
                                           const typename K::R* f) {
  using C = typename K::C;
  using T = typename K::T;
  constexpr int Q = K::Q;
  auto wb = [&](int q) { return const_cast<T*>(qbase<T>(A, stride, q)); };
  st<NTL>(wb(0), n, C::enc(f[0], s[0]));
  sfor<(Q - 1) / 2>([&](auto Pr) {
    constexpr int i = 2 * Pr + 1;
    const unsigned nb = unsigned(int(n) + nb_off<Q, i>(o));
    if constexpr (P == 0) {
      st<kNtRemote>(wb(i + 1), nb, C::enc(f[i], s[i + 1]));
      st<NTL>(wb(i), n, C::enc(f[i + 1], s[i]));
    } else {
      st<kNtRemote>(wb(i), nb, C::enc(f[i], s[i]));
      st<NTL>(wb(i + 1), n, C::enc(f[i + 1], s[i + 1]));
    }
  });
}

// Blocks are dispatched round-robin over the 8 XCDs (block b runs on XCD
// b % 8).  The tile a block computes is chosen so that runs of consecutive
// tiles -- the x-chunks of one row, then the next rows -- land on one XCD and
// share its L2; the cache lines at chunk boundaries, which two tiles each
// write in part, are then merged in one L2 instead of reaching HBM twice.
//   run == 0 : tile = b (hardware order)
//   run  < 0 : XCD k takes the k-th contiguous eighth of the tiles
//   run  > 0 : XCD k takes every 8th run of `run` consecutive tiles; the
//              m mod run tiles left per XCD after whole rounds are dealt as
//              one contiguous block each
// A permutation of [0, G) for every G and run (tested through
// m3lb_tile_order_map).
__host__ __device__ __forceinline__ unsigned tile_of(unsigned b, unsigned G, int run) {
  if (run == 0) return b;
  const unsigned full = G & ~7u;
  if (b >= full) return b;
  const unsigned xcd = b & 7u, k = b >> 3, m = full >> 3;
  if (run < 0) return xcd * m + k;
  const unsigned r = unsigned(run);
  if ((r & (r - 1u)) == 0u) {
    // a power of two (every order the library picks): shifts, which stay in
    // scalar instructions, where a division by a run-time value would take
    // a vector reciprocal sequence per wave
    const unsigned sh = unsigned(__builtin_ctz(r));
    const unsigned whole = (m >> sh) << sh;
    if (k < whole) return (((k >> sh) * 8u + xcd) << sh) + (k & (r - 1u));
    return whole * 8u + xcd * (m - whole) + (k - whole);
  }
  const unsigned whole = (m / r) * r;  // tiles per XCD dealt in whole rounds
  if (k < whole) return ((k / r) * 8u + xcd) * r + (k % r);
  return whole * 8u + xcd * (m - whole) + (k - whole);
}

// Row and x of the node a thread updates in tile `tile` of a box: one
// 256-wide x-chunk of a row when chunks > 1, else 256/bx whole rows (bx >= nx).
__device__ __forceinline__ void tile_node(unsigned tile, int chunks, int bx_log2, int& row,
                                          int& x) {
  if (chunks > 1) {
    const unsigned c = unsigned(chunks);
    // a power of two (every box edge up to 1024 that is a multiple of 256,
    // e.g. 512: two chunks) needs a shift, not a run-time division
    row = (c & (c - 1u)) == 0u ? int(tile >> unsigned(__builtin_ctz(c))) : int(tile / c);
    x = int(tile - unsigned(row) * c) * kBlock + int(threadIdx.x);
  } else {
    row = int(tile) * (kBlock >> bx_log2) + int(threadIdx.x >> bx_log2);
    x = int(threadIdx.x & ((1u << bx_log2) - 1u));
  }
}

// The per-node body of run_step<P, StoreMacro> (FluidSolver.hpp:712-1038),
// Fluid and Solid paths: solid and excluded nodes do nothing (implicit
// bounce-back, :759-761); a fluid node loads its incoming set, collides and
// stores it back crossed.  Shared by the single-box and the batched kernels,
// so a box computes bit-identical results in either.
template <class K, int P, bool GEOM, bool NTL>
__device__ __forceinline__ void node_update(typename K::T* __restrict__ A, uint64_t stride,
                                            const uint8_t* flags, const Dom& d, int x, int y,
                                            int z, const Params<typename K::R>& p0,
                                            const float* scale, const float* inv_scale,
                                            void* const* macro, const void* ff) {
  using R = typename K::R;
  const unsigned n = unsigned(((z + d.hz) * d.sy + (y + d.hy)) * d.sx + d.hx + x);
  if constexpr (GEOM) {
    if (flags[n] != M3LB_FLUID) return;
  }
  Params<R> p = p0;
  force_at<K>(p, ff, stride, n);
  const Offs o = offsets(d, x, y, z);
  R f[K::Q];
  uint64_t stride_st = stride;
  if constexpr (K::C::scaled) {
    // 16-bit storage: the step is bound by arithmetic, and the scalar
    // registers of 27 direction bases and 54 scales do not fit at once (the
    // compiler spills them through vector lanes, 60-70 extra vector
    // instructions).  The scales are powers of two, so 1/s is 2^-k from
    // the bits of s, and the store bases are recomputed after the collision
    // (the empty asm makes the stride opaque from the point f[0] exists on),
    // so neither stays live across it.
    float inv[K::Q];
    for (int q = 0; q < K::Q; ++q)
      inv[q] = __builtin_bit_cast(float, 0x7F000000u - __builtin_bit_cast(uint32_t, scale[q]));
    load_node<K, P, NTL>(A, stride, n, o, inv, f);
  } else {
    load_node<K, P, NTL>(A, stride, n, o, inv_scale, f);
  }
  const Macro<R> m = collide_fused<K>(f, p);
  uint32_t k1 = 0;
  if constexpr (K::C::scaled) asm volatile("" : "+s"(stride_st), "+s"(k1) : "v"(f[0]));
  store_node<K, P, NTL>(A, stride_st, n, o, scale + k1, f);
  if (macro && macro[0]) {
    static_cast<R*>(macro[0])[n] = density<K>(m);
    static_cast<R*>(macro[1])[n] = m.ux;
    static_cast<R*>(macro[2])[n] = m.uy;
    static_cast<R*>(macro[3])[n] = m.uz;
  }
}

// run_step<P, StoreMacro> over one box: one thread per node, tiles of 256.
// NTL: nontemporal accesses to the local slots (working sets far beyond the
// caches).
template <class K, int P, bool GEOM, bool NTL>
__global__ void __launch_bounds__(kBlock) k_step(StepArgs a) {
  const Dom& d = a.d;
  const unsigned tile = tile_of(blockIdx.x, gridDim.x, a.remap);
  const Params<typename K::R> p = params_of<K>(a);
  if (a.chunks > 1) {
    // rows of more than one 256-node chunk: the whole block is on one row,
    // so row, y, z and the row part of every address stay scalar (a copy of
    // the node body that saves about 30 vector instructions per node; the
    // rows-per-block path below has them per lane)
    const unsigned c = unsigned(a.chunks);
    const unsigned sh = unsigned(__builtin_ctz(c));
    const int row = (c & (c - 1u)) == 0u ? int(tile >> sh) : int(tile / c);
    const int x = int(tile - unsigned(row) * c) * kBlock + int(threadIdx.x);
    if (row >= a.rows || x >= d.nx) return;
    const int z = int(fd_div(a.div_ny, unsigned(row)));
    node_update<K, P, GEOM, NTL>(static_cast<typename K::T*>(a.f), a.stride, a.flags, d, x,
                            row - z * d.ny, z, p, a.scale, a.inv_scale, a.macro, a.ff);
    return;
  }
  int row, x;
  tile_node(tile, a.chunks, a.bx_log2, row, x);
  if (row >= a.rows || x >= d.nx) return;
  const int z = int(fd_div(a.div_ny, unsigned(row)));
  node_update<K, P, GEOM, NTL>(static_cast<typename K::T*>(a.f), a.stride, a.flags, d, x,
                          row - z * d.ny, z, p, a.scale, a.inv_scale, a.macro, a.ff);
}

// ---------------------------------------------------------------------------
// Batched step: every resident box of a server advances in one launch.  The
// tiles of all boxes form one packed list (tile_start = prefix sum of the
// boxes' tile counts); a persistent grid of G blocks (G a multiple of 8) walks
// it in virtual-tile order v = block + k*G, mapped through tile_of so that
// runs of consecutive tiles stay on one XCD.  Step j of a chunk of steps is
// launch j; box b takes part while j < steps, at parity (parity0 + j) & 1.
// The boxes are packed in order of descending steps, so launch j walks only
// the prefix of the list that steps at j (the tiles of boxes that have
// stopped are not visited).  A box's step barrier is the kernel boundary.
// Box count and tile totals are read from device memory, so one captured
// sequence of launches serves every chunk.  The per-node work is node_update, as in k_step, so a box gets
// bit-identical results alone, in any mix of boxes, and in the single-box
// solver.
// ---------------------------------------------------------------------------
enum TileShape { kShapeRows = 0, kShapeLinear = 1 };

struct BoxDesc {
  void* f;
  uint64_t stride;
  const uint8_t* flags;
  Dom d;
  int shape;        // kShapeRows: tile_node's rows / x-chunks; kShapeLinear:
                    // 256 consecutive interior nodes in x-fastest order
  int chunks, bx_log2, rows;
  unsigned nodes;   // nx * ny * nz
  FastDiv div_x, div_y, div_xy;
  int parity0, steps;
  double omega, omega_m, omega_b, F[3];
  float scale[27], inv_scale[27];
};

struct BatchArgs {
  const BoxDesc* desc;
  const unsigned* hdr;  // {boxes, total tiles, tile_start[0..boxes], stepping[0..K-1]}:
                        // the boxes of a chunk in order of descending steps, and
                        // stepping[j] = tiles of the boxes that step at j, a
                        // prefix of the packed list (serve_plan.hpp pack_chunk)
  int j;                // step within the chunk
  int run;              // tile order, see tile_of()
  int cached;           // 1: plain accesses everywhere (resident set in cache)
};

// One tile of box b at step j of the chunk.
template <class K, bool NTL>
__device__ __forceinline__ void batched_tile(const BoxDesc& b, unsigned local, int j) {
  using R = typename K::R;
  using T = typename K::T;
  int x, y, z;
  if (b.shape == kShapeLinear) {
    const unsigned i = local * unsigned(kBlock) + threadIdx.x;
    if (i >= b.nodes) return;
    const unsigned zz = fd_div(b.div_xy, i);
    const unsigned r = i - zz * (unsigned(b.d.nx) * unsigned(b.d.ny));
    const unsigned yy = fd_div(b.div_x, r);
    x = int(r - yy * unsigned(b.d.nx));
    y = int(yy);
    z = int(zz);
  } else {
    int row;
    tile_node(local, b.chunks, b.bx_log2, row, x);
    if (row >= b.rows || x >= b.d.nx) return;
    z = int(fd_div(b.div_y, unsigned(row)));
    y = row - z * b.d.ny;
  }
  Params<R> p;
  p.omega = R(b.omega); p.omega_m = R(b.omega_m); p.omega_b = R(b.omega_b);
  p.F[0] = R(b.F[0]); p.F[1] = R(b.F[1]); p.F[2] = R(b.F[2]);
  T* A = static_cast<T*>(b.f);
  // Flags are always checked: a second (flag-free) variant of the node body
  // next to the two parities would take the kernel from 67 to 113 VGPRs.
  if (((b.parity0 + j) & 1) == 0)
    node_update<K, 0, true, NTL>(A, b.stride, b.flags, b.d, x, y, z, p, b.scale, b.inv_scale, nullptr, nullptr);
  else
    node_update<K, 1, true, NTL>(A, b.stride, b.flags, b.d, x, y, z, p, b.scale, b.inv_scale, nullptr, nullptr);
}

// Tile order of the batched kernel: run >= -1 as tile_of (virtual tile
// v = block + k * grid, one box search per tile); kRunBlockRanges: block b
// takes the contiguous range [b * per, (b + 1) * per) of the stepping prefix,
// searches once and walks across box boundaries.
constexpr int kRunBlockRanges = -2;

// RANGES: the kRunBlockRanges order (a separate instance, so that neither
// order's loop raises the other's register count).
// The header and the descriptors do not change during a launch and every
// index into them is uniform over a block, so they are read through the
// constant address space: scalar loads, and the box's pointer, stride and
// scales in scalar registers (read through generic pointers they were vector
// loads, the slots were addressed by 64-bit flat instructions and every
// scale constant took vector instructions).
using ConstBoxDesc = const __attribute__((address_space(4))) BoxDesc;
using ConstWord = const __attribute__((address_space(4))) unsigned;

__device__ __forceinline__ const BoxDesc& box_at(const BoxDesc* desc, unsigned i) {
  return *(const BoxDesc*)((ConstBoxDesc*)desc + i);
}

template <class K, bool RANGES, bool NTL>
__global__ void __launch_bounds__(kBlock) k_step_batched(BatchArgs a) {
  ConstWord* hdr = (ConstWord*)a.hdr;
  const unsigned nbox = hdr[0];
  ConstWord* start = hdr + 2;  // start[nbox] == all tiles of the chunk
  // only the tiles of boxes that step at j: the list's first `total`
  const unsigned total = hdr[3 + nbox + unsigned(a.j)];
  if constexpr (RANGES) {
    const unsigned per = (total + gridDim.x - 1) / gridDim.x;
    unsigned t = blockIdx.x * per;
    const unsigned end = min(total, t + per);
    if (t >= end) return;
    unsigned lo = 0, hi = nbox;
    while (hi - lo > 1) {
      const unsigned mid = (lo + hi) >> 1;
      if (start[mid] <= t) lo = mid;
      else hi = mid;
    }
    while (t < end) {
      while (start[lo + 1] <= t) ++lo;
      const BoxDesc& b = box_at(a.desc, lo);
      const unsigned stop = min(end, start[lo + 1]);
      for (; t < stop; ++t) batched_tile<K, NTL>(b, t - start[lo], a.j);
    }
    return;
  }
  for (unsigned v = blockIdx.x; v < total; v += gridDim.x) {
    const unsigned t = tile_of(v, total, a.run);
    unsigned lo = 0, hi = nbox;  // largest lo with start[lo] <= t
    while (hi - lo > 1) {
      const unsigned mid = (lo + hi) >> 1;
      if (start[mid] <= t) lo = mid;
      else hi = mid;
    }
    batched_tile<K, NTL>(box_at(a.desc, lo), t - start[lo], a.j);
  }
}


// ---------------------------------------------------------------------------
// Seeding.  The reference's scatter<P> (FluidSolver.hpp:1231-1241) walks the
// padded box in ascending node order and writes population i of node n into
// the slot load() reads, with neighbours clamped at non-periodic faces; where
// clamping makes two nodes write one slot the later (higher-indexed) one is
// what remains.  Here every slot is visited once and takes the value of its
// highest-indexed writer, which is that same result with no write conflict.
// ---------------------------------------------------------------------------
enum SeedMode { kSeedUniform = 0, kSeedField = 1, kSeedArray = 2, kSeedTaylorGreen = 3 };

struct SeedArgs {
  void* f;
  uint64_t stride;
  const uint8_t* flags;
  Dom d;
  int parity;
  int mode;
  double rho0, u0[3];
  const double* field[4];  // rho, ux, uy, uz (padded), kSeedField
  const double* arr;       // (n, q) logical populations, kSeedArray
  float scale[27];
};

// Along one axis: the largest coordinate v with shift(v + e) == m, or -1.
__device__ __forceinline__ int writer_axis(int m, int e, int n, int s, int per) {
  if (e == 0) return m;
  if (per) {
    int v = m - e;
    if (v < 0) v += n;
    if (v >= n) v -= n;
    return v;
  }
  int best = -1;
  const int v = m - e;
  if (v >= 0 && v < s) best = v;
  if ((e < 0 && m == 0) || (e > 0 && m == s - 1)) best = best > m ? best : m;
  return best;
}

// Taylor-Green vortex of validation/tgv3d_bench.cpp:79-90 (w = 0, the
// pressure-carrying density) at padded node n, with one wavenumber per axis,
// k_a = 2 pi / n_a, so that boxes with unequal edges stay periodic; for a cube
// it is the reference's field.
__device__ __forceinline__ void taylor_green_at(const Dom& d, int n, double u0, double& rho,
                                                double& ux, double& uy) {
  const int x = n % d.sx, y = (n / d.sx) % d.sy, z = n / (d.sx * d.sy);
  const double pi2 = 6.283185307179586;
  const double X = (pi2 / d.nx) * double(x - d.hx);
  const double Y = (pi2 / d.ny) * double(y - d.hy);
  const double Z = (pi2 / d.nz) * double(z - d.hz);
  const double dr = 3.0 * u0 * u0 / 16.0;
  rho = 1.0 + dr * (cos(2.0 * X) + cos(2.0 * Y)) * (cos(2.0 * Z) + 2.0);
  ux = u0 * sin(X) * cos(Y) * cos(Z);
  uy = -u0 * cos(X) * sin(Y) * cos(Z);
}

template <class K>
__device__ __forceinline__ void seed_value_of(const SeedArgs& a, int n, typename K::R* v) {
  using R = typename K::R;
  const bool fluid = fluid_like(a.flags[n]);
  if (a.mode == kSeedUniform) {
    if (fluid) seed<K>(v, R(a.rho0), R(a.u0[0]), R(a.u0[1]), R(a.u0[2]));
    else seed<K>(v, R(a.rho0), R(0), R(0), R(0));
  } else if (a.mode == kSeedTaylorGreen) {
    if (fluid) {
      double rho, ux, uy;
      taylor_green_at(a.d, n, a.u0[0], rho, ux, uy);
      seed<K>(v, R(rho), R(ux), R(uy), R(0));
    } else {
      seed<K>(v, R(1), R(0), R(0), R(0));
    }
  } else {
    if (fluid)
      seed<K>(v, R(a.field[0][n]), R(a.field[1][n]), R(a.field[2][n]), R(a.field[3][n]));
    else seed<K>(v, R(1), R(0), R(0), R(0));
  }
}

template <class K>
__global__ void __launch_bounds__(kBlock) k_seed(SeedArgs a) {
  using R = typename K::R;
  using T = typename K::T;
  using C = typename K::C;
  constexpr int Q = K::Q;
  const int m = blockIdx.x * kBlock + threadIdx.x;
  const Dom& d = a.d;
  if (m >= d.np) return;
  const int mx = m % d.sx, my = (m / d.sx) % d.sy, mz = m / (d.sx * d.sy);
  T* A = static_cast<T*>(a.f);
  R own[Q];
  if (a.mode != kSeedArray) seed_value_of<K>(a, m, own);
  sfor<Q>([&](auto Qs) {
    constexpr int q = Qs;
    // direction i written into slot q, and the displacement e with
    // nbr(writer, e) == m for remote slots (EsotericPull.hpp:104-111)
    int i = q, ex = 0, ey = 0, ez = 0;
    bool remote = false;
    if constexpr (q != 0) {
      if (a.parity == 0) {
        if constexpr ((q & 1) == 0) {
          remote = true;
          ex = Lat<Q>::cx[q - 1]; ey = Lat<Q>::cy[q - 1]; ez = Lat<Q>::cz[q - 1];
        }
      } else {
        if constexpr ((q & 1) == 0) {
          i = q - 1;
        } else {
          i = q + 1;
          remote = true;
          ex = Lat<Q>::cx[q]; ey = Lat<Q>::cy[q]; ez = Lat<Q>::cz[q];
        }
      }
    }
    int w = m;
    if (remote) {
      const int wx = writer_axis(mx, ex, d.nx, d.sx, d.px);
      const int wy = writer_axis(my, ey, d.ny, d.sy, d.py);
      const int wz = writer_axis(mz, ez, d.nz, d.sz, d.pz);
      w = (wx < 0 || wy < 0 || wz < 0) ? -1 : (wz * d.sy + wy) * d.sx + wx;
    }
    if (w < 0) return;  // no node writes this slot: it keeps its value
    R val;
    if (a.mode == kSeedArray) {
      val = R(a.arr[uint64_t(w) * Q + i]);
    } else if (w == m) {
      val = R(0);
      sfor<Q>([&](auto J) { if (int(J) == i) val = own[int(J)]; });
    } else {
      R other[Q];
      seed_value_of<K>(a, w, other);
      val = R(0);
      sfor<Q>([&](auto J) { if (int(J) == i) val = other[int(J)]; });
    }
    A[uint64_t(q) * a.stride + unsigned(m)] = C::enc(val, a.scale[q]);
  });
}

// ---------------------------------------------------------------------------
// Logical readout: gather<P> (FluidSolver.hpp:1243-1253) through
// Access<P>::load (EsotericPull.hpp:65-72) with the reference's neighbour
// rule, clamping included, for every padded node.
// ---------------------------------------------------------------------------
struct ReadArgs {
  const void* f;
  uint64_t stride;
  const uint8_t* flags;
  Dom d;
  int parity;
  double* out;       // gather: (n, q); macro: rho, ux, uy, uz planes
  double* out2[4];
  float inv_scale[27];
  double omega, omega_m, omega_b, F[3];
  const void* ff;  // per-node force planes (FIELD only)
  // regularised walls (null without): normal code, table index, corner rho
  // (compute type) per padded node, and the (ux, uy, uz, rho) table
  const uint8_t* bc_nrm;
  const uint16_t* bc_tag;
  const void* bc_rho;
  const double* wtab;
  const unsigned* bc_don;   // outflow donors (NrmOutFree)
  const uint8_t* bc_onrm;   // their axis face codes
  const void* wall_un;      // per-node wall velocity planes (compute type) or null
  double un_scale, outflow_rho;
  int out_order;
  const uint8_t* spec_faces;  // on-node specular face masks, or null
};

__device__ __forceinline__ int shift_axis(int v, int h, int n, int s, int per) {
  if (per) {
    if (v < h) return v + n;
    if (v >= h + n) return v - n;
    return v;
  }
  if (v < 0) return 0;
  if (v >= s) return s - 1;
  return v;
}

// Stored value of the population travelling in direction i that arrives at
// padded node n = (x, y, z): Access<P>::load's slot for (n, i).
template <class K, int i>
__device__ __forceinline__ typename K::R arriving(const ReadArgs& a, int n, int x, int y, int z) {
  using T = typename K::T;
  using C = typename K::C;
  constexpr int Q = K::Q;
  const Dom& d = a.d;
  const T* A = static_cast<const T*>(a.f);
  int node = n, slot = i;
  if constexpr (i != 0) {
    if constexpr ((i & 1) == 0) {  // read from the odd partner's neighbour
      constexpr int j = i - 1;
      const int nx_ = shift_axis(x + Lat<Q>::cx[j], d.hx, d.nx, d.sx, d.px);
      const int ny_ = shift_axis(y + Lat<Q>::cy[j], d.hy, d.ny, d.sy, d.py);
      const int nz_ = shift_axis(z + Lat<Q>::cz[j], d.hz, d.nz, d.sz, d.pz);
      node = (nz_ * d.sy + ny_) * d.sx + nx_;
      slot = a.parity == 0 ? i : i - 1;
    } else {
      slot = a.parity == 0 ? i : i + 1;
    }
  }
  return C::dec(A[uint64_t(slot) * a.stride + unsigned(node)], a.inv_scale[slot]);
}

template <class K>
__global__ void __launch_bounds__(kBlock) k_gather(ReadArgs a) {
  constexpr int Q = K::Q;
  const int n = blockIdx.x * kBlock + threadIdx.x;
  const Dom& d = a.d;
  if (n >= d.np) return;
  const int x = n % d.sx, y = (n / d.sx) % d.sy, z = n / (d.sx * d.sy);
  sfor<Q>([&](auto I) {
    constexpr int i = I;
    a.out[uint64_t(n) * Q + i] = double(arriving<K, i>(a, n, x, y, z));
  });
}

// Raw slot image (EsotericPull::storage()), (n, q) ordered.
template <class K>
__global__ void __launch_bounds__(kBlock) k_storage(ReadArgs a) {
  using T = typename K::T;
  using C = typename K::C;
  constexpr int Q = K::Q;
  const int n = blockIdx.x * kBlock + threadIdx.x;
  if (n >= a.d.np) return;
  const T* A = static_cast<const T*>(a.f);
  sfor<Q>([&](auto I) {
    constexpr int q = I;
    a.out[uint64_t(n) * Q + q] = double(C::dec(A[uint64_t(q) * a.stride + unsigned(n)],
                                               a.inv_scale[q]));
  });
}

// macro_kernel<P> fluid path (FluidSolver.hpp:1053-1056, 1192-1199): rho and u
// at the parity the next step reads; 0 on every other node.
template <class K, int P>
__device__ __forceinline__ void interior_coords(const Dom& d, int n, int& x, int& y, int& z) {
  x = n % d.sx - d.hx;
  y = (n / d.sx) % d.sy - d.hy;
  z = n / (d.sx * d.sy) - d.hz;
}

template <class K>
__device__ void wall_macro(const ReadArgs& a, unsigned n, double& r, double& ux, double& uy,
                           double& uz);

template <class K>
__global__ void __launch_bounds__(kBlock) k_macro(ReadArgs a) {
  using R = typename K::R;
  using T = typename K::T;
  const int n = blockIdx.x * kBlock + threadIdx.x;
  const Dom& d = a.d;
  if (n >= d.np) return;
  double r = 0, ux = 0, uy = 0, uz = 0;
  if (a.flags[n] == M3LB_FLUID) {
    int x, y, z;
    interior_coords<K, 0>(d, n, x, y, z);
    const Offs o = offsets(d, x, y, z);
    Params<R> p;
    p.omega = R(a.omega); p.omega_m = R(a.omega_m); p.omega_b = R(a.omega_b);
    p.F[0] = R(a.F[0]); p.F[1] = R(a.F[1]); p.F[2] = R(a.F[2]);
    force_at<K>(p, a.ff, a.stride, unsigned(n));
    R f[K::Q];
    const T* A = static_cast<const T*>(a.f);
    if (a.parity == 0) load_node<K, 0>(A, a.stride, unsigned(n), o, a.inv_scale, f);
    else load_node<K, 1>(A, a.stride, unsigned(n), o, a.inv_scale, f);
    const Macro<R> m = macroscopic<K>(f, p);
    r = double(density<K>(m)); ux = double(m.ux); uy = double(m.uy); uz = double(m.uz);
  } else if (a.flags[n] == M3LB_REGWALL && a.bc_nrm) {
    wall_macro<K>(a, unsigned(n), r, ux, uy, uz);
  } else if (a.flags[n] == M3LB_SPECNODE && a.spec_faces) {
    // macro_kernel's SpecNode path (FluidSolver.hpp:1057-1077): raw moments
    // of the mirrored incoming set
    int x, y, z;
    interior_coords<K, 0>(d, n, x, y, z);
    const Offs o = offsets(d, x, y, z);
    R f[K::Q];
    const T* A = static_cast<const T*>(a.f);
    if (a.parity == 0) load_node<K, 0>(A, a.stride, unsigned(n), o, a.inv_scale, f);
    else load_node<K, 1>(A, a.stride, unsigned(n), o, a.inv_scale, f);
    mirror_unknowns_faces<K>(f, a.spec_faces[n]);
    if constexpr (K::SHIFTED) sfor<K::Q>([&](auto I) { f[int(I)] += weight<K>(int(I)); });
    R sm = R(0), mx = R(0), my = R(0), mz = R(0);
    sfor<K::Q>([&](auto I) {
      constexpr int i = I;
      sm += f[i];
      mx += f[i] * R(cvel<K::Q>(i, 0));
      my += f[i] * R(cvel<K::Q>(i, 1));
      mz += f[i] * R(cvel<K::Q>(i, 2));
    });
    const R ir = R(1) / sm;
    r = double(sm); ux = double(mx * ir); uy = double(my * ir); uz = double(mz * ir);
  }
  if (a.out2[0]) a.out2[0][n] = r;
  if (a.out2[1]) a.out2[1][n] = ux;
  if (a.out2[2]) a.out2[2][n] = uy;
  if (a.out2[3]) a.out2[3][n] = uz;
}

// ---------------------------------------------------------------------------
// Reductions.  Wave64 butterfly, then LDS across the block's waves; one
// partial per block, summed on the host in block order (deterministic).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double* out) {
  __shared__ double lds[NV][kBlock / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    const double s = wave_sum(v[k]);
    if (lane == 0) lds[k][wv] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = 0;
      for (int w = 0; w < kBlock / 64; ++w) s += lds[k][w];
      out[uint64_t(blockIdx.x) * NV + k] = s;
    }
  }
}

// total_mass / total_momentum (reduce<P>, FluidSolver.hpp:1205-1229): per
// fluid node rho, rho ux, rho uy, rho uz, all four at once.
template <class K>
__global__ void __launch_bounds__(kBlock) k_reduce_mass(ReadArgs a) {
  using R = typename K::R;
  using T = typename K::T;
  const Dom& d = a.d;
  double acc[4] = {0, 0, 0, 0};
  for (int n = blockIdx.x * kBlock + threadIdx.x; n < d.np; n += gridDim.x * kBlock) {
    if (!fluid_like(a.flags[n])) continue;
    int x, y, z;
    interior_coords<K, 0>(d, n, x, y, z);
    const Offs o = offsets(d, x, y, z);
    Params<R> p;
    p.omega = R(a.omega); p.omega_m = R(a.omega_m); p.omega_b = R(a.omega_b);
    p.F[0] = R(a.F[0]); p.F[1] = R(a.F[1]); p.F[2] = R(a.F[2]);
    force_at<K>(p, a.ff, a.stride, unsigned(n));
    R f[K::Q];
    const T* A = static_cast<const T*>(a.f);
    if (a.parity == 0) load_node<K, 0>(A, a.stride, unsigned(n), o, a.inv_scale, f);
    else load_node<K, 1>(A, a.stride, unsigned(n), o, a.inv_scale, f);
    // the streamed unknowns of a SpecNode hold nothing until mirrored (:1218-1220)
    if (a.flags[n] == M3LB_SPECNODE && a.spec_faces) mirror_unknowns_faces<K>(f, a.spec_faces[n]);
    const Macro<R> m = macroscopic<K>(f, p);
    const R r = density<K>(m);
    acc[0] += double(r);
    acc[1] += double(r * m.ux);
    acc[2] += double(r * m.uy);
    acc[3] += double(r * m.uz);
  }
  block_sum<4>(acc, a.out);
}

// Momentum exchange with a body (validation/square_cylinder.cpp:27-50,
// 340-350): over every link from a Fluid node n to a neighbour n + c_i (the
// reference's neighbour rule) that belongs to the body,
//     F += 2 c_i f(n, opp(i)),
// f(n, opp(i)) being the population arriving at n along opp(i) -- the
// post-collision f_i that halfway bounce-back returned -- unshifted in the
// compute type as unshift<C> does.  body: padded bytes, nonzero = in the body.
// Three FP64 partials per block.
// ---------------------------------------------------------------------------
// Regularised velocity walls (Latt et al. 2008): src/boundary/Regularized.hpp
// with FluidSolver.hpp:663-710 (corner density), 793-1033 (the wall path of
// run_step) and 1076-1185 (macro_kernel).  Straight walls and corners.  Every
// node reads and writes only its own slots, so the wall nodes are updated by
// k_walls over a compact list in the same step as k_step updates the fluid
// nodes (which skips them): the result is the reference's single-pass step.
// k_walls runs first, so a corner's finite-difference stencil reads a fluid
// neighbour's incoming set from before the step.
// ---------------------------------------------------------------------------
constexpr int kNrmCorner = 7;
constexpr int kNrmOutXp = 8;
constexpr int kNrmOutFree = 9;
constexpr int kNrmOutEq = 10;
constexpr int kWallReads = 6;  // neighbour sets a wall node may read (3 axes x 2)

struct WallArgs {
  void* f;
  uint64_t stride;
  const uint8_t* flags;
  Dom d;
  const unsigned* list;  // padded index of each wall node of this launch
  int nwall;
  int fd_corners;
  const uint8_t* bc_nrm;
  const uint16_t* bc_tag;
  const uint32_t* bc_unk;  // bit i: direction i did not stream from a fluid-like node
  const uint8_t* bc_ext;   // corners: extrapolation axis 1..6, 0 = none
  void* bc_rho;            // compute type
  const double* wtab;      // (ux, uy, uz, rho) per table entry
  const unsigned* bc_don;  // NrmOutFree donors
  const void* wall_un;     // per-node wall velocity planes (compute type) or null
  double un_scale, outflow_rho;
  int out_order, force_bc;
  // Where each of a wall node's neighbour reads comes from: -1 = the live
  // slots now, k >= 0 = row k of snap (that neighbour's incoming set taken
  // before the step); kWallReads entries per list entry.
  const int* rd;
  const void* snap;
  double omega, omega_m, omega_b, F[3];
  const void* ff;
  void* macro[4];  // step(store_macroscopic): rho, ux, uy, uz or null
  float scale[27], inv_scale[27];
};

__host__ __device__ __forceinline__ void normal_of(int code, int n[3]) {
  n[0] = n[1] = n[2] = 0;
  if (code == 1 || code == kNrmOutXp) n[0] = 1;
  else if (code == 2) n[0] = -1;
  else if (code == 3) n[1] = 1;
  else if (code == 4) n[1] = -1;
  else if (code == 5) n[2] = 1;
  else if (code == 6) n[2] = -1;
}

// EquilibriumOf<Collision>::eq (Regularized.hpp:14-19): the operator's own
// equilibrium for BGK and TRT, the product form for the moment operators.
template <class K, int i>
__device__ __forceinline__ typename K::R wall_eq(typename K::R rho, typename K::R ux,
                                                 typename K::R uy, typename K::R uz) {
  using R = typename K::R;
  constexpr bool product = K::OP != M3LB_OP_BGK && K::OP != M3LB_OP_TRT;
  R phi;
  if constexpr (product) phi = phi_product<R, K::Q, i>(ux, uy, uz);
  else phi = phi_second<R, K::Q, i>(ux, uy, uz);
  return weight<K>(i) * rho * (R(1) + phi);
}

// Regularized::density and, with rho given, normal_velocity (the same
// partition of the set inverted)
template <class K>
__device__ __forceinline__ void bc_partition(const typename K::R* f, const int nrm[3],
                                             typename K::R& out, typename K::R& tang) {
  using R = typename K::R;
  out = R(0);
  tang = R(0);
  sfor<K::Q>([&](auto I) {
    constexpr int i = I;
    const int cn = cvel<K::Q>(i, 0) * nrm[0] + cvel<K::Q>(i, 1) * nrm[1] + cvel<K::Q>(i, 2) * nrm[2];
    if (cn > 0) out += f[i];
    else if (cn == 0) tang += f[i];
  });
}
template <class K>
__device__ __forceinline__ typename K::R bc_density(const typename K::R* f, const int nrm[3],
                                                    const typename K::R u[3]) {
  using R = typename K::R;
  R out, tang;
  bc_partition<K>(f, nrm, out, tang);
  const R uperp = u[0] * R(nrm[0]) + u[1] * R(nrm[1]) + u[2] * R(nrm[2]);
  return (R(2) * out + tang) / (R(1) + uperp);
}
template <class K>
__device__ __forceinline__ typename K::R bc_normal_velocity(const typename K::R* f,
                                                            const int nrm[3], typename K::R rho) {
  using R = typename K::R;
  R out, tang;
  bc_partition<K>(f, nrm, out, tang);
  return (R(2) * out + tang) / rho - R(1);
}

template <class K, int i>
__device__ __forceinline__ typename K::R qpi(const typename K::R Pi[6]) {
  using R = typename K::R;
  const R cs2v = R(1) / R(3);
  const R cx = R(cvel<K::Q>(i, 0)), cy = R(cvel<K::Q>(i, 1)), cz = R(cvel<K::Q>(i, 2));
  return (cx * cx - cs2v) * Pi[0] + (cy * cy - cs2v) * Pi[1] +
         ((K::D == 3) ? (cz * cz - cs2v) * Pi[2] : R(0)) +
         R(2) * (cx * cy * Pi[3] + ((K::D == 3) ? (cx * cz * Pi[4] + cy * cz * Pi[5]) : R(0)));
}

// Regularized::apply: unknown directions from the bounced non-equilibrium
// part (with the Guo odd part), then every direction rebuilt from (rho, u, Pi).
template <class K>
__device__ __forceinline__ void bc_apply(typename K::R* f, typename K::R rho,
                                         const typename K::R u[3], uint32_t unk,
                                         const typename K::R Fv[3]) {
  using R = typename K::R;
  constexpr int Q = K::Q;
  const R cs2v = R(1) / R(3), cs4v = cs2v * cs2v;
  R feq[Q];
  sfor<Q>([&](auto I) { feq[int(I)] = wall_eq<K, int(I)>(rho, u[0], u[1], u[2]); });
  const R hodd = R(0.5) / cs2v;
  sfor<Q>([&](auto I) {
    constexpr int i = I;
    constexpr int io = opp<Q>(i);
    if (!(unk & (1u << i))) return;
    const R cF = R(cvel<Q>(i, 0)) * Fv[0] + R(cvel<Q>(i, 1)) * Fv[1] + R(cvel<Q>(i, 2)) * Fv[2];
    const R odd = weight<K>(i) * hodd * cF;
    f[i] = (unk & (1u << io)) ? feq[i] - odd : feq[i] + (f[io] - feq[io]) - R(2) * odd;
  });
  R Pi[6] = {R(0), R(0), R(0), R(0), R(0), R(0)};
  sfor<Q>([&](auto I) {
    constexpr int i = I;
    const R dlt = f[i] - feq[i];
    const R cx = R(cvel<Q>(i, 0)), cy = R(cvel<Q>(i, 1)), cz = R(cvel<Q>(i, 2));
    Pi[0] += dlt * cx * cx; Pi[1] += dlt * cy * cy; Pi[2] += dlt * cz * cz;
    Pi[3] += dlt * cx * cy; Pi[4] += dlt * cx * cz; Pi[5] += dlt * cy * cz;
  });
  sfor<Q>([&](auto I) {
    constexpr int i = I;
    const R cF = R(cvel<Q>(i, 0)) * Fv[0] + R(cvel<Q>(i, 1)) * Fv[1] + R(cvel<Q>(i, 2)) * Fv[2];
    f[i] = feq[i] - weight<K>(i) * hodd * cF + weight<K>(i) / (R(2) * cs4v) * qpi<K, i>(Pi);
  });
}

// Regularized::apply_with_stress
template <class K>
__device__ __forceinline__ void bc_apply_with_stress(typename K::R* f, typename K::R rho,
                                                     const typename K::R u[3],
                                                     const typename K::R Pi[6]) {
  using R = typename K::R;
  const R cs2v = R(1) / R(3), cs4v = cs2v * cs2v;
  sfor<K::Q>([&](auto I) {
    constexpr int i = I;
    f[i] = wall_eq<K, i>(rho, u[0], u[1], u[2]) + weight<K>(i) / (R(2) * cs4v) * qpi<K, i>(Pi);
  });
}

// The table velocity of wall node n (the finite-difference stencil's wall
// neighbours always use it, FluidSolver.hpp:961-966)
template <class K>
__device__ __forceinline__ void table_velocity(const double* wtab, const uint16_t* tag,
                                               unsigned n, typename K::R u[3]) {
  const double* w = wtab + 4 * size_t(tag[n]);
  u[0] = typename K::R(w[0]);
  u[1] = typename K::R(w[1]);
  u[2] = typename K::R(w[2]);
}

// The wall velocity of node n: the per-node field when one is set
// (set_wall_velocity_field), else its table entry (FluidSolver.hpp:686-689).
template <class K>
__device__ __forceinline__ void wall_velocity(const double* wtab, const uint16_t* tag,
                                              const void* un, double un_scale, int np,
                                              unsigned n, typename K::R u[3]) {
  using R = typename K::R;
  if (un) {
    const R* w = static_cast<const R*>(un);
    const R s = R(un_scale);
    u[0] = s * w[n];
    u[1] = s * w[size_t(np) + n];
    u[2] = s * w[2 * size_t(np) + n];
    return;
  }
  table_velocity<K>(wtab, tag, n, u);
}

// The incoming set of padded node n at parity P, unshifted to f.
template <class K, int P>
__device__ __forceinline__ void load_raw(const void* f, uint64_t stride, const Dom& d,
                                         unsigned n, const float* inv, typename K::R* out) {
  int x, y, z;
  interior_coords<K, 0>(d, int(n), x, y, z);
  const Offs o = offsets(d, x, y, z);
  load_node<K, P>(static_cast<const typename K::T*>(f), stride, n, o, inv, out);
  if constexpr (K::SHIFTED) sfor<K::Q>([&](auto I) { out[int(I)] += weight<K>(int(I)); });
}

// Sum and mean velocity of an (unshifted) set; guard: 0 instead of dividing
// by a non-positive sum
template <class K>
__device__ __forceinline__ typename K::R set_velocity(const typename K::R* g, typename K::R v[3],
                                                      bool guard) {
  using R = typename K::R;
  R sm = R(0), mx = R(0), my = R(0), mz = R(0);
  sfor<K::Q>([&](auto I) {
    constexpr int i = I;
    sm += g[i];
    mx += g[i] * R(cvel<K::Q>(i, 0));
    my += g[i] * R(cvel<K::Q>(i, 1));
    mz += g[i] * R(cvel<K::Q>(i, 2));
  });
  const R ir = guard ? ((sm > R(0)) ? R(1) / sm : R(0)) : R(1) / sm;
  v[0] = mx * ir; v[1] = my * ir; v[2] = mz * ir;
  return sm;
}

__device__ __forceinline__ unsigned interior_id(const Dom& d, int x, int y, int z) {
  return unsigned(((z + d.hz) * d.sy + (y + d.hy)) * d.sx + x + d.hx);
}

// corner_density<P>, first half: rho of every straight-wall node (imposed at
// the outflow codes).
template <class K, int P>
__global__ void __launch_bounds__(kBlock) k_wall_density(WallArgs a) {
  using R = typename K::R;
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= a.nwall) return;
  const unsigned n = a.list[k];
  const int code = a.bc_nrm[n];
  if (code == kNrmCorner || code == 0) return;
  R* rho = static_cast<R*>(a.bc_rho);
  if (code == kNrmOutXp || code == kNrmOutEq) {
    rho[n] = R(a.wtab[4 * size_t(a.bc_tag[n]) + 3]);
    return;
  }
  if (code == kNrmOutFree) {
    rho[n] = R(a.outflow_rho);
    return;
  }
  R f[K::Q];
  load_raw<K, P>(a.f, a.stride, a.d, n, a.inv_scale, f);
  int nrm[3];
  normal_of(code, nrm);
  R uw[3];
  wall_velocity<K>(a.wtab, a.bc_tag, a.wall_un, a.un_scale, a.d.np, n, uw);
  rho[n] = bc_density<K>(f, nrm, uw);
}

// corner_density<P>, second half: each corner 2 rho(n1) - rho(n2) along its
// axis, or 1 without a stencil.
template <class K>
__global__ void __launch_bounds__(kBlock) k_corner_density(WallArgs a) {
  using R = typename K::R;
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= a.nwall) return;
  const unsigned n = a.list[k];
  if (a.bc_nrm[n] != kNrmCorner) return;
  R* rho = static_cast<R*>(a.bc_rho);
  const int e = a.bc_ext[n];
  if (e == 0) {
    rho[n] = R(1);
    return;
  }
  const Dom& d = a.d;
  int x, y, z;
  interior_coords<K, 0>(d, int(n), x, y, z);
  const int ax = (e - 1) >> 1, sg = (e & 1) ? 1 : -1;  // 1:+x 2:-x 3:+y 4:-y 5:+z 6:-z
  int v[3] = {0, 0, 0};
  v[ax] = sg;
  const unsigned n1 = interior_id(d, x + v[0], y + v[1], z + v[2]);
  const unsigned n2 = interior_id(d, x + 2 * v[0], y + 2 * v[1], z + 2 * v[2]);
  rho[n] = R(2) * rho[n1] - rho[n2];
}

// The neighbour sets the wall nodes read, taken before the step (see rd).
struct SnapArgs {
  const void* f;
  uint64_t stride;
  Dom d;
  const unsigned* nodes;
  int count;
  void* out;  // count x Q, compute type, unshifted
  float inv_scale[27];
};

template <class K, int P>
__global__ void __launch_bounds__(kBlock) k_snapshot(SnapArgs a) {
  using R = typename K::R;
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= a.count) return;
  R g[K::Q];
  load_raw<K, P>(a.f, a.stride, a.d, a.nodes[k], a.inv_scale, g);
  R* o = static_cast<R*>(a.out) + size_t(k) * K::Q;
  sfor<K::Q>([&](auto I) { o[int(I)] = g[int(I)]; });
}

// The wall path of run_step for one wall node (FluidSolver.hpp:793-1033):
// velocity walls, corners and the outflow codes; rebuild, collide, store
// (the population-extrapolation and equilibrium outlets store without
// colliding).
template <class K, int P>
__global__ void __launch_bounds__(kBlock) k_walls(WallArgs a) {
  using R = typename K::R;
  using T = typename K::T;
  constexpr int Q = K::Q;
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= a.nwall) return;
  const unsigned n = a.list[k];
  const Dom& d = a.d;
  int x, y, z;
  interior_coords<K, 0>(d, int(n), x, y, z);
  const Offs o = offsets(d, x, y, z);
  T* A = static_cast<T*>(a.f);
  R f[Q];
  load_node<K, P>(A, a.stride, n, o, a.inv_scale, f);
  if constexpr (K::SHIFTED) sfor<Q>([&](auto I) { f[int(I)] += weight<K>(int(I)); });
  const int code = a.bc_nrm[n];
  int nrm[3];
  normal_of(code, nrm);
  const size_t tg = a.bc_tag[n];
  R uw[3];
  wall_velocity<K>(a.wtab, a.bc_tag, a.wall_un, a.un_scale, d.np, n, uw);
  Params<R> p;
  p.omega = R(a.omega); p.omega_m = R(a.omega_m); p.omega_b = R(a.omega_b);
  p.F[0] = R(a.F[0]); p.F[1] = R(a.F[1]); p.F[2] = R(a.F[2]);
  force_at<K>(p, a.ff, a.stride, n);
  R Fv[3] = {R(0), R(0), R(0)};
  if constexpr (K::FORCE) {
    if (a.force_bc) { Fv[0] = p.F[0]; Fv[1] = p.F[1]; Fv[2] = p.F[2]; }
  }
  // neighbour read r of this node: the snapshot row or the live slots
  auto incoming = [&](int r, unsigned m, R g[Q]) {
    const int src = a.rd ? a.rd[size_t(k) * kWallReads + r] : -1;
    if (src >= 0) {
      const R* sp = static_cast<const R*>(a.snap) + size_t(src) * Q;
      sfor<Q>([&](auto I) { g[int(I)] = sp[int(I)]; });
    } else {
      load_raw<K, P>(a.f, a.stride, d, m, a.inv_scale, g);
    }
  };
  R ur[3] = {uw[0], uw[1], uw[2]};
  R rw = R(0);
  bool collide_it = true;
  if (code == kNrmOutFree) {
    // the donor's distribution rescaled to the imposed density (:843-880)
    R g[Q];
    incoming(0, a.bc_don[n], g);
    R sm = R(0);
    sfor<Q>([&](auto I) { sm += g[int(I)]; });
    if (sm > R(0)) {
      const R sc = R(a.outflow_rho) / sm;
      sfor<Q>([&](auto I) { f[int(I)] = g[int(I)] * sc; });
    }
    collide_it = false;
  } else if (code == kNrmOutEq) {
    // equilibrium at the imposed rho and the upstream velocity (:881-911)
    const R rw2 = R(a.wtab[4 * tg + 3]);
    R g[Q], mu[3];
    incoming(0, interior_id(d, x - 1, y, z), g);
    set_velocity<K>(g, mu, true);
    sfor<Q>([&](auto I) { f[int(I)] = wall_eq<K, int(I)>(rw2, mu[0], mu[1], mu[2]); });
    collide_it = false;
  } else if (code == kNrmOutXp) {
    // back-pressure outlet (:912-951): rho imposed, the normal velocity from
    // the inverted closure, the tangential ones from upstream
    rw = R(a.wtab[4 * tg + 3]);
    R g[Q], ut[3];
    incoming(0, interior_id(d, x - 1, y, z), g);
    set_velocity<K>(g, ut, false);
    if (a.out_order >= 2) {
      R v2[3];
      incoming(1, interior_id(d, x - 2, y, z), g);
      set_velocity<K>(g, v2, false);
      for (int c = 0; c < 3; ++c) ut[c] = R(2) * ut[c] - v2[c];
    }
    const R un = bc_normal_velocity<K>(f, nrm, rw);
    R utn = R(0);
    for (int c = 0; c < 3; ++c) utn += ut[c] * R(nrm[c]);
    for (int c = 0; c < 3; ++c) ur[c] = ut[c] + (un - utn) * R(nrm[c]);
  } else {
    rw = (code == kNrmCorner) ? static_cast<const R*>(a.bc_rho)[n] : bc_density<K>(f, nrm, ur);
  }
  if (collide_it) {
    if (a.fd_corners && code == kNrmCorner) {
      const int ic[3] = {x, y, z}, nn[3] = {d.nx, d.ny, d.nz}, per[3] = {d.px, d.py, d.pz};
      R grad[3][3] = {{R(0), R(0), R(0)}, {R(0), R(0), R(0)}, {R(0), R(0), R(0)}};
      // velocity at a stencil node: imposed (table) at a wall node, else from
      // its incoming set
      auto vel_at = [&](int r, int cx, int cy, int cz, R out[3]) {
        const unsigned m = interior_id(d, cx, cy, cz);
        if (a.flags[m] == M3LB_REGWALL) {
          table_velocity<K>(a.wtab, a.bc_tag, m, out);
          return;
        }
        R g[Q];
        incoming(r, m, g);
        set_velocity<K>(g, out, false);
      };
      for (int ax = 0; ax < K::D; ++ax) {
        if (nn[ax] < 3) continue;
        int dir = 0;
        if (!per[ax]) {
          if (ic[ax] == 0) dir = 1;
          else if (ic[ax] == nn[ax] - 1) dir = -1;
        }
        int c1[3] = {ic[0], ic[1], ic[2]}, c2[3] = {ic[0], ic[1], ic[2]};
        R u1[3], u2[3];
        if (dir != 0) {
          c1[ax] += dir; c2[ax] += 2 * dir;
          vel_at(2 * ax, c1[0], c1[1], c1[2], u1);
          vel_at(2 * ax + 1, c2[0], c2[1], c2[2], u2);
          for (int b = 0; b < K::D; ++b)
            grad[ax][b] = R(dir) * (R(-1.5) * uw[b] + R(2) * u1[b] - R(0.5) * u2[b]);
        } else {
          c1[ax] = (ic[ax] + 1) % nn[ax];
          c2[ax] = (ic[ax] - 1 + nn[ax]) % nn[ax];
          vel_at(2 * ax, c1[0], c1[1], c1[2], u1);
          vel_at(2 * ax + 1, c2[0], c2[1], c2[2], u2);
          for (int b = 0; b < K::D; ++b) grad[ax][b] = R(0.5) * (u1[b] - u2[b]);
        }
      }
      // Regularized::stress_from_gradient with the shear rate
      const R kf = -R(2) * (R(1) / R(3)) * rw / p.omega;
      const R Pi[6] = {kf * grad[0][0], kf * grad[1][1], kf * grad[2][2],
                       kf * R(0.5) * (grad[0][1] + grad[1][0]), kf * R(0.5) * (grad[0][2] + grad[2][0]),
                       kf * R(0.5) * (grad[1][2] + grad[2][1])};
      bc_apply_with_stress<K>(f, rw, ur, Pi);
    } else {
      bc_apply<K>(f, rw, ur, a.bc_unk[n], Fv);
    }
  }
  if constexpr (K::SHIFTED) sfor<Q>([&](auto I) { f[int(I)] -= weight<K>(int(I)); });
  if (collide_it) {
    const Macro<R> m = collide_fused<K>(f, p);
    if (a.macro[0]) {
      static_cast<R*>(a.macro[0])[n] = density<K>(m);
      static_cast<R*>(a.macro[1])[n] = m.ux;
      static_cast<R*>(a.macro[2])[n] = m.uy;
      static_cast<R*>(a.macro[3])[n] = m.uz;
    }
  }
  store_node<K, P>(A, a.stride, n, o, a.scale, f);
}

// Specular cells over a compact list (FluidSolver.hpp:745-756, 782-793).  A
// ghost (SpecWall) permutes its incoming set about its normal's axis and
// stores it back without colliding; an on-node node (SpecNode) mirrors its
// unknown directions and collides as fluid.  Both touch only their own slots.
struct SpecArgs {
  void* f;
  uint64_t stride;
  const uint8_t* flags;
  Dom d;
  const unsigned* list;
  int count;
  const uint8_t* spec_nrm;
  const uint8_t* spec_faces;
  double omega, omega_m, omega_b, F[3];
  const void* ff;
  void* macro[4];
  float scale[27], inv_scale[27];
};

template <class K, int P>
__global__ void __launch_bounds__(kBlock) k_specular(SpecArgs a) {
  using R = typename K::R;
  using T = typename K::T;
  constexpr int Q = K::Q;
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= a.count) return;
  const unsigned n = a.list[k];
  const Dom& d = a.d;
  int x, y, z;
  interior_coords<K, 0>(d, int(n), x, y, z);
  const Offs o = offsets(d, x, y, z);
  T* A = static_cast<T*>(a.f);
  R f[Q];
  load_node<K, P>(A, a.stride, n, o, a.inv_scale, f);
  if (a.flags[n] == M3LB_SPECWALL) {
    // the weights are invariant under an axis mirror, so the permutation
    // commutes with the shift
    int nv[3];
    normal_of(a.spec_nrm[n], nv);
    const int ax = nv[0] ? 0 : (nv[1] ? 1 : 2);
    R g[Q];
    sfor<Q>([&](auto I) {
      constexpr int i = I;
      g[i] = ax == 0 ? f[mirror_of<Q, i, 1, 0, 0>()]
                     : (ax == 1 ? f[mirror_of<Q, i, 0, 1, 0>()] : f[mirror_of<Q, i, 0, 0, 1>()]);
    });
    store_node<K, P>(A, a.stride, n, o, a.scale, g);
    return;
  }
  mirror_unknowns_faces<K>(f, a.spec_faces[n]);
  Params<R> p;
  p.omega = R(a.omega); p.omega_m = R(a.omega_m); p.omega_b = R(a.omega_b);
  p.F[0] = R(a.F[0]); p.F[1] = R(a.F[1]); p.F[2] = R(a.F[2]);
  force_at<K>(p, a.ff, a.stride, n);
  const Macro<R> m = collide_fused<K>(f, p);
  store_node<K, P>(A, a.stride, n, o, a.scale, f);
  if (a.macro[0]) {
    static_cast<R*>(a.macro[0])[n] = density<K>(m);
    static_cast<R*>(a.macro[1])[n] = m.ux;
    static_cast<R*>(a.macro[2])[n] = m.uy;
    static_cast<R*>(a.macro[3])[n] = m.uz;
  }
}

// macro_kernel's RegWall path (FluidSolver.hpp:1078-1185): the imposed (or
// outflow) velocity and the closure (or imposed) density; a corner reports
// the last pass's rho.
template <class K>
__device__ void wall_macro(const ReadArgs& a, unsigned n, double& r, double& ux, double& uy,
                           double& uz) {
  using R = typename K::R;
  constexpr int Q = K::Q;
  const Dom& d = a.d;
  const int code = a.bc_nrm[n];
  const size_t tg = a.bc_tag[n];
  R uw[3];
  wall_velocity<K>(a.wtab, a.bc_tag, a.wall_un, a.un_scale, d.np, n, uw);
  R fw[Q];
  auto raw = [&](unsigned m, R* g) {
    if (a.parity == 0) load_raw<K, 0>(a.f, a.stride, d, m, a.inv_scale, g);
    else load_raw<K, 1>(a.f, a.stride, d, m, a.inv_scale, g);
  };
  raw(n, fw);
  int nrm[3];
  normal_of(code, nrm);
  int x, y, z;
  interior_coords<K, 0>(d, int(n), x, y, z);
  R rw;
  if (code == kNrmOutFree) {
    rw = R(a.outflow_rho);
    int onr[3];
    normal_of(a.bc_onrm[n], onr);
    R g[Q], v[3];
    raw(a.bc_don[n], g);
    set_velocity<K>(g, v, true);
    const R un = bc_normal_velocity<K>(fw, onr, rw);
    R utn = R(0);
    for (int c = 0; c < 3; ++c) utn += v[c] * R(onr[c]);
    for (int c = 0; c < 3; ++c) uw[c] = v[c] + (un - utn) * R(onr[c]);
  } else if (code == kNrmOutEq) {
    rw = R(a.wtab[4 * tg + 3]);
    R g[Q];
    raw(interior_id(d, x - 1, y, z), g);
    set_velocity<K>(g, uw, true);
  } else if (code == kNrmOutXp) {
    rw = R(a.wtab[4 * tg + 3]);
    R g[Q], ut[3];
    raw(interior_id(d, x - 1, y, z), g);
    set_velocity<K>(g, ut, false);
    if (a.out_order >= 2) {
      R v2[3];
      raw(interior_id(d, x - 2, y, z), g);
      set_velocity<K>(g, v2, false);
      for (int c = 0; c < 3; ++c) ut[c] = R(2) * ut[c] - v2[c];
    }
    const R un = bc_normal_velocity<K>(fw, nrm, rw);
    R utn = R(0);
    for (int c = 0; c < 3; ++c) utn += ut[c] * R(nrm[c]);
    for (int c = 0; c < 3; ++c) uw[c] = ut[c] + (un - utn) * R(nrm[c]);
  } else {
    rw = (code == kNrmCorner) ? static_cast<const R*>(a.bc_rho)[n] : bc_density<K>(fw, nrm, uw);
  }
  r = double(rw);
  ux = double(uw[0]); uy = double(uw[1]); uz = double(uw[2]);
}

template <class K>
__global__ void __launch_bounds__(kBlock) k_momentum_exchange(ReadArgs a, const uint8_t* body) {
  using R = typename K::R;
  constexpr int Q = K::Q;
  const Dom& d = a.d;
  double acc[3] = {0, 0, 0};
  for (int n = blockIdx.x * kBlock + threadIdx.x; n < d.np; n += gridDim.x * kBlock) {
    if (a.flags[n] != M3LB_FLUID) continue;
    const int x = n % d.sx, y = (n / d.sx) % d.sy, z = n / (d.sx * d.sy);
    sfor<Q - 1>([&](auto I) {
      constexpr int i = int(I) + 1;
      constexpr int io = opp<Q>(i);
      const int mx = shift_axis(x + Lat<Q>::cx[i], d.hx, d.nx, d.sx, d.px);
      const int my = shift_axis(y + Lat<Q>::cy[i], d.hy, d.ny, d.sy, d.py);
      const int mz = shift_axis(z + Lat<Q>::cz[i], d.hz, d.nz, d.sz, d.pz);
      if (!body[(mz * d.sy + my) * d.sx + mx]) return;
      R f = arriving<K, io>(a, n, x, y, z);
      if constexpr (K::SHIFTED) f = f + weight<K>(io);
      const double fd = double(f);
      acc[0] += 2.0 * Lat<Q>::cx[i] * fd;
      acc[1] += 2.0 * Lat<Q>::cy[i] * fd;
      acc[2] += 2.0 * Lat<Q>::cz[i] * fd;
    });
  }
  block_sum<3>(acc, a.out);
}

// Exact sum of every slot: value * 2^100 truncated toward zero, summed as a
// 128-bit two's-complement integer (two 64-bit limbs with carry).
struct U128 {
  uint64_t hi, lo;
};

__device__ __forceinline__ U128 add128(U128 a, U128 b) {
  U128 r;
  r.lo = a.lo + b.lo;
  r.hi = a.hi + b.hi + (r.lo < a.lo ? 1 : 0);
  return r;
}

__device__ __forceinline__ U128 fixed100(double v) {
  U128 r{0, 0};
  if (v == 0.0 || !isfinite(v)) return r;
  const uint64_t bits = __double_as_longlong(v);
  const int ebits = int((bits >> 52) & 0x7ff);
  uint64_t mant = bits & 0xfffffffffffffull;
  int e;  // |v| = mant * 2^e
  if (ebits == 0) {
    e = -1074;
  } else {
    mant |= (1ull << 52);
    e = ebits - 1075;
  }
  const int sh = e + 100;
  if (sh >= 0) {
    if (sh + 53 > 126) {
      r.hi = 1ull << 62;
    } else if (sh >= 64) {
      r.hi = mant << (sh - 64);
    } else if (sh == 0) {
      r.lo = mant;
    } else {
      r.lo = mant << sh;
      r.hi = mant >> (64 - sh);
    }
  } else if (sh > -64) {
    r.lo = mant >> (-sh);
  }
  if (v < 0) {
    r.lo = ~r.lo;
    r.hi = ~r.hi;
    r = add128(r, U128{0, 1});
  }
  return r;
}

template <class K>
__global__ void __launch_bounds__(kBlock) k_sum_exact(ReadArgs a) {
  using T = typename K::T;
  using C = typename K::C;
  constexpr int Q = K::Q;
  const T* A = static_cast<const T*>(a.f);
  U128 acc{0, 0};
  for (int n = blockIdx.x * kBlock + threadIdx.x; n < a.d.np; n += gridDim.x * kBlock) {
    sfor<Q>([&](auto I) {
      constexpr int q = I;
      acc = add128(acc, fixed100(double(C::dec(A[uint64_t(q) * a.stride + unsigned(n)],
                                                a.inv_scale[q]))));
    });
  }
  __shared__ uint64_t lh[kBlock], ll[kBlock];
  lh[threadIdx.x] = acc.hi;
  ll[threadIdx.x] = acc.lo;
  __syncthreads();
  if (threadIdx.x == 0) {
    U128 s{0, 0};
    for (int k = 0; k < kBlock; ++k) s = add128(s, U128{lh[k], ll[k]});
    uint64_t* o = reinterpret_cast<uint64_t*>(a.out);
    o[2 * blockIdx.x] = s.hi;
    o[2 * blockIdx.x + 1] = s.lo;
  }
}

// Per-slot maximum |stored value| (decoded), for the 16-bit scale choice.
template <class K>
__global__ void __launch_bounds__(kBlock) k_absmax(ReadArgs a) {
  using T = typename K::T;
  using C = typename K::C;
  constexpr int Q = K::Q;
  const T* A = static_cast<const T*>(a.f);
  double mx[Q];
  sfor<Q>([&](auto I) { mx[int(I)] = 0.0; });
  for (int n = blockIdx.x * kBlock + threadIdx.x; n < a.d.np; n += gridDim.x * kBlock) {
    sfor<Q>([&](auto I) {
      constexpr int q = I;
      const double v = fabs(double(C::dec(A[uint64_t(q) * a.stride + unsigned(n)],
                                          a.inv_scale[q])));
      mx[q] = v > mx[q] ? v : mx[q];
    });
  }
  __shared__ double lds[kBlock];
  sfor<Q>([&](auto I) {
    constexpr int q = I;
    lds[threadIdx.x] = mx[q];
    __syncthreads();
    if (threadIdx.x == 0) {
      double m = 0;
      for (int k = 0; k < kBlock; ++k) m = lds[k] > m ? lds[k] : m;
      a.out[uint64_t(blockIdx.x) * Q + q] = m;
    }
    __syncthreads();
  });
}

// Re-encode 16-bit storage from old to new scales.  The ratio is a power of
// two, so a larger scale is exact unless the code saturates.  A smaller scale
// drops low-order bits wherever the form has a fixed step: every I16 code, and
// E2M13 codes below 32 (step 2^-8) and FP16 subnormals.  Those values are
// rounded to the nearest code of the new scale, an error of at most half its
// step; the floating forms' normal codes stay exact.
struct Ratio {
  float r[27];
};

template <class K>
__global__ void __launch_bounds__(kBlock) k_rescale(void* f, uint64_t stride, int np,
                                                    Ratio ratio) {
  using T = typename K::T;
  constexpr int Q = K::Q;
  using C = typename K::C;
  T* A = static_cast<T*>(f);
  for (int n = blockIdx.x * kBlock + threadIdx.x; n < np; n += gridDim.x * kBlock) {
    sfor<Q>([&](auto I) {
      constexpr int q = I;
      T& c = A[uint64_t(q) * stride + unsigned(n)];
      c = C::enc(C::dec(c, 1.0f) * ratio.r[q], 1.0f);
    });
  }
}

}  // namespace m3lb
