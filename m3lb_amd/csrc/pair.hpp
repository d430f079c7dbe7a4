// pair.hpp -- two esoteric-pull steps in one pass over HBM.
//
// Step t (parity P) and step t+1 (parity 1-P) of the same nodes run in one
// persistent launch, step t+1 following step t plane by plane a few z-planes
// behind it, so that step t+1 reads what step t wrote while it is still in
// the caches instead of from HBM.  The bit pattern of the result is that of
// two separate k_step launches: every node runs the same node update on the
// same inputs (kernels.hpp node_update).
//
// Why it is exact.  In the esoteric pull every slot is read and written by
// exactly one node per step (EsotericPull.hpp:5-25), and a node's update at
// step t+1 reads only slots written at step t by itself and its 26
// neighbours, and overwrites only those.  So step t+1 may run at node n as
// soon as step t has run at n and its neighbours, and at no other time
// constraint.
//
// Work split.  The interior rows are cut into S stripes along y.  Blocks read
// the XCD they run on (hardware register XCC_ID) and each stripe is claimed by
// one XCD, whose blocks do all of the stripe's launch-1 work.  Within a
// stripe, tickets taken from one atomic counter hand out tiles in the order
//     A(0) .. A(L), then A(k), B(k-L) for k = L+1 .. nz-1, then B(nz-L) .. B(nz-2)
// where A(z) is step t on plane z of the stripe and B(z) step t+1 on the
// stripe's inner rows of plane z (all rows but its first and last).  A B(z)
// tile waits until the A tiles of planes z-1, z and z+1 have all finished.
// Every tile it waits on has an earlier ticket, held by a running block, so
// the earliest unfinished ticket can always proceed: no deadlock.
//
// What is deferred.  Step t+1 of the stripes' first and last rows and of the
// planes 0 and nz-1 (their neighbours lie in another stripe, or across the
// periodic wrap, which step t reaches last) runs in a second, ordinary launch
// (k_pair_tail) after the first has finished.
//
// Coherence.  The 8 XCDs' L2 caches are not coherent with each other within a
// launch, but the CUs of one XCD share its L2.  In launch 1 every slot is
// accessed by the blocks of one stripe only (the slot sets of different
// stripes' nodes are disjoint; a cache line shared between stripes is only
// ever written in disjoint bytes, which the L2s write back by byte mask).
// Within a stripe a finished tile publishes its stores to the L2 (wait for
// the stores, barrier, then a relaxed agent-scope counter increment), and a
// waiting tile invalidates its CU's L1 before it reads (the workgroup-scope
// acquire of split-workgroup mode, an XCD-wide scope here).  Across the two
// launches the kernel boundary orders everything.
//
// Every wait is bounded: after spin_limit polls (kPairSpinLimit, about a
// second) the block records an error word, skips the tile and carries on, and
// every later wait gives up at once, so the launch always drains and the host
// reports the failure.
#pragma once

#include "kernels.hpp"
#include "pair_plan.hpp"

namespace m3lb {

struct PairArgs {
  StepArgs a;
  int stripes;                    // S
  int lag;                        // L: step t+1 follows L planes behind step t
  int group;                      // consecutive 256-node tiles per ticket
  unsigned spin_limit;            // polls before a wait gives up
  unsigned wait_extra;            // 0; 1 makes every wait unsatisfiable (failure-path test)
  int y0[kPairMaxStripes + 1];    // stripe s holds interior rows [y0[s], y0[s+1])
  unsigned* sync;                 // pair_sync_words(nz, S) words, zero before launch 1
};

__device__ __forceinline__ unsigned xcc_id() {
  // HW_REG_XCC_ID (id 20), bits [3:0]: the XCD this wave runs on
  return __builtin_amdgcn_s_getreg((20) | (0 << 6) | ((4 - 1) << 11)) & 0xFu;
}

// The per-node body of a step with separate cache hints for loads and stores
// of the node's own slots (node_update uses one hint for both).
template <class K, int P, bool GEOM, bool NT_LD, bool NT_ST>
__device__ __forceinline__ void node_update_hinted(typename K::T* __restrict__ A,
                                                   const StepArgs& a, int x, int y, int z,
                                                   const Params<typename K::R>& p0) {
  using R = typename K::R;
  const Dom& d = a.d;
  const unsigned n = unsigned(((z + d.hz) * d.sy + (y + d.hy)) * d.sx + d.hx + x);
  if constexpr (GEOM) {
    if (a.flags[n] != M3LB_FLUID) return;
  }
  Params<R> p = p0;
  force_at<K>(p, a.ff, a.stride, n);
  const Offs o = offsets(d, x, y, z);
  R f[K::Q];
  if constexpr (K::C::scaled) {
    // 16-bit storage: the node body runs in the persistent loop, where the
    // direction bases and the 27 scales would otherwise be held in scalar
    // registers across every iteration (and spilled through vector lanes).
    // The empty asm statements make the stride and a zero index opaque, so
    // that bases and scales are formed again, from the kernel arguments, for
    // the loads and, after the collision, for the stores; 1/s comes from
    // the bits of the power-of-two scale.
    uint64_t stride_ld = a.stride;
    uint32_t k0 = 0;
    asm volatile("" : "+s"(stride_ld), "+s"(k0));
    float inv[K::Q];
    for (int q = 0; q < K::Q; ++q)
      inv[q] = __builtin_bit_cast(float,
                                  0x7F000000u - __builtin_bit_cast(uint32_t, a.scale[k0 + q]));
    load_node<K, P, NT_LD>(A, stride_ld, n, o, inv, f);
    (void)collide_fused<K>(f, p);
    uint64_t stride_st = a.stride;
    uint32_t k1 = 0;
    asm volatile("" : "+s"(stride_st), "+s"(k1) : "v"(f[0]));
    store_node<K, P, NT_ST>(A, stride_st, n, o, a.scale + k1, f);
  } else {
    load_node<K, P, NT_LD>(A, a.stride, n, o, a.inv_scale, f);
    (void)collide_fused<K>(f, p);
    store_node<K, P, NT_ST>(A, a.stride, n, o, a.scale, f);
  }
}

// Launch 1 of a pair.  P is the parity of step t.
template <class K, int P, bool GEOM>
__global__ void __launch_bounds__(kBlock) k_pair(PairArgs pa) {
  __shared__ int sh[2];
  const StepArgs& a = pa.a;
  const Dom& d = a.d;
  const int S = pa.stripes, nz = d.nz, chunks = a.chunks, G = pa.group;
  unsigned* const err = pa.sync;
  const unsigned xcc = xcc_id();
  const Params<typename K::R> prm = params_of<K>(a);
  typename K::T* const A = static_cast<typename K::T*>(a.f);
  int cur = -1;        // stripe this block works on (thread 0's view)
  int scan = 0;        // next candidate stripe to look at (thread 0)
  for (;;) {
    if (threadIdx.x == 0) {
      int ticket = -1, stripe = -1;
      for (;;) {
        if (cur >= 0) {
          const int rows = pa.y0[cur + 1] - pa.y0[cur];
          const unsigned total =
              pair_tickets(nz, pair_per_ticket(pair_tiles_a(rows, chunks), G),
                           pair_per_ticket(pair_tiles_b(rows, chunks), G));
          const unsigned t = __hip_atomic_fetch_add(pa.sync + pair_ticket_word(nz, cur), 1u,
                                                    __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if (t < total) {
            ticket = int(t);
            stripe = cur;
            break;
          }
          cur = -1;
        }
        // next stripe: first those congruent to this XCD, then any unclaimed
        bool found = false;
        while (scan < 2 * S && !found) {
          const int k = scan++;
          int c;
          if (k < S) {
            c = int(xcc % 8u) + 8 * k;
            if (c >= S) {
              scan = S;  // no more preferred stripes
              continue;
            }
          } else {
            c = k - S;
          }
          unsigned* own = pa.sync + pair_owner_word(nz, c);
          unsigned o = __hip_atomic_load(own, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if (o == 0u) {
            unsigned expect = 0u;
            __hip_atomic_compare_exchange_strong(own, &expect, xcc + 1u, __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            o = expect == 0u ? xcc + 1u : expect;
          }
          if (o == xcc + 1u) {
            cur = c;
            found = true;
          }
        }
        if (!found) break;
      }
      sh[0] = ticket;
      sh[1] = stripe;
    }
    __syncthreads();
    const int ticket = sh[0], stripe = sh[1];
    __syncthreads();
    if (stripe < 0) return;

    const int y0 = pa.y0[stripe], rows = pa.y0[stripe + 1] - y0;
    const unsigned na = pair_tiles_a(rows, chunks), nb = pair_tiles_b(rows, chunks);
    const unsigned ta = pair_per_ticket(na, G), tb = pair_per_ticket(nb, G);
    const PairTile pt = pair_decode(unsigned(ticket), nz, pa.lag, ta, tb);
    // tiles [first, last) of the plane
    const unsigned first = pt.idx * unsigned(G);
    if (pt.kind == 0) {
      // step t: inputs come from the previous launch; stores stay in cache
      const unsigned last = first + unsigned(G) < na ? first + unsigned(G) : na;
      for (unsigned tile = first; tile < last; ++tile) {
        const int row = int(tile / unsigned(chunks));
        const int x = int(tile - unsigned(row) * unsigned(chunks)) * kBlock + int(threadIdx.x);
        if (x < d.nx) node_update_hinted<K, P, GEOM, true, false>(A, a, x, y0 + row, pt.z, prm);
      }
      __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0): this wave's stores are in L2
      __syncthreads();
      if (threadIdx.x == 0)
        __hip_atomic_fetch_add(pa.sync + pair_done_word(nz, stripe, pt.z), 1u, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
    } else {
      if (threadIdx.x == 0) {
        unsigned spins = 0;
        bool ok = true;
        for (int dz = 1; dz >= -1 && ok; --dz) {
          const unsigned* c = pa.sync + pair_done_word(nz, stripe, pt.z + dz);
          while (__hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <
                 ta + pa.wait_extra) {
            // give up after the limit, or at once when another block has
            // already given up: the launch then drains in about one limit
            if (++spins >= pa.spin_limit ||
                ((spins & 63u) == 0u &&
                 __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u)) {
              __hip_atomic_fetch_or(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              ok = false;
              break;
            }
            __builtin_amdgcn_s_sleep(8);
          }
        }
        sh[0] = ok ? 1 : 0;
      }
      __syncthreads();
      const bool ok = sh[0] != 0;
      __syncthreads();
      // L1 invalidate: the step-t values of neighbouring tiles were written
      // through other CUs' L1s into this XCD's L2
      asm volatile("buffer_inv sc0" ::: "memory");
      if (ok) {
        const unsigned last = first + unsigned(G) < nb ? first + unsigned(G) : nb;
        for (unsigned tile = first; tile < last; ++tile) {
          const int row = int(tile / unsigned(chunks));
          const int x = int(tile - unsigned(row) * unsigned(chunks)) * kBlock + int(threadIdx.x);
          if (x < d.nx)
            node_update_hinted<K, 1 - P, GEOM, false, true>(A, a, x, y0 + 1 + row, pt.z, prm);
        }
      }
    }
  }
}

// Launch 2 of a pair: step t+1 of the deferred rows.  It also clears the
// synchronisation area for the next pair (launch 1 has finished) and moves a
// recorded failure into the host word the solver checks after synchronising.
template <class K, int P, bool GEOM>
__global__ void __launch_bounds__(kBlock) k_pair_tail(PairArgs pa, unsigned nclear,
                                                      unsigned* host_err) {
  const StepArgs& a = pa.a;
  const Dom& d = a.d;
  const unsigned g = blockIdx.x * kBlock + threadIdx.x;
  if (g == 0u) {
    const unsigned e = pa.sync[0];
    if (e) {
      if (host_err) *host_err = e;
      pa.sync[0] = 0u;
    }
  } else if (g < nclear) {
    pa.sync[g] = 0u;
  }
  const unsigned r = blockIdx.x / unsigned(a.chunks);
  const int x = int(blockIdx.x - r * unsigned(a.chunks)) * kBlock + int(threadIdx.x);
  if (r >= pair_tail_rows(d.ny, d.nz, pa.stripes) || x >= d.nx) return;
  int y, z;
  pair_tail_row(r, d.ny, d.nz, pa.stripes, pa.y0, y, z);
  const Params<typename K::R> prm = params_of<K>(a);
  node_update_hinted<K, 1 - P, GEOM, kNtLocal, kNtLocal>(static_cast<typename K::T*>(a.f), a,
                                                          x, y, z, prm);
}

}  // namespace m3lb
