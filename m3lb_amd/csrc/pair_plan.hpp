// pair_plan.hpp -- the index arithmetic of the two-step pass (pair.hpp):
// ticket order within a stripe, the deferred rows of the second launch and
// the layout of the synchronisation words.  Plain C++ (host and device), so
// that the schedule can be checked on the host (tests/cpp/pair_plan_test.cpp).
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define M3LB_PLAN_HD __host__ __device__
#else
#define M3LB_PLAN_HD
#endif

namespace m3lb {

constexpr int kPairMaxStripes = 64;
constexpr unsigned kPairSpinLimit = 1u << 20;  // polls of one wait, each followed by s_sleep

// Words of the synchronisation area, each counter on its own 128-byte line
// (polls of one counter then never delay the atomics on another):
//   [0]                      error word
//   per stripe s, at 32 + s * pair_stripe_words(nz):
//     [0]       ticket counter
//     [32]      owning XCD + 1 (0: unclaimed)
//     [64 + 32z] step-t tickets of plane z finished
M3LB_PLAN_HD constexpr unsigned pair_stripe_words(int nz) { return 64u + 32u * unsigned(nz); }
M3LB_PLAN_HD constexpr unsigned pair_sync_words(int nz, int stripes) {
  return 32u + unsigned(stripes) * pair_stripe_words(nz);
}
M3LB_PLAN_HD constexpr unsigned pair_ticket_word(int nz, int s) {
  return 32u + unsigned(s) * pair_stripe_words(nz);
}
M3LB_PLAN_HD constexpr unsigned pair_owner_word(int nz, int s) {
  return pair_ticket_word(nz, s) + 32u;
}
M3LB_PLAN_HD constexpr unsigned pair_done_word(int nz, int s, int z) {
  return pair_ticket_word(nz, s) + 64u + 32u * unsigned(z);
}

// Tickets of one plane of a stripe with R rows, each row cut into `chunks`
// tiles of 256 nodes and `group` consecutive tiles per ticket: step t (A)
// covers every row, step t+1 (B) the inner rows.
M3LB_PLAN_HD inline unsigned pair_tiles_a(int rows, int chunks) {
  return unsigned(rows) * unsigned(chunks);
}
M3LB_PLAN_HD inline unsigned pair_tiles_b(int rows, int chunks) {
  return rows > 2 ? unsigned(rows - 2) * unsigned(chunks) : 0u;
}
M3LB_PLAN_HD inline unsigned pair_per_ticket(unsigned tiles, int group) {
  return (tiles + unsigned(group) - 1u) / unsigned(group);
}
M3LB_PLAN_HD inline unsigned pair_tickets(int nz, unsigned ta, unsigned tb) {
  return ta * unsigned(nz) + tb * unsigned(nz - 2);
}

// Decodes ticket t of a stripe, with ta and tb tickets per plane: kind
// (0 = A, 1 = B), plane z and ticket index within the plane.  Valid for
// 1 <= lag <= nz - 3 and tb >= 1.
struct PairTile {
  int kind, z;
  unsigned idx;  // ticket within the plane
};
M3LB_PLAN_HD inline PairTile pair_decode(unsigned t, int nz, int lag, unsigned ta, unsigned tb) {
  PairTile r;
  const unsigned head = ta * unsigned(lag + 1);
  if (t < head) {
    r.kind = 0;
    r.z = int(t / ta);
    r.idx = t - unsigned(r.z) * ta;
    return r;
  }
  t -= head;
  const unsigned per = ta + tb, mid = unsigned(nz - 1 - lag);
  const unsigned k = t / per;
  if (k < mid) {
    const unsigned w = t - k * per;
    if (w < ta) {
      r.kind = 0;
      r.z = lag + 1 + int(k);
      r.idx = w;
    } else {
      r.kind = 1;
      r.z = 1 + int(k);
      r.idx = w - ta;
    }
    return r;
  }
  t -= mid * per;
  r.kind = 1;
  const unsigned j = t / tb;
  r.z = nz - lag + int(j);
  r.idx = t - j * tb;
  return r;
}

// Launch 2: the deferred step-t+1 nodes, as tiles of 256.  Planes 0 and nz-1
// whole (2 * ny rows), then for planes 1 .. nz-2 the first and last row of
// every stripe.  Returns the row (interior y) and plane of tile u's row.
M3LB_PLAN_HD inline unsigned pair_tail_rows(int ny, int nz, int stripes) {
  return unsigned(2 * ny) + unsigned(nz - 2) * unsigned(2 * stripes);
}
M3LB_PLAN_HD inline void pair_tail_row(unsigned r, int ny, int nz, int stripes, const int* y0,
                                  int& y, int& z) {
  const unsigned edge = unsigned(2 * ny);
  if (r < edge) {
    z = r < unsigned(ny) ? 0 : nz - 1;
    y = int(r % unsigned(ny));
    return;
  }
  r -= edge;
  const unsigned per = unsigned(2 * stripes);
  z = 1 + int(r / per);
  const unsigned j = r % per;
  const int s = int(j >> 1);
  y = (j & 1u) ? y0[s + 1] - 1 : y0[s];
}

}  // namespace m3lb
