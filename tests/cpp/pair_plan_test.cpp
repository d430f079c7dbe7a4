// Host-side checks of the two-step pass's schedule (m3lb_amd/csrc/pair_plan.hpp)
// on small D3Q27 boxes, with the esoteric-pull slot sets of every node
// (EsotericPull.hpp:5-25):
//   1. the tickets of every stripe hand out each step-t tile once and each
//      inner-row step-t+1 tile once, and the step-t+1 tiles of launch 1 and the
//      rows of launch 2 together cover every node once;
//   2. every step-t tile a step-t+1 tile waits on has an earlier ticket (so
//      the earliest unfinished ticket can always proceed);
//   3. every slot's step-t accessor and its step-t+1 accessor, when both exist
//      and the latter runs in launch 1, are in the same stripe and the latter waits for the
//      former's plane: no race within a launch, and each slot is touched by
//      one stripe (one XCD) only in launch 1;
//   4. running the tiles in ticket order, stripes interleaved at random, then
//      the deferred rows, leaves the same array as two whole steps.
// Built with g++ by tests/test_pair_cpu.py; prints "ok" and exits 0.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../../m3lb_amd/csrc/pair_plan.hpp"

static int failures = 0;
#define CHECK(c)                                                   \
  do {                                                             \
    if (!(c)) {                                                    \
      std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #c);      \
      if (++failures > 20) std::exit(1);                           \
    }                                                              \
  } while (0)

// D3Q27 in esoteric-pull order (src/lattice/Lattices.hpp:125-159)
static const int CX[27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1,
                           0, 0, 0, 0, 1, -1, 1, -1, 1, -1, -1, 1};
static const int CY[27] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0,
                           1, -1, 1, -1, 1, -1, 1, -1, -1, 1, 1, -1};
static const int CZ[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1,
                           1, -1, -1, 1, 1, -1, -1, 1, 1, -1, 1, -1};

struct Box {
  int nx, ny, nz, per[3], h[3], s[3];
  int64_t np;
  Box(int x, int y, int z, int px, int py, int pz) : nx(x), ny(y), nz(z), per{px, py, pz} {
    const int n[3] = {x, y, z};
    for (int a = 0; a < 3; ++a) {
      h[a] = (n[a] <= 1 || per[a]) ? 0 : 1;  // Domain.hpp:104-108
      s[a] = n[a] + 2 * h[a];
    }
    np = int64_t(s[0]) * s[1] * s[2];
  }
  int64_t id(int x, int y, int z) const {
    return (int64_t(z + h[2]) * s[1] + (y + h[1])) * s[0] + (x + h[0]);
  }
  // neighbour of an interior node: periodic wrap inside the interior, else
  // the halo node
  int64_t nb(int x, int y, int z, int i) const {
    int c[3] = {x + CX[i], y + CY[i], z + CZ[i]};
    const int n[3] = {nx, ny, nz};
    for (int a = 0; a < 3; ++a)
      if (per[a]) c[a] = (c[a] + n[a]) % n[a];
    return id(c[0], c[1], c[2]);
  }
  // the 27 slots (q * np + m) node (x, y, z) reads and writes at parity P
  void slots(int x, int y, int z, int P, int64_t* out) const {
    const int64_t n = id(x, y, z);
    out[0] = n;
    for (int i = 1; i < 27; i += 2) {
      const int64_t m = nb(x, y, z, i);
      if (P == 0) {
        out[i] = int64_t(i) * np + n;
        out[i + 1] = int64_t(i + 1) * np + m;
      } else {
        out[i] = int64_t(i + 1) * np + n;
        out[i + 1] = int64_t(i) * np + m;
      }
    }
  }
};

// The node update used for the data check: loads through the slot map,
// mixes the 27 values with the node index (any deterministic function of the
// inputs will do), stores crossed into the same slots.
static void update(const Box& b, std::vector<uint64_t>& A, int x, int y, int z, int P) {
  int64_t sl[27];
  b.slots(x, y, z, P, sl);
  uint64_t f[27];
  // load: f_i from its slot (EsotericPull.hpp load_pair / store_pair)
  for (int q = 0; q < 27; ++q) f[q] = A[size_t(sl[q])];
  uint64_t h = uint64_t(b.id(x, y, z)) * 0x9E3779B97F4A7C15ull;
  uint64_t g[27];
  for (int q = 0; q < 27; ++q) {
    h ^= f[q] + 0x632BE59BD9B4E019ull + (h << 6) + (h >> 2);
    g[q] = h ^ f[(q * 5 + 3) % 27];
  }
  // store crossed: the pair (i, i+1) trades slots
  A[size_t(sl[0])] = g[0];
  for (int i = 1; i < 27; i += 2) {
    A[size_t(sl[i])] = g[i + 1];
    A[size_t(sl[i + 1])] = g[i];
  }
}

struct Tile {
  int kind, stripe, z, row, chunk;  // row: interior y
};

static void run_case(int nx, int ny, int nz, int px, int py, int pz, int S, int lag, int group,
                     uint64_t seed) {
  const Box b(nx, ny, nz, px, py, pz);
  const int chunks = (nx + 255) / 256;
  int y0[m3lb::kPairMaxStripes + 1];
  for (int s = 0; s <= S; ++s) y0[s] = int(int64_t(s) * ny / S);
  // --- 1. coverage -------------------------------------------------------
  std::vector<std::vector<Tile>> order(S);
  std::vector<int> covA(size_t(ny) * nz, 0), covB(size_t(ny) * nz, 0);
  std::vector<std::vector<int>> ticket_of(S);  // ticket of every entry of order[s]
  for (int s = 0; s < S; ++s) {
    const int rows = y0[s + 1] - y0[s];
    const unsigned na = m3lb::pair_tiles_a(rows, chunks), nb = m3lb::pair_tiles_b(rows, chunks);
    const unsigned ta = m3lb::pair_per_ticket(na, group), tb = m3lb::pair_per_ticket(nb, group);
    const unsigned total = m3lb::pair_tickets(nz, ta, tb);
    for (unsigned t = 0; t < total; ++t) {
      const m3lb::PairTile pt = m3lb::pair_decode(t, nz, lag, ta, tb);
      CHECK(pt.z >= 0 && pt.z < nz);
      CHECK(pt.idx < (pt.kind ? tb : ta));
      if (pt.kind == 1) CHECK(pt.z >= 1 && pt.z <= nz - 2);
      const unsigned n = pt.kind ? nb : na;
      for (unsigned tile = pt.idx * unsigned(group); tile < n && tile < (pt.idx + 1) * unsigned(group);
           ++tile) {
        Tile T;
        T.kind = pt.kind;
        T.stripe = s;
        T.z = pt.z;
        const int r = int(tile / unsigned(chunks));
        T.chunk = int(tile % unsigned(chunks));
        T.row = y0[s] + (pt.kind ? 1 + r : r);
        if (pt.kind == 0) covA[size_t(T.z) * ny + T.row] += 1;
        else covB[size_t(T.z) * ny + T.row] += 1;
        order[s].push_back(T);
        ticket_of[s].push_back(int(t));
      }
    }
  }
  // every counter on its own 128-byte line, inside the area
  {
    std::vector<unsigned> w = {0u};
    for (int s = 0; s < S; ++s) {
      w.push_back(m3lb::pair_ticket_word(nz, s));
      w.push_back(m3lb::pair_owner_word(nz, s));
      for (int z = 0; z < nz; ++z) w.push_back(m3lb::pair_done_word(nz, s, z));
    }
    std::vector<char> line(m3lb::pair_sync_words(nz, S) / 32 + 1, 0);
    for (unsigned x : w) {
      CHECK(x % 32u == 0u && x < m3lb::pair_sync_words(nz, S));
      CHECK(!line[x / 32u]);
      line[x / 32u] = 1;
    }
  }
  const unsigned tail = m3lb::pair_tail_rows(ny, nz, S);
  std::vector<Tile> tails;
  for (unsigned r = 0; r < tail; ++r) {
    int y, z;
    m3lb::pair_tail_row(r, ny, nz, S, y0, y, z);
    CHECK(y >= 0 && y < ny && z >= 0 && z < nz);
    covB[size_t(z) * ny + y] += chunks;
    for (int c = 0; c < chunks; ++c) tails.push_back({2, -1, z, y, c});
  }
  for (size_t k = 0; k < covA.size(); ++k) {
    CHECK(covA[k] == chunks);
    CHECK(covB[k] == chunks);
  }
  // --- 2. waits point to earlier tickets ----------------------------------
  for (int s = 0; s < S; ++s) {
    for (size_t t = 0; t < order[s].size(); ++t) {
      const Tile& T = order[s][t];
      if (T.kind != 1) continue;
      for (int dz = -1; dz <= 1; ++dz) {
        // every A tile of the plane belongs to an earlier ticket
        int seen = 0;
        for (size_t u = 0; u < order[s].size(); ++u)
          if (order[s][u].kind == 0 && order[s][u].z == T.z + dz) {
            ++seen;
            CHECK(ticket_of[s][u] < ticket_of[s][t]);
          }
        CHECK(seen == int(m3lb::pair_tiles_a(y0[s + 1] - y0[s], chunks)));
      }
    }
  }
  // --- 3. slot accessors --------------------------------------------------
  // launch-1 accessor of every slot at each step: stripe, plane (-1: none)
  const int P = int(seed & 1u);
  std::vector<int> aStripe(size_t(27 * b.np), -1), aPlane(size_t(27 * b.np), -1);
  std::vector<int> bStripe(size_t(27 * b.np), -1), bPlane(size_t(27 * b.np), -1);
  int64_t sl[27];
  for (int s = 0; s < S; ++s)
    for (const Tile& T : order[s])
      for (int x = T.chunk * 256; x < nx && x < (T.chunk + 1) * 256; ++x) {
        b.slots(x, T.row, T.z, T.kind == 0 ? P : 1 - P, sl);
        for (int q = 0; q < 27; ++q) {
          std::vector<int>& st = T.kind == 0 ? aStripe : bStripe;
          std::vector<int>& pl = T.kind == 0 ? aPlane : bPlane;
          CHECK(st[size_t(sl[q])] == -1);  // one accessor per slot per step
          st[size_t(sl[q])] = s;
          pl[size_t(sl[q])] = T.z;
        }
      }
  // (a slot of a halo node that no interior node accesses at step t has no
  // step-t accessor: halo nodes never run, so nothing races on it)
  for (size_t k = 0; k < aStripe.size(); ++k) {
    if (bStripe[k] < 0 || aStripe[k] < 0) continue;
    CHECK(aStripe[k] == bStripe[k]);
    CHECK(aPlane[k] >= bPlane[k] - 1 && aPlane[k] <= bPlane[k] + 1);
  }
  // --- 4. data ------------------------------------------------------------
  std::mt19937_64 rng(seed);
  std::vector<uint64_t> A0(size_t(27 * b.np));
  for (auto& v : A0) v = rng();
  std::vector<uint64_t> ref = A0, got = A0;
  for (int step = 0; step < 2; ++step)
    for (int z = 0; z < nz; ++z)
      for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) update(b, ref, x, y, z, step == 0 ? P : 1 - P);
  std::vector<size_t> next(S, 0);
  for (;;) {
    std::vector<int> live;
    for (int s = 0; s < S; ++s)
      if (next[s] < order[s].size()) live.push_back(s);
    if (live.empty()) break;
    const int s = live[rng() % live.size()];
    const Tile& T = order[s][next[s]++];
    for (int x = T.chunk * 256; x < nx && x < (T.chunk + 1) * 256; ++x)
      update(b, got, x, T.row, T.z, T.kind == 0 ? P : 1 - P);
  }
  for (const Tile& T : tails)
    for (int x = T.chunk * 256; x < nx && x < (T.chunk + 1) * 256; ++x)
      update(b, got, x, T.row, T.z, 1 - P);
  CHECK(got == ref);
}

int main() {
  // periodic cubes and boxes, uneven stripes, halos on y and z, two chunks
  run_case(6, 24, 8, 1, 1, 1, 8, 2, 1, 1);
  run_case(6, 24, 8, 1, 1, 1, 8, 1, 1, 2);
  run_case(5, 17, 9, 1, 1, 1, 4, 3, 2, 3);
  run_case(7, 12, 6, 1, 0, 1, 3, 3, 3, 4);
  run_case(4, 13, 7, 1, 0, 0, 4, 2, 2, 5);
  run_case(300, 9, 5, 1, 1, 1, 3, 2, 1, 6);
  run_case(300, 13, 6, 1, 1, 1, 3, 2, 3, 9);
  run_case(9, 48, 11, 0, 1, 1, 16, 4, 4, 7);
  run_case(8, 30, 6, 1, 1, 0, 10, 3, 5, 8);
  if (failures) return 1;
  std::printf("ok\n");
  return 0;
}
