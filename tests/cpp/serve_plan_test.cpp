// Host-side checks of the batched server's planning code and of the exact
// divider the batched kernel decodes node coordinates with.  Built with g++
// by tests/test_serve_cpu.py; prints "ok" and exits 0 when every check holds.
#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>
#include <set>
#include <vector>

#include "../../m3lb_amd/csrc/fastdiv.hpp"
#include "../../m3lb_amd/csrc/serve_plan.hpp"

static int failures = 0;
#define CHECK(c)                                                   \
  do {                                                             \
    if (!(c)) {                                                    \
      std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #c);      \
      if (++failures > 20) std::exit(1);                           \
    }                                                              \
  } while (0)

static void fastdiv_checks() {
  std::mt19937_64 rng(7);
  // every divisor up to 4096 against dividends near multiples, small values,
  // and the top of the 32-bit range
  for (uint32_t d = 1; d <= 4096; ++d) {
    const m3lb::FastDiv f = m3lb::make_fastdiv(d);
    for (uint32_t n = 0; n < 3 * d + 5; ++n) CHECK(m3lb::fd_div(f, n) == n / d);
    for (uint32_t k = 1; k < 64; ++k) {
      const uint64_t m = uint64_t(k) * 65599u * d;
      for (int e = -1; e <= 1; ++e) {
        const uint64_t n = m + uint64_t(int64_t(e));
        if (n <= 0xFFFFFFFFull) CHECK(m3lb::fd_div(f, uint32_t(n)) == uint32_t(n) / d);
      }
    }
    for (uint32_t n : {0xFFFFFFFFu, 0xFFFFFFFEu, 0x80000000u, 0x7FFFFFFFu})
      CHECK(m3lb::fd_div(f, n) == n / d);
  }
  // large and random divisors (box extents and xy-plane sizes)
  for (int t = 0; t < 200000; ++t) {
    const uint32_t d = uint32_t(rng() % 0xFFFFFFFFull) + 1;
    const uint32_t n = uint32_t(rng());
    CHECK(m3lb::fd_div(m3lb::make_fastdiv(d), n) == n / d);
  }
  // exhaustive over every node index of a 192^3 box for the divisors the
  // batched kernel uses (nx, ny, nx*ny)
  for (uint32_t d : {32u, 33u, 63u, 100u, 191u, 192u, 33u * 191u, 192u * 192u}) {
    const m3lb::FastDiv f = m3lb::make_fastdiv(d);
    for (uint32_t n = 0; n < 192u * 192u * 192u; ++n)
      if (m3lb::fd_div(f, n) != n / d) {
        CHECK(false);
        break;
      }
  }
}

struct Range {
  uint64_t off, len;
};

static void arena_checks() {
  // random alloc/free against a byte-granular model: ranges never overlap,
  // stay inside the capacity, and freeing everything leaves one free range
  const uint64_t cap = 1 << 24;
  m3lb::Arena a(cap);
  CHECK(a.capacity() == cap && a.used() == 0 && a.fragments() == 1);
  std::mt19937_64 rng(11);
  std::vector<Range> live;
  for (int it = 0; it < 20000; ++it) {
    if (live.empty() || rng() % 3 != 0) {
      const uint64_t want = 1 + rng() % (cap / 64);
      const uint64_t off = a.alloc(want);
      if (off == m3lb::Arena::kNone) {
        CHECK(a.largest_free() < m3lb::Arena::round(want));
        continue;
      }
      const uint64_t len = m3lb::Arena::round(want);
      CHECK(off % m3lb::Arena::kAlign == 0);
      CHECK(off + len <= cap);
      for (const Range& r : live) CHECK(off + len <= r.off || r.off + r.len <= off);
      live.push_back({off, want});
      live.back().len = len;
    } else {
      const size_t k = rng() % live.size();
      a.release(live[k].off, live[k].len);
      live.erase(live.begin() + long(k));
    }
    uint64_t sum = 0;
    for (const Range& r : live) sum += r.len;
    CHECK(a.used() == sum);
    CHECK(a.peak() >= a.used());
  }
  for (const Range& r : live) a.release(r.off, r.len);
  CHECK(a.used() == 0);
  CHECK(a.fragments() == 1);
  CHECK(a.largest_free() == cap);
  // first fit, lowest address, and coalescing on both sides
  m3lb::Arena b(4096);
  const uint64_t x = b.alloc(1000), y = b.alloc(1000), z = b.alloc(1000);
  CHECK(x == 0 && y == 1024 && z == 2048);
  CHECK(b.alloc(2000) == m3lb::Arena::kNone);
  b.release(y, 1000);
  CHECK(b.fragments() == 2);
  CHECK(b.alloc(1) == 1024);  // reuses the hole
  b.release(1024, 1);
  b.release(x, 1000);
  b.release(z, 1000);
  CHECK(b.fragments() == 1 && b.largest_free() == 4096);
}

static void admission_checks() {
  // strict FIFO when nothing may pass the head
  {
    m3lb::Admission q(0);
    for (int i = 1; i <= 5; ++i) q.push(i);
    std::set<int64_t> fits = {1, 2, 4, 5};  // 3 does not fit
    auto got = q.run(10, [&](int64_t id) { return fits.count(id) > 0; });
    CHECK(got.size() == 2 && got[0] == 1 && got[1] == 2);
    CHECK(q.size() == 3);
  }
  // bounded backfill: two may pass a head that does not fit, then none
  {
    m3lb::Admission q(2);
    for (int i = 1; i <= 6; ++i) q.push(i);
    auto got = q.run(10, [&](int64_t id) { return id != 1; });
    CHECK(got.size() == 2 && got[0] == 2 && got[1] == 3);
    CHECK(q.bypassed() == 2);
    got = q.run(10, [&](int64_t id) { return id != 1; });
    CHECK(got.empty());
    // the head fits now: it goes first and the count resets
    got = q.run(10, [&](int64_t) { return true; });
    CHECK(got.size() == 4 && got[0] == 1 && got[1] == 4);
    CHECK(q.bypassed() == 0 && q.size() == 0);
  }
  // slots bound the number admitted; arrival order is kept
  {
    m3lb::Admission q(64);
    for (int i = 1; i <= 6; ++i) q.push(i);
    auto got = q.run(3, [&](int64_t) { return true; });
    CHECK(got.size() == 3 && got[0] == 1 && got[2] == 3);
    CHECK(q.remove(5) && !q.remove(5));
    got = q.run(3, [&](int64_t) { return true; });
    CHECK(got.size() == 2 && got[0] == 4 && got[1] == 6);
  }
}

// Backfill that never delays the head: fit_horizon against a step-by-step
// replay, may_pass in Admission::run, and a lockstep simulation in which every
// head is admitted no later than the horizon it had when it first waited.
static void horizon_checks() {
  using m3lb::Holding;
  // fits now; node cap; ties leave together; bytes; empty server
  CHECK(m3lb::fit_horizon({{5, 100, 256}}, 0, 1000, 1 << 20, 100, 256) == 0);
  const std::vector<Holding> r = {{5, 100, 256}, {3, 200, 256}, {8, 50, 256}, {3, 10, 256}};
  CHECK(m3lb::fit_horizon(r, 0, 300, 1 << 20, 150, 256) == 3);
  CHECK(m3lb::fit_horizon(r, 0, 300, 1 << 20, 250, 256) == 5);
  CHECK(m3lb::fit_horizon(r, 0, 300, 1 << 20, 10000, 256) == 8);  // empty server
  CHECK(m3lb::fit_horizon(r, 0, 0, 4 * 256 + 256, 1, 512) == 3);   // bytes only
  CHECK(m3lb::fit_horizon(r, 1 << 20, 0, 1 << 20, 1, 256) == m3lb::kNoHorizon);
  // against a replay: step until the request fits
  std::mt19937_64 rng(11);
  for (int t = 0; t < 20000; ++t) {
    std::vector<Holding> res(rng() % 7);
    for (auto& h : res) h = {int64_t(1 + rng() % 20), 1 + rng() % 100, 256 * (1 + rng() % 8)};
    const uint64_t cap = rng() % 3 ? 50 + rng() % 300 : 0;
    const uint64_t budget = 256 * (8 + rng() % 40), fixed = 256 * (rng() % 4);
    const uint64_t nn = 1 + rng() % 150, nb = 256 * (1 + rng() % 8);
    int64_t want = m3lb::kNoHorizon;
    for (int64_t step = 0; step <= 21; ++step) {
      uint64_t nodes = 0, bytes = fixed;
      for (auto& h : res)
        if (h.steps_left > step) {
          nodes += h.nodes;
          bytes += h.bytes;
        }
      if ((cap == 0 || nodes == 0 || nodes + nn <= cap) && bytes + nb <= budget) {
        want = step;
        break;
      }
    }
    CHECK(m3lb::fit_horizon(res, fixed, cap, budget, nn, nb) == want);
  }
  // may_pass: behind a head that does not fit, only the allowed ones are tried
  {
    m3lb::Admission q(64);
    for (int i = 1; i <= 6; ++i) q.push(i);
    std::vector<int64_t> asked;
    auto got = q.run(
        10, [&](int64_t id) { return id != 1; },
        [&](int64_t head, int64_t id) {
          CHECK(head == 1);
          asked.push_back(id);
          return id % 2 == 0;
        });
    CHECK((got == std::vector<int64_t>{2, 4, 6}));
    CHECK((asked == std::vector<int64_t>{2, 3, 4, 5, 6}));
    CHECK(q.size() == 3 && q.bypassed() == 3);
  }
  // lockstep simulation under a node cap
  for (int trial = 0; trial < 200; ++trial) {
    struct Req {
      uint64_t nodes;
      int64_t steps;
    };
    std::vector<Req> reqs(60);
    for (auto& q : reqs) q = {1 + rng() % 120, int64_t(1 + rng() % 40)};
    const uint64_t cap = 150 + rng() % 200;
    m3lb::Admission adm(1000);
    std::map<int64_t, int64_t> left;  // resident id -> steps left
    std::map<int64_t, int64_t> due;   // waiting head -> latest admission step
    size_t next = 0, finished = 0;
    for (int64_t step = 0; finished < reqs.size() && step < 100000; ++step) {
      while (next < reqs.size() && next <= size_t(step / 3)) adm.push(int64_t(next++));
      uint64_t nodes = 0;
      for (auto& kv : left) nodes += reqs[size_t(kv.first)].nodes;
      auto holdings = [&]() {
        std::vector<Holding> h;
        for (auto& kv : left) h.push_back({kv.second, reqs[size_t(kv.first)].nodes, 0});
        return h;
      };
      int64_t head = -1, horizon = 0;
      auto got = adm.run(
          1000,
          [&](int64_t id) {
            const uint64_t bn = reqs[size_t(id)].nodes;
            if (nodes > 0 && nodes + bn > cap) return false;
            nodes += bn;
            left[id] = reqs[size_t(id)].steps;
            return true;
          },
          [&](int64_t h, int64_t id) {
            if (h != head) {
              head = h;
              horizon = m3lb::fit_horizon(holdings(), 0, cap, ~uint64_t(0) >> 1,
                                          reqs[size_t(h)].nodes, 0);
              if (!due.count(h)) due[h] = step + horizon;
            }
            return reqs[size_t(id)].steps <= horizon;
          });
      for (int64_t id : got) {
        if (due.count(id)) CHECK(step <= due[id]);
      }
      // one step of every resident box
      for (auto it = left.begin(); it != left.end();) {
        if (--it->second == 0) {
          it = left.erase(it);
          ++finished;
        } else {
          ++it;
        }
      }
    }
    CHECK(finished == reqs.size());
  }
}

// The chunk layout: descending steps (stable), prefix sums, and at every step
// the stepping boxes are exactly a prefix whose tiles hdr records.
static void pack_checks() {
  std::mt19937_64 rng(5);
  for (int t = 0; t < 5000; ++t) {
    const int K = 1 + int(rng() % 12);
    const size_t n = rng() % 20;
    std::vector<unsigned> tiles(n);
    std::vector<int> steps(n);
    for (size_t k = 0; k < n; ++k) {
      tiles[k] = unsigned(rng() % 50);
      steps[k] = 1 + int(rng() % unsigned(K));
    }
    std::vector<unsigned> hdr(m3lb::pack_words(n, K), 0xDEADu);
    std::vector<size_t> order;
    m3lb::pack_chunk(tiles, steps, K, order, hdr.data());
    CHECK(order.size() == n && hdr[0] == n);
    std::set<size_t> seen(order.begin(), order.end());
    CHECK(seen.size() == n);
    unsigned total = 0;
    for (size_t k = 0; k < n; ++k) {
      CHECK(hdr[2 + k] == total);
      total += tiles[order[k]];
      if (k > 0) {
        CHECK(steps[order[k - 1]] >= steps[order[k]]);
        if (steps[order[k - 1]] == steps[order[k]]) CHECK(order[k - 1] < order[k]);
      }
    }
    CHECK(hdr[1] == total && hdr[2 + n] == total);
    for (int j = 0; j < K; ++j) {
      unsigned want = 0;
      for (size_t k = 0; k < n; ++k)
        if (steps[k] > j) want += tiles[k];
      const unsigned act = hdr[3 + n + size_t(j)];
      CHECK(act == want);
      // the tiles [0, act) are exactly those of the boxes stepping at j
      for (size_t k = 0; k < n; ++k) {
        const bool in = hdr[2 + k] < act || (tiles[order[k]] == 0 && hdr[2 + k] <= act &&
                                             steps[order[k]] > j);
        if (tiles[order[k]]) CHECK(in == (steps[order[k]] > j));
      }
    }
  }
}

// A seeding failure during admission: the failing box and every box admitted
// after it are handed back unseeded, in order, and the error comes back; the
// boxes before it stay.  With an arena behind it, the memory of the handed-back
// boxes is free again.
static void seed_failure_checks() {
  m3lb::Arena a(1 << 20);
  std::vector<int64_t> ids = {11, 12, 13, 14, 15};
  std::map<int64_t, uint64_t> off;
  for (int64_t id : ids) off[id] = a.alloc(4096);
  std::vector<int64_t> seeded, undone;
  const int rc = m3lb::seed_admitted(
      ids,
      [&](int64_t id) {
        if (id == 13) return -3;  // the injected failure
        seeded.push_back(id);
        return 0;
      },
      [&](int64_t id) {
        undone.push_back(id);
        a.release(off[id], 4096);
      });
  CHECK(rc == -3);
  CHECK((seeded == std::vector<int64_t>{11, 12}));
  CHECK((undone == std::vector<int64_t>{13, 14, 15}));
  CHECK(a.used() == 2 * 4096);
  // no failure: every box seeded, nothing undone
  seeded.clear();
  undone.clear();
  CHECK(m3lb::seed_admitted(ids, [&](int64_t id) { seeded.push_back(id); return 0; },
                            [&](int64_t id) { undone.push_back(id); }) == 0);
  CHECK(seeded == ids && undone.empty());
  // the first box fails: all of them are handed back
  undone.clear();
  CHECK(m3lb::seed_admitted(ids, [&](int64_t) { return -1; },
                            [&](int64_t id) { undone.push_back(id); }) == -1);
  CHECK(undone == ids);
}

int main() {
  fastdiv_checks();
  arena_checks();
  admission_checks();
  horizon_checks();
  pack_checks();
  seed_failure_checks();
  if (failures) return 1;
  std::printf("ok\n");
  return 0;
}
