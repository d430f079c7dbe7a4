// Host-side checks of the batched server's planning code and of the exact
// divider the batched kernel decodes node coordinates with.  Built with g++
// by tests/test_serve_cpu.py; prints "ok" and exits 0 when every check holds.
#include <cstdio>
#include <cstdlib>
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
  seed_failure_checks();
  if (failures) return 1;
  std::printf("ok\n");
  return 0;
}
