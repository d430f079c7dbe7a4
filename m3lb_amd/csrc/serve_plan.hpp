// serve_plan.hpp -- host-side bookkeeping of the batched server, plain C++ so
// that it can be tested without a GPU:
//   Arena      first-fit allocator over one device allocation (the HBM budget),
//              256-byte granularity, freed ranges coalesce with their neighbours;
//   Admission  the order in which queued requests are admitted: first come,
//              first served, with a bounded number of smaller requests allowed
//              past a queue head that does not fit yet (backfill);
//   seed_admitted  seeds the boxes of one admission pass and, at the first
//              failure, hands that box and every later one back unseeded.
// The reference has no serving layer; nothing here has a counterpart in it.
#pragma once

#include <cstdint>
#include <deque>
#include <iterator>
#include <map>
#include <vector>

namespace m3lb {

class Arena {
 public:
  static constexpr uint64_t kAlign = 256;
  static constexpr uint64_t kNone = ~uint64_t(0);

  explicit Arena(uint64_t capacity = 0) { reset(capacity); }

  void reset(uint64_t capacity) {
    cap_ = capacity / kAlign * kAlign;
    free_.clear();
    if (cap_) free_[0] = cap_;
    used_ = peak_ = 0;
  }

  static uint64_t round(uint64_t bytes) {
    const uint64_t b = (bytes + kAlign - 1) / kAlign * kAlign;
    return b ? b : kAlign;
  }

  // Offset of a free range of `bytes` (rounded up), lowest address first, or
  // kNone when no free range is large enough.
  uint64_t alloc(uint64_t bytes) {
    bytes = round(bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      const uint64_t off = it->first, rest = it->second - bytes;
      free_.erase(it);
      if (rest) free_[off + bytes] = rest;
      used_ += bytes;
      if (used_ > peak_) peak_ = used_;
      return off;
    }
    return kNone;
  }

  // Returns a range obtained from alloc() with the same size argument.
  void release(uint64_t off, uint64_t bytes) {
    bytes = round(bytes);
    used_ -= bytes;
    uint64_t start = off, len = bytes;
    auto next = free_.lower_bound(off);
    if (next != free_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == off) {
        start = prev->first;
        len += prev->second;
        free_.erase(prev);
      }
    }
    if (next != free_.end() && off + bytes == next->first) {
      len += next->second;
      free_.erase(next);
    }
    free_[start] = len;
  }

  uint64_t capacity() const { return cap_; }
  uint64_t used() const { return used_; }
  uint64_t peak() const { return peak_; }
  size_t fragments() const { return free_.size(); }
  uint64_t largest_free() const {
    uint64_t m = 0;
    for (const auto& kv : free_) m = kv.second > m ? kv.second : m;
    return m;
  }

 private:
  std::map<uint64_t, uint64_t> free_;  // offset -> length, disjoint, never adjacent
  uint64_t cap_ = 0, used_ = 0, peak_ = 0;
};

// Queue of request ids in arrival order.  run() walks it front to back and
// calls try_admit(id) for each while slots are free; try_admit reserves what
// the request needs and returns true, or returns false when it does not fit
// now.  A queue head that does not fit lets at most max_bypass later requests
// be admitted before it; after that nothing more is admitted until it fits.
class Admission {
 public:
  explicit Admission(int max_bypass = 64) : max_bypass_(max_bypass) {}

  void push(int64_t id) { q_.push_back(id); }

  bool remove(int64_t id) {
    for (auto it = q_.begin(); it != q_.end(); ++it)
      if (*it == id) {
        if (it == q_.begin()) bypass_ = 0;
        q_.erase(it);
        return true;
      }
    return false;
  }

  size_t size() const { return q_.size(); }
  int bypassed() const { return bypass_; }
  void set_max_bypass(int m) { max_bypass_ = m; }

  template <class TryAdmit>
  std::vector<int64_t> run(size_t slots_free, TryAdmit&& try_admit) {
    std::vector<int64_t> admitted;
    auto it = q_.begin();
    bool head = true;  // `it` is the current queue head
    while (it != q_.end() && admitted.size() < slots_free) {
      if (!head && bypass_ >= max_bypass_) break;
      if (try_admit(*it)) {
        admitted.push_back(*it);
        if (head) bypass_ = 0;
        else ++bypass_;
        it = q_.erase(it);
      } else {
        ++it;
        head = false;
      }
    }
    return admitted;
  }

 private:
  std::deque<int64_t> q_;
  int max_bypass_;
  int bypass_ = 0;  // admissions past the current head since it became head
};

// Seeds the boxes admitted in one pass, in admission order.  seed(id) returns
// 0 or an error code.  At the first failure undo(id) is called for that box and
// then for every later one (none of which has been seeded), so no box is left
// resident on memory it never initialised, and the error is returned.
template <class Seed, class Undo>
int seed_admitted(const std::vector<int64_t>& ids, Seed&& seed, Undo&& undo) {
  for (size_t k = 0; k < ids.size(); ++k) {
    const int rc = seed(ids[k]);
    if (rc != 0) {
      for (size_t j = k; j < ids.size(); ++j) undo(ids[j]);
      return rc;
    }
  }
  return 0;
}

}  // namespace m3lb
