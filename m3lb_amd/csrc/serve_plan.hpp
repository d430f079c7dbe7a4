// serve_plan.hpp -- host-side bookkeeping of the batched server, plain C++ so
// that it can be tested without a GPU:
//   Arena      first-fit allocator over one device allocation (the HBM budget),
//              256-byte granularity, freed ranges coalesce with their neighbours;
//   Admission  the order in which queued requests are admitted: first come,
//              first served; a later request may pass a queue head that does
//              not fit yet when it will have finished before the head can be
//              admitted (backfill that never delays the head), at most
//              max_bypass of them per head;
//   fit_horizon  when that head can be admitted, counted in steps;
//   pack_chunk  the packed tile list of one chunk, with the boxes that stop
//              early at its end, so that step j covers a prefix of it;
//   seed_admitted  seeds the boxes of one admission pass and, at the first
//              failure, hands that box and every later one back unseeded.
// The reference has no serving layer; nothing here has a counterpart in it.
#pragma once

#include <algorithm>
#include <cstdint>
#include <deque>
#include <iterator>
#include <limits>
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

// A resident box as admission sees it: steps not yet issued, interior nodes
// and arena bytes.
struct Holding {
  int64_t steps_left;
  uint64_t nodes, bytes;
};

constexpr int64_t kNoHorizon = std::numeric_limits<int64_t>::max();

// Every resident box advances one step per step of the server, and leaves
// (returning its nodes and bytes) once its steps run out.  Returns the number
// of steps after which a request of need_nodes nodes and need_bytes bytes
// fits: under the node cap (0: none; a request always fits when no nodes are
// resident) and within the budget, counting fixed_bytes held by no resident
// box (kept results).  0 when it fits now; kNoHorizon when the resident boxes
// leaving is not enough.  Bytes are counted in total, so an arena too
// fragmented for the request can still make it wait past the horizon.
inline int64_t fit_horizon(std::vector<Holding> res, uint64_t fixed_bytes, uint64_t node_cap,
                           uint64_t budget, uint64_t need_nodes, uint64_t need_bytes) {
  need_bytes = Arena::round(need_bytes);
  uint64_t nodes = 0, bytes = fixed_bytes;
  for (const Holding& h : res) {
    nodes += h.nodes;
    bytes += Arena::round(h.bytes);
  }
  auto fits = [&]() {
    const bool node_ok = node_cap == 0 || nodes == 0 || nodes + need_nodes <= node_cap;
    return node_ok && bytes + need_bytes <= budget;
  };
  if (fits()) return 0;
  std::sort(res.begin(), res.end(),
            [](const Holding& a, const Holding& b) { return a.steps_left < b.steps_left; });
  for (size_t k = 0; k < res.size();) {
    const int64_t s = res[k].steps_left;
    for (; k < res.size() && res[k].steps_left == s; ++k) {
      nodes -= res[k].nodes;
      bytes -= Arena::round(res[k].bytes);
    }
    if (fits()) return s;
  }
  return kNoHorizon;
}

// Queue of request ids in arrival order.  run() walks it front to back and
// calls try_admit(id) for each while slots are free; try_admit reserves what
// the request needs and returns true, or returns false when it does not fit
// now.  Behind a queue head that does not fit, a later request is tried only
// when may_pass(head, id) allows it, and at most max_bypass are admitted
// before the head; after that nothing more is admitted until it fits.
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

  template <class TryAdmit, class MayPass>
  std::vector<int64_t> run(size_t slots_free, TryAdmit&& try_admit, MayPass&& may_pass) {
    std::vector<int64_t> admitted;
    auto it = q_.begin();
    bool head = true;  // `it` is the current queue head
    int64_t waiting = -1;
    while (it != q_.end() && admitted.size() < slots_free) {
      if (!head && bypass_ >= max_bypass_) break;
      if (!head && !may_pass(waiting, *it)) {
        ++it;
        continue;
      }
      if (try_admit(*it)) {
        admitted.push_back(*it);
        if (head) bypass_ = 0;
        else ++bypass_;
        it = q_.erase(it);
      } else {
        if (head) waiting = *it;
        ++it;
        head = false;
      }
    }
    return admitted;
  }

  // Any later request may pass (up to max_bypass).
  template <class TryAdmit>
  std::vector<int64_t> run(size_t slots_free, TryAdmit&& try_admit) {
    return run(slots_free, try_admit, [](int64_t, int64_t) { return true; });
  }

 private:
  std::deque<int64_t> q_;
  int max_bypass_;
  int bypass_ = 0;  // admissions past the current head since it became head
};

// The packed tile list of one chunk of K steps.  Box k of the input has
// tiles[k] tiles and takes part in steps[k] (1..K) steps of the chunk.  The
// boxes are laid out in order of descending steps (ties keep their input
// order), so that the boxes still stepping at step j are a prefix of the list.
// order receives the input index of each laid-out box; hdr receives
//   hdr[0] = boxes, hdr[1] = total tiles, hdr[2 + k] = first tile of laid-out
//   box k (k = 0..boxes, hdr[2 + boxes] = total), and
//   hdr[3 + boxes + j] = tiles of the boxes stepping at step j (j = 0..K-1),
// pack_words(boxes, K) words in all.
inline size_t pack_words(size_t boxes, int K) { return boxes + 3 + size_t(K); }

inline void pack_chunk(const std::vector<unsigned>& tiles, const std::vector<int>& steps, int K,
                       std::vector<size_t>& order, unsigned* hdr) {
  const size_t n = tiles.size();
  order.resize(n);
  for (size_t k = 0; k < n; ++k) order[k] = k;
  std::stable_sort(order.begin(), order.end(),
                   [&](size_t a, size_t b) { return steps[a] > steps[b]; });
  unsigned total = 0;
  for (size_t k = 0; k < n; ++k) {
    hdr[2 + k] = total;
    total += tiles[order[k]];
  }
  hdr[0] = unsigned(n);
  hdr[1] = total;
  hdr[2 + n] = total;
  size_t m = n;  // boxes stepping at step j: the first m of the layout
  for (int j = 0; j < K; ++j) {
    while (m > 0 && steps[order[m - 1]] <= j) --m;
    hdr[3 + n + size_t(j)] = hdr[2 + m];
  }
}

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
