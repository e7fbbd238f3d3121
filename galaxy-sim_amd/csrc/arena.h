// Free lists of an arena pool (host side, no HIP): the particle range and the
// node range of each occupied slot are taken from a budget-sized arena when a
// galaxy is placed in the slot and given back when the slot is emptied.
#pragma once
#include <stdint.h>

#include <iterator>
#include <map>

namespace gs {

// Free ranges [start, start + len) of one resource of `total` units, kept
// coalesced (no two free ranges touch).  take() is best fit: the smallest
// free range that holds len, the lowest start among equals, which keeps
// large ranges whole for large galaxies.
class RangeList {
 public:
  void reset(uint64_t total) {
    free_.clear();
    total_ = total;
    if (total) free_[0] = total;
  }

  // a range of len > 0 units, or false when no free range holds it
  bool take(uint64_t len, uint64_t* start) {
    if (len == 0 || len > total_) return false;
    auto best = free_.end();
    for (auto it = free_.begin(); it != free_.end(); ++it)
      if (it->second >= len && (best == free_.end() || it->second < best->second)) best = it;
    if (best == free_.end()) return false;
    *start = best->first;
    const uint64_t rest = best->second - len, after = best->first + len;
    free_.erase(best);
    if (rest) free_[after] = rest;
    return true;
  }

  // returns a taken range; false (nothing changes) for a range outside the
  // resource or overlapping a free range
  bool give(uint64_t start, uint64_t len) {
    if (len == 0) return true;
    if (start >= total_ || len > total_ - start) return false;
    auto next = free_.lower_bound(start);
    if (next != free_.end() && next->first < start + len) return false;
    if (next != free_.begin() && std::prev(next)->first + std::prev(next)->second > start)
      return false;
    uint64_t s = start, l = len;
    if (next != free_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == start) {
        s = prev->first;
        l += prev->second;
        free_.erase(prev);
      }
    }
    if (next != free_.end() && next->first == start + len) {
      l += next->second;
      free_.erase(next);
    }
    free_[s] = l;
    return true;
  }

  uint64_t total() const { return total_; }
  uint64_t free_units() const {
    uint64_t t = 0;
    for (const auto& r : free_) t += r.second;
    return t;
  }
  uint64_t largest() const {
    uint64_t m = 0;
    for (const auto& r : free_) m = r.second > m ? r.second : m;
    return m;
  }
  uint64_t pieces() const { return free_.size(); }
  const std::map<uint64_t, uint64_t>& ranges() const { return free_; }

 private:
  std::map<uint64_t, uint64_t> free_;  // start -> length
  uint64_t total_ = 0;
};

}  // namespace gs
