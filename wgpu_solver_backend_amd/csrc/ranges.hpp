// ranges.hpp -- free-range bookkeeping of a device arena (host only).
//
// Blocks are 256-byte aligned and taken first fit (lowest offset whose free
// range is long enough); a returned block merges with free neighbours, so the
// free ranges never touch each other.
#pragma once
#include <cstdint>
#include <iterator>
#include <map>

namespace hbj {

class FreeRanges {
 public:
  static uint64_t round(uint64_t bytes) { return (bytes + 255) & ~255ull; }

  void reset(uint64_t size) {
    free_.clear();
    if (size) free_[0] = size;
  }
  bool take(uint64_t bytes, uint64_t* off) {
    bytes = round(bytes ? bytes : 1);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      *off = it->first;
      const uint64_t rest = it->second - bytes;
      free_.erase(it);
      if (rest) free_[*off + bytes] = rest;
      return true;
    }
    return false;
  }
  void give(uint64_t off, uint64_t bytes) {
    bytes = round(bytes ? bytes : 1);
    auto it = free_.emplace(off, bytes).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  const std::map<uint64_t, uint64_t>& ranges() const { return free_; }

 private:
  std::map<uint64_t, uint64_t> free_;  // offset -> length
};

}  // namespace hbj
