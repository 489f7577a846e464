// Range bookkeeping for the batch engine's device arena: best-fit allocation
// of byte ranges inside one fixed-size chunk, with coalescing on release.
// Pure host code (offsets only, no device calls) so it is unit-tested on the
// CPU; the device memory behind each chunk belongs to the caller.
#pragma once
#include <cstdint>
#include <iterator>
#include <map>
#include <stdexcept>

namespace mgbfs {

constexpr uint64_t kArenaAlign = 256;

inline uint64_t arena_round(uint64_t bytes) {
  return bytes == 0 ? kArenaAlign : (bytes + kArenaAlign - 1) / kArenaAlign * kArenaAlign;
}

class RangeSet {
 public:
  explicit RangeSet(uint64_t size = 0) : size_(size) {
    if (size) free_[0] = size;
  }

  // Best fit: the smallest free range that holds `bytes`; the lowest offset
  // among equals.  Returns false when no range is large enough.
  bool take(uint64_t bytes, uint64_t* offset) {
    bytes = arena_round(bytes);
    auto best = free_.end();
    for (auto it = free_.begin(); it != free_.end(); ++it)
      if (it->second >= bytes && (best == free_.end() || it->second < best->second)) best = it;
    if (best == free_.end()) return false;
    const uint64_t off = best->first, len = best->second;
    free_.erase(best);
    if (len > bytes) free_[off + bytes] = len - bytes;
    used_ += bytes;
    *offset = off;
    return true;
  }

  // Returns [offset, offset + bytes) and merges it with free neighbours.
  void give(uint64_t offset, uint64_t bytes) {
    bytes = arena_round(bytes);
    if (offset % kArenaAlign || offset + bytes > size_ || bytes > used_) throw std::logic_error("ARENA_RANGE");
    const uint64_t returned = bytes;
    auto next = free_.lower_bound(offset);
    if (next != free_.end() && next->first < offset + bytes) throw std::logic_error("ARENA_DOUBLE_FREE");
    if (next != free_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second > offset) throw std::logic_error("ARENA_DOUBLE_FREE");
      if (prev->first + prev->second == offset) {
        offset = prev->first;
        bytes += prev->second;
        free_.erase(prev);
      }
    }
    if (next != free_.end() && next->first == offset + bytes) {
      bytes += next->second;
      free_.erase(next);
    }
    free_[offset] = bytes;
    used_ -= returned;
  }

  uint64_t size() const { return size_; }
  uint64_t used() const { return used_; }
  bool empty() const { return used_ == 0; }
  uint64_t largest_free() const {
    uint64_t m = 0;
    for (const auto& r : free_) m = r.second > m ? r.second : m;
    return m;
  }
  size_t free_ranges() const { return free_.size(); }

 private:
  uint64_t size_ = 0, used_ = 0;
  std::map<uint64_t, uint64_t> free_;  // offset -> length, disjoint, never adjacent
};

}  // namespace mgbfs
