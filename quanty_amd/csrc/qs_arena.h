// Sub-allocator of the library's device memory (host-side bookkeeping only,
// no HIP calls, so the host tests compile it with g++).
//
// Memory comes from the driver in large chunks that stay with the process:
// the driver clears memory when it hands it out (tens of milliseconds per GB
// on the MI355X, and other threads' allocations wait meanwhile), so a serving
// process that released and re-allocated its multi-GB Krylov bases and
// assembled operators for every request would spend seconds in the driver.
// Blocks are carved from the chunks best-fit (the smallest free block that
// holds the request) and freed blocks merge with free neighbours of the same
// chunk. Not thread-safe: the caller holds a lock.
#pragma once

#include <cstddef>
#include <cstdint>
#include <iterator>
#include <map>
#include <unordered_map>
#include <vector>

class QsArena {
 public:
  static constexpr size_t kAlign = 256;

  static size_t round(size_t bytes) { return bytes == 0 ? kAlign : (bytes + kAlign - 1) / kAlign * kAlign; }

  // A new chunk [base, base + size) of free memory.
  void add_chunk(uintptr_t base, size_t size) {
    chunks_[base] = size;
    insert_free(base, size);
    free_bytes_ += size;
  }

  // Best-fit block of round(bytes) bytes, or 0 when no free block is large enough.
  uintptr_t take(size_t bytes) {
    const size_t want = round(bytes);
    auto it = by_size_.lower_bound(want);
    if (it == by_size_.end()) return 0;
    const size_t have = it->first;
    const uintptr_t p = it->second;
    by_size_.erase(it);
    by_addr_.erase(p);
    if (have > want) insert_free(p + want, have - want);
    live_[p] = want;
    live_bytes_ += want;
    free_bytes_ -= want;
    return p;
  }

  // A block of round(bytes) bytes at the top of the highest free block that
  // holds it, or 0. For long-lived data: kept at the top of the address
  // range, it leaves the free space below in few large pieces for the
  // short-lived blocks best-fit places there.
  uintptr_t take_top(size_t bytes) {
    const size_t want = round(bytes);
    for (auto it = by_addr_.rbegin(); it != by_addr_.rend(); ++it) {
      if (it->second < want) continue;
      const uintptr_t start = it->first;
      const size_t have = it->second;
      erase_free(std::next(it).base());
      if (have > want) insert_free(start, have - want);
      const uintptr_t p = start + (have - want);
      live_[p] = want;
      live_bytes_ += want;
      free_bytes_ -= want;
      return p;
    }
    return 0;
  }

  // Return a block from take(); false when p is not a live block of this arena.
  bool give(uintptr_t p) {
    auto it = live_.find(p);
    if (it == live_.end()) return false;
    size_t size = it->second;
    live_.erase(it);
    live_bytes_ -= size;
    free_bytes_ += size;
    uintptr_t start = p;
    const auto ch = chunk_of(p);
    // merge with the free block that ends at p, and the one that starts at its end, inside the chunk
    auto next = by_addr_.lower_bound(start);
    if (next != by_addr_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == start && prev->first >= ch.first) {
        start = prev->first;
        size += prev->second;
        erase_free(prev);
      }
    }
    next = by_addr_.find(start + size);
    if (next != by_addr_.end() && next->first < ch.first + ch.second) {
      size += next->second;
      erase_free(next);
    }
    insert_free(start, size);
    return true;
  }

  bool owns(uintptr_t p) const { return live_.count(p) != 0; }
  size_t live_bytes() const { return live_bytes_; }
  size_t free_bytes() const { return free_bytes_; }
  size_t chunk_count() const { return chunks_.size(); }

  // Remove wholly free chunks, largest first, until at most keep bytes are
  // free; returns them (base, size) for the caller to hand back to the driver.
  std::vector<std::pair<uintptr_t, size_t>> release_free_chunks(size_t keep) {
    std::multimap<size_t, uintptr_t> whole;
    for (auto &c : chunks_) {
      auto f = by_addr_.find(c.first);
      if (f != by_addr_.end() && f->second == c.second) whole.insert({c.second, c.first});
    }
    std::vector<std::pair<uintptr_t, size_t>> out;
    for (auto it = whole.rbegin(); it != whole.rend() && free_bytes_ > keep; ++it) {
      erase_free(by_addr_.find(it->second));
      chunks_.erase(it->second);
      free_bytes_ -= it->first;
      out.push_back({it->second, it->first});
    }
    return out;
  }

 private:
  // the chunk holding address p: (base, size)
  std::pair<uintptr_t, size_t> chunk_of(uintptr_t p) const {
    auto it = chunks_.upper_bound(p);
    --it;
    return *it;
  }
  void insert_free(uintptr_t p, size_t size) {
    by_addr_[p] = size;
    by_size_.insert({size, p});
  }
  void erase_free(std::map<uintptr_t, size_t>::iterator it) {
    auto range = by_size_.equal_range(it->second);
    for (auto s = range.first; s != range.second; ++s)
      if (s->second == it->first) {
        by_size_.erase(s);
        break;
      }
    by_addr_.erase(it);
  }

  std::map<uintptr_t, size_t> chunks_;        // base -> size
  std::map<uintptr_t, size_t> by_addr_;       // free blocks: start -> size
  std::multimap<size_t, uintptr_t> by_size_;  // free blocks: size -> start
  std::unordered_map<uintptr_t, size_t> live_;
  size_t live_bytes_ = 0, free_bytes_ = 0;
};
