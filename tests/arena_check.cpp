// Randomized check of the device-memory sub-allocator's bookkeeping (host only):
// live blocks never overlap and lie inside their chunks, freed memory merges
// back so that wholly free chunks are found again, and the byte counts add up.
#include <cstdio>
#include <cstdlib>
#include <map>
#include <random>
#include <vector>

#include "../quanty_amd/csrc/qs_arena.h"

#define CHECK(c)                                                   \
  do {                                                             \
    if (!(c)) {                                                    \
      std::fprintf(stderr, "check failed line %d: %s\n", __LINE__, #c); \
      return 1;                                                    \
    }                                                              \
  } while (0)

int main() {
  std::mt19937_64 rng(7);
  for (int trial = 0; trial < 20; ++trial) {
    QsArena a;
    // chunks at adjacent addresses: merges must stop at chunk edges
    std::vector<std::pair<uintptr_t, size_t>> chunks;
    uintptr_t base = 1 << 20;
    size_t total = 0;
    for (int c = 0; c < 4; ++c) {
      const size_t sz = (size_t)(1 + rng() % 64) << 16;
      a.add_chunk(base, sz);
      chunks.push_back({base, sz});
      base += sz;
      total += sz;
    }
    std::map<uintptr_t, size_t> live;
    for (int step = 0; step < 4000; ++step) {
      if (live.empty() || rng() % 3) {
        const size_t want = 1 + rng() % (1 << 18);
        const bool top = rng() % 4 == 0;  // long-lived blocks from the top
        const uintptr_t p = top ? a.take_top(want) : a.take(want);
        if (p == 0) continue;
        const size_t sz = QsArena::round(want);
        CHECK(p % QsArena::kAlign == 0);
        bool inside = false;
        for (auto &c : chunks) inside |= p >= c.first && p + sz <= c.first + c.second;
        CHECK(inside);
        auto nx = live.lower_bound(p);
        if (nx != live.end()) CHECK(p + sz <= nx->first);
        if (nx != live.begin()) {
          auto pv = std::prev(nx);
          CHECK(pv->first + pv->second <= p);
        }
        live[p] = sz;
      } else {
        auto it = live.begin();
        std::advance(it, rng() % live.size());
        CHECK(a.give(it->first));
        live.erase(it);
      }
      size_t lb = 0;
      for (auto &l : live) lb += l.second;
      CHECK(a.live_bytes() == lb);
      CHECK(a.live_bytes() + a.free_bytes() == total);
    }
    CHECK(!a.give(12345));
    for (auto &l : live) CHECK(a.give(l.first));
    CHECK(a.live_bytes() == 0 && a.free_bytes() == total);
    // everything merged back: every chunk is whole again
    auto out = a.release_free_chunks(0);
    CHECK(out.size() == chunks.size());
    CHECK(a.free_bytes() == 0 && a.chunk_count() == 0);
    // a block as large as the largest chunk fits again after re-adding the chunks
    for (auto &c : chunks) a.add_chunk(c.first, c.second);
    size_t biggest = 0;
    for (auto &c : chunks) biggest = std::max(biggest, c.second);
    CHECK(a.take(biggest) != 0);
    CHECK(a.take(total) == 0);
  }
  {
    // take_top: the block ends where the highest free block that holds it ends
    QsArena a;
    a.add_chunk(1 << 20, 1 << 20);
    a.add_chunk(4 << 20, 1 << 20);
    const uintptr_t t = a.take_top(1000);
    CHECK(t == (5 << 20) - QsArena::round(1000));
    const uintptr_t b = a.take(1000);  // best-fit: the smaller (first chunk's) block... any block
    CHECK(b != 0 && b != t);
    const uintptr_t big = a.take_top((1 << 20) - 512);  // only the first chunk still holds this much
    CHECK(big == (1 << 20) + 512 || big == 0);
    CHECK(a.give(t) && a.give(b));
    if (big) CHECK(a.give(big));
    CHECK(a.live_bytes() == 0);
  }
  std::printf("arena ok\n");
  return 0;
}
