// Randomised check of the batch arena's range bookkeeping against a
// byte-occupancy model: no two live ranges overlap, every range stays inside
// the chunk, the chosen range is a best fit, releases coalesce (a fully
// released chunk is one free range again), and misuse is rejected.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>
#include "arena.hpp"

using mgbfs::RangeSet;
using mgbfs::kArenaAlign;

#define CHECK(c)                                                   \
  do {                                                             \
    if (!(c)) {                                                    \
      std::fprintf(stderr, "check failed line %d: %s\n", __LINE__, #c); \
      return 1;                                                    \
    }                                                              \
  } while (0)

int main() {
  const uint64_t units = 512, size = units * kArenaAlign;
  std::mt19937_64 rng(12345);
  for (int trial = 0; trial < 200; ++trial) {
    RangeSet r(size);
    std::vector<int> owner(units, -1);  // unit -> live allocation id
    struct Live { uint64_t off, len; };
    std::vector<Live> live;
    for (int op = 0; op < 400; ++op) {
      if (live.empty() || rng() % 3 != 0) {
        const uint64_t want = 1 + rng() % (48 * kArenaAlign);
        const uint64_t len = (want + kArenaAlign - 1) / kArenaAlign * kArenaAlign;
        // Model: smallest free run that fits, lowest offset among equals.
        uint64_t best_off = ~0ull, best_len = ~0ull, u = 0;
        while (u < units) {
          if (owner[u] >= 0) { ++u; continue; }
          uint64_t v = u;
          while (v < units && owner[v] < 0) ++v;
          const uint64_t run = (v - u) * kArenaAlign;
          if (run >= len && run < best_len) { best_len = run; best_off = u * kArenaAlign; }
          u = v;
        }
        uint64_t off = 0;
        const bool ok = r.take(want, &off);
        CHECK(ok == (best_off != ~0ull));
        if (!ok) continue;
        CHECK(off == best_off);
        CHECK(off % kArenaAlign == 0 && off + len <= size);
        for (uint64_t k = off / kArenaAlign; k < (off + len) / kArenaAlign; ++k) {
          CHECK(owner[k] < 0);
          owner[k] = static_cast<int>(live.size());
        }
        live.push_back({off, want});
      } else {
        const size_t i = rng() % live.size();
        const Live x = live[i];
        r.give(x.off, x.len);
        const uint64_t len = (x.len + kArenaAlign - 1) / kArenaAlign * kArenaAlign;
        for (uint64_t k = x.off / kArenaAlign; k < (x.off + len) / kArenaAlign; ++k) owner[k] = -1;
        live.erase(live.begin() + i);
        for (auto& o : owner) if (o > static_cast<int>(i)) --o;
        bool threw = false;  // the same range again is a double free
        try { r.give(x.off, x.len); } catch (const std::logic_error&) { threw = true; }
        CHECK(threw);
      }
      uint64_t used = 0;
      for (const Live& l : live) used += (l.len + kArenaAlign - 1) / kArenaAlign * kArenaAlign;
      CHECK(r.used() == used);
      // Coalescing: the number of free ranges equals the number of free runs.
      size_t runs = 0;
      for (uint64_t k = 0; k < units; ++k) runs += owner[k] < 0 && (k == 0 || owner[k - 1] >= 0);
      CHECK(r.free_ranges() == runs);
    }
    for (const Live& l : live) r.give(l.off, l.len);
    CHECK(r.empty() && r.free_ranges() == 1 && r.largest_free() == size);
  }
  bool threw = false;
  try { RangeSet r(size); r.give(0, 256); } catch (const std::logic_error&) { threw = true; }
  CHECK(threw);
  std::printf("arena ok\n");
  return 0;
}
