// Host check of the sorted backend's key mixer (multigpubfs_amd/csrc/common.hpp):
// keyunmix(keymix(x)) == x for several key widths, and the top bits used
// as bucket ids are near-uniform over all packed permutations of S9 and
// over the children (swap of positions 0 and 1) of a sample of parents
// spread across the mixed order.  Also range_start (the first top-32 value
// of a sub-bucket, computed without a 64-bit division) against the exact
// ceil(f 2^32 / n).
#include <algorithm>
#include <cstdio>
#include <numeric>
#include <random>
#include <vector>
#include "common.hpp"

using namespace mgbfs;

static double max_over_mean(const std::vector<uint64_t>& keys, const KeyMix& km, uint32_t bbits) {
  std::vector<uint64_t> count(1u << bbits, 0);
  for (uint64_t k : keys) ++count[keymix(k, km) >> (km.bits - bbits)];
  const double mean = static_cast<double>(keys.size()) / count.size();
  return *std::max_element(count.begin(), count.end()) / mean;
}

static bool range_start_exact(uint32_t f, uint32_t n) {
  const uint64_t want = ((static_cast<uint64_t>(f) << 32) + n - 1) / n;
  return range_start(f, n, range_inverse(n)) == want;
}

int main() {
  std::mt19937_64 rng(7);
  for (uint32_t n = 1; n <= 4096; ++n)
    for (uint32_t f = 0; f < n; f += 1 + n / 97)
      if (!range_start_exact(f, n) || !range_start_exact(n - 1, n)) {
        std::printf("range_start failed n=%u f=%u\n", n, f);
        return 1;
      }
  for (int i = 0; i < 2000000; ++i) {
    const uint32_t n = 1 + static_cast<uint32_t>(rng() % (1u << 20));
    const uint32_t f = static_cast<uint32_t>(rng() % n);
    if (!range_start_exact(f, n)) {
      std::printf("range_start failed n=%u f=%u\n", n, f);
      return 1;
    }
  }
  for (uint32_t n : {3u, 5120u, 131071u, 131072u, 1048575u, 1048576u})
    for (uint32_t f : {0u, 1u, n / 2, n - 2, n - 1})
      if (f < n && !range_start_exact(f, n)) {
        std::printf("range_start failed n=%u f=%u\n", n, f);
        return 1;
      }
  for (uint32_t bits : {8u, 28u, 36u, 52u, 60u, 64u}) {
    const KeyMix km = make_keymix(bits);
    for (int i = 0; i < 200000; ++i) {
      const uint64_t x = rng() & km.mask;
      const uint64_t y = keymix(x, km);
      if (y > km.mask || keyunmix(y, km) != x) {
        std::printf("round trip failed bits=%u x=%llx\n", bits, static_cast<unsigned long long>(x));
        return 1;
      }
    }
  }
  // All of S9, packed 4 bits per position.
  std::vector<int> p(9);
  std::iota(p.begin(), p.end(), 0);
  std::vector<uint64_t> all;
  do {
    uint64_t k = 0;
    for (int i = 0; i < 9; ++i) k |= static_cast<uint64_t>(p[i]) << (4 * i);
    all.push_back(k);
  } while (std::next_permutation(p.begin(), p.end()));
  const KeyMix km = make_keymix(36);
  const double r_all = max_over_mean(all, km, 6);
  // Children by "swap positions 0 and 1" of 4096 parents taken at a
  // stride across the mixed order.
  std::vector<std::pair<uint64_t, uint64_t>> order;
  for (uint64_t k : all) order.push_back({keymix(k, km), k});
  std::sort(order.begin(), order.end());
  std::vector<uint64_t> kids;
  for (size_t i = 0; i < 4096; ++i) {
    const uint64_t k = order[i * (order.size() / 4096)].second;
    const uint64_t a = k & 15, b = (k >> 4) & 15;
    kids.push_back((k & ~0xffull) | (a << 4) | b);
  }
  const double r_kids = max_over_mean(kids, km, 4);
  std::printf("max/mean: all S9 %.3f, children of spread parents %.3f\n", r_all, r_kids);
  if (r_all > 1.1 || r_kids > 1.35) return 1;
  std::printf("keymix ok\n");
  return 0;
}
