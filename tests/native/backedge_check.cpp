// Host check of the sorted backend's back-edge masks
// (fill_back_edge_masks in multigpubfs_amd/csrc/common.hpp).
//
// 1. Mask values: for LRX on S5 given as [L,R,X], [L,R,X,X] and [L,R,X,L],
//    child_mask[g] holds every index whose generator is g's inverse.
// 2. A CPU model of the pruned layer loop (each stored state carries the OR
//    of its discoverers' masks, expansion skips masked generators, and only
//    the current layer is probed) reproduces a plain BFS on those sets and
//    on random inverse-closed sets with repeated generators.
// 3. The same loop with single-index tags (the first inverse only) does not
//    reproduce the plain BFS on [L,R,X,X]: the check can fail.
#include <cstdio>
#include <map>
#include <random>
#include <set>
#include <vector>
#include "common.hpp"

using namespace mgbfs;

static PackedGenerators make(const std::vector<std::vector<uint8_t>>& gens, uint32_t n) {
  PackedGenerators pg{};
  pg.count = static_cast<uint32_t>(gens.size());
  pg.degree = n;
  for (uint32_t g = 0; g < pg.count; ++g)
    for (uint32_t i = 0; i < n; ++i) pg.src[g][i] = gens[g][i];
  return pg;
}

static uint64_t identity(uint32_t n) {
  uint64_t k = 0;
  for (uint32_t i = 0; i < n; ++i) k |= static_cast<uint64_t>(i) << (4 * i);
  return k;
}

static std::vector<uint64_t> plain_bfs(const PackedGenerators& pg) {
  std::set<uint64_t> seen{identity(pg.degree)};
  std::vector<uint64_t> front{identity(pg.degree)}, sizes{1};
  while (true) {
    std::set<uint64_t> next;
    for (uint64_t s : front)
      for (uint32_t g = 0; g < pg.count; ++g) {
        const uint64_t c = apply_packed_host(s, pg.src[g], pg.degree);
        if (!seen.count(c)) next.insert(c);
      }
    if (next.empty()) return sizes;
    seen.insert(next.begin(), next.end());
    front.assign(next.begin(), next.end());
    sizes.push_back(front.size());
  }
}

// Pruned loop: layer d maps state -> back-edge mask; children of d are
// probed against layer d only.  mask_of[g] is the mask a child made by g gets.
static std::vector<uint64_t> pruned_bfs(const PackedGenerators& pg, const std::vector<uint32_t>& mask_of,
                                        size_t max_layers) {
  std::map<uint64_t, uint32_t> cur{{identity(pg.degree), 0u}};
  std::vector<uint64_t> sizes{1};
  while (sizes.size() < max_layers) {
    std::map<uint64_t, uint32_t> next;
    for (const auto& [s, skip] : cur)
      for (uint32_t g = 0; g < pg.count; ++g) {
        if ((skip >> g) & 1u) continue;
        const uint64_t c = apply_packed_host(s, pg.src[g], pg.degree);
        if (cur.count(c)) continue;
        next[c] |= mask_of[g];
      }
    if (next.empty()) break;
    sizes.push_back(next.size());
    cur.swap(next);
  }
  return sizes;
}

static int failures = 0;
static void expect(bool ok, const char* what) {
  if (!ok) {
    std::printf("FAILED: %s\n", what);
    ++failures;
  }
}

int main() {
  const uint32_t n = 5;
  std::vector<uint8_t> L(n), R(n), X(n);
  for (uint32_t i = 0; i < n; ++i) {
    L[i] = static_cast<uint8_t>((i + 1) % n);
    R[i] = static_cast<uint8_t>((i + n - 1) % n);
    X[i] = static_cast<uint8_t>(i);
  }
  X[0] = 1;
  X[1] = 0;

  PackedGenerators lrx = make({L, R, X}, n);
  expect(fill_back_edge_masks(lrx), "LRX is inverse-closed");
  expect(lrx.child_mask[0] == 2 && lrx.child_mask[1] == 1 && lrx.child_mask[2] == 4, "LRX masks");

  PackedGenerators lrxx = make({L, R, X, X}, n);
  expect(fill_back_edge_masks(lrxx), "LRXX is inverse-closed");
  expect(lrxx.child_mask[2] == 12 && lrxx.child_mask[3] == 12, "repeated involution: both copies");
  expect(lrxx.child_mask[0] == 2 && lrxx.child_mask[1] == 1, "LRXX rotation masks");

  PackedGenerators lrxl = make({L, R, X, L}, n);
  expect(fill_back_edge_masks(lrxl), "LRXL is inverse-closed");
  expect(lrxl.child_mask[1] == 9 && lrxl.child_mask[0] == 2 && lrxl.child_mask[3] == 2, "repeated L: R's mask");

  PackedGenerators lx = make({L, X}, n);
  expect(!fill_back_edge_masks(lx), "LX is not inverse-closed");

  const std::vector<uint64_t> s5 = {1, 3, 6, 10, 16, 24, 29, 21, 6, 3, 1};
  for (PackedGenerators* pg : {&lrx, &lrxx, &lrxl}) {
    std::vector<uint32_t> masks(pg->count);
    for (uint32_t g = 0; g < pg->count; ++g) masks[g] = pg->child_mask[g];
    expect(plain_bfs(*pg) == s5, "plain BFS of LRX on S5");
    expect(pruned_bfs(*pg, masks, 64) == s5, "pruned loop with masks equals plain BFS");
  }

  // Single-index tags (first inverse only) on [L,R,X,X]: the second copy of
  // X leads back into the previous layer, which is not probed.
  {
    std::vector<uint32_t> first(lrxx.count);
    for (uint32_t g = 0; g < lrxx.count; ++g) first[g] = lrxx.child_mask[g] & (0u - lrxx.child_mask[g]);
    expect(pruned_bfs(lrxx, first, 64) != s5, "single-index tags are detected as wrong");
  }

  // Random inverse-closed sets on S6 with repeated generators.
  std::mt19937_64 rng(11);
  for (int trial = 0; trial < 40; ++trial) {
    const uint32_t m = 6;
    std::vector<std::vector<uint8_t>> gens;
    const int pairs = 1 + static_cast<int>(rng() % 3);
    for (int p = 0; p < pairs; ++p) {
      std::vector<uint8_t> a(m), inv(m);
      for (uint32_t i = 0; i < m; ++i) a[i] = static_cast<uint8_t>(i);
      for (uint32_t i = m - 1; i > 0; --i) std::swap(a[i], a[rng() % (i + 1)]);
      for (uint32_t i = 0; i < m; ++i) inv[a[i]] = static_cast<uint8_t>(i);
      gens.push_back(a);
      gens.push_back(inv);
    }
    const int dups = static_cast<int>(rng() % 3);
    for (int d = 0; d < dups; ++d) gens.push_back(gens[rng() % gens.size()]);
    for (size_t i = gens.size() - 1; i > 0; --i) std::swap(gens[i], gens[rng() % (i + 1)]);
    PackedGenerators pg = make(gens, m);
    expect(fill_back_edge_masks(pg), "random set is inverse-closed");
    std::vector<uint32_t> masks(pg.count);
    for (uint32_t g = 0; g < pg.count; ++g) masks[g] = pg.child_mask[g];
    const std::vector<uint64_t> want = plain_bfs(pg);
    expect(pruned_bfs(pg, masks, want.size() + 8) == want, "random set: pruned loop equals plain BFS");
  }

  if (failures) return 1;
  std::printf("backedge ok\n");
  return 0;
}
