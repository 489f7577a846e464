// Host-side layer geometry of the sorted backend, shared by the engine
// (sorted_engine.hip) and the kernel-level layer ABI (sorted_abi.hip).
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include "kernels.hpp"

namespace mgbfs {

inline uint32_t ceil_log2(uint64_t x) {
  uint32_t b = 0;
  while ((1ull << b) < x) ++b;
  return b;
}

// Average candidates per sub-bucket: large sub-buckets amortise each
// workgroup's fixed work (bounds, table clear, fine-count scan); the cap is
// kUniqueCap = 5120 and counts stay within a few percent of the mean.
constexpr uint64_t kTargetBucket = 5000;
constexpr uint32_t kLaneParts = 32;  // slabs per owner lane (cursor contention / P)

// Slab capacity for `mean` expected entries: mean + slack * (mean / 4 +
// 16 sqrt(mean) + 1024), rounded to 64 entries.  A lane part's children come
// from contiguous tiles of a frontier ordered by mixed-key bits, and the
// mixer's top bits are linear in the key, so lane counts spread wider than
// a binomial; slack doubles after an overflow.
inline uint32_t slab_capacity(double mean, uint32_t slack) {
  const double c = mean + slack * (mean / 4.0 + 16.0 * std::sqrt(mean) + 1024.0);
  const uint64_t r = (static_cast<uint64_t>(c) + 63) / 64 * 64;
  return static_cast<uint32_t>(std::min<uint64_t>(r, 0xffffffc0ull));
}

// Sub-buckets wanted for an estimated candidate count: any count, not a
// power of two, so the mean fill stays near kTargetBucket on every layer.
inline uint64_t sub_buckets_for(uint64_t Cest) {
  const uint64_t kMaxSub = 1ull << kMaxBucketBits;
  return std::min<uint64_t>(kMaxSub, Cest <= kTargetBucket ? 1 : (Cest + kTargetBucket - 1) / kTargetBucket);
}

// Fine-bucket bits beyond the layer's sub-bucket bits for G generators.
inline uint32_t fine_margin(uint32_t G) { return std::min<uint32_t>(kMaxFineBits, 2 * (ceil_log2(G) + 1)); }

struct SortedGeometry {
  uint32_t K, K2, nsub, lanes, NB1, P, lcap, fb, nbo;
  uint64_t lslabs, lane_entries, sub_entries;
};

// Geometry of one layer pass.  K bits cover the sub-buckets.  Two routing
// passes above kMaxPassBits: expansion routes to lanes of 2^K2 consecutive
// sub-buckets, taking one bit more than half of K (measured on S13:
// partition 43-46 -> 39 ms, expand +2 ms); the sub-bucket count is rounded
// up to a whole number of lanes.  Direct (K2 = 0): the lanes are the
// sub-buckets, one slab each.  Otherwise every lane part receives the
// children of NB1 / P blocks (NB1 is a multiple of P), i.e.
// Cest / (lanes * P) on average.  The layer written is partitioned into
// nbo = nsub << fb fine buckets (at most 2^kMaxBucketBits), nested in the
// sub-buckets.
inline SortedGeometry sorted_geometry(uint64_t want, uint64_t F, uint64_t Cest, uint32_t slack, uint32_t margin) {
  SortedGeometry g{};
  g.K = ceil_log2(want);
  const uint32_t K1 = g.K <= kMaxPassBits ? g.K : std::min<uint32_t>(kMaxPassBits, (g.K + 1) / 2 + 1);
  g.K2 = g.K - K1;
  g.nsub = static_cast<uint32_t>(((want + (1ull << g.K2) - 1) >> g.K2) << g.K2);
  g.lanes = g.nsub >> g.K2;
  g.NB1 = static_cast<uint32_t>(std::clamp<uint64_t>((F + 2047) / 2048, 1, 8192));
  g.P = g.K2 ? std::min<uint32_t>(kLaneParts, g.NB1) : 1;
  g.NB1 = (g.NB1 + g.P - 1) / g.P * g.P;
  g.lcap = g.K2 ? slab_capacity(static_cast<double>(Cest) / (static_cast<double>(g.lanes) * g.P), slack)
                : kUniqueCap;
  g.lslabs = static_cast<uint64_t>(g.lanes) * g.P;
  g.lane_entries = g.K2 ? g.lslabs * g.lcap : 0;
  g.sub_entries = static_cast<uint64_t>(g.nsub) * kSlabStride;
  g.fb = std::min<uint32_t>(margin, kMaxBucketBits - g.K);
  g.nbo = g.nsub << g.fb;
  return g;
}

}  // namespace mgbfs
