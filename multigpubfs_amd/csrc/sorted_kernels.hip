// Sort-unique backend with three rotating VRAM banks (inverse-closed sets).
//
// States are stored as h = mix64(key) (a bijection, so equality of h is
// full-key equality).  A bank holds one BFS layer, partitioned by the top
// Ko bits of h ("fine buckets"), with an offset table of 2^Ko + 1 entries.
// One layer d -> d+1:
//   expand_count / expand_scatter : children of bank[d], bucketed by the top
//        K1 bits into owner lanes (LDS counting sort per 4096-element tile,
//        contiguous run writes).                              8F + 8F + 8C
//   partition_count / _scatter    : each lane split by the next K2 bits.
//                                                             8C + 16C
//   unique_diff                   : one workgroup per sub-bucket (K = K1+K2
//        bits): LDS open-addressing set of the candidates (in-layer dedup),
//        probe with the same sub-bucket of bank[d] and bank[d-1] (exact
//        three-bank retention for inverse-closed generators), survivors
//        regrouped by Ko fine bits and written back in place. 8C + 8F + 8F' + 8N
//   compact                       : survivors -> bank[d+1].   16N
// Replaces the reference's sorted-run owner (cuda/segmented_hash_batch.cu:
// 157-215: merge sort by (hash,state), binary-search probe of bank runs,
// exclusive-scan compaction) and its batch owner routing
// (cuda/segmented_batch_route.cu:52-151) with LDS-resident bucket work.
#include <algorithm>
#include "common.hpp"
#include "kernels.hpp"

namespace mgbfs {

constexpr int kRouteThreads = 512;
constexpr int kTile = 8192;                            // elements per LDS routing tile
constexpr int kTilePerThread = kTile / kRouteThreads;  // 16 elements held in registers
constexpr int kUniqueThreads = 512;
constexpr int kUniquePerThread = kUniqueCap / kUniqueThreads;  // candidates held in registers
constexpr int kPrefetch = 4;  // independent loads issued before dependent LDS work

__device__ __forceinline__ uint32_t top_bits(uint64_t h, uint32_t bits) {
  return bits ? static_cast<uint32_t>(h >> (64u - bits)) : 0u;
}

__device__ __forceinline__ uint32_t mid_bits(uint64_t h, uint32_t skip, uint32_t bits) {
  // bits [64-skip-bits, 64-skip) of h
  return bits ? static_cast<uint32_t>((h << skip) >> (64u - bits)) : 0u;
}

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const uint32_t lane = threadIdx.x & (kWave - 1);
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const uint32_t t = __shfl_up(v, off, kWave);
    if (lane >= static_cast<uint32_t>(off)) v += t;
  }
  return v;
}

// Exclusive scan of `count` (<= 4 * blockDim) LDS counters into `out`.
__device__ void lds_exclusive_scan(const uint32_t* in, uint32_t* out, uint32_t count, uint32_t* wave_sums) {
  const uint32_t per = (count + blockDim.x - 1) / blockDim.x;  // <= 4
  const uint32_t base = threadIdx.x * per;
  uint32_t v[4];
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t i = base + k;
    v[k] = (k < static_cast<int>(per) && i < count) ? in[i] : 0u;
    s += v[k];
  }
  const uint32_t incl = wave_incl_scan(s);
  const uint32_t lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  if (lane == kWave - 1) wave_sums[wave] = incl;
  __syncthreads();
  uint32_t run = incl - s;
  for (uint32_t w = 0; w < wave; ++w) run += wave_sums[w];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t i = base + k;
    if (k < static_cast<int>(per) && i < count) out[i] = run;
    run += v[k];
  }
  __syncthreads();
}

// ------------------------------------------------------------ expansion

// Stored value of a child made by generator g from `key`.
__device__ __forceinline__ uint64_t child_value(uint64_t key, const PackedGenerators& gens, int g,
                                                const SortedLayout& L) {
  const uint64_t child = apply_groups(key, gens, g);
  return (keymix(child, L.km) << L.tagbits) | gens.child_tag[g];
}

__device__ __forceinline__ uint32_t parent_skip(uint64_t v, const SortedLayout& L) {
  // generator index + 1 to skip (0: none)
  return L.prune ? static_cast<uint32_t>(v & L.tagmask) : 0u;
}

template <int G>
__global__ __launch_bounds__(kRouteThreads) void expand_count_kernel(const uint64_t* __restrict__ bank, uint64_t F,
                                                                      PackedGenerators gens, SortedLayout L,
                                                                      uint32_t K1, uint32_t NB,
                                                                      uint32_t* __restrict__ counts) {
  __shared__ uint32_t hist[1u << kMaxPassBits];
  const uint32_t buckets = 1u << K1;
  for (uint32_t i = threadIdx.x; i < buckets; i += blockDim.x) hist[i] = 0;
  __syncthreads();
  const uint64_t p0 = F * blockIdx.x / NB, p1 = F * (blockIdx.x + 1) / NB;
  for (uint64_t base = p0 + threadIdx.x; base < p1; base += kPrefetch * kRouteThreads) {
    uint64_t pv[kPrefetch];
#pragma unroll
    for (int k = 0; k < kPrefetch; ++k) {
      const uint64_t p = base + static_cast<uint64_t>(k) * kRouteThreads;
      pv[k] = p < p1 ? bank[p] : 0;
    }
#pragma unroll
    for (int k = 0; k < kPrefetch; ++k) {
      if (base + static_cast<uint64_t>(k) * kRouteThreads >= p1) break;
      const uint64_t v = pv[k];
      const uint64_t key = keyunmix(v >> L.tagbits, L.km);
      const uint32_t skip = parent_skip(v, L);
#pragma unroll
      for (int g = 0; g < G; ++g)
        if (static_cast<uint32_t>(g) + 1 != skip) atomicAdd(&hist[top_bits(child_value(key, gens, g, L), K1)], 1u);
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < buckets; i += blockDim.x) counts[static_cast<uint64_t>(i) * NB + blockIdx.x] = hist[i];
}

// Shared tile machinery.  Each thread holds up to PER tile elements in
// registers (bit k of `valid` marks v[k]); they are ranked within their
// bucket by LDS atomics, placed bucket-contiguously in t.s, and written out
// as contiguous runs at out[running[b] + offset within the tile's run].
struct TileLds {
  uint64_t s[kTile];
  uint32_t lhist[1u << kMaxPassBits];
  uint32_t lstart[1u << kMaxPassBits];
  uint32_t running[1u << kMaxPassBits];
  uint32_t wave_sums[kRouteThreads / kWave];
};

template <int PER, typename BucketFn>
__device__ __forceinline__ void route_regs(TileLds& t, const uint64_t (&v)[PER], uint32_t valid, uint32_t buckets,
                                           BucketFn bkt, uint64_t* __restrict__ out) {
  uint32_t rank[PER];
#pragma unroll
  for (int k = 0; k < PER; ++k) rank[k] = ((valid >> k) & 1u) ? atomicAdd(&t.lhist[bkt(v[k])], 1u) : 0u;
  __syncthreads();
  lds_exclusive_scan(t.lhist, t.lstart, buckets, t.wave_sums);
  const uint32_t ne = t.lstart[buckets - 1] + t.lhist[buckets - 1];
#pragma unroll
  for (int k = 0; k < PER; ++k)
    if ((valid >> k) & 1u) t.s[t.lstart[bkt(v[k])] + rank[k]] = v[k];
  __syncthreads();
  for (uint32_t j = threadIdx.x; j < ne; j += kRouteThreads) {
    const uint64_t h = t.s[j];
    const uint32_t b = bkt(h);
    out[static_cast<uint64_t>(t.running[b]) + (j - t.lstart[b])] = h;
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < buckets; i += kRouteThreads) {
    t.running[i] += t.lhist[i];
    t.lhist[i] = 0;
  }
  __syncthreads();
}

// TAGGED: every parent carries a tag (all layers after the start of an
// inverse-closed set with pruning), so each yields exactly G-1 children and
// twice as many parents fit a tile.
template <int G, bool TAGGED>
struct ExpandTile {
  static constexpr bool kCompact = TAGGED && G > 1;          // drop the skipped child by selects
  static constexpr int kPerParent = kCompact ? G - 1 : G;
  static constexpr int kParentsPerThread = (kTilePerThread / kPerParent) > 0 ? (kTilePerThread / kPerParent) : 1;
  static constexpr int kRegs = kParentsPerThread * kPerParent;  // <= 32 (valid mask width)
  static constexpr int kParents = kParentsPerThread * kRouteThreads;
};

template <int G, bool TAGGED>
__global__ __launch_bounds__(kRouteThreads) void expand_scatter_kernel(const uint64_t* __restrict__ bank, uint64_t F,
                                                                        PackedGenerators gens, SortedLayout L,
                                                                        uint32_t K1, uint32_t NB,
                                                                        const uint32_t* __restrict__ offsets,
                                                                        uint64_t* __restrict__ out) {
  using T = ExpandTile<G, TAGGED>;
  static_assert(T::kRegs <= 32, "valid mask is 32 bits");
  __shared__ TileLds t;
  const uint32_t buckets = 1u << K1;
  for (uint32_t i = threadIdx.x; i < buckets; i += kRouteThreads) {
    t.running[i] = offsets[static_cast<uint64_t>(i) * NB + blockIdx.x];
    t.lhist[i] = 0;
  }
  __syncthreads();
  const uint64_t p0 = F * blockIdx.x / NB, p1 = F * (blockIdx.x + 1) / NB;
  auto bkt = [K1](uint64_t h) { return top_bits(h, K1); };
  for (uint64_t t0 = p0; t0 < p1; t0 += T::kParents) {
    uint64_t pv[T::kParentsPerThread];
#pragma unroll
    for (int m = 0; m < T::kParentsPerThread; ++m) {
      const uint64_t q = t0 + static_cast<uint64_t>(m) * kRouteThreads + threadIdx.x;
      pv[m] = q < p1 ? bank[q] : 0;
    }
    uint64_t v[T::kRegs];
    uint32_t valid = 0;
#pragma unroll
    for (int m = 0; m < T::kParentsPerThread; ++m) {
      const bool live = t0 + static_cast<uint64_t>(m) * kRouteThreads + threadIdx.x < p1;
      const uint64_t key = keyunmix(pv[m] >> L.tagbits, L.km);
      const uint32_t skip = parent_skip(pv[m], L);
      if constexpr (T::kCompact) {
        // Every parent skips exactly generator skip-1; keep the other G-1.
        uint64_t c[G];
#pragma unroll
        for (int g = 0; g < G; ++g) c[g] = child_value(key, gens, g, L);
#pragma unroll
        for (int j = 0; j < G - 1; ++j) v[m * (G - 1) + j] = static_cast<uint32_t>(j) + 1 < skip ? c[j] : c[j + 1];
        if (live) valid |= ((1u << (G - 1)) - 1u) << (m * (G - 1));
      } else {
#pragma unroll
        for (int g = 0; g < G; ++g) {
          v[m * G + g] = child_value(key, gens, g, L);
          if (live && static_cast<uint32_t>(g) + 1 != skip) valid |= 1u << (m * G + g);
        }
      }
    }
    route_regs(t, v, valid, buckets, bkt, out);
  }
}

// ------------------------------------------------------------ partition

__global__ __launch_bounds__(kRouteThreads) void partition_count_kernel(const uint64_t* __restrict__ in,
                                                                         const uint32_t* __restrict__ scan1, uint32_t NB1,
                                                                         uint32_t K1, uint32_t K2, uint32_t NC,
                                                                         uint32_t* __restrict__ counts) {
  __shared__ uint32_t hist[1u << kMaxPassBits];
  const uint32_t buckets = 1u << K2;
  for (uint32_t i = threadIdx.x; i < buckets; i += blockDim.x) hist[i] = 0;
  __syncthreads();
  const uint32_t j = blockIdx.x / NC, c = blockIdx.x % NC;
  const uint64_t gs = scan1[static_cast<uint64_t>(j) * NB1], ge = scan1[static_cast<uint64_t>(j + 1) * NB1];
  const uint64_t r0 = gs + (ge - gs) * c / NC, r1 = gs + (ge - gs) * (c + 1) / NC;
  for (uint64_t base = r0 + threadIdx.x; base < r1; base += kPrefetch * kRouteThreads) {
    uint64_t v[kPrefetch];
#pragma unroll
    for (int k = 0; k < kPrefetch; ++k) {
      const uint64_t i = base + static_cast<uint64_t>(k) * kRouteThreads;
      v[k] = i < r1 ? in[i] : 0;
    }
#pragma unroll
    for (int k = 0; k < kPrefetch; ++k)
      if (base + static_cast<uint64_t>(k) * kRouteThreads < r1) atomicAdd(&hist[mid_bits(v[k], K1, K2)], 1u);
  }
  __syncthreads();
  for (uint32_t b = threadIdx.x; b < buckets; b += blockDim.x)
    counts[((static_cast<uint64_t>(j) << K2) + b) * NC + c] = hist[b];
}

__global__ __launch_bounds__(kRouteThreads) void partition_scatter_kernel(const uint64_t* __restrict__ in,
                                                                           const uint32_t* __restrict__ scan1,
                                                                           uint32_t NB1, uint32_t K1, uint32_t K2,
                                                                           uint32_t NC,
                                                                           const uint32_t* __restrict__ offsets,
                                                                           uint64_t* __restrict__ out) {
  __shared__ TileLds t;
  const uint32_t buckets = 1u << K2;
  const uint32_t j = blockIdx.x / NC, c = blockIdx.x % NC;
  for (uint32_t b = threadIdx.x; b < buckets; b += kRouteThreads) {
    t.running[b] = offsets[((static_cast<uint64_t>(j) << K2) + b) * NC + c];
    t.lhist[b] = 0;
  }
  __syncthreads();
  const uint64_t gs = scan1[static_cast<uint64_t>(j) * NB1], ge = scan1[static_cast<uint64_t>(j + 1) * NB1];
  const uint64_t r0 = gs + (ge - gs) * c / NC, r1 = gs + (ge - gs) * (c + 1) / NC;
  auto bkt = [K1, K2](uint64_t h) { return mid_bits(h, K1, K2); };
  for (uint64_t t0 = r0; t0 < r1; t0 += kTile) {
    uint64_t v[kTilePerThread];
    uint32_t valid = 0;
#pragma unroll
    for (int k = 0; k < kTilePerThread; ++k) {
      const uint64_t i = t0 + static_cast<uint64_t>(k) * kRouteThreads + threadIdx.x;
      v[k] = i < r1 ? in[i] : 0;
      if (i < r1) valid |= 1u << k;
    }
    route_regs(t, v, valid, buckets, bkt, out);
  }
}

// ------------------------------------------------------ unique + difference

__device__ __forceinline__ bool same_state(uint64_t a, uint64_t b, uint32_t tagbits) {
  return ((a ^ b) >> tagbits) == 0;
}

struct UniqueLds {
  uint64_t table[kUniqueSlots];
  uint32_t found[kUniqueSlots / 32];
  uint32_t fhist[1u << kMaxFineBits];
  uint32_t fstart[1u << kMaxFineBits];
  uint32_t wave_sums[kUniqueThreads / kWave];
  uint32_t flags;  // bit0: candidate ~0 present, bit1: ~0 found in a bank
  uint32_t unique;
};

constexpr uint64_t kEmpty = ~0ull;

__device__ __forceinline__ int lds_find(const UniqueLds& u, uint64_t h, uint32_t tagbits, uint32_t mask) {
  uint32_t slot = static_cast<uint32_t>(h >> tagbits) & mask;
  for (uint32_t probe = 0; probe <= mask; ++probe) {
    const uint64_t v = u.table[slot];
    if (v == kEmpty) return -1;
    if (same_state(v, h, tagbits)) return static_cast<int>(slot);
    slot = (slot + 1) & mask;
  }
  return -1;
}

// Marks table entries present in bank[lo, hi).  The first kPrefetch
// elements of this thread (pre) were loaded before the table was built;
// later chunks are loaded in batches.
__device__ __forceinline__ void probe_value(UniqueLds& u, uint64_t h, uint32_t tagbits, uint32_t mask) {
  if (h == kEmpty) {
    atomicOr(&u.flags, 2u);
    return;
  }
  const int s = lds_find(u, h, tagbits, mask);
  if (s >= 0) atomicOr(&u.found[s >> 5], 1u << (s & 31));
}

__device__ __forceinline__ void prefetch_slice(const uint64_t* __restrict__ bank, uint64_t lo, uint64_t hi,
                                               uint64_t (&pre)[kPrefetch]) {
#pragma unroll
  for (int k = 0; k < kPrefetch; ++k) {
    const uint64_t i = lo + threadIdx.x + static_cast<uint64_t>(k) * kUniqueThreads;
    pre[k] = i < hi ? bank[i] : 0;
  }
}

__device__ __forceinline__ void probe_bank(UniqueLds& u, const uint64_t* __restrict__ bank, uint64_t lo, uint64_t hi,
                                           const uint64_t (&pre)[kPrefetch], uint32_t tagbits, uint32_t mask) {
#pragma unroll
  for (int k = 0; k < kPrefetch; ++k)
    if (lo + threadIdx.x + static_cast<uint64_t>(k) * kUniqueThreads < hi) probe_value(u, pre[k], tagbits, mask);
  for (uint64_t base = lo + threadIdx.x + kPrefetch * kUniqueThreads; base < hi; base += kPrefetch * kUniqueThreads) {
    uint64_t v[kPrefetch];
#pragma unroll
    for (int k = 0; k < kPrefetch; ++k) {
      const uint64_t i = base + static_cast<uint64_t>(k) * kUniqueThreads;
      v[k] = i < hi ? bank[i] : 0;
    }
#pragma unroll
    for (int k = 0; k < kPrefetch; ++k)
      if (base + static_cast<uint64_t>(k) * kUniqueThreads < hi) probe_value(u, v[k], tagbits, mask);
  }
}

// Bounds of sub-bucket f: candidates [s, e), current-bank slice [clo, chi),
// previous-bank slice [plo, phi).
struct BucketBounds {
  uint32_t s, e, clo, chi, plo, phi;
};

struct UniqueArgs {
  uint64_t* cand;
  const uint32_t* cscan;
  const uint64_t* cur;
  const uint32_t* cur_off;
  const uint64_t* prev;
  const uint32_t* prev_off;
  uint32_t* fine_counts;
  uint32_t* error;
  unsigned long long* unique_total;
  uint32_t NC, K, Kc, Kp, Ko, tagbits, buckets;
};

__device__ __forceinline__ BucketBounds load_bounds(const UniqueArgs& a, uint32_t f) {
  BucketBounds b{0, 0, 0, 0, 0, 0};
  if (f >= a.buckets) return b;
  b.s = a.cscan[static_cast<uint64_t>(f) * a.NC];
  b.e = a.cscan[static_cast<uint64_t>(f + 1) * a.NC];
  if (a.cur) {
    b.clo = a.cur_off[static_cast<uint64_t>(f) << (a.Kc - a.K)];
    b.chi = a.cur_off[static_cast<uint64_t>(f + 1) << (a.Kc - a.K)];
  }
  if (a.prev) {
    b.plo = a.prev_off[static_cast<uint64_t>(f) << (a.Kp - a.K)];
    b.phi = a.prev_off[static_cast<uint64_t>(f + 1) << (a.Kp - a.K)];
  }
  return b;
}

struct BucketData {
  uint64_t v[kUniquePerThread];
  uint64_t cpre[kPrefetch], ppre[kPrefetch];
};

__device__ __forceinline__ void load_data(const UniqueArgs& a, const BucketBounds& b, BucketData& d) {
  const bool fits = b.e - b.s <= kUniqueCap;
#pragma unroll
  for (int k = 0; k < kUniquePerThread; ++k) {
    const uint32_t i = b.s + k * kUniqueThreads + threadIdx.x;
    d.v[k] = (fits && i < b.e) ? a.cand[i] : 0;
  }
  prefetch_slice(a.cur, b.clo, b.chi, d.cpre);
  prefetch_slice(a.prev, b.plo, b.phi, d.ppre);
}

// Persistent blocks: block x handles sub-buckets x, x + grid, ...  Loads run
// two stages ahead of the LDS work: bounds of bucket i+2 and candidates /
// first bank chunks of bucket i+1 are in flight while bucket i is deduped.
// Each thread keeps its (at most kUniquePerThread) candidates and their
// slots in registers, so only the candidates it inserted first are revisited.
__global__ __launch_bounds__(kUniqueThreads) __attribute__((amdgpu_waves_per_eu(4, 4))) void unique_diff_kernel(
    UniqueArgs a) {
  __shared__ UniqueLds u;
  const uint32_t fine_bits = a.Ko - a.K, fine = 1u << fine_bits;
  const uint32_t step = gridDim.x;
  uint32_t f = blockIdx.x;
  BucketBounds bcur = load_bounds(a, f);
  BucketBounds bnext = load_bounds(a, f + step);
  BucketData dcur;
  load_data(a, bcur, dcur);
  unsigned long long unique_block = 0;
  for (; f < a.buckets; f += step) {
    const BucketBounds bafter = load_bounds(a, f + 2 * step);
    BucketData dnext;
    load_data(a, bnext, dnext);

    const uint32_t s = bcur.s, e = bcur.e;
    const uint32_t c = e - s;
    if (c > kUniqueCap) {
      if (threadIdx.x == 0) atomicOr(a.error, kErrBucketCap);
      for (uint32_t i = threadIdx.x; i < fine; i += kUniqueThreads)
        a.fine_counts[(static_cast<uint64_t>(f) << fine_bits) + i] = 0;
    } else {
      // Table sized to the bucket: slots = next_pow2(2c) (load factor <= 0.5).
      uint32_t slots = 64;
      while (slots < 2 * c) slots <<= 1;
      const uint32_t mask = slots - 1;
      for (uint32_t i = threadIdx.x; i < slots; i += kUniqueThreads) u.table[i] = kEmpty;
      for (uint32_t i = threadIdx.x; i < slots / 32 + 1; i += kUniqueThreads) u.found[i] = 0;
      for (uint32_t i = threadIdx.x; i < fine; i += kUniqueThreads) u.fhist[i] = 0;
      if (threadIdx.x == 0) {
        u.flags = 0;
        u.unique = 0;
      }
      __syncthreads();
      // In-layer dedup: insert every candidate (state equality ignores tags).
      uint32_t slot_of[kUniquePerThread];
      uint32_t owned = 0;  // bit k: this thread inserted v[k] first
      uint32_t special = 0;
#pragma unroll
      for (int k = 0; k < kUniquePerThread; ++k) {
        slot_of[k] = 0;
        if (s + k * kUniqueThreads + threadIdx.x >= e) continue;
        const uint64_t h = dcur.v[k];
        if (h == kEmpty) {
          if (!(atomicOr(&u.flags, 1u) & 1u)) special = 1;
          continue;
        }
        uint32_t slot = static_cast<uint32_t>(h >> a.tagbits) & mask;
        while (true) {
          const uint64_t old = atomicCAS(reinterpret_cast<unsigned long long*>(&u.table[slot]), kEmpty, h);
          if (old == kEmpty) {
            owned |= 1u << k;
            slot_of[k] = slot;
            break;
          }
          if (same_state(old, h, a.tagbits)) break;
          slot = (slot + 1) & mask;
        }
      }
      const uint32_t mine = __popc(owned) + special;
      if (mine) atomicAdd(&u.unique, mine);
      __syncthreads();
      // Difference against the two retained layers (same sub-bucket only).
      probe_bank(u, a.cur, bcur.clo, bcur.chi, dcur.cpre, a.tagbits, mask);
      probe_bank(u, a.prev, bcur.plo, bcur.phi, dcur.ppre, a.tagbits, mask);
      __syncthreads();
      const bool max_survives = special && !(u.flags & 2u);
      uint32_t survive = 0;
#pragma unroll
      for (int k = 0; k < kUniquePerThread; ++k) {
        if ((owned >> k) & 1u) {
          const uint32_t sl = slot_of[k];
          if (!((u.found[sl >> 5] >> (sl & 31)) & 1u)) {
            survive |= 1u << k;
            atomicAdd(&u.fhist[mid_bits(dcur.v[k], a.K, fine_bits)], 1u);
          }
        }
      }
      if (max_survives) atomicAdd(&u.fhist[fine - 1], 1u);
      __syncthreads();
      lds_exclusive_scan(u.fhist, u.fstart, fine, u.wave_sums);
      for (uint32_t i = threadIdx.x; i < fine; i += kUniqueThreads) {
        a.fine_counts[(static_cast<uint64_t>(f) << fine_bits) + i] = u.fhist[i];
        u.fhist[i] = u.fstart[i];
      }
      __syncthreads();
      // Every candidate is in registers: write survivors back in place.
#pragma unroll
      for (int k = 0; k < kUniquePerThread; ++k)
        if ((survive >> k) & 1u) a.cand[s + atomicAdd(&u.fhist[mid_bits(dcur.v[k], a.K, fine_bits)], 1u)] = dcur.v[k];
      if (max_survives) a.cand[s + atomicAdd(&u.fhist[fine - 1], 1u)] = kEmpty;
      if (threadIdx.x == 0) unique_block += u.unique;
      __syncthreads();  // LDS reuse by the next bucket
    }
    bcur = bnext;
    bnext = bafter;
    dcur = dnext;
  }
  if (threadIdx.x == 0 && unique_block) atomicAdd(a.unique_total, unique_block);
}

// grid = 2^K blocks: survivors of sub-bucket f -> bank rows.
__global__ __launch_bounds__(kRouteThreads) void compact_kernel(const uint64_t* __restrict__ cand,
                                                                const uint32_t* __restrict__ cscan, uint32_t NC,
                                                                uint32_t K, uint32_t Ko,
                                                                const uint32_t* __restrict__ off,
                                                                uint64_t* __restrict__ bank) {
  const uint32_t f = blockIdx.x;
  const uint64_t s = cscan[static_cast<uint64_t>(f) * NC];
  const uint32_t sh = Ko - K;
  const uint64_t lo = off[static_cast<uint64_t>(f) << sh], hi = off[static_cast<uint64_t>(f + 1) << sh];
  for (uint64_t i = threadIdx.x; i < hi - lo; i += kRouteThreads) bank[lo + i] = cand[s + i];
}

// Start layer: bank[0] = mix(start); offsets at Ko bits.
__global__ void sorted_seed_kernel(uint64_t start_key, SortedLayout L, uint64_t* bank, uint32_t* off, uint32_t Ko) {
  const uint64_t h = keymix(start_key, L.km) << L.tagbits;
  const uint32_t b = top_bits(h, Ko);
  const uint64_t n = (1ull << Ko) + 1;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    off[i] = i <= b ? 0u : 1u;
  if (blockIdx.x == 0 && threadIdx.x == 0) bank[0] = h;
}

// ------------------------------------------------------------ launchers

hipError_t sorted_seed(uint64_t start_key, const SortedLayout& L, uint64_t* bank, uint32_t* off, uint32_t Ko,
                       hipStream_t stream) {
  const uint64_t n = (1ull << Ko) + 1;
  const uint32_t blocks = static_cast<uint32_t>(std::min<uint64_t>((n + 255) / 256, 4096));
  hipLaunchKernelGGL(sorted_seed_kernel, dim3(blocks), dim3(256), 0, stream, start_key, L, bank, off, Ko);
  return hipGetLastError();
}

template <int G, bool TAGGED>
static hipError_t expand_route_g(const uint64_t* bank, uint64_t F, const PackedGenerators& gens, const SortedLayout& L,
                                 uint32_t K1, uint32_t NB, uint32_t* counts, uint32_t* scan, uint32_t* scan_ws,
                                 uint64_t* out, hipStream_t stream) {
  hipLaunchKernelGGL(expand_count_kernel<G>, dim3(NB), dim3(kRouteThreads), 0, stream, bank, F, gens, L, K1, NB,
                     counts);
  hipError_t e = exclusive_scan_u32(counts, (static_cast<uint64_t>(1) << K1) * NB, scan, scan_ws, stream);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((expand_scatter_kernel<G, TAGGED>), dim3(NB), dim3(kRouteThreads), 0, stream, bank, F, gens, L,
                     K1, NB, scan, out);
  return hipGetLastError();
}

hipError_t sorted_expand_route(const uint64_t* bank, uint64_t F, const PackedGenerators& gens, const SortedLayout& L,
                               bool all_tagged, uint32_t K1, uint32_t NB, uint32_t* counts, uint32_t* scan,
                               uint32_t* scan_ws, uint64_t* out, hipStream_t stream) {
  if (K1 > kMaxPassBits || NB == 0) return hipErrorInvalidValue;
  switch (gens.count) {
#define MGBFS_G(k)                                                                                        \
  case k:                                                                                                 \
    return all_tagged ? expand_route_g<k, true>(bank, F, gens, L, K1, NB, counts, scan, scan_ws, out, stream) \
                      : expand_route_g<k, false>(bank, F, gens, L, K1, NB, counts, scan, scan_ws, out, stream);
    MGBFS_G(1) MGBFS_G(2) MGBFS_G(3) MGBFS_G(4) MGBFS_G(5) MGBFS_G(6) MGBFS_G(7) MGBFS_G(8)
    MGBFS_G(9) MGBFS_G(10) MGBFS_G(11) MGBFS_G(12) MGBFS_G(13) MGBFS_G(14) MGBFS_G(15) MGBFS_G(16)
#undef MGBFS_G
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t sorted_partition(const uint64_t* in, const uint32_t* scan1, uint32_t NB1, uint32_t K1, uint32_t K2,
                            uint32_t NC, uint32_t* counts, uint32_t* scan2, uint32_t* scan_ws, uint64_t* out,
                            hipStream_t stream) {
  if (K2 == 0 || K2 > kMaxPassBits || K1 + K2 > kMaxBucketBits || NC == 0) return hipErrorInvalidValue;
  const uint32_t blocks = (1u << K1) * NC;
  hipLaunchKernelGGL(partition_count_kernel, dim3(blocks), dim3(kRouteThreads), 0, stream, in, scan1, NB1, K1, K2, NC,
                     counts);
  hipError_t e = exclusive_scan_u32(counts, (static_cast<uint64_t>(1) << (K1 + K2)) * NC, scan2, scan_ws, stream);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(partition_scatter_kernel, dim3(blocks), dim3(kRouteThreads), 0, stream, in, scan1, NB1, K1, K2,
                     NC, scan2, out);
  return hipGetLastError();
}

hipError_t sorted_unique_diff(uint64_t* cand, const uint32_t* cscan, uint32_t NC, uint32_t K, const uint64_t* cur,
                              const uint32_t* cur_off, uint32_t Kc, const uint64_t* prev, const uint32_t* prev_off,
                              uint32_t Kp, uint32_t Ko, uint32_t* fine_counts, uint32_t* error,
                              unsigned long long* unique_total, uint32_t tagbits, uint32_t grid_limit,
                              hipStream_t stream) {
  if (K > Ko || Ko - K > kMaxFineBits || (cur && K > Kc) || (prev && K > Kp) || grid_limit == 0)
    return hipErrorInvalidValue;
  UniqueArgs a{cand, cscan, cur, cur_off, prev, prev_off, fine_counts, error, unique_total,
               NC, K, Kc, Kp, Ko, tagbits, 1u << K};
  const uint32_t grid = std::min<uint32_t>(1u << K, grid_limit);
  hipLaunchKernelGGL(unique_diff_kernel, dim3(grid), dim3(kUniqueThreads), 0, stream, a);
  return hipGetLastError();
}

hipError_t sorted_compact(const uint64_t* cand, const uint32_t* cscan, uint32_t NC, uint32_t K, uint32_t Ko,
                          const uint32_t* off, uint64_t* bank, hipStream_t stream) {
  hipLaunchKernelGGL(compact_kernel, dim3(1u << K), dim3(kRouteThreads), 0, stream, cand, cscan, NC, K, Ko, off, bank);
  return hipGetLastError();
}

}  // namespace mgbfs
