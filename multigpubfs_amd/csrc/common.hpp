// Shared device/host helpers for the MI355X exact BFS engine (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace mgbfs {

constexpr int kWave = 64;
constexpr uint32_t kMaxDegree = 16;     // nibble-packed permutations: n <= 16
constexpr uint32_t kMaxGenerators = 16; // packed engine generator table size

// Sticky device error word bits (caller-zeroed, OR-accumulated).
constexpr uint32_t kErrTableFull = 1u;      // open-addressing table has no free slot
constexpr uint32_t kErrVisitedCap = 2u;     // visited/row arena capacity exceeded
constexpr uint32_t kErrFutureCap = 4u;      // next-frontier capacity exceeded
constexpr uint32_t kErrBadRef = 8u;         // invalid row reference
constexpr uint32_t kErrCount = 32u;         // received count above bound
constexpr uint32_t kErrBucketCap = 64u;     // sorted engine: bucket above LDS capacity
constexpr uint32_t kErrDomain = 128u;       // state outside the packed/rank domain
constexpr uint32_t kErrLaneSlab = 256u;     // sorted engine: owner-lane slab full (retried with more room)

// Generator table passed by value: wave-uniform indices become scalar loads.
//
// Parent-edge pruning: a stored state may carry a small tag t = 1 + index of
// the inverse of the generator that produced it (0 = no tag).  Expanding that
// state skips generator t - 1: its child is the recorded parent, which lies
// in the previous layer and is already visited.  Any one parent edge is a
// valid tag, so the choice among duplicates does not matter.
struct PackedGenerators {
  uint32_t count;   // generators
  uint32_t degree;  // n
  uint8_t src[kMaxGenerators][kMaxDegree];  // child[i] = parent[src[g][i]]
  uint8_t child_tag[kMaxGenerators];        // tag written into a child made by g
  // Back-edge mask of a child made by g: bit b set for every b whose
  // generator is g's inverse (a repeated generator appears once per copy),
  // so the OR over a state's discoverers is exactly its set of generators
  // leading back to the previous layer.
  uint16_t child_mask[kMaxGenerators];
  uint32_t tag_shift;                        // bitmap backend: tag bit position (64 = off)
  // The same permutation as shift groups: child nibble i takes parent nibble
  // i + delta for every destination nibble in the group's mask, so a
  // generator costs one shift and one mask per distinct displacement
  // (2 for a rotation, 3 for a transposition of neighbours).  Groups with
  // delta >= 0 (right shifts) come first, `right` of them; the shift
  // amounts 4|delta| are packed four to a word so that a uniform group
  // index reads them with one scalar load.
  uint8_t groups[kMaxGenerators];
  uint8_t right[kMaxGenerators];
  uint32_t shift4[kMaxGenerators][kMaxDegree / 4];
  uint64_t gmask[kMaxGenerators][kMaxDegree];
};

__host__ __device__ inline uint32_t group_shift(const PackedGenerators& pg, int g, uint32_t j) {
  return (pg.shift4[g][j >> 2] >> (8u * (j & 3u))) & 0xffu;
}

__host__ inline void fill_shift_groups(PackedGenerators& pg) {
  for (uint32_t g = 0; g < pg.count; ++g) {
    int delta[kMaxDegree];
    uint64_t mask[kMaxDegree];
    uint32_t groups = 0;
    for (uint32_t i = 0; i < pg.degree; ++i) {
      const int d = static_cast<int>(pg.src[g][i]) - static_cast<int>(i);
      uint32_t j = 0;
      while (j < groups && delta[j] != d) ++j;
      if (j == groups) {
        delta[j] = d;
        mask[j] = 0;
        ++groups;
      }
      mask[j] |= 15ull << (4 * i);
    }
    pg.groups[g] = static_cast<uint8_t>(groups);
    pg.right[g] = 0;
    for (uint32_t w = 0; w < kMaxDegree / 4; ++w) pg.shift4[g][w] = 0;
    uint32_t out = 0;
    for (int pass = 0; pass < 2; ++pass)  // right shifts (delta >= 0) first
      for (uint32_t j = 0; j < groups; ++j) {
        if ((delta[j] >= 0) != (pass == 0)) continue;
        const uint32_t sh = 4u * static_cast<uint32_t>(delta[j] >= 0 ? delta[j] : -delta[j]);
        pg.shift4[g][out >> 2] |= sh << (8u * (out & 3u));
        pg.gmask[g][out] = mask[j];
        if (pass == 0) ++pg.right[g];
        ++out;
      }
  }
}

// Bijective mixer on B-bit keys (B = 4n): an odd multiplier mod 2^B and an
// xorshift by B/2+1 (self-inverse within B bits).  The sorted backend stores
// (mix(key) << (64 - B)) | tag, so the top bits are uniform bucket ids.
struct KeyMix {
  uint32_t bits, shift;
  uint64_t mask, m1, m2, i1, i2;
};

__host__ __device__ inline uint32_t nib(uint64_t key, uint32_t i) {
  return static_cast<uint32_t>((key >> (4u * i)) & 15u);
}

// child[i] = parent[src[i]] on nibble-packed keys, done on 32-bit halves.
__device__ __forceinline__ uint64_t apply_packed(uint64_t key, const uint8_t* src, uint32_t n) {
  const uint32_t lo = static_cast<uint32_t>(key);
  const uint32_t hi = static_cast<uint32_t>(key >> 32);
  uint32_t clo = 0, chi = 0;
#pragma unroll
  for (uint32_t i = 0; i < kMaxDegree; ++i) {
    if (i < n) {
      const uint32_t s = src[i];
      const uint32_t word = s < 8 ? lo : hi;
      const uint32_t v = (word >> (4u * (s & 7u))) & 15u;
      if (i < 8) clo |= v << (4u * i);
      else chi |= v << (4u * (i - 8));
    }
  }
  return (static_cast<uint64_t>(chi) << 32) | clo;
}

// (a & m) | c on both 32-bit halves, one v_and_or_b32 each (the compiler
// emits a separate AND and OR for this pattern); m is uniform.
__device__ __forceinline__ uint64_t and_or64(uint64_t a, uint64_t m, uint64_t c) {
  uint32_t lo, hi;
  asm("v_and_or_b32 %0, %1, %2, %3" : "=v"(lo) : "v"(static_cast<uint32_t>(a)), "s"(static_cast<uint32_t>(m)),
      "v"(static_cast<uint32_t>(c)));
  asm("v_and_or_b32 %0, %1, %2, %3" : "=v"(hi) : "v"(static_cast<uint32_t>(a >> 32)),
      "s"(static_cast<uint32_t>(m >> 32)), "v"(static_cast<uint32_t>(c >> 32)));
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

// Same result as apply_packed, from the shift groups of generator g, for N
// keys at once: each group's shift and mask are uniform (scalar loads) and
// are applied to all N keys before the next group is read.
template <int N>
__device__ __forceinline__ void apply_groups_n(const uint64_t (&key)[N], uint64_t (&child)[N],
                                               const PackedGenerators& pg, int g) {
#pragma unroll
  for (int m = 0; m < N; ++m) child[m] = 0;
  const uint32_t groups = pg.groups[g], right = pg.right[g];
  uint32_t j = 0;
  for (; j < right; ++j) {
    const uint32_t sh = group_shift(pg, g, j);
    const uint64_t mask = pg.gmask[g][j];
    if (sh == 0) {  // fixed nibbles (a uniform branch: a 64-bit shift takes two issue slots)
#pragma unroll
      for (int m = 0; m < N; ++m) child[m] = and_or64(key[m], mask, child[m]);
    } else {
#pragma unroll
      for (int m = 0; m < N; ++m) child[m] = and_or64(key[m] >> sh, mask, child[m]);
    }
  }
  for (; j < groups; ++j) {
    const uint32_t sh = group_shift(pg, g, j);
    const uint64_t mask = pg.gmask[g][j];
#pragma unroll
    for (int m = 0; m < N; ++m) child[m] = and_or64(key[m] << sh, mask, child[m]);
  }
}

__device__ __forceinline__ uint64_t apply_groups(uint64_t key, const PackedGenerators& pg, int g) {
  const uint64_t k[1] = {key};
  uint64_t c[1];
  apply_groups_n<1>(k, c, pg, g);
  return c[0];
}

__host__ inline uint64_t apply_packed_host(uint64_t key, const uint8_t* src, uint32_t n) {
  uint64_t child = 0;
  for (uint32_t i = 0; i < n; ++i) child |= ((key >> (4u * src[i])) & 15u) << (4u * i);
  return child;
}

// child_mask[a] = { b : src[b] o src[a] = identity }.  Returns false when
// some generator has no inverse in the set (the set is not inverse-closed).
__host__ inline bool fill_back_edge_masks(PackedGenerators& pg) {
  bool closed = true;
  for (uint32_t a = 0; a < pg.count; ++a) {
    uint32_t m = 0;
    for (uint32_t b = 0; b < pg.count; ++b) {
      bool inverse = true;
      for (uint32_t i = 0; i < pg.degree && inverse; ++i) inverse = pg.src[b][pg.src[a][i]] == i;
      if (inverse) m |= 1u << b;
    }
    pg.child_mask[a] = static_cast<uint16_t>(m);
    closed &= m != 0;
  }
  return closed;
}

// Lexicographic rank (position 0 most significant) of a nibble-packed
// permutation of 0..n-1.  Digit i = symbol - #smaller symbols already used.
// Sets *bad when the key is not a permutation of 0..n-1.
__host__ __device__ __forceinline__ uint64_t lehmer_rank_packed(uint64_t key, uint32_t n, bool* bad) {
  uint32_t seen = 0;
  uint64_t r = 0;
  bool invalid = false;
#pragma unroll
  for (uint32_t i = 0; i < kMaxDegree; ++i) {
    if (i < n) {
      const uint32_t s = nib(key, i);
      const uint32_t below = seen & ((1u << s) - 1u);
#if defined(__HIP_DEVICE_COMPILE__)
      const uint32_t d = s - static_cast<uint32_t>(__popc(below));
#else
      const uint32_t d = s - static_cast<uint32_t>(__builtin_popcount(below));
#endif
      invalid |= (s >= n) | ((seen >> s) & 1u);
      seen |= 1u << s;
      r = r * (n - i) + d;
    }
  }
  *bad = invalid;
  return r;
}

// Bijective 64-bit mixer (murmur3 fmix64 constants) and its inverse; the
// sorted engine orders states by mix(key) so buckets are balanced.
__host__ __device__ __forceinline__ uint64_t mix64(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdull;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ull;
  x ^= x >> 33;
  return x;
}

__host__ __device__ constexpr uint64_t inverse_odd(uint64_t a) {
  uint64_t x = a;  // Newton iteration for a^-1 mod 2^64 (a odd)
  for (int i = 0; i < 6; ++i) x *= 2 - a * x;
  return x;
}

__host__ __device__ __forceinline__ uint64_t unmix64(uint64_t x) {
  constexpr uint64_t i1 = inverse_odd(0xff51afd7ed558ccdull);
  constexpr uint64_t i2 = inverse_odd(0xc4ceb9fe1a85ec53ull);
  x ^= x >> 33;
  x *= i2;
  x ^= x >> 33;
  x *= i1;
  x ^= x >> 33;
  return x;
}

// Multiplicative (Fibonacci-style) hashing mod 2^B puts every key bit into
// the top bits (the bucket ids); the final xorshift folds the top half into
// the low bits used as LDS slot hashes.  Both steps are bijections on B bits.
// Bijection on `bits`-bit keys: multiply by an odd constant mod 2^bits,
// then xorshift by shift >= bits/2 (its own inverse).  The top bits (bucket
// ids) are those of the product (Fibonacci hashing), which spreads whole
// layers evenly; callers must not rely on keys that differ only in a few
// low nibbles landing far apart (the sorted backend samples its routing
// tiles across the whole frontier for that reason).
// Ranges of the top 32 bits: sub-bucket f of n holds x with
// floor(x n / 2^32) = f, i.e. x >= ceil(f 2^32 / n).  range_start computes
// that bound for 0 <= f < n <= 2^31 without a 64-bit division (dozens of
// scalar instructions on the GPU), from inv = range_inverse(n) =
// floor((2^64 - 1) / n): the product estimate floor(f inv / 2^32) is
// floor(f 2^32 / n) or one less, one comparison fixes it, a second rounds
// it up.
__host__ __device__ __forceinline__ uint64_t range_inverse(uint32_t n) { return ~0ull / n; }

__host__ __device__ __forceinline__ uint32_t range_start(uint32_t f, uint32_t n, uint64_t inv) {
  const uint64_t whole = static_cast<uint64_t>(f) << 32;
  const uint64_t lo = static_cast<uint64_t>(f) * static_cast<uint32_t>(inv);
  uint64_t x = static_cast<uint64_t>(f) * static_cast<uint32_t>(inv >> 32) + (lo >> 32);
  if ((x + 1) * n <= whole) ++x;
  if (x * n < whole) ++x;
  return static_cast<uint32_t>(x);
}

__host__ __device__ __forceinline__ uint64_t keymix(uint64_t x, const KeyMix& k) {
  x = (x * k.m1) & k.mask;
  x ^= x >> k.shift;
  return x;
}

__host__ __device__ __forceinline__ uint64_t keyunmix(uint64_t x, const KeyMix& k) {
  x ^= x >> k.shift;
  x = (x * k.i1) & k.mask;
  return x;
}

__host__ inline KeyMix make_keymix(uint32_t bits) {
  KeyMix k;
  k.bits = bits;
  k.shift = bits / 2 + 1;
  k.mask = bits >= 64 ? ~0ull : ((1ull << bits) - 1);
  // 2^64 / golden ratio, rounded to odd.
  k.m1 = 0x9e3779b97f4a7c15ull & k.mask;
  k.i1 = inverse_odd(0x9e3779b97f4a7c15ull) & k.mask;
  k.m2 = k.i2 = 0;
  return k;
}

__host__ inline uint64_t factorial_u64(uint32_t n) {
  uint64_t f = 1;
  for (uint32_t i = 2; i <= n; ++i) f *= i;
  return f;
}

}  // namespace mgbfs
