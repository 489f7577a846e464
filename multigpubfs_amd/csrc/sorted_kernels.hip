// Sort-unique backend with three rotating VRAM banks (inverse-closed sets).
//
// States are stored as h = (keymix(key) << tagbits) | tag (keymix is a
// bijection, so equality of the high bits is full-key equality).  Buckets
// are ranges of the top 32 bits x: sub-bucket floor(x nsub / 2^32) of a
// layer's nsub (any count, a multiple of 2^K2), fine bucket
// floor(x nbo / 2^32) with nbo = nsub << fb nested in it.  A bank holds one
// BFS layer grouped by fine bucket, with an offset table of nbo + 1
// entries.  One layer d -> d+1:
//   expand_scatter    : children of bank[d], bucketed into owner lanes of
//        2^K2 consecutive sub-buckets (LDS counting sort per 8192-element
//        tile).  Each tile
//        reserves its run in a lane slab with one atomic per lane; every
//        lane is split into P parts (block b uses part b % P) so no cursor
//        sees more than 1/P of the tiles.                    8F + 8C
//   partition_scatter : each lane part split into its 2^K2 sub-bucket
//        slabs of kUniqueCap entries (atomic run reservation).
//                                                             16C
//   unique_diff       : one workgroup per sub-bucket: LDS
//        open-addressing set of the candidates (in-layer dedup), probe with
//        the same sub-bucket of bank[d] and bank[d-1] (exact three-bank
//        retention for inverse-closed generators), survivors regrouped by
//        fine bucket and written back in place.           8C + 8F + 8F' + 8N
//   compact           : survivors -> bank[d+1].              16N
// A run that would overflow its slab is dropped and flagged; the host
// retries the layer with more slab room (lanes) or twice the sub-buckets
// (sub-buckets).  Order inside a slab depends on atomic order; the layer
// sets do not.
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
static_assert(kRouteDump >= kTile, "a dropped run fits the dump area");
constexpr int kUniqueThreads = 512;
constexpr int kUniquePerThread = kUniqueCap / kUniqueThreads;  // candidates held in registers
constexpr int kPrefetch = 4;  // independent loads issued before dependent LDS work

// Index of h among n equal ranges of its top 32 bits: floor(top32 * n / 2^32).
// Monotone in h, so a bank ordered by its top bits is ordered by any such
// index, and with n' = n << b every range of n' nests in one range of n.
__device__ __forceinline__ uint32_t range_index(uint64_t h, uint32_t n) {
  return __umulhi(static_cast<uint32_t>(h >> 32), n);
}

// Inclusive prefix sum over the 64 lanes of a wave: DPP row shifts by 1, 2,
// 4 and 8 (lanes shifted in from outside the 16-lane row read 0) give each
// row's prefix; row broadcasts of lane 15 into rows 1 and 3, then of lane
// 31 into rows 2 and 3, carry the row totals.  No lane predicates, no LDS.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  int x = static_cast<int>(v);
  x += __builtin_amdgcn_update_dpp(0, x, 0x111, 0xf, 0xf, true);   // row_shr:1
  x += __builtin_amdgcn_update_dpp(0, x, 0x112, 0xf, 0xf, true);   // row_shr:2
  x += __builtin_amdgcn_update_dpp(0, x, 0x114, 0xf, 0xf, true);   // row_shr:4
  x += __builtin_amdgcn_update_dpp(0, x, 0x118, 0xf, 0xf, true);   // row_shr:8
  x += __builtin_amdgcn_update_dpp(0, x, 0x142, 0xa, 0xf, false);  // row_bcast:15
  x += __builtin_amdgcn_update_dpp(0, x, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return static_cast<uint32_t>(x);
}

// ------------------------------------------------------------ expansion

// Shared tile machinery.  Each thread holds up to PER tile elements in
// registers (bit k of `valid` marks v[k]); they are ranked within their
// bucket by LDS atomics and placed bucket-contiguously in t.s.  Bucket b's
// run of n elements is reserved in its slab (which starts at slab_base(b))
// through runs.cursor(b); kDropRun: the slab is full and the run is
// dropped.  The run is then written contiguously.  Run offsets stay 32-bit
// so two 512-thread tiles fit one CU's LDS.
constexpr uint32_t kDropRun = ~0u;

struct TileLds {
  uint64_t s[kTile + kWave];  // + one parking slot per lane for invalid elements
  alignas(8) uint32_t off[1u << kMaxPassBits];
  // Bucket counters, then one spare counter per lane that ranks invalid
  // elements (its values are never used).
  alignas(8) uint32_t lhist[(1u << kMaxPassBits) + kWave];
  alignas(8) uint32_t lstart[1u << kMaxPassBits];
  alignas(16) uint32_t wave_sums[kRouteThreads / kWave];
};
static_assert((1u << kMaxPassBits) == 2 * kRouteThreads, "each routing thread owns two bucket counters");

// Clears every bucket counter (two per thread); entries at or above a
// pass's bucket count stay zero, so the tile scan runs over all of them.
__device__ __forceinline__ void clear_tile_counters(TileLds& t) {
  reinterpret_cast<uint2*>(t.lhist)[threadIdx.x] = make_uint2(0u, 0u);
}

// Exclusive scan of all bucket counters into t.lstart; thread t owns
// buckets 2t and 2t+1, returns their counts in c and clears them (no other
// thread reads them: the next tile's increments come after later barriers).
// Returns the tile's total; start receives the two buckets' starts.
__device__ __forceinline__ uint32_t tile_scan(TileLds& t, uint2& c, uint2& start) {
  c = reinterpret_cast<const uint2*>(t.lhist)[threadIdx.x];
  reinterpret_cast<uint2*>(t.lhist)[threadIdx.x] = make_uint2(0u, 0u);
  const uint32_t s = c.x + c.y;
  const uint32_t incl = wave_incl_scan(s);
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint32_t wave = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(threadIdx.x)) / kWave;  // uniform
  if (lane == kWave - 1) t.wave_sums[wave] = incl;
  __syncthreads();
  uint32_t before = 0, total = 0;
#pragma unroll
  for (uint32_t w = 0; w < kRouteThreads / kWave; ++w) {
    const uint32_t ws = t.wave_sums[w];
    before += w < wave ? ws : 0u;
    total += ws;
  }
  const uint32_t run = before + incl - s;
  start = make_uint2(run, run + c.x);
  reinterpret_cast<uint2*>(t.lstart)[threadIdx.x] = start;
  __syncthreads();
  return total;
}

// Destination indexing.  When the output array (plus a kTile-entry dump
// area at index `dump`) has fewer than 2^32 entries, t.off[b] holds bucket
// b's 32-bit destination index minus the run's LDS start, so the write-out
// is one add per element; a dropped run is pointed into the dump area (the
// layer is redone anyway, the writes only have to stay in bounds).
// Otherwise (dump = kWideOut) t.off[b] is the run's slab offset and the
// write-out computes 64-bit addresses and skips dropped runs.
constexpr uint32_t kWideOut = ~0u;

// Run cursors of a routing pass: cursor(b) counts the entries reserved in
// bucket b's slab; settle(off, n) checks a run reserved at off against the
// slab capacity (kDropRun and the error bit when it does not fit, or when
// n = 0).
struct RunCursors {
  uint32_t* base;
  uint64_t stride;  // cursor of bucket b at base[b * stride]
  uint32_t cap;
  uint32_t* error;
  uint32_t err_bit;
  __device__ __forceinline__ uint32_t* cursor(uint32_t b) const { return base + static_cast<uint64_t>(b) * stride; }
  __device__ __forceinline__ uint32_t settle(uint32_t off, uint32_t n) const {
    if (n == 0) return kDropRun;
    if (static_cast<uint64_t>(off) + n > cap) {
      atomicOr(error, err_bit);
      return kDropRun;
    }
    return off;
  }
};

// `between` runs after the tile is placed in LDS and before its write-out:
// the caller issues the next tile's loads there (v is no longer read), so
// they are in flight during the write-out.
template <int PER, typename BucketFn, typename BaseFn, typename BetweenFn>
__device__ __forceinline__ void route_regs(TileLds& t, const uint64_t (&v)[PER], uint32_t valid,
                                           BucketFn bkt, const RunCursors& runs, BaseFn slab_base,
                                           uint64_t* __restrict__ out, uint32_t dump, BetweenFn between) {
  // No branches per element: an invalid element takes its rank from the
  // lane's spare counter and is parked in the lane's slot past the tile.
  const uint32_t lane = threadIdx.x & (kWave - 1);
  uint32_t rank[PER];
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    uint32_t b = bkt(v[k]);
    asm volatile("" : "+v"(b));  // keeps the choice below a select, not a branch
    rank[k] = atomicAdd(&t.lhist[((valid >> k) & 1u) ? b : (1u << kMaxPassBits) + lane], 1u);
  }
  __syncthreads();
  uint2 n, st;
  const uint32_t ne = tile_scan(t, n, st);
  // Run reservations first: both atomics go out before the LDS placement
  // and their offsets are checked after it, so the round trips overlap the
  // placement.  (A check right after each atomic branches on its result and
  // waits for it there.)  Buckets without a run issue no atomic: a tile's
  // children cluster into few buckets.
  const uint32_t b0 = 2 * threadIdx.x, b1 = b0 + 1;
  uint32_t off0 = 0, off1 = 0;
  if (n.x) off0 = atomicAdd(runs.cursor(b0), n.x);
  if (n.y) off1 = atomicAdd(runs.cursor(b1), n.y);
  // All bucket starts are read first (in bounds for any element) and pinned,
  // so the reads go out together instead of one guarded read per element.
  uint32_t pos[PER];
#pragma unroll
  for (int k = 0; k < PER; ++k) pos[k] = t.lstart[bkt(v[k])];
#pragma unroll
  for (int k = 0; k < PER; ++k) asm volatile("" : "+v"(pos[k]));
  // The element predicates are taken afresh from the mask here: reusing the
  // rank step's keeps 2 x PER SGPRs live across the scan, and they spill.
  uint32_t vmask = valid;
  asm volatile("" : "+v"(vmask));
#pragma unroll
  for (int k = 0; k < PER; ++k) t.s[((vmask >> k) & 1u) ? pos[k] + rank[k] : kTile + lane] = v[k];
  asm volatile("" : "+v"(off0), "+v"(off1));  // the offsets are used only from here on
  off0 = runs.settle(off0, n.x);
  off1 = runs.settle(off1, n.y);
  if (dump != kWideOut) {
    auto dest = [&](uint32_t b, uint32_t off, uint32_t start) {
      return (off == kDropRun ? dump : static_cast<uint32_t>(slab_base(b)) + off) - start;
    };
    reinterpret_cast<uint2*>(t.off)[threadIdx.x] = make_uint2(dest(b0, off0, st.x), dest(b1, off1, st.y));
  } else {
    reinterpret_cast<uint2*>(t.off)[threadIdx.x] = make_uint2(off0, off1);
  }
  __syncthreads();
  between();
  if (dump != kWideOut) {
    // Rows go out in rounds of kRound per thread: the round's rows are read
    // at once (rows at or past ne hold stale entries whose bucket index is
    // still in range), then their run offsets, then the rows below ne are
    // stored.  A loop over single rows waits on two dependent LDS reads per
    // row; a single round of all kTilePerThread rows costs the registers of
    // the second workgroup per CU.
    constexpr int kRound = kTilePerThread / 2;
    for (uint32_t r = 0; r < kTile && r < ne; r += kRound * kRouteThreads) {
      uint64_t h[kRound];
      uint32_t o[kRound];
#pragma unroll
      for (int u = 0; u < kRound; ++u) h[u] = t.s[r + threadIdx.x + u * kRouteThreads];
#pragma unroll
      for (int u = 0; u < kRound; ++u) o[u] = t.off[bkt(h[u])];
#pragma unroll
      for (int u = 0; u < kRound; ++u) {
        const uint32_t j = r + threadIdx.x + u * kRouteThreads;
        if (j < ne) out[o[u] + j] = h[u];  // 32-bit index arithmetic, wraps as intended
      }
    }
  } else {
    for (uint32_t j = threadIdx.x; j < ne; j += kRouteThreads) {
      const uint64_t h = t.s[j];
      const uint32_t b = bkt(h);
      const uint32_t off = t.off[b];
      if (off != kDropRun) out[slab_base(b) + off + (j - t.lstart[b])] = h;
    }
  }
  // No barrier here: the next tile's placement and scan come after its
  // first barrier, which every thread reaches only after its write-out.
}

// A parent's key for the shift-group expansion: keyunmix without its final
// mask, so bits above the key's 4n bits are left as the multiply leaves
// them.  Every shift group keeps only destination nibbles whose source
// nibble lies below n, so those bits never reach a child.
__device__ __forceinline__ uint64_t parent_key(uint64_t h, const SortedLayout& L) {
  uint64_t x = h >> L.tagbits;
  x ^= x >> L.km.shift;
  return x * L.km.i1;
}

// The stored value (keymix(child) << tagbits) | tag.  The layout's key
// width is at least 32 bits, so keymix's mask is applied to the high word
// only (its low word is all ones).
__device__ __forceinline__ uint64_t stored_value(uint64_t child, uint64_t tag, const SortedLayout& L) {
  uint64_t x = child * L.km.m1;
  x &= (static_cast<uint64_t>(static_cast<uint32_t>(L.km.mask >> 32)) << 32) | 0xffffffffull;
  x ^= x >> L.km.shift;
  return (x << L.tagbits) | tag;
}

// MASKED: every parent carries a non-empty back-edge mask (all layers after
// the start when back-edge pruning is on) and yields at most G-1 children,
// so kTilePerThread / (G-1) parents fit a tile; each parent keeps G child
// slots in registers, marked valid unless its generator is in the mask.
template <int G, bool MASKED>
struct ExpandTile {
  static constexpr int kMaxKids = MASKED ? (G > 1 ? G - 1 : 1) : G;
  static constexpr int kParentsPerThread = (kTilePerThread / kMaxKids) > 0 ? (kTilePerThread / kMaxKids) : 1;
  static constexpr int kRegs = kParentsPerThread * G;  // <= 32 (valid mask width)
  static constexpr int kParents = kParentsPerThread * kRouteThreads;
  static_assert(kParentsPerThread * kMaxKids <= kTilePerThread, "valid children fit the LDS tile");
};

template <int G, bool MASKED>
__global__ __launch_bounds__(kRouteThreads) void expand_scatter_kernel(const uint64_t* __restrict__ bank, uint64_t F,
                                                                        PackedGenerators gens, SortedLayout L,
                                                                        uint32_t nsub, uint32_t K2, uint32_t NB,
                                                                        uint32_t P,
                                                                        uint32_t cap, uint32_t stride,
                                                                        uint32_t err_bit,
                                                                        uint32_t* __restrict__ cursors,
                                                                        uint32_t* __restrict__ error,
                                                                        uint64_t* __restrict__ out,
                                                                        uint32_t dump) {
  using T = ExpandTile<G, MASKED>;
  static_assert(T::kRegs <= 32, "valid mask is 32 bits");
  __shared__ TileLds t;
  const uint32_t part = blockIdx.x % P;
  clear_tile_counters(t);
  __syncthreads();
  const RunCursors runs{cursors + part, P, cap, error, err_bit};
  auto slab_base = [=](uint32_t b) { return (static_cast<uint64_t>(b) * P + part) * stride; };
  auto bkt = [nsub, K2](uint64_t h) { return range_index(h, nsub) >> K2; };
  // Tiles of contiguous parents go to blocks round-robin (tile i to block
  // i mod NB), so every lane part samples the whole frontier.
  const uint64_t tiles = (F + T::kParents - 1) / T::kParents;
  // Parents are loaded at the top of their own tile: loading the next
  // tile's during the write-out (as the partition does) takes
  // expand_scatter past 128 VGPRs, one workgroup per CU.
  auto next = [] {};
  for (uint64_t tile = blockIdx.x; tile < tiles; tile += NB) {
    const uint64_t t0 = tile * T::kParents;
    // Loads past the frontier re-read its last parent (F >= 1), so every
    // load is in bounds and none is predicated; `live` drops them.
    uint64_t pv[T::kParentsPerThread];
    uint32_t live = 0;
#pragma unroll
    for (int m = 0; m < T::kParentsPerThread; ++m) {
      const uint64_t q = t0 + static_cast<uint64_t>(m) * kRouteThreads + threadIdx.x;
      live |= (q < F ? 1u : 0u) << m;
      pv[m] = bank[q < F ? q : F - 1];
    }
    uint32_t valid = 0;
    bool empty_mask = false;
    uint64_t key[T::kParentsPerThread];
#pragma unroll
    for (int m = 0; m < T::kParentsPerThread; ++m) {
      key[m] = parent_key(pv[m], L);
      uint32_t want = (1u << G) - 1u;
      if (MASKED) {
        const uint32_t skip = static_cast<uint32_t>(pv[m] & L.tagmask);
        // A discovered state always has a back-edge; an empty mask would
        // let G children into a tile sized for G-1 per parent.
        empty_mask |= ((live >> m) & 1u) && skip == 0;
        want = skip == 0 ? 0u : (~skip & want);
      }
      valid |= ((live >> m) & 1u ? want : 0u) << (m * G);
    }
    if (MASKED && empty_mask) atomicOr(error, kErrDomain);
    // Generator-major: each shift group is read once and applied to all of
    // this thread's parents.  With back-edge masks the child's tag is
    // child_mask[g]: the bits of every copy of g's inverse, the generators
    // leading back here.
    if constexpr (MASKED && G == 3) {
      // Two slots per parent (its mask leaves at most two children): slot 0
      // takes the first unmasked child, slot 1 the second, so only those are
      // mixed and routed (a third of the child slots fewer).
      constexpr int PP = T::kParentsPerThread;
      uint64_t w[2 * PP];   // child keys, then stored values
      uint32_t wt[2 * PP];  // their tags
      uint32_t wvalid = 0;
      uint64_t child[PP];
      const uint32_t tag0 = gens.child_mask[0], tag1 = gens.child_mask[1], tag2 = gens.child_mask[2];
      apply_groups_n(key, child, gens, 0);
#pragma unroll
      for (int m = 0; m < PP; ++m) w[2 * m] = child[m];
      apply_groups_n(key, child, gens, 1);
#pragma unroll
      for (int m = 0; m < PP; ++m) {
        const bool first = !((valid >> (3 * m)) & 1u);  // child 0 masked: child 1 takes slot 0
        w[2 * m] = first ? child[m] : w[2 * m];
        wt[2 * m] = first ? tag1 : tag0;
        w[2 * m + 1] = child[m];
      }
      apply_groups_n(key, child, gens, 2);
#pragma unroll
      for (int m = 0; m < PP; ++m) {
        const uint32_t ok = (valid >> (3 * m)) & 7u;
        const bool only = ok == 4u, second = (ok & 4u) && !only;
        w[2 * m] = only ? child[m] : w[2 * m];
        wt[2 * m] = only ? tag2 : wt[2 * m];
        w[2 * m + 1] = second ? child[m] : w[2 * m + 1];
        wt[2 * m + 1] = second ? tag2 : tag1;
        wvalid |= ((1u << __popc(ok)) - 1u) << (2 * m);  // <= 2 children (non-empty mask)
      }
#pragma unroll
      for (int k = 0; k < 2 * PP; ++k) w[k] = stored_value(w[k], wt[k], L);
      route_regs(t, w, wvalid, bkt, runs, slab_base, out, dump, next);
    } else {
      uint64_t v[T::kRegs];
#pragma unroll
      for (int g = 0; g < G; ++g) {
        uint64_t child[T::kParentsPerThread];
        apply_groups_n(key, child, gens, g);
        const uint64_t tag = gens.child_mask[g];
#pragma unroll
        for (int m = 0; m < T::kParentsPerThread; ++m) v[m * G + g] = stored_value(child[m], tag, L);
      }
      route_regs(t, v, valid, bkt, runs, slab_base, out, dump, next);
    }
  }
}

// ------------------------------------------------------------ partition

// One block per lane part (lane = blockIdx.x / P): its lcount[] elements at
// slab (lane * P + part) * lcap are routed by the next K2 bits into the
// sub-bucket slabs ((lane << K2) + b) * kSlabStride.
__global__ __launch_bounds__(kRouteThreads) void partition_scatter_kernel(const uint64_t* __restrict__ in,
                                                                           const uint32_t* __restrict__ lcount,
                                                                           uint32_t lcap, uint32_t P, uint32_t nsub,
                                                                           uint32_t K2,
                                                                           uint32_t* __restrict__ scount,
                                                                           uint32_t* __restrict__ error,
                                                                           uint64_t* __restrict__ out,
                                                                           uint32_t dump) {
  __shared__ TileLds t;
  clear_tile_counters(t);
  __syncthreads();
  const uint64_t slab = blockIdx.x;
  const uint64_t r0 = slab * lcap, r1 = r0 + min(lcount[slab], lcap);
  auto bkt = [nsub, K2](uint64_t h) { return range_index(h, nsub) & ((1u << K2) - 1u); };
  const uint64_t first = static_cast<uint64_t>(blockIdx.x / P) << K2;  // first sub-bucket of this lane
  const RunCursors runs{scount + first, 1, kUniqueCap, error, kErrBucketCap};
  auto slab_base = [=](uint32_t b) { return (first + b) * kSlabStride; };
  // Loads past the slab's count re-read its last entry (t0 < r1), so none
  // is under a branch and each rank atomic waits only for its own load.  A
  // tile's entries are loaded during the previous tile's write-out.
  uint64_t v[kTilePerThread];
  uint32_t valid = 0;
  auto load_tile = [&](uint64_t t0) {
    valid = 0;
#pragma unroll
    for (int k = 0; k < kTilePerThread; ++k) {
      const uint64_t i = t0 + static_cast<uint64_t>(k) * kRouteThreads + threadIdx.x;
      v[k] = in[i < r1 ? i : r1 - 1];
      if (i < r1) valid |= 1u << k;
    }
  };
  if (r0 < r1) load_tile(r0);
  for (uint64_t t0 = r0; t0 < r1; t0 += kTile) {
    auto next = [&] {
      if (t0 + kTile < r1) load_tile(t0 + kTile);
    };
    route_regs(t, v, valid, bkt, runs, slab_base, out, dump, next);
  }
}

// ------------------------------------------------------ unique + difference

constexpr uint32_t kFilterWords = 2048;  // 64 Ki bits

struct UniqueLds {
  alignas(16) uint64_t table[kUniqueSlots];  // read two slots at a time
  alignas(16) uint32_t filter[kFilterWords];  // one bit per candidate state (key bits 13..28)
  uint32_t fhist[(1u << kMaxFineBits) + kWave];  // survivors, then write cursors, per fine bucket;
                                                  // then one spare counter per lane
  uint32_t unique;
};

struct UniqueArgs {
  uint64_t* cand;          // sub-bucket slabs of kUniqueCap entries
  const uint32_t* scount;  // entries reserved per slab (> kUniqueCap: overflowed)
  const uint64_t* cur;
  const uint32_t* cur_off;
  const uint64_t* prev;
  const uint32_t* prev_off;
  uint32_t* fine_counts;
  uint32_t* error;
  unsigned long long* unique_total;
  unsigned long long* cand_total;  // candidates seen (pruned expansions vary per parent)
  uint64_t tagmask;  // (1 << tagbits) - 1
  uint64_t mark;     // top tag bit, never set in a candidate: "present in a bank"
  uint32_t merge;    // OR the back-edge masks of duplicates into the kept entry
  uint32_t fb, nbo_c, nbo_p, tagbits, buckets;  // buckets: sub-buckets of this layer
  uint64_t range_inv;                           // range_inverse(buckets)
};

constexpr uint64_t kEmpty = ~0ull;

// Bank entry h: if its state is in the table, set `mark` in that slot's
// tag.  Each slot matches at most one bank entry (a bank holds unique
// states; two banks are disjoint layers), so a plain store suffices;
// concurrent probes passing the slot see the same state bits either way.
__device__ __forceinline__ void probe_value(UniqueLds& u, uint64_t h, const UniqueArgs& a, uint32_t mask) {
  const uint64_t key = h | a.tagmask;
  uint32_t slot = static_cast<uint32_t>(h >> a.tagbits) & mask;
  for (uint32_t probe = 0; probe <= mask; ++probe) {
    const uint64_t v = u.table[slot];
    if (v == kEmpty) return;
    if ((v | a.tagmask) == key) {
      u.table[slot] = v | a.mark;
      return;
    }
    slot = (slot + 1) & mask;
  }
}

// Bank entries [lo + skip, hi), skip = entries per thread already probed
// from the prefetch, in chunks of kPrefetch loads per thread.
__device__ __forceinline__ void probe_rest(UniqueLds& u, const uint64_t* __restrict__ bank, uint64_t lo, uint64_t hi,
                                           uint32_t skip, const UniqueArgs& a, uint32_t mask) {
  for (uint64_t base = lo + threadIdx.x + static_cast<uint64_t>(skip) * kUniqueThreads; base < hi;
       base += kPrefetch * kUniqueThreads) {
    uint64_t v[kPrefetch];
#pragma unroll
    for (int k = 0; k < kPrefetch; ++k) {
      const uint64_t i = base + static_cast<uint64_t>(k) * kUniqueThreads;
      v[k] = i < hi ? bank[i] : 0;
    }
#pragma unroll
    for (int k = 0; k < kPrefetch; ++k)
      if (base + static_cast<uint64_t>(k) * kUniqueThreads < hi) probe_value(u, v[k], a, mask);
  }
}

// Bounds of sub-bucket f: candidates [s, e) of its slab, current-bank slice
// [clo, chi), previous-bank slice [plo, phi).
struct BucketBounds {
  uint64_t s, e;
  uint32_t clo, chi, plo, phi;
};

// A sub-bucket's bounds in two steps: the loads, then the arithmetic on
// their values, for a sub-bucket f < a.buckets with a current bank (and,
// when kPrev, a previous one).  Unconditional loads: a load under a branch leaves a
// register copy at the merge, and the copy waits for the load.
struct RawBounds {
  uint32_t count, clo, chi, plo, phi;
};

template <bool kPrev>
__device__ __forceinline__ RawBounds load_raw_bounds(const UniqueArgs& a, uint32_t f) {
  // Sub-bucket f holds top-32 values [range_start(f), range_start(f + 1));
  // a bank's slice is every fine bucket meeting that range (a boundary
  // bucket may add states of a neighbour, which match nothing here).
  const uint32_t n = a.buckets;
  const uint32_t xlo = range_start(f, n, a.range_inv);
  const bool last = f + 1 == n;
  const uint32_t xhi = last ? 0u : range_start(f + 1, n, a.range_inv) - 1u;
  RawBounds r{a.scount[f], a.cur_off[__umulhi(xlo, a.nbo_c)],
              a.cur_off[last ? a.nbo_c : __umulhi(xhi, a.nbo_c) + 1], 0, 0};
  if constexpr (kPrev) {
    r.plo = a.prev_off[__umulhi(xlo, a.nbo_p)];
    r.phi = a.prev_off[last ? a.nbo_p : __umulhi(xhi, a.nbo_p) + 1];
  }
  return r;
}

__device__ __forceinline__ BucketBounds finish_bounds(uint32_t f, const RawBounds& r) {
  const uint64_t s = static_cast<uint64_t>(f) * kSlabStride;
  return BucketBounds{s, s + r.count, r.clo, r.chi, r.plo, r.phi};
}

// The first bank entries: 2 kPrefetch per thread, all from the current
// bank when there is no previous one (back-edge masks), else kPrefetch from
// each.  Slot k holds entry slice_lo + threadIdx.x + j * kUniqueThreads of
// its bank.
constexpr int kBankPrefetch = 2 * kPrefetch;

struct BucketData {
  uint64_t v[kUniquePerThread];
  uint64_t bpre[kBankPrefetch];
};

__device__ __forceinline__ int cur_prefetch(const UniqueArgs& a) { return a.prev ? kPrefetch : kBankPrefetch; }

__device__ __forceinline__ bool bank_slot(const UniqueArgs& a, const BucketBounds& b, int k, uint64_t* index) {
  const int ncur = cur_prefetch(a);
  const bool cur = k < ncur;
  const uint64_t i = (cur ? b.clo : b.plo) + threadIdx.x + static_cast<uint64_t>(cur ? k : k - ncur) * kUniqueThreads;
  *index = i;
  return i < (cur ? b.chi : b.phi);
}

// All kUniqueCap slots of slab f: the slab start is known without the
// bounds, so these loads go out beside the bounds loads and one dependent
// round trip leaves the critical path.  Slots past the count hold stale
// values and are ignored.
__device__ __forceinline__ void load_slab(const UniqueArgs& a, uint32_t f, BucketData& d) {
  const uint64_t s = static_cast<uint64_t>(f) * kSlabStride;
#pragma unroll
  for (int k = 0; k < kUniquePerThread; ++k) d.v[k] = a.cand[s + k * kUniqueThreads + threadIdx.x];
}

// Unconditional loads: a slot past its slice reads the sub-bucket's own
// slab instead (always allocated; the value is never used).  With the loads
// under branches the number in flight at the insert loop is unknown there,
// so its first use of a candidate register waited for every bank load too.
__device__ __forceinline__ void load_bank(const UniqueArgs& a, const BucketBounds& b, BucketData& d) {
  const int ncur = cur_prefetch(a);
  const uint64_t* spare = a.cand + b.s + threadIdx.x;
#pragma unroll
  for (int k = 0; k < kBankPrefetch; ++k) {
    uint64_t i;
    const bool live = bank_slot(a, b, k, &i);
    const uint64_t* bank = k < ncur ? a.cur : a.prev;
    d.bpre[k] = *(live ? bank + i : spare);
  }
}

// Filter bit of a stored value: key bits 13..28, above the table-slot bits.
__device__ __forceinline__ uint32_t filter_bit(uint64_t h, uint32_t tagbits) {
  return static_cast<uint32_t>((h >> tagbits) >> 13) & (kFilterWords * 32 - 1);
}

// Entry i & (kBankPrefetch - 1) of a lane's prefetched bank entries, by a
// select tree on the index bits (register arrays cannot be indexed per lane).
__device__ __forceinline__ uint64_t pick_entry(const uint64_t (&e)[kBankPrefetch], uint32_t i) {
  static_assert(kBankPrefetch == 8, "three index bits");
  const bool b0 = i & 1u, b1 = i & 2u, b2 = i & 4u;
  const uint64_t a0 = b0 ? e[1] : e[0], a1 = b0 ? e[3] : e[2], a2 = b0 ? e[5] : e[4], a3 = b0 ? e[7] : e[6];
  const uint64_t c0 = b1 ? a1 : a0, c1 = b1 ? a3 : a2;
  return b2 ? c1 : c0;
}

constexpr int kPairsPerThread = kUniqueSlots / 2 / kUniqueThreads;  // table slot pairs per thread

// Block x handles sub-buckets x, x + grid, ... (the engine launches one
// block per sub-bucket).  The candidates and the first bank chunks are
// loaded before the LDS table is built.  Phases: insert every candidate
// (in-layer dedup; state equality ignores tags) -> probe the retained
// banks, marking matched slots -> the occupied, unmarked slots are the
// survivors: count them per fine bucket, scan, and write them back into
// the slab grouped by fine bucket.
template <bool kPrev>
__global__ __launch_bounds__(kUniqueThreads) __attribute__((amdgpu_waves_per_eu(4, 4))) void unique_diff_kernel(
    UniqueArgs a) {
  __shared__ UniqueLds u;
  const uint32_t fine_bits = a.fb, fine = 1u << fine_bits;
  const uint32_t nbo = a.buckets << fine_bits;  // fine buckets of the layer written
  // Fine bucket of a survivor of sub-bucket f, within f.  A layer whose
  // lane slab overflowed holds stale entries of other sub-buckets (the
  // layer is redone); clamping keeps their counter index in range.
  auto fine_of = [nbo, fine_bits, fine](uint64_t h, uint32_t f) {
    const uint32_t d = range_index(h, nbo) - (f << fine_bits);
    return d < fine ? d : fine - 1u;
  };
  unsigned long long unique_block = 0, cand_block = 0;
  for (uint32_t f = blockIdx.x; f < a.buckets; f += gridDim.x) {
    // The bounds' loads go out first, then the slab's, and only then are the
    // bounds used: waiting for them (to address the bank slice) leaves the
    // 40 KB slab in flight, so the dependent bank loads go out while it
    // streams in.
    RawBounds raw = load_raw_bounds<kPrev>(a, f);
    BucketData dcur;
    load_slab(a, f, dcur);
    // The empty asm "redefines" the bounds after the slab loads are issued,
    // so no use of them (and no wait for them) is scheduled above the loads.
    if constexpr (kPrev)
      asm volatile("" : "+v"(raw.count), "+v"(raw.clo), "+v"(raw.chi), "+v"(raw.plo), "+v"(raw.phi)::"memory");
    else
      asm volatile("" : "+v"(raw.count), "+v"(raw.clo), "+v"(raw.chi)::"memory");
    const BucketBounds bcur = finish_bounds(f, raw);
    load_bank(a, bcur, dcur);
    const uint64_t s = bcur.s, e = bcur.e;
    const uint64_t c = e - s;
    if (c > kUniqueCap) {
      if (threadIdx.x == 0) atomicOr(a.error, kErrBucketCap);
      for (uint32_t i = threadIdx.x; i < fine; i += kUniqueThreads)
        a.fine_counts[(static_cast<uint64_t>(f) << fine_bits) + i] = 0;
      continue;
    }
    // Table sized to the bucket: slots = next_pow2(2c) (load factor <= 0.5),
    // at most kUniqueSlots (load factor <= kUniqueCap / kUniqueSlots).
    uint32_t slots = 64;
    while (slots < 2 * c && slots < kUniqueSlots) slots <<= 1;
    const uint32_t mask = slots - 1;
    ulonglong2* table2 = reinterpret_cast<ulonglong2*>(u.table);
    for (uint32_t i = threadIdx.x; i < slots / 2; i += kUniqueThreads) table2[i] = make_ulonglong2(kEmpty, kEmpty);
    for (uint32_t i = threadIdx.x; i < kFilterWords / 4; i += kUniqueThreads)
      reinterpret_cast<uint4*>(u.filter)[i] = make_uint4(0, 0, 0, 0);
    for (uint32_t i = threadIdx.x; i < fine; i += kUniqueThreads) u.fhist[i] = 0;
    if (threadIdx.x == 0) u.unique = 0;
    __syncthreads();
    // Each lane works through its own list of candidates, one CAS per round,
    // moving to its next candidate as soon as the current one is placed, so
    // a wave's rounds follow its busiest lane's total probe count instead of
    // the sum of per-candidate maxima.  No stored value is all ones (the
    // mark bit is clear in candidates and bank entries), so kEmpty is only
    // ever an empty slot.
    {
      uint32_t left = c > threadIdx.x ? (static_cast<uint32_t>(c) - threadIdx.x + kUniqueThreads - 1) / kUniqueThreads : 0;
      // Every candidate's filter bit, before any bank entry tests it.
#pragma unroll
      for (int k = 0; k < kUniquePerThread; ++k)
        if (k < static_cast<int>(left)) {
          const uint32_t b = filter_bit(dcur.v[k], a.tagbits);
          atomicOr(&u.filter[b >> 5], 1u << (b & 31));
        }
      uint64_t h = dcur.v[0];
      uint32_t slot = static_cast<uint32_t>(h >> a.tagbits) & mask;
      while (left) {
        const uint64_t old = atomicCAS(reinterpret_cast<unsigned long long*>(&u.table[slot]), kEmpty, h);
        const bool dup = old != kEmpty && (old | a.tagmask) == (h | a.tagmask);
        if (dup && a.merge) atomicOr(reinterpret_cast<unsigned long long*>(&u.table[slot]), h & a.tagmask);
        // The list advance as selects: a branch per round costs more scalar
        // exec-mask work than the selects cost vector work.
        const bool done = old == kEmpty || dup;
        left -= done ? 1u : 0u;
#pragma unroll
        for (int k = 0; k + 1 < kUniquePerThread; ++k) dcur.v[k] = done ? dcur.v[k + 1] : dcur.v[k];
        h = dcur.v[0];
        slot = done ? static_cast<uint32_t>(h >> a.tagbits) & mask : (slot + 1) & mask;
      }
    }
    __syncthreads();
    // Difference against the retained layers (same sub-bucket only).  Of
    // the prefetched entries only the live ones are probed: inside a bank's
    // slice and with their filter bit set (some candidate may have that
    // state).  Each round probes the lane's lowest live entry, picked by its
    // index, so a lane's rounds follow its live entries and their probe
    // lengths, not its list length.
    {
      uint32_t live = 0;
#pragma unroll
      for (int k = 0; k < kBankPrefetch; ++k) {
        uint64_t i;
        if (bank_slot(a, bcur, k, &i)) {
          const uint32_t b = filter_bit(dcur.bpre[k], a.tagbits);
          if ((u.filter[b >> 5] >> (b & 31)) & 1u) live |= 1u << k;
        }
      }
      uint64_t h = pick_entry(dcur.bpre, __builtin_ctz(live | (1u << kBankPrefetch)));
      uint32_t slot = static_cast<uint32_t>(h >> a.tagbits) & mask;
      while (live) {
        const uint64_t v = u.table[slot];
        const bool match = v != kEmpty && (v | a.tagmask) == (h | a.tagmask);
        if (match) u.table[slot] = v | a.mark;
        const bool done = v == kEmpty || match;
        live = done ? live & (live - 1) : live;
        h = done ? pick_entry(dcur.bpre, __builtin_ctz(live | (1u << kBankPrefetch))) : h;
        slot = done ? static_cast<uint32_t>(h >> a.tagbits) & mask : (slot + 1) & mask;
      }
    }
    const int ncur = cur_prefetch(a);
    probe_rest(u, a.cur, bcur.clo, bcur.chi, ncur, a, mask);
    if (a.prev) probe_rest(u, a.prev, bcur.plo, bcur.phi, kBankPrefetch - ncur, a, mask);
    __syncthreads();
    // Survivors are the occupied, unmarked slots; each entry carries the
    // merged back-edge mask of its duplicates.  Thread t owns slot pairs
    // t + j * kUniqueThreads (conflict-free 16-byte reads).  All its pairs
    // are read first (pairs past a smaller table hold stale entries and are
    // masked out), then every slot's counting atomic goes out: a survivor's
    // on its fine bucket's counter, whose return value is its rank there
    // (kept, 16 bits each, two per pair, for the write-out), any other
    // slot's on the lane's spare counter.  One guarded read-then-atomics
    // step per pair waited on three dependent LDS round trips per pair.
    const uint32_t lane = threadIdx.x & (kWave - 1);
    const uint32_t spare = (1u << kMaxFineBits) + lane;
    uint32_t survive = 0, occupied = 0;
    uint32_t rank[kPairsPerThread];
    ulonglong2 pairs[kPairsPerThread];
#pragma unroll
    for (int j = 0; j < kPairsPerThread; ++j) pairs[j] = table2[threadIdx.x + j * kUniqueThreads];
#pragma unroll
    for (int j = 0; j < kPairsPerThread; ++j) {
      const bool in = 2 * (threadIdx.x + j * kUniqueThreads) < slots;
      const ulonglong2 p = pairs[j];
      const bool ox = in && p.x != kEmpty, oy = in && p.y != kEmpty;
      const bool sx = ox && !(p.x & a.mark), sy = oy && !(p.y & a.mark);
      occupied += (ox ? 1u : 0u) + (oy ? 1u : 0u);
      survive |= ((sx ? 1u : 0u) | (sy ? 2u : 0u)) << (2 * j);
      const uint32_t rx = atomicAdd(&u.fhist[sx ? fine_of(p.x, f) : spare], 1u);
      const uint32_t ry = atomicAdd(&u.fhist[sy ? fine_of(p.y, f) : spare], 1u);
      rank[j] = (rx & 0xffffu) | (ry << 16);
    }
    // One LDS atomic per wave for the occupied-slot count (the last lane of
    // the wave's prefix sum holds the wave's total).
    occupied = wave_incl_scan(occupied);
    if ((threadIdx.x & (kWave - 1)) == kWave - 1 && occupied) atomicAdd(&u.unique, occupied);
    __syncthreads();
    // One wave turns the fine counts into the fine buckets' first rows
    // (exclusive scan by shuffles, <= 16 counters per lane) while the
    // others wait at a single barrier.
    if (threadIdx.x < kWave) {
      const uint32_t per = (fine + kWave - 1) / kWave, base = threadIdx.x * per;
      uint32_t sum = 0;
      for (uint32_t j = 0; j < per; ++j)
        if (base + j < fine) sum += u.fhist[base + j];
      uint32_t run = wave_incl_scan(sum) - sum;
      for (uint32_t j = 0; j < per; ++j)
        if (base + j < fine) {
          const uint32_t c = u.fhist[base + j];
          a.fine_counts[(static_cast<uint64_t>(f) << fine_bits) + base + j] = c;
          u.fhist[base + j] = run;
          run += c;
        }
    }
    __syncthreads();
    // Survivors are written back in place at their fine bucket's first row
    // plus their rank: every slot's first-row read goes out first (a stale
    // or empty slot's fine index is clamped into range), then the stores.
    uint32_t first[2 * kPairsPerThread];
#pragma unroll
    for (int j = 0; j < kPairsPerThread; ++j) {
      first[2 * j] = u.fhist[fine_of(pairs[j].x, f)];
      first[2 * j + 1] = u.fhist[fine_of(pairs[j].y, f)];
    }
#pragma unroll
    for (int j = 0; j < kPairsPerThread; ++j) {
      if ((survive >> (2 * j)) & 1u) a.cand[s + first[2 * j] + (rank[j] & 0xffffu)] = pairs[j].x;
      if ((survive >> (2 * j)) & 2u) a.cand[s + first[2 * j + 1] + (rank[j] >> 16)] = pairs[j].y;
    }
    if (threadIdx.x == 0) {
      unique_block += u.unique;
      cand_block += c;
    }
    __syncthreads();  // LDS reuse by the next bucket
  }
  if (threadIdx.x == 0 && unique_block) atomicAdd(a.unique_total, unique_block);
  if (threadIdx.x == 0 && cand_block) atomicAdd(a.cand_total, cand_block);
}

// grid = 2^K blocks: survivors of sub-bucket f (at the start of its slab)
// -> bank rows.  A layer with an error bit (offsets not trustworthy) or with
// more rows than the bank holds is left to the host.
__global__ __launch_bounds__(kRouteThreads) void compact_kernel(const uint64_t* __restrict__ cand, uint32_t nsub,
                                                                uint32_t fb, const uint32_t* __restrict__ off,
                                                                const uint32_t* __restrict__ error,
                                                                uint64_t bank_rows, uint64_t* __restrict__ bank) {
  if (*error != 0 || off[static_cast<uint64_t>(nsub) << fb] > bank_rows) return;
  const uint32_t f = blockIdx.x;
  const uint64_t s = static_cast<uint64_t>(f) * kSlabStride;
  const uint32_t sh = fb;
  const uint64_t lo = off[static_cast<uint64_t>(f) << sh], hi = off[static_cast<uint64_t>(f + 1) << sh];
  for (uint64_t i = threadIdx.x; i < hi - lo; i += kRouteThreads) bank[lo + i] = cand[s + i];
}

// Start layer: bank[0] = mix(start); offsets over 2^Ko fine buckets.
__global__ void sorted_seed_kernel(uint64_t start_key, SortedLayout L, uint64_t* bank, uint32_t* off, uint32_t Ko) {
  const uint64_t h = keymix(start_key, L.km) << L.tagbits;
  const uint32_t b = range_index(h, 1u << Ko);
  const uint64_t n = (1ull << Ko) + 1;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    off[i] = i <= b ? 0u : 1u;
  if (blockIdx.x == 0 && threadIdx.x == 0) bank[0] = h;
}

// Sets kErrFutureCap when the layer total off[n] exceeds `rows` (the
// compaction then writes nothing).
__global__ void capacity_check_kernel(const uint32_t* off, uint64_t n, uint64_t rows, uint32_t* error) {
  if (off[n] > rows) atomicOr(error, kErrFutureCap);
}

// ------------------------------------------------------------ launchers

hipError_t sorted_capacity_check(const uint32_t* off, uint64_t n, uint64_t rows, uint32_t* error,
                                 hipStream_t stream) {
  hipLaunchKernelGGL(capacity_check_kernel, dim3(1), dim3(1), 0, stream, off, n, rows, error);
  return hipGetLastError();
}

hipError_t sorted_seed(uint64_t start_key, const SortedLayout& L, uint64_t* bank, uint32_t* off, uint32_t Ko,
                       hipStream_t stream) {
  const uint64_t n = (1ull << Ko) + 1;
  const uint32_t blocks = static_cast<uint32_t>(std::min<uint64_t>((n + 255) / 256, 4096));
  hipLaunchKernelGGL(sorted_seed_kernel, dim3(blocks), dim3(256), 0, stream, start_key, L, bank, off, Ko);
  return hipGetLastError();
}

// Dump-area index for 32-bit destination indexing, or kWideOut when the
// output (entries, plus kTile dump entries) does not fit 32-bit indices.
static uint32_t dump_index(uint64_t out_entries) {
  return out_entries + kTile < kWideOut ? static_cast<uint32_t>(out_entries) : kWideOut;
}

template <int G, bool MASKED>
static hipError_t expand_route_g(const uint64_t* bank, uint64_t F, const PackedGenerators& gens, const SortedLayout& L,
                                 uint32_t nsub, uint32_t K2, uint32_t NB, uint32_t P, uint32_t cap, uint32_t stride,
                                 uint32_t err_bit, uint32_t* cursors, uint32_t* error, uint64_t* out,
                                 uint64_t out_entries, hipStream_t stream) {
  hipLaunchKernelGGL((expand_scatter_kernel<G, MASKED>), dim3(NB), dim3(kRouteThreads), 0, stream, bank, F, gens, L,
                     nsub, K2, NB, P, cap, stride, err_bit, cursors, error, out, dump_index(out_entries));
  return hipGetLastError();
}

hipError_t sorted_expand_route(const uint64_t* bank, uint64_t F, const PackedGenerators& gens, const SortedLayout& L,
                               bool masked, uint32_t nsub, uint32_t K2, uint32_t NB, uint32_t P, uint32_t cap,
                               uint32_t stride,
                               uint32_t err_bit, uint32_t* cursors, uint32_t* error, uint64_t* out,
                               uint64_t out_entries, hipStream_t stream) {
  if (nsub == 0 || K2 > kMaxPassBits || (nsub & ((1u << K2) - 1u)) || (nsub >> K2) > (1u << kMaxPassBits) ||
      NB == 0 || P == 0 || cap == 0 || stride < cap)
    return hipErrorInvalidValue;
  switch (gens.count) {
#define MGBFS_G(k)                                                                                               \
  case k:                                                                                                        \
    return masked                                                                                                \
               ? expand_route_g<k, true>(bank, F, gens, L, nsub, K2, NB, P, cap, stride, err_bit, cursors, error, out, \
                                         out_entries, stream)                                               \
               : expand_route_g<k, false>(bank, F, gens, L, nsub, K2, NB, P, cap, stride, err_bit, cursors, error, out, \
                                          out_entries, stream);
    MGBFS_G(1) MGBFS_G(2) MGBFS_G(3) MGBFS_G(4) MGBFS_G(5) MGBFS_G(6) MGBFS_G(7) MGBFS_G(8)
    MGBFS_G(9) MGBFS_G(10) MGBFS_G(11) MGBFS_G(12) MGBFS_G(13) MGBFS_G(14) MGBFS_G(15) MGBFS_G(16)
#undef MGBFS_G
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t sorted_partition(const uint64_t* in, const uint32_t* lcount, uint32_t lcap, uint32_t P, uint32_t nsub,
                            uint32_t K2, uint32_t* scount, uint32_t* error, uint64_t* out, uint64_t out_entries,
                            hipStream_t stream) {
  if (K2 == 0 || K2 > kMaxPassBits || (nsub & ((1u << K2) - 1u)) || nsub == 0 || nsub > (1u << kMaxBucketBits) ||
      (nsub >> K2) > (1u << kMaxPassBits) || P == 0 || lcap == 0)
    return hipErrorInvalidValue;
  const uint32_t blocks = (nsub >> K2) * P;
  hipLaunchKernelGGL(partition_scatter_kernel, dim3(blocks), dim3(kRouteThreads), 0, stream, in, lcount, lcap, P, nsub,
                     K2, scount, error, out, dump_index(out_entries));
  return hipGetLastError();
}

hipError_t sorted_unique_diff(uint64_t* cand, const uint32_t* scount, uint32_t nsub, const uint64_t* cur,
                              const uint32_t* cur_off, uint32_t nbo_c, const uint64_t* prev, const uint32_t* prev_off,
                              uint32_t nbo_p, uint32_t fb, uint32_t* fine_counts, uint32_t* error,
                              unsigned long long* unique_total, unsigned long long* cand_total, uint32_t tagbits,
                              bool merge, uint32_t grid_limit, hipStream_t stream) {
  if (nsub == 0 || !cur || !cur_off || fb > kMaxFineBits || (static_cast<uint64_t>(nsub) << fb) > (1ull << kMaxBucketBits) ||
      (cur && (nbo_c == 0 || nbo_c > (1u << kMaxBucketBits))) ||
      (prev && (nbo_p == 0 || nbo_p > (1u << kMaxBucketBits))) || grid_limit == 0)
    return hipErrorInvalidValue;
  if (tagbits < 4 || tagbits > 60) return hipErrorInvalidValue;
  const uint64_t tagmask = (1ull << tagbits) - 1;
  // Candidate tags use at most bits 0 .. tagbits-3 (back-edge masks need
  // G <= tagbits - 2), so the mark bit tagbits-1 is free and a marked entry
  // never equals the all-ones empty marker.
  UniqueArgs a{cand, scount, cur, cur_off, prev, prev_off, fine_counts, error, unique_total, cand_total, tagmask,
               1ull << (tagbits - 1), merge ? 1u : 0u, fb, nbo_c, nbo_p, tagbits, nsub, range_inverse(nsub)};
  const uint32_t grid = std::min<uint32_t>(nsub, grid_limit);
  if (prev)
    hipLaunchKernelGGL(unique_diff_kernel<true>, dim3(grid), dim3(kUniqueThreads), 0, stream, a);
  else
    hipLaunchKernelGGL(unique_diff_kernel<false>, dim3(grid), dim3(kUniqueThreads), 0, stream, a);
  return hipGetLastError();
}

hipError_t sorted_compact(const uint64_t* cand, uint32_t nsub, uint32_t fb, const uint32_t* off, const uint32_t* error,
                          uint64_t bank_rows, uint64_t* bank, hipStream_t stream) {
  hipLaunchKernelGGL(compact_kernel, dim3(nsub), dim3(kRouteThreads), 0, stream, cand, nsub, fb, off, error,
                     bank_rows, bank);
  return hipGetLastError();
}

}  // namespace mgbfs
