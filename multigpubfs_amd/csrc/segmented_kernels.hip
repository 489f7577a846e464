// Segmented open-addressing history for states of any width.
//
// Replaces the reference's segmented history (cuda/segmented_history.cu:
// 63-157 accept, :173-279 grouped accept, :287-360 retire/rehash;
// crates/mgbfs-runtime/src/segmented_history.rs:95-177 layout) for one GPU.
// States live in element-major SoA arenas, one per bank (a BFS layer, or the
// whole history for directed graphs), addressed by 32-bit local rows.  Each
// bank owns a list of segments: power-of-two tables of u64 slots filled to
// at most half.  Every segment but the newest of the bank being written is
// sealed (read-only); a child probes every sealed segment of the banks it
// must be checked against, then claims a slot in the active segment.
//
// Slot tokens: EMPTY = ~0; published = [0 | fp:29 | bank:2 | row:32];
// pending = [1 | fp:29 | 0:2 | child:32] while its claimant writes the
// state (claim-then-publish, as cuda/generic_visited.cu:27-45: peers compare
// a pending token by regenerating the child from its immutable origin).
// fp is 29 bits of the state hash; full-state equality decides.
//
// A segment that reaches its fill limit refuses further claims (error bit
// kErrSegmentFull).  The host then seals it, adds a larger segment and
// replays the batch: children already claimed are found in the sealed
// segment, so a replay never adds a state twice.  Retiring a bank drops its
// segments, so no tombstones or rehash passes are needed.
#include "common.hpp"
#include "kernels.hpp"
#include "segmented.hpp"

namespace mgbfs {

namespace {

constexpr uint64_t kSegEmpty = ~0ull;
constexpr uint64_t kSegPending = 1ull << 63;
constexpr uint64_t kSegFpMask = 0x7ffffffc00000000ull;  // 29 hash bits above bank and row
constexpr int kSegThreads = 256;

__device__ __forceinline__ uint64_t seg_fold(uint64_t h, uint64_t v) {
  h ^= v + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
  h *= 0xd6e8feb86659fd93ull;
  return h ^ (h >> 32);
}

template <typename S>
struct SegAction {
  uint32_t kind, elements, rows, cols, generators, count;
  uint64_t stride;    // row stride of the parent arena
  uint64_t first;     // first parent row
  const S* parents;
  const uint32_t* perms;
  const int64_t* mats;
  const uint32_t* moduli;

  __device__ __forceinline__ int64_t parent_value(uint32_t p, uint32_t e) const {
    return static_cast<int64_t>(parents[static_cast<uint64_t>(e) * stride + first + p]);
  }
  // Element e of generator g applied to parent p: child[i] = parent[g[i]],
  // or the int64 matrix product wrapped mod 2^64, then rem_euclid by the
  // generator's modulus (0 = no reduction).
  __device__ __forceinline__ int64_t value(uint32_t g, uint32_t p, uint32_t e) const {
    if (kind == 0) return parent_value(p, perms[static_cast<uint64_t>(g) * elements + e]);
    const uint32_t r = e / cols, c = e % cols;
    const int64_t* m = mats + static_cast<uint64_t>(g) * rows * rows + static_cast<uint64_t>(r) * rows;
    uint64_t acc = 0;
    for (uint32_t k = 0; k < rows; ++k)
      acc += static_cast<uint64_t>(m[k]) * static_cast<uint64_t>(parent_value(p, k * cols + c));
    int64_t v = static_cast<int64_t>(acc);
    const uint32_t mod = moduli[g];
    if (mod) {
      v %= static_cast<int64_t>(mod);
      if (v < 0) v += mod;
    }
    return v;
  }
  __device__ __forceinline__ uint64_t hash(uint32_t g, uint32_t p, uint32_t bits) const {
    uint64_t h = 0x19f856a2ull;
    for (uint32_t e = 0; e < elements; ++e) h = seg_fold(h, static_cast<uint64_t>(value(g, p, e)) ^ (uint64_t(e) << 56));
    h = mix64(h);
    return bits >= 64 ? h : (bits ? (h & ((1ull << bits) - 1)) : 0ull);
  }
};

template <typename S>
__device__ __forceinline__ uint64_t row_hash(const S* arena, uint64_t stride, uint32_t elements, uint32_t row,
                                             uint32_t bits) {
  uint64_t h = 0x19f856a2ull;
  for (uint32_t e = 0; e < elements; ++e)
    h = seg_fold(h, static_cast<uint64_t>(static_cast<int64_t>(arena[static_cast<uint64_t>(e) * stride + row])) ^
                        (uint64_t(e) << 56));
  h = mix64(h);
  return bits >= 64 ? h : (bits ? (h & ((1ull << bits) - 1)) : 0ull);
}

// Probes one sealed segment for the child (g, p); true when found.
template <typename S>
__device__ bool probe_sealed(const SegAction<S>& a, const SegSegment& seg, const SegBanks<S>& banks, uint32_t g,
                             uint32_t p, uint64_t h, uint32_t* error) {
  const uint64_t fp = (h << 34) & kSegFpMask;
  uint64_t slot = h & seg.mask;
  for (uint64_t probe = 0; probe <= seg.mask; ++probe, slot = (slot + 1) & seg.mask) {
    const uint64_t t = seg.slots[slot];
    if (t == kSegEmpty) return false;
    if ((t & kSegFpMask) != fp) continue;
    const uint32_t bank = static_cast<uint32_t>(t >> 32) & 3u, row = static_cast<uint32_t>(t);
    if ((t & kSegPending) || bank >= kSegBanks || row >= banks.rows[bank]) {
      atomicOr(error, kErrBadRef);
      return true;
    }
    const S* arena = banks.arena[bank];
    bool equal = true;
    for (uint32_t e = 0; e < a.elements && equal; ++e)
      equal = static_cast<int64_t>(arena[static_cast<uint64_t>(e) * banks.stride[bank] + row]) == a.value(g, p, e);
    if (equal) return true;
  }
  return false;
}

// One batch of parents: every child is looked up in the sealed segments of
// the probed banks, then claimed in the active segment of the bank being
// written; a claimed child gets the next row of that bank.
template <typename S>
__global__ __launch_bounds__(kSegThreads) void segmented_expand_kernel(SegAction<S> a, SegBanks<S> banks,
                                                                       SegProbeSet probe, SegSegment active,
                                                                       uint32_t target, uint32_t bits,
                                                                       uint32_t* control, uint32_t* error) {
  uint32_t* rows_used = control;      // rows claimed in the target bank
  uint32_t* fill = control + 1;       // slots claimed in the active segment
  const uint32_t children = a.count * a.generators;
  for (uint32_t child = blockIdx.x * blockDim.x + threadIdx.x; child < children; child += blockDim.x * gridDim.x) {
    if (__hip_atomic_load(error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    // Generator-major: a wave's lanes share a generator over consecutive parents.
    const uint32_t g = child / a.count, p = child % a.count;
    const uint64_t h = a.hash(g, p, bits);
    bool found = false;
    for (uint32_t s = 0; s < probe.count && !found; ++s) found = probe_sealed(a, probe.seg[s], banks, g, p, h, error);
    if (found) continue;
    const uint64_t fp = (h << 34) & kSegFpMask;
    const uint64_t pending = kSegPending | fp | child;
    uint64_t slot = h & active.mask;
    bool settled = false;
    for (uint64_t step = 0; step <= active.mask && !settled; ++step, slot = (slot + 1) & active.mask) {
      uint64_t seen = __hip_atomic_load(active.slots + slot, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
      if (seen == kSegEmpty) {
        // Fill limit first: a refused claim leaves the child for the replay.
        if (atomicAdd(fill, 1u) >= active.limit) {
          atomicOr(error, kErrSegmentFull);
          return;
        }
        if (__hip_atomic_compare_exchange_strong(active.slots + slot, &seen, pending, __ATOMIC_ACQ_REL,
                                                 __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT)) {
          const uint32_t row = atomicAdd(rows_used, 1u);
          if (row >= banks.capacity[target]) {
            atomicOr(error, kErrVisitedCap);
            return;
          }
          S* arena = banks.arena[target];
          for (uint32_t e = 0; e < a.elements; ++e)
            arena[static_cast<uint64_t>(e) * banks.stride[target] + row] = static_cast<S>(a.value(g, p, e));
          __hip_atomic_store(active.slots + slot, fp | (static_cast<uint64_t>(target) << 32) | row, __ATOMIC_RELEASE,
                             __HIP_MEMORY_SCOPE_AGENT);
          settled = true;
          break;
        }
        // Lost the race for this slot: `seen` now holds the winner's token.
      }
      if ((seen & kSegFpMask) != fp) continue;
      const bool transient = (seen & kSegPending) != 0;
      const uint32_t ref = static_cast<uint32_t>(seen);
      bool equal = true;
      if (transient) {
        if (ref >= children) {
          atomicOr(error, kErrBadRef);
          return;
        }
        const uint32_t og = ref / a.count, op = ref % a.count;
        for (uint32_t e = 0; e < a.elements && equal; ++e) equal = a.value(og, op, e) == a.value(g, p, e);
      } else {
        const uint32_t bank = static_cast<uint32_t>(seen >> 32) & 3u;
        if (bank != target || ref >= banks.capacity[target]) {
          atomicOr(error, kErrBadRef);
          return;
        }
        const S* arena = banks.arena[target];
        for (uint32_t e = 0; e < a.elements && equal; ++e)
          equal = static_cast<int64_t>(arena[static_cast<uint64_t>(e) * banks.stride[target] + ref]) ==
                  a.value(g, p, e);
      }
      if (equal) settled = true;
    }
    if (!settled) atomicOr(error, kErrTableFull);
  }
}

// Inserts rows [first, first + count) of bank `bank` into a segment (the
// start state, or rows handed over from another backend).  Rows are
// distinct states, so no equality test is needed.
template <typename S>
__global__ __launch_bounds__(kSegThreads) void segmented_seed_kernel(SegBanks<S> banks, uint32_t bank, uint32_t first,
                                                                     uint32_t count, uint32_t elements, SegSegment seg,
                                                                     uint32_t bits, uint32_t* fill, uint32_t* error) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += blockDim.x * gridDim.x) {
    const uint32_t row = first + i;
    const uint64_t h = row_hash(banks.arena[bank], banks.stride[bank], elements, row, bits);
    const uint64_t token = ((h << 34) & kSegFpMask) | (static_cast<uint64_t>(bank) << 32) | row;
    if (atomicAdd(fill, 1u) >= seg.limit) {
      atomicOr(error, kErrSegmentFull);
      return;
    }
    uint64_t slot = h & seg.mask;
    bool placed = false;
    for (uint64_t step = 0; step <= seg.mask && !placed; ++step, slot = (slot + 1) & seg.mask) {
      uint64_t expected = kSegEmpty;
      placed = __hip_atomic_compare_exchange_strong(seg.slots + slot, &expected, token, __ATOMIC_ACQ_REL,
                                                    __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (!placed) atomicOr(error, kErrTableFull);
  }
}

// Sorted-backend bank values (mix(key) << tagbits | tag) -> nibble rows of
// a u8 SoA arena.
__global__ void segmented_adopt_packed_kernel(const uint64_t* values, uint64_t count, KeyMix km, uint32_t tagbits,
                                              uint32_t degree, uint8_t* arena, uint64_t stride) {
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count;
       i += static_cast<uint64_t>(blockDim.x) * gridDim.x) {
    const uint64_t key = keyunmix(values[i] >> tagbits, km);
    for (uint32_t e = 0; e < degree; ++e) arena[static_cast<uint64_t>(e) * stride + i] = static_cast<uint8_t>((key >> (4 * e)) & 15u);
  }
}

uint32_t seg_grid(uint64_t work) {
  const uint64_t n = (work + kSegThreads - 1) / kSegThreads;
  return static_cast<uint32_t>(n > 65536 ? 65536 : (n ? n : 1));
}

}  // namespace

template <typename S>
hipError_t segmented_expand(const SegBatch<S>& b, hipStream_t stream) {
  if (b.count == 0) return hipSuccess;
  if (static_cast<uint64_t>(b.count) * b.generators >= 0x7fffffffull || b.probe.count > kSegMaxProbe ||
      b.target >= kSegBanks)
    return hipErrorInvalidValue;
  SegAction<S> a{b.kind, b.elements, b.rows, b.cols, b.generators, b.count, b.parent_stride, b.first_parent,
                 b.parents, b.perms, b.mats, b.moduli};
  hipLaunchKernelGGL(segmented_expand_kernel<S>, dim3(seg_grid(static_cast<uint64_t>(b.count) * b.generators)),
                     dim3(kSegThreads), 0, stream, a, b.banks, b.probe, b.active, b.target, b.hash_bits, b.control,
                     b.error);
  return hipGetLastError();
}

template <typename S>
hipError_t segmented_seed(const SegBanks<S>& banks, uint32_t bank, uint32_t first, uint32_t count, uint32_t elements,
                          const SegSegment& seg, uint32_t bits, uint32_t* fill, uint32_t* error, hipStream_t stream) {
  if (count == 0) return hipSuccess;
  hipLaunchKernelGGL(segmented_seed_kernel<S>, dim3(seg_grid(count)), dim3(kSegThreads), 0, stream, banks, bank, first,
                     count, elements, seg, bits, fill, error);
  return hipGetLastError();
}

hipError_t segmented_adopt_packed(const uint64_t* values, uint64_t count, const KeyMix& km, uint32_t tagbits,
                                  uint32_t degree, uint8_t* arena, uint64_t stride, hipStream_t stream) {
  if (count == 0) return hipSuccess;
  hipLaunchKernelGGL(segmented_adopt_packed_kernel, dim3(seg_grid(count)), dim3(kSegThreads), 0, stream, values, count,
                     km, tagbits, degree, arena, stride);
  return hipGetLastError();
}

template hipError_t segmented_expand<uint8_t>(const SegBatch<uint8_t>&, hipStream_t);
template hipError_t segmented_expand<int64_t>(const SegBatch<int64_t>&, hipStream_t);
template hipError_t segmented_seed<uint8_t>(const SegBanks<uint8_t>&, uint32_t, uint32_t, uint32_t, uint32_t,
                                            const SegSegment&, uint32_t, uint32_t*, uint32_t*, hipStream_t);
template hipError_t segmented_seed<int64_t>(const SegBanks<int64_t>&, uint32_t, uint32_t, uint32_t, uint32_t,
                                            const SegSegment&, uint32_t, uint32_t*, uint32_t*, hipStream_t);

}  // namespace mgbfs
