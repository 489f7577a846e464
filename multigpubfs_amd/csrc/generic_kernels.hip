// Generic exact path: states of any width stored as element-major SoA planes
// (int64 or uint8), permutation or int64 matrix actions, and an exact
// all-visited open-addressing table.  Drop-in for the reference's C ABI
// (cuda/generic_visited.h:3-11, cuda/generic_graph.h:3-6) with the same
// argument meaning and error bits.
//
// Work layout: child c = g * count + parent (generator-major), so the 64
// lanes of a wavefront handle 64 consecutive parents under one generator and
// every element read of a permutation gather is a coalesced SoA access.
//
// Table slots (64 bit): EMPTY = ~0; published = [0 | fp:31 | row:32];
// pending = [1 | fp:31 | child:32].  A claimant CASes EMPTY -> pending (its
// origin is immutable for the launch, so peers compare by regenerating the
// child instead of waiting), allocates a row, writes the state, then
// release-stores the published token (the claim-then-publish protocol of
// cuda/generic_visited.cu:27-45).  fp is a fingerprint of the state hash;
// equality of full states always decides membership.
#include "common.hpp"
#include "kernels.hpp"
#include "../../include/mgbfs_hip.h"

namespace mgbfs {

namespace {

constexpr uint64_t kSlotEmpty = ~0ull;
constexpr uint64_t kPendingBit = 1ull << 63;
constexpr uint64_t kFpMask = 0x7fffffff00000000ull;
constexpr int kGenericThreads = 256;

__device__ __forceinline__ uint64_t fold64(uint64_t h, uint64_t v) {
  h ^= v + 0x9e3779b97f4a7c15ull + (h << 6) + (h >> 2);
  h *= 0xd6e8feb86659fd93ull;
  return h ^ (h >> 32);
}

__device__ __forceinline__ uint64_t finish(uint64_t h, uint32_t bits) {
  h = mix64(h);
  return bits >= 64 ? h : (bits ? (h & ((1ull << bits) - 1)) : 0ull);
}

template <typename S>
struct SoaAction {
  uint32_t kind, elements, rows, cols, generators, count;
  uint64_t stride;
  const S* parents;
  const uint32_t* perms;
  const int64_t* mats;
  const uint32_t* moduli;

  __device__ __forceinline__ int64_t parent_value(uint32_t parent, uint32_t e) const {
    return static_cast<int64_t>(parents[static_cast<uint64_t>(e) * stride + parent]);
  }
  // Element e of the child (generator g applied to parent p).
  __device__ __forceinline__ int64_t value(uint32_t g, uint32_t p, uint32_t e) const {
    if (kind == 0) return parent_value(p, perms[static_cast<uint64_t>(g) * elements + e]);
    const uint32_t r = e / cols, c = e % cols;
    const int64_t* m = mats + static_cast<uint64_t>(g) * rows * rows + static_cast<uint64_t>(r) * rows;
    uint64_t acc = 0;  // two's-complement wrap, as int64 arithmetic in the reference
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
  __device__ __forceinline__ uint64_t hash(uint32_t g, uint32_t p, uint64_t seed, uint32_t bits) const {
    uint64_t h = seed;
    for (uint32_t e = 0; e < elements; ++e) h = fold64(h, static_cast<uint64_t>(value(g, p, e)) ^ (uint64_t(e) << 56));
    return finish(h, bits);
  }
};

template <typename S>
__global__ __launch_bounds__(kGenericThreads) void generic_seed_kernel(uint32_t elements, const S* states,
                                                                        uint32_t stride, uint32_t count,
                                                                        uint64_t* slots, uint64_t capacity,
                                                                        uint64_t seed, uint32_t bits,
                                                                        uint32_t* error) {
  for (uint32_t row = blockIdx.x * blockDim.x + threadIdx.x; row < count; row += blockDim.x * gridDim.x) {
    uint64_t h = seed;
    for (uint32_t e = 0; e < elements; ++e)
      h = fold64(h, static_cast<uint64_t>(static_cast<int64_t>(states[static_cast<uint64_t>(e) * stride + row])) ^
                        (uint64_t(e) << 56));
    h = finish(h, bits);
    const uint64_t token = (h & kFpMask) | row;
    uint64_t slot = h & (capacity - 1);
    bool placed = false;
    for (uint64_t probe = 0; probe < capacity; ++probe, slot = (slot + 1) & (capacity - 1)) {
      uint64_t expected = kSlotEmpty;
      if (__hip_atomic_compare_exchange_strong(slots + slot, &expected, token, __ATOMIC_ACQ_REL, __ATOMIC_ACQUIRE,
                                               __HIP_MEMORY_SCOPE_AGENT)) {
        placed = true;
        break;
      }
    }
    if (!placed) atomicOr(error, kErrTableFull);
  }
}

template <typename S>
__global__ __launch_bounds__(kGenericThreads) void generic_expand_kernel(
    SoaAction<S> a, uint64_t* slots, uint64_t slot_capacity, S* visited, uint32_t visited_capacity,
    uint32_t* visited_count, uint32_t* future, uint32_t future_capacity, uint32_t* future_count, uint64_t seed,
    uint32_t bits, uint32_t* error) {
  const uint32_t children = a.count * a.generators;
  const uint32_t origin_limit = children;
  for (uint32_t child = blockIdx.x * blockDim.x + threadIdx.x; child < children; child += blockDim.x * gridDim.x) {
    if (__hip_atomic_load(error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
    const uint32_t g = child / a.count, p = child % a.count;
    const uint64_t h = a.hash(g, p, seed, bits);
    const uint64_t fp = h & kFpMask;
    const uint64_t pending = kPendingBit | fp | child;
    uint64_t slot = h & (slot_capacity - 1);
    bool settled = false;
    for (uint64_t probe = 0; probe < slot_capacity && !settled; ++probe, slot = (slot + 1) & (slot_capacity - 1)) {
      uint64_t seen = kSlotEmpty;
      if (__hip_atomic_compare_exchange_strong(slots + slot, &seen, pending, __ATOMIC_ACQ_REL, __ATOMIC_ACQUIRE,
                                               __HIP_MEMORY_SCOPE_AGENT)) {
        // Claimed: allocate a row, write the state, publish.
        const uint32_t row = atomicAdd(visited_count, 1u);
        if (row >= visited_capacity) {
          atomicOr(error, kErrVisitedCap);
          return;
        }
        for (uint32_t e = 0; e < a.elements; ++e)
          visited[static_cast<uint64_t>(e) * visited_capacity + row] = static_cast<S>(a.value(g, p, e));
        __hip_atomic_store(slots + slot, fp | row, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t pos = atomicAdd(future_count, 1u);
        if (pos >= future_capacity) {
          atomicOr(error, kErrFutureCap);
          return;
        }
        future[pos] = row;
        settled = true;
        break;
      }
      if ((seen & kFpMask) != fp) continue;
      const bool transient = (seen & kPendingBit) != 0;
      const uint32_t ref = static_cast<uint32_t>(seen);
      if ((transient && ref >= origin_limit) || (!transient && ref >= visited_capacity)) {
        atomicOr(error, kErrBadRef);
        return;
      }
      bool equal = true;
      const uint32_t og = ref / a.count, op = ref % a.count;
      for (uint32_t e = 0; e < a.elements && equal; ++e) {
        const int64_t other = transient ? a.value(og, op, e)
                                        : static_cast<int64_t>(visited[static_cast<uint64_t>(e) * visited_capacity + ref]);
        equal = other == a.value(g, p, e);
      }
      if (equal) settled = true;
    }
    if (!settled) atomicOr(error, kErrTableFull);
  }
}

template <typename S>
__global__ void generic_gather_kernel(uint32_t elements, const S* source, uint32_t source_stride,
                                      const uint32_t* indices, uint32_t count, S* output, uint32_t output_stride) {
  const uint64_t total = static_cast<uint64_t>(elements) * count;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
       i += static_cast<uint64_t>(blockDim.x) * gridDim.x) {
    const uint32_t e = static_cast<uint32_t>(i / count), row = static_cast<uint32_t>(i % count);
    output[static_cast<uint64_t>(e) * output_stride + row] = source[static_cast<uint64_t>(e) * source_stride + indices[row]];
  }
}

// selected[j] = parent * generators + generator (the reference's encoding).
template <typename S>
__global__ void generic_generate_kernel(SoaAction<S> a, const uint64_t* selected, uint32_t output_count, S* output,
                                        uint32_t output_stride, uint32_t* error) {
  const uint64_t total = static_cast<uint64_t>(a.elements) * output_count;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
       i += static_cast<uint64_t>(blockDim.x) * gridDim.x) {
    const uint32_t e = static_cast<uint32_t>(i / output_count), j = static_cast<uint32_t>(i % output_count);
    const uint64_t c = selected[j];
    if (c >= static_cast<uint64_t>(a.count) * a.generators) {
      atomicOr(error, kErrBadRef);
      continue;
    }
    const uint32_t p = static_cast<uint32_t>(c / a.generators), g = static_cast<uint32_t>(c % a.generators);
    output[static_cast<uint64_t>(e) * output_stride + j] = static_cast<S>(a.value(g, p, e));
  }
}

uint32_t grid_for(uint64_t work) {
  const uint64_t n = (work + kGenericThreads - 1) / kGenericThreads;
  return static_cast<uint32_t>(n > 65536 ? 65536 : (n ? n : 1));
}

bool pow2(uint64_t x) { return x && !(x & (x - 1)); }

template <typename S>
int seed_impl(uint32_t elements, const S* states, uint32_t stride, uint32_t count, uint64_t* slots, uint64_t capacity,
              uint64_t seed, uint32_t bits, uint32_t* error, void* stream) {
  if (!elements || count > stride || !pow2(capacity) || bits > 64 || !error || !slots || (!states && count) ||
      count >= 0x80000000u)
    return hipErrorInvalidValue;
  if (!count) return 0;
  hipLaunchKernelGGL(generic_seed_kernel<S>, dim3(grid_for(count)), dim3(kGenericThreads), 0,
                     static_cast<hipStream_t>(stream), elements, states, stride, count, slots, capacity, seed, bits,
                     error);
  return hipGetLastError();
}

template <typename S>
int expand_impl(uint32_t kind, uint32_t elements, uint32_t n, uint32_t m, uint32_t generators, const S* parents,
                uint32_t count, uint32_t stride, const uint32_t* perms, const int64_t* mats, const uint32_t* moduli,
                uint64_t* slots, uint64_t slot_capacity, S* visited, uint32_t visited_capacity,
                uint32_t* visited_count, uint32_t* future, uint32_t future_capacity, uint32_t* future_count,
                uint64_t seed, uint32_t bits, uint32_t* error, void* stream) {
  if (kind > 1 || !elements || !n || !m || !generators || count > stride ||
      static_cast<uint64_t>(count) * generators >= 0x7fffffffull || !pow2(slot_capacity) || !visited_capacity ||
      visited_capacity >= 0x80000000u || bits > 64 || !visited_count || !future_count || !error || !slots ||
      !visited || !future || !parents || (kind == 0 && (!perms || elements != n || m != 1)) ||
      (kind == 1 && (!mats || !moduli || static_cast<uint64_t>(n) * m != elements)) ||
      (kind == 1 && sizeof(S) == 1))
    return hipErrorInvalidValue;
  if (!count) return 0;
  SoaAction<S> a{kind, elements, n, m, generators, count, stride, parents, perms, mats, moduli};
  hipLaunchKernelGGL(generic_expand_kernel<S>, dim3(grid_for(static_cast<uint64_t>(count) * generators)),
                     dim3(kGenericThreads), 0, static_cast<hipStream_t>(stream), a, slots, slot_capacity, visited,
                     visited_capacity, visited_count, future, future_capacity, future_count, seed, bits, error);
  return hipGetLastError();
}

template <typename S>
int gather_impl(uint32_t elements, const S* source, uint32_t stride, const uint32_t* indices, uint32_t count,
                S* output, uint32_t output_stride, void* stream) {
  if (!elements || count > output_stride || (!source && count) || (!indices && count) || (!output && count))
    return hipErrorInvalidValue;
  if (!count) return 0;
  hipLaunchKernelGGL(generic_gather_kernel<S>, dim3(grid_for(static_cast<uint64_t>(elements) * count)),
                     dim3(kGenericThreads), 0, static_cast<hipStream_t>(stream), elements, source, stride, indices,
                     count, output, output_stride);
  return hipGetLastError();
}

template <typename S>
int generate_impl(uint32_t kind, uint32_t elements, uint32_t n, uint32_t m, uint32_t generators, const S* parents,
                  uint32_t count, uint32_t stride, const uint32_t* perms, const int64_t* mats, const uint32_t* moduli,
                  const uint64_t* selected, uint32_t output_count, S* output, uint32_t output_stride, uint32_t* error,
                  void* stream) {
  if (kind > 1 || !elements || !generators || count > stride || output_count > output_stride || !error ||
      (count && !parents) || (output_count && (!selected || !output)) ||
      (kind == 0 && (!perms || elements != n || m != 1)) ||
      (kind == 1 && (!mats || !moduli || static_cast<uint64_t>(n) * m != elements)) || (kind == 1 && sizeof(S) == 1))
    return hipErrorInvalidValue;
  if (!output_count) return 0;
  SoaAction<S> a{kind, elements, n, m, generators, count, stride, parents, perms, mats, moduli};
  hipLaunchKernelGGL(generic_generate_kernel<S>, dim3(grid_for(static_cast<uint64_t>(elements) * output_count)),
                     dim3(kGenericThreads), 0, static_cast<hipStream_t>(stream), a, selected, output_count, output,
                     output_stride, error);
  return hipGetLastError();
}

}  // namespace
}  // namespace mgbfs

using namespace mgbfs;

extern "C" int mgbfs_generic_seed_i64_wide(uint32_t elements, const int64_t* states, uint32_t state_stride,
                                           uint32_t state_count, uint64_t* slots, uint64_t slot_capacity,
                                           uint64_t seed, uint32_t hash_bits, uint32_t* error, void* stream) {
  return seed_impl<int64_t>(elements, states, state_stride, state_count, slots, slot_capacity, seed, hash_bits, error,
                            stream);
}

extern "C" int mgbfs_generic_seed_u8_wide(uint32_t elements, const uint8_t* states, uint32_t stride, uint32_t count,
                                          uint64_t* slots, uint64_t capacity, uint64_t seed, uint32_t bits,
                                          uint32_t* error, void* stream) {
  return seed_impl<uint8_t>(elements, states, stride, count, slots, capacity, seed, bits, error, stream);
}

extern "C" int mgbfs_generic_expand_i64_wide(uint32_t kind, uint32_t elements, uint32_t rows, uint32_t cols,
                                             uint32_t generators, const int64_t* parents, uint32_t parent_count,
                                             uint32_t parent_stride, const uint32_t* permutation_tables,
                                             const int64_t* matrix_tables, const uint32_t* moduli, uint64_t* slots,
                                             uint64_t slot_capacity, int64_t* visited, uint32_t visited_capacity,
                                             uint32_t* visited_count, uint32_t* future, uint32_t future_capacity,
                                             uint32_t* future_count, uint64_t seed, uint32_t hash_bits,
                                             uint32_t* error, void* stream) {
  return expand_impl<int64_t>(kind, elements, rows, cols, generators, parents, parent_count, parent_stride,
                              permutation_tables, matrix_tables, moduli, slots, slot_capacity, visited,
                              visited_capacity, visited_count, future, future_capacity, future_count, seed, hash_bits,
                              error, stream);
}

extern "C" int mgbfs_generic_expand_u8_wide(uint32_t kind, uint32_t elements, uint32_t n, uint32_t m,
                                            uint32_t generators, const uint8_t* parents, uint32_t count,
                                            uint32_t stride, const uint32_t* permutations, const int64_t* matrices,
                                            const uint32_t* moduli, uint64_t* slots, uint64_t slot_capacity,
                                            uint8_t* visited, uint32_t visited_capacity, uint32_t* visited_count,
                                            uint32_t* future, uint32_t future_capacity, uint32_t* future_count,
                                            uint64_t seed, uint32_t bits, uint32_t* error, void* stream) {
  return expand_impl<uint8_t>(kind, elements, n, m, generators, parents, count, stride, permutations, matrices, moduli,
                              slots, slot_capacity, visited, visited_capacity, visited_count, future, future_capacity,
                              future_count, seed, bits, error, stream);
}

extern "C" int mgbfs_generic_gather_i64(uint32_t elements, const int64_t* source, uint32_t source_stride,
                                        const uint32_t* indices, uint32_t count, int64_t* output,
                                        uint32_t output_stride, void* stream) {
  return gather_impl<int64_t>(elements, source, source_stride, indices, count, output, output_stride, stream);
}

extern "C" int mgbfs_generic_gather_u8(uint32_t elements, const uint8_t* source, uint32_t stride,
                                       const uint32_t* indices, uint32_t count, uint8_t* output,
                                       uint32_t output_stride, void* stream) {
  return gather_impl<uint8_t>(elements, source, stride, indices, count, output, output_stride, stream);
}

extern "C" int mgbfs_generic_generate_i64(uint32_t kind, uint32_t elements, uint32_t rows, uint32_t cols,
                                          uint32_t generators, const int64_t* parents, uint32_t parent_count,
                                          uint32_t parent_stride, const uint32_t* permutation_tables,
                                          const int64_t* matrix_tables, const uint32_t* moduli,
                                          const uint64_t* selected_children, uint32_t output_count, int64_t* output,
                                          uint32_t output_stride, uint32_t* device_error, void* stream) {
  return generate_impl<int64_t>(kind, elements, rows, cols, generators, parents, parent_count, parent_stride,
                                permutation_tables, matrix_tables, moduli, selected_children, output_count, output,
                                output_stride, device_error, stream);
}

extern "C" int mgbfs_generic_generate_u8(uint32_t kind, uint32_t elements, uint32_t rows, uint32_t cols,
                                         uint32_t generators, const uint8_t* parents, uint32_t parent_count,
                                         uint32_t parent_stride, const uint32_t* permutation_tables,
                                         const int64_t* matrix_tables, const uint32_t* moduli,
                                         const uint64_t* selected_children, uint32_t output_count, uint8_t* output,
                                         uint32_t output_stride, uint32_t* device_error, void* raw_stream) {
  return generate_impl<uint8_t>(kind, elements, rows, cols, generators, parents, parent_count, parent_stride,
                                permutation_tables, matrix_tables, moduli, selected_children, output_count, output,
                                output_stride, device_error, raw_stream);
}
