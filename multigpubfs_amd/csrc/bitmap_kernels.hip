// Permutation-bitmap backend: fused expand -> Lehmer rank -> bit claim ->
// wave/workgroup-aggregated append.  One pass over the frontier per layer.
//
// Replaces the reference's bitmap_route + bitmap_accept pair
// (cuda/permutation_bitmap.cu:27-48, :105-133) and the fused prototype
// expand_kernel (gpu/cayley_bfs.cu:84-124) with an MI355X-first kernel:
//  * frontier keys are read fully coalesced (tile-major: item*256 + tid),
//  * the generator table is a by-value kernel argument (scalar registers),
//  * a plain load of the bitmap word filters already-visited children before
//    any atomic (bits only go 0 -> 1, so a stale read can only under-report),
//  * new children are appended with one global atomic per workgroup.
#include "common.hpp"
#include "kernels.hpp"

namespace mgbfs {

constexpr int kBitmapThreads = 256;
constexpr int kChildrenPerThread = 24;

template <int G>
constexpr int bitmap_items() { return (kChildrenPerThread / G) > 0 ? (kChildrenPerThread / G) : 1; }

// One tile of kBitmapThreads x kItems parents (tile index `tile_index`).
template <int G>
__device__ __forceinline__ void bitmap_tile(
    uint64_t tile_index, const uint64_t* __restrict__ frontier, uint64_t count, const PackedGenerators& gens,
    uint32_t* __restrict__ bitmap, uint64_t bitmap_bits, uint64_t* __restrict__ next,
    unsigned long long* __restrict__ next_count, uint64_t count_base, uint64_t next_cap,
    uint32_t* __restrict__ error) {
  constexpr int kItems = bitmap_items<G>();
  __shared__ uint64_t parents[kItems][kBitmapThreads];
  __shared__ uint32_t wave_sums[kBitmapThreads / kWave];
  __shared__ unsigned long long wg_base;
  const uint32_t n = gens.degree;
  const uint64_t tile = tile_index * kBitmapThreads * kItems;
  uint32_t newmask = 0;
  uint32_t errbits = 0;

  // Phase 1: claim.  Children are not kept; the few new ones are recomputed
  // from the LDS copy of the parent when they are written.
  for (int it = 0; it < kItems; ++it) {
    const uint64_t p = tile + static_cast<uint64_t>(it) * kBitmapThreads + threadIdx.x;
    const bool valid = p < count;
    const uint64_t key = valid ? frontier[p] : 0;
    parents[it][threadIdx.x] = key;
    if (!valid) continue;
    const uint32_t skip = gens.tag_shift < 64 ? static_cast<uint32_t>(key >> gens.tag_shift) : 0u;
    // A parent's probes go out together: every child's rank, then every
    // bitmap word load, then the atomics for the clear bits.
    uint64_t word[G];
    uint32_t bit[G], seen[G];
    uint32_t live = 0;
#pragma unroll
    for (int g = 0; g < G; ++g) {
      word[g] = 0;
      bit[g] = 0;
      if (static_cast<uint32_t>(g) + 1 == skip) continue;
      const uint64_t c = apply_groups(key, gens, g);
      bool bad;
      const uint64_t r = lehmer_rank_packed(c, n, &bad);
      if (bad || r >= bitmap_bits) {
        errbits |= kErrDomain;
        continue;
      }
      word[g] = r >> 5;
      bit[g] = 1u << (r & 31u);
      live |= 1u << g;
    }
#pragma unroll
    for (int g = 0; g < G; ++g) seen[g] = (live >> g) & 1u ? bitmap[word[g]] : 0u;
#pragma unroll
    for (int g = 0; g < G; ++g)
      if (((live >> g) & 1u) && !(seen[g] & bit[g]) &&
          !(__hip_atomic_fetch_or(bitmap + word[g], bit[g], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & bit[g]))
        newmask |= 1u << (it * G + g);
  }

  // Phase 2: workgroup exclusive scan of per-thread new-child counts and one
  // global reservation per workgroup.
  const uint32_t mine = __popc(newmask);
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint32_t wave = threadIdx.x / kWave;
  uint32_t incl = mine;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const uint32_t v = __shfl_up(incl, off, kWave);
    if (lane >= static_cast<uint32_t>(off)) incl += v;
  }
  if (lane == kWave - 1) wave_sums[wave] = incl;
  __syncthreads();
  uint32_t wave_off = 0, total = 0;
#pragma unroll
  for (int w = 0; w < kBitmapThreads / kWave; ++w) {
    const uint32_t s = wave_sums[w];
    if (w < static_cast<int>(wave)) wave_off += s;
    total += s;
  }
  if (threadIdx.x == 0)
    wg_base = total ? atomicAdd(next_count, static_cast<unsigned long long>(total)) - count_base : 0ull;
  __syncthreads();

  // Phase 3: write this thread's new children contiguously.
  if (mine) {
    uint64_t pos = wg_base + wave_off + (incl - mine);
    if (pos + mine > next_cap) {
      errbits |= kErrFutureCap;
    } else {
      uint32_t m = newmask;
      while (m) {
        const uint32_t b = __ffs(m) - 1;
        m &= m - 1;
        const uint32_t it = b / G, g = b - it * G;
        uint64_t c = apply_groups(parents[it][threadIdx.x], gens, g);
        if (gens.tag_shift < 64) c |= static_cast<uint64_t>(gens.child_tag[g]) << gens.tag_shift;
        next[pos++] = c;
      }
    }
  }
  if (errbits) atomicOr(error, errbits);
}

template <int G>
__global__ __launch_bounds__(kBitmapThreads) void bitmap_expand_kernel(
    const uint64_t* __restrict__ frontier, uint64_t count, PackedGenerators gens,
    uint32_t* __restrict__ bitmap, uint64_t bitmap_bits, uint64_t* __restrict__ next,
    unsigned long long* __restrict__ next_count, uint64_t count_base, uint64_t next_cap,
    uint32_t* __restrict__ error) {
  bitmap_tile<G>(blockIdx.x, frontier, count, gens, bitmap, bitmap_bits, next, next_count, count_base, next_cap,
                 error);
}

// A layer whose size is read from the device, so layers can be queued
// without the host reading each size: marks[0] and marks[1] are the layer
// counter before and after the previous layer was produced, so the frontier
// holds marks[1] - marks[0] states and the children are counted from
// marks[1].  A fixed grid walks the tiles.
template <int G>
__global__ __launch_bounds__(kBitmapThreads) void bitmap_expand_marked_kernel(
    const uint64_t* __restrict__ frontier, const unsigned long long* __restrict__ marks, PackedGenerators gens,
    uint32_t* __restrict__ bitmap, uint64_t bitmap_bits, uint64_t* __restrict__ next,
    unsigned long long* __restrict__ next_count, uint64_t next_cap, uint32_t* __restrict__ error) {
  const uint64_t base = marks[1], count = base - marks[0];
  const uint64_t per_block = static_cast<uint64_t>(kBitmapThreads) * bitmap_items<G>();
  const uint64_t tiles = (count + per_block - 1) / per_block;
  for (uint64_t t = blockIdx.x; t < tiles; t += gridDim.x) {
    bitmap_tile<G>(t, frontier, count, gens, bitmap, bitmap_bits, next, next_count, base, next_cap, error);
    __syncthreads();  // the tile's shared arrays are reused by the next one
  }
}

// mark = the layer counter once the layers before it on the stream are done.
__global__ void bitmap_mark_kernel(const unsigned long long* __restrict__ counter, unsigned long long* __restrict__ mark) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *mark = *counter;
}

__global__ void bitmap_seed_kernel(uint32_t degree, uint64_t start_key, uint32_t* bitmap, uint64_t bitmap_bits,
                                   uint64_t* frontier, uint32_t* error) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  bool bad;
  const uint64_t r = lehmer_rank_packed(start_key, degree, &bad);
  if (bad || r >= bitmap_bits) {
    atomicOr(error, kErrDomain);
    return;
  }
  atomicOr(bitmap + (r >> 5), 1u << (r & 31u));
  frontier[0] = start_key;
}

template <int G>
static hipError_t launch_expand(const uint64_t* frontier, uint64_t count, const PackedGenerators& gens,
                                uint32_t* bitmap, uint64_t bits, uint64_t* next, unsigned long long* next_count,
                                uint64_t count_base, uint64_t next_cap, uint32_t* error, hipStream_t stream) {
  constexpr int kItems = bitmap_items<G>();
  const uint64_t per_block = static_cast<uint64_t>(kBitmapThreads) * kItems;
  const uint64_t blocks = (count + per_block - 1) / per_block;
  if (blocks == 0) return hipSuccess;
  if (blocks > 0x7fffffffull) return hipErrorInvalidValue;
  hipLaunchKernelGGL(bitmap_expand_kernel<G>, dim3(static_cast<uint32_t>(blocks)), dim3(kBitmapThreads), 0, stream,
                     frontier, count, gens, bitmap, bits, next, next_count, count_base, next_cap, error);
  return hipGetLastError();
}

hipError_t bitmap_expand(const uint64_t* frontier, uint64_t count, const PackedGenerators& gens, uint32_t* bitmap,
                         uint64_t bitmap_bits, uint64_t* next, unsigned long long* next_count, uint64_t count_base,
                         uint64_t next_cap, uint32_t* error, hipStream_t stream) {
  switch (gens.count) {
#define MGBFS_G(k)                                                                                            \
  case k:                                                                                                     \
    return launch_expand<k>(frontier, count, gens, bitmap, bitmap_bits, next, next_count, count_base, next_cap, \
                            error, stream);
    MGBFS_G(1) MGBFS_G(2) MGBFS_G(3) MGBFS_G(4) MGBFS_G(5) MGBFS_G(6) MGBFS_G(7) MGBFS_G(8)
    MGBFS_G(9) MGBFS_G(10) MGBFS_G(11) MGBFS_G(12) MGBFS_G(13) MGBFS_G(14) MGBFS_G(15) MGBFS_G(16)
#undef MGBFS_G
    default:
      return hipErrorInvalidValue;
  }
}

hipError_t bitmap_expand_marked(const uint64_t* frontier, const unsigned long long* marks,
                                const PackedGenerators& gens, uint32_t* bitmap, uint64_t bitmap_bits, uint64_t* next,
                                unsigned long long* next_count, uint64_t next_cap, uint32_t* error, uint32_t grid,
                                hipStream_t stream) {
  if (grid == 0) return hipErrorInvalidValue;
  switch (gens.count) {
#define MGBFS_G(k)                                                                                             \
  case k:                                                                                                      \
    hipLaunchKernelGGL(bitmap_expand_marked_kernel<k>, dim3(grid), dim3(kBitmapThreads), 0, stream, frontier,   \
                       marks, gens, bitmap, bitmap_bits, next, next_count, next_cap, error);                    \
    break;
    MGBFS_G(1) MGBFS_G(2) MGBFS_G(3) MGBFS_G(4) MGBFS_G(5) MGBFS_G(6) MGBFS_G(7) MGBFS_G(8)
    MGBFS_G(9) MGBFS_G(10) MGBFS_G(11) MGBFS_G(12) MGBFS_G(13) MGBFS_G(14) MGBFS_G(15) MGBFS_G(16)
#undef MGBFS_G
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t bitmap_mark(const unsigned long long* counter, unsigned long long* mark, hipStream_t stream) {
  hipLaunchKernelGGL(bitmap_mark_kernel, dim3(1), dim3(64), 0, stream, counter, mark);
  return hipGetLastError();
}

hipError_t bitmap_seed(uint32_t degree, uint64_t start_key, uint32_t* bitmap, uint64_t bitmap_bits, uint64_t* frontier,
                       uint32_t* error, hipStream_t stream) {
  hipLaunchKernelGGL(bitmap_seed_kernel, dim3(1), dim3(64), 0, stream, degree, start_key, bitmap, bitmap_bits,
                     frontier, error);
  return hipGetLastError();
}

}  // namespace mgbfs
