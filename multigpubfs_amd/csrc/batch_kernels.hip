// Batched multi-instance BFS kernels.  One layer of every active instance
// is four launches, all capturable in a hipGraph (no host readback, no
// memset nodes):
//   batch_tiles  : one block scans the per-instance tile counts
//                  ceil(count / tile) of the active instances -> prefix,
//                  and the clear work of rolling instances' next tables.
//   batch_clear  : empties the table each rolling instance's next layer
//                  is claimed in (it held layer depth - 2).
//   batch_expand : a persistent grid walks the global tile index space;
//                  tile t belongs to the instance i with prefix[i] <= t <
//                  prefix[i+1].  The instance descriptor (generator tables
//                  included) is staged in LDS, children are claimed in the
//                  instance's own visited set (Lehmer-rank bitmap or open
//                  addressing hash with full-key equality) and new ones are
//                  appended with one atomic per workgroup.
//   batch_finish : one thread per instance records the layer size, swaps the
//                  frontier buffers and retires finished instances.
#include <algorithm>
#include "batch.hpp"
#include "kernels.hpp"

namespace mgbfs {

namespace {

constexpr int kBatchItems = 2;  // parents per thread per tile (tile = 2 x threads parents)
constexpr uint32_t kTileRun = 16;  // consecutive tiles one block takes in turn
constexpr uint32_t kClearChunk = 8192;   // table slots per clear work item
constexpr uint64_t kRollingProbes = 512;  // probe bound of rolling tables (overflow -> restart)
constexpr uint32_t kErrorCheckTiles = 8;  // tiles between reads of a rolling instance's error bits

__device__ __forceinline__ int64_t pymod64(int64_t x, int64_t m) {
  int64_t r = x % m;
  return r < 0 ? r + m : r;
}

// Device pointers read out of descriptors (in LDS or global memory) carry no
// address space, so accesses through them compile to FLAT instructions; a
// FLAT access also counts against the LDS wait counter, so every wait for an
// LDS read would wait for the bitmap and table probes in flight as well.
// These casts make them global accesses.
template <typename T>
using global_ptr = __attribute__((address_space(1))) T*;
template <typename T>
__device__ __forceinline__ global_ptr<T> gp(T* p) {
  return (global_ptr<T>)p;
}

// child nibble i = parent nibble src[i], the sources read from one packed word.
// The loops over elements stay rolled (n is uniform): the expand kernel's
// code must fit the instruction cache, and unrolled copies for 16 elements x
// 16 generators x 2 parents did not (waves stalled on instruction fetch).
__device__ __forceinline__ uint64_t apply_word(uint64_t key, uint64_t sw, uint32_t n) {
  uint64_t child = 0;
#pragma unroll 1
  for (uint32_t i = 0; i < n; ++i) child |= ((key >> (4u * ((sw >> (4u * i)) & 15u))) & 15u) << (4u * i);
  return child;
}

// Lehmer rank of the child apply_word(key, sw, n) without forming it; R is
// uint32_t when n! < 2^32 (n <= 12).
template <typename R>
__device__ __forceinline__ R child_rank(uint64_t key, uint64_t sw, uint32_t n, bool& bad) {
  uint32_t seen = 0;
  R r = 0;
#pragma unroll 1
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t v = static_cast<uint32_t>(key >> (4u * ((sw >> (4u * i)) & 15u))) & 15u;
    bad |= (v >= n) | ((seen >> v) & 1u);
    const uint32_t d = v - static_cast<uint32_t>(__popc(seen & ((1u << v) - 1u)));
    seen |= 1u << v;
    r = r * static_cast<R>(n - i) + static_cast<R>(d);
  }
  return r;
}

__device__ __forceinline__ uint64_t batch_apply(const BatchInstance& s, uint64_t key, uint32_t g) {
  if (s.action == kActPerm) return apply_word(key, s.src_word[g], s.degree);
  const uint64_t emask = (1ull << s.ebits) - 1;
  uint64_t child = 0;
  for (uint32_t i = 0; i < s.rows; ++i)
    for (uint32_t j = 0; j < s.cols; ++j) {
      uint64_t acc = 0;  // int64 arithmetic with two's-complement wrap
      for (uint32_t k = 0; k < s.rows; ++k) {
        const int64_t e = static_cast<int64_t>((key >> ((k * s.cols + j) * s.ebits)) & emask);
        acc += static_cast<uint64_t>(static_cast<int64_t>(s.mat[g][i * s.rows + k])) * static_cast<uint64_t>(e);
      }
      const uint64_t v = static_cast<uint64_t>(pymod64(static_cast<int64_t>(acc), s.modulus));
      child |= v << ((i * s.cols + j) * s.ebits);
    }
  return child;
}

// Bitmap claims of the children of one parent (generators other than the
// one its tag names), degree <= 12 so a rank is one 32-bit word.  Generators
// go in groups of kClaimGroup: the group's ranks, then its bitmap word
// loads, then the atomics for the clear bits, so a thread keeps a group's
// probes in flight together.  Returns the claimed generators as a bit mask.
constexpr uint32_t kClaimGroup = 4;

__device__ __forceinline__ uint32_t batch_claim_bitmap(const BatchInstance& s, uint64_t key, uint32_t& err) {
  constexpr uint32_t kNone = ~0u;  // above every rank (12! < 2^32 - 1)
  const uint32_t G = s.generators, n = s.degree;
  const uint32_t skip = s.tag_shift < 64 ? static_cast<uint32_t>(key >> s.tag_shift) : 0u;
  const global_ptr<uint32_t> bitmap = gp(s.bitmap);
  uint32_t claimed = 0;
#pragma unroll 1
  for (uint32_t g0 = 0; g0 < G; g0 += kClaimGroup) {
    uint32_t rank[kClaimGroup], seen[kClaimGroup];
#pragma unroll
    for (uint32_t k = 0; k < kClaimGroup; ++k) {
      const uint32_t g = g0 + k;
      rank[k] = kNone;
      if (g >= G || g + 1 == skip) continue;
      bool bad = false;
      const uint32_t r = child_rank<uint32_t>(key, s.src_word[g], n, bad);
      if (bad || r >= s.bitmap_bits) err |= kErrDomain;
      else rank[k] = r;
    }
#pragma unroll
    for (uint32_t k = 0; k < kClaimGroup; ++k) seen[k] = rank[k] != kNone ? bitmap[rank[k] >> 5] : 0u;
#pragma unroll
    for (uint32_t k = 0; k < kClaimGroup; ++k) {
      if (rank[k] == kNone) continue;
      const uint32_t bit = 1u << (rank[k] & 31u);
      if (!(seen[k] & bit) &&
          !(__hip_atomic_fetch_or(bitmap + (rank[k] >> 5), bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & bit))
        claimed |= 1u << (g0 + k);
    }
  }
  return claimed;
}

__device__ __forceinline__ uint64_t home_slot(uint64_t key, uint64_t slots) { return __umul64hi(mix64(key), slots); }

// Rolling tables never hold a run longer than kRollingProbes while their
// layer fits its capacity (load <= hash_load_pct); a longer run means the
// layer outgrew it, and the claim fails with kErrTableFull so the instance
// restarts with more room.  A lookup that reaches the bound has seen every
// slot a claim of `key` could have used.
__device__ __forceinline__ bool table_has(const uint64_t* table, uint64_t slots, uint64_t key, uint64_t limit) {
  const global_ptr<const uint64_t> t = gp(table);
  uint64_t slot = home_slot(key, slots);
  for (uint64_t probe = 0; probe < limit; ++probe) {
    const uint64_t seen = t[slot];
    if (seen == key) return true;
    if (seen == ~0ull) return false;
    slot = slot + 1 == slots ? 0 : slot + 1;
  }
  return false;
}

// Claim-then-compare open addressing: true when this thread inserted `key`
// (at most `limit` probes).
__device__ __forceinline__ bool table_claim(uint64_t* table, uint64_t slots, uint64_t key, uint32_t& err,
                                            uint64_t limit) {
  const global_ptr<uint64_t> t = gp(table);
  uint64_t slot = home_slot(key, slots);
  for (uint64_t probe = 0; probe < limit; ++probe) {
    const uint64_t seen = t[slot];
    if (seen == key) return false;
    if (seen == ~0ull) {
      uint64_t old = ~0ull;
      if (__hip_atomic_compare_exchange_strong(t + slot, &old, key, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT))
        return true;
      if (old == key) return false;
    }
    slot = slot + 1 == slots ? 0 : slot + 1;
  }
  err |= kErrTableFull;
  return false;
}

// Returns true when this thread claimed `key` (first visit).
__device__ __forceinline__ bool batch_claim(const BatchInstance& s, uint64_t key, uint32_t& err) {
  if (s.visited == kVisBitmap) {
    bool bad;
    const uint64_t r = lehmer_rank_packed(key, s.degree, &bad);
    if (bad || r >= s.bitmap_bits) {
      err |= kErrDomain;
      return false;
    }
    const global_ptr<uint32_t> word = gp(s.bitmap) + (r >> 5);
    const uint32_t bit = 1u << (r & 31u);
    if (*word & bit) return false;
    return !(__hip_atomic_fetch_or(word, bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & bit);
  }
  if (s.rolling) {
    const uint32_t d = s.depth % 3;
    const uint64_t limit = s.table_slots < kRollingProbes ? s.table_slots : kRollingProbes;
    if (table_has(s.table + ((d + 2) % 3) * s.table_stride, s.table_slots, key, limit) ||
        table_has(s.table + d * s.table_stride, s.table_slots, key, limit))
      return false;
    return table_claim(s.table + ((d + 1) % 3) * s.table_stride, s.table_slots, key, err, limit);
  }
  return table_claim(s.table, s.table_slots, key, err, s.table_slots);
}

__global__ void batch_reset_flag_kernel(uint32_t* flag) {
  if (threadIdx.x == 0) *flag = 0;
}

// Chunk t of the clear prefix: the instance with cpre[i] <= t < cpre[i+1]
// clears slots [k, k + kClearChunk) of table (depth + 1) % 3, k = (t -
// cpre[i]) * kClearChunk.  Runs before the expansion of the same layer.
__global__ __launch_bounds__(256) void batch_clear_kernel(const BatchInstance* inst, uint32_t slots,
                                                          const uint32_t* __restrict__ cpre) {
  const uint32_t total = cpre[slots];
  for (uint32_t t = blockIdx.x; t < total; t += gridDim.x) {
    uint32_t lo = 0, hi = slots - 1;
    while (lo < hi) {
      const uint32_t mid = (lo + hi + 1) / 2;
      if (cpre[mid] <= t) lo = mid;
      else hi = mid - 1;
    }
    const BatchInstance& s = inst[lo];
    const global_ptr<uint64_t> table = gp(s.table) + ((s.depth + 1) % 3) * s.table_stride;
    const uint64_t k0 = static_cast<uint64_t>(t - cpre[lo]) * kClearChunk;
    const uint64_t k1 = k0 + kClearChunk < s.table_stride ? k0 + kClearChunk : s.table_stride;
    for (uint64_t k = k0 + threadIdx.x; k < k1; k += blockDim.x) table[k] = ~0ull;
  }
}

// Exclusive scan of v over the first `slots` threads of a 1024-thread
// block: out[i] = sum of v below i, out[slots] = total.
__device__ __forceinline__ void block_scan_out(uint32_t v, uint32_t slots, uint32_t* wave_sums, uint32_t* out) {
  const uint32_t i = threadIdx.x, lane = i & (kWave - 1), wave = i / kWave;
  uint32_t incl = v;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const uint32_t t = __shfl_up(incl, off, kWave);
    if (lane >= static_cast<uint32_t>(off)) incl += t;
  }
  if (lane == kWave - 1) wave_sums[wave] = incl;
  __syncthreads();
  uint32_t before = 0, total = 0;
  for (uint32_t w = 0; w < blockDim.x / kWave; ++w) {
    if (w < wave) before += wave_sums[w];
    total += wave_sums[w];
  }
  if (i < slots) out[i] = before + incl - v;
  if (i == 0) out[slots] = total;
  __syncthreads();  // wave_sums reuse
}

// prefix[0..slots]: expansion tiles of the active instances; prefix[slots +
// 1 ..]: clear chunks (kClearChunk slots) of the table each active rolling
// instance claims its next layer in (table (depth + 1) % 3, from depth 2).
__global__ __launch_bounds__(1024) void batch_tiles_kernel(const BatchInstance* inst, uint32_t slots, uint32_t tile,
                                                           uint32_t* prefix) {
  __shared__ uint32_t wave_sums[1024 / kWave];
  const uint32_t i = threadIdx.x;
  uint32_t v = 0, c = 0;
  if (i < slots && (inst[i].status & kInstActive) && !inst[i].error) {
    v = static_cast<uint32_t>((inst[i].count + tile - 1) / tile);
    if (inst[i].rolling && inst[i].depth >= 2)
      c = static_cast<uint32_t>((inst[i].table_stride + kClearChunk - 1) / kClearChunk);
  }
  block_scan_out(v, slots, wave_sums, prefix);
  block_scan_out(c, slots, wave_sums, prefix + slots + 1);
}

// Eight waves per SIMD (64 VGPRs, no spills): the grid of 8 blocks per CU
// is resident at once and the probes' latency is hidden across waves.
template <int kBatchThreads>
__global__ __launch_bounds__(kBatchThreads) __attribute__((amdgpu_waves_per_eu(8))) void batch_expand_kernel(
    BatchInstance* inst, uint32_t slots, const uint32_t* __restrict__ prefix, uint32_t tile) {
  __shared__ uint32_t spre[kBatchMaxSlots + 1];
  __shared__ BatchInstance si;
  __shared__ uint64_t parents[kBatchItems][kBatchThreads];
  __shared__ uint32_t wave_sums[kBatchThreads / kWave];
  __shared__ unsigned long long wg_base;
  __shared__ int staged;
  __shared__ uint32_t tile_error[2];  // alternating per tile: a slot is rewritten only after a barrier
  for (uint32_t k = threadIdx.x; k <= slots; k += kBatchThreads) spre[k] = prefix[k];
  if (threadIdx.x == 0) staged = -1;
  __syncthreads();
  const uint32_t total = spre[slots];
  // Runs of up to kTileRun consecutive tiles are dealt to blocks round-robin:
  // an instance's descriptor is staged once per run at most, and every
  // block still samples all instances of the layer.
  // Short layers take one tile per block so every block has work.
  const uint32_t run = min(kTileRun, max(1u, total / gridDim.x));
  uint32_t i = 0, phase = 0, since_check = 0;
  bool retired = false;
  for (uint32_t t0 = blockIdx.x * run; t0 < total; t0 += gridDim.x * run)
  for (uint32_t t = t0; t < t0 + run && t < total; ++t) {
    // Instance owning tile t: the last i with spre[i] <= t.
    if (!(spre[i] <= t && t < spre[i + 1])) {
      uint32_t lo = 0, hi = slots - 1;
      while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) / 2;
        if (spre[mid] <= t) lo = mid;
        else hi = mid - 1;
      }
      i = lo;
    }
    if (static_cast<int>(i) != staged) {
      __syncthreads();
      const uint32_t* src = reinterpret_cast<const uint32_t*>(inst + i);
      uint32_t* dst = reinterpret_cast<uint32_t*>(&si);
      for (uint32_t w = threadIdx.x; w < sizeof(BatchInstance) / 4; w += kBatchThreads) dst[w] = src[w];
      if (threadIdx.x == 0) staged = static_cast<int>(i);
      __syncthreads();
      since_check = kErrorCheckTiles;
    }
    // A rolling instance that outgrew its tables has an error bit set (it is
    // retired by the finish kernel and restarted by the host); its remaining
    // tiles are skipped so the full tables are not probed further.  The bit
    // is read once per kErrorCheckTiles tiles: a read and barrier per tile
    // doubled the expansion time of large bitmap instances.
    if (si.rolling) {
      if (since_check++ >= kErrorCheckTiles - 1) {
        since_check = 0;
        if (threadIdx.x == 0)
          tile_error[phase & 1] = __hip_atomic_load(&inst[i].error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        retired = tile_error[phase++ & 1] != 0;
      }
      if (retired) continue;
    }
    const uint64_t p0 = static_cast<uint64_t>(t - spre[i]) * tile;
    const uint64_t p1 = p0 + tile < si.count ? p0 + tile : si.count;
    const uint32_t G = si.generators;
    uint32_t newmask = 0, err = 0;
#pragma unroll
    for (int it = 0; it < kBatchItems; ++it) {
      const uint64_t p = p0 + static_cast<uint64_t>(it) * kBatchThreads + threadIdx.x;
      const bool valid = p < p1;
      const uint64_t key = valid ? gp(si.front)[p] : 0;
      parents[it][threadIdx.x] = key;
      if (!valid) continue;
      if (si.visited == kVisBitmap && si.degree <= 12) {
        newmask |= batch_claim_bitmap(si, key, err) << (it * kBatchMaxGens);
      } else {
        for (uint32_t g = 0; g < G; ++g)
          if (batch_claim(si, batch_apply(si, key, g), err)) newmask |= 1u << (it * kBatchMaxGens + g);
      }
    }
    // Workgroup scan of new-child counts, one reservation per tile.
    const uint32_t mine = __popc(newmask);
    const uint32_t lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    uint32_t incl = mine;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
      const uint32_t v = __shfl_up(incl, off, kWave);
      if (lane >= static_cast<uint32_t>(off)) incl += v;
    }
    if (lane == kWave - 1) wave_sums[wave] = incl;
    __syncthreads();
    uint32_t before = 0, sum = 0;
#pragma unroll
    for (int w = 0; w < kBatchThreads / kWave; ++w) {
      if (w < static_cast<int>(wave)) before += wave_sums[w];
      sum += wave_sums[w];
    }
    if (threadIdx.x == 0) wg_base = sum ? atomicAdd(&inst[i].next_count, static_cast<unsigned long long>(sum)) : 0ull;
    __syncthreads();
    if (mine) {
      uint64_t pos = wg_base + before + incl - mine;
      if (pos + mine > si.cap) {
        err |= kErrFutureCap;
      } else {
        uint32_t m = newmask;
        while (m) {
          const uint32_t b = __ffs(m) - 1;
          m &= m - 1;
          const uint32_t g = b % kBatchMaxGens;
          const uint64_t tag = si.tag_shift < 64 ? static_cast<uint64_t>(si.child_tag[g]) << si.tag_shift : 0ull;
          gp(si.next)[pos++] = batch_apply(si, parents[b / kBatchMaxGens][threadIdx.x], g) | tag;
        }
      }
    }
    if (err) atomicOr(&inst[i].error, err);
    __syncthreads();  // wave_sums / wg_base / parents reuse
  }
}

__global__ void batch_finish_kernel(BatchInstance* inst, uint32_t slots, uint32_t* any_active) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= slots) return;
  BatchInstance& s = inst[i];
  if (!(s.status & kInstActive)) return;
  if (s.error) {
    s.status = kInstDone;
    return;
  }
  const uint64_t nc = s.next_count;
  if (nc == 0) {
    s.status = kInstDone;
    return;
  }
  if (s.depth + 1 >= s.max_layers) {
    s.error |= kErrCount;
    s.status = kInstDone;
    return;
  }
  s.layer_sizes[s.depth + 1] = nc;
  uint64_t* f = s.front;
  s.front = s.next;
  s.next = f;
  s.count = nc;
  s.next_count = 0;
  s.depth += 1;
  atomicOr(any_active, 1u);
}

// Layer tables of finished instances into one staging buffer: request q is
// (slot, offset, count) and copies layer_sizes[0, count) to out[offset...].
__global__ void batch_gather_sizes_kernel(const BatchInstance* inst, const uint32_t* req, uint32_t n, uint64_t* out) {
  for (uint32_t q = blockIdx.x; q < n; q += gridDim.x) {
    const global_ptr<const uint64_t> sizes = gp(static_cast<const uint64_t*>(inst[req[3 * q]].layer_sizes));
    const uint32_t off = req[3 * q + 1], count = req[3 * q + 2];
    for (uint32_t k = threadIdx.x; k < count; k += blockDim.x) out[off + k] = sizes[k];
  }
}

__device__ void seed_instance(BatchInstance& s, uint64_t start_key) {
  uint32_t err = 0;
  const bool claimed = s.visited == kVisHash && s.rolling ? table_claim(s.table, s.table_slots, start_key, err, s.table_slots)
                                                          : batch_claim(s, start_key, err);
  if (!claimed) err |= kErrDomain;
  s.front[0] = start_key;
  s.count = 1;
  s.next_count = 0;
  s.depth = 0;
  s.layer_sizes[0] = 1;
  s.error = err;
  s.status = err ? kInstDone : kInstActive;
}

__global__ void batch_seed_kernel(BatchInstance* inst, uint32_t slot, uint64_t start_key) {
  if (threadIdx.x == 0) seed_instance(inst[slot], start_key);
}

// Blocks (x, y) clear the visited set of request y (and y + gridDim.y ...)
// when it seeds: a Lehmer bitmap of ceil(n! / 32) words to 0, or the hash
// table(s), (rolling ? 3 : 1) x table_slots slots, to ~0.
__global__ __launch_bounds__(256) void batch_clear_visited_kernel(const BatchInstance* staged,
                                                                   const BatchInstall* req, uint32_t n) {
  for (uint32_t q = blockIdx.y; q < n; q += gridDim.y) {
    if (!req[q].seed) continue;
    const BatchInstance& d = staged[q];
    const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
    const uint64_t first = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (d.visited == kVisBitmap) {
      const global_ptr<uint32_t> bitmap = gp(d.bitmap);
      const uint64_t words = (d.bitmap_bits + 31) / 32;
      for (uint64_t k = first; k < words; k += stride) bitmap[k] = 0u;
    } else {
      const global_ptr<uint64_t> table = gp(d.table);
      const uint64_t slots = (d.rolling ? 3ull : 1ull) * d.table_slots;
      for (uint64_t k = first; k < slots; k += stride) table[k] = ~0ull;
    }
  }
}

// Block q writes staged descriptor q into its slot and, when asked, seeds it.
__global__ void batch_install_kernel(BatchInstance* inst, const BatchInstance* staged, const BatchInstall* req,
                                     uint32_t n) {
  for (uint32_t q = blockIdx.x; q < n; q += gridDim.x) {
    const uint32_t* src = reinterpret_cast<const uint32_t*>(staged + q);
    uint32_t* dst = reinterpret_cast<uint32_t*>(inst + req[q].slot);
    for (uint32_t w = threadIdx.x; w < sizeof(BatchInstance) / 4; w += blockDim.x) dst[w] = src[w];
    __syncthreads();
    if (threadIdx.x == 0 && req[q].seed) seed_instance(inst[req[q].slot], req[q].start_key);
    __syncthreads();
  }
}

}  // namespace

hipError_t batch_reset_flag(uint32_t* flag, hipStream_t stream) {
  hipLaunchKernelGGL(batch_reset_flag_kernel, dim3(1), dim3(64), 0, stream, flag);
  return hipGetLastError();
}

hipError_t batch_tiles(BatchInstance* inst, uint32_t slots, uint32_t tile, uint32_t* prefix, hipStream_t stream) {
  if (slots == 0 || slots > kBatchMaxSlots || tile == 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(batch_tiles_kernel, dim3(1), dim3(1024), 0, stream, inst, slots, tile, prefix);
  return hipGetLastError();
}

hipError_t batch_expand(BatchInstance* inst, uint32_t slots, const uint32_t* prefix, uint32_t tile, uint32_t grid,
                        uint32_t threads, hipStream_t stream) {
  if (slots == 0 || slots > kBatchMaxSlots || tile != threads * kBatchItems || grid == 0) return hipErrorInvalidValue;
  switch (threads) {
    case 128:
      hipLaunchKernelGGL(batch_expand_kernel<128>, dim3(grid), dim3(128), 0, stream, inst, slots, prefix, tile);
      break;
    case 256:
      hipLaunchKernelGGL(batch_expand_kernel<256>, dim3(grid), dim3(256), 0, stream, inst, slots, prefix, tile);
      break;
    case 512:
      hipLaunchKernelGGL(batch_expand_kernel<512>, dim3(grid), dim3(512), 0, stream, inst, slots, prefix, tile);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t batch_clear(const BatchInstance* inst, uint32_t slots, const uint32_t* prefix, uint32_t grid,
                       hipStream_t stream) {
  if (slots == 0 || slots > kBatchMaxSlots || grid == 0) return hipErrorInvalidValue;
  hipLaunchKernelGGL(batch_clear_kernel, dim3(grid), dim3(256), 0, stream, inst, slots, prefix + slots + 1);
  return hipGetLastError();
}

hipError_t batch_finish(BatchInstance* inst, uint32_t slots, uint32_t* any_active, hipStream_t stream) {
  hipLaunchKernelGGL(batch_finish_kernel, dim3((slots + 255) / 256), dim3(256), 0, stream, inst, slots, any_active);
  return hipGetLastError();
}

hipError_t batch_gather_sizes(const BatchInstance* inst, const uint32_t* req, uint32_t n, uint64_t* out,
                              hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(batch_gather_sizes_kernel, dim3(std::min<uint32_t>(n, 1024)), dim3(64), 0, stream, inst, req, n,
                     out);
  return hipGetLastError();
}

hipError_t batch_install(BatchInstance* inst, const BatchInstance* staged, const BatchInstall* req, uint32_t n,
                         hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(batch_install_kernel, dim3(std::min<uint32_t>(n, 1024)), dim3(64), 0, stream, inst, staged, req,
                     n);
  return hipGetLastError();
}

hipError_t batch_clear_visited(const BatchInstance* staged, const BatchInstall* req, uint32_t n, uint64_t max_bytes,
                               hipStream_t stream) {
  if (n == 0 || max_bytes == 0) return hipSuccess;
  // About 64 KB of the largest set per workgroup (256 bytes per thread).
  const uint32_t x = static_cast<uint32_t>(std::min<uint64_t>(1024, (max_bytes + 65535) / 65536));
  hipLaunchKernelGGL(batch_clear_visited_kernel, dim3(x, std::min<uint32_t>(n, 1024)), dim3(256), 0, stream, staged,
                     req, n);
  return hipGetLastError();
}

hipError_t batch_seed(BatchInstance* inst, uint32_t slot, uint64_t start_key, hipStream_t stream) {
  hipLaunchKernelGGL(batch_seed_kernel, dim3(1), dim3(64), 0, stream, inst, slot, start_key);
  return hipGetLastError();
}

}  // namespace mgbfs
