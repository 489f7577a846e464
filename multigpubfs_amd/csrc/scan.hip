// Exclusive prefix sum of uint32 counts (three passes: block sums, a carry
// scan of the block sums, block-local scans).  out has n + 1 entries; out[n]
// is the total.  Totals must fit in uint32 (callers bound their capacities).
#include "common.hpp"
#include "kernels.hpp"

namespace mgbfs {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;  // 4096 counts per block

__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v) {
  const uint32_t lane = threadIdx.x & (kWave - 1);
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const uint32_t t = __shfl_up(v, off, kWave);
    if (lane >= static_cast<uint32_t>(off)) v += t;
  }
  return v;
}

// Block-wide exclusive scan of one value per thread; returns the exclusive
// prefix and writes the block total to *total (all threads).
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* wave_sums, uint32_t* total) {
  const uint32_t lane = threadIdx.x & (kWave - 1);
  const uint32_t wave = threadIdx.x / kWave;
  const uint32_t incl = wave_inclusive_scan(v);
  if (lane == kWave - 1) wave_sums[wave] = incl;
  __syncthreads();
  uint32_t before = 0, all = 0;
  const uint32_t waves = blockDim.x / kWave;
  for (uint32_t w = 0; w < waves; ++w) {
    const uint32_t s = wave_sums[w];
    if (w < wave) before += s;
    all += s;
  }
  __syncthreads();
  *total = all;
  return before + incl - v;
}

__global__ __launch_bounds__(kScanThreads) void scan_reduce_kernel(const uint32_t* __restrict__ in, uint64_t n,
                                                                   uint32_t* __restrict__ block_sums) {
  __shared__ uint32_t wave_sums[kScanThreads / kWave];
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kScanTile;
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const uint64_t i = base + static_cast<uint64_t>(k) * kScanThreads + threadIdx.x;
    if (i < n) s += in[i];
  }
  uint32_t total;
  block_exclusive_scan(s, wave_sums, &total);
  if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

// Single block: exclusive scan of the block sums in place (any count).
__global__ __launch_bounds__(kScanThreads) void scan_carry_kernel(uint32_t* __restrict__ sums, uint32_t count) {
  __shared__ uint32_t wave_sums[kScanThreads / kWave];
  uint32_t carry = 0;
  for (uint32_t base = 0; base < count; base += kScanThreads) {
    const uint32_t i = base + threadIdx.x;
    const uint32_t v = i < count ? sums[i] : 0;
    uint32_t total;
    const uint32_t ex = block_exclusive_scan(v, wave_sums, &total);
    if (i < count) sums[i] = carry + ex;
    carry += total;
  }
  if (threadIdx.x == 0) sums[count] = carry;
}

__global__ __launch_bounds__(kScanThreads) void scan_downsweep_kernel(const uint32_t* __restrict__ in, uint64_t n,
                                                                      const uint32_t* __restrict__ block_offsets,
                                                                      uint32_t* __restrict__ out) {
  __shared__ uint32_t wave_sums[kScanThreads / kWave];
  // Thread t owns the contiguous run [base + t*items, base + (t+1)*items).
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kScanTile + static_cast<uint64_t>(threadIdx.x) * kScanItems;
  uint32_t v[kScanItems];
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const uint64_t i = base + k;
    v[k] = i < n ? in[i] : 0;
    s += v[k];
  }
  uint32_t total;
  uint32_t run = block_exclusive_scan(s, wave_sums, &total) + block_offsets[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const uint64_t i = base + k;
    if (i < n) out[i] = run;
    run += v[k];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = block_offsets[gridDim.x];
}

uint64_t scan_workspace_words(uint64_t n) { return (n + kScanTile - 1) / kScanTile + 1; }

hipError_t exclusive_scan_u32(const uint32_t* in, uint64_t n, uint32_t* out, uint32_t* workspace,
                              hipStream_t stream) {
  if (n == 0) return hipErrorInvalidValue;
  const uint64_t blocks = (n + kScanTile - 1) / kScanTile;
  if (blocks > 0x7fffffffull) return hipErrorInvalidValue;
  hipLaunchKernelGGL(scan_reduce_kernel, dim3(static_cast<uint32_t>(blocks)), dim3(kScanThreads), 0, stream, in, n,
                     workspace);
  hipLaunchKernelGGL(scan_carry_kernel, dim3(1), dim3(kScanThreads), 0, stream, workspace,
                     static_cast<uint32_t>(blocks));
  hipLaunchKernelGGL(scan_downsweep_kernel, dim3(static_cast<uint32_t>(blocks)), dim3(kScanThreads), 0, stream, in, n,
                     workspace, out);
  return hipGetLastError();
}

}  // namespace mgbfs
