// Small device utilities: fills (used instead of hipMemsetAsync so that
// layer loops stay capturable in hipGraphs) and device queries.
#include "common.hpp"
#include "kernels.hpp"
#include "../../include/mgbfs_hip.h"

namespace mgbfs {

template <typename T>
__global__ void fill_kernel(T* dst, T value, uint64_t count) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride)
    dst[i] = value;
}

template <typename T>
static hipError_t fill(T* dst, T value, uint64_t count, hipStream_t stream) {
  if (count == 0) return hipSuccess;
  uint64_t blocks = (count + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(fill_kernel<T>, dim3(static_cast<uint32_t>(blocks)), dim3(256), 0, stream, dst, value, count);
  return hipGetLastError();
}

hipError_t fill_u64(uint64_t* dst, uint64_t value, uint64_t count, hipStream_t stream) {
  return fill<uint64_t>(dst, value, count, stream);
}
hipError_t fill_u32(uint32_t* dst, uint32_t value, uint64_t count, hipStream_t stream) {
  return fill<uint32_t>(dst, value, count, stream);
}

// Three zero fills in one launch, over the concatenated index space.
__global__ void zero3_kernel(uint32_t* a, uint64_t na, uint32_t* b, uint64_t nb, uint32_t* c, uint64_t nc) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < na + nb + nc; i += stride) {
    if (i < na) a[i] = 0;
    else if (i < na + nb) b[i - na] = 0;
    else c[i - na - nb] = 0;
  }
}

hipError_t zero_u32x3(uint32_t* a, uint64_t na, uint32_t* b, uint64_t nb, uint32_t* c, uint64_t nc,
                      hipStream_t stream) {
  const uint64_t count = na + nb + nc;
  if (count == 0) return hipSuccess;
  uint64_t blocks = (count + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(zero3_kernel, dim3(static_cast<uint32_t>(blocks)), dim3(256), 0, stream, a, na, b, nb, c, nc);
  return hipGetLastError();
}

}  // namespace mgbfs

extern "C" uint32_t mgbfs_hip_abi_version(void) { return MGBFS_HIP_ABI_VERSION; }

extern "C" int mgbfs_hip_device_info(int32_t device, uint32_t* compute_units, uint32_t* wave_size,
                                     uint32_t* lds_bytes, uint64_t* total_bytes, uint64_t* free_bytes) {
  hipDeviceProp_t prop;
  hipError_t e = hipGetDeviceProperties(&prop, device);
  if (e != hipSuccess) return e;
  int old = 0;
  (void)hipGetDevice(&old);
  if ((e = hipSetDevice(device)) != hipSuccess) return e;
  size_t fr = 0, tot = 0;
  e = hipMemGetInfo(&fr, &tot);
  (void)hipSetDevice(old);
  if (e != hipSuccess) return e;
  if (compute_units) *compute_units = static_cast<uint32_t>(prop.multiProcessorCount);
  if (wave_size) *wave_size = static_cast<uint32_t>(prop.warpSize);
  if (lds_bytes) *lds_bytes = static_cast<uint32_t>(prop.sharedMemPerBlock);
  if (total_bytes) *total_bytes = tot;
  if (free_bytes) *free_bytes = fr;
  return 0;
}

extern "C" int mgbfs_generic_hardware_info(int device, uint64_t* values) {
  if (!values) return hipErrorInvalidValue;
  hipDeviceProp_t prop;
  const hipError_t e = hipGetDeviceProperties(&prop, device);
  if (e != hipSuccess) return e;
  values[0] = static_cast<uint64_t>(prop.multiProcessorCount);
  values[1] = static_cast<uint64_t>(prop.major);
  values[2] = static_cast<uint64_t>(prop.minor);
  values[3] = static_cast<uint64_t>(prop.l2CacheSize);
  values[4] = static_cast<uint64_t>(prop.warpSize);
  values[5] = static_cast<uint64_t>(prop.maxThreadsPerMultiProcessor);
  values[6] = static_cast<uint64_t>(prop.maxThreadsPerBlock);
  values[7] = static_cast<uint64_t>(prop.maxSharedMemoryPerMultiProcessor);
  return 0;
}
