// Internal launch wrappers shared by the engine and the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include "common.hpp"

namespace mgbfs {

hipError_t bitmap_seed(uint32_t degree, uint64_t start_key, uint32_t* bitmap, uint64_t bitmap_bits,
                       uint64_t* frontier, uint32_t* error, hipStream_t stream);
hipError_t bitmap_expand(const uint64_t* frontier, uint64_t count, const PackedGenerators& gens, uint32_t* bitmap,
                         uint64_t bitmap_bits, uint64_t* next, unsigned long long* next_count, uint64_t next_cap,
                         uint32_t* error, hipStream_t stream);
hipError_t fill_u64(uint64_t* dst, uint64_t value, uint64_t count, hipStream_t stream);
hipError_t fill_u32(uint32_t* dst, uint32_t value, uint64_t count, hipStream_t stream);

}  // namespace mgbfs
