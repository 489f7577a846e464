// Launch interface of the segmented open-addressing history
// (segmented_kernels.hip) shared with its engine (segmented_engine.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include "common.hpp"

namespace mgbfs {

constexpr uint32_t kSegBanks = 3;       // rolling layer banks (bank 0 only for directed graphs)
constexpr uint32_t kSegMaxProbe = 24;   // sealed segments one lookup may probe
constexpr uint32_t kErrSegmentFull = 1024u;  // active segment at its fill limit: seal, grow, replay

// One segment: 2^k u64 slots, claims refused beyond `limit` (half the slots).
struct SegSegment {
  uint64_t* slots;
  uint64_t mask;
  uint32_t limit;
  uint32_t bank;
};

struct SegProbeSet {
  uint32_t count;
  SegSegment seg[kSegMaxProbe];
};

// SoA row arenas of the banks: element e of row r at arena[e * stride + r].
template <typename S>
struct SegBanks {
  S* arena[kSegBanks];
  uint64_t stride[kSegBanks];
  uint32_t rows[kSegBanks];      // rows published before this launch (bounds published references)
  uint32_t capacity[kSegBanks];  // rows the arena holds
};

// One expansion launch: parents [first_parent, first_parent + count) of a
// bank arena, children claimed into bank `target` through `active`.
template <typename S>
struct SegBatch {
  uint32_t kind, elements, rows, cols, generators, count;
  uint64_t parent_stride, first_parent;
  const S* parents;
  const uint32_t* perms;
  const int64_t* mats;
  const uint32_t* moduli;
  SegBanks<S> banks;
  SegProbeSet probe;
  SegSegment active;
  uint32_t target, hash_bits;
  uint32_t* control;  // [rows claimed in target, slots claimed in active]
  uint32_t* error;
};

template <typename S>
hipError_t segmented_expand(const SegBatch<S>& b, hipStream_t stream);
template <typename S>
hipError_t segmented_seed(const SegBanks<S>& banks, uint32_t bank, uint32_t first, uint32_t count, uint32_t elements,
                          const SegSegment& seg, uint32_t bits, uint32_t* fill, uint32_t* error, hipStream_t stream);
// Sorted-backend bank values -> nibble rows of a u8 SoA arena.
hipError_t segmented_adopt_packed(const uint64_t* values, uint64_t count, const KeyMix& km, uint32_t tagbits,
                                  uint32_t degree, uint8_t* arena, uint64_t stride, hipStream_t stream);

}  // namespace mgbfs
