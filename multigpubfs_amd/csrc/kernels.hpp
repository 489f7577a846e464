// Internal launch wrappers shared by the engine and the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include "common.hpp"

namespace mgbfs {

hipError_t bitmap_seed(uint32_t degree, uint64_t start_key, uint32_t* bitmap, uint64_t bitmap_bits,
                       uint64_t* frontier, uint32_t* error, hipStream_t stream);
// New children go to next[*next_count - count_base + ...]: the counter may
// keep running across layers, count_base being its value at the layer start.
hipError_t bitmap_expand(const uint64_t* frontier, uint64_t count, const PackedGenerators& gens, uint32_t* bitmap,
                         uint64_t bitmap_bits, uint64_t* next, unsigned long long* next_count, uint64_t count_base,
                         uint64_t next_cap, uint32_t* error, hipStream_t stream);
// The same layer with its size and counter base read from the device
// (marks[0], marks[1]: the counter before and after the previous layer), on
// a fixed grid; bitmap_mark then stores the counter into the next mark.
hipError_t bitmap_expand_marked(const uint64_t* frontier, const unsigned long long* marks,
                                const PackedGenerators& gens, uint32_t* bitmap, uint64_t bitmap_bits, uint64_t* next,
                                unsigned long long* next_count, uint64_t next_cap, uint32_t* error, uint32_t grid,
                                hipStream_t stream);
hipError_t bitmap_mark(const unsigned long long* counter, unsigned long long* mark, hipStream_t stream);
// Sorted-bank backend limits.
constexpr uint32_t kMaxPassBits = 10;     // buckets per routing pass (LDS counters)
constexpr uint32_t kMaxBucketBits = 20;   // sub-buckets per layer: 2^20
constexpr uint32_t kMaxFineBits = 10;     // bank partition bits beyond the layer's K
constexpr uint32_t kUniqueSlots = 8192;   // LDS hash slots per sub-bucket
constexpr uint32_t kUniqueCap = 5120;     // candidates per sub-bucket (table load factor <= 0.625)
constexpr uint32_t kSlabStride = 5120;    // entries between sub-bucket slab starts (>= kUniqueCap)

// Stored-value layout of the sorted backend: (keymix(key) << tagbits) | tag,
// keymix on b = max(32, 4n) bits (at least 32, so keymix's mask leaves the
// low word whole) and tagbits = 64 - b.
inline uint32_t sorted_key_bits(uint32_t degree) { return degree >= 8 ? 4 * degree : 32u; }

struct SortedLayout {
  KeyMix km;
  uint32_t tagbits;  // 64 - sorted_key_bits(n)
  uint32_t prune;    // back-edge masks in use: tag bit i = skip generator i
  uint64_t tagmask;  // (1 << tagbits) - 1
};

uint64_t scan_workspace_words(uint64_t n);
hipError_t exclusive_scan_u32(const uint32_t* in, uint64_t n, uint32_t* out, uint32_t* workspace,
                              hipStream_t stream);
hipError_t sorted_seed(uint64_t start_key, const SortedLayout& L, uint64_t* bank, uint32_t* off, uint32_t Ko,
                       hipStream_t stream);
// Children of bank[0..F) routed by lane (sub-bucket >> K2) into slabs: slab
// (lane * P + blockIdx % P) of `cap` entries, `stride` entries apart,
// cursors[] counts per slab
// (zeroed by the caller); a run beyond `cap` is dropped and sets err_bit.
// Routing outputs: `out` holds out_entries slab entries followed by a dump
// area of kRouteDump entries (a slab overflow is written there; the layer
// is then redone).
constexpr uint64_t kRouteDump = 8192;
// Sub-bucket of a stored value: floor(top32 * nsub / 2^32) (any nsub, a
// multiple of 2^K2); lane = sub-bucket >> K2 (K2 = 0: lanes are sub-buckets).
hipError_t sorted_expand_route(const uint64_t* bank, uint64_t F, const PackedGenerators& gens, const SortedLayout& L,
                               bool masked, uint32_t nsub, uint32_t K2, uint32_t NB, uint32_t P, uint32_t cap,
                               uint32_t stride,
                               uint32_t err_bit, uint32_t* cursors, uint32_t* error, uint64_t* out,
                               uint64_t out_entries, hipStream_t stream);
// Lane slabs (lcount per slab, capacity and stride lcap) -> sub-bucket
// slabs of kUniqueCap entries, kSlabStride apart (scount per sub-bucket,
// zeroed by the caller).
hipError_t sorted_partition(const uint64_t* in, const uint32_t* lcount, uint32_t lcap, uint32_t P, uint32_t nsub,
                            uint32_t K2, uint32_t* scount, uint32_t* error, uint64_t* out, uint64_t out_entries,
                            hipStream_t stream);
// In-layer dedup of the sub-bucket slabs and difference against the
// current bank and, when given, the previous one (banks partitioned into
// nbo_c / nbo_p fine buckets); survivors are written back grouped by the
// layer's nsub << fb fine buckets.  merge: OR duplicates' back-edge masks.
hipError_t sorted_unique_diff(uint64_t* cand, const uint32_t* scount, uint32_t nsub, const uint64_t* cur,
                              const uint32_t* cur_off, uint32_t nbo_c, const uint64_t* prev, const uint32_t* prev_off,
                              uint32_t nbo_p, uint32_t fb, uint32_t* fine_counts, uint32_t* error,
                              unsigned long long* unique_total, unsigned long long* cand_total, uint32_t tagbits,
                              bool merge, uint32_t grid_limit, hipStream_t stream);
// Writes nothing unless *error == 0 and the layer total off[nsub << fb]
// fits bank_rows, so it may be queued before the host has read either.
hipError_t sorted_compact(const uint64_t* cand, uint32_t nsub, uint32_t fb, const uint32_t* off, const uint32_t* error,
                          uint64_t bank_rows, uint64_t* bank, hipStream_t stream);

// Sets kErrFutureCap in *error when off[n] > rows.
hipError_t sorted_capacity_check(const uint32_t* off, uint64_t n, uint64_t rows, uint32_t* error,
                                 hipStream_t stream);

hipError_t fill_u64(uint64_t* dst, uint64_t value, uint64_t count, hipStream_t stream);
hipError_t fill_u32(uint32_t* dst, uint32_t value, uint64_t count, hipStream_t stream);
hipError_t zero_u32x3(uint32_t* a, uint64_t na, uint32_t* b, uint64_t nb, uint32_t* c, uint64_t nc,
                      hipStream_t stream);

}  // namespace mgbfs
