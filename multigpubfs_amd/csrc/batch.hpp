// Batched multi-instance BFS: device-resident instance table shared by the
// batch kernels and the batch engine.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include "common.hpp"

namespace mgbfs {

constexpr uint32_t kBatchMaxSlots = 1024;   // instances per batch (one scan block)
constexpr uint32_t kBatchMaxGens = 16;
constexpr uint32_t kBatchMaxEntries = 9;    // matrix states: rows*cols <= 9 packed entries

enum BatchAction : uint32_t { kActPerm = 0, kActMatrix = 1 };
enum BatchVisited : uint32_t { kVisBitmap = 0, kVisHash = 1 };

// Instance status bits.
constexpr uint32_t kInstActive = 1u;
constexpr uint32_t kInstDone = 2u;

struct BatchInstance {
  // Control (device-updated by the finish kernel).
  uint32_t status;        // kInstActive / kInstDone
  uint32_t depth;         // index of the current layer
  uint32_t error;         // sticky error bits (kErr*)
  uint32_t pad0;
  uint64_t count;         // current layer size
  unsigned long long next_count;
  uint64_t* front;        // current layer keys
  uint64_t* next;         // next layer keys
  uint64_t cap;           // capacity of front/next
  uint64_t* layer_sizes;  // [max_layers]
  uint32_t max_layers;
  // Visited set.
  uint32_t visited;       // kVisBitmap / kVisHash
  uint32_t* bitmap;       // kVisBitmap: n! bits (Lehmer rank)
  uint64_t bitmap_bits;
  uint64_t* table;        // kVisHash: open addressing, EMPTY = ~0, key in slot
  uint64_t table_slots;   // slots per table (home slot = high half of mix64(key) * slots)
  // Three rolling tables (inverse-closed hash instances): layer d's states
  // are claimed in table d % 3 at table + (d % 3) * table_stride; a child of
  // layer d is new iff it is in neither table (d - 1) % 3 nor d % 3, which
  // is exact because an undirected edge joins states at most one layer
  // apart.  Table (d + 1) % 3 is cleared before layer d is expanded.
  uint32_t rolling;
  uint32_t pad1;
  uint64_t table_stride;
  // Action.
  uint32_t action;        // kActPerm / kActMatrix
  uint32_t degree;        // permutation degree n
  uint32_t generators;
  uint32_t rows, cols;    // matrix action: rows x cols state, rows x rows generators
  uint32_t ebits;         // matrix: bits per packed entry
  int64_t modulus;        // matrix: common modulus (> 0)
  uint8_t src[kBatchMaxGens][kMaxDegree];         // permutation tables
  uint64_t src_word[kBatchMaxGens];               // the same tables, src[g][i] in nibble i
  int32_t mat[kBatchMaxGens][kBatchMaxEntries];   // matrix generators (row-major rows x rows)
  // Parent-edge tags (bitmap instances of degree <= 15): a stored key
  // carries 1 + the index of the inverse of the generator that made it in
  // nibble 15, and expanding it skips that generator (its child is the
  // parent, already visited).  tag_shift 64 = no tags.
  uint8_t child_tag[kBatchMaxGens];
  uint32_t tag_shift;
};

// Deferred descriptor upload: slot <- staged descriptor, then (seed != 0)
// the start state is claimed and becomes layer 0.
struct BatchInstall {
  uint32_t slot, seed;
  uint64_t start_key;
};

hipError_t batch_reset_flag(uint32_t* flag, hipStream_t stream);
// prefix: 2 * (slots + 1) words (expansion tiles, then clear work items).
hipError_t batch_tiles(BatchInstance* inst, uint32_t slots, uint32_t tile, uint32_t* prefix, hipStream_t stream);
// threads: 128, 256 or 512 per workgroup; tile = 2 * threads parents.
hipError_t batch_expand(BatchInstance* inst, uint32_t slots, const uint32_t* prefix, uint32_t tile, uint32_t grid,
                        uint32_t threads, hipStream_t stream);
// Clears the table the next layer of every active rolling instance claims in
// (work items from the second half of batch_tiles' prefix).
hipError_t batch_clear(const BatchInstance* inst, uint32_t slots, const uint32_t* prefix, uint32_t grid,
                       hipStream_t stream);
hipError_t batch_finish(BatchInstance* inst, uint32_t slots, uint32_t* any_active, hipStream_t stream);
hipError_t batch_gather_sizes(const BatchInstance* inst, const uint32_t* req, uint32_t n, uint64_t* out,
                              hipStream_t stream);
// Empties the visited set of every seeding request's staged descriptor
// (bitmap words to 0, hash slots to ~0); runs before batch_install.
// max_bytes: the largest of those visited sets (sizes the grid).
hipError_t batch_clear_visited(const BatchInstance* staged, const BatchInstall* req, uint32_t n, uint64_t max_bytes,
                               hipStream_t stream);
hipError_t batch_install(BatchInstance* inst, const BatchInstance* staged, const BatchInstall* req, uint32_t n,
                         hipStream_t stream);
hipError_t batch_seed(BatchInstance* inst, uint32_t slot, uint64_t start_key, hipStream_t stream);

}  // namespace mgbfs
