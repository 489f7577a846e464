/*
 * mgbfs_hip.h — C ABI of libmgbfs_hip.so, the MI355X (gfx950) exact BFS
 * engine for implicit Cayley graphs.
 *
 * This library is the drop-in replacement for what the reference's thin FFI
 * crate crates/mgbfs-cuda binds (libmgbfs_cuda.so), plus the host runtime the
 * reference keeps in crates/mgbfs-runtime.  Conventions follow the
 * reference's ABI (cuda/mgbfs_cuda.h:10-18, cuda/generic_visited.h:13-21):
 *
 *  - Return value 0 = success.  For kernel-level entry points (those taking a
 *    `void* stream`) 0 means "successfully enqueued", not completion; a
 *    nonzero value is a hipError_t (hipErrorInvalidValue on argument
 *    violations).  `stream` is a hipStream_t cast to void* (NULL = default).
 *  - Kernel-level entry points never allocate and never synchronise; every
 *    buffer is caller-owned device memory.  `error` is a caller-zeroed sticky
 *    device word; bits: 1 table full, 2 visited capacity, 4 future capacity,
 *    8 bad row reference, 32 count above bound, 64 bucket above LDS capacity,
 *    128 state outside the packed/rank domain, 256 owner-lane slab full
 *    (sorted backend; the layer is repeated with more room), 1024 segment
 *    at its fill limit (segmented backend; the batch is replayed with a
 *    larger segment).  On any error bit the future
 *    layer is invalid and the arena must not be resumed.
 *  - Engine entry points (mgbfs_engine_*) own their device memory, run on
 *    their own stream, and are not reentrant: one engine per host thread.
 *
 * Packed states: a permutation state of degree n <= 16 whose values are all
 * < 16 is the uint64 key  sum(state[i] << 4*i)  (gpu/cayley_bfs.cu:50-56,
 * scripts/state_layout.py:9-16).  A generator acts as child[i] = parent[g[i]]
 * (multigpubfs/graph_definition.py:74).
 */
#ifndef MGBFS_HIP_H
#define MGBFS_HIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MGBFS_HIP_ABI_VERSION 1u

/* Engine status values returned by mgbfs_engine_advance(). */
#define MGBFS_LAYER 0    /* a new non-empty layer was produced          */
#define MGBFS_COMPLETE 1 /* frontier exhausted: the BFS is complete      */
#define MGBFS_RESOURCE 2 /* device capacity reached: INCOMPLETE RESOURCE */

/* Backends (mgbfs_engine_options.backend). */
#define MGBFS_BACKEND_AUTO 0
#define MGBFS_BACKEND_BITMAP 1   /* exact Lehmer-rank permutation bitmap          */
#define MGBFS_BACKEND_SORTED 2   /* sort-unique + three rotating sorted banks     */
#define MGBFS_BACKEND_SEGMENTED 3 /* segmented open-addressing hash, any state width:
                                     three rolling banks (inverse-closed) or full history */
#define MGBFS_BACKEND_GENERIC 4  /* SoA int64/u8 states, all-visited exact hash   */

/* Option flags. */
#define MGBFS_FLAG_TEST_COARSE_BUCKETS 1u /* sorted backend: start 3 bucket bits coarse
                                             so large layers take the overflow-retry path */
#define MGBFS_FLAG_TEST_COLLIDE_HASH 2u   /* generic backend: hash bits 0, every state
                                             collides and full-state equality decides */
#define MGBFS_FLAG_TEST_TIGHT_SLABS 4u    /* sorted backend: owner-lane slabs start with
                                             no slack, so lane overflows take the retry path */
#define MGBFS_FLAG_TEST_NO_PRUNE 8u       /* sorted backend: no back-edge masks; every
                                             child is generated and both banks are probed */
#define MGBFS_FLAG_MAX_CAPACITY 16u       /* keep 256 MiB of free VRAM in reserve instead of
                                             2 GiB (reference memory_preset "max_capacity") */
#define MGBFS_FLAG_TEST_WIDE_ROUTING 32u  /* sorted backend: route with 64-bit slab addresses,
                                             the path taken by outputs of 2^32 entries or more */
#define MGBFS_FLAG_TEST_TINY_SEGMENTS 128u /* segmented backend: every layer starts with a
                                             1024-slot segment, so large layers spill into
                                             growing segments */
#define MGBFS_FLAG_TEST_SORTED_CAP 256u   /* sorted backend: layers of more than 2^18 children
                                             count as beyond its limits, so the run is handed
                                             to the segmented backend (as at F*G > 2^32) */

/* Graph description: the schema-2 graph JSON of the reference
 * (multigpubfs/graph_definition.py:46-69) as plain arrays.  All pointers are
 * HOST pointers, consumed before mgbfs_engine_create returns. */
typedef struct mgbfs_graph_desc {
  uint32_t kind;        /* 0 permutation, 1 matrix                          */
  uint32_t elements;    /* state length (degree, or rows*cols)              */
  uint32_t rows, cols;  /* matrix shape (kind 1)                            */
  uint32_t generators;  /* number of generators                             */
  uint32_t inverse_closed; /* 1 if every generator's inverse is in the set
                              (graph_definition.py:84-108); selects three-bank
                              retention, else full visited history          */
  const uint32_t* permutations; /* kind 0: [generators][elements]           */
  const int64_t* matrices;      /* kind 1: [generators][rows*rows]          */
  const uint32_t* moduli;       /* kind 1: [generators], 0 = no reduction   */
  const int64_t* start;         /* [elements]                               */
  uint64_t expected_states;     /* expected_max_unique_states (0 = none)    */
} mgbfs_graph_desc;

typedef struct mgbfs_engine_options {
  int32_t device;        /* HIP device ordinal                                 */
  uint32_t backend;      /* MGBFS_BACKEND_*                                    */
  uint64_t capacity;     /* max states per layer (0 = plan from free VRAM)     */
  uint64_t vram_budget;  /* bytes the engine may allocate (0 = free - reserve) */
  uint32_t flags;        /* MGBFS_FLAG_* (0 for normal runs)                   */
  uint32_t batch;        /* generic backend: parents per expansion launch (0 = 2^20) */
} mgbfs_engine_options;

typedef struct mgbfs_engine_info {
  uint32_t backend;        /* backend actually used                          */
  uint32_t history_layers; /* 3 (inverse-closed) or 0 (full history)         */
  uint32_t state_bytes;    /* 8 for packed keys                              */
  uint32_t depth;          /* index of the current layer                     */
  uint64_t current_count;  /* states in the current layer                    */
  uint64_t total_states;   /* states discovered so far                       */
  uint64_t device_bytes;   /* device memory held by the engine               */
  uint64_t peak_layer;     /* largest layer so far                           */
  uint64_t candidates;     /* generated children so far (edges expanded)     */
  uint64_t unique_candidates; /* distinct children after in-layer dedup     */
  double kernel_seconds;   /* summed device time of the hot kernels          */
  uint64_t kernel_launches;   /* launches of the hot kernels                 */
  uint64_t algorithmic_bytes; /* modelled streaming HBM bytes of those kernels */
  uint64_t retries;        /* sorted backend: layer passes redone at finer buckets */
} mgbfs_engine_info;

/* Memory plan of a graph on one device, computed without allocating: the
 * admission figures of the reference's GenericMemoryPlan /
 * GenericDistributedMemoryPlan (crates/mgbfs-runtime/src/generic_memory.rs:
 * 36-49, generic_distributed_memory.rs:7-13) for the backend this library
 * would pick.  With options->vram_budget = 0 the budget is the device's free
 * memory minus the 2 GiB (or 256 MiB, MGBFS_FLAG_MAX_CAPACITY) reserve. */
typedef struct mgbfs_engine_plan_info {
  uint32_t backend;          /* MGBFS_BACKEND_* chosen                               */
  uint32_t history_layers;   /* retention rule: 3 inverse-closed, 1 full history     */
  uint32_t retained_banks;   /* layer banks kept (sorted: 2 with back-edge masks)    */
  uint32_t state_bytes;      /* bytes per stored element (8 packed/int64, 1 u8 SoA)  */
  uint64_t capacity;         /* states per layer (packed) or arena rows (generic)    */
  uint64_t batch;            /* parents expanded per launch                          */
  uint64_t queue_capacity;   /* children generated per launch at most                */
  uint64_t table_slots;      /* visited-set slots (hash slots; bitmap: bits; sorted: 0) */
  uint64_t device_bytes;     /* planned device bytes                                 */
  uint64_t budget_bytes;     /* the budget the plan was sized against                */
  uint64_t state_bound;      /* state-space bound (0 = unknown)                      */
} mgbfs_engine_plan_info;
int mgbfs_engine_plan(const mgbfs_graph_desc* graph, const mgbfs_engine_options* options,
                      mgbfs_engine_plan_info* plan, char* error_text, size_t error_capacity);

/* ------------------------------------------------------------ engine */
/* Host runtime of the single-GPU path; replaces GenericNativeBfs
 * (crates/mgbfs-runtime/src/generic_native.rs:58-145: automatic_plan, new,
 * advance, sample, previous_small_sample) and SegmentedDistributedBfs::advance
 * at world = 1 (segmented_distributed_native.rs:1116-1315). */
int mgbfs_engine_create(const mgbfs_graph_desc* graph, const mgbfs_engine_options* options, void** engine,
                        char* error_text, size_t error_capacity);
/* Expands the current layer.  Returns MGBFS_LAYER / MGBFS_COMPLETE /
 * MGBFS_RESOURCE (fatal bits in *fatal) or a negative value on a host/HIP
 * failure.  *next_count receives the size of the new current layer. */
int mgbfs_engine_advance(void* engine, uint64_t* next_count, uint32_t* fatal);
/* Runs advance() until completion, a resource stop, max_layers layers or
 * max_seconds (0 = unlimited).  Returns the last advance() status.  With
 * neither limit the permutation-bitmap backend may queue several layers per
 * host wait (same layers, same final state). */
int mgbfs_engine_run(void* engine, uint32_t max_layers, double max_seconds);
int mgbfs_engine_query(void* engine, mgbfs_engine_info* info);
/* Per-phase device seconds and modelled streaming HBM bytes of the hot
 * kernels, summed over the layers run so far.  Sorted backend phases:
 * 0 expand+route, 1 partition, 2 unique+difference, 3 compaction.
 * Bitmap backend: phase 0 is the fused expand kernel. */
int mgbfs_engine_phases(void* engine, double* seconds, uint64_t* bytes, uint32_t count);
/* Clears the visited state and re-seeds the start state, keeping every
 * allocation (repeat runs of the same graph, e.g. timing loops). */
int mgbfs_engine_reset(void* engine);
/* Switches the engine to another graph of the same shape (packed backends:
 * same degree; the sorted backend also needs an inverse-closed set) and
 * resets it, keeping every device allocation.  For sweeps and resident
 * workers that run many graphs of one degree. */
int mgbfs_engine_rebind(void* engine, const mgbfs_graph_desc* graph);
/* Layer sizes so far, depth 0 first (the report.json `layer_sizes`). */
int mgbfs_engine_layer_sizes(void* engine, uint64_t* sizes, uint32_t capacity, uint32_t* count);
/* Copies up to `limit` states of the current layer (which = 0) or of the
 * previous layer (which = 1) as int64 rows [limit][elements] in host memory. */
int mgbfs_engine_sample(void* engine, uint32_t which, uint64_t limit, int64_t* rows, uint64_t* count);
/* Packed backends: copies the current layer's packed keys (host memory). */
int mgbfs_engine_layer_keys(void* engine, uint64_t* keys, uint64_t capacity, uint64_t* count);
const char* mgbfs_engine_backend_name(void* engine);
void mgbfs_engine_destroy(void* engine);

/* -------------------------------------------------- packed bitmap kernels */
/* Seeds the start key: sets its rank bit and writes frontier[0].
 * Replaces mgbfs_bitmap_seed (cuda/permutation_bitmap.cu:87-99). */
int mgbfs_packed_bitmap_seed(uint32_t degree, uint64_t start_key, uint32_t* bitmap, uint64_t bitmap_bits,
                             uint64_t* frontier, uint32_t* error, void* stream);
/* One fused layer: children of frontier[0..count) under `generators`
 * permutations (host table [generators][degree]), Lehmer rank, bit claim,
 * append of newly claimed children to next[] (atomic *next_count).
 * Replaces mgbfs_bitmap_route + mgbfs_bitmap_accept
 * (cuda/permutation_bitmap.cu:27-48, :105-133) at world = 1. */
int mgbfs_packed_bitmap_expand(uint32_t degree, uint32_t generators, const uint32_t* permutations,
                               const uint64_t* frontier, uint64_t count, uint32_t* bitmap, uint64_t bitmap_bits,
                               uint64_t* next, uint64_t* next_count, uint64_t next_capacity, uint32_t* error,
                               void* stream);

/* ------------------------------------------------ sorted-bank layers */
/* The sort-unique owner as one call per BFS layer, for inverse-closed
 * permutation generators on nibble-packed states (degree 1..15, at most 16
 * generators).  Replaces the reference's sorted-run owner
 * mgbfs_hash_batch_accept_timed (cuda/segmented_hash_batch.cu:157-215) and
 * its batch owner route (cuda/segmented_batch_route.cu:52-151): children of
 * a bank are routed into sub-buckets, deduplicated in LDS, differenced
 * against the bank (and, without back-edge masks, the bank before it; the
 * three-bank rule of crates/mgbfs-runtime/src/generic_distributed_run.rs:40)
 * and compacted into the next bank.
 *
 * Bank format: one uint64 per state, (mix(key) << (64 - b)) | tag with
 * b = max(32, 4*degree), where mix is a bijection on b bits and the tag holds the state's
 * back-edge mask; a bank is grouped by `buckets` fine buckets with an offset
 * table of buckets + 1 uint32 entries (offsets[buckets] = states).  Offset
 * tables must hold MGBFS_SORTED_OFFSETS entries.  The seed bank has
 * MGBFS_SORTED_SEED_BUCKETS fine buckets; a layer call writes
 * plan.next_buckets.  mgbfs_sorted_unpack turns bank values into packed keys.
 *
 * After a layer call the next layer holds next_offsets[next_buckets] states
 * (a device word).  Error bits, after which `next` is untouched and the
 * layer is repeated with the error word and counters zeroed again:
 *   64  a sub-bucket above the LDS capacity: pass sub_buckets = twice the
 *       plan's sub_buckets;
 *   256 an owner-lane slab overflowed: pass slack = twice the slack;
 *   4   next_capacity below the layer: grow `next` to
 *       next_offsets[next_buckets] rows. */
#define MGBFS_SORTED_OFFSETS 1048577u     /* 2^20 + 1 */
#define MGBFS_SORTED_SEED_BUCKETS 1048576u /* 2^20 */

typedef struct mgbfs_sorted_plan_info {
  uint64_t workspace_bytes;  /* device workspace mgbfs_sorted_layer needs        */
  uint64_t candidates_bound; /* children generated at most                       */
  uint32_t sub_buckets;      /* sub-buckets of the pass                           */
  uint32_t next_buckets;     /* fine buckets of the bank written                  */
  uint32_t masked;           /* 1: back-edge masks, `previous` is not read        */
  uint32_t history_banks;    /* banks the caller rotates: 2 with masks, 3 without */
} mgbfs_sorted_plan_info;

typedef struct mgbfs_sorted_layer_args {
  uint32_t degree, generators;
  const uint32_t* permutations; /* HOST table [generators][degree]               */
  uint32_t depth;               /* depth of `bank` (0 = the seed bank)           */
  uint32_t sub_buckets;         /* 0 = sized from count                          */
  uint32_t slack;               /* lane-slab slack, 0 or 1 = default             */
  uint32_t buckets;             /* fine buckets of `bank`                        */
  const uint64_t* bank;         /* layer `depth`, `count` states                 */
  const uint32_t* offsets;
  uint64_t count;
  const uint64_t* previous;     /* layer depth-1 (read only without masks)       */
  const uint32_t* previous_offsets;
  uint32_t previous_buckets, reserved;
  uint64_t* next;               /* receives layer depth+1                        */
  uint32_t* next_offsets;
  uint64_t next_capacity;       /* rows `next` holds                             */
  uint64_t* counters;           /* device [2], caller-zeroed: += distinct
                                   children, += children generated               */
  uint32_t* error;              /* device sticky word, caller-zeroed             */
} mgbfs_sorted_layer_args;

/* Host-only: workspace size and geometry of one layer call. */
int mgbfs_sorted_plan(uint32_t degree, uint32_t generators, const uint32_t* permutations, uint32_t depth,
                      uint64_t count, uint32_t sub_buckets, uint32_t slack, mgbfs_sorted_plan_info* plan);
/* Writes the one-state seed bank of `start_key` (MGBFS_SORTED_SEED_BUCKETS
 * fine buckets). */
int mgbfs_sorted_seed(uint32_t degree, uint64_t start_key, uint64_t* bank, uint32_t* offsets, void* stream);
/* Enqueues one layer; `workspace` (256-byte aligned) holds at least
 * plan.workspace_bytes.  Calls sharing a workspace must not overlap. */
int mgbfs_sorted_layer(const mgbfs_sorted_layer_args* args, void* workspace, uint64_t workspace_bytes,
                       void* stream);
/* mgbfs_sorted_layer without a workspace argument, for callers that cannot
 * allocate (code under hipGraph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  The workspace is a
 * static device array of MGBFS_SORTED_SCRATCH_BYTES per GPU, set up on the
 * first call on each device and held until the process ends; the call
 * allocates nothing.  Layers whose plan.workspace_bytes exceed it are
 * rejected with hipErrorOutOfMemory (about 5 million children per layer
 * fit).  Every call on a GPU shares that GPU's array: calls that overlap on
 * one GPU (different streams or threads) race, so serialise them. */
#define MGBFS_SORTED_SCRATCH_BYTES (64ull << 20)
int mgbfs_sorted_layer_noalloc(const mgbfs_sorted_layer_args* args, void* stream);
/* Host-only: bank values (host memory) -> nibble-packed keys. */
int mgbfs_sorted_unpack(uint32_t degree, const uint64_t* values, uint64_t count, uint64_t* keys);

/* ------------------------------------------------- batched instances */
/* Many independent BFS instances advanced together, one layer of every
 * active instance per step, by a hipGraph-captured layer loop; instances are
 * admitted and released at layer boundaries (continuous batching).  The
 * reference has no batched kernel (its nearest relative is the sequential
 * resident worker of crates/mgbfs-runtime/src/generic_session.rs:6-71 and
 * multigpubfs/generic_session.py:56-144); parity is each instance's layer
 * table, identical to the one-at-a-time engine's. */
#define MGBFS_BATCH_FULL 3 /* no free slot or the planned VRAM exceeds the budget */

typedef struct mgbfs_batch_options {
  int32_t device;
  uint32_t max_slots;        /* <= 1024 (0 = 256)                              */
  uint64_t vram_budget;      /* bytes for admitted jobs (0 = free - 4 GiB)      */
  uint32_t layers_per_graph; /* layers captured per hipGraph launch (0 = 4)     */
  uint32_t grid;             /* persistent expand grid (0 = 8 x CUs)            */
  uint32_t threads;          /* expand workgroup: 128, 256 or 512 (0 = 256)     */
  uint32_t hash_load_pct;    /* hash tables filled to at most this % (0 = 50)   */
  uint32_t history;          /* inverse-closed hash instances: 0 = full history
                                when it fits the free budget, else three
                                rolling tables; 1 = always full history;
                                2 = always three rolling tables                */
  uint32_t layer_pct;        /* rolling instances: layer capacity as % of the
                                state bound (0 = 30); a layer that outgrows it
                                restarts the instance with twice the room      */
  uint32_t bitmap_bias_permille; /* visited 0: bitmap when its bytes are at most
                                this many thousandths of the hash tables'
                                (0 = 1000)                                     */
} mgbfs_batch_options;

int mgbfs_batch_create(const mgbfs_batch_options* options, void** batch, char* error_text, size_t error_capacity);
/* Planned device bytes of a job; visited 0 = autotuned, 1 = bitmap, 2 = hash.
 * *chosen_visited receives 1 (Lehmer-rank bitmap) or 2 (open-addressing hash). */
int mgbfs_batch_plan(void* batch, const mgbfs_graph_desc* graph, uint32_t visited, uint32_t max_layers,
                     uint64_t* bytes, uint32_t* chosen_visited);
/* Admits a job into a free slot (returns MGBFS_BATCH_FULL when none fits). */
int mgbfs_batch_admit(void* batch, const mgbfs_graph_desc* graph, uint32_t visited, uint32_t max_layers,
                      int32_t* slot);
/* Runs up to `layers` layers of all active instances; *active = still active. */
int mgbfs_batch_step(void* batch, uint32_t layers, uint32_t* active);
/* Steps until no instance is active. */
int mgbfs_batch_run(void* batch);
/* Slots that finished since the last poll, at most `capacity` of them (later
   polls report the rest); their layer tables are read back here, so
   mgbfs_batch_result makes no device access for them. */
int mgbfs_batch_poll(void* batch, int32_t* slots, uint32_t capacity, uint32_t* count);
/* Layer table of a slot (depth 0 first); status 2 = done; error = sticky bits. */
int mgbfs_batch_result(void* batch, int32_t slot, uint64_t* sizes, uint32_t capacity, uint32_t* count,
                       uint32_t* status, uint32_t* error);
/* Current-layer states of a slot as int64 rows. */
int mgbfs_batch_sample(void* batch, int32_t slot, uint64_t limit, int64_t* rows, uint64_t* count);
int mgbfs_batch_release(void* batch, int32_t slot);
int mgbfs_batch_usage(void* batch, uint64_t* used, uint64_t* budget, uint32_t* slots);
/* An admitted slot's visited set (1 bitmap, 2 hash), retained layers (3 =
   rolling tables, 0 = full history), layer capacity and restarts. */
int mgbfs_batch_instance_info(void* batch, int32_t slot, uint32_t* visited, uint32_t* history_layers,
                              uint64_t* layer_capacity, uint32_t* restarts);
void mgbfs_batch_destroy(void* batch);

/* --------------------------------------------- generic SoA kernel ABI */
/* Same signatures and argument meaning as the reference's libmgbfs_cuda
 * (cuda/generic_visited.h:3-11 and :22-44, cuda/generic_graph.h:3-6):
 * states are element-major SoA planes (int64 or uint8) with a row stride;
 * `slots` is a power-of-two open-addressing table initialised to all ones;
 * `future` receives arena row indices; counts and `error` are device words.
 * kind 0 = permutation (rows = degree, cols = 1), kind 1 = int64 matrix
 * action rows x cols with per-generator modulus (0 = wrap only).  The uint8
 * entry points accept permutation actions only.  hash_bits 0..64 (0 forces
 * every state to collide; equality always decides). */
int mgbfs_generic_seed_i64_wide(uint32_t elements, const int64_t* states, uint32_t state_stride, uint32_t state_count,
                                uint64_t* slots, uint64_t slot_capacity, uint64_t seed, uint32_t hash_bits,
                                uint32_t* error, void* stream);
int mgbfs_generic_seed_u8_wide(uint32_t elements, const uint8_t* states, uint32_t stride, uint32_t count,
                               uint64_t* slots, uint64_t capacity, uint64_t seed, uint32_t bits, uint32_t* error,
                               void* stream);
int mgbfs_generic_expand_i64_wide(uint32_t kind, uint32_t elements, uint32_t rows, uint32_t cols, uint32_t generators,
                                  const int64_t* parents, uint32_t parent_count, uint32_t parent_stride,
                                  const uint32_t* permutation_tables, const int64_t* matrix_tables,
                                  const uint32_t* moduli, uint64_t* slots, uint64_t slot_capacity, int64_t* visited,
                                  uint32_t visited_capacity, uint32_t* visited_count, uint32_t* future,
                                  uint32_t future_capacity, uint32_t* future_count, uint64_t seed, uint32_t hash_bits,
                                  uint32_t* error, void* stream);
int mgbfs_generic_expand_u8_wide(uint32_t kind, uint32_t elements, uint32_t n, uint32_t m, uint32_t generators,
                                 const uint8_t* parents, uint32_t count, uint32_t stride, const uint32_t* permutations,
                                 const int64_t* matrices, const uint32_t* moduli, uint64_t* slots,
                                 uint64_t slot_capacity, uint8_t* visited, uint32_t visited_capacity,
                                 uint32_t* visited_count, uint32_t* future, uint32_t future_capacity,
                                 uint32_t* future_count, uint64_t seed, uint32_t bits, uint32_t* error, void* stream);
int mgbfs_generic_gather_i64(uint32_t elements, const int64_t* source, uint32_t source_stride,
                             const uint32_t* indices, uint32_t count, int64_t* output, uint32_t output_stride,
                             void* stream);
int mgbfs_generic_gather_u8(uint32_t elements, const uint8_t* source, uint32_t stride, const uint32_t* indices,
                            uint32_t count, uint8_t* output, uint32_t output_stride, void* stream);
/* Materialises selected children (index = parent * generators + generator). */
int mgbfs_generic_generate_i64(uint32_t kind, uint32_t elements, uint32_t rows, uint32_t cols, uint32_t generators,
                               const int64_t* parents, uint32_t parent_count, uint32_t parent_stride,
                               const uint32_t* permutation_tables, const int64_t* matrix_tables,
                               const uint32_t* moduli, const uint64_t* selected_children, uint32_t output_count,
                               int64_t* output, uint32_t output_stride, uint32_t* device_error, void* stream);
int mgbfs_generic_generate_u8(uint32_t kind, uint32_t elements, uint32_t rows, uint32_t cols, uint32_t generators,
                              const uint8_t* parents, uint32_t parent_count, uint32_t parent_stride,
                              const uint32_t* permutation_tables, const int64_t* matrix_tables,
                              const uint32_t* moduli, const uint64_t* selected_children, uint32_t output_count,
                              uint8_t* output, uint32_t output_stride, uint32_t* device_error, void* raw_stream);

/* ---------------------------------------------------------- versions */
/* Device properties in the reference's order (cuda/generic_graph.cu:78-82):
 * values[8] = {compute units, compute major, compute minor, L2 bytes, wave
 * size, max threads per CU, max threads per workgroup, LDS bytes per CU}. */
int mgbfs_generic_hardware_info(int device, uint64_t* values);
uint32_t mgbfs_hip_abi_version(void);
/* Device properties: CU count, wave size, LDS bytes per workgroup, HBM bytes. */
int mgbfs_hip_device_info(int32_t device, uint32_t* compute_units, uint32_t* wave_size, uint32_t* lds_bytes,
                          uint64_t* total_bytes, uint64_t* free_bytes);

#ifdef __cplusplus
}
#endif
#endif /* MGBFS_HIP_H */
