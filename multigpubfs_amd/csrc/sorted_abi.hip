// Kernel-level C ABI of the sorted backend: one call per BFS layer on
// caller-owned banks and a caller-owned workspace.
//
// Replaces the reference's sorted-run owner, mgbfs_hash_batch_accept_timed
// (cuda/segmented_hash_batch.cu:157-215: prepare -> merge sort by (hash,
// state) -> unique_and_probe against bank runs -> exclusive scan ->
// materialize), together with its batch owner route
// (cuda/segmented_batch_route.cu:52-151).  Here one call enqueues the
// sorted engine's layer pipeline: expand+route into sub-bucket slabs
// (through owner-lane slabs above 2^10 sub-buckets), LDS unique+difference
// against the bank (and the bank before it without back-edge masks), a scan
// of the fine-bucket counts into the next bank's offsets, a capacity check,
// and the compaction into the next bank.  Like every kernel-level entry
// point it allocates nothing and never synchronises; overflows are reported
// through the sticky error word and the caller repeats the layer.
#include <cstring>
#include <mutex>
#include "../../include/mgbfs_hip.h"
#include "common.hpp"
#include "kernels.hpp"
#include "sorted_plan.hpp"

namespace mgbfs {

namespace {

constexpr uint64_t kAlign = 256;

uint64_t align_up(uint64_t x) { return (x + kAlign - 1) / kAlign * kAlign; }

// Generators, stored-value layout and one layer's geometry.
struct LayerPlan {
  PackedGenerators gens;
  SortedLayout layout;
  SortedGeometry geo;
  bool masked;
  uint64_t cbound;
  // Workspace carve-up (byte offsets) and its size.
  uint64_t o_scount, o_lcount, o_fine, o_scan, o_lane, o_sub, bytes;
};

// Validates the generator table and derives the plan; false on an invalid
// argument (degree outside 1..15, generator count outside 1..16, a table
// that is not a permutation, a set that is not inverse-closed, or a layer
// whose children do not fit 32-bit counts).
bool make_plan(uint32_t degree, uint32_t generators, const uint32_t* permutations, uint32_t depth, uint64_t count,
               uint32_t sub_buckets, uint32_t slack, LayerPlan& p) {
  if (degree == 0 || degree > 15 || generators == 0 || generators > kMaxGenerators || !permutations || count == 0)
    return false;
  std::memset(&p.gens, 0, sizeof(p.gens));
  p.gens.count = generators;
  p.gens.degree = degree;
  p.gens.tag_shift = 64;
  for (uint32_t g = 0; g < generators; ++g) {
    uint32_t seen = 0;
    for (uint32_t i = 0; i < degree; ++i) {
      const uint32_t v = permutations[g * degree + i];
      if (v >= degree || ((seen >> v) & 1u)) return false;
      seen |= 1u << v;
      p.gens.src[g][i] = static_cast<uint8_t>(v);
    }
  }
  fill_shift_groups(p.gens);
  if (!fill_back_edge_masks(p.gens)) return false;
  p.layout.km = make_keymix(sorted_key_bits(degree));
  p.layout.tagbits = 64 - sorted_key_bits(degree);
  p.layout.tagmask = (1ull << p.layout.tagbits) - 1;
  p.layout.prune = generators + 2 <= p.layout.tagbits ? 1u : 0u;
  if (!p.layout.prune)
    for (uint32_t g = 0; g < generators; ++g) p.gens.child_mask[g] = 0;
  p.masked = p.layout.prune && depth > 0;
  p.cbound = p.masked ? count * (generators > 1 ? generators - 1 : 1) : count * generators;
  if (count > 0xffffffffull || p.cbound > 0xfffffff0ull) return false;
  const uint64_t want = sub_buckets ? std::min<uint64_t>(sub_buckets, 1ull << kMaxBucketBits) : sub_buckets_for(p.cbound);
  p.geo = sorted_geometry(want, count, p.cbound, slack ? slack : 1, fine_margin(generators));
  uint64_t at = 0;
  p.o_scount = at;
  at = align_up(at + 4ull * p.geo.nsub);
  p.o_lcount = at;
  at = align_up(at + 4ull * (p.geo.K2 ? p.geo.lslabs : 0));
  p.o_fine = at;
  at = align_up(at + 4ull * (p.geo.nbo + 1));
  p.o_scan = at;
  at = align_up(at + 4ull * (scan_workspace_words(p.geo.nbo) + 16));
  p.o_lane = at;
  at = align_up(at + 8ull * (p.geo.K2 ? p.geo.lane_entries + kRouteDump : 0));
  p.o_sub = at;
  at = align_up(at + 8ull * (p.geo.sub_entries + kRouteDump));
  p.bytes = at;
  return true;
}

// Enqueues one layer with the workspace at `ws` (at least plan.bytes).
hipError_t sorted_layer_enqueue(const mgbfs_sorted_layer_args& a, const LayerPlan& p, uint8_t* ws,
                                hipStream_t stream) {
  const SortedGeometry& g = p.geo;
  auto* scount = reinterpret_cast<uint32_t*>(ws + p.o_scount);
  auto* lcount = reinterpret_cast<uint32_t*>(ws + p.o_lcount);
  auto* fine = reinterpret_cast<uint32_t*>(ws + p.o_fine);
  auto* scan = reinterpret_cast<uint32_t*>(ws + p.o_scan);
  auto* lane = reinterpret_cast<uint64_t*>(ws + p.o_lane);
  auto* sub = reinterpret_cast<uint64_t*>(ws + p.o_sub);
  const bool probe_prev = !p.layout.prune && a.depth > 0;
  hipError_t e;
  if ((e = zero_u32x3(scount, g.nsub, lcount, g.K2 ? g.lslabs : 0, nullptr, 0, stream)) != hipSuccess) return e;
  e = sorted_expand_route(a.bank, a.count, p.gens, p.layout, p.masked, g.nsub, g.K2, g.NB1, g.P, g.lcap,
                          g.K2 ? g.lcap : kSlabStride, g.K2 ? kErrLaneSlab : kErrBucketCap, g.K2 ? lcount : scount,
                          a.error, g.K2 ? lane : sub, g.K2 ? g.lane_entries : g.sub_entries, stream);
  if (e != hipSuccess) return e;
  if (g.K2 && (e = sorted_partition(lane, lcount, g.lcap, g.P, g.nsub, g.K2, scount, a.error, sub, g.sub_entries,
                                    stream)) != hipSuccess)
    return e;
  e = sorted_unique_diff(sub, scount, g.nsub, a.bank, a.offsets, a.buckets, probe_prev ? a.previous : nullptr,
                         probe_prev ? a.previous_offsets : nullptr, probe_prev ? a.previous_buckets : 0, g.fb, fine,
                         a.error, reinterpret_cast<unsigned long long*>(a.counters),
                         reinterpret_cast<unsigned long long*>(a.counters) + 1, p.layout.tagbits, p.layout.prune != 0,
                         1u << kMaxBucketBits, stream);
  if (e != hipSuccess) return e;
  if ((e = exclusive_scan_u32(fine, g.nbo, a.next_offsets, scan, stream)) != hipSuccess) return e;
  if ((e = sorted_capacity_check(a.next_offsets, g.nbo, a.next_capacity, a.error, stream)) != hipSuccess) return e;
  return sorted_compact(sub, g.nsub, g.fb, a.next_offsets, a.error, a.next_capacity, a.next, stream);
}

// Argument checks shared by the layer entry points; fills the plan.
bool sorted_layer_check(const mgbfs_sorted_layer_args* a, LayerPlan& p) {
  if (!a || !make_plan(a->degree, a->generators, a->permutations, a->depth, a->count, a->sub_buckets, a->slack, p))
    return false;
  if (!a->bank || !a->offsets || a->buckets == 0 || a->buckets > (1u << kMaxBucketBits) || !a->next ||
      !a->next_offsets || !a->counters || !a->error)
    return false;
  const bool probe_prev = !p.layout.prune && a->depth > 0;
  if (probe_prev && (!a->previous || !a->previous_offsets || a->previous_buckets == 0 ||
                     a->previous_buckets > (1u << kMaxBucketBits)))
    return false;
  return true;
}

}  // namespace

// Scratch of mgbfs_sorted_layer_noalloc: one array per GPU, loaded with the
// code object and held until the process ends.  Its name is a host-side
// stand-in; kernels get the device address that hipGetSymbolAddress
// returns for the device, looked up once per device.
static_assert(MGBFS_SORTED_SCRATCH_BYTES % 256 == 0, "scratch size keeps 256-byte alignment");
static __device__ __align__(256) uint8_t scratch[MGBFS_SORTED_SCRATCH_BYTES];

namespace {

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxScratchDevices] = {};

// Device address of `scratch` on the stream's device (the current device
// for the null stream), looked up with that device current.
hipError_t scratch_address(hipStream_t stream, uint8_t** out) {
  int current = 0, device = 0;
  hipError_t e = hipGetDevice(&current);
  if (e != hipSuccess) return e;
  device = current;
  if (stream && (e = hipStreamGetDevice(stream, &device)) != hipSuccess) return e;
  if (device < 0 || device >= kMaxScratchDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[device]) {
    if (device != current && (e = hipSetDevice(device)) != hipSuccess) return e;
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (device != current) {
      const hipError_t r = hipSetDevice(current);
      if (e == hipSuccess) e = r;
    }
    if (e != hipSuccess) return e;
    if (!ptr) return hipErrorInvalidSymbol;
    g_scratch_ptr[device] = ptr;
  }
  *out = static_cast<uint8_t*>(g_scratch_ptr[device]);
  return hipSuccess;
}

}  // namespace

}  // namespace mgbfs

using namespace mgbfs;

extern "C" int mgbfs_sorted_plan(uint32_t degree, uint32_t generators, const uint32_t* permutations, uint32_t depth,
                                 uint64_t count, uint32_t sub_buckets, uint32_t slack, mgbfs_sorted_plan_info* plan) {
  LayerPlan p;
  if (!plan || !make_plan(degree, generators, permutations, depth, count, sub_buckets, slack, p))
    return hipErrorInvalidValue;
  plan->workspace_bytes = p.bytes;
  plan->candidates_bound = p.cbound;
  plan->sub_buckets = p.geo.nsub;
  plan->next_buckets = p.geo.nbo;
  plan->masked = p.masked ? 1u : 0u;
  plan->history_banks = p.layout.prune ? 2u : 3u;
  return 0;
}

extern "C" int mgbfs_sorted_seed(uint32_t degree, uint64_t start_key, uint64_t* bank, uint32_t* offsets,
                                 void* stream) {
  if (degree == 0 || degree > 15 || !bank || !offsets || (start_key >> (4 * degree)) != 0) return hipErrorInvalidValue;
  SortedLayout L{};
  L.km = make_keymix(sorted_key_bits(degree));
  L.tagbits = 64 - sorted_key_bits(degree);
  L.tagmask = (1ull << L.tagbits) - 1;
  return sorted_seed(start_key, L, bank, offsets, kMaxBucketBits, static_cast<hipStream_t>(stream));
}

extern "C" int mgbfs_sorted_layer(const mgbfs_sorted_layer_args* args, void* workspace, uint64_t workspace_bytes,
                                  void* stream) {
  LayerPlan p;
  if (!sorted_layer_check(args, p) || !workspace || workspace_bytes < p.bytes ||
      reinterpret_cast<uintptr_t>(workspace) % kAlign)
    return hipErrorInvalidValue;
  return sorted_layer_enqueue(*args, p, static_cast<uint8_t*>(workspace), static_cast<hipStream_t>(stream));
}

extern "C" int mgbfs_sorted_layer_noalloc(const mgbfs_sorted_layer_args* args, void* stream) {
  LayerPlan p;
  if (!sorted_layer_check(args, p)) return hipErrorInvalidValue;
  if (p.bytes > MGBFS_SORTED_SCRATCH_BYTES) return hipErrorOutOfMemory;
  uint8_t* ws = nullptr;
  const hipError_t e = scratch_address(static_cast<hipStream_t>(stream), &ws);
  if (e != hipSuccess) return e;
  return sorted_layer_enqueue(*args, p, ws, static_cast<hipStream_t>(stream));
}

extern "C" int mgbfs_sorted_unpack(uint32_t degree, const uint64_t* values, uint64_t count, uint64_t* keys) {
  if (degree == 0 || degree > 15 || (count && (!values || !keys))) return hipErrorInvalidValue;
  const KeyMix km = make_keymix(sorted_key_bits(degree));
  const uint32_t tagbits = 64 - sorted_key_bits(degree);
  for (uint64_t i = 0; i < count; ++i) keys[i] = keyunmix(values[i] >> tagbits, km);
  return 0;
}
