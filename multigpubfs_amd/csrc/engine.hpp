// Host runtime of the single-GPU exact BFS path (C++ side of the C ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>
#include "common.hpp"
#include "../../include/mgbfs_hip.h"

namespace mgbfs {

struct HipError : std::runtime_error {
  hipError_t code;
  HipError(hipError_t c, const std::string& what) : std::runtime_error(what), code(c) {}
};

#define MGBFS_HIP(call)                                                                          \
  do {                                                                                           \
    hipError_t _e = (call);                                                                      \
    if (_e != hipSuccess)                                                                        \
      throw ::mgbfs::HipError(_e, std::string(#call) + ": " + hipGetErrorString(_e));            \
  } while (0)

// Owned device allocation (freed on destruction, on the owning device).
class DeviceBuffer {
 public:
  DeviceBuffer() = default;
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  DeviceBuffer(DeviceBuffer&& o) noexcept { swap(o); }
  DeviceBuffer& operator=(DeviceBuffer&& o) noexcept { swap(o); return *this; }
  ~DeviceBuffer() { release(); }
  void allocate(size_t bytes) {
    release();
    if (bytes == 0) bytes = 8;
    MGBFS_HIP(hipMalloc(&ptr_, bytes));
    bytes_ = bytes;
  }
  void release() {
    if (ptr_) (void)hipFree(ptr_);
    ptr_ = nullptr;
    bytes_ = 0;
  }
  void swap(DeviceBuffer& o) noexcept {
    std::swap(ptr_, o.ptr_);
    std::swap(bytes_, o.bytes_);
  }
  template <typename T> T* as() const { return static_cast<T*>(ptr_); }
  size_t bytes() const { return bytes_; }

 private:
  void* ptr_ = nullptr;
  size_t bytes_ = 0;
};

// Host copy of the graph description (validated).
struct Graph {
  uint32_t kind = 0, elements = 0, rows = 0, cols = 0, generators = 0;
  bool inverse_closed = false;
  std::vector<uint32_t> perms;    // [G][elements]
  std::vector<int64_t> matrices;  // [G][rows*rows]
  std::vector<uint32_t> moduli;   // [G]
  std::vector<int64_t> start;     // [elements]
  uint64_t expected_states = 0;   // expected_max_unique_states (0 = none)

  bool packable() const;      // permutation, n <= 16, all start values in 0..15
  bool distinct_perm() const; // start is a permutation of 0..n-1
  uint64_t start_key() const;
  PackedGenerators packed_generators() const;
};

class Engine {
 public:
  virtual ~Engine() = default;
  virtual int advance(uint64_t* next_count, uint32_t* fatal) = 0;
  // Runs to the end (no layer or time limit): advance() until terminal,
  // returning the last status.  A backend may queue several layers per
  // host wait instead.
  virtual int run_to_end() {
    int s = 0;
    while (!terminal) {
      uint64_t nc;
      uint32_t f;
      s = advance(&nc, &f);
    }
    return s;
  }
  virtual void reset() = 0;
  // Switches to another graph of the same shape, keeping every allocation,
  // and resets (backends that support it).
  virtual void rebind(const Graph&) { throw std::runtime_error("ENGINE_REBIND_UNSUPPORTED"); }
  // Sorted backend: describes its last two layers after advance() returned
  // kAdvanceHandover (its limits are exceeded) and frees its scratch.
  virtual bool handover(struct PackedHandover&) { return false; }
  virtual void layer_keys(std::vector<uint64_t>& out) = 0;
  virtual void sample(uint32_t which, uint64_t limit, std::vector<int64_t>& rows, uint64_t& count) = 0;
  virtual const char* name() const = 0;
  virtual uint64_t retry_count() const { return 0; }
  // Waits for device work that the last advance left queued on the engine
  // stream and folds its timing into the counters below.
  virtual void settle() {}
  virtual void fill_info(mgbfs_engine_info& info) const;
  // Per-phase device seconds and modelled streaming bytes (backend-defined).
  virtual void fill_phases(double* seconds, uint64_t* bytes, uint32_t n) const {
    for (uint32_t i = 0; i < n; ++i) {
      seconds[i] = i == 0 ? kernel_seconds : 0.0;
      bytes[i] = i == 0 ? streamed_bytes : 0;
    }
  }

  std::vector<uint64_t> layers;  // layer sizes, depth 0 first
  uint64_t candidates = 0, unique_candidates = 0, device_bytes = 0;
  double kernel_seconds = 0.0;
  uint64_t launches = 0;        // hot-kernel launches
  uint64_t streamed_bytes = 0;  // algorithmic HBM bytes of the hot kernels
  uint32_t backend = 0, history_layers = 0, state_bytes = 8;
  bool terminal = false;

 protected:
  int device_ = 0;
  hipStream_t stream_ = nullptr;
};

// Internal advance() status: the layer exceeds the sorted backend's limits
// and the run continues on the segmented backend (never returned by the C ABI).
constexpr int kAdvanceHandover = 3;

Engine* make_bitmap_engine(const Graph& g, const mgbfs_engine_options& o);
Engine* make_sorted_engine(const Graph& g, const mgbfs_engine_options& o);
Engine* make_generic_engine(const Graph& g, const mgbfs_engine_options& o, uint64_t state_bound);
Engine* make_segmented_engine(const Graph& g, const mgbfs_engine_options& o, uint64_t state_bound);
// Layers d-1 and d of a sorted-backend run, handed to the segmented backend
// when a layer exceeds the sorted backend's limits (device pointers to bank
// values (mix(key) << tagbits) | tag; prev_count 0 when there is no d-1).
struct PackedHandover {
  const uint64_t* prev = nullptr;
  uint64_t prev_count = 0;
  const uint64_t* cur = nullptr;
  uint64_t cur_count = 0;
  KeyMix km{};
  uint32_t tagbits = 0;
  std::vector<uint64_t> layers;
};
Engine* make_segmented_from_packed(const Graph& g, const mgbfs_engine_options& o, uint64_t state_bound,
                                   const PackedHandover& h);
void set_last_error(const char* what);
Graph intake_graph(const mgbfs_graph_desc* d);
uint64_t vram_budget(const mgbfs_engine_options& o);
// Mathematical bound on reachable states (0 = unknown), capped at 2^63.
uint64_t state_space_bound(const Graph& g);
// Index of each generator's inverse in the set (-1 when absent).
std::vector<int> inverse_index(const PackedGenerators& gens);
// Generic backend sizing (generic_engine.hip).
uint64_t generic_plan_bytes(uint32_t elements, uint64_t capacity, uint32_t state_bytes, uint64_t* slots_out);
uint32_t generic_state_bytes(const Graph& g);
uint64_t generic_capacity(const Graph& g, uint64_t budget, uint64_t state_bound);
uint32_t generic_batch(uint32_t generators, uint64_t capacity);
// Backend choice and memory plan of a graph, without allocating.
mgbfs_engine_plan_info plan_engine(const Graph& g, const mgbfs_engine_options& o);

}  // namespace mgbfs
