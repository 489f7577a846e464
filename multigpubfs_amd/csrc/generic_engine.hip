// Host runtime of the generic SoA path (any state width, permutation or
// int64 matrix action, exact all-visited table).  Restates the layer loop of
// GenericNativeBfs (crates/mgbfs-runtime/src/generic_native.rs:106-126):
// gather parents -> expand (batches of <= 2^20 parents) -> one 12-byte
// readback per layer [visited_used, future_count, error] -> swap.
// Memory plan: crates/mgbfs-runtime/src/generic_memory.rs:36-41 (arena
// elements x capacity x state_bytes, next_pow2(2 capacity) table slots).
#include <algorithm>
#include <cstring>
#include "engine.hpp"
#include "kernels.hpp"

extern "C" {
int mgbfs_generic_seed_i64_wide(uint32_t, const int64_t*, uint32_t, uint32_t, uint64_t*, uint64_t, uint64_t, uint32_t,
                                uint32_t*, void*);
int mgbfs_generic_seed_u8_wide(uint32_t, const uint8_t*, uint32_t, uint32_t, uint64_t*, uint64_t, uint64_t, uint32_t,
                               uint32_t*, void*);
int mgbfs_generic_expand_i64_wide(uint32_t, uint32_t, uint32_t, uint32_t, uint32_t, const int64_t*, uint32_t, uint32_t,
                                  const uint32_t*, const int64_t*, const uint32_t*, uint64_t*, uint64_t, int64_t*,
                                  uint32_t, uint32_t*, uint32_t*, uint32_t, uint32_t*, uint64_t, uint32_t, uint32_t*,
                                  void*);
int mgbfs_generic_expand_u8_wide(uint32_t, uint32_t, uint32_t, uint32_t, uint32_t, const uint8_t*, uint32_t, uint32_t,
                                 const uint32_t*, const int64_t*, const uint32_t*, uint64_t*, uint64_t, uint8_t*,
                                 uint32_t, uint32_t*, uint32_t*, uint32_t, uint32_t*, uint64_t, uint32_t, uint32_t*,
                                 void*);
int mgbfs_generic_gather_i64(uint32_t, const int64_t*, uint32_t, const uint32_t*, uint32_t, int64_t*, uint32_t, void*);
int mgbfs_generic_gather_u8(uint32_t, const uint8_t*, uint32_t, const uint32_t*, uint32_t, uint8_t*, uint32_t, void*);
}

namespace mgbfs {

namespace {
constexpr uint64_t kHashSeed = 0x19f856a2ull;  // generic_run.rs:24 seeds the visited hash with this value
constexpr uint32_t kMaxArena = 0x7ffffffeu;
constexpr uint32_t kBatch = 1u << 20;

#define MGBFS_ABI(call)                                                                      \
  do {                                                                                       \
    const int _s = (call);                                                                   \
    if (_s) throw HipError(static_cast<hipError_t>(_s), std::string(#call) + " rejected");   \
  } while (0)
}  // namespace

// Device bytes of a plan (generic_memory.rs:36-41).
uint64_t generic_plan_bytes(uint32_t elements, uint64_t capacity, uint32_t state_bytes, uint64_t* slots_out) {
  uint64_t slots = 1;
  while (slots < 2 * capacity) slots <<= 1;
  if (slots_out) *slots_out = slots;
  return static_cast<uint64_t>(elements) * capacity * state_bytes * 2 + capacity * 8 + slots * 8 + 24;
}

uint32_t generic_state_bytes(const Graph& g) {
  return g.kind == 0 && std::all_of(g.start.begin(), g.start.end(), [](int64_t v) { return v >= 0 && v <= 255; })
             ? 1u
             : 8u;
}

// Largest arena capacity whose plan fits `budget` (generic_native.rs:59-64),
// capped by the state-space bound; 0 when not even one row fits.
uint64_t generic_capacity(const Graph& g, uint64_t budget, uint64_t state_bound) {
  const uint32_t sb = generic_state_bytes(g);
  uint64_t lo = 1, hi = std::min<uint64_t>(kMaxArena, state_bound ? state_bound : kMaxArena);
  if (generic_plan_bytes(g.elements, 1, sb, nullptr) > budget) return 0;
  while (lo < hi) {
    const uint64_t mid = lo + (hi - lo + 1) / 2;
    if (generic_plan_bytes(g.elements, mid, sb, nullptr) <= budget) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

uint32_t generic_batch(uint32_t generators, uint64_t capacity) {
  const uint64_t b = std::max<uint32_t>(1, std::min<uint32_t>(kBatch, 0x7ffffffeu / generators));
  return static_cast<uint32_t>(std::min<uint64_t>(b, capacity));
}

class GenericEngine final : public Engine {
 public:
  GenericEngine(const Graph& g, const mgbfs_engine_options& o, uint64_t state_bound) : g_(g) {
    device_ = o.device;
    backend = MGBFS_BACKEND_GENERIC;
    history_layers = 0;
    MGBFS_HIP(hipSetDevice(device_));
    MGBFS_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    MGBFS_HIP(hipEventCreate(&ev0_));
    MGBFS_HIP(hipEventCreate(&ev1_));
    state_bytes = generic_state_bytes(g);
    elements_ = g.elements;
    uint64_t cap = o.capacity;
    if (!cap) {
      cap = generic_capacity(g, vram_budget(o), state_bound);
      if (!cap) throw std::runtime_error("GENERIC_VRAM_CAPACITY");
    }
    if (cap == 0 || cap > kMaxArena) throw std::runtime_error("GENERIC_ARENA_SHAPE");
    cap_ = static_cast<uint32_t>(cap);
    generic_plan_bytes(elements_, cap_, state_bytes, &slots_);
    visited_.allocate(static_cast<uint64_t>(elements_) * cap_ * state_bytes);
    parents_.allocate(static_cast<uint64_t>(elements_) * cap_ * state_bytes);
    front_.allocate(static_cast<uint64_t>(cap_) * 4);
    future_.allocate(static_cast<uint64_t>(cap_) * 4);
    table_.allocate(slots_ * 8);
    control_.allocate(64);
    perms_.allocate(std::max<size_t>(8, g.perms.size() * 4));
    mats_.allocate(std::max<size_t>(8, g.matrices.size() * 8));
    mods_.allocate(std::max<size_t>(8, g.moduli.size() * 4));
    if (!g.perms.empty())
      MGBFS_HIP(hipMemcpy(perms_.as<void>(), g.perms.data(), g.perms.size() * 4, hipMemcpyHostToDevice));
    if (!g.matrices.empty())
      MGBFS_HIP(hipMemcpy(mats_.as<void>(), g.matrices.data(), g.matrices.size() * 8, hipMemcpyHostToDevice));
    if (!g.moduli.empty())
      MGBFS_HIP(hipMemcpy(mods_.as<void>(), g.moduli.data(), g.moduli.size() * 4, hipMemcpyHostToDevice));
    MGBFS_HIP(hipHostMalloc(reinterpret_cast<void**>(&host_), 64, hipHostMallocDefault));
    hash_bits_ = (o.flags & MGBFS_FLAG_TEST_COLLIDE_HASH) ? 0u : 64u;
    batch_ = generic_batch(g.generators, o.batch ? o.batch : kMaxArena);
    device_bytes = visited_.bytes() + parents_.bytes() + front_.bytes() + future_.bytes() + table_.bytes() +
                   control_.bytes() + perms_.bytes() + mats_.bytes() + mods_.bytes();
    reset();
  }
  ~GenericEngine() override {
    (void)hipSetDevice(device_);
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (host_) (void)hipHostFree(host_);
    if (ev0_) (void)hipEventDestroy(ev0_);
    if (ev1_) (void)hipEventDestroy(ev1_);
    if (stream_) (void)hipStreamDestroy(stream_);
  }

  void reset() override {
    MGBFS_HIP(hipSetDevice(device_));
    MGBFS_HIP(fill_u64(table_.as<uint64_t>(), ~0ull, slots_, stream_));
    MGBFS_HIP(fill_u32(control_.as<uint32_t>(), 0u, 16, stream_));
    MGBFS_HIP(hipStreamSynchronize(stream_));
    // Start state = arena row 0 (SoA), visited_used = 1, front = [0].
    for (uint32_t e = 0; e < elements_; ++e) {
      if (state_bytes == 1) {
        const uint8_t b = static_cast<uint8_t>(g_.start[e]);
        MGBFS_HIP(hipMemcpy(visited_.as<uint8_t>() + static_cast<uint64_t>(e) * cap_, &b, 1, hipMemcpyHostToDevice));
      } else {
        MGBFS_HIP(hipMemcpy(visited_.as<int64_t>() + static_cast<uint64_t>(e) * cap_, &g_.start[e], 8,
                            hipMemcpyHostToDevice));
      }
    }
    const uint32_t one = 1, zero = 0;
    MGBFS_HIP(hipMemcpy(control_.as<uint32_t>(), &one, 4, hipMemcpyHostToDevice));
    MGBFS_HIP(hipMemcpy(front_.as<uint32_t>(), &zero, 4, hipMemcpyHostToDevice));
    if (state_bytes == 1)
      MGBFS_ABI(mgbfs_generic_seed_u8_wide(elements_, visited_.as<uint8_t>(), cap_, 1, table_.as<uint64_t>(), slots_,
                                           kHashSeed, hash_bits_, control_.as<uint32_t>() + 2, stream_));
    else
      MGBFS_ABI(mgbfs_generic_seed_i64_wide(elements_, visited_.as<int64_t>(), cap_, 1, table_.as<uint64_t>(), slots_,
                                            kHashSeed, hash_bits_, control_.as<uint32_t>() + 2, stream_));
    MGBFS_HIP(hipStreamSynchronize(stream_));
    count_ = 1;
    visited_used_ = 1;
    current_start_ = 0;
    prev_start_ = prev_count_ = 0;
    layers.assign(1, 1);
    candidates = unique_candidates = 0;
    kernel_seconds = 0;
    launches = 0;
    streamed_bytes = 0;
    terminal = false;
  }

  int advance(uint64_t* next_count, uint32_t* fatal) override {
    if (terminal) throw std::runtime_error("ENGINE_TERMINAL");
    MGBFS_HIP(hipSetDevice(device_));
    const uint32_t G = g_.generators;
    const uint32_t rows = g_.kind == 0 ? elements_ : g_.rows, cols = g_.kind == 0 ? 1 : g_.cols;
    MGBFS_HIP(hipEventRecord(ev0_, stream_));
    gather(front_.as<uint32_t>(), count_, parents_, cap_);
    MGBFS_HIP(fill_u32(control_.as<uint32_t>() + 1, 0u, 1, stream_));
    const uint32_t batch = batch_;
    for (uint32_t begin = 0; begin < count_; begin += batch) {
      const uint32_t c = std::min(batch, count_ - begin);
      ++launches;
      if (state_bytes == 1)
        MGBFS_ABI(mgbfs_generic_expand_u8_wide(
            g_.kind, elements_, rows, cols, G, parents_.as<uint8_t>() + begin, c, cap_, perms_.as<uint32_t>(),
            mats_.as<int64_t>(), mods_.as<uint32_t>(), table_.as<uint64_t>(), slots_, visited_.as<uint8_t>(), cap_,
            control_.as<uint32_t>(), future_.as<uint32_t>(), cap_, control_.as<uint32_t>() + 1, kHashSeed,
            hash_bits_, control_.as<uint32_t>() + 2, stream_));
      else
        MGBFS_ABI(mgbfs_generic_expand_i64_wide(
            g_.kind, elements_, rows, cols, G, parents_.as<int64_t>() + begin, c, cap_, perms_.as<uint32_t>(),
            mats_.as<int64_t>(), mods_.as<uint32_t>(), table_.as<uint64_t>(), slots_, visited_.as<int64_t>(), cap_,
            control_.as<uint32_t>(), future_.as<uint32_t>(), cap_, control_.as<uint32_t>() + 1, kHashSeed,
            hash_bits_, control_.as<uint32_t>() + 2, stream_));
    }
    MGBFS_HIP(hipEventRecord(ev1_, stream_));
    // One scalar layer-boundary readback: [visited_used, future_count, error].
    MGBFS_HIP(hipMemcpyAsync(host_, control_.as<uint32_t>(), 12, hipMemcpyDeviceToHost, stream_));
    MGBFS_HIP(hipStreamSynchronize(stream_));
    float ms = 0;
    MGBFS_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
    kernel_seconds += ms * 1e-3;
    candidates += static_cast<uint64_t>(count_) * G;
    if (host_[2]) {
      *fatal = host_[2];
      terminal = true;
      return MGBFS_RESOURCE;
    }
    const uint32_t next = host_[1];
    *next_count = next;
    if (next == 0) {
      terminal = true;
      return MGBFS_COMPLETE;
    }
    prev_start_ = current_start_;
    prev_count_ = count_;
    current_start_ = visited_used_;
    visited_used_ = host_[0];
    std::swap(front_, future_);
    count_ = next;
    layers.push_back(next);
    return MGBFS_LAYER;
  }

  void layer_keys(std::vector<uint64_t>&) override { throw std::runtime_error("GENERIC_NO_PACKED_KEYS"); }

  void sample(uint32_t which, uint64_t limit, std::vector<int64_t>& rows, uint64_t& count) override {
    uint64_t c = 0;
    std::vector<uint32_t> idx;
    if (which == 0) {
      c = std::min<uint64_t>(limit, count_);
      gather(front_.as<uint32_t>(), static_cast<uint32_t>(c), parents_, cap_);
    } else {
      if (layers.size() < 2) {
        rows.clear();
        count = 0;
        return;
      }
      // Layer d-1 occupies the contiguous arena rows [prev_start, prev_start + prev_count).
      c = std::min<uint64_t>(limit, prev_count_);
      idx.resize(c);
      for (uint64_t i = 0; i < c; ++i) idx[i] = static_cast<uint32_t>(prev_start_ + i);
      DeviceBuffer tmp;
      tmp.allocate(std::max<uint64_t>(4, c * 4));
      if (c) MGBFS_HIP(hipMemcpy(tmp.as<void>(), idx.data(), c * 4, hipMemcpyHostToDevice));
      gather(tmp.as<uint32_t>(), static_cast<uint32_t>(c), parents_, cap_);
    }
    MGBFS_HIP(hipStreamSynchronize(stream_));
    rows.assign(c * elements_, 0);
    if (state_bytes == 1) {
      std::vector<uint8_t> raw(elements_ * c);
      for (uint32_t e = 0; e < elements_; ++e)
        if (c) MGBFS_HIP(hipMemcpy(raw.data() + e * c, parents_.as<uint8_t>() + static_cast<uint64_t>(e) * cap_, c,
                                   hipMemcpyDeviceToHost));
      for (uint64_t i = 0; i < c; ++i)
        for (uint32_t e = 0; e < elements_; ++e) rows[i * elements_ + e] = raw[e * c + i];
    } else {
      std::vector<int64_t> raw(elements_ * c);
      for (uint32_t e = 0; e < elements_; ++e)
        if (c) MGBFS_HIP(hipMemcpy(raw.data() + e * c, parents_.as<int64_t>() + static_cast<uint64_t>(e) * cap_, c * 8,
                                   hipMemcpyDeviceToHost));
      for (uint64_t i = 0; i < c; ++i)
        for (uint32_t e = 0; e < elements_; ++e) rows[i * elements_ + e] = raw[e * c + i];
    }
    count = c;
  }

  const char* name() const override { return "GENERIC_EXACT_ALL_VISITED"; }

 private:
  void gather(const uint32_t* idx, uint32_t c, DeviceBuffer& out, uint32_t out_stride) {
    if (!c) return;
    if (state_bytes == 1)
      MGBFS_ABI(mgbfs_generic_gather_u8(elements_, visited_.as<uint8_t>(), cap_, idx, c, out.as<uint8_t>(), out_stride,
                                        stream_));
    else
      MGBFS_ABI(mgbfs_generic_gather_i64(elements_, visited_.as<int64_t>(), cap_, idx, c, out.as<int64_t>(),
                                         out_stride, stream_));
  }

  Graph g_;
  uint32_t elements_ = 0, cap_ = 0, hash_bits_ = 64;
  uint64_t slots_ = 0;
  DeviceBuffer visited_, parents_, front_, future_, table_, control_, perms_, mats_, mods_;
  uint32_t count_ = 0, visited_used_ = 0, current_start_ = 0, prev_start_ = 0, prev_count_ = 0, batch_ = 1;
  uint32_t* host_ = nullptr;
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
};

Engine* make_generic_engine(const Graph& g, const mgbfs_engine_options& o, uint64_t state_bound) {
  return new GenericEngine(g, o, state_bound);
}

}  // namespace mgbfs
