// Host runtime of the segmented open-addressing backend (states of any
// width: u8 rows for permutations with labels <= 255, else int64).
//
// Inverse-closed graphs keep three rolling banks (history_layers = 3,
// crates/mgbfs-runtime/src/generic_distributed_run.rs:40): layer d+1 is
// written into the bank that held layer d-2, whose segments are dropped
// first (the reference's retire_rows / reseed_rows, cuda/generic_rolling.
// cuh:3-47, without tombstones); a child is looked up in the segments of
// layers d and d-1 and claimed in those of layer d+1.  Directed graphs keep
// every layer in bank 0 (full history), its segments growing as it fills.
//
// Segments grow lazily (segmented_distributed_native.rs:520-680): a bank
// starts with one segment sized from the layer estimate; when it reaches
// half full the batch is replayed after a segment twice as large is added.
// A bank with more than two segments after its layer is consolidated into
// one table (the reference's rehash, cuda/segmented_history.cu:314-334), so
// lookups probe few segments.
#include <algorithm>
#include <cstring>
#include <memory>
#include "engine.hpp"
#include "kernels.hpp"
#include "segmented.hpp"

namespace mgbfs {

namespace {

constexpr uint32_t kSegBatchChildren = 1u << 22;  // children per expansion launch
constexpr uint64_t kMinSegmentSlots = 1024;

uint64_t next_pow2(uint64_t x) {
  uint64_t p = 1;
  while (p < x) p <<= 1;
  return p;
}

struct Segment {
  DeviceBuffer slots;
  uint64_t mask = 0;
  uint32_t limit = 0;
  SegSegment view(uint32_t bank) const { return SegSegment{slots.as<uint64_t>(), mask, limit, bank}; }
};

struct Bank {
  DeviceBuffer arena;
  uint64_t capacity = 0;  // rows
  uint32_t used = 0;      // rows published
  std::vector<std::unique_ptr<Segment>> segs;  // the last one is the active segment while writing
};

}  // namespace

template <typename S>
class SegmentedEngine final : public Engine {
 public:
  // o.capacity bounds the rows of a layer (three banks) or of the whole
  // history (directed), as the arena capacity does on the generic backend.
  SegmentedEngine(const Graph& g, const mgbfs_engine_options& o, uint64_t state_bound)
      : g_(g), elements_(g.elements), bound_(state_bound) {
    device_ = o.device;
    backend = MGBFS_BACKEND_SEGMENTED;
    state_bytes = sizeof(S);
    closed_ = g.inverse_closed;
    history_layers = closed_ ? 3 : 0;
    flags_ = o.flags;
    hash_bits_ = (o.flags & MGBFS_FLAG_TEST_COLLIDE_HASH) ? 0u : 64u;
    MGBFS_HIP(hipSetDevice(device_));
    MGBFS_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    MGBFS_HIP(hipEventCreate(&ev0_));
    MGBFS_HIP(hipEventCreate(&ev1_));
    budget_ = vram_budget(o);
    cap_limit_ = o.capacity ? o.capacity : 0xfffffffeull;
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
  }

  ~SegmentedEngine() override {
    (void)hipSetDevice(device_);
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (host_) (void)hipHostFree(host_);
    if (ev0_) (void)hipEventDestroy(ev0_);
    if (ev1_) (void)hipEventDestroy(ev1_);
    if (stream_) (void)hipStreamDestroy(stream_);
  }

  // Start state in row 0 of bank 0.
  void reset() override {
    MGBFS_HIP(hipSetDevice(device_));
    for (auto& b : banks_) {
      b.segs.clear();
      b.used = 0;
    }
    ensure_arena(0, closed_ ? 1 : 1024);
    std::vector<S> col(1);
    for (uint32_t e = 0; e < elements_; ++e) {
      col[0] = static_cast<S>(g_.start[e]);
      MGBFS_HIP(hipMemcpy(banks_[0].arena.as<S>() + static_cast<uint64_t>(e) * banks_[0].capacity, col.data(),
                          sizeof(S), hipMemcpyHostToDevice));
    }
    banks_[0].used = 1;
    add_segment(0, 2);
    seed_rows(0, 0, 1);
    cur_ = 0;
    prev_ = -1;
    layer_start_ = 0;
    prev_start_ = prev_count_ = 0;
    count_ = 1;
    layers.assign(1, 1);
    reset_counters();
  }

  // Continues a BFS whose layers d-1 (prev) and d (cur) are nibble-packed
  // states in sorted-backend banks (the sorted backend's hand-over when a
  // layer exceeds its limits); `table` is the layer table so far.
  void adopt_packed(const uint64_t* prev_vals, uint64_t prev_count, const uint64_t* cur_vals, uint64_t cur_count,
                    const KeyMix& km, uint32_t tagbits, const std::vector<uint64_t>& table) {
    static_assert(sizeof(S) == 1 || sizeof(S) == 8, "state width");
    if (!closed_ || sizeof(S) != 1) throw std::runtime_error("SEGMENTED_ADOPT_SHAPE");
    MGBFS_HIP(hipSetDevice(device_));
    for (auto& b : banks_) {
      b.segs.clear();
      b.used = 0;
    }
    const int pb = prev_count ? 0 : -1, cb = 1;
    if (prev_count) adopt_bank(0, prev_vals, prev_count, km, tagbits);
    adopt_bank(1, cur_vals, cur_count, km, tagbits);
    prev_ = pb;
    cur_ = cb;
    count_ = cur_count;
    layers = table;
    reset_counters();
  }

  int advance(uint64_t* next_count, uint32_t* fatal) override {
    if (terminal) throw std::runtime_error("ENGINE_TERMINAL");
    MGBFS_HIP(hipSetDevice(device_));
    MGBFS_HIP(hipEventRecord(ev0_, stream_));
    const uint64_t F = count_;
    const uint64_t G = g_.generators;
    int target;
    if (closed_) {
      target = next_bank();
      Bank& nb = banks_[target];
      nb.segs.clear();  // retire layer d-2
      nb.used = 0;
      const uint64_t most = std::min<uint64_t>({G * F, bound_ ? bound_ : G * F, cap_limit_});
      if (!ensure_arena(target, most)) return stop(fatal, kErrVisitedCap);
      const uint64_t guess = (flags_ & MGBFS_FLAG_TEST_TINY_SEGMENTS) ? 16 : std::min<uint64_t>(most, 2 * F + 64);
      if (!add_segment(target, guess)) return stop(fatal, kErrVisitedCap);
    } else {
      // Full history: room for every child of this layer beside the rows
      // kept, before any claim (a claim never runs out of rows).
      target = 0;
      const uint64_t most = std::min<uint64_t>(banks_[0].used + G * F, bound_ ? bound_ : ~0ull);
      if (!grow_history(std::min<uint64_t>(most, cap_limit_))) return stop(fatal, kErrVisitedCap);
      if (banks_[0].segs.empty() && !add_segment(0, 2 * F + 64)) return stop(fatal, kErrVisitedCap);
    }
    const uint64_t per_batch = std::max<uint64_t>(1, kSegBatchChildren / G);
    for (uint64_t begin = 0; begin < F; begin += per_batch) {
      const uint32_t count = static_cast<uint32_t>(std::min<uint64_t>(per_batch, F - begin));
      while (true) {
        set_control(banks_[target].used);
        MGBFS_HIP(launch(target, begin, count));
        const uint32_t err = read_control(target);
        if (err & kErrSegmentFull) {
          // Seal the active segment and add one twice as large; replay.
          ++spills_;
          const uint64_t grow = 2 * (banks_[target].segs.back()->mask + 1);
          if (total_segments() >= kSegMaxProbe || !add_segment(target, grow / 2)) return stop(fatal, kErrTableFull);
          continue;
        }
        if (err) return stop(fatal, err);
        break;
      }
    }
    candidates += F * G;
    const uint64_t N = closed_ ? banks_[target].used : banks_[0].used - (layer_start_ + F);
    // A bank left with many segments is rebuilt as one table.
    if (banks_[target].segs.size() > 2 && !consolidate(target)) return stop(fatal, kErrVisitedCap);
    MGBFS_HIP(hipEventRecord(ev1_, stream_));
    MGBFS_HIP(hipEventSynchronize(ev1_));
    float ms = 0.f;
    MGBFS_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
    kernel_seconds += ms * 1e-3;
    launches += 1;
    streamed_bytes += (F + N) * elements_ * sizeof(S);
    unique_candidates += N;
    if (N == 0) {
      *next_count = 0;
      terminal = true;
      return MGBFS_COMPLETE;
    }
    if (N > cap_limit_) return stop(fatal, kErrFutureCap);
    if (closed_) {
      prev_ = cur_;
      cur_ = target;
    } else {
      prev_start_ = layer_start_;
      prev_count_ = F;
      layer_start_ += F;
    }
    count_ = N;
    layers.push_back(N);
    *next_count = N;
    return MGBFS_LAYER;
  }

  void layer_keys(std::vector<uint64_t>& out) override {
    std::vector<int64_t> rows;
    read_rows(closed_ ? cur_ : 0, closed_ ? 0 : layer_start_, count_, rows);
    out.resize(count_);
    for (uint64_t r = 0; r < count_; ++r) {
      uint64_t k = 0;
      for (uint32_t e = 0; e < elements_ && e < 16; ++e) k |= (static_cast<uint64_t>(rows[r * elements_ + e]) & 15u) << (4 * e);
      out[r] = k;
    }
  }

  void sample(uint32_t which, uint64_t limit, std::vector<int64_t>& rows, uint64_t& count) override {
    rows.clear();
    count = 0;
    if (which == 0) {
      count = std::min(limit, count_);
      read_rows(closed_ ? cur_ : 0, closed_ ? 0 : layer_start_, count, rows);
    } else if (layers.size() >= 2) {
      if (closed_ && prev_ >= 0) {
        count = std::min<uint64_t>(limit, banks_[prev_].used);
        read_rows(prev_, 0, count, rows);
      } else if (!closed_) {
        count = std::min<uint64_t>(limit, prev_count_);
        read_rows(0, prev_start_, count, rows);
      }
    }
  }

  const char* name() const override { return "SEGMENTED_OPEN_ADDRESSING_THREE_BANK"; }
  uint64_t retry_count() const override { return spills_; }

 private:
  // Device control words [rows used in the bank written, claims in its
  // active segment, error, 0], uploaded before and read back after each
  // launch (the engine waits on every batch).
  void set_control(uint32_t rows) {
    host_[8] = rows;
    host_[9] = active_fill_;
    host_[10] = host_[11] = 0;
    MGBFS_HIP(hipMemcpyAsync(control_.as<uint32_t>(), host_ + 8, 16, hipMemcpyHostToDevice, stream_));
  }
  uint32_t read_control(int bank) {
    MGBFS_HIP(hipMemcpyAsync(host_, control_.as<uint32_t>(), 16, hipMemcpyDeviceToHost, stream_));
    MGBFS_HIP(hipStreamSynchronize(stream_));
    banks_[bank].used = std::min<uint32_t>(host_[0], static_cast<uint32_t>(banks_[bank].capacity));
    active_fill_ = host_[1];
    return host_[2];
  }
  int stop(uint32_t* fatal, uint32_t bits) {
    *fatal = bits;
    terminal = true;
    return MGBFS_RESOURCE;
  }
  void reset_counters() {
    candidates = unique_candidates = 0;
    kernel_seconds = 0.0;
    launches = 0;
    streamed_bytes = 0;
    spills_ = 0;
    terminal = false;
    update_bytes();
  }
  int next_bank() const {
    for (int b = 0; b < 3; ++b)
      if (b != cur_ && b != prev_) return b;
    return -1;
  }
  uint32_t total_segments() const {
    uint32_t n = 0;
    for (const auto& b : banks_) n += static_cast<uint32_t>(b.segs.size());
    return n;
  }
  // Bank 0 of a directed run with room for `rows` rows, its published rows
  // kept (each element plane copied into the larger arena).
  bool grow_history(uint64_t rows) {
    Bank& bank = banks_[0];
    if (bank.capacity >= rows) return true;
    if (rows > 0xfffffffeull) return false;
    uint64_t want = std::min<uint64_t>(std::max<uint64_t>(rows, 2 * bank.capacity), 0xfffffffeull);
    if (!fits(want * elements_ * sizeof(S))) want = rows;
    if (!fits(want * elements_ * sizeof(S))) return false;
    DeviceBuffer fresh;
    fresh.allocate(want * elements_ * sizeof(S));
    MGBFS_HIP(hipMemcpy2DAsync(fresh.as<S>(), want * sizeof(S), bank.arena.template as<S>(), bank.capacity * sizeof(S),
                               static_cast<size_t>(bank.used) * sizeof(S), elements_, hipMemcpyDeviceToDevice, stream_));
    MGBFS_HIP(hipStreamSynchronize(stream_));
    bank.arena = std::move(fresh);
    bank.capacity = want;
    update_bytes();
    return true;
  }
  bool fits(uint64_t extra) const { return device_bytes + extra <= budget_; }
  // Arena of bank b with room for `rows` rows (contents are not kept).
  bool ensure_arena(int b, uint64_t rows) {
    rows = std::max<uint64_t>(rows, 1);
    if (rows > 0xfffffffeull) return false;
    Bank& bank = banks_[b];
    if (bank.capacity >= rows) return true;
    uint64_t want = std::max<uint64_t>(rows, bank.capacity + bank.capacity / 4);
    want = std::min<uint64_t>(want, 0xfffffffeull);
    const uint64_t extra = (want - bank.capacity) * elements_ * sizeof(S);
    if (!fits(extra)) {
      want = rows;
      if (!fits((want - bank.capacity) * elements_ * sizeof(S))) return false;
    }
    MGBFS_HIP(hipStreamSynchronize(stream_));
    bank.arena.release();
    bank.arena.allocate(want * elements_ * sizeof(S));
    bank.capacity = want;
    update_bytes();
    return true;
  }
  // New active segment for bank b holding `states` at load factor <= 0.5.
  bool add_segment(int b, uint64_t states) {
    const uint64_t slots = std::max<uint64_t>(kMinSegmentSlots, next_pow2(2 * std::max<uint64_t>(states, 1)));
    if (!fits(slots * 8)) return false;
    auto seg = std::make_unique<Segment>();
    seg->slots.allocate(slots * 8);
    seg->mask = slots - 1;
    seg->limit = static_cast<uint32_t>(std::min<uint64_t>(slots / 2, 0xffffffffull));
    MGBFS_HIP(fill_u64(seg->slots.as<uint64_t>(), ~0ull, slots, stream_));
    banks_[b].segs.push_back(std::move(seg));
    active_fill_ = 0;
    update_bytes();
    return true;
  }
  SegBanks<S> bank_views() const {
    SegBanks<S> v{};
    for (uint32_t b = 0; b < kSegBanks; ++b) {
      v.arena[b] = banks_[b].arena.template as<S>();
      v.stride[b] = banks_[b].capacity;
      v.rows[b] = banks_[b].used;
      v.capacity[b] = static_cast<uint32_t>(banks_[b].capacity);
    }
    return v;
  }
  // Inserts rows [first, first + count) of bank b into its active segment.
  void seed_rows(int b, uint32_t first, uint32_t count) {
    set_control(banks_[b].used);
    MGBFS_HIP(segmented_seed<S>(bank_views(), static_cast<uint32_t>(b), first, count, elements_,
                                banks_[b].segs.back()->view(b), hash_bits_, control_.as<uint32_t>() + 1,
                                control_.as<uint32_t>() + 2, stream_));
    if (read_control(b)) throw std::runtime_error("SEGMENTED_SEED");
  }
  // Rebuilds bank b's segments as one table of all its rows.
  bool consolidate(int b) {
    Bank& bank = banks_[b];
    std::vector<std::unique_ptr<Segment>> old;
    old.swap(bank.segs);
    if (!add_segment(b, bank.used)) {
      bank.segs.swap(old);
      return true;  // keep probing the segments as they are
    }
    old.clear();
    update_bytes();
    seed_rows(b, 0, bank.used);
    return true;
  }
  void adopt_bank(int b, const uint64_t* vals, uint64_t count, const KeyMix& km, uint32_t tagbits) {
    if (!ensure_arena(b, count)) throw std::runtime_error("SEGMENTED_ADOPT_CAPACITY");
    MGBFS_HIP(segmented_adopt_packed(vals, count, km, tagbits, elements_,
                                     reinterpret_cast<uint8_t*>(banks_[b].arena.template as<S>()), banks_[b].capacity,
                                     stream_));
    banks_[b].used = static_cast<uint32_t>(count);
    if (!add_segment(b, count)) throw std::runtime_error("SEGMENTED_ADOPT_CAPACITY");
    seed_rows(b, 0, static_cast<uint32_t>(count));
  }
  hipError_t launch(int target, uint64_t begin, uint32_t count) {
    SegBatch<S> bt{};
    bt.kind = g_.kind;
    bt.elements = elements_;
    bt.rows = g_.kind == 0 ? elements_ : g_.rows;
    bt.cols = g_.kind == 0 ? 1 : g_.cols;
    bt.generators = g_.generators;
    bt.count = count;
    const int pb = closed_ ? cur_ : 0;
    bt.parent_stride = banks_[pb].capacity;
    bt.first_parent = (closed_ ? 0 : layer_start_) + begin;
    bt.parents = banks_[pb].arena.template as<S>();
    bt.perms = perms_.as<uint32_t>();
    bt.mats = mats_.as<int64_t>();
    bt.moduli = mods_.as<uint32_t>();
    bt.banks = bank_views();
    // Sealed segments: layers d and d-1 (rolling) or all history but the
    // active segment, plus the target bank's sealed segments.
    bt.probe.count = 0;
    auto add = [&](int b, size_t upto) {
      for (size_t s = 0; s < upto && bt.probe.count < kSegMaxProbe; ++s)
        bt.probe.seg[bt.probe.count++] = banks_[b].segs[s]->view(b);
    };
    if (closed_) {
      add(cur_, banks_[cur_].segs.size());
      if (prev_ >= 0) add(prev_, banks_[prev_].segs.size());
    }
    add(target, banks_[target].segs.size() - 1);
    bt.active = banks_[target].segs.back()->view(target);
    bt.target = static_cast<uint32_t>(target);
    bt.hash_bits = hash_bits_;
    bt.control = control_.as<uint32_t>();
    bt.error = control_.as<uint32_t>() + 2;
    return segmented_expand<S>(bt, stream_);
  }
  void read_rows(int b, uint64_t first, uint64_t count, std::vector<int64_t>& out) {
    out.assign(count * elements_, 0);
    if (!count || b < 0) return;
    std::vector<S> col(count);
    for (uint32_t e = 0; e < elements_; ++e) {
      MGBFS_HIP(hipMemcpy(col.data(), banks_[b].arena.template as<S>() + static_cast<uint64_t>(e) * banks_[b].capacity + first,
                          count * sizeof(S), hipMemcpyDeviceToHost));
      for (uint64_t r = 0; r < count; ++r) out[r * elements_ + e] = static_cast<int64_t>(col[r]);
    }
  }
  void update_bytes() {
    device_bytes = control_.bytes() + perms_.bytes() + mats_.bytes() + mods_.bytes();
    for (const auto& b : banks_) {
      device_bytes += b.arena.bytes();
      for (const auto& s : b.segs) device_bytes += s->slots.bytes();
    }
  }

  Graph g_;
  uint32_t elements_;
  uint64_t bound_, budget_ = 0, cap_limit_ = 0;
  bool closed_ = false;
  uint32_t flags_ = 0, hash_bits_ = 64;
  Bank banks_[kSegBanks];
  int cur_ = 0, prev_ = -1;
  uint64_t count_ = 0, layer_start_ = 0, prev_start_ = 0, prev_count_ = 0, spills_ = 0;
  uint32_t active_fill_ = 0;  // claims in the newest segment (persists across batches)
  DeviceBuffer control_, perms_, mats_, mods_;
  uint32_t* host_ = nullptr;
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
};

Engine* make_segmented_engine(const Graph& g, const mgbfs_engine_options& o, uint64_t state_bound) {
  const bool narrow = g.kind == 0 && std::all_of(g.start.begin(), g.start.end(), [](int64_t v) { return v >= 0 && v <= 255; });
  Engine* e = narrow ? static_cast<Engine*>(new SegmentedEngine<uint8_t>(g, o, state_bound))
                     : static_cast<Engine*>(new SegmentedEngine<int64_t>(g, o, state_bound));
  try {
    e->reset();
  } catch (...) {
    delete e;
    throw;
  }
  return e;
}

Engine* make_segmented_from_packed(const Graph& g, const mgbfs_engine_options& o, uint64_t state_bound,
                                   const PackedHandover& h) {
  auto* e = new SegmentedEngine<uint8_t>(g, o, state_bound);
  try {
    e->adopt_packed(h.prev, h.prev_count, h.cur, h.cur_count, h.km, h.tagbits, h.layers);
  } catch (...) {
    delete e;
    throw;
  }
  return e;
}

}  // namespace mgbfs
