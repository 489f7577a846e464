// Host runtime: graph intake, backend selection, the bitmap engine and the
// C ABI of mgbfs_engine_*.
//
// Layer loop contract (crates/mgbfs-runtime/src/generic_native.rs:106-126):
// every advance() enqueues the layer's kernels on the engine stream and reads
// back one small control block (next count + sticky error word) per layer.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>
#include "engine.hpp"
#include "kernels.hpp"

namespace mgbfs {

bool Graph::packable() const {
  if (kind != 0 || elements == 0 || elements > kMaxDegree || generators == 0 || generators > kMaxGenerators)
    return false;
  for (int64_t v : start)
    if (v < 0 || v > 15) return false;
  return true;
}

bool Graph::distinct_perm() const {
  if (kind != 0 || elements > kMaxDegree) return false;
  uint32_t seen = 0;
  for (int64_t v : start) {
    if (v < 0 || v >= static_cast<int64_t>(elements) || ((seen >> v) & 1u)) return false;
    seen |= 1u << v;
  }
  return true;
}

uint64_t Graph::start_key() const {
  uint64_t key = 0;
  for (uint32_t i = 0; i < elements; ++i) key |= (static_cast<uint64_t>(start[i]) & 15u) << (4u * i);
  return key;
}

PackedGenerators Graph::packed_generators() const {
  PackedGenerators pg;
  std::memset(&pg, 0, sizeof(pg));
  pg.count = generators;
  pg.degree = elements;
  pg.tag_shift = 64;
  for (uint32_t g = 0; g < generators; ++g)
    for (uint32_t i = 0; i < elements; ++i) pg.src[g][i] = static_cast<uint8_t>(perms[g * elements + i]);
  fill_shift_groups(pg);
  return pg;
}

std::vector<int> inverse_index(const PackedGenerators& gens) {
  std::vector<int> inv(gens.count, -1);
  for (uint32_t a = 0; a < gens.count; ++a)
    for (uint32_t b = 0; b < gens.count && inv[a] < 0; ++b) {
      bool ok = true;
      for (uint32_t i = 0; i < gens.degree && ok; ++i) ok = gens.src[b][gens.src[a][i]] == i;
      if (ok) inv[a] = static_cast<int>(b);
    }
  return inv;
}

void Engine::fill_info(mgbfs_engine_info& info) const {
  std::memset(&info, 0, sizeof(info));
  info.backend = backend;
  info.history_layers = history_layers;
  info.state_bytes = state_bytes;
  info.depth = layers.empty() ? 0 : static_cast<uint32_t>(layers.size() - 1);
  info.current_count = layers.empty() ? 0 : layers.back();
  for (uint64_t s : layers) {
    info.total_states += s;
    info.peak_layer = std::max(info.peak_layer, s);
  }
  info.device_bytes = device_bytes;
  info.candidates = candidates;
  info.unique_candidates = unique_candidates;
  info.kernel_seconds = kernel_seconds;
  info.kernel_launches = launches;
  info.algorithmic_bytes = streamed_bytes;
  info.retries = retry_count();
}

// Device bytes an engine may use: the caller's budget, else free memory
// minus a reserve of 2 GiB, or 256 MiB under MGBFS_FLAG_MAX_CAPACITY (the
// cold and warm reserves of generic_memory.rs:9-19).
uint64_t vram_budget(const mgbfs_engine_options& o) {
  if (o.vram_budget) return o.vram_budget;
  size_t fr = 0, tot = 0;
  MGBFS_HIP(hipMemGetInfo(&fr, &tot));
  const uint64_t reserve = (o.flags & MGBFS_FLAG_MAX_CAPACITY) ? (256ull << 20) : (2ull << 30);
  return fr > reserve ? fr - reserve : 0;
}

static void unpack_rows(const std::vector<uint64_t>& keys, uint32_t n, std::vector<int64_t>& rows) {
  rows.resize(keys.size() * n);
  for (size_t k = 0; k < keys.size(); ++k)
    for (uint32_t i = 0; i < n; ++i) rows[k * n + i] = static_cast<int64_t>((keys[k] >> (4u * i)) & 15u);
}

// ------------------------------------------------------------ bitmap engine

class BitmapEngine final : public Engine {
 public:
  BitmapEngine(const Graph& g, const mgbfs_engine_options& o) : n_(g.elements) {
    device_ = o.device;
    set_generators(g);
    key_mask_ = n_ >= 16 ? ~0ull : ((1ull << (4 * n_)) - 1);
    backend = MGBFS_BACKEND_BITMAP;
    MGBFS_HIP(hipSetDevice(device_));
    MGBFS_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    MGBFS_HIP(hipEventCreate(&ev0_));
    MGBFS_HIP(hipEventCreate(&ev1_));
    budget_ = vram_budget(o);
    domain_ = factorial_u64(n_);
    const uint64_t words = (domain_ + 31) / 32;
    const uint64_t bitmap_bytes = words * 4;
    if (bitmap_bytes + 64 * 1024 > budget_) throw std::runtime_error("BITMAP_VRAM_CAPACITY");
    bitmap_.allocate(bitmap_bytes);
    control_.allocate(64);
    stash_.allocate(kStashRows * 8);
    marks_.allocate((kAhead + 2) * 8);
    MGBFS_HIP(hipHostMalloc(reinterpret_cast<void**>(&host_ctl_), 64, hipHostMallocDefault));
    MGBFS_HIP(hipHostMalloc(reinterpret_cast<void**>(&host_marks_), (kAhead + 2) * 8, hipHostMallocDefault));
    cap_limit_ = o.capacity ? o.capacity : domain_;
    // Frontier buffers grow on demand up to what the budget admits.
    const uint64_t frontier_budget = (budget_ - bitmap_bytes) / 2 / 8;
    cap_limit_ = std::min(cap_limit_, frontier_budget);
    if (cap_limit_ == 0) throw std::runtime_error("BITMAP_VRAM_CAPACITY");
    grow(front_, front_cap_, std::min<uint64_t>(cap_limit_, 1u << 16));
    grow(next_, next_cap_, std::min<uint64_t>(cap_limit_, 1u << 16));
    reset();
  }

  void reset() override {
    MGBFS_HIP(hipSetDevice(device_));
    MGBFS_HIP(fill_u32(bitmap_.as<uint32_t>(), 0u, (domain_ + 31) / 32, stream_));
    MGBFS_HIP(fill_u64(control_.as<uint64_t>(), 0ull, 8, stream_));
    MGBFS_HIP(bitmap_seed(n_, start_key_, bitmap_.as<uint32_t>(), domain_, front_.as<uint64_t>(),
                          error_word(), stream_));
    read_control();
    if (host_ctl_[1]) throw std::runtime_error("BITMAP_START_DOMAIN");
    count_ = 1;
    counter_base_ = 0;
    prev_count_ = stash_count_ = 0;
    complete_ = false;
    batch_failed_ = false;
    layers.assign(1, 1);
    candidates = unique_candidates = 0;
    kernel_seconds = 0.0;
    launches = 0;
    streamed_bytes = 0;
    terminal = false;
    update_bytes();
  }
  ~BitmapEngine() override {
    (void)hipSetDevice(device_);
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (host_ctl_) (void)hipHostFree(host_ctl_);
    if (host_marks_) (void)hipHostFree(host_marks_);
    if (ev0_) (void)hipEventDestroy(ev0_);
    if (ev1_) (void)hipEventDestroy(ev1_);
    if (stream_) (void)hipStreamDestroy(stream_);
  }

  int advance(uint64_t* next_count, uint32_t* fatal) override {
    if (terminal) throw std::runtime_error("ENGINE_TERMINAL");
    MGBFS_HIP(hipSetDevice(device_));
    const uint64_t remaining = domain_ - total();
    const uint64_t want = std::min(gens_.count * count_, remaining);
    if (want > next_cap_) grow_next(std::min(cap_limit_, std::max(want, next_cap_ * 2)));
    // The layer counter keeps running: this layer's children are counted
    // from counter_base_, its value at the start of the layer.
    // Chunk the frontier only when the next buffer cannot hold the worst case.
    const uint64_t chunk = want <= next_cap_ ? count_ : std::max<uint64_t>(1, next_cap_ / gens_.count);
    // A chunked expansion can stop part way with next_ (which holds the
    // previous layer) partly overwritten: keep its first kStashRows rows,
    // which a sample taken after the failed advance reads.  An unchunked one
    // cannot run out of room.
    stash_count_ = chunk != count_ ? std::min<uint64_t>(prev_count_, kStashRows) : 0;
    if (stash_count_)
      MGBFS_HIP(hipMemcpyAsync(stash_.as<uint64_t>(), next_.as<uint64_t>(), stash_count_ * 8,
                               hipMemcpyDeviceToDevice, stream_));
    MGBFS_HIP(hipEventRecord(ev0_, stream_));
    uint64_t produced = 0;
    for (uint64_t begin = 0; begin < count_; begin += chunk) {
      const uint64_t c = std::min(chunk, count_ - begin);
      if (chunk != count_) {
        read_control();
        produced = host_ctl_[0] - counter_base_;
        if (produced + c * gens_.count > next_cap_) {
          *fatal = kErrFutureCap;
          terminal = true;
          return MGBFS_RESOURCE;
        }
      }
      ++launches;
      MGBFS_HIP(bitmap_expand(front_.as<uint64_t>() + begin, c, gens_, bitmap_.as<uint32_t>(), domain_,
                              next_.as<uint64_t>(), reinterpret_cast<unsigned long long*>(control_.as<uint64_t>()),
                              counter_base_, next_cap_, error_word(), stream_));
    }
    MGBFS_HIP(hipEventRecord(ev1_, stream_));
    read_control();  // the readback queues behind the layer's kernels: one wait for both
    float ms = 0.f;
    MGBFS_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
    kernel_seconds += ms * 1e-3;
    candidates += count_ * gens_.count;
    streamed_bytes += 8 * count_;
    const uint32_t err = static_cast<uint32_t>(host_ctl_[1]);
    if (err) {
      *fatal = err;
      terminal = true;
      return MGBFS_RESOURCE;
    }
    const uint64_t produced_all = host_ctl_[0] - counter_base_;
    counter_base_ = host_ctl_[0];
    *next_count = produced_all;
    if (produced_all == 0) {
      terminal = true;
      complete_ = true;
      return MGBFS_COMPLETE;
    }
    std::swap(front_, next_);
    std::swap(front_cap_, next_cap_);
    prev_count_ = count_;
    count_ = produced_all;
    layers.push_back(count_);
    streamed_bytes += 8 * count_;
    update_bytes();
    return MGBFS_LAYER;
  }

  // Runs to the end with up to kAhead layers queued per host wait, when
  // both frontier buffers can hold every state not yet reached (so no layer
  // can outgrow them, and none is chunked): each layer reads its size and
  // counter base from marks_ (the running counter before and after the
  // layer that produced it) and a one-thread kernel records the counter
  // after it.  Otherwise, and for bigger state spaces, one layer per wait.
  int run_to_end() override {
    if (terminal) return MGBFS_LAYER;
    const uint64_t remaining = domain_ - total();
    if (remaining > cap_limit_ || remaining > kAheadStates) return Engine::run_to_end();
    MGBFS_HIP(hipSetDevice(device_));
    if (next_cap_ < remaining) grow_next(remaining);
    if (front_cap_ < remaining) grow_front(remaining);
    update_bytes();
    stash_count_ = 0;  // no layer here is chunked
    int status = MGBFS_LAYER;
    while (!terminal) {
      host_marks_[0] = counter_base_ - count_;
      host_marks_[1] = counter_base_;
      MGBFS_HIP(hipMemcpyAsync(marks_.as<uint64_t>(), host_marks_, 16, hipMemcpyHostToDevice, stream_));
      MGBFS_HIP(hipEventRecord(ev0_, stream_));
      auto* marks = reinterpret_cast<unsigned long long*>(marks_.as<uint64_t>());
      auto* counter = reinterpret_cast<unsigned long long*>(control_.as<uint64_t>());
      for (uint32_t i = 0; i < kAhead; ++i) {
        const bool even = (i & 1u) == 0;
        uint64_t* src = (even ? front_ : next_).as<uint64_t>();
        uint64_t* dst = (even ? next_ : front_).as<uint64_t>();
        MGBFS_HIP(bitmap_expand_marked(src, marks + i, gens_, bitmap_.as<uint32_t>(), domain_, dst, counter,
                                       remaining, error_word(), kAheadGrid, stream_));
        MGBFS_HIP(bitmap_mark(counter, marks + i + 2, stream_));
      }
      MGBFS_HIP(hipEventRecord(ev1_, stream_));
      MGBFS_HIP(hipMemcpyAsync(host_marks_, marks, (kAhead + 2) * 8, hipMemcpyDeviceToHost, stream_));
      read_control();  // queued behind the layers: one wait for all of them
      float ms = 0.f;
      MGBFS_HIP(hipEventElapsedTime(&ms, ev0_, ev1_));
      kernel_seconds += ms * 1e-3;
      const uint32_t err = static_cast<uint32_t>(host_ctl_[1]);
      if (err) {
        // Later layers of the batch may have overwritten both frontiers:
        // no samples from here on.
        terminal = true;
        batch_failed_ = true;
        return MGBFS_RESOURCE;
      }
      // Layer i + 1 of the batch holds marks[i + 2] - marks[i + 1] states;
      // the batch ends at the first empty one.
      uint32_t produced = 0;
      for (uint32_t i = 0; i < kAhead; ++i) {
        const uint64_t parents = host_marks_[i + 1] - host_marks_[i];
        const uint64_t kids = host_marks_[i + 2] - host_marks_[i + 1];
        ++launches;
        candidates += parents * gens_.count;
        streamed_bytes += 8 * parents + 8 * kids;
        if (kids == 0) {
          terminal = true;
          complete_ = true;
          status = MGBFS_COMPLETE;
          break;
        }
        layers.push_back(kids);
        prev_count_ = parents;
        count_ = kids;
        ++produced;
      }
      // Layer k of the batch sits in front_ when k is even, else in next_.
      if (produced & 1u) {
        std::swap(front_, next_);
        std::swap(front_cap_, next_cap_);
      }
      counter_base_ = host_marks_[produced + 1];
      if (!terminal) status = MGBFS_LAYER;
    }
    return status;
  }

  // Another graph of the same degree reuses the bitmap and frontiers.
  void rebind(const Graph& g) override {
    if (g.elements != n_ || !g.packable() || !g.distinct_perm()) throw std::runtime_error("ENGINE_REBIND_SHAPE");
    set_generators(g);
    reset();
  }

  void layer_keys(std::vector<uint64_t>& out) override {
    if (batch_failed_)
      out.clear();
    else
      copy_keys(front_, count_, count_, out);
  }

  void sample(uint32_t which, uint64_t limit, std::vector<int64_t>& rows, uint64_t& count) override {
    std::vector<uint64_t> keys;
    if (batch_failed_) {
      // Neither frontier holds a known layer.
    } else if (which == 0) {
      copy_keys(front_, count_, limit, keys);
    } else if (terminal && !complete_) {
      // A failed advance wrote children over next_: the stash holds the
      // previous layer's first rows.
      copy_keys(stash_, stash_count_, limit, keys);
    } else if (prev_count_ && layers.size() >= 2) {
      // After a swap the previous layer stays intact in next_ until the next advance.
      copy_keys(next_, prev_count_, limit, keys);
    }
    unpack_rows(keys, n_, rows);
    count = keys.size();
  }
  const char* name() const override { return "PERMUTATION_BITMAP_LEHMER"; }

 private:
  uint32_t* error_word() { return reinterpret_cast<uint32_t*>(control_.as<uint64_t>() + 1); }
  // Generator table, parent-edge tags (nibble 15, when the state uses at
  // most 15 nibbles) and start state of a graph.
  void set_generators(const Graph& g) {
    gens_ = g.packed_generators();
    if (n_ <= 15 && gens_.count <= 14) {
      const std::vector<int> inv = inverse_index(gens_);
      gens_.tag_shift = 60;
      for (uint32_t a = 0; a < gens_.count; ++a) gens_.child_tag[a] = inv[a] >= 0 ? static_cast<uint8_t>(inv[a] + 1) : 0;
    }
    history_layers = g.inverse_closed ? 3 : 0;
    start_key_ = g.start_key();
  }
  uint64_t total() const {
    uint64_t t = 0;
    for (uint64_t s : layers) t += s;
    return t;
  }
  void read_control() {
    MGBFS_HIP(hipMemcpyAsync(host_ctl_, control_.as<uint64_t>(), 16, hipMemcpyDeviceToHost, stream_));
    MGBFS_HIP(hipStreamSynchronize(stream_));
    host_ctl_[1] &= 0xffffffffull;
  }
  void grow(DeviceBuffer& b, uint64_t& cap, uint64_t want) {
    MGBFS_HIP(hipStreamSynchronize(stream_));
    b.allocate(want * 8);
    cap = want;
  }
  // Grows next_ (which holds the previous layer between advances) keeping
  // the first rows that the previous-layer sample can still ask for.
  // Grows front_ keeping the current layer.
  void grow_front(uint64_t want) {
    DeviceBuffer fresh;
    MGBFS_HIP(hipStreamSynchronize(stream_));
    fresh.allocate(want * 8);
    if (count_)
      MGBFS_HIP(hipMemcpyAsync(fresh.as<uint64_t>(), front_.as<uint64_t>(), count_ * 8, hipMemcpyDeviceToDevice, stream_));
    MGBFS_HIP(hipStreamSynchronize(stream_));
    front_ = std::move(fresh);
    front_cap_ = want;
  }
  void grow_next(uint64_t want) {
    DeviceBuffer fresh;
    MGBFS_HIP(hipStreamSynchronize(stream_));
    fresh.allocate(want * 8);
    const uint64_t keep = std::min<uint64_t>(prev_count_, 1000);
    if (keep)
      MGBFS_HIP(hipMemcpyAsync(fresh.as<uint64_t>(), next_.as<uint64_t>(), keep * 8, hipMemcpyDeviceToDevice, stream_));
    MGBFS_HIP(hipStreamSynchronize(stream_));
    next_ = std::move(fresh);
    next_cap_ = want;
  }
  void copy_keys(const DeviceBuffer& b, uint64_t available, uint64_t limit, std::vector<uint64_t>& out) {
    const uint64_t c = std::min(available, limit);
    out.resize(c);
    if (c) {
      MGBFS_HIP(hipMemcpyAsync(out.data(), b.as<uint64_t>(), c * 8, hipMemcpyDeviceToHost, stream_));
      MGBFS_HIP(hipStreamSynchronize(stream_));
      for (auto& k : out) k &= key_mask_;
    }
  }
  void update_bytes() {
    device_bytes = bitmap_.bytes() + front_.bytes() + next_.bytes() + control_.bytes() + stash_.bytes() + marks_.bytes();
  }

  // Rows of the previous layer kept across an advance (the states.json
  // previous-layer sample holds at most 1000 rows).
  static constexpr uint64_t kStashRows = 1000;
  // Layers queued per host wait by run_to_end, the fixed grid of those
  // layers, and the largest unreached state count it runs with (two
  // frontier buffers of that many states: 2 GiB).
  static constexpr uint32_t kAhead = 16;
  static constexpr uint32_t kAheadGrid = 2048;
  static constexpr uint64_t kAheadStates = 1ull << 27;
  PackedGenerators gens_;
  uint32_t n_;
  uint64_t domain_ = 0, budget_ = 0, cap_limit_ = 0;
  DeviceBuffer bitmap_, front_, next_, control_, stash_, marks_;
  uint64_t front_cap_ = 0, next_cap_ = 0, count_ = 0, prev_count_ = 0, stash_count_ = 0, start_key_ = 0,
           key_mask_ = ~0ull;
  uint64_t counter_base_ = 0;  // value of the running child counter at the start of the layer
  bool complete_ = false;
  bool batch_failed_ = false;  // run_to_end stopped on a device error
  uint64_t* host_ctl_ = nullptr;
  uint64_t* host_marks_ = nullptr;
  hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
};

Engine* make_bitmap_engine(const Graph& g, const mgbfs_engine_options& o) { return new BitmapEngine(g, o); }

// ----------------------------------------------------------- graph intake

static Graph intake(const mgbfs_graph_desc* d) {
  if (!d || !d->start || d->elements == 0 || d->generators == 0) throw std::runtime_error("GRAPH_SCHEMA");
  Graph g;
  g.kind = d->kind;
  g.elements = d->elements;
  g.rows = d->rows;
  g.cols = d->cols;
  g.generators = d->generators;
  g.inverse_closed = d->inverse_closed != 0;
  g.expected_states = d->expected_states;
  g.start.assign(d->start, d->start + d->elements);
  if (d->kind == 0) {
    if (!d->permutations) throw std::runtime_error("GRAPH_PERMUTATION_GENERATOR");
    g.perms.assign(d->permutations, d->permutations + static_cast<size_t>(d->generators) * d->elements);
    for (uint32_t k = 0; k < d->generators; ++k) {
      std::vector<uint8_t> used(d->elements, 0);
      for (uint32_t i = 0; i < d->elements; ++i) {
        const uint32_t v = g.perms[k * d->elements + i];
        if (v >= d->elements || used[v]) throw std::runtime_error("GRAPH_PERMUTATION_GENERATOR");
        used[v] = 1;
      }
    }
  } else if (d->kind == 1) {
    if (!d->matrices || !d->moduli || d->rows == 0 || d->cols == 0 ||
        static_cast<uint64_t>(d->rows) * d->cols != d->elements)
      throw std::runtime_error("GRAPH_MATRIX_SHAPE");
    g.matrices.assign(d->matrices, d->matrices + static_cast<size_t>(d->generators) * d->rows * d->rows);
    g.moduli.assign(d->moduli, d->moduli + d->generators);
    for (uint32_t m : g.moduli)
      if (m == 1 || m > (1u << 31)) throw std::runtime_error("GRAPH_MATRIX_MODULUS");
  } else {
    throw std::runtime_error("GRAPH_ACTION_KIND");
  }
  return g;
}

// Bound used for admission only (crates/mgbfs-runtime/src/generic_memory.rs:82-108):
// permutations cannot move a label between position orbits, so the number of
// arrangements of the start labels within each orbit bounds the state count;
// a single nonzero modulus bounds matrices by m^elements (+1 for a
// non-canonical start).  The caller's expected_max_unique_states caps both.
uint64_t state_space_bound(const Graph& g) {
  const uint64_t limit = 1ull << 63;
  uint64_t bound = 0;
  if (g.kind == 0) {
    std::vector<uint32_t> parent(g.elements);
    for (uint32_t i = 0; i < g.elements; ++i) parent[i] = i;
    auto root = [&](uint32_t x) {
      while (parent[x] != x) x = parent[x] = parent[parent[x]];
      return x;
    };
    for (uint32_t k = 0; k < g.generators; ++k)
      for (uint32_t i = 0; i < g.elements; ++i) parent[root(i)] = root(g.perms[k * g.elements + i]);
    // Multinomial per orbit, computed as a product of binomials.
    std::vector<std::vector<int64_t>> labels(g.elements);
    for (uint32_t i = 0; i < g.elements; ++i) labels[root(i)].push_back(g.start[i]);
    long double total = 1;
    for (auto& l : labels) {
      if (l.empty()) continue;
      std::sort(l.begin(), l.end());
      long double ways = 1;
      uint64_t placed = 0;
      for (size_t a = 0; a < l.size();) {
        size_t b = a;
        while (b < l.size() && l[b] == l[a]) ++b;
        for (size_t t = 1; t <= b - a; ++t) ways = ways * static_cast<long double>(placed + t) / static_cast<long double>(t);
        placed += b - a;
        a = b;
      }
      total *= ways;
      if (total > static_cast<long double>(limit)) break;
    }
    bound = total > static_cast<long double>(limit) ? limit : static_cast<uint64_t>(total + 0.5L);
  } else {
    const uint32_t m = g.moduli.empty() ? 0 : g.moduli[0];
    bool same = m != 0;
    for (uint32_t x : g.moduli) same &= x == m;
    if (same) {
      long double total = 1;
      for (uint32_t e = 0; e < g.elements && total <= static_cast<long double>(limit); ++e) total *= m;
      bound = total > static_cast<long double>(limit) ? limit : static_cast<uint64_t>(total);
      for (int64_t v : g.start)
        if (v < 0 || v >= static_cast<int64_t>(m)) {
          bound = bound < limit ? bound + 1 : limit;
          break;
        }
    }
  }
  if (g.expected_states) bound = bound ? std::min(bound, g.expected_states) : g.expected_states;
  return bound;
}

static uint32_t choose_backend(const Graph& g, uint32_t requested) {
  uint32_t backend = requested;
  if (backend == MGBFS_BACKEND_AUTO) {
    // The sorted backend keeps at least 4 tag bits beside its key bits;
    // other inverse-closed graphs keep three rolling banks in segmented
    // tables; other directed graphs keep their full history.
    if (g.packable() && g.inverse_closed && g.elements <= 15) backend = MGBFS_BACKEND_SORTED;
    else if (g.packable() && g.distinct_perm() && g.elements <= 15) backend = MGBFS_BACKEND_BITMAP;
    else if (g.inverse_closed) backend = MGBFS_BACKEND_SEGMENTED;
    else backend = MGBFS_BACKEND_GENERIC;
  }
  switch (backend) {
    case MGBFS_BACKEND_BITMAP:
      if (!g.packable() || !g.distinct_perm()) throw std::runtime_error("BITMAP_UNSUPPORTED_GRAPH");
      return backend;
    case MGBFS_BACKEND_SORTED:
      if (!g.packable() || g.elements > 15) throw std::runtime_error("SORTED_UNSUPPORTED_GRAPH");
      if (!g.inverse_closed) throw std::runtime_error("SORTED_REQUIRES_INVERSE_CLOSED");
      return backend;
    case MGBFS_BACKEND_GENERIC:
    case MGBFS_BACKEND_SEGMENTED:
      return backend;
    default:
      throw std::runtime_error("ENGINE_BACKEND_UNAVAILABLE");
  }
}

// The sorted backend, handing the run to the segmented backend when a
// layer exceeds its limits (its last two layers become the segmented
// backend's banks; the sorted engine and its memory are released).
class HandoverEngine final : public Engine {
 public:
  HandoverEngine(const Graph& g, const mgbfs_engine_options& o) : g_(g), o_(o) {
    inner_.reset(make_sorted_engine(g, o));
    mirror();
  }
  int advance(uint64_t* next_count, uint32_t* fatal) override {
    int s = inner_->advance(next_count, fatal);
    if (s == kAdvanceHandover) {
      PackedHandover h;
      if (!inner_->handover(h)) throw std::runtime_error("ENGINE_HANDOVER");
      const uint64_t cand = inner_->candidates, uniq = inner_->unique_candidates, launches_before = inner_->launches;
      const double secs = inner_->kernel_seconds;
      const uint64_t streamed = inner_->streamed_bytes;
      std::unique_ptr<Engine> next(make_segmented_from_packed(g_, o_, state_space_bound(g_), h));
      inner_ = std::move(next);
      carried_ = {cand, uniq, launches_before, streamed, secs};
      ++handovers_;
      s = inner_->advance(next_count, fatal);
    }
    mirror();
    return s;
  }
  void reset() override {
    if (handovers_) {
      inner_.reset();
      inner_.reset(make_sorted_engine(g_, o_));
      carried_ = {};
    } else {
      inner_->reset();
    }
    mirror();
  }
  void rebind(const Graph& g) override {
    if (handovers_) throw std::runtime_error("ENGINE_REBIND_AFTER_HANDOVER");
    inner_->rebind(g);
    g_ = g;
    mirror();
  }
  void layer_keys(std::vector<uint64_t>& out) override { inner_->layer_keys(out); }
  void sample(uint32_t which, uint64_t limit, std::vector<int64_t>& rows, uint64_t& count) override {
    inner_->sample(which, limit, rows, count);
  }
  const char* name() const override { return inner_->name(); }
  uint64_t retry_count() const override { return inner_->retry_count(); }
  void settle() override {
    inner_->settle();
    mirror();
  }
  void fill_info(mgbfs_engine_info& info) const override {
    Engine::fill_info(info);
    info.retries = inner_->retry_count();
  }
  void fill_phases(double* seconds, uint64_t* bytes, uint32_t n) const override { inner_->fill_phases(seconds, bytes, n); }

 private:
  struct Carried {
    uint64_t candidates = 0, unique = 0, launches = 0, streamed = 0;
    double seconds = 0.0;
  };
  // Counters of the sorted part of a handed-over run are added to the
  // segmented engine's.
  void mirror() {
    layers = inner_->layers;
    candidates = inner_->candidates + carried_.candidates;
    unique_candidates = inner_->unique_candidates + carried_.unique;
    device_bytes = inner_->device_bytes;
    kernel_seconds = inner_->kernel_seconds + carried_.seconds;
    launches = inner_->launches + carried_.launches;
    streamed_bytes = inner_->streamed_bytes + carried_.streamed;
    backend = inner_->backend;
    history_layers = inner_->history_layers;
    state_bytes = inner_->state_bytes;
    terminal = inner_->terminal;
  }
  Graph g_;
  mgbfs_engine_options o_;
  std::unique_ptr<Engine> inner_;
  Carried carried_;
  uint32_t handovers_ = 0;
};

// Planned bytes per state of the widest layer on the sorted backend: three
// banks grown by a quarter, plus per child (G-1 with back-edge masks, else
// G) a sub-bucket slab entry (8 B at the mean fill of 5000 of 5120) and an
// owner-lane slab entry with its slack (10 B).
static uint64_t sorted_bytes_per_state(const Graph& g) {
  const uint32_t G = g.generators;
  const bool masks = G + 2 <= 64 - 4 * g.elements;
  const uint32_t children = masks && G > 1 ? G - 1 : G;
  return 30 + 18ull * children;
}

mgbfs_engine_plan_info plan_engine(const Graph& g, const mgbfs_engine_options& o) {
  mgbfs_engine_plan_info p{};
  p.backend = choose_backend(g, o.backend);
  p.history_layers = g.inverse_closed ? 3 : 1;
  p.state_bound = state_space_bound(g);
  if (!o.vram_budget) MGBFS_HIP(hipSetDevice(o.device));  // a given budget needs no device
  p.budget_bytes = vram_budget(o);
  const uint64_t G = g.generators;
  switch (p.backend) {
    case MGBFS_BACKEND_BITMAP: {
      // BitmapEngine: n!-bit bitmap plus two frontier arrays of capacity rows.
      const uint64_t domain = factorial_u64(g.elements);
      const uint64_t bitmap = (domain + 31) / 32 * 4;
      if (bitmap + 64 * 1024 > p.budget_bytes) throw std::runtime_error("BITMAP_VRAM_CAPACITY");
      p.state_bytes = 8;
      p.retained_banks = 2;
      p.capacity = std::min<uint64_t>(o.capacity ? o.capacity : domain, (p.budget_bytes - bitmap) / 16);
      p.batch = p.capacity;
      p.table_slots = domain;
      p.device_bytes = bitmap + 16 * p.capacity + 8000 + 64;
      break;
    }
    case MGBFS_BACKEND_SORTED: {
      const uint64_t per = sorted_bytes_per_state(g);
      const uint64_t fixed = 5 * ((1ull << 20) + 1) * 4 + (1ull << 20);
      if (fixed + per > p.budget_bytes) throw std::runtime_error("SORTED_VRAM_CAPACITY");
      uint64_t cap = (p.budget_bytes - fixed) / per;
      if (p.state_bound) cap = std::min(cap, p.state_bound);
      cap = std::min<uint64_t>(cap, 0xfffffff0ull / std::max<uint64_t>(1, G - (G > 1 ? 1 : 0)));
      if (o.capacity) cap = std::min(cap, o.capacity);
      p.state_bytes = 8;
      p.retained_banks = G + 2 <= 64 - 4 * g.elements ? 2 : 3;
      p.capacity = cap;
      p.batch = cap;
      p.table_slots = 0;
      p.device_bytes = fixed + per * cap;
      break;
    }
    case MGBFS_BACKEND_SEGMENTED: {
      // Rows of a layer (three banks) or of the whole history, with the
      // segments at load factor 0.5 (16 bytes per row).
      p.state_bytes = generic_state_bytes(g);
      p.retained_banks = g.inverse_closed ? 3 : 0;
      const uint64_t per = static_cast<uint64_t>(g.elements) * p.state_bytes + 16;
      uint64_t cap = p.budget_bytes / (g.inverse_closed ? 3 * per : per);
      if (p.state_bound) cap = std::min(cap, p.state_bound);
      cap = std::min<uint64_t>(cap, 0xfffffffeull);
      if (o.capacity) cap = std::min(cap, o.capacity);
      if (!cap) throw std::runtime_error("SEGMENTED_VRAM_CAPACITY");
      p.capacity = cap;
      p.batch = std::max<uint64_t>(1, std::min<uint64_t>(cap, (1ull << 22) / G));
      p.table_slots = 2 * cap;
      p.device_bytes = (g.inverse_closed ? 3 : 1) * per * cap;
      break;
    }
    default: {
      p.state_bytes = generic_state_bytes(g);
      p.retained_banks = 0;
      p.capacity = o.capacity ? o.capacity : generic_capacity(g, p.budget_bytes, p.state_bound);
      if (!p.capacity) throw std::runtime_error("GENERIC_VRAM_CAPACITY");
      p.batch = generic_batch(g.generators, o.batch ? std::min<uint64_t>(o.batch, p.capacity) : p.capacity);
      p.device_bytes = generic_plan_bytes(g.elements, p.capacity, p.state_bytes, &p.table_slots) +
                       g.perms.size() * 4 + g.matrices.size() * 8 + g.moduli.size() * 4;
      break;
    }
  }
  p.queue_capacity = p.batch * G;
  return p;
}

static Engine* create(const Graph& g, const mgbfs_engine_options& o) {
  switch (choose_backend(g, o.backend)) {
    case MGBFS_BACKEND_BITMAP:
      return make_bitmap_engine(g, o);
    case MGBFS_BACKEND_SORTED:
      return new HandoverEngine(g, o);
    case MGBFS_BACKEND_SEGMENTED:
      return make_segmented_engine(g, o, state_space_bound(g));
    default:
      return make_generic_engine(g, o, state_space_bound(g));
  }
}

}  // namespace mgbfs

using mgbfs::Engine;

static void put_error(char* text, size_t cap, const char* what) {
  if (text && cap) {
    std::strncpy(text, what, cap - 1);
    text[cap - 1] = 0;
  }
}

static thread_local std::string g_last_error;

namespace mgbfs {
void set_last_error(const char* what) { g_last_error = what; }
Graph intake_graph(const mgbfs_graph_desc* d) { return intake(d); }
}  // namespace mgbfs

#define MGBFS_GUARD(...)                                    \
  try {                                                     \
    __VA_ARGS__                                             \
  } catch (const mgbfs::HipError& e) {                      \
    g_last_error = e.what();                                \
    return -static_cast<int>(e.code == 0 ? 1 : e.code);     \
  } catch (const std::exception& e) {                       \
    g_last_error = e.what();                                \
    return -1;                                              \
  }

extern "C" int mgbfs_engine_create(const mgbfs_graph_desc* graph, const mgbfs_engine_options* options,
                                   void** engine, char* error_text, size_t error_capacity) {
  if (!engine || !options) return hipErrorInvalidValue;
  *engine = nullptr;
  try {
    mgbfs::Graph g = mgbfs::intake(graph);
    *engine = mgbfs::create(g, *options);
    return 0;
  } catch (const std::exception& e) {
    put_error(error_text, error_capacity, e.what());
    g_last_error = e.what();
    return -1;
  }
}

extern "C" int mgbfs_engine_plan(const mgbfs_graph_desc* graph, const mgbfs_engine_options* options,
                                 mgbfs_engine_plan_info* plan, char* error_text, size_t error_capacity) {
  if (!plan || !options) return hipErrorInvalidValue;
  int old = 0;
  (void)hipGetDevice(&old);
  try {
    const mgbfs::Graph g = mgbfs::intake(graph);
    *plan = mgbfs::plan_engine(g, *options);
    (void)hipSetDevice(old);
    return 0;
  } catch (const std::exception& e) {
    (void)hipSetDevice(old);
    put_error(error_text, error_capacity, e.what());
    g_last_error = e.what();
    return -1;
  }
}

extern "C" int mgbfs_engine_advance(void* engine, uint64_t* next_count, uint32_t* fatal) {
  if (!engine) return hipErrorInvalidValue;
  uint64_t nc = 0;
  uint32_t f = 0;
  MGBFS_GUARD({
    const int s = static_cast<Engine*>(engine)->advance(&nc, &f);
    if (next_count) *next_count = nc;
    if (fatal) *fatal = f;
    return s;
  })
}

extern "C" int mgbfs_engine_run(void* engine, uint32_t max_layers, double max_seconds) {
  if (!engine) return hipErrorInvalidValue;
  MGBFS_GUARD({
    Engine* e = static_cast<Engine*>(engine);
    if (!max_layers && max_seconds <= 0 && !e->terminal) {
      const int s = e->run_to_end();
      e->settle();
      return s;
    }
    const auto t0 = std::chrono::steady_clock::now();
    int s = MGBFS_LAYER;
    uint32_t done = 0;
    while (!e->terminal) {
      if (max_layers && done >= max_layers) break;
      if (max_seconds > 0 &&
          std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() >= max_seconds)
        break;
      uint64_t nc;
      uint32_t f;
      s = e->advance(&nc, &f);
      ++done;
    }
    e->settle();
    return s;
  })
}

extern "C" int mgbfs_engine_query(void* engine, mgbfs_engine_info* info) {
  if (!engine || !info) return hipErrorInvalidValue;
  MGBFS_GUARD({
    static_cast<Engine*>(engine)->settle();
    static_cast<Engine*>(engine)->fill_info(*info);
    return 0;
  })
}

extern "C" int mgbfs_engine_phases(void* engine, double* seconds, uint64_t* bytes, uint32_t count) {
  if (!engine || !seconds || !bytes) return hipErrorInvalidValue;
  MGBFS_GUARD({
    static_cast<Engine*>(engine)->settle();
    static_cast<Engine*>(engine)->fill_phases(seconds, bytes, count);
    return 0;
  })
}

extern "C" int mgbfs_engine_reset(void* engine) {
  if (!engine) return hipErrorInvalidValue;
  MGBFS_GUARD({
    static_cast<Engine*>(engine)->reset();
    return 0;
  })
}

extern "C" int mgbfs_engine_rebind(void* engine, const mgbfs_graph_desc* graph) {
  if (!engine) return hipErrorInvalidValue;
  MGBFS_GUARD(static_cast<Engine*>(engine)->rebind(mgbfs::intake(graph)); return 0;)
}

extern "C" int mgbfs_engine_layer_sizes(void* engine, uint64_t* sizes, uint32_t capacity, uint32_t* count) {
  if (!engine || !count) return hipErrorInvalidValue;
  const auto& l = static_cast<Engine*>(engine)->layers;
  *count = static_cast<uint32_t>(l.size());
  if (sizes)
    for (uint32_t i = 0; i < capacity && i < l.size(); ++i) sizes[i] = l[i];
  return 0;
}

extern "C" int mgbfs_engine_sample(void* engine, uint32_t which, uint64_t limit, int64_t* rows, uint64_t* count) {
  if (!engine || !count) return hipErrorInvalidValue;
  MGBFS_GUARD({
    std::vector<int64_t> r;
    uint64_t c = 0;
    static_cast<Engine*>(engine)->sample(which, limit, r, c);
    if (rows && !r.empty()) std::memcpy(rows, r.data(), r.size() * sizeof(int64_t));
    *count = c;
    return 0;
  })
}

extern "C" int mgbfs_engine_layer_keys(void* engine, uint64_t* keys, uint64_t capacity, uint64_t* count) {
  if (!engine || !count) return hipErrorInvalidValue;
  MGBFS_GUARD({
    std::vector<uint64_t> k;
    static_cast<Engine*>(engine)->layer_keys(k);
    *count = k.size();
    if (k.size() > capacity) return static_cast<int>(hipErrorInvalidValue);
    if (keys && !k.empty()) std::memcpy(keys, k.data(), k.size() * 8);
    return 0;
  })
}

extern "C" const char* mgbfs_engine_backend_name(void* engine) {
  return engine ? static_cast<Engine*>(engine)->name() : g_last_error.c_str();
}

extern "C" void mgbfs_engine_destroy(void* engine) { delete static_cast<Engine*>(engine); }

// ---------------------------------------------------- kernel-level ABI

extern "C" int mgbfs_packed_bitmap_seed(uint32_t degree, uint64_t start_key, uint32_t* bitmap, uint64_t bitmap_bits,
                                        uint64_t* frontier, uint32_t* error, void* stream) {
  if (degree < 1 || degree > mgbfs::kMaxDegree || !bitmap || !frontier || !error) return hipErrorInvalidValue;
  return mgbfs::bitmap_seed(degree, start_key, bitmap, bitmap_bits, frontier, error,
                            static_cast<hipStream_t>(stream));
}

extern "C" int mgbfs_packed_bitmap_expand(uint32_t degree, uint32_t generators, const uint32_t* permutations,
                                          const uint64_t* frontier, uint64_t count, uint32_t* bitmap,
                                          uint64_t bitmap_bits, uint64_t* next, uint64_t* next_count,
                                          uint64_t next_capacity, uint32_t* error, void* stream) {
  if (degree < 1 || degree > mgbfs::kMaxDegree || generators < 1 || generators > mgbfs::kMaxGenerators ||
      !permutations || !bitmap || !next || !next_count || !error || (count && !frontier))
    return hipErrorInvalidValue;
  mgbfs::PackedGenerators pg;
  std::memset(&pg, 0, sizeof(pg));
  pg.count = generators;
  pg.degree = degree;
  pg.tag_shift = 64;
  for (uint32_t g = 0; g < generators; ++g)
    for (uint32_t i = 0; i < degree; ++i) {
      const uint32_t v = permutations[g * degree + i];
      if (v >= degree) return hipErrorInvalidValue;
      pg.src[g][i] = static_cast<uint8_t>(v);
    }
  mgbfs::fill_shift_groups(pg);
  return mgbfs::bitmap_expand(frontier, count, pg, bitmap, bitmap_bits, next,
                              reinterpret_cast<unsigned long long*>(next_count), 0, next_capacity, error,
                              static_cast<hipStream_t>(stream));
}
