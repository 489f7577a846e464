// Host side of the batched multi-instance engine: slot admission with VRAM
// planning, a pooled device allocator, and a hipGraph-captured layer loop
// (reset flag, then K x [tiles, clear, expand, finish]) replayed until no instance
// is active.  Instances join and leave at layer boundaries, so the same
// object serves both the fixed sweep and continuous-batching serving.
#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <vector>
#include "arena.hpp"
#include "batch.hpp"
#include "engine.hpp"
#include "kernels.hpp"

namespace mgbfs {

namespace {

constexpr uint32_t kLayersPerGraph = 4;
constexpr uint32_t kDefaultMaxLayers = 4096;

uint64_t next_pow2(uint64_t x) {
  uint64_t p = 1;
  while (p < x) p <<= 1;
  return p;
}

// Device memory for admitted jobs: a few large chunks, sub-allocated by
// best-fit ranges, so admission costs no hipMalloc once the arena has grown
// to the working set.  Chunks grow geometrically (kArenaFirst doubling to
// kArenaMax) within the engine's budget; a chunk that is entirely free is
// handed back to the device only when the arena holds more than the budget.
constexpr uint64_t kArenaFirst = 1ull << 30;
constexpr uint64_t kArenaMax = 32ull << 30;

struct Block {
  char* ptr = nullptr;
  uint32_t chunk = 0;
  uint64_t offset = 0, bytes = 0;
};

class DeviceArena {
 public:
  Block take(uint64_t bytes, uint64_t budget) {
    bytes = arena_round(bytes);
    for (uint32_t c = 0; c < chunks_.size(); ++c) {
      uint64_t off = 0;
      if (chunks_[c].buf.bytes() && chunks_[c].ranges.take(bytes, &off))
        return Block{chunks_[c].buf.as<char>() + off, c, off, bytes};
    }
    uint64_t size = std::max(bytes, std::min(next_, kArenaMax));
    if (reserved_ + size > budget && reserved_ + bytes <= budget) size = budget - reserved_;
    size = std::max(bytes, size / kArenaAlign * kArenaAlign);
    uint32_t c = 0;
    while (c < chunks_.size() && chunks_[c].buf.bytes()) ++c;
    if (c == chunks_.size()) chunks_.emplace_back();
    chunks_[c].buf.allocate(size);
    chunks_[c].ranges = RangeSet(size);
    reserved_ += size;
    next_ = std::min(2 * next_, kArenaMax);
    uint64_t off = 0;
    chunks_[c].ranges.take(bytes, &off);
    return Block{chunks_[c].buf.as<char>() + off, c, off, bytes};
  }
  void give(Block& b, uint64_t budget) {
    if (!b.ptr) return;
    Chunk& ch = chunks_[b.chunk];
    ch.ranges.give(b.offset, b.bytes);
    if (ch.ranges.empty() && reserved_ > budget) {
      reserved_ -= ch.buf.bytes();
      ch.buf.release();
      ch.ranges = RangeSet();
    }
    b = Block{};
  }
  void trim() {
    for (Chunk& ch : chunks_)
      if (ch.buf.bytes() && ch.ranges.empty()) {
        reserved_ -= ch.buf.bytes();
        ch.buf.release();
        ch.ranges = RangeSet();
      }
  }
  uint64_t reserved() const { return reserved_; }

 private:
  struct Chunk {
    DeviceBuffer buf;
    RangeSet ranges;
  };
  std::vector<Chunk> chunks_;
  uint64_t reserved_ = 0, next_ = kArenaFirst;
};

struct Slot {
  bool used = false;
  bool reported = false;
  uint64_t planned = 0;
  Block block;  // front, next, layer sizes, visited set (256-byte aligned parts)
};

}  // namespace

struct BatchPlan {
  uint32_t visited = kVisBitmap;
  bool rolling = false;  // three rolling hash tables, layers of at most `cap` states
  uint64_t bound = 0, cap = 0, table_slots = 0, visited_bytes = 0, bytes = 0;
  uint32_t max_layers = kDefaultMaxLayers;
};

class BatchEngine {
 public:
  BatchEngine(const mgbfs_batch_options& o) : device_(o.device) {
    MGBFS_HIP(hipSetDevice(device_));
    MGBFS_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    slots_n_ = o.max_slots ? std::min(o.max_slots, kBatchMaxSlots) : 256;
    budget_ = o.vram_budget;
    if (!budget_) {
      size_t fr = 0, tot = 0;
      MGBFS_HIP(hipMemGetInfo(&fr, &tot));
      budget_ = fr > (4ull << 30) ? fr - (4ull << 30) : 0;
    }
    hipDeviceProp_t prop;
    MGBFS_HIP(hipGetDeviceProperties(&prop, device_));
    grid_ = o.grid ? o.grid : static_cast<uint32_t>(prop.multiProcessorCount) * 8;
    clear_grid_ = static_cast<uint32_t>(prop.multiProcessorCount);
    layers_per_graph_ = o.layers_per_graph ? o.layers_per_graph : kLayersPerGraph;
    threads_ = o.threads ? o.threads : 256;
    if (threads_ != 128 && threads_ != 256 && threads_ != 512) throw std::runtime_error("BATCH_OPTION_THREADS");
    hash_load_pct_ = o.hash_load_pct ? o.hash_load_pct : 50;
    if (hash_load_pct_ < 10 || hash_load_pct_ > 90) throw std::runtime_error("BATCH_OPTION_HASH_LOAD");
    if (o.history > 2) throw std::runtime_error("BATCH_OPTION_HISTORY");
    history_ = o.history;
    layer_pct_ = o.layer_pct ? o.layer_pct : 30;
    if (layer_pct_ > 100) throw std::runtime_error("BATCH_OPTION_LAYER_PCT");
    bitmap_bias_permille_ = o.bitmap_bias_permille ? o.bitmap_bias_permille : 1000;
    inst_.allocate(sizeof(BatchInstance) * slots_n_);
    prefix_.allocate(8 * (slots_n_ + 1));
    flag_.allocate(64);
    MGBFS_HIP(hipHostMalloc(reinterpret_cast<void**>(&host_flag_), 64, hipHostMallocDefault));
    MGBFS_HIP(hipHostMalloc(reinterpret_cast<void**>(&host_mirror_), sizeof(BatchInstance) * slots_n_,
                            hipHostMallocDefault));
    finished_.resize(slots_n_);
    jobs_.resize(slots_n_);
    pending_.assign(slots_n_, kNothing);
    pending_key_.assign(slots_n_, 0);
    const size_t stage_bytes = (sizeof(BatchInstance) + sizeof(BatchInstall)) * slots_n_;
    MGBFS_HIP(hipHostMalloc(reinterpret_cast<void**>(&host_stage_), stage_bytes, hipHostMallocDefault));
    stage_.allocate(stage_bytes);
    host_inst_.assign(slots_n_, BatchInstance{});
    std::memset(host_inst_.data(), 0, sizeof(BatchInstance) * slots_n_);
    MGBFS_HIP(hipMemcpy(inst_.as<void>(), host_inst_.data(), sizeof(BatchInstance) * slots_n_, hipMemcpyHostToDevice));
    slots_.resize(slots_n_);
    capture();
  }
  ~BatchEngine() {
    (void)hipSetDevice(device_);
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (exec_) (void)hipGraphExecDestroy(exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    if (host_flag_) (void)hipHostFree(host_flag_);
    if (host_mirror_) (void)hipHostFree(host_mirror_);
    if (host_stage_) (void)hipHostFree(host_stage_);
    if (host_req_) (void)hipHostFree(host_req_);
    if (host_out_) (void)hipHostFree(host_out_);
    slots_.clear();
    arena_.trim();
    if (stream_) (void)hipStreamDestroy(stream_);
  }

  // Planned footprint (the admission unit) and the visited-set choice.
  // Hash instances of inverse-closed graphs may keep three rolling tables
  // (history_layers 3) sized, like their frontiers, for layers of up to
  // layer_pct % of the state bound, times `scale` after restarts; otherwise
  // one table holds the whole orbit.  History policy 2 always rolls them,
  // 1 never does, and 0 rolls them only when the full-history footprint
  // does not fit the free budget (three extra lookups per child buy the
  // memory).  Tables hold at most hash_load_pct % of their slots.  The
  // automatic visited choice takes the Lehmer bitmap when its bytes are at
  // most bitmap_bias_permille / 1000 of the hash tables'.
  BatchPlan plan(const Graph& g, uint32_t visited, uint32_t max_layers, uint32_t scale = 1) const {
    if (history_ == 0 && g.inverse_closed) {
      const BatchPlan full = plan_with(g, visited, max_layers, false, 1);
      if (full.visited == kVisBitmap || used_ + full.bytes <= budget_) return full;
      return plan_with(g, visited, max_layers, true, scale);
    }
    return plan_with(g, visited, max_layers, history_ == 2 && g.inverse_closed, scale);
  }

  BatchPlan plan_with(const Graph& g, uint32_t visited, uint32_t max_layers, bool rolling, uint32_t scale) const {
    BatchPlan p;
    p.max_layers = max_layers ? max_layers : kDefaultMaxLayers;
    if (g.kind == 0) {
      if (!g.packable() || g.elements > 15) throw std::runtime_error("BATCH_UNSUPPORTED_GRAPH");
    } else {
      if (g.elements > kBatchMaxEntries || g.rows > 3 || g.generators > kBatchMaxGens)
        throw std::runtime_error("BATCH_UNSUPPORTED_GRAPH");
      const uint32_t m = g.moduli[0];
      for (uint32_t x : g.moduli)
        if (x != m || m < 2) throw std::runtime_error("BATCH_UNSUPPORTED_GRAPH");
      for (int64_t v : g.start)
        if (v < 0 || v >= static_cast<int64_t>(m)) throw std::runtime_error("BATCH_UNSUPPORTED_GRAPH");
      for (int64_t v : g.matrices)
        if (v > INT32_MAX || v < INT32_MIN) throw std::runtime_error("BATCH_UNSUPPORTED_GRAPH");
      uint32_t eb = 1;
      while ((1ull << eb) < m) ++eb;
      if (eb * g.elements > 63) throw std::runtime_error("BATCH_UNSUPPORTED_GRAPH");
    }
    if (g.generators > kBatchMaxGens) throw std::runtime_error("BATCH_UNSUPPORTED_GRAPH");
    p.bound = state_space_bound(g);
    if (!p.bound || p.bound > (1ull << 33)) throw std::runtime_error("BATCH_STATE_BOUND");
    const bool bitmap_ok = g.kind == 0 && g.distinct_perm() && factorial_u64(g.elements) <= (1ull << 34);
    const uint64_t bitmap_bytes = bitmap_ok ? (factorial_u64(g.elements) + 31) / 32 * 4 : ~0ull;
    uint64_t cap = p.bound;
    if (rolling) {
      const uint64_t want = (p.bound * layer_pct_ + 99) / 100 * scale;
      cap = std::min(p.bound, std::max<uint64_t>(1024, want));
    }
    const uint64_t slots = std::max<uint64_t>(64, (cap * 100 + hash_load_pct_ - 1) / hash_load_pct_);
    const uint64_t hash_bytes = (rolling ? 3 : 1) * slots * 8;
    if (visited == 0)
      p.visited = bitmap_ok && static_cast<double>(bitmap_bytes) * 1000.0 <=
                                   static_cast<double>(hash_bytes) * bitmap_bias_permille_
                      ? kVisBitmap
                      : kVisHash;
    else
      p.visited = visited == 1 ? kVisBitmap : kVisHash;
    if (p.visited == kVisBitmap && !bitmap_ok) throw std::runtime_error("BATCH_BITMAP_UNSUPPORTED");
    if (p.visited == kVisBitmap) {
      p.visited_bytes = bitmap_bytes;
      p.cap = p.bound;
    } else {
      p.rolling = rolling;
      p.table_slots = slots;
      p.visited_bytes = hash_bytes;
      p.cap = cap;
    }
    p.bytes = 2 * p.cap * 8 + p.visited_bytes + 8ull * p.max_layers;
    return p;
  }

  int admit(const Graph& g, uint32_t visited, uint32_t max_layers, int32_t* slot_out) {
    const BatchPlan p = plan(g, visited, max_layers);
    if (used_ + p.bytes > budget_) return MGBFS_BATCH_FULL;
    int32_t slot = -1;
    for (uint32_t s = 0; s < slots_n_; ++s)
      if (!slots_[s].used) {
        slot = static_cast<int32_t>(s);
        break;
      }
    if (slot < 0) return MGBFS_BATCH_FULL;
    MGBFS_HIP(hipSetDevice(device_));
    jobs_[slot] = Job{g, visited, max_layers, 1, 0};
    install(slot, p);
    *slot_out = slot;
    return 0;
  }

  // Per-slot facts: visited set (1 bitmap, 2 hash), retained layers (0 =
  // all), planned layer capacity and restarts after an outgrown layer.
  void instance_info(int32_t slot, uint32_t& visited, uint32_t& history, uint64_t& cap, uint32_t& restarts) const {
    check_slot(slot);
    const BatchInstance& d = host_inst_[slot];
    visited = d.visited == kVisBitmap ? 1 : 2;
    history = d.rolling ? 3 : 0;
    cap = d.cap;
    restarts = jobs_[slot].restarts;
  }

  // Lays out a slot's block (frontiers, layer sizes, visited set) from its
  // plan and queues the descriptor for installation and seeding.
  void install(int32_t slot, const BatchPlan& p) {
    const Graph& g = jobs_[slot].graph;
    Slot& sl = slots_[slot];
    const uint64_t keys_bytes = arena_round(p.cap * 8), sizes_bytes = arena_round(8ull * p.max_layers);
    sl.block = arena_.take(2 * keys_bytes + sizes_bytes + p.visited_bytes, budget_);
    char* base = sl.block.ptr;
    BatchInstance& d = host_inst_[slot];
    std::memset(&d, 0, sizeof d);
    d.front = reinterpret_cast<uint64_t*>(base);
    d.next = reinterpret_cast<uint64_t*>(base + keys_bytes);
    d.cap = p.cap;
    d.layer_sizes = reinterpret_cast<uint64_t*>(base + 2 * keys_bytes);
    char* visited_base = base + 2 * keys_bytes + sizes_bytes;
    d.max_layers = p.max_layers;
    d.visited = p.visited;
    d.generators = g.generators;
    uint64_t start_key = 0;
    if (p.visited == kVisBitmap) {
      d.bitmap = reinterpret_cast<uint32_t*>(visited_base);
      d.bitmap_bits = factorial_u64(g.elements);
    } else {
      d.table = reinterpret_cast<uint64_t*>(visited_base);
      d.table_slots = p.table_slots;
      d.table_stride = p.table_slots;
      d.rolling = p.rolling ? 1u : 0u;
    }
    d.tag_shift = 64;
    if (g.kind == 0) {
      d.action = kActPerm;
      d.degree = g.elements;
      for (uint32_t k = 0; k < g.generators; ++k)
        for (uint32_t i = 0; i < g.elements; ++i) {
          d.src[k][i] = static_cast<uint8_t>(g.perms[k * g.elements + i]);
          d.src_word[k] |= static_cast<uint64_t>(g.perms[k * g.elements + i] & 15) << (4 * i);
        }
      start_key = g.start_key();
      if (p.visited == kVisBitmap && g.elements <= 15 && g.generators <= 14) {
        const std::vector<int> inv = inverse_index(g.packed_generators());
        d.tag_shift = 60;
        for (uint32_t k = 0; k < g.generators; ++k) d.child_tag[k] = inv[k] >= 0 ? static_cast<uint8_t>(inv[k] + 1) : 0;
      }
    } else {
      d.action = kActMatrix;
      d.rows = g.rows;
      d.cols = g.cols;
      d.modulus = g.moduli[0];
      uint32_t eb = 1;
      while ((1ull << eb) < static_cast<uint64_t>(d.modulus)) ++eb;
      d.ebits = eb;
      for (uint32_t k = 0; k < g.generators; ++k)
        for (uint32_t e = 0; e < g.rows * g.rows; ++e)
          d.mat[k][e] = static_cast<int32_t>(g.matrices[static_cast<size_t>(k) * g.rows * g.rows + e]);
      for (uint32_t e = 0; e < g.elements; ++e) start_key |= static_cast<uint64_t>(g.start[e]) << (e * eb);
    }
    // The visited set is emptied, and the descriptor installed and seeded,
    // by the next flush(), which the next step runs before its graph.
    pending_[slot] = kInstallSeed;
    pending_key_[slot] = start_key;
    sl.used = true;
    sl.reported = false;
    sl.planned = p.bytes;
    used_ += p.bytes;
    kinds_[slot] = g.kind;
    rows_[slot] = g.rows;
    cols_[slot] = g.cols;
  }

  // Runs up to `layers` layers of every active instance; returns how many
  // instances are still active afterwards.
  uint32_t step(uint32_t layers) {
    MGBFS_HIP(hipSetDevice(device_));
    const uint32_t launches = std::max<uint32_t>(1, (layers + layers_per_graph_ - 1) / layers_per_graph_);
    flush();
    for (uint32_t k = 0; k < launches; ++k) MGBFS_HIP(hipGraphLaunch(exec_, stream_));
    refresh();
    uint32_t active = 0;
    for (uint32_t s = 0; s < slots_n_; ++s) {
      if (!slots_[s].used) continue;
      if (host_inst_[s].status & kInstDone) restart_if_outgrown(s);
      if ((host_inst_[s].status & kInstActive) || pending_[s] == kInstallSeed) ++active;
    }
    return active;
  }

  // A rolling instance whose layer outgrew its planned capacity (next
  // frontier or table full) starts over with twice the room, up to the
  // state bound, when the budget allows; otherwise its error stands.
  void restart_if_outgrown(uint32_t s) {
    const BatchInstance& d = host_inst_[s];
    if (!d.rolling || !(d.error & (kErrFutureCap | kErrTableFull)) || slots_[s].reported) return;
    Job& job = jobs_[s];
    if (d.cap >= state_space_bound(job.graph)) return;
    const BatchPlan p = plan_with(job.graph, job.visited, job.max_layers, true, job.scale * 2);
    if (used_ - slots_[s].planned + p.bytes > budget_) return;
    job.scale *= 2;
    ++job.restarts;
    arena_.give(slots_[s].block, budget_);
    used_ -= slots_[s].planned;
    install(static_cast<int32_t>(s), p);
  }

  void run_all() {
    while (step(64)) {
    }
  }

  // Slots whose instance finished and has not been reported yet (at most
  // `limit`; the rest are reported by later polls).  Their layer tables are
  // read back together, so result() needs no device access.
  std::vector<int32_t> poll(uint32_t limit) {
    std::vector<int32_t> out;
    for (uint32_t s = 0; s < slots_n_ && out.size() < limit; ++s)
      if (slots_[s].used && !slots_[s].reported && (host_inst_[s].status & kInstDone)) {
        slots_[s].reported = true;
        out.push_back(static_cast<int32_t>(s));
      }
    gather(out);
    return out;
  }

  void result(int32_t slot, std::vector<uint64_t>& sizes, uint32_t& status, uint32_t& error) {
    check_slot(slot);
    const BatchInstance& d = host_inst_[slot];
    if (finished_[slot].ready) {
      sizes = finished_[slot].sizes;
    } else {
      sizes.assign(d.depth + 1, 0);
      MGBFS_HIP(hipMemcpy(sizes.data(), d.layer_sizes, sizes.size() * 8, hipMemcpyDeviceToHost));
    }
    status = d.status;
    error = d.error;
  }

  // Current-layer states of a slot as int64 rows (at most limit).
  void sample(int32_t slot, uint64_t limit, std::vector<int64_t>& rows, uint64_t& count) {
    check_slot(slot);
    const BatchInstance& d = host_inst_[slot];
    count = std::min<uint64_t>(limit, d.count);
    std::vector<uint64_t> keys(count);
    if (count) MGBFS_HIP(hipMemcpy(keys.data(), d.front, count * 8, hipMemcpyDeviceToHost));
    const uint32_t w = d.action == kActPerm ? d.degree : d.rows * d.cols;
    rows.assign(count * w, 0);
    for (uint64_t r = 0; r < count; ++r)
      for (uint32_t e = 0; e < w; ++e) {
        const uint32_t bits = d.action == kActPerm ? 4 : d.ebits;
        rows[r * w + e] = static_cast<int64_t>((keys[r] >> (e * bits)) & ((1ull << bits) - 1));
      }
  }

  void release(int32_t slot) {
    check_slot(slot);
    MGBFS_HIP(hipSetDevice(device_));
    Slot& sl = slots_[slot];
    BatchInstance& d = host_inst_[slot];
    std::memset(&d, 0, sizeof d);
    // The cleared descriptor is installed by the next flush, before any
    // graph reads descriptors again, so the range can be reused at once.
    pending_[slot] = kInstallOnly;
    arena_.give(sl.block, budget_);
    used_ -= sl.planned;
    sl = Slot{};
    finished_[slot] = Finished{};
    jobs_[slot] = Job{};
  }

  uint64_t used() const { return used_; }
  uint64_t budget() const { return budget_; }
  uint32_t slots() const { return slots_n_; }

 private:
  void check_slot(int32_t slot) const {
    if (slot < 0 || static_cast<uint32_t>(slot) >= slots_n_ || !slots_[slot].used)
      throw std::runtime_error("BATCH_SLOT");
  }
  void capture() {
    BatchInstance* inst = inst_.as<BatchInstance>();
    uint32_t* prefix = prefix_.as<uint32_t>();
    uint32_t* flag = flag_.as<uint32_t>();
    MGBFS_HIP(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    MGBFS_HIP(batch_reset_flag(flag, stream_));
    for (uint32_t k = 0; k < layers_per_graph_; ++k) {
      MGBFS_HIP(batch_tiles(inst, slots_n_, 2 * threads_, prefix, stream_));
      MGBFS_HIP(batch_clear(inst, slots_n_, prefix, clear_grid_, stream_));
      MGBFS_HIP(batch_expand(inst, slots_n_, prefix, 2 * threads_, grid_, threads_, stream_));
      MGBFS_HIP(batch_finish(inst, slots_n_, flag, stream_));
    }
    MGBFS_HIP(hipStreamEndCapture(stream_, &graph_));
    MGBFS_HIP(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
  }
  // Installs every descriptor admitted or released since the last flush
  // (the latest content per slot) with one upload and two kernels: the
  // visited sets of the admitted ones are emptied, then the descriptors are
  // installed and seeded.  Only the latest descriptor of a slot is staged,
  // so memory an instance released before the flush is cleared, if reused,
  // only as its new owner's visited set.  The pinned staging area is reused
  // only after refresh() has waited.
  void flush() {
    uint32_t n = 0;
    uint64_t clear_bytes = 0;  // largest visited set to empty
    BatchInstall* req = reinterpret_cast<BatchInstall*>(host_stage_ + slots_n_);
    for (uint32_t s = 0; s < slots_n_; ++s) {
      if (pending_[s] == kNothing) continue;
      const BatchInstance& d = host_inst_[s];
      host_stage_[n] = d;
      req[n] = BatchInstall{s, pending_[s] == kInstallSeed ? 1u : 0u, pending_key_[s]};
      if (req[n].seed)
        clear_bytes = std::max<uint64_t>(clear_bytes, d.visited == kVisBitmap
                                                          ? (d.bitmap_bits + 31) / 32 * 4
                                                          : (d.rolling ? 3ull : 1ull) * d.table_slots * 8);
      pending_[s] = kNothing;
      ++n;
    }
    if (n == 0) return;
    const size_t bytes = sizeof(BatchInstance) * slots_n_ + sizeof(BatchInstall) * n;
    MGBFS_HIP(hipMemcpyAsync(stage_.as<void>(), host_stage_, bytes, hipMemcpyHostToDevice, stream_));
    const BatchInstance* staged = stage_.as<BatchInstance>();
    const BatchInstall* reqs = reinterpret_cast<const BatchInstall*>(staged + slots_n_);
    MGBFS_HIP(batch_clear_visited(staged, reqs, n, clear_bytes, stream_));
    MGBFS_HIP(batch_install(inst_.as<BatchInstance>(), staged, reqs, n, stream_));
  }
  void refresh() {
    MGBFS_HIP(hipMemcpyAsync(host_mirror_, inst_.as<void>(), sizeof(BatchInstance) * slots_n_, hipMemcpyDeviceToHost,
                             stream_));
    MGBFS_HIP(hipStreamSynchronize(stream_));
    std::memcpy(host_inst_.data(), host_mirror_, sizeof(BatchInstance) * slots_n_);
  }
  // Layer tables of the given finished slots: one request upload, one gather
  // kernel and one download, then a single wait.
  void gather(const std::vector<int32_t>& done) {
    if (done.empty()) return;
    uint64_t total = 0;
    for (int32_t slot : done) total += host_inst_[slot].depth + 1ull;
    if (total > 0xffffffffull) throw std::runtime_error("BATCH_GATHER_SIZE");
    const size_t req_bytes = 12 * done.size(), out_bytes = 8 * total;
    if (req_bytes > req_cap_) {
      if (host_req_) MGBFS_HIP(hipHostFree(host_req_));
      host_req_ = nullptr;
      MGBFS_HIP(hipHostMalloc(reinterpret_cast<void**>(&host_req_), req_bytes, hipHostMallocDefault));
      gather_req_.allocate(req_bytes);
      req_cap_ = req_bytes;
    }
    if (out_bytes > out_cap_) {
      if (host_out_) MGBFS_HIP(hipHostFree(host_out_));
      host_out_ = nullptr;
      MGBFS_HIP(hipHostMalloc(reinterpret_cast<void**>(&host_out_), out_bytes, hipHostMallocDefault));
      gather_out_.allocate(out_bytes);
      out_cap_ = out_bytes;
    }
    uint32_t off = 0;
    for (size_t q = 0; q < done.size(); ++q) {
      const uint32_t count = host_inst_[done[q]].depth + 1;
      host_req_[3 * q] = static_cast<uint32_t>(done[q]);
      host_req_[3 * q + 1] = off;
      host_req_[3 * q + 2] = count;
      off += count;
    }
    MGBFS_HIP(hipMemcpyAsync(gather_req_.as<void>(), host_req_, req_bytes, hipMemcpyHostToDevice, stream_));
    MGBFS_HIP(batch_gather_sizes(inst_.as<BatchInstance>(), gather_req_.as<uint32_t>(),
                                 static_cast<uint32_t>(done.size()), gather_out_.as<uint64_t>(), stream_));
    MGBFS_HIP(hipMemcpyAsync(host_out_, gather_out_.as<void>(), out_bytes, hipMemcpyDeviceToHost, stream_));
    MGBFS_HIP(hipStreamSynchronize(stream_));
    for (size_t q = 0; q < done.size(); ++q) {
      Finished& f = finished_[done[q]];
      f.sizes.assign(host_out_ + host_req_[3 * q + 1], host_out_ + host_req_[3 * q + 1] + host_req_[3 * q + 2]);
      f.ready = true;
    }
  }

  struct Finished {
    std::vector<uint64_t> sizes;
    bool ready = false;
  };

  struct Job {
    Graph graph;
    uint32_t visited = 0, max_layers = 0, scale = 1, restarts = 0;
  };

  int device_ = 0;
  hipStream_t stream_ = nullptr;
  uint32_t slots_n_ = 0, grid_ = 0, clear_grid_ = 0, layers_per_graph_ = kLayersPerGraph;
  uint32_t threads_ = 256, hash_load_pct_ = 50, layer_pct_ = 30, bitmap_bias_permille_ = 1000;
  uint32_t history_ = 0;
  std::vector<Job> jobs_;
  uint64_t budget_ = 0, used_ = 0;
  DeviceBuffer inst_, prefix_, flag_;
  uint32_t* host_flag_ = nullptr;
  BatchInstance* host_mirror_ = nullptr;  // pinned target of refresh()
  enum Pending : uint8_t { kNothing = 0, kInstallSeed = 1, kInstallOnly = 2 };
  std::vector<uint8_t> pending_;          // per slot: descriptor awaiting flush()
  std::vector<uint64_t> pending_key_;
  BatchInstance* host_stage_ = nullptr;   // pinned: slots_n_ descriptors, then the requests
  DeviceBuffer stage_;
  uint32_t* host_req_ = nullptr;          // pinned gather requests
  uint64_t* host_out_ = nullptr;          // pinned gathered layer tables
  size_t req_cap_ = 0, out_cap_ = 0;
  DeviceBuffer gather_req_, gather_out_;
  std::vector<Finished> finished_;
  std::vector<BatchInstance> host_inst_;
  std::vector<Slot> slots_;
  std::map<int32_t, uint32_t> kinds_, rows_, cols_;
  DeviceArena arena_;
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t exec_ = nullptr;
};

Graph intake_graph(const mgbfs_graph_desc* d);

}  // namespace mgbfs

using mgbfs::BatchEngine;

#define MGBFS_BATCH_GUARD(...)                          \
  try {                                                 \
    __VA_ARGS__                                         \
  } catch (const mgbfs::HipError& e) {                  \
    mgbfs::set_last_error(e.what());                    \
    return -static_cast<int>(e.code == 0 ? 1 : e.code); \
  } catch (const std::exception& e) {                   \
    mgbfs::set_last_error(e.what());                    \
    return -1;                                          \
  }

extern "C" int mgbfs_batch_create(const mgbfs_batch_options* options, void** batch, char* error_text,
                                  size_t error_capacity) {
  if (!options || !batch) return hipErrorInvalidValue;
  *batch = nullptr;
  try {
    *batch = new BatchEngine(*options);
    return 0;
  } catch (const std::exception& e) {
    if (error_text && error_capacity) {
      std::strncpy(error_text, e.what(), error_capacity - 1);
      error_text[error_capacity - 1] = 0;
    }
    mgbfs::set_last_error(e.what());
    return -1;
  }
}

extern "C" int mgbfs_batch_plan(void* batch, const mgbfs_graph_desc* graph, uint32_t visited, uint32_t max_layers,
                                uint64_t* bytes, uint32_t* chosen_visited) {
  if (!batch || !graph) return hipErrorInvalidValue;
  MGBFS_BATCH_GUARD({
    const mgbfs::BatchPlan p =
        static_cast<BatchEngine*>(batch)->plan(mgbfs::intake_graph(graph), visited, max_layers);
    if (bytes) *bytes = p.bytes;
    if (chosen_visited) *chosen_visited = p.visited == mgbfs::kVisBitmap ? 1u : 2u;
    return 0;
  })
}

extern "C" int mgbfs_batch_admit(void* batch, const mgbfs_graph_desc* graph, uint32_t visited, uint32_t max_layers,
                                 int32_t* slot) {
  if (!batch || !graph || !slot) return hipErrorInvalidValue;
  MGBFS_BATCH_GUARD({
    return static_cast<BatchEngine*>(batch)->admit(mgbfs::intake_graph(graph), visited, max_layers, slot);
  })
}

extern "C" int mgbfs_batch_step(void* batch, uint32_t layers, uint32_t* active) {
  if (!batch) return hipErrorInvalidValue;
  MGBFS_BATCH_GUARD({
    const uint32_t a = static_cast<BatchEngine*>(batch)->step(layers);
    if (active) *active = a;
    return 0;
  })
}

extern "C" int mgbfs_batch_run(void* batch) {
  if (!batch) return hipErrorInvalidValue;
  MGBFS_BATCH_GUARD({
    static_cast<BatchEngine*>(batch)->run_all();
    return 0;
  })
}

extern "C" int mgbfs_batch_poll(void* batch, int32_t* slots, uint32_t capacity, uint32_t* count) {
  if (!batch || !count) return hipErrorInvalidValue;
  MGBFS_BATCH_GUARD({
    if (!slots && capacity) return static_cast<int>(hipErrorInvalidValue);
    const std::vector<int32_t> done = static_cast<BatchEngine*>(batch)->poll(capacity);
    *count = static_cast<uint32_t>(done.size());
    for (size_t k = 0; k < done.size(); ++k) slots[k] = done[k];
    return 0;
  })
}

extern "C" int mgbfs_batch_instance_info(void* batch, int32_t slot, uint32_t* visited, uint32_t* history_layers,
                                         uint64_t* layer_capacity, uint32_t* restarts) {
  if (!batch) return hipErrorInvalidValue;
  MGBFS_BATCH_GUARD({
    uint32_t v = 0, h = 0, r = 0;
    uint64_t c = 0;
    static_cast<BatchEngine*>(batch)->instance_info(slot, v, h, c, r);
    if (visited) *visited = v;
    if (history_layers) *history_layers = h;
    if (layer_capacity) *layer_capacity = c;
    if (restarts) *restarts = r;
    return 0;
  })
}

extern "C" int mgbfs_batch_result(void* batch, int32_t slot, uint64_t* sizes, uint32_t capacity, uint32_t* count,
                                  uint32_t* status, uint32_t* error) {
  if (!batch || !count) return hipErrorInvalidValue;
  MGBFS_BATCH_GUARD({
    std::vector<uint64_t> s;
    uint32_t st = 0, er = 0;
    static_cast<BatchEngine*>(batch)->result(slot, s, st, er);
    *count = static_cast<uint32_t>(s.size());
    if (status) *status = st;
    if (error) *error = er;
    if (sizes)
      for (uint32_t k = 0; k < capacity && k < s.size(); ++k) sizes[k] = s[k];
    return 0;
  })
}

extern "C" int mgbfs_batch_sample(void* batch, int32_t slot, uint64_t limit, int64_t* rows, uint64_t* count) {
  if (!batch || !count) return hipErrorInvalidValue;
  MGBFS_BATCH_GUARD({
    std::vector<int64_t> r;
    uint64_t c = 0;
    static_cast<BatchEngine*>(batch)->sample(slot, limit, r, c);
    if (rows && !r.empty()) std::memcpy(rows, r.data(), r.size() * 8);
    *count = c;
    return 0;
  })
}

extern "C" int mgbfs_batch_release(void* batch, int32_t slot) {
  if (!batch) return hipErrorInvalidValue;
  MGBFS_BATCH_GUARD({
    static_cast<BatchEngine*>(batch)->release(slot);
    return 0;
  })
}

extern "C" int mgbfs_batch_usage(void* batch, uint64_t* used, uint64_t* budget, uint32_t* slots) {
  if (!batch) return hipErrorInvalidValue;
  auto* b = static_cast<BatchEngine*>(batch);
  if (used) *used = b->used();
  if (budget) *budget = b->budget();
  if (slots) *slots = b->slots();
  return 0;
}

extern "C" void mgbfs_batch_destroy(void* batch) { delete static_cast<BatchEngine*>(batch); }
