// Host runtime of the sort-unique backend with rotating VRAM banks.
//
// Layer loop (one host readback per layer, plus a retry when a slab
// overflows): expand+route into owner-lane slabs -> partition into
// sub-bucket slabs -> unique_diff -> scan of fine counts -> compact into
// the next bank (queued before the host reads N) -> read N -> rotate.
// Retention is exact for inverse-closed generator sets: every neighbour of
// layer d lies in layer d-1, d or d+1 (graph_definition.py:84-92; retention
// rule history_layers = 3, crates/mgbfs-runtime/src/generic_distributed_run.rs:40).
// With back-edge masks (G <= tagbits - 2) each stored state carries the set
// of generators leading back to layer d-1 -- the OR over all the duplicate
// candidates that discovered it -- and expansion skips them, so every child
// lies in layer d or d+1 and only the current bank is probed.  Otherwise
// nothing is skipped and both retained banks are probed.
#include <algorithm>
#include <cmath>
#include <cstring>
#include "engine.hpp"
#include "kernels.hpp"
#include "sorted_plan.hpp"

namespace mgbfs {

namespace {

constexpr uint32_t kPhaseCount = 4;  // expand, partition, unique, compact

}  // namespace

class SortedEngine final : public Engine {
 public:
  SortedEngine(const Graph& g, const mgbfs_engine_options& o) : gens_(g.packed_generators()), n_(g.elements) {
    device_ = o.device;
    backend = MGBFS_BACKEND_SORTED;
    history_layers = 3;
    if (!g.inverse_closed) throw std::runtime_error("SORTED_REQUIRES_INVERSE_CLOSED");
    MGBFS_HIP(hipSetDevice(device_));
    MGBFS_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    for (auto& e : ev_) MGBFS_HIP(hipEventCreate(&e));
    for (auto& p : cev_)
      for (auto& e : p) MGBFS_HIP(hipEventCreate(&e));
    budget_ = vram_budget(o);
    cap_limit_ = o.capacity ? o.capacity : ~0ull;
    flags_ = o.flags;
    // Stored-value layout and parent-edge tags.
    layout_.km = make_keymix(sorted_key_bits(n_));
    layout_.tagbits = 64 - sorted_key_bits(n_);
    layout_.tagmask = layout_.tagbits ? ((1ull << layout_.tagbits) - 1) : 0;
    set_generators(g);
    // One workgroup per sub-bucket: persistent grids of 2 or 8 workgroups
    // per CU, prefetching the next sub-bucket's bounds, measured slower
    // (127 and 122 vs 120 ms per S13 BFS).
    unique_grid_ = 1u << kMaxBucketBits;
    control_.allocate(64);
    fine_counts_.allocate(((1ull << kMaxBucketBits) + 1) * 4);
    for (int b = 0; b < 3; ++b) off_[b].allocate(((1ull << kMaxBucketBits) + 1) * 4);
    MGBFS_HIP(hipHostMalloc(reinterpret_cast<void**>(&host_), 64, hipHostMallocDefault));
    grow_bank(0, 1024);
    reset();
  }

  ~SortedEngine() override {
    (void)hipSetDevice(device_);
    if (stream_) (void)hipStreamSynchronize(stream_);
    if (host_) (void)hipHostFree(host_);
    for (auto& e : ev_)
      if (e) (void)hipEventDestroy(e);
    for (auto& p : cev_)
      for (auto& e : p)
        if (e) (void)hipEventDestroy(e);
    if (stream_) (void)hipStreamDestroy(stream_);
  }

  void reset() override {
    MGBFS_HIP(hipSetDevice(device_));
    cur_ = 0;
    prev_ = -1;
    slack_ = (flags_ & MGBFS_FLAG_TEST_TIGHT_SLABS) ? 0 : 1;
    last_ratio_ = 0.0;
    for (int b = 0; b < 3; ++b) count_[b] = 0;
    count_[0] = 1;
    nbo_[0] = 1u << kMaxBucketBits;
    MGBFS_HIP(sorted_seed(start_key_, layout_, bank_[0].as<uint64_t>(), off_[0].as<uint32_t>(), kMaxBucketBits,
                          stream_));
    MGBFS_HIP(fill_u64(control_.as<uint64_t>(), 0ull, 8, stream_));
    MGBFS_HIP(hipStreamSynchronize(stream_));
    cpending_[0] = cpending_[1] = false;
    layers.assign(1, 1);
    candidates = unique_candidates = 0;
    kernel_seconds = 0.0;
    launches = 0;
    streamed_bytes = 0;
    for (auto& t : phase_seconds) t = 0.0;
    for (auto& b : phase_bytes) b = 0;
    retries = 0;
    terminal = false;
    update_bytes();
  }

  int advance(uint64_t* next_count, uint32_t* fatal) override {
    if (terminal) throw std::runtime_error("ENGINE_TERMINAL");
    MGBFS_HIP(hipSetDevice(device_));
    const uint64_t F = count_[cur_];
    const uint32_t G = gens_.count;
    // Every parent yields at most G children, at most G-1 when it carries a
    // back-edge mask (every state but the start does); the slabs, the
    // sub-bucket count and the 32-bit candidate bound use that bound.
    const bool masked = layout_.prune && layers.size() > 1;
    const uint64_t Cmax = masked ? F * (G > 1 ? G - 1 : 1) : F * G;
    // Beyond 32-bit candidate counts the run continues on the segmented
    // backend (a test flag lowers the limit to 2^18 children).
    if (Cmax > ((flags_ & MGBFS_FLAG_TEST_SORTED_CAP) ? (1ull << 18) : 0xfffffff0ull)) return kAdvanceHandover;
    const int nxt = next_bank();
    uint64_t Cest = Cmax;
    // Masked layers skip every back-edge, so the previous layer's measured
    // children per parent (plus 10%) predicts this one closely; a low guess
    // costs one retry of the layer (slab or sub-bucket overflow).
    if (masked && last_ratio_ > 0)
      Cest = std::min<uint64_t>(Cest, static_cast<uint64_t>(static_cast<double>(F) * last_ratio_ * 1.1) + 4096);
    const uint64_t kMaxSub = 1ull << kMaxBucketBits;
    uint64_t want = sub_buckets_for(Cest);
    // Test hook: start eight times coarse so large layers exercise the
    // bucket-overflow retry path.
    if ((flags_ & MGBFS_FLAG_TEST_COARSE_BUCKETS) && want >= 8) want >>= 3;
    const bool probe_prev = prev_ >= 0 && !layout_.prune;
    while (true) {
      const SortedGeometry geo = sorted_geometry(want, F, Cest, slack_, margin_);
      const uint32_t K2 = geo.K2, nsub = geo.nsub, NB1 = geo.NB1, P = geo.P, lcap = geo.lcap;
      const uint64_t lslabs = geo.lslabs, lane_entries = geo.lane_entries, sub_entries = geo.sub_entries;
      if (!ensure_slabs(lane_entries + kRouteDump, sub_entries + kRouteDump)) {
        *fatal = kErrFutureCap;
        terminal = true;
        return MGBFS_RESOURCE;
      }
      ensure_u32(lcount_, lslabs);
      ensure_u32(scount_, nsub);
      ensure_u32(scan_ws_, scan_workspace_words(1ull << kMaxBucketBits) + 16);
      const uint32_t fb = geo.fb, nbo = geo.nbo;
      // Counters (unique, error, candidates), sub-bucket and lane cursors.
      MGBFS_HIP(zero_u32x3(reinterpret_cast<uint32_t*>(control_.as<uint64_t>()), 6, scount_.as<uint32_t>(), nsub,
                           lcount_.as<uint32_t>(), K2 ? lslabs : 0, stream_));
      MGBFS_HIP(hipEventRecord(ev_[0], stream_));
      uint64_t* sub = sub_.as<uint64_t>();
      MGBFS_HIP(sorted_expand_route(bank_[cur_].as<uint64_t>(), F, gens_, layout_, masked, nsub, K2, NB1, P, lcap,
                                    K2 ? lcap : kSlabStride, K2 ? kErrLaneSlab : kErrBucketCap,
                                    K2 ? lcount_.as<uint32_t>() : scount_.as<uint32_t>(), error_word(),
                                    K2 ? lane_.as<uint64_t>() : sub, routed(K2 ? lane_entries : sub_entries),
                                    stream_));
      MGBFS_HIP(hipEventRecord(ev_[1], stream_));
      if (K2)
        MGBFS_HIP(sorted_partition(lane_.as<uint64_t>(), lcount_.as<uint32_t>(), lcap, P, nsub, K2,
                                   scount_.as<uint32_t>(), error_word(), sub, routed(sub_entries), stream_));
      MGBFS_HIP(hipEventRecord(ev_[2], stream_));
      auto* control = reinterpret_cast<unsigned long long*>(control_.as<uint64_t>());
      MGBFS_HIP(sorted_unique_diff(
          sub, scount_.as<uint32_t>(), nsub, bank_[cur_].as<uint64_t>(), off_[cur_].as<uint32_t>(), nbo_[cur_],
          probe_prev ? bank_[prev_].as<uint64_t>() : nullptr, probe_prev ? off_[prev_].as<uint32_t>() : nullptr,
          probe_prev ? nbo_[prev_] : 0, fb, fine_counts_.as<uint32_t>(), error_word(), control, control + 2,
          layout_.tagbits, layout_.prune != 0, unique_grid_, stream_));
      MGBFS_HIP(hipEventRecord(ev_[4], stream_));
      MGBFS_HIP(exclusive_scan_u32(fine_counts_.as<uint32_t>(), nbo, off_[nxt].as<uint32_t>(),
                                   scan_ws_.as<uint32_t>(), stream_));
      MGBFS_HIP(hipEventRecord(ev_[3], stream_));
      // Readback: unique count, error word, candidate count, N (last offset).
      MGBFS_HIP(hipMemcpyAsync(host_, control_.as<uint64_t>(), 24, hipMemcpyDeviceToHost, stream_));
      MGBFS_HIP(hipMemcpyAsync(reinterpret_cast<uint32_t*>(host_ + 3), off_[nxt].as<uint32_t>() + nbo, 4,
                               hipMemcpyDeviceToHost, stream_));
      MGBFS_HIP(hipEventRecord(ev_[5], stream_));
      // The compaction is queued before the host has read N or the error
      // word; it writes nothing unless the layer is clean and fits the bank
      // as allocated now, and the host relaunches it after growing the bank.
      // The host wakes on the readback alone, so its turnaround and the next
      // layer's launches overlap the compaction.
      const uint64_t early_rows = bank_cap_[nxt];
      if (early_rows) compact(sub, nsub, fb, nxt);
      MGBFS_HIP(hipEventSynchronize(ev_[5]));
      const uint32_t err = static_cast<uint32_t>(host_[1]);
      const uint64_t C = host_[2];
      const uint64_t N = static_cast<uint32_t>(host_[3]);
      account(F, C, N, probe_prev ? count_[prev_] : 0, K2 != 0);
      if (err & kErrLaneSlab) {  // lane slab overflow: retry with twice the slack
        slack_ = slack_ ? 2 * slack_ : 1;
        ++retries;
        continue;
      }
      if (err & kErrBucketCap) {  // a sub-bucket overflowed: twice as many
        if (want < kMaxSub) {
          want = std::min(kMaxSub, 2 * want);
          ++retries;
          continue;
        }
      }
      if (err) {
        *fatal = err;
        terminal = true;
        return MGBFS_RESOURCE;
      }
      candidates += C;
      last_ratio_ = masked ? static_cast<double>(C) / static_cast<double>(F) : 0.0;
      unique_candidates += host_[0];
      if (N == 0) {
        *next_count = 0;
        terminal = true;
        return MGBFS_COMPLETE;
      }
      if (N > cap_limit_ || !grow_bank(nxt, N)) {
        *fatal = kErrFutureCap;
        terminal = true;
        return MGBFS_RESOURCE;
      }
      if (N > early_rows) compact(sub, nsub, fb, nxt);
      phase_bytes[3] += 16 * N;
      streamed_bytes += 16 * N;
      launches += 1;
      nbo_[nxt] = nbo;
      count_[nxt] = N;
      prev_ = cur_;
      cur_ = nxt;
      layers.push_back(N);
      *next_count = N;
      update_bytes();
      return MGBFS_LAYER;
    }
  }

  // Another inverse-closed graph of the same degree reuses every bank and slab.
  void rebind(const Graph& g) override {
    if (g.elements != n_ || !g.packable() || !g.inverse_closed) throw std::runtime_error("ENGINE_REBIND_SHAPE");
    gens_ = g.packed_generators();
    set_generators(g);
    reset();
  }

  bool handover(PackedHandover& h) override {
    MGBFS_HIP(hipStreamSynchronize(stream_));
    settle();
    lane_.release();
    sub_.release();
    lane_cap_ = sub_cap_ = 0;
    h.cur = bank_[cur_].as<uint64_t>();
    h.cur_count = count_[cur_];
    h.prev = prev_ >= 0 ? bank_[prev_].as<uint64_t>() : nullptr;
    h.prev_count = prev_ >= 0 ? count_[prev_] : 0;
    h.km = layout_.km;
    h.tagbits = layout_.tagbits;
    h.layers = layers;
    update_bytes();
    return true;
  }

  void layer_keys(std::vector<uint64_t>& out) override { read_bank(cur_, count_[cur_], out); }

  void sample(uint32_t which, uint64_t limit, std::vector<int64_t>& rows, uint64_t& count) override {
    std::vector<uint64_t> keys;
    if (which == 0) read_bank(cur_, std::min(limit, count_[cur_]), keys);
    else if (prev_ >= 0) read_bank(prev_, std::min(limit, count_[prev_]), keys);
    rows.resize(keys.size() * n_);
    for (size_t k = 0; k < keys.size(); ++k)
      for (uint32_t i = 0; i < n_; ++i) rows[k * n_ + i] = static_cast<int64_t>((keys[k] >> (4u * i)) & 15u);
    count = keys.size();
  }

  const char* name() const override { return "SORTED_THREE_BANK_LDS_UNIQUE"; }
  uint64_t retry_count() const override { return retries; }

  void settle() override {
    for (int s = 0; s < 2; ++s) read_compact(s);
  }

  double phase_seconds[kPhaseCount] = {0, 0, 0, 0};
  uint64_t phase_bytes[kPhaseCount] = {0, 0, 0, 0};
  uint64_t retries = 0;

  void fill_phases(double* seconds, uint64_t* bytes, uint32_t n) const override {
    for (uint32_t i = 0; i < n && i < kPhaseCount; ++i) {
      seconds[i] = phase_seconds[i];
      bytes[i] = phase_bytes[i];
    }
  }

 private:
  uint32_t* error_word() { return reinterpret_cast<uint32_t*>(control_.as<uint64_t>() + 1); }
  // Back-edge masks of gens_ (tag bits 0..G-1; bit tagbits-1 is the in-bank
  // mark and bit tagbits-2 stays clear, so no entry is all ones), the
  // fine-bucket margin and the start state.
  void set_generators(const Graph& g) {
    if (!fill_back_edge_masks(gens_)) throw std::runtime_error("SORTED_REQUIRES_INVERSE_CLOSED");
    layout_.prune = gens_.count + 2 <= layout_.tagbits && !(flags_ & MGBFS_FLAG_TEST_NO_PRUNE);
    if (!layout_.prune)
      for (uint32_t a = 0; a < gens_.count; ++a) gens_.child_mask[a] = 0;
    history_layers = layout_.prune ? 2 : 3;
    margin_ = fine_margin(gens_.count);
    start_key_ = g.start_key();
  }
  // Compactions run unwaited; each is bracketed by one of two event pairs,
  // read (and charged to phase 3) when the pair is reused or on settle().
  // A pair is reused two compactions later, after the host has waited on
  // the stream past it (a later readback, or the sync before a bank grows),
  // so that read does not block.
  void compact(uint64_t* sub, uint32_t nsub, uint32_t fb, int nxt) {
    const int s = cslot_ ^= 1;
    read_compact(s);
    MGBFS_HIP(hipEventRecord(cev_[s][0], stream_));
    MGBFS_HIP(sorted_compact(sub, nsub, fb, off_[nxt].as<uint32_t>(), error_word(), bank_cap_[nxt],
                             bank_[nxt].as<uint64_t>(), stream_));
    MGBFS_HIP(hipEventRecord(cev_[s][1], stream_));
    cpending_[s] = true;
  }
  void read_compact(int s) {
    if (!cpending_[s]) return;
    MGBFS_HIP(hipEventSynchronize(cev_[s][1]));
    float ms = 0.f;
    MGBFS_HIP(hipEventElapsedTime(&ms, cev_[s][0], cev_[s][1]));
    phase_seconds[3] += ms * 1e-3;
    kernel_seconds += ms * 1e-3;
    cpending_[s] = false;
  }
  // Routing output size as passed to the kernels; the test flag reports an
  // oversized output so they take their 64-bit addressing path.
  uint64_t routed(uint64_t entries) const {
    return (flags_ & MGBFS_FLAG_TEST_WIDE_ROUTING) ? ~0ull : entries;
  }
  int next_bank() const {
    for (int b = 0; b < 3; ++b)
      if (b != cur_ && b != prev_) return b;
    return -1;
  }
  // Phase times (events on the engine stream) and modelled streaming bytes
  // (see the kernel file header).  Phase 2 is unique_diff alone; the fine
  // count scan is charged to phase 3 with the compaction it feeds.
  void account(uint64_t F, uint64_t C, uint64_t N, uint64_t Fp, bool partitioned) {
    float a = 0, b = 0, c = 0, d = 0;
    MGBFS_HIP(hipEventElapsedTime(&a, ev_[0], ev_[1]));
    MGBFS_HIP(hipEventElapsedTime(&b, ev_[1], ev_[2]));
    MGBFS_HIP(hipEventElapsedTime(&c, ev_[2], ev_[4]));
    MGBFS_HIP(hipEventElapsedTime(&d, ev_[4], ev_[3]));
    phase_seconds[0] += a * 1e-3;
    phase_seconds[1] += b * 1e-3;
    phase_seconds[2] += c * 1e-3;
    phase_seconds[3] += d * 1e-3;
    kernel_seconds += (a + b + c + d) * 1e-3;
    const uint64_t be = 8 * F + 8 * C, bp = partitioned ? 16 * C : 0, bu = 8 * C + 8 * F + 8 * Fp + 8 * N;
    phase_bytes[0] += be;
    phase_bytes[1] += bp;
    phase_bytes[2] += bu;
    streamed_bytes += be + bp + bu;
    launches += partitioned ? 3 : 2;
  }
  void ensure_u32(DeviceBuffer& b, uint64_t words) {
    if (b.bytes() < words * 4) {
      MGBFS_HIP(hipStreamSynchronize(stream_));
      b.allocate(std::max<uint64_t>(words * 4, b.bytes() * 2));
    }
  }
  // Lane slabs (entries) and sub-bucket slabs (entries); grown, never shrunk.
  bool ensure_slabs(uint64_t lane_entries, uint64_t sub_entries) {
    if (lane_cap_ >= lane_entries && sub_cap_ >= sub_entries) return true;
    uint64_t lane = std::max(lane_cap_, lane_entries), sub = std::max(sub_cap_, sub_entries);
    if (!fits((lane - lane_cap_ + sub - sub_cap_) * 8)) return false;
    MGBFS_HIP(hipStreamSynchronize(stream_));
    if (lane > lane_cap_) {
      lane_.release();
      lane_.allocate(lane * 8);
      lane_cap_ = lane;
    }
    if (sub > sub_cap_) {
      sub_.release();
      sub_.allocate(sub * 8);
      sub_cap_ = sub;
    }
    update_bytes();
    return true;
  }
  bool grow_bank(int b, uint64_t rows) {
    if (bank_cap_[b] >= rows) return true;
    uint64_t want = std::max<uint64_t>(rows, bank_cap_[b] + bank_cap_[b] / 4);
    if (!fits((want - bank_cap_[b]) * 8)) want = rows;
    if (!fits((want - bank_cap_[b]) * 8)) return false;
    MGBFS_HIP(hipStreamSynchronize(stream_));
    bank_[b].release();
    bank_[b].allocate(want * 8);
    bank_cap_[b] = want;
    update_bytes();
    return true;
  }
  bool fits(uint64_t extra) const { return device_bytes + extra <= budget_; }
  void read_bank(int b, uint64_t count, std::vector<uint64_t>& out) {
    std::vector<uint64_t> raw(count);
    if (count) {
      MGBFS_HIP(hipMemcpyAsync(raw.data(), bank_[b].as<uint64_t>(), count * 8, hipMemcpyDeviceToHost, stream_));
      MGBFS_HIP(hipStreamSynchronize(stream_));
    }
    out.resize(count);
    for (uint64_t i = 0; i < count; ++i) out[i] = keyunmix(raw[i] >> layout_.tagbits, layout_.km);
  }
  void update_bytes() {
    device_bytes = control_.bytes() + fine_counts_.bytes() + lane_.bytes() + sub_.bytes() + lcount_.bytes() +
                   scount_.bytes() + scan_ws_.bytes();
    for (int b = 0; b < 3; ++b) device_bytes += bank_[b].bytes() + off_[b].bytes();
  }

  PackedGenerators gens_;
  SortedLayout layout_{};
  uint32_t n_, margin_ = 0, flags_ = 0, unique_grid_ = 2048, slack_ = 1;
  double last_ratio_ = 0.0;  // children per parent of the last masked layer
  uint64_t start_key_ = 0, budget_ = 0, cap_limit_ = 0, lane_cap_ = 0, sub_cap_ = 0;
  DeviceBuffer bank_[3], off_[3], lane_, sub_, lcount_, scount_, scan_ws_, fine_counts_, control_;
  uint64_t bank_cap_[3] = {0, 0, 0}, count_[3] = {0, 0, 0};
  uint32_t nbo_[3] = {0, 0, 0};  // fine buckets of each bank
  int cur_ = 0, prev_ = -1;
  uint64_t* host_ = nullptr;
  hipEvent_t ev_[6] = {}, cev_[2][2] = {};
  bool cpending_[2] = {false, false};
  int cslot_ = 0;
};

Engine* make_sorted_engine(const Graph& g, const mgbfs_engine_options& o) { return new SortedEngine(g, o); }

}  // namespace mgbfs
