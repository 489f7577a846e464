// Host runtime of the sort-unique backend with three rotating VRAM banks.
//
// Layer loop (one host readback per layer, plus a retry when a sub-bucket
// exceeds the LDS capacity): expand+route -> partition -> unique_diff ->
// scan of fine counts -> read N -> compact into the next bank -> rotate.
// Retention is exact for inverse-closed generator sets: every neighbour of
// layer d lies in layer d-1, d or d+1 (graph_definition.py:84-92; retention
// rule history_layers = 3, crates/mgbfs-runtime/src/generic_distributed_run.rs:40).
#include <algorithm>
#include <cmath>
#include <cstring>
#include "engine.hpp"
#include "kernels.hpp"

namespace mgbfs {

namespace {

uint32_t ceil_log2(uint64_t x) {
  uint32_t b = 0;
  while ((1ull << b) < x) ++b;
  return b;
}

constexpr uint64_t kTargetBucket = 3072;  // average candidates per sub-bucket (cap 4096)
constexpr uint32_t kPhaseCount = 4;       // expand, partition, unique, compact

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
    budget_ = o.vram_budget;
    if (!budget_) {
      size_t fr = 0, tot = 0;
      MGBFS_HIP(hipMemGetInfo(&fr, &tot));
      budget_ = fr > (2ull << 30) ? fr - (2ull << 30) : 0;
    }
    cap_limit_ = o.capacity ? o.capacity : ~0ull;
    flags_ = o.flags;
    // Stored-value layout and parent-edge tags.
    layout_.km = make_keymix(4 * n_);
    layout_.tagbits = 64 - 4 * n_;
    layout_.tagmask = layout_.tagbits ? ((1ull << layout_.tagbits) - 1) : 0;
    const std::vector<int> inv = inverse_index(gens_);
    for (int x : inv)
      if (x < 0) throw std::runtime_error("SORTED_REQUIRES_INVERSE_CLOSED");
    // Tags 1..G must leave the all-ones tag value unused (kEmpty sentinel).
    layout_.prune = layout_.tagbits >= 5 && gens_.count + 1 < (1ull << layout_.tagbits) - 1;
    for (uint32_t a = 0; a < gens_.count; ++a)
      gens_.child_tag[a] = (layout_.prune && inv[a] >= 0) ? static_cast<uint8_t>(inv[a] + 1) : 0;
    margin_ = std::min<uint32_t>(kMaxFineBits, 2 * (ceil_log2(gens_.count) + 1));
    start_key_ = g.start_key();
    // One workgroup per sub-bucket measured faster than a persistent grid
    // with next-bucket prefetch (the kernel is bound by LDS atomics, not by
    // load latency), so the grid is not capped.
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
    if (stream_) (void)hipStreamDestroy(stream_);
  }

  void reset() override {
    MGBFS_HIP(hipSetDevice(device_));
    cur_ = 0;
    prev_ = -1;
    for (int b = 0; b < 3; ++b) count_[b] = 0;
    count_[0] = 1;
    ko_[0] = kMaxBucketBits;
    MGBFS_HIP(sorted_seed(start_key_, layout_, bank_[0].as<uint64_t>(), off_[0].as<uint32_t>(), kMaxBucketBits,
                          stream_));
    MGBFS_HIP(fill_u64(control_.as<uint64_t>(), 0ull, 8, stream_));
    MGBFS_HIP(hipStreamSynchronize(stream_));
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
    const uint64_t Cmax = F * G;
    if (Cmax > 0xfffffff0ull) {
      *fatal = kErrFutureCap;
      terminal = true;
      return MGBFS_RESOURCE;
    }
    if (!ensure_candidates(Cmax)) {
      *fatal = kErrFutureCap;
      terminal = true;
      return MGBFS_RESOURCE;
    }
    const int nxt = next_bank();
    // With parent-edge tags every tagged parent skips exactly one generator.
    const uint64_t Cest = (layout_.prune && layers.size() > 1) ? F * (G - 1 ? G - 1 : 1) : Cmax;
    uint32_t K = Cest <= kTargetBucket ? 0 : ceil_log2((Cest + kTargetBucket - 1) / kTargetBucket);
    // Test hook: start three bits coarse so large layers exercise the
    // bucket-overflow retry path.
    if ((flags_ & MGBFS_FLAG_TEST_COARSE_BUCKETS) && K >= 3) K -= 3;
    uint32_t Klimit = std::min<uint32_t>(kMaxBucketBits, ko_[cur_]);
    if (prev_ >= 0) Klimit = std::min(Klimit, ko_[prev_]);
    K = std::min(K, Klimit);
    while (true) {
      const uint32_t K1 = K <= kMaxPassBits ? K : (K + 1) / 2;
      const uint32_t K2 = K - K1;
      const uint32_t NB1 = static_cast<uint32_t>(std::clamp<uint64_t>((F + 2047) / 2048, 1, 2048));
      const uint64_t n1 = (1ull << K1) * NB1;
      uint32_t NC = 1;
      if (K2) NC = static_cast<uint32_t>(std::clamp<uint64_t>((Cest >> K1) / 262144 + 1, 1, 16));
      const uint64_t n2 = K2 ? (1ull << K) * NC : 0;
      ensure_u32(counts1_, n1);
      ensure_u32(scan1_, n1 + 1);
      if (K2) {
        ensure_u32(counts2_, n2);
        ensure_u32(scan2_, n2 + 1);
      }
      ensure_u32(scan_ws_, scan_workspace_words(std::max<uint64_t>({n1, n2, 1ull << kMaxBucketBits})) + 16);
      const uint32_t Ko = std::min<uint32_t>(kMaxBucketBits, K + margin_);
      MGBFS_HIP(fill_u64(control_.as<uint64_t>(), 0ull, 2, stream_));
      MGBFS_HIP(hipEventRecord(ev_[0], stream_));
      const bool all_tagged = layout_.prune && layers.size() > 1;
      MGBFS_HIP(sorted_expand_route(bank_[cur_].as<uint64_t>(), F, gens_, layout_, all_tagged, K1, NB1,
                                    counts1_.as<uint32_t>(),
                                    scan1_.as<uint32_t>(), scan_ws_.as<uint32_t>(), cand1_.as<uint64_t>(), stream_));
      MGBFS_HIP(hipEventRecord(ev_[1], stream_));
      uint64_t* seg = cand1_.as<uint64_t>();
      const uint32_t* segscan = scan1_.as<uint32_t>();
      uint32_t segNC = NB1;
      if (K2) {
        MGBFS_HIP(sorted_partition(cand1_.as<uint64_t>(), scan1_.as<uint32_t>(), NB1, K1, K2, NC,
                                   counts2_.as<uint32_t>(), scan2_.as<uint32_t>(), scan_ws_.as<uint32_t>(),
                                   cand2_.as<uint64_t>(), stream_));
        seg = cand2_.as<uint64_t>();
        segscan = scan2_.as<uint32_t>();
        segNC = NC;
      }
      MGBFS_HIP(hipEventRecord(ev_[2], stream_));
      const bool has_prev = prev_ >= 0;
      MGBFS_HIP(sorted_unique_diff(
          seg, segscan, segNC, K, bank_[cur_].as<uint64_t>(), off_[cur_].as<uint32_t>(), ko_[cur_],
          has_prev ? bank_[prev_].as<uint64_t>() : nullptr, has_prev ? off_[prev_].as<uint32_t>() : nullptr,
          has_prev ? ko_[prev_] : 0, Ko, fine_counts_.as<uint32_t>(), error_word(),
          reinterpret_cast<unsigned long long*>(control_.as<uint64_t>()), layout_.tagbits, unique_grid_, stream_));
      MGBFS_HIP(exclusive_scan_u32(fine_counts_.as<uint32_t>(), 1ull << Ko, off_[nxt].as<uint32_t>(),
                                   scan_ws_.as<uint32_t>(), stream_));
      MGBFS_HIP(hipEventRecord(ev_[3], stream_));
      // Readback: unique count, error word, N (last offset), candidate total.
      MGBFS_HIP(hipMemcpyAsync(host_, control_.as<uint64_t>(), 16, hipMemcpyDeviceToHost, stream_));
      MGBFS_HIP(hipMemcpyAsync(reinterpret_cast<uint32_t*>(host_ + 2), off_[nxt].as<uint32_t>() + (1ull << Ko), 4,
                               hipMemcpyDeviceToHost, stream_));
      MGBFS_HIP(hipMemcpyAsync(reinterpret_cast<uint32_t*>(host_ + 3), scan1_.as<uint32_t>() + n1, 4,
                               hipMemcpyDeviceToHost, stream_));
      MGBFS_HIP(hipStreamSynchronize(stream_));
      const uint32_t err = static_cast<uint32_t>(host_[1]);
      const uint64_t N = static_cast<uint32_t>(host_[2]);
      const uint64_t C = static_cast<uint32_t>(host_[3]);
      account(0, 1, 2, 3, F, C, N, has_prev ? count_[prev_] : 0, K2 != 0);
      if (err & kErrBucketCap) {
        if (K < Klimit) {
          ++K;
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
      MGBFS_HIP(hipEventRecord(ev_[4], stream_));
      MGBFS_HIP(sorted_compact(seg, segscan, segNC, K, Ko, off_[nxt].as<uint32_t>(), bank_[nxt].as<uint64_t>(),
                               stream_));
      MGBFS_HIP(hipEventRecord(ev_[5], stream_));
      MGBFS_HIP(hipEventSynchronize(ev_[5]));
      float ms = 0.f;
      MGBFS_HIP(hipEventElapsedTime(&ms, ev_[4], ev_[5]));
      phase_seconds[3] += ms * 1e-3;
      kernel_seconds += ms * 1e-3;
      phase_bytes[3] += 16 * N;
      streamed_bytes += 16 * N;
      launches += 1;
      ko_[nxt] = Ko;
      count_[nxt] = N;
      prev_ = cur_;
      cur_ = nxt;
      layers.push_back(N);
      *next_count = N;
      update_bytes();
      return MGBFS_LAYER;
    }
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
  int next_bank() const {
    for (int b = 0; b < 3; ++b)
      if (b != cur_ && b != prev_) return b;
    return -1;
  }
  // Modelled streaming bytes per phase (see the kernel file header).
  void account(int e0, int e1, int e2, int e3, uint64_t F, uint64_t C, uint64_t N, uint64_t Fp, bool partitioned) {
    float a = 0, b = 0, c = 0;
    MGBFS_HIP(hipEventElapsedTime(&a, ev_[e0], ev_[e1]));
    MGBFS_HIP(hipEventElapsedTime(&b, ev_[e1], ev_[e2]));
    MGBFS_HIP(hipEventElapsedTime(&c, ev_[e2], ev_[e3]));
    phase_seconds[0] += a * 1e-3;
    phase_seconds[1] += b * 1e-3;
    phase_seconds[2] += c * 1e-3;
    kernel_seconds += (a + b + c) * 1e-3;
    const uint64_t be = 16 * F + 8 * C, bp = partitioned ? 24 * C : 0, bu = 8 * C + 8 * F + 8 * Fp + 8 * N;
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
  bool ensure_candidates(uint64_t c) {
    if (cand_cap_ >= c) return true;
    uint64_t want = std::max<uint64_t>(c, cand_cap_ + cand_cap_ / 2);
    if (!fits(2 * (want - cand_cap_) * 8)) want = c;
    if (!fits(2 * (want - cand_cap_) * 8)) return false;
    MGBFS_HIP(hipStreamSynchronize(stream_));
    cand1_.release();
    cand2_.release();
    cand1_.allocate(want * 8);
    cand2_.allocate(want * 8);
    cand_cap_ = want;
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
    device_bytes = control_.bytes() + fine_counts_.bytes() + cand1_.bytes() + cand2_.bytes() + counts1_.bytes() +
                   scan1_.bytes() + counts2_.bytes() + scan2_.bytes() + scan_ws_.bytes();
    for (int b = 0; b < 3; ++b) device_bytes += bank_[b].bytes() + off_[b].bytes();
  }

  PackedGenerators gens_;
  SortedLayout layout_{};
  uint32_t n_, margin_ = 0, flags_ = 0, unique_grid_ = 2048;
  uint64_t start_key_ = 0, budget_ = 0, cap_limit_ = 0, cand_cap_ = 0;
  DeviceBuffer bank_[3], off_[3], cand1_, cand2_, counts1_, scan1_, counts2_, scan2_, scan_ws_, fine_counts_, control_;
  uint64_t bank_cap_[3] = {0, 0, 0}, count_[3] = {0, 0, 0};
  uint32_t ko_[3] = {0, 0, 0};
  int cur_ = 0, prev_ = -1;
  uint64_t* host_ = nullptr;
  hipEvent_t ev_[6] = {};
};

Engine* make_sorted_engine(const Graph& g, const mgbfs_engine_options& o) { return new SortedEngine(g, o); }

}  // namespace mgbfs
