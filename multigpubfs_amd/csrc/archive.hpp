// Host-only writer of the layer archive (no HIP): the CLI streams each
// finished layer's sample into it, and the CPU tests build it with g++.
#pragma once
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "host_util.hpp"

namespace mgbfs::host {

// Layer archive, selected-samples form (`MGBFSAS2`; the reference reads it
// with multigpubfs/specialized.py:54-79 and writes it in
// crates/mgbfs-runtime/src/archive.rs:436-690).  Header: magic, state width
// (u64 LE), 32-byte configuration digest (here: the graph digest).  Frames:
// 80-byte head {"MGBFSFR1", kind, depth, count, payload bytes, sequence (u64
// LE each), previous chain hash}, payload, sha256(head + payload) -- which
// is the next chain hash.  Kinds: 1 records (<= 1000 states of the layer,
// one byte per element), 2 layer end (count = its records), 3 commit (depth
// = layers, count = all records).  Every run that stops cleanly (complete,
// deadline, cancel, capacity) commits the layers it finished; a failed run
// leaves an uncommitted prefix, which readers reject.
class LayerArchive {
 public:
  LayerArchive(const std::string& path, uint64_t width, const std::string& config_hex) : width_(width) {
    if (width == 0 || width > 33025) throw std::runtime_error("ARCHIVE_SHAPE");
    if (config_hex.size() != 64 || config_hex.find_first_not_of("0123456789abcdefABCDEF") != std::string::npos)
      throw std::runtime_error("ARCHIVE_CONFIG");
    f_ = std::fopen(path.c_str(), "wb");
    if (!f_) throw std::runtime_error("ARCHIVE_OPEN");
    uint8_t head[48] = {'M', 'G', 'B', 'F', 'S', 'A', 'S', '2'};
    put64(head + 8, width);
    for (int i = 0; i < 32; ++i) head[16 + i] = static_cast<uint8_t>(std::stoi(config_hex.substr(2 * i, 2), nullptr, 16));
    write(head, sizeof head);
    Sha256 h;
    h.update(head, sizeof head);
    h.bytes(chain_);
  }
  ~LayerArchive() {
    if (f_) std::fclose(f_);
  }
  // Records of layer `depth`: rows x width int64 elements, each 0..255.
  void layer(uint64_t depth, const std::vector<int64_t>& rows, uint64_t count) {
    live();
    if (depth != layers_) throw std::runtime_error("ARCHIVE_LAYER_ORDER");
    if (count > 1000) throw std::runtime_error("ARCHIVE_RECORD_BOUND");
    if (rows.size() / width_ < count) throw std::runtime_error("ARCHIVE_RECORD_SHAPE");
    if (count) {
      std::vector<uint8_t> payload(count * width_);
      for (size_t i = 0; i < payload.size(); ++i) {
        if (rows[i] < 0 || rows[i] > 255) throw std::runtime_error("ARCHIVE_STATE_WIDTH");
        payload[i] = static_cast<uint8_t>(rows[i]);
      }
      frame(1, depth, count, payload.data(), payload.size());
    }
    frame(2, depth, count, nullptr, 0);
    total_ += count;
    layers_ = depth + 1;
    std::fflush(f_);
  }
  // Seals the archive; at least one layer must have been written.
  void commit() {
    live();
    if (layers_ == 0) throw std::runtime_error("ARCHIVE_UNCOMMITTED_LAYER");
    frame(3, layers_, total_, nullptr, 0);
    if (std::fflush(f_) || std::fclose(f_)) {
      f_ = nullptr;
      throw std::runtime_error("ARCHIVE_WRITE");
    }
    f_ = nullptr;
  }

 private:
  void live() const {
    if (!f_) throw std::runtime_error("ARCHIVE_CLOSED");
  }
  static void put64(uint8_t* p, uint64_t v) {
    for (int k = 0; k < 8; ++k) p[k] = static_cast<uint8_t>(v >> (8 * k));
  }
  void write(const void* p, size_t n) {
    if (n && std::fwrite(p, 1, n, f_) != n) throw std::runtime_error("ARCHIVE_WRITE");
  }
  void frame(uint64_t kind, uint64_t depth, uint64_t count, const uint8_t* payload, size_t size) {
    uint8_t head[80] = {'M', 'G', 'B', 'F', 'S', 'F', 'R', '1'};
    put64(head + 8, kind);
    put64(head + 16, depth);
    put64(head + 24, count);
    put64(head + 32, size);
    put64(head + 40, sequence_++);
    std::memcpy(head + 48, chain_, 32);
    Sha256 h;
    h.update(head, sizeof head);
    if (size) h.update(payload, size);
    h.bytes(chain_);
    write(head, sizeof head);
    write(payload, size);
    write(chain_, 32);
  }

  std::FILE* f_ = nullptr;
  uint64_t width_, sequence_ = 0, total_ = 0, layers_ = 0;
  uint8_t chain_[32];
};

}  // namespace mgbfs::host
