// Host-only helpers for the command-line host: a small JSON value/parser/
// writer (objects, arrays, strings, 64-bit integers, booleans, null) and
// SHA-256 (FIPS 180-4).  Object keys are kept sorted, so dump() yields the
// canonical form json.dumps(sort_keys=True, separators=(',', ':')).
#pragma once
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace mgbfs::host {

struct Json {
  enum Kind { Null, Bool, Int, UInt, Real, Str, Arr, Obj } kind = Null;
  bool b = false;
  int64_t i = 0;
  uint64_t u = 0;
  double d = 0;
  std::string s;
  std::vector<Json> a;
  std::map<std::string, Json> o;

  Json() = default;
  static Json integer(int64_t v) { Json j; j.kind = Int; j.i = v; return j; }
  static Json unsigned_(uint64_t v) { Json j; j.kind = UInt; j.u = v; return j; }
  static Json real(double v) { Json j; j.kind = Real; j.d = v; return j; }
  static Json str(std::string v) { Json j; j.kind = Str; j.s = std::move(v); return j; }
  static Json boolean(bool v) { Json j; j.kind = Bool; j.b = v; return j; }
  static Json array() { Json j; j.kind = Arr; return j; }
  static Json object() { Json j; j.kind = Obj; return j; }

  bool is_int() const { return kind == Int || (kind == UInt && u <= static_cast<uint64_t>(INT64_MAX)); }
  int64_t as_int() const {
    if (kind == Int) return i;
    if (kind == UInt && u <= static_cast<uint64_t>(INT64_MAX)) return static_cast<int64_t>(u);
    throw std::runtime_error("GRAPH_INT64_RANGE");
  }
  bool has(const std::string& k) const { return kind == Obj && o.count(k); }
  const Json& at(const std::string& k) const {
    auto it = o.find(k);
    if (kind != Obj || it == o.end()) throw std::runtime_error("GRAPH_SCHEMA");
    return it->second;
  }
  Json& operator[](const std::string& k) {
    kind = Obj;
    return o[k];
  }
  void push(Json v) {
    kind = Arr;
    a.push_back(std::move(v));
  }
};

class JsonParser {
 public:
  explicit JsonParser(const std::string& t) : t_(t) {}
  Json parse() {
    Json v = value();
    ws();
    if (p_ != t_.size()) fail();
    return v;
  }

 private:
  [[noreturn]] void fail() const { throw std::runtime_error("JSON_SYNTAX"); }
  void ws() {
    while (p_ < t_.size() && (t_[p_] == ' ' || t_[p_] == '\n' || t_[p_] == '\r' || t_[p_] == '\t')) ++p_;
  }
  bool eat(char c) {
    ws();
    if (p_ < t_.size() && t_[p_] == c) {
      ++p_;
      return true;
    }
    return false;
  }
  Json value() {
    ws();
    if (p_ >= t_.size()) fail();
    const char c = t_[p_];
    if (c == '{') return object();
    if (c == '[') return array();
    if (c == '"') return Json::str(string());
    if (t_.compare(p_, 4, "null") == 0) { p_ += 4; return Json(); }
    if (t_.compare(p_, 4, "true") == 0) { p_ += 4; return Json::boolean(true); }
    if (t_.compare(p_, 5, "false") == 0) { p_ += 5; return Json::boolean(false); }
    return number();
  }
  Json object() {
    Json j = Json::object();
    ++p_;
    if (eat('}')) return j;
    do {
      ws();
      if (p_ >= t_.size() || t_[p_] != '"') fail();
      std::string k = string();
      if (!eat(':')) fail();
      j.o[k] = value();
    } while (eat(','));
    if (!eat('}')) fail();
    return j;
  }
  Json array() {
    Json j = Json::array();
    ++p_;
    if (eat(']')) return j;
    do j.a.push_back(value());
    while (eat(','));
    if (!eat(']')) fail();
    return j;
  }
  std::string string() {
    std::string out;
    ++p_;
    while (p_ < t_.size() && t_[p_] != '"') {
      char c = t_[p_++];
      if (c == '\\') {
        if (p_ >= t_.size()) fail();
        const char e = t_[p_++];
        switch (e) {
          case 'n': out += '\n'; break;
          case 't': out += '\t'; break;
          case 'r': out += '\r'; break;
          case 'b': out += '\b'; break;
          case 'f': out += '\f'; break;
          case 'u': {
            if (p_ + 4 > t_.size()) fail();
            const unsigned cp = std::stoul(t_.substr(p_, 4), nullptr, 16);
            p_ += 4;
            if (cp < 0x80) out += static_cast<char>(cp);
            else if (cp < 0x800) { out += static_cast<char>(0xC0 | (cp >> 6)); out += static_cast<char>(0x80 | (cp & 63)); }
            else { out += static_cast<char>(0xE0 | (cp >> 12)); out += static_cast<char>(0x80 | ((cp >> 6) & 63)); out += static_cast<char>(0x80 | (cp & 63)); }
            break;
          }
          default: out += e;
        }
      } else {
        out += c;
      }
    }
    if (p_ >= t_.size()) fail();
    ++p_;
    return out;
  }
  Json number() {
    const size_t start = p_;
    bool neg = false, real = false;
    if (t_[p_] == '-') { neg = true; ++p_; }
    while (p_ < t_.size() && (isdigit(static_cast<unsigned char>(t_[p_])) || t_[p_] == '.' || t_[p_] == 'e' ||
                              t_[p_] == 'E' || t_[p_] == '+' || t_[p_] == '-')) {
      if (!isdigit(static_cast<unsigned char>(t_[p_]))) real = true;
      ++p_;
    }
    const std::string tok = t_.substr(start, p_ - start);
    if (tok.empty() || tok == "-") fail();
    if (real) return Json::real(std::stod(tok));
    errno = 0;
    if (neg) {
      const long long v = std::strtoll(tok.c_str(), nullptr, 10);
      if (errno) throw std::runtime_error("GRAPH_INT64_RANGE");
      return Json::integer(v);
    }
    const unsigned long long v = std::strtoull(tok.c_str(), nullptr, 10);
    if (errno) throw std::runtime_error("GRAPH_INT64_RANGE");
    return Json::unsigned_(v);
  }
  const std::string& t_;
  size_t p_ = 0;
};

inline void dump_string(const std::string& s, std::string& out) {
  out += '"';
  for (unsigned char c : s) {
    switch (c) {
      case '"': out += "\\\""; break;
      case '\\': out += "\\\\"; break;
      case '\n': out += "\\n"; break;
      case '\r': out += "\\r"; break;
      case '\t': out += "\\t"; break;
      case '\b': out += "\\b"; break;
      case '\f': out += "\\f"; break;
      default:
        if (c < 0x20 || c >= 0x7f) {  // ensure_ascii for bytes outside printable ASCII
          char buf[8];
          std::snprintf(buf, sizeof buf, "\\u%04x", c);
          out += buf;
        } else {
          out += static_cast<char>(c);
        }
    }
  }
  out += '"';
}

inline void dump(const Json& j, std::string& out) {
  switch (j.kind) {
    case Json::Null: out += "null"; break;
    case Json::Bool: out += j.b ? "true" : "false"; break;
    case Json::Int: out += std::to_string(j.i); break;
    case Json::UInt: out += std::to_string(j.u); break;
    case Json::Real: {
      char buf[40];
      std::snprintf(buf, sizeof buf, "%.17g", j.d);
      out += buf;
      break;
    }
    case Json::Str: dump_string(j.s, out); break;
    case Json::Arr:
      out += '[';
      for (size_t k = 0; k < j.a.size(); ++k) {
        if (k) out += ',';
        dump(j.a[k], out);
      }
      out += ']';
      break;
    case Json::Obj: {
      out += '{';
      bool first = true;
      for (const auto& kv : j.o) {
        if (!first) out += ',';
        first = false;
        dump_string(kv.first, out);
        out += ':';
        dump(kv.second, out);
      }
      out += '}';
      break;
    }
  }
}

inline std::string dumps(const Json& j) {
  std::string out;
  dump(j, out);
  return out;
}

// ----------------------------------------------------------------- SHA-256

class Sha256 {
 public:
  Sha256() { reset(); }
  void reset() {
    static const uint32_t init[8] = {0x6a09e667, 0xbb67ae85, 0x3c6ef372, 0xa54ff53a,
                                     0x510e527f, 0x9b05688c, 0x1f83d9ab, 0x5be0cd19};
    std::memcpy(h_, init, sizeof h_);
    len_ = 0;
    fill_ = 0;
  }
  void update(const void* data, size_t n) {
    const uint8_t* p = static_cast<const uint8_t*>(data);
    len_ += n;
    while (n) {
      const size_t take = std::min<size_t>(64 - fill_, n);
      std::memcpy(buf_ + fill_, p, take);
      fill_ += take;
      p += take;
      n -= take;
      if (fill_ == 64) {
        block(buf_);
        fill_ = 0;
      }
    }
  }
  // Finishes the hash (call once).
  void bytes(uint8_t out[32]) {
    uint8_t pad[72] = {0x80};
    const uint64_t bits = len_ * 8;
    const size_t padlen = (fill_ < 56) ? 56 - fill_ : 120 - fill_;
    uint8_t lenb[8];
    for (int k = 0; k < 8; ++k) lenb[k] = static_cast<uint8_t>(bits >> (56 - 8 * k));
    update(pad, padlen);
    update(lenb, 8);
    for (int w = 0; w < 8; ++w)
      for (int k = 0; k < 4; ++k) out[4 * w + k] = static_cast<uint8_t>(h_[w] >> (24 - 8 * k));
  }
  std::string hex() {
    uint8_t b[32];
    bytes(b);
    static const char* digits = "0123456789abcdef";
    std::string out;
    for (uint8_t byte : b) {
      out += digits[byte >> 4];
      out += digits[byte & 15];
    }
    return out;
  }

 private:
  static uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }
  void block(const uint8_t* b) {
    static const uint32_t k[64] = {
        0x428a2f98, 0x71374491, 0xb5c0fbcf, 0xe9b5dba5, 0x3956c25b, 0x59f111f1, 0x923f82a4, 0xab1c5ed5,
        0xd807aa98, 0x12835b01, 0x243185be, 0x550c7dc3, 0x72be5d74, 0x80deb1fe, 0x9bdc06a7, 0xc19bf174,
        0xe49b69c1, 0xefbe4786, 0x0fc19dc6, 0x240ca1cc, 0x2de92c6f, 0x4a7484aa, 0x5cb0a9dc, 0x76f988da,
        0x983e5152, 0xa831c66d, 0xb00327c8, 0xbf597fc7, 0xc6e00bf3, 0xd5a79147, 0x06ca6351, 0x14292967,
        0x27b70a85, 0x2e1b2138, 0x4d2c6dfc, 0x53380d13, 0x650a7354, 0x766a0abb, 0x81c2c92e, 0x92722c85,
        0xa2bfe8a1, 0xa81a664b, 0xc24b8b70, 0xc76c51a3, 0xd192e819, 0xd6990624, 0xf40e3585, 0x106aa070,
        0x19a4c116, 0x1e376c08, 0x2748774c, 0x34b0bcb5, 0x391c0cb3, 0x4ed8aa4a, 0x5b9cca4f, 0x682e6ff3,
        0x748f82ee, 0x78a5636f, 0x84c87814, 0x8cc70208, 0x90befffa, 0xa4506ceb, 0xbef9a3f7, 0xc67178f2};
    uint32_t w[64];
    for (int t = 0; t < 16; ++t)
      w[t] = (uint32_t(b[4 * t]) << 24) | (uint32_t(b[4 * t + 1]) << 16) | (uint32_t(b[4 * t + 2]) << 8) | b[4 * t + 3];
    for (int t = 16; t < 64; ++t) {
      const uint32_t s0 = rotr(w[t - 15], 7) ^ rotr(w[t - 15], 18) ^ (w[t - 15] >> 3);
      const uint32_t s1 = rotr(w[t - 2], 17) ^ rotr(w[t - 2], 19) ^ (w[t - 2] >> 10);
      w[t] = w[t - 16] + s0 + w[t - 7] + s1;
    }
    uint32_t a = h_[0], bb = h_[1], c = h_[2], d = h_[3], e = h_[4], f = h_[5], g = h_[6], hh = h_[7];
    for (int t = 0; t < 64; ++t) {
      const uint32_t S1 = rotr(e, 6) ^ rotr(e, 11) ^ rotr(e, 25);
      const uint32_t ch = (e & f) ^ (~e & g);
      const uint32_t t1 = hh + S1 + ch + k[t] + w[t];
      const uint32_t S0 = rotr(a, 2) ^ rotr(a, 13) ^ rotr(a, 22);
      const uint32_t mj = (a & bb) ^ (a & c) ^ (bb & c);
      const uint32_t t2 = S0 + mj;
      hh = g; g = f; f = e; e = d + t1; d = c; c = bb; bb = a; a = t1 + t2;
    }
    h_[0] += a; h_[1] += bb; h_[2] += c; h_[3] += d; h_[4] += e; h_[5] += f; h_[6] += g; h_[7] += hh;
  }
  uint32_t h_[8];
  uint64_t len_;
  uint8_t buf_[64];
  size_t fill_;
};

inline std::string sha256_hex(const std::string& s) {
  Sha256 h;
  h.update(s.data(), s.size());
  return h.hex();
}

}  // namespace mgbfs::host
