// mgbfs — command-line host of the MI355X exact BFS engine.
//
// Command grammar and output contract of the reference executable
// (crates/mgbfs-cli/src/main.rs:20-199, crates/mgbfs-runtime/src/generic_run.rs:9-43,
// generic_distributed_run.rs:13-73):
//   mgbfs graph <graph.json> <out-dir> [--device N] [--capacity N] [--seconds N] [--backend NAME]
//   mgbfs graph-info <graph.json> <devices> <shards> [capacity|auto] [batch]
//   mgbfs graph-plan <graph.json> <inventory.json> <shards> [capacity|auto] [batch]
//   mgbfs graph-local-plan <graph.json> <device> <world> <shards> <capacity|auto> <batch>
//   mgbfs graph-rank <graph.json> <launch.json> <rank> <bootstrap> <out-dir> <seconds>   (world 1)
//   mgbfs graph-count
//   mgbfs graph-digest <graph.json>      (validation, digest, retention rule; no device)
//   mgbfs key-info
//   mgbfs version
// Errors: nonzero exit and a last stderr line {"status":"ERROR","error":"CODE"}.
// `graph` writes report.json (exact layer sizes) and states.json (<=1000 rows
// of the last layer, the previous layer iff it has < 1000 rows); SIGTERM or
// SIGINT ends the run as INCOMPLETE/CANCELLED with the completed prefix.
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <chrono>
#include <csignal>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <iostream>
#include <sstream>
#include <string>
#include <sys/stat.h>
#include <unistd.h>
#include <vector>
#include <hip/hip_runtime.h>
#include "../../include/mgbfs_hip.h"
#include "archive.hpp"
#include "host_util.hpp"

using mgbfs::host::Json;
using mgbfs::host::LayerArchive;

namespace {

volatile std::sig_atomic_t g_cancel = 0;
extern "C" void on_signal(int) { g_cancel = 1; }

struct CliError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

const int64_t kI64Max = INT64_MAX;
const uint64_t kReserveBytes = 2ull << 30;  // cold admission reserve (generic_memory.rs:9-19)

int64_t wrap_mul_add(int64_t acc, int64_t a, int64_t b) {
  return static_cast<int64_t>(static_cast<uint64_t>(acc) + static_cast<uint64_t>(a) * static_cast<uint64_t>(b));
}

int64_t pymod(int64_t x, int64_t m) {
  int64_t r = x % m;
  return r < 0 ? r + m : r;
}

// Validated schema-2 graph (multigpubfs/graph_definition.py:46-69).
struct GraphDoc {
  Json raw;
  uint32_t kind = 0, elements = 0, rows = 0, cols = 0;
  std::vector<std::vector<uint32_t>> perms;
  std::vector<std::vector<int64_t>> mats;
  std::vector<uint32_t> moduli;
  std::vector<int64_t> start;
  uint64_t expected = 0;

  static std::vector<int64_t> ints(const Json& j) {
    if (j.kind != Json::Arr) throw CliError("GRAPH_VECTOR");
    std::vector<int64_t> v;
    for (const Json& x : j.a) {
      if (!x.is_int()) throw CliError("GRAPH_INT64_RANGE");
      v.push_back(x.as_int());
    }
    return v;
  }

  explicit GraphDoc(const Json& j) : raw(j) {
    static const char* keys[] = {"action", "expected_max_unique_states", "generator_names", "name", "schema", "start"};
    if (j.kind != Json::Obj || j.o.size() != 6) throw CliError("GRAPH_SCHEMA");
    for (const char* k : keys)
      if (!j.has(k)) throw CliError("GRAPH_SCHEMA");
    if (!j.at("schema").is_int() || j.at("schema").as_int() != 2) throw CliError("GRAPH_SCHEMA");
    if (j.at("name").kind != Json::Str) throw CliError("GRAPH_NAME");
    start = ints(j.at("start"));
    const Json& a = j.at("action");
    if (a.kind != Json::Obj || !a.has("generators") || a.at("generators").kind != Json::Arr ||
        a.at("generators").a.empty())
      throw CliError("GRAPH_GENERATORS");
    const Json& names = j.at("generator_names");
    if (names.kind != Json::Arr || names.a.size() != a.at("generators").a.size()) throw CliError("GRAPH_GENERATOR_NAMES");
    for (const Json& n : names.a)
      if (n.kind != Json::Str) throw CliError("GRAPH_GENERATOR_NAMES");
    const Json& bound = j.at("expected_max_unique_states");
    if (bound.kind != Json::Null) {
      if (bound.kind == Json::UInt && bound.u >= 1) expected = bound.u;
      else if (bound.kind == Json::Int && bound.i >= 1) expected = static_cast<uint64_t>(bound.i);
      else throw CliError("GRAPH_ORBIT_BOUND");
    }
    const std::string kind_s = a.has("kind") && a.at("kind").kind == Json::Str ? a.at("kind").s : "";
    if (kind_s == "permutation") {
      if (a.o.size() != 3 || !a.has("degree") || !a.at("degree").is_int()) throw CliError("GRAPH_PERMUTATION_SHAPE");
      const int64_t n = a.at("degree").as_int();
      if (n < 1 || n > 0xffffffffll || static_cast<size_t>(n) != start.size()) throw CliError("GRAPH_PERMUTATION_SHAPE");
      kind = 0;
      elements = rows = static_cast<uint32_t>(n);
      cols = 1;
      for (const Json& gj : a.at("generators").a) {
        std::vector<int64_t> g = ints(gj);
        std::vector<int64_t> sorted_g = g;
        std::sort(sorted_g.begin(), sorted_g.end());
        for (int64_t i = 0; i < n; ++i)
          if (static_cast<int64_t>(sorted_g.size()) != n || sorted_g[i] != i) throw CliError("GRAPH_PERMUTATION_GENERATOR");
        perms.emplace_back(g.begin(), g.end());
      }
    } else if (kind_s == "matrix") {
      if (a.o.size() != 4 || !a.has("rows") || !a.has("cols") || !a.at("rows").is_int() || !a.at("cols").is_int())
        throw CliError("GRAPH_MATRIX_SHAPE");
      const int64_t r = a.at("rows").as_int(), c = a.at("cols").as_int();
      if (r < 1 || c < 1 || r * c > 0xffffffffll || static_cast<size_t>(r * c) != start.size())
        throw CliError("GRAPH_MATRIX_SHAPE");
      kind = 1;
      rows = static_cast<uint32_t>(r);
      cols = static_cast<uint32_t>(c);
      elements = rows * cols;
      for (const Json& gj : a.at("generators").a) {
        if (gj.kind != Json::Obj || gj.o.size() != 2 || !gj.has("matrix") || !gj.has("modulo"))
          throw CliError("GRAPH_MATRIX_GENERATOR");
        std::vector<int64_t> m = ints(gj.at("matrix"));
        if (m.size() != static_cast<size_t>(r * r)) throw CliError("GRAPH_MATRIX_GENERATOR");
        if (!gj.at("modulo").is_int()) throw CliError("GRAPH_MATRIX_MODULUS");
        const int64_t mod = gj.at("modulo").as_int();
        if (mod != 0 && (mod < 2 || mod > (1ll << 31))) throw CliError("GRAPH_MATRIX_MODULUS");
        mats.push_back(m);
        moduli.push_back(static_cast<uint32_t>(mod));
      }
    } else {
      throw CliError("GRAPH_ACTION_KIND");
    }
  }

  // sha256 of canonical JSON {schema, action, start} (graph_definition.py:43-45).
  std::string digest() const {
    Json core = Json::object();
    core["schema"] = raw.at("schema");
    core["action"] = raw.at("action");
    core["start"] = raw.at("start");
    return mgbfs::host::sha256_hex(mgbfs::host::dumps(core));
  }

  std::vector<int64_t> compose(const std::vector<int64_t>& g, const std::vector<int64_t>& h, int64_t mod) const {
    const uint32_t n = rows;
    std::vector<int64_t> right(h);
    if (mod)
      for (auto& v : right) v = pymod(v, mod);
    std::vector<int64_t> out(n * n);
    for (uint32_t i = 0; i < n; ++i)
      for (uint32_t j = 0; j < n; ++j) {
        int64_t acc = 0;
        for (uint32_t k = 0; k < n; ++k) acc = wrap_mul_add(acc, g[i * n + k], right[k * n + j]);
        out[i * n + j] = mod ? pymod(acc, mod) : acc;
      }
    return out;
  }

  // Retention rule of graph_definition.py:84-108.
  bool inverse_closed() const {
    if (kind == 0) {
      for (const auto& g : perms) {
        std::vector<uint32_t> inv(g.size());
        for (size_t i = 0; i < g.size(); ++i) inv[g[i]] = static_cast<uint32_t>(i);
        if (std::find(perms.begin(), perms.end(), inv) == perms.end()) return false;
      }
      return true;
    }
    for (uint32_t m : moduli)
      if (m != moduli[0]) return false;
    const int64_t mod = moduli[0];
    const uint32_t n = rows;
    if (mod && (mod & (mod - 1))) {
      for (const auto& g : mats)
        for (uint32_t i = 0; i < n; ++i) {
          long double sum = 0;
          for (uint32_t k = 0; k < n; ++k) sum += std::abs(static_cast<long double>(g[i * n + k])) * (mod - 1);
          if (sum > static_cast<long double>(kI64Max)) return false;
        }
    }
    std::vector<int64_t> eye(n * n, 0);
    for (uint32_t i = 0; i < n; ++i) eye[i * n + i] = 1;
    for (const auto& g : mats) {
      bool found = false;
      for (const auto& h : mats)
        if (compose(g, h, mod) == eye && compose(h, g, mod) == eye) {
          found = true;
          break;
        }
      if (!found) return false;
    }
    if (mod)
      for (int64_t v : start)
        if (v < 0 || v >= mod) return false;
    return true;
  }

  mgbfs_graph_desc desc(std::vector<uint32_t>& pflat, std::vector<int64_t>& mflat) const {
    mgbfs_graph_desc d{};
    d.kind = kind;
    d.elements = elements;
    d.rows = rows;
    d.cols = cols;
    d.generators = static_cast<uint32_t>(kind == 0 ? perms.size() : mats.size());
    d.inverse_closed = inverse_closed() ? 1u : 0u;
    for (const auto& g : perms) pflat.insert(pflat.end(), g.begin(), g.end());
    for (const auto& m : mats) mflat.insert(mflat.end(), m.begin(), m.end());
    d.permutations = pflat.empty() ? nullptr : pflat.data();
    d.matrices = mflat.empty() ? nullptr : mflat.data();
    d.moduli = moduli.empty() ? nullptr : moduli.data();
    d.start = start.data();
    d.expected_states = expected;
    return d;
  }
};

std::string read_file(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw CliError("CLI_GRAPH_READ");
  std::stringstream ss;
  ss << f.rdbuf();
  return ss.str();
}

void write_file(const std::string& path, const std::string& text) {
  const std::string tmp = path + ".tmp";
  {
    std::ofstream f(tmp, std::ios::binary);
    if (!f) throw CliError("CLI_OUTPUT_WRITE");
    f << text;
  }
  if (std::rename(tmp.c_str(), path.c_str())) throw CliError("CLI_OUTPUT_WRITE");
}

uint64_t parse_u64(const std::string& s, const char* code) {
  if (s.empty() || !std::all_of(s.begin(), s.end(), [](char c) { return std::isdigit(static_cast<unsigned char>(c)); }))
    throw CliError(code);
  errno = 0;
  const unsigned long long v = std::strtoull(s.c_str(), nullptr, 10);
  if (errno) throw CliError(code);
  return v;
}

uint32_t backend_code(const std::string& name) {
  if (name == "auto") return MGBFS_BACKEND_AUTO;
  if (name == "permutation_bitmap" || name == "bitmap") return MGBFS_BACKEND_BITMAP;
  if (name == "sorted") return MGBFS_BACKEND_SORTED;
  if (name == "segmented") return MGBFS_BACKEND_SEGMENTED;
  if (name == "generic" || name == "legacy") return MGBFS_BACKEND_GENERIC;
  throw CliError("CLI_GRAPH_BACKEND");
}

Json rows_json(const std::vector<int64_t>& flat, uint64_t count, uint32_t width) {
  Json arr = Json::array();
  for (uint64_t r = 0; r < count; ++r) {
    Json row = Json::array();
    for (uint32_t e = 0; e < width; ++e) row.push(Json::integer(flat[r * width + e]));
    arr.push(row);
  }
  return arr;
}

// An engine kept alive between resident-session jobs: a job whose graph,
// device and options match the previous one resets it instead of
// allocating a new one.
struct EngineCache {
  std::string key;
  void* engine = nullptr;
  uint64_t device_bytes = 0;
  mgbfs_engine_plan_info plan{};
  void release() {
    if (engine) mgbfs_engine_destroy(engine);
    engine = nullptr;
    key.clear();
    device_bytes = 0;
  }
  ~EngineCache() { release(); }
};

bool path_exists(const std::string& path) {
  struct stat st;
  return !path.empty() && stat(path.c_str(), &st) == 0;
}

// Reference names of this library's visited-set algorithms and layouts.
const char* history_algorithm(uint32_t backend) {
  switch (backend) {
    case MGBFS_BACKEND_BITMAP: return "PERMUTATION_BITMAP";
    case MGBFS_BACKEND_SORTED: return "SORTED_RUNS";
    case MGBFS_BACKEND_SEGMENTED: return "SEGMENTED_HASH";
    default: return "HASH";
  }
}
const char* table_layout(uint32_t backend) {
  switch (backend) {
    case MGBFS_BACKEND_BITMAP: return "LEHMER_RANK_BITMAP";
    case MGBFS_BACKEND_SORTED: return "SUB_BUCKET_LDS_TABLES";
    case MGBFS_BACKEND_SEGMENTED: return "BANK_SEGMENTS_U64_SLOTS";
    default: return "OPEN_ADDRESSING_U64_SLOTS";
  }
}
const char* backend_key(uint32_t backend) {
  switch (backend) {
    case MGBFS_BACKEND_BITMAP: return "permutation_bitmap";
    case MGBFS_BACKEND_SORTED: return "sorted";
    case MGBFS_BACKEND_SEGMENTED: return "segmented";
    default: return "generic";
  }
}

// Generator table bytes as the reference counts them (generic_distributed_run.rs:10).
uint64_t generator_bytes(const GraphDoc& g) {
  uint64_t b = 0;
  for (const auto& p : g.perms) b += 4 * p.size();
  for (const auto& m : g.mats) b += 8 * m.size() + 4;
  return b;
}

// A rank plan with the fields of GenericDistributedMemoryPlan
// (crates/mgbfs-runtime/src/generic_distributed_memory.rs:7-13), from this
// library's engine plan.
Json plan_json(const GraphDoc& graph, const mgbfs_engine_plan_info& p, uint32_t world, uint32_t shards) {
  Json plan = Json::object();
  plan["generator_backend"] = Json::str("hip");
  plan["history_algorithm"] = Json::str(history_algorithm(p.backend));
  plan["owner_lanes"] = Json::integer(0);
  plan["sorted_owner_bytes"] = Json::integer(0);
  plan["driver_graph_reserve_bytes"] = Json::integer(0);
  plan["parent_transport"] = Json::boolean(false);
  plan["sort_candidates"] = Json::boolean(p.backend == MGBFS_BACKEND_SORTED);
  plan["history_layers"] = Json::integer(p.history_layers);
  plan["retained_banks"] = Json::integer(p.retained_banks);
  plan["state_bytes"] = Json::integer(p.state_bytes);
  plan["elements"] = Json::integer(graph.elements);
  plan["world"] = Json::integer(world);
  plan["shards"] = Json::integer(shards);
  plan["capacity"] = Json::unsigned_(p.capacity);
  plan["batch"] = Json::unsigned_(p.batch);
  plan["table_layout"] = Json::str(table_layout(p.backend));
  plan["queue_capacity"] = Json::unsigned_(p.queue_capacity);
  plan["table_slots"] = Json::unsigned_(p.table_slots);
  plan["generator_bytes"] = Json::unsigned_(generator_bytes(graph));
  plan["device_bytes"] = Json::unsigned_(p.device_bytes);
  plan["state_bound"] = Json::unsigned_(p.state_bound);
  plan["backend"] = Json::str(backend_key(p.backend));
  return plan;
}

void write_job_phase(const std::string& path, const char* phase) {
  if (path.empty()) return;
  const double unix_s = std::chrono::duration<double>(std::chrono::system_clock::now().time_since_epoch()).count();
  Json v = Json::object();
  if (path_exists(path)) {
    try {
      v = mgbfs::host::JsonParser(read_file(path)).parse();
    } catch (const std::exception&) {
      v = Json::object();
    }
  }
  if (v.kind != Json::Obj || !v.has("entered")) {
    v = Json::object();
    v["schema"] = Json::integer(1);
    v["pid"] = Json::integer(static_cast<int64_t>(getpid()));
    const char* seq = std::getenv("MGBFS_JOB_SEQUENCE");
    v["sequence"] = Json::str(seq ? seq : "");
    v["entered"] = Json::object();
  }
  v["phase"] = Json::str(phase);
  v["entered"][phase] = Json::real(unix_s);
  v["updated_unix"] = Json::real(unix_s);
  write_file(path, mgbfs::host::dumps(v));
}

// What one BFS run asks for.
struct BfsRequest {
  int32_t device = 0;
  uint64_t capacity = 0, seconds = 0;
  uint32_t batch = 0, flags = 0;
  std::string backend = "auto", archive_path, phase_path;
};

// What it produced: layer table, stop reason, timings, engine figures and
// the states.json text.
struct BfsOutcome {
  std::vector<uint64_t> sizes;
  std::vector<double> times;
  std::string status = "COMPLETE", reason = "FRONTIER_EXHAUSTED", backend;
  double setup_seconds = 0, bfs_seconds = 0;
  mgbfs_engine_info info{};
  mgbfs_engine_plan_info plan{};
  std::string states_text;
};

// Runs one BFS to completion, a stop (deadline, cancel, capacity) or a
// failure.  With a cache the engine survives the run; cancel_path (if it
// appears between layers) ends the run INCOMPLETE / CANCELLED like SIGTERM.
// sample_key names the states.json sample-limit field.
BfsOutcome run_bfs(const GraphDoc& graph, const BfsRequest& req, EngineCache* cache, const std::string& cancel_path,
                   bool* reused, const char* sample_key) {
  BfsOutcome r;
  std::vector<uint32_t> pflat;
  std::vector<int64_t> mflat;
  const mgbfs_graph_desc desc = graph.desc(pflat, mflat);
  mgbfs_engine_options opts{};
  opts.device = req.device;
  opts.backend = backend_code(req.backend);
  opts.capacity = req.capacity;
  opts.flags = req.flags;
  opts.batch = req.batch;
  const std::string key = graph.digest() + "|" + std::to_string(req.device) + "|" + req.backend + "|" +
                          std::to_string(req.capacity) + "|" + std::to_string(req.flags) + "|" +
                          std::to_string(req.batch);
  void* engine = nullptr;
  char err[256] = {0};
  const auto setup0 = std::chrono::steady_clock::now();
  if (reused) *reused = false;
  if (cache && cache->engine && cache->key == key) {
    if (mgbfs_engine_reset(cache->engine)) {
      cache->release();
      throw CliError("ENGINE_RESET");
    }
    engine = cache->engine;
    if (reused) *reused = true;
  } else {
    if (cache) cache->release();
    if (mgbfs_engine_plan(&desc, &opts, &r.plan, err, sizeof err)) throw CliError(err[0] ? err : "ENGINE_PLAN");
    if (mgbfs_engine_create(&desc, &opts, &engine, err, sizeof err)) throw CliError(err[0] ? err : "ENGINE_CREATE");
    if (cache) {
      cache->engine = engine;
      cache->key = key;
      cache->plan = r.plan;
    }
  }
  if (cache) r.plan = cache->plan;
  r.setup_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - setup0).count();
  std::signal(SIGTERM, on_signal);
  std::signal(SIGINT, on_signal);
  std::cerr << "MGBFS_GRAPH_READY device=" << req.device << std::endl;
  write_job_phase(req.phase_path, "bfs");
  std::unique_ptr<LayerArchive> archive;
  std::vector<int64_t> arows(1000 * static_cast<size_t>(graph.elements));
  auto archive_layer = [&](uint64_t depth) {
    uint64_t n = 0;
    if (mgbfs_engine_sample(engine, 0, 1000, arows.data(), &n)) throw CliError("ARCHIVE_SAMPLE");
    archive->layer(depth, arows, n);
  };
  if (!req.archive_path.empty()) {
    if (graph.kind != 0) throw CliError("ARCHIVE_STATE_WIDTH");
    archive = std::make_unique<LayerArchive>(req.archive_path, graph.elements, graph.digest());
    archive_layer(0);
  }
  const auto t0 = std::chrono::steady_clock::now();
  while (true) {
    if (g_cancel || path_exists(cancel_path)) {
      r.status = "INCOMPLETE";
      r.reason = "CANCELLED";
      break;
    }
    if (req.seconds && std::chrono::steady_clock::now() - t0 >= std::chrono::seconds(req.seconds)) {
      r.status = "INCOMPLETE";
      r.reason = "DEADLINE";
      break;
    }
    const auto l0 = std::chrono::steady_clock::now();
    uint64_t next = 0;
    uint32_t fatal = 0;
    const int s = mgbfs_engine_advance(engine, &next, &fatal);
    if (s < 0) {
      if (cache) cache->release();
      else mgbfs_engine_destroy(engine);
      throw CliError(std::string("ENGINE_ADVANCE: ") + mgbfs_engine_backend_name(nullptr));
    }
    if (s == MGBFS_LAYER) {
      r.times.push_back(std::chrono::duration<double>(std::chrono::steady_clock::now() - l0).count());
      if (archive) archive_layer(r.times.size());
      continue;
    }
    if (s == MGBFS_RESOURCE) {
      r.status = "INCOMPLETE";
      r.reason = "RESOURCE_" + std::to_string(fatal);
    }
    break;
  }
  r.bfs_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  write_job_phase(req.phase_path, "retiring");
  if (archive) archive->commit();
  uint32_t nl = 0;
  mgbfs_engine_layer_sizes(engine, nullptr, 0, &nl);
  r.sizes.resize(nl);
  mgbfs_engine_layer_sizes(engine, r.sizes.data(), nl, &nl);
  mgbfs_engine_query(engine, &r.info);
  r.backend = mgbfs_engine_backend_name(engine);
  // Samples: <= 1000 rows of the last completed layer; the previous layer only if < 1000.
  std::vector<int64_t> cur(1000 * static_cast<size_t>(graph.elements)), prev;
  uint64_t ncur = 0, nprev = 0;
  if (mgbfs_engine_sample(engine, 0, 1000, cur.data(), &ncur)) ncur = 0;
  if (nl >= 2 && r.sizes[nl - 2] < 1000) {
    prev.resize(1000 * static_cast<size_t>(graph.elements));
    if (mgbfs_engine_sample(engine, 1, 1000, prev.data(), &nprev)) nprev = 0;
  }
  Json states = Json::object();
  states["schema"] = Json::integer(2);
  states["state_encoding"] = Json::str("signed_int64_vectors");
  states["current"] = rows_json(cur, ncur, graph.elements);
  states["previous_small"] = rows_json(prev, nprev, graph.elements);
  states[sample_key] = Json::integer(1000);
  r.states_text = mgbfs::host::dumps(states);
  if (cache) cache->device_bytes = r.info.device_bytes;
  else mgbfs_engine_destroy(engine);
  return r;
}

Json layer_json(const BfsOutcome& r, Json& seconds) {
  Json ls = Json::array();
  for (uint64_t v : r.sizes) ls.push(Json::unsigned_(v));
  seconds = Json::array();
  for (double v : r.times) seconds.push(Json::real(v));
  return ls;
}

// One `graph` job: writes <out>/states.json and <out>/report.json and returns
// the report.  default_device applies when no --device is given.
Json run_graph_job(const std::vector<std::string>& args, int32_t default_device, EngineCache* cache,
                   const std::string& cancel_path, bool* reused) {
  if (args.size() < 2)
    throw CliError(
        "CLI_USAGE: mgbfs graph <graph.json> <output-dir> [--device N] [--capacity N] [--seconds N] "
        "[--archive PATH]");
  BfsRequest req;
  req.device = default_device;
  for (size_t i = 2; i < args.size(); i += 2) {
    if (i + 1 >= args.size()) throw CliError("CLI_GRAPH_OPTION_VALUE");
    const std::string& k = args[i];
    if (k == "--device") req.device = static_cast<int32_t>(parse_u64(args[i + 1], "CLI_GRAPH_DEVICE"));
    else if (k == "--capacity") req.capacity = parse_u64(args[i + 1], "CLI_GRAPH_CAPACITY");
    else if (k == "--seconds") req.seconds = parse_u64(args[i + 1], "CLI_GRAPH_SECONDS");
    else if (k == "--backend") req.backend = args[i + 1];
    else if (k == "--archive") req.archive_path = args[i + 1];
    else if (k == "--memory-preset") {
      if (args[i + 1] == "max_capacity") req.flags |= MGBFS_FLAG_MAX_CAPACITY;
      else if (args[i + 1] != "balanced") throw CliError("INVALID_MEMORY_PRESET");
    } else throw CliError("CLI_GRAPH_UNKNOWN_OPTION");
  }
  if (cache && req.device != default_device) throw CliError("GENERIC_SESSION_DEVICE");
  const char* world = std::getenv("WORLD_SIZE");
  if (world && std::string(world) != "1") throw CliError("GENERIC_DISTRIBUTED_INTEGRATION_NOT_READY");
  GraphDoc graph(mgbfs::host::JsonParser(read_file(args[0])).parse());
  const std::string out = args[1];
  if (path_exists(out)) throw CliError("CLI_GRAPH_OUTPUT_EXISTS");
  if (mkdir(out.c_str(), 0755)) throw CliError("CLI_GRAPH_OUTPUT_CREATE");
  const BfsOutcome r = run_bfs(graph, req, cache, cancel_path, reused, "current_sample_limit");

  Json report = Json::object();
  report["schema"] = Json::integer(2);
  report["status"] = Json::str(r.status);
  report["reason"] = Json::str(r.reason);
  report["backend"] = Json::str(r.backend);
  report["device"] = Json::integer(req.device);
  report["world"] = Json::integer(1);
  report["graph_digest"] = Json::str(graph.digest());
  Json seconds;
  report["layer_sizes"] = layer_json(r, seconds);
  report["layer_seconds"] = seconds;
  report["setup_seconds"] = Json::real(r.setup_seconds);
  report["bfs_seconds"] = Json::real(r.bfs_seconds);
  report["capacity"] = Json::unsigned_(r.plan.capacity);
  report["requested_capacity"] = Json::unsigned_(req.capacity);
  report["arena_bytes"] = Json::unsigned_(r.info.device_bytes);
  report["plan"] = plan_json(graph, r.plan, 1, 1);
  Json devices = Json::array();
  devices.push(Json::integer(req.device));
  report["devices"] = devices;
  report["state_bytes"] = Json::integer(r.info.state_bytes);
  report["states_sha256"] = Json::str(mgbfs::host::sha256_hex(r.states_text));
  report["candidates_generated"] = Json::unsigned_(r.info.candidates);  // after back-edge pruning
  report["kernel_seconds"] = Json::real(r.info.kernel_seconds);
  report["retention"] = Json::str("current sample <=1000 and previous layer only if <1000");
  report["scope"] = Json::str("single-device MI355X engine");
  write_file(out + "/states.json", r.states_text);
  write_file(out + "/report.json", mgbfs::host::dumps(report));
  return report;
}

int cmd_graph(const std::vector<std::string>& args) {
  const Json report = run_graph_job(args, 0, nullptr, "", nullptr);
  std::cout << mgbfs::host::dumps(report) << std::endl;
  return 0;
}

// ------------------------------------------------------------ resident session

std::string seq_name(const std::string& root, const char* stem, uint64_t sequence, const char* ext) {
  char buf[32];
  std::snprintf(buf, sizeof buf, "%08llu", static_cast<unsigned long long>(sequence));
  return root + "/" + stem + "-" + buf + ext;
}

void write_atomically(const std::string& path, const std::string& text) {
  const std::string tmp = path + ".tmp";
  write_file(tmp, text);
  if (std::rename(tmp.c_str(), path.c_str())) throw CliError("GENERIC_SESSION_WRITE");
}

void set_phase(const std::string& path, uint64_t sequence, const char* phase) {
  if (path.empty()) return;
  Json v = Json::object();
  v["schema"] = Json::integer(1);
  v["pid"] = Json::integer(static_cast<int64_t>(getpid()));
  v["sequence"] = Json::str(std::to_string(sequence));
  v["phase"] = Json::str(phase);
  write_atomically(path, mgbfs::host::dumps(v));
}

Json inventory_json(int32_t device, const EngineCache& cache) {
  uint32_t cus = 0, wave = 0, lds = 0;
  uint64_t total = 0, free_bytes = 0;
  if (mgbfs_hip_device_info(device, &cus, &wave, &lds, &total, &free_bytes)) throw CliError("GENERIC_SESSION_INVENTORY");
  Json v = Json::object();
  v["device"] = Json::integer(device);
  v["free_bytes"] = Json::unsigned_(std::min(total, free_bytes + cache.device_bytes));
  v["total_bytes"] = Json::unsigned_(total);
  v["cached_reusable_bytes"] = Json::unsigned_(cache.device_bytes);
  v["additional_reserve_bytes"] = Json::unsigned_(2ull << 30);
  v["sm_count"] = Json::integer(cus);
  v["warp_size"] = Json::integer(wave);
  v["shared_bytes_per_block"] = Json::integer(lds);
  return v;
}

// Resident worker (multigpubfs/generic_session.py protocol, schema 1): jobs
// arrive as <root>/job-NNNNNNNN.json {schema, sequence, command, args, env};
// each gets <root>/response-NNNNNNNN.json {schema, sequence, ok, result |
// error, cache}.  A `shutdown` file ends the loop.  Commands: graph (the
// `mgbfs graph` arguments; the device is the worker's), inventory,
// release-memory.  Job env keys must start with MGBFS_ or NCCL_ and are
// undone before the next job.  After a failed job the worker exits: a rank
// is never reused after an error.
int cmd_graph_session(const std::vector<std::string>& args) {
  if (args.size() != 3) throw CliError("CLI_USAGE: mgbfs graph-session <root> <rank> <device>");
  const std::string root = args[0];
  const uint64_t rank = parse_u64(args[1], "GENERIC_SESSION_RANK");
  const int32_t device = static_cast<int32_t>(parse_u64(args[2], "GENERIC_SESSION_DEVICE"));
  (void)rank;
  EngineCache cache;
  std::map<std::string, std::pair<bool, std::string>> previous;  // name -> (was set, old value)
  std::signal(SIGTERM, on_signal);
  std::signal(SIGINT, on_signal);
  for (uint64_t sequence = 0;; ++sequence) {
    const std::string job_path = seq_name(root, "job", sequence, ".json");
    const std::string shutdown = root + "/shutdown";
    uint32_t nap_us = 200;
    while (!path_exists(job_path)) {
      if (path_exists(shutdown) || g_cancel) return 0;
      usleep(nap_us);
      nap_us = std::min<uint32_t>(nap_us * 2, 20000);
    }
    bool ok = true;
    std::string error;
    Json result = Json::object();
    std::string phase_path;
    try {
      const Json job = mgbfs::host::JsonParser(read_file(job_path)).parse();
      if (!job.has("schema") || !job.at("schema").is_int() || job.at("schema").as_int() != 1 ||
          !job.has("sequence") || !job.at("sequence").is_int() ||
          job.at("sequence").as_int() != static_cast<int64_t>(sequence))
        throw CliError("GENERIC_SESSION_SEQUENCE");
      for (const auto& kv : previous) {
        if (kv.second.first) setenv(kv.first.c_str(), kv.second.second.c_str(), 1);
        else unsetenv(kv.first.c_str());
      }
      previous.clear();
      if (job.has("env")) {
        const Json& env = job.at("env");
        if (env.kind != Json::Obj) throw CliError("GENERIC_SESSION_ENV");
        for (const auto& kv : env.o)
          if (kv.second.kind != Json::Str || (kv.first.rfind("MGBFS_", 0) != 0 && kv.first.rfind("NCCL_", 0) != 0))
            throw CliError("GENERIC_SESSION_ENV");
        for (const auto& kv : env.o) {
          const char* old = std::getenv(kv.first.c_str());
          previous[kv.first] = {old != nullptr, old ? old : ""};
          setenv(kv.first.c_str(), kv.second.s.c_str(), 1);
        }
      }
      if (const char* p = std::getenv("MGBFS_JOB_PHASE_PATH")) phase_path = p;
      set_phase(phase_path, sequence, "running");
      const std::string command = job.has("command") && job.at("command").kind == Json::Str ? job.at("command").s : "";
      std::vector<std::string> cargs;
      if (job.has("args")) {
        if (job.at("args").kind != Json::Arr) throw CliError("GENERIC_SESSION_ARGS");
        for (const Json& a : job.at("args").a) {
          if (a.kind != Json::Str) throw CliError("GENERIC_SESSION_ARGS");
          cargs.push_back(a.s);
        }
      }
      if (command == "graph") {
        bool reused = false;
        const Json report = run_graph_job(cargs, device, &cache, seq_name(root, "cancel", sequence, ""), &reused);
        result["pid"] = Json::integer(static_cast<int64_t>(getpid()));
        result["status"] = report.at("status");
        result["reason"] = report.at("reason");
        result["output"] = Json::str(cargs.size() > 1 ? cargs[1] : "");
        result["engine_reused"] = Json::boolean(reused);
      } else if (command == "inventory") {
        result = inventory_json(device, cache);
      } else if (command == "release-memory") {
        cache.release();
        result["released"] = Json::boolean(true);
      } else {
        throw CliError("GENERIC_SESSION_COMMAND");
      }
    } catch (const std::exception& e) {
      ok = false;
      error = e.what();
    }
    Json response = Json::object();
    response["schema"] = Json::integer(1);
    response["sequence"] = Json::unsigned_(sequence);
    response["ok"] = Json::boolean(ok);
    if (ok) response["result"] = result;
    else response["error"] = Json::str(error);
    Json cache_stats = Json::object();
    cache_stats["cached_engines"] = Json::integer(cache.engine ? 1 : 0);
    cache_stats["cached_bytes"] = Json::unsigned_(cache.device_bytes);
    response["cache"] = cache_stats;
    try {
      set_phase(phase_path, sequence, ok ? "done" : "failed");
    } catch (const std::exception&) {
    }
    write_atomically(seq_name(root, "response", sequence, ".json"), mgbfs::host::dumps(response));
    if (!ok) {
      cache.release();
      Json err = Json::object();
      err["status"] = Json::str("ERROR");
      err["error"] = Json::str(error);
      std::cerr << mgbfs::host::dumps(err) << std::endl;
      return 1;
    }
  }
}

int cmd_graph_count() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  std::cout << n << std::endl;
  return 0;
}

// Device list of graph-info: "auto" = every visible device, else a comma
// list of distinct visible ordinals (generic_distributed_run.rs:24-28).
std::vector<int32_t> parse_devices(const std::string& spec) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) throw CliError("NO_VISIBLE_HIP_DEVICES");
  std::vector<int32_t> devs;
  if (spec == "auto") {
    for (int d = 0; d < n; ++d) devs.push_back(d);
  } else {
    std::stringstream ss(spec);
    std::string tok;
    while (std::getline(ss, tok, ',')) devs.push_back(static_cast<int32_t>(parse_u64(tok, "CLI_GRAPH_DEVICES")));
  }
  bool bad = devs.empty() || devs.size() > 128;
  for (size_t i = 0; i < devs.size() && !bad; ++i)
    bad = devs[i] >= n || std::find(devs.begin(), devs.begin() + i, devs[i]) != devs.begin() + i;
  if (bad) throw CliError("CLI_GRAPH_DEVICE_SELECTION");
  return devs;
}

Json inventory_entry(int32_t device) {
  uint32_t cu = 0, ws = 0, lds = 0;
  uint64_t tot = 0, fr = 0, hw[8] = {0};
  if (mgbfs_hip_device_info(device, &cu, &ws, &lds, &tot, &fr)) throw CliError("GRAPH_INFO_DEVICE");
  if (mgbfs_generic_hardware_info(device, hw)) throw CliError("GRAPH_INFO_HARDWARE");
  Json e = Json::object();
  e["device"] = Json::integer(device);
  e["free_bytes"] = Json::unsigned_(fr);
  e["total_bytes"] = Json::unsigned_(tot);
  static const char* names[8] = {"sm_count",     "compute_major",  "compute_minor",     "l2_bytes",
                                 "warp_size",    "threads_per_sm", "threads_per_block", "shared_bytes_per_sm"};
  for (int i = 0; i < 8; ++i) e[names[i]] = Json::unsigned_(hw[i]);
  return e;
}

// Admission (generic_distributed_run.rs:33-73): one rank plan per inventory
// entry, sized against that entry's free bytes less the 2 GiB reserve (a
// live query when budgets are 0), the smallest as `plan`, owner cut points
// proportional to the rank capacities (none when a capacity was requested).
// A requested capacity or batch that a budget cannot hold is refused with
// the reference's codes.
Json admission(const GraphDoc& graph, const Json& inventory, const std::vector<uint64_t>& budgets, uint64_t shards,
               uint64_t requested, uint64_t batch) {
  std::vector<uint32_t> pflat;
  std::vector<int64_t> mflat;
  const mgbfs_graph_desc desc = graph.desc(pflat, mflat);
  const uint32_t world = static_cast<uint32_t>(inventory.a.size());
  Json out = Json::object();
  out["size_profile_capability"] = Json::integer(1);
  out["graph_digest"] = Json::str(graph.digest());
  Json dj = Json::array(), plans = Json::array();
  std::vector<mgbfs_engine_plan_info> infos;
  size_t smallest = 0;
  for (size_t i = 0; i < inventory.a.size(); ++i) {
    const Json& entry = inventory.a[i];
    mgbfs_engine_options opts{};
    opts.device = static_cast<int32_t>(entry.at("device").as_int());
    opts.capacity = requested;
    opts.vram_budget = budgets[i];
    opts.batch = static_cast<uint32_t>(std::min<uint64_t>(batch, 0xffffffffull));
    dj.push(Json::integer(opts.device));
    mgbfs_engine_plan_info p{};
    char err[256] = {0};
    if (mgbfs_engine_plan(&desc, &opts, &p, err, sizeof err)) throw CliError(err[0] ? err : "GRAPH_INFO_PLAN");
    if (requested && (p.capacity != requested || p.device_bytes > p.budget_bytes))
      throw CliError("REQUESTED_CAPACITY_EXCEEDS_ADMISSION");
    if (batch) {
      if (batch > p.capacity) throw CliError("REQUESTED_BATCH_EXCEEDS_ADMISSION");
      p.batch = batch;
      p.queue_capacity = batch * desc.generators;
    }
    infos.push_back(p);
    if (p.capacity < infos[smallest].capacity) smallest = infos.size() - 1;
  }
  uint64_t total = 0;
  for (const auto& p : infos) {
    plans.push(plan_json(graph, p, world, static_cast<uint32_t>(shards)));
    total += p.capacity;
  }
  out["devices"] = dj;
  out["inventory"] = inventory;
  out["plan"] = plan_json(graph, infos[smallest], world, static_cast<uint32_t>(shards));
  out["rank_plans"] = plans;
  if (requested) {
    out["owner_cuts"] = Json();
  } else {
    Json cuts = Json::array();
    cuts.push(Json::integer(0));
    unsigned __int128 cumulative = 0;
    for (const auto& p : infos) {
      cumulative += p.capacity;
      cuts.push(Json::unsigned_(static_cast<uint64_t>(((cumulative << 32) + total - 1) / total)));
    }
    out["owner_cuts"] = cuts;
  }
  out["additional_reserve_bytes"] = Json::unsigned_(kReserveBytes);
  out["reserve_policy"] = Json::str("ABSOLUTE_BYTES_COLD_2GIB_WARM_256MIB");
  out["profile_status"] = Json::str("MEMORY_ADMITTED_NOT_THROUGHPUT_TUNED");
  return out;
}

// Optional [capacity|auto] [batch] arguments of the admission commands.
void parse_capacity_batch(const std::vector<std::string>& args, size_t at, uint64_t& requested, uint64_t& batch) {
  requested = args.size() > at && args[at] != "auto" ? parse_u64(args[at], "CLI_GRAPH_CAPACITY") : 0;
  batch = args.size() > at + 1 ? parse_u64(args[at + 1], "CLI_GRAPH_BATCH") : 0;
  if (args.size() > at && args[at] != "auto" && requested == 0) throw CliError("CLI_GRAPH_CAPACITY");
  if (args.size() > at + 1 && batch == 0) throw CliError("CLI_GRAPH_BATCH");
}

uint64_t parse_shards(const std::string& s) {
  const uint64_t shards = parse_u64(s, "CLI_GRAPH_SHARDS");
  if (shards == 0 || shards > 4096) throw CliError("CLI_GRAPH_SHARDS");
  return shards;
}

// graph-info <graph.json> <devices> <shards> [capacity|auto] [batch]: the
// admission of the selected local devices, from live free memory.
int cmd_graph_info(const std::vector<std::string>& args) {
  if (args.size() < 3 || args.size() > 5) throw CliError("CLI_GRAPH_INFO_ARGUMENTS");
  GraphDoc graph(mgbfs::host::JsonParser(read_file(args[0])).parse());
  const uint64_t shards = parse_shards(args[2]);
  const std::vector<int32_t> devs = parse_devices(args[1]);
  uint64_t requested = 0, batch = 0;
  parse_capacity_batch(args, 3, requested, batch);
  Json inv = Json::array();
  for (int32_t d : devs) inv.push(inventory_entry(d));
  std::vector<uint64_t> budgets(devs.size(), 0);  // live hipMemGetInfo per device
  std::cout << mgbfs::host::dumps(admission(graph, inv, budgets, shards, requested, batch)) << std::endl;
  return 0;
}

// graph-plan <graph.json> <inventory.json> <shards> [capacity|auto] [batch]:
// the same admission from an inventory list written by graph-info or by
// another host (generic_distributed_run.rs:91-96); no device is queried.
int cmd_graph_plan(const std::vector<std::string>& args) {
  if (args.size() < 3 || args.size() > 5) throw CliError("CLI_GRAPH_GLOBAL_ARGUMENTS");
  GraphDoc graph(mgbfs::host::JsonParser(read_file(args[0])).parse());
  const Json inventory = mgbfs::host::JsonParser(read_file(args[1])).parse();
  const uint64_t shards = parse_shards(args[2]);
  uint64_t requested = 0, batch = 0;
  parse_capacity_batch(args, 3, requested, batch);
  if (inventory.kind != Json::Arr || inventory.a.empty() || inventory.a.size() > 128)
    throw CliError("GRAPH_GLOBAL_INVENTORY_SHAPE");
  std::vector<uint64_t> budgets;
  for (const Json& e : inventory.a) {
    if (e.kind != Json::Obj || !e.has("device") || !e.at("device").is_int() || e.at("device").as_int() < 0)
      throw CliError("GRAPH_GLOBAL_DEVICE");
    if (!e.has("free_bytes") || !e.at("free_bytes").is_int()) throw CliError("GRAPH_GLOBAL_FREE_BYTES");
    if (!e.has("total_bytes") || !e.at("total_bytes").is_int()) throw CliError("GRAPH_GLOBAL_TOTAL_BYTES");
    const uint64_t fr = static_cast<uint64_t>(e.at("free_bytes").as_int());
    const uint64_t tot = static_cast<uint64_t>(e.at("total_bytes").as_int());
    if (fr == 0 || fr > tot) throw CliError("GRAPH_GLOBAL_MEMORY_RANGE");
    if (fr <= kReserveBytes) throw CliError("GENERIC_DISTRIBUTED_HEADROOM");
    budgets.push_back(fr - kReserveBytes);
  }
  std::cout << mgbfs::host::dumps(admission(graph, inventory, budgets, shards, requested, batch)) << std::endl;
  return 0;
}

// graph-local-plan <graph.json> <device> <world> <shards> <capacity|auto> <batch>:
// this device's plan and inventory figures only (generic_distributed_run.rs:75-90).
int cmd_graph_local_plan(const std::vector<std::string>& args) {
  if (args.size() != 6) throw CliError("CLI_GRAPH_LOCAL_PLAN_ARGUMENTS");
  GraphDoc graph(mgbfs::host::JsonParser(read_file(args[0])).parse());
  const int32_t device = static_cast<int32_t>(parse_u64(args[1], "CLI_GRAPH_DEVICE"));
  const uint64_t world = parse_u64(args[2], "CLI_GRAPH_WORLD");
  const uint64_t shards = parse_shards(args[3]);
  if (world == 0 || world > 128) throw CliError("CLI_GRAPH_LOCAL_PLAN_GEOMETRY");
  uint64_t requested = 0, batch = 0;
  parse_capacity_batch(args, 4, requested, batch);
  if (batch == 0) throw CliError("CLI_GRAPH_BATCH");
  Json inv = Json::array();
  inv.push(inventory_entry(device));
  const Json adm = admission(graph, inv, std::vector<uint64_t>(1, 0), shards, requested, batch);
  Json out = inv.a[0];
  out["size_profile_capability"] = Json::integer(1);
  out["graph_digest"] = Json::str(graph.digest());
  Json plan = adm.at("plan");
  plan["world"] = Json::unsigned_(world);
  out["plan"] = plan;
  std::cout << mgbfs::host::dumps(out) << std::endl;
  return 0;
}

// Rank worker of the reference launcher (generic_distributed_run.rs:109-232)
// for a world of one device: reads the admission written by graph-info,
// runs the BFS on devices[rank] with the rank plan's backend, capacity and
// batch, and writes <out>/report.json (with the rank plan echoed verbatim)
// and <out>/states.json.  Worlds above one need the owner exchange, which
// this build does not provide.
int cmd_graph_rank(const std::vector<std::string>& args) {
  if (args.size() != 6) throw CliError("CLI_GRAPH_RANK_ARGUMENTS");
  const Json launch = mgbfs::host::JsonParser(read_file(args[1])).parse();
  const uint64_t rank = parse_u64(args[2], "CLI_GRAPH_RANK");
  GraphDoc graph(mgbfs::host::JsonParser(read_file(args[0])).parse());
  if (launch.kind != Json::Obj || !launch.has("graph_digest") || !launch.has("devices") || !launch.has("plan") ||
      launch.at("devices").kind != Json::Arr || launch.at("plan").kind != Json::Obj)
    throw CliError("GRAPH_LAUNCH_IDENTITY_OR_GEOMETRY");
  const Json& devices = launch.at("devices");
  const Json& base = launch.at("plan");
  const uint64_t world = base.has("world") && base.at("world").is_int() ? static_cast<uint64_t>(base.at("world").as_int()) : 0;
  if (launch.at("graph_digest").kind != Json::Str || launch.at("graph_digest").s != graph.digest() ||
      devices.a.size() != world || rank >= world)
    throw CliError("GRAPH_LAUNCH_IDENTITY_OR_GEOMETRY");
  if (world != 1) throw CliError("GRAPH_RANK_WORLD_UNSUPPORTED");
  const Json& rank_plan =
      launch.has("rank_plans") && launch.at("rank_plans").kind == Json::Arr && !launch.at("rank_plans").a.empty()
          ? launch.at("rank_plans").a.at(rank)
          : base;
  auto field = [&](const char* k) -> uint64_t {
    if (!rank_plan.has(k) || !rank_plan.at(k).is_int() || rank_plan.at(k).as_int() < 0) throw CliError("GRAPH_RANK_PLANS");
    return static_cast<uint64_t>(rank_plan.at(k).as_int());
  };
  BfsRequest req;
  req.device = static_cast<int32_t>(devices.a[rank].as_int());
  req.capacity = field("capacity");
  req.batch = static_cast<uint32_t>(std::min<uint64_t>(field("batch"), 0xffffffffull));
  req.backend = rank_plan.has("backend") && rank_plan.at("backend").kind == Json::Str ? rank_plan.at("backend").s : "auto";
  req.seconds = parse_u64(args[5], "CLI_GRAPH_SECONDS");
  const char* phase = std::getenv("MGBFS_JOB_PHASE_PATH");
  req.phase_path = phase ? phase : "";
  const std::string out = args[4];
  if (path_exists(out)) throw CliError("CLI_GRAPH_OUTPUT_EXISTS");
  if (mkdir(out.c_str(), 0755)) throw CliError("CLI_GRAPH_OUTPUT_CREATE");
  const BfsOutcome r = run_bfs(graph, req, nullptr, "", nullptr, "current_sample_global_limit");
  Json report = Json::object();
  report["schema"] = Json::integer(2);
  report["rank"] = Json::unsigned_(rank);
  report["world"] = Json::unsigned_(world);
  report["device"] = Json::integer(req.device);
  report["status"] = Json::str(r.status);
  report["reason"] = Json::str(r.reason);
  report["graph_digest"] = Json::str(graph.digest());
  report["states_sha256"] = Json::str(mgbfs::host::sha256_hex(r.states_text));
  Json seconds;
  report["layer_sizes"] = layer_json(r, seconds);
  report["layer_seconds"] = seconds;
  report["bfs_seconds"] = Json::real(r.bfs_seconds);
  report["setup_seconds"] = Json::real(r.setup_seconds);
  report["source_retries"] = Json::unsigned_(r.info.retries);
  report["plan"] = rank_plan;
  report["backend"] = Json::str(r.backend);
  report["candidates_generated"] = Json::unsigned_(r.info.candidates);
  write_file(out + "/states.json", r.states_text);
  write_file(out + "/report.json", mgbfs::host::dumps(report));
  write_job_phase(req.phase_path, "done");
  return 0;
}

// Graph contract without a device: digest, validation, retention rule.
int cmd_graph_digest(const std::vector<std::string>& args) {
  if (args.size() != 1) throw CliError("CLI_USAGE: mgbfs graph-digest <graph.json>");
  GraphDoc graph(mgbfs::host::JsonParser(read_file(args[0])).parse());
  Json out = Json::object();
  out["graph_digest"] = Json::str(graph.digest());
  out["inverse_closed"] = Json::boolean(graph.inverse_closed());
  out["elements"] = Json::integer(graph.elements);
  out["generators"] = Json::integer(graph.kind == 0 ? graph.perms.size() : graph.mats.size());
  std::cout << mgbfs::host::dumps(out) << std::endl;
  return 0;
}

int cmd_key_info() {
  Json out = Json::object();
  out["schema"] = Json::integer(1);
  out["abi_version"] = Json::integer(mgbfs_hip_abi_version());
  Json b = Json::array();
  for (const char* x : {"sorted", "permutation_bitmap", "segmented", "generic"}) b.push(Json::str(x));
  out["backends"] = b;
  Json c = Json::array();
  for (const char* x : {"graph", "graph-info", "graph-plan", "graph-local-plan", "graph-rank", "graph-session",
                        "graph-count", "graph-digest", "key-info", "version"})
    c.push(Json::str(x));
  out["commands"] = c;
  out["max_packed_degree"] = Json::integer(16);
  out["state_encoding"] = Json::str("signed_int64_vectors");
  out["device_arch"] = Json::str("gfx950");
  std::cout << mgbfs::host::dumps(out) << std::endl;
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  std::vector<std::string> args(argv + 1, argv + argc);
  try {
    if (args.empty())
      throw CliError(
          "CLI_USAGE: mgbfs <graph|graph-session|graph-info|graph-plan|graph-local-plan|graph-rank|graph-count|"
          "graph-digest|key-info|version> ...");
    const std::string cmd = args[0];
    std::vector<std::string> rest(args.begin() + 1, args.end());
    if (cmd == "graph") return cmd_graph(rest);
    if (cmd == "graph-session") return cmd_graph_session(rest);
    if (cmd == "graph-info") return cmd_graph_info(rest);
    if (cmd == "graph-rank") return cmd_graph_rank(rest);
    if (cmd == "graph-plan") return cmd_graph_plan(rest);
    if (cmd == "graph-local-plan") return cmd_graph_local_plan(rest);
    if (cmd == "graph-count") return cmd_graph_count();
    if (cmd == "graph-digest") return cmd_graph_digest(rest);
    if (cmd == "key-info") return cmd_key_info();
    if (cmd == "version") {
      std::cout << "mgbfs-hip abi " << mgbfs_hip_abi_version() << std::endl;
      return 0;
    }
    throw CliError("CLI_UNKNOWN_COMMAND");
  } catch (const std::exception& e) {
    std::string code = e.what();
    Json err = Json::object();
    err["status"] = Json::str("ERROR");
    err["error"] = Json::str(code);
    std::cerr << mgbfs::host::dumps(err) << std::endl;
    return 1;
  }
}
