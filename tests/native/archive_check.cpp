// Drives the layer-archive writer (multigpubfs_amd/csrc/archive.hpp) from a
// text script on stdin, so the Python reader can check the bytes it writes.
//   line 1:  <path> <state width n> <64 hex digits of config digest>
//   then:    L <count>  followed by <count> lines of n integers  (one layer)
//            C                                                   (commit)
// On a writer error prints "ERROR <code>" and exits with status 2.
#include <iostream>
#include <string>
#include <vector>

#include "archive.hpp"

int main() {
  std::string path, config, op;
  uint64_t n = 0;
  if (!(std::cin >> path >> n >> config)) return 1;
  try {
    mgbfs::host::LayerArchive archive(path, n, config);
    uint64_t depth = 0;
    while (std::cin >> op) {
      if (op == "C") {
        archive.commit();
        break;
      }
      uint64_t count = 0;
      std::cin >> count;
      std::vector<int64_t> rows(count * n);
      for (auto& v : rows) std::cin >> v;
      archive.layer(depth++, rows, count);
    }
  } catch (const std::exception& e) {
    std::cout << "ERROR " << e.what() << std::endl;
    return 2;
  }
  std::cout << "archive ok" << std::endl;
  return 0;
}
