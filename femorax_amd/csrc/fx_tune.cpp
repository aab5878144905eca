// The persistent autotune cache of lattice batches: the reference's choose_force_settings
// (src/exec/force_settings.cpp:157-205) keeps the tuned force settings per machine and model in a
// text file (one "key<TAB>value" line each) and tunes again only for an unknown key. Here the
// key names the device and the library, and the model as the tile kernels see it (lattice cell
// count in powers of two, whether mass or damping terms appear, the tile depth when the caller
// fixed it); the value is the tile depth and the kernel variant that measured fastest.
#include <hip/hip_runtime.h>

#include <unistd.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <map>
#include <mutex>
#include <string>

#include "fx_internal.hpp"

namespace fx {

namespace {

constexpr const char* kLibraryVersion = "femorax_amd-2";
std::mutex cache_mutex;

std::filesystem::path cache_path() {
  if (const char* p = std::getenv("FX_TUNE_CACHE"); p && *p) return p;
  if (const char* x = std::getenv("XDG_CACHE_HOME"); x && *x)
    return std::filesystem::path(x) / "femorax_amd" / "autotune.txt";
  if (const char* h = std::getenv("HOME"); h && *h)
    return std::filesystem::path(h) / ".cache" / "femorax_amd" / "autotune.txt";
  return {};
}

std::map<std::string, std::string> read_all(const std::filesystem::path& p) {
  std::map<std::string, std::string> m;
  std::ifstream in(p);
  for (std::string line; std::getline(in, line);) {
    const auto tab = line.find('\t');
    if (tab != std::string::npos) m[line.substr(0, tab)] = line.substr(tab + 1);
  }
  return m;
}

}  // namespace

std::string tune_machine_key(int device) {
  hipDeviceProp_t prop{};
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
    (void)hipGetLastError();
    return std::string("device") + std::to_string(device) + "," + kLibraryVersion;
  }
  return std::string(prop.gcnArchName) + "," + std::to_string(prop.multiProcessorCount) + "cu," + kLibraryVersion;
}

std::string tune_model_key(int64_t cells, bool mass, bool damping, int forced_depth) {
  const int log2n = cells > 0 ? static_cast<int>(std::floor(std::log2(static_cast<double>(cells)))) : 0;
  return std::string("stdBrick+ElasticIsotropic,lattice,cells=2^") + std::to_string(log2n) +
         ",mass=" + (mass ? "1" : "0") + ",damping=" + (damping ? "1" : "0") +
         ",depth=" + (forced_depth ? std::to_string(forced_depth) : std::string("auto"));
}

bool tune_cache_get(const std::string& key, int& depth, int& variant) {
  std::lock_guard<std::mutex> lock(cache_mutex);
  const auto path = cache_path();
  if (path.empty()) return false;
  const auto m = read_all(path);
  const auto it = m.find(key);
  if (it == m.end()) return false;
  int d = 0, v = -1;
  if (std::sscanf(it->second.c_str(), "depth=%d variant=%d", &d, &v) != 2) return false;
  if ((d != 4 && d != 8 && d != 16) || v < 0 || v > 15) return false;
  depth = d;
  variant = v;
  return true;
}

void tune_cache_put(const std::string& key, int depth, int variant) {
  std::lock_guard<std::mutex> lock(cache_mutex);
  const auto path = cache_path();
  if (path.empty()) return;
  auto m = read_all(path);
  m[key] = "depth=" + std::to_string(depth) + " variant=" + std::to_string(variant);
  std::error_code ec;
  std::filesystem::create_directories(path.parent_path(), ec);
  const auto tmp = path.string() + ".tmp" + std::to_string(static_cast<long long>(::getpid()));
  {
    std::ofstream out(tmp);
    if (!out) return;  // no writable cache: tune again next time
    for (const auto& [k, v] : m) out << k << '\t' << v << '\n';
  }
  std::filesystem::rename(tmp, path, ec);
}

}  // namespace fx
