// tune.hip -- persistent cache of the autotuners' choices (host code).
#include "tune.h"

#include <hip/hip_runtime.h>

#include <dlfcn.h>
#include <fcntl.h>
#include <sys/file.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <sstream>

#include "../../include/fomo_hip.h"

namespace fomo {

namespace {

std::mutex g_mu;
int g_last_source = 0;

// Per tuning file (keyed by its path; "" when the file is off): the records this process timed or
// looked up, and the file's records, read once.
struct Store {
  std::map<std::string, std::vector<int>> mem, file;
  bool loaded = false;
};
std::map<std::string, Store> g_stores;

std::string cache_path() {
  const char* env = std::getenv("FOMO_TUNE_CACHE");
  if (env) return std::strcmp(env, "off") == 0 ? std::string() : std::string(env);
  std::string base;
  if (const char* x = std::getenv("XDG_CACHE_HOME"); x && *x)
    base = x;
  else if (const char* h = std::getenv("HOME"); h && *h)
    base = std::string(h) + "/.cache";
  else
    return std::string();
  return base + "/fomo_hip/tune.txt";
}

void make_parent_dirs(const std::string& path) {
  for (size_t p = path.find('/', 1); p != std::string::npos; p = path.find('/', p + 1)) {
    const std::string dir = path.substr(0, p);
    if (mkdir(dir.c_str(), 0755) != 0 && errno != EEXIST) return;
  }
}

// Records of the file (the last one of a key wins), read once per process.
void load_file_locked(Store& st, const std::string& path) {
  if (st.loaded) return;
  st.loaded = true;
  if (path.empty()) return;
  FILE* f = std::fopen(path.c_str(), "r");
  if (!f) return;
  char line[1024];
  while (std::fgets(line, sizeof line, f)) {
    const char* eq = std::strstr(line, " = ");
    if (!eq) continue;
    const std::string key(line, eq - line);
    std::vector<int> vals;
    std::istringstream in(eq + 3);
    int v;
    while (in >> v) vals.push_back(v);
    if (!key.empty() && !vals.empty()) st.file[key] = vals;
  }
  std::fclose(f);
}

}  // namespace

std::string tune_device_tag() {
  int dev = 0;
  hipDeviceProp_t prop{};
  std::string arch = "unknown";
  int cus = 0;
  if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess) {
    arch = prop.gcnArchName;
    const size_t colon = arch.find(':');
    if (colon != std::string::npos) arch = arch.substr(0, colon);
    cus = prop.multiProcessorCount;
  }
  // the build (the library file's size and modification time): any rebuild may change which
  // shape is fastest
  std::string build = fomo_version();
  Dl_info info{};
  struct stat st {};
  if (dladdr((void*)&tune_device_tag, &info) && info.dli_fname && stat(info.dli_fname, &st) == 0)
    build += "/" + std::to_string((long long)st.st_size) + "/" + std::to_string((long long)st.st_mtime);
  return arch + "/" + std::to_string(cus) + "cu/" + build;
}

bool tune_lookup(const std::string& key, std::vector<int>& vals) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_last_source = 0;
  const std::string path = cache_path();
  Store& st = g_stores[path];
  auto it = st.mem.find(key);
  if (it != st.mem.end()) {
    vals = it->second;
    g_last_source = 1;
    return true;
  }
  load_file_locked(st, path);
  it = st.file.find(key);
  if (it != st.file.end()) {
    vals = it->second;
    st.mem[key] = vals;
    g_last_source = 2;
    return true;
  }
  return false;
}

void tune_store(const std::string& key, const std::vector<int>& vals) {
  std::lock_guard<std::mutex> lk(g_mu);
  const std::string path = cache_path();
  g_stores[path].mem[key] = vals;
  if (path.empty() || key.find('\n') != std::string::npos) return;
  make_parent_dirs(path);
  const int fd = open(path.c_str(), O_WRONLY | O_CREAT | O_APPEND, 0644);
  if (fd < 0) return;
  std::string line = key + " =";
  for (int v : vals) line += " " + std::to_string(v);
  line += "\n";
  // one write of one line under an exclusive lock: concurrent processes append whole records
  if (flock(fd, LOCK_EX) == 0) {
    ssize_t w = write(fd, line.data(), line.size());
    (void)w;
    flock(fd, LOCK_UN);
  }
  close(fd);
}

int tune_last_source() {
  std::lock_guard<std::mutex> lk(g_mu);
  return g_last_source;
}

}  // namespace fomo
