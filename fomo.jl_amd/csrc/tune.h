// tune.h -- the autotuners' choices, kept per process and in a file across processes.
#pragma once
#include <string>
#include <vector>

namespace fomo {

// A tuning record: `key` names the device, the library build and the problem; `vals` is the
// choice.  The file holds one record per line ("key = v0 v1 ..."), appended; the last record of a
// key wins.  Path: FOMO_TUNE_CACHE, else $XDG_CACHE_HOME/fomo_hip/tune.txt, else
// $HOME/.cache/fomo_hip/tune.txt; FOMO_TUNE_CACHE=off keeps choices in the process only.  The
// process keeps the records of each file path apart (a process that switches FOMO_TUNE_CACHE
// reads the new file).  A file that cannot be read or written is skipped silently: the choice only
// affects speed, never bits.
bool tune_lookup(const std::string& key, std::vector<int>& vals);
void tune_store(const std::string& key, const std::vector<int>& vals);
// Prefix of every key: architecture, compute units and the library build.
std::string tune_device_tag();
// Where the last lookup found its record: 0 none, 1 this process, 2 the file (for introspection).
int tune_last_source();

}  // namespace fomo
