// Launch-shape autotuner (see kpc_tune.h).
#include "kpc_tune.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <sstream>
#include <sys/stat.h>

namespace kpc {

namespace {
std::mutex g_tune_mutex;
}

Tuner& Tuner::instance() {
  static Tuner t;
  return t;
}

Tuner::Tuner() {
  const char* p = std::getenv("KPC_TUNE_CACHE");
  if (p && *p) {
    path_ = p;
  } else if (const char* home = std::getenv("HOME")) {
    path_ = std::string(home) + "/.cache/kpcmi/tune.txt";
  }
  if (path_ == "off") path_.clear();
}

Tuner::Entry* Tuner::find(const std::string& key) {
  for (auto& kv : table_)
    if (kv.first == key) return &kv.second;
  return nullptr;
}

void Tuner::load() {
  loaded_ = true;
  if (path_.empty()) return;
  FILE* f = std::fopen(path_.c_str(), "r");
  if (!f) return;
  char line[512];
  while (std::fgets(line, sizeof(line), f)) {
    char key[400];
    int value = 0;
    if (std::sscanf(line, "%399s %d", key, &value) == 2) {
      Entry e;
      e.chosen = value;
      table_.emplace_back(key, e);
    }
  }
  std::fclose(f);
}

void Tuner::save() {
  if (path_.empty()) return;
  // create the directory chain of the cache file (best effort)
  std::string dir = path_.substr(0, path_.find_last_of('/'));
  for (size_t i = 1; i <= dir.size(); ++i)
    if (i == dir.size() || dir[i] == '/') (void)mkdir(dir.substr(0, i).c_str(), 0755);
  const std::string tmp = path_ + ".tmp";
  FILE* f = std::fopen(tmp.c_str(), "w");
  if (!f) return;
  for (auto& kv : table_)
    if (kv.second.chosen >= 0) std::fprintf(f, "%s %d\n", kv.first.c_str(), kv.second.chosen);
  std::fclose(f);
  (void)std::rename(tmp.c_str(), path_.c_str());
}

int Tuner::choose(const std::string& key, const std::vector<int>& candidates, bool* trial) {
  std::lock_guard<std::mutex> lock(g_tune_mutex);
  if (!loaded_) load();
  *trial = false;
  Entry* e = find(key);
  if (!e) {
    table_.emplace_back(key, Entry());
    e = &table_.back().second;
  }
  if (e->candidates.empty()) {
    e->candidates = candidates;
    e->per_unit.assign(candidates.size(), -1.0);
  }
  if (e->chosen >= 0) {
    for (int c : candidates)
      if (c == e->chosen) return c;
    e->chosen = -1;  // cached value no longer offered: measure again
  }
  for (size_t i = 0; i < e->candidates.size(); ++i)
    if (e->per_unit[i] < 0.0) {
      *trial = true;
      return e->candidates[i];
    }
  return e->candidates[0];
}

void Tuner::record(const std::string& key, int value, double seconds, double work) {
  std::lock_guard<std::mutex> lock(g_tune_mutex);
  Entry* e = find(key);
  if (!e || e->chosen >= 0 || work <= 0.0) return;
  for (size_t i = 0; i < e->candidates.size(); ++i)
    if (e->candidates[i] == value && e->per_unit[i] < 0.0) e->per_unit[i] = seconds / work;
  size_t best = 0;
  for (size_t i = 0; i < e->candidates.size(); ++i) {
    if (e->per_unit[i] < 0.0) return;
    if (e->per_unit[i] < e->per_unit[best]) best = i;
  }
  e->chosen = e->candidates[best];
  save();
}

std::string Tuner::report() {
  std::lock_guard<std::mutex> lock(g_tune_mutex);
  std::ostringstream os;
  for (auto& kv : table_) {
    os << kv.first << " value=";
    if (kv.second.chosen >= 0)
      os << kv.second.chosen;
    else
      os << "?";
    for (size_t i = 0; i < kv.second.candidates.size(); ++i)
      os << " " << kv.second.candidates[i] << ":" << kv.second.per_unit[i];
    os << "\n";
  }
  return os.str();
}

void Tuner::reset() {
  std::lock_guard<std::mutex> lock(g_tune_mutex);
  table_.clear();
  loaded_ = true;  // a reset table is not re-read from the file
}

}  // namespace kpc
