// Launch-shape autotuner of the gfx950 kernels (host side).
//
// A knob is a launch parameter whose value changes only how work is split
// over workgroups, never the arithmetic a result sees (replicates per
// workgroup, generation width): every candidate produces bit-identical
// outputs, so tuning is safe to do on the live workload.  The tuner is online:
// the first launches of a knob (per shape key) each try the next candidate,
// timed with HIP events and normalised by the work they carried; once every
// candidate has a time, the fastest is used from then on and written to the
// cache file (KPC_TUNE_CACHE, default $HOME/.cache/kpcmi/tune.txt; "off"
// disables the file), which later processes read instead of re-measuring.
#pragma once
#include <string>
#include <vector>

namespace kpc {

class Tuner {
 public:
  static Tuner& instance();
  // Candidate to use for this launch of `key` (index into `candidates`);
  // `trial` is set when the launch is a measurement.
  int choose(const std::string& key, const std::vector<int>& candidates, bool* trial);
  // Report the timed trial: seconds for `work` units with candidate value.
  void record(const std::string& key, int value, double seconds, double work);
  // One line per key: "key value=<chosen or ?> <value>:<sec per unit> ...".
  std::string report();
  void reset();

 private:
  Tuner();
  struct Entry {
    std::vector<int> candidates;
    std::vector<double> per_unit;  // < 0 = not measured yet
    int chosen = -1;               // value, -1 while trials run
  };
  void load();
  void save();
  std::string path_;
  bool loaded_ = false;
  std::vector<std::pair<std::string, Entry>> table_;
  Entry* find(const std::string& key);
};

}  // namespace kpc
