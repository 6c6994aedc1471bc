// Optional per-stage timing with hipEvents recorded on the launching stream.
// Disabled by default; when enabled every instrumented launch is bracketed by two events and
// nusht_timing_collect() synchronizes and accumulates the elapsed time per stage name.
#pragma once

#include "common.h"

#include <map>
#include <mutex>
#include <string>
#include <vector>

namespace nusht {

class StageTimer {
 public:
  static StageTimer& get();
  bool enabled() const { return enabled_; }
  void enable(bool on);
  void begin(const char* name, hipStream_t st);
  void end(hipStream_t st);
  // synchronizes outstanding events; returns number of stages
  int collect();
  int count() const { return (int)names_.size(); }
  const char* name(int i) const { return names_[i].c_str(); }
  double total_ms(int i) const { return total_[i]; }
  long calls(int i) const { return calls_[i]; }
  void reset();

 private:
  struct Pending {
    int stage;
    hipEvent_t a, b;
  };
  hipEvent_t take_event();
  bool enabled_ = false;
  std::mutex mu_;
  std::vector<std::string> names_;
  std::map<std::string, int> index_;
  std::vector<double> total_;
  std::vector<long> calls_;
  std::vector<Pending> pending_;
  std::vector<hipEvent_t> pool_;
  int open_stage_ = -1;
  hipEvent_t open_a_ = nullptr;
};

// RAII bracket around one launch (no-op when timing is disabled)
struct StageScope {
  hipStream_t st;
  bool on;
  StageScope(const char* name, hipStream_t s) : st(s), on(StageTimer::get().enabled()) {
    if (on) StageTimer::get().begin(name, st);
  }
  ~StageScope() {
    if (on) StageTimer::get().end(st);
  }
};

}  // namespace nusht
