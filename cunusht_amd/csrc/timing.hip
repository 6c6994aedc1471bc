// Stage timing with hipEvents (see timing.h).
#include "timing.h"

namespace nusht {

StageTimer& StageTimer::get() {
  static StageTimer t;
  return t;
}

void StageTimer::enable(bool on) {
  std::lock_guard<std::mutex> g(mu_);
  enabled_ = on;
}

hipEvent_t StageTimer::take_event() {
  if (!pool_.empty()) {
    hipEvent_t e = pool_.back();
    pool_.pop_back();
    return e;
  }
  hipEvent_t e;
  NUSHT_HIP(hipEventCreate(&e));
  return e;
}

void StageTimer::begin(const char* name, hipStream_t st) {
  std::lock_guard<std::mutex> g(mu_);
  auto it = index_.find(name);
  int idx;
  if (it == index_.end()) {
    idx = (int)names_.size();
    index_[name] = idx;
    names_.push_back(name);
    total_.push_back(0.0);
    calls_.push_back(0);
  } else {
    idx = it->second;
  }
  open_stage_ = idx;
  open_a_ = take_event();
  NUSHT_HIP(hipEventRecord(open_a_, st));
}

void StageTimer::end(hipStream_t st) {
  std::lock_guard<std::mutex> g(mu_);
  if (open_stage_ < 0) return;
  hipEvent_t b = take_event();
  NUSHT_HIP(hipEventRecord(b, st));
  pending_.push_back(Pending{open_stage_, open_a_, b});
  open_stage_ = -1;
  open_a_ = nullptr;
}

int StageTimer::collect() {
  std::lock_guard<std::mutex> g(mu_);
  for (auto& p : pending_) {
    NUSHT_HIP(hipEventSynchronize(p.b));
    float ms = 0.f;
    NUSHT_HIP(hipEventElapsedTime(&ms, p.a, p.b));
    total_[p.stage] += ms;
    calls_[p.stage] += 1;
    pool_.push_back(p.a);
    pool_.push_back(p.b);
  }
  pending_.clear();
  return (int)names_.size();
}

void StageTimer::reset() {
  collect();
  std::lock_guard<std::mutex> g(mu_);
  for (auto& t : total_) t = 0.0;
  for (auto& c : calls_) c = 0;
}

}  // namespace nusht
