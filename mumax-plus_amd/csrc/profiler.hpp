// Per-kernel timing with HIP events on the stream the kernels run on.
//
// When enabled, every launch site brackets its kernel with two events and
// records the kernel's algorithmic HBM bytes; flush() (called after a stream
// synchronisation) folds the elapsed times into per-name totals.
#pragma once

#include <map>
#include <string>
#include <vector>

#include "common.hpp"

namespace mpa {

class Profiler {
 public:
  struct Stat {
    double total_ms = 0.0;
    long long count = 0;
    double bytes = 0.0;  // algorithmic bytes summed over launches
  };

  ~Profiler() {
    for (auto& r : pending_) {
      pool_.push_back(r.a);
      pool_.push_back(r.b);
    }
    for (auto e : pool_) (void)hipEventDestroy(e);
  }

  bool enabled = false;

  // Returns a ticket for end(); -1 when profiling is off.
  int begin(const char* name, double bytes, hipStream_t s) {
    if (!enabled) return -1;
    if (pending_.size() >= 8192) flush_sync(s);
    Rec r{name, bytes, get(), get()};
    HIP_CHECK(hipEventRecord(r.a, s));
    pending_.push_back(r);
    return (int)pending_.size() - 1;
  }
  void end(int ticket, hipStream_t s) noexcept {
    if (ticket < 0) return;
    (void)hipEventRecord(pending_[ticket].b, s);
  }
  void flush_sync(hipStream_t s) {
    if (pending_.empty()) return;
    HIP_CHECK(hipStreamSynchronize(s));
    for (auto& r : pending_) {
      float ms = 0.f;
      HIP_CHECK(hipEventElapsedTime(&ms, r.a, r.b));
      Stat& st = stats[r.name];
      st.total_ms += ms;
      st.count += 1;
      st.bytes += r.bytes;
      pool_.push_back(r.a);
      pool_.push_back(r.b);
    }
    pending_.clear();
  }
  void reset() { stats.clear(); }

  std::map<std::string, Stat> stats;

 private:
  struct Rec {
    std::string name;
    double bytes;
    hipEvent_t a, b;
  };
  hipEvent_t get() {
    if (!pool_.empty()) {
      hipEvent_t e = pool_.back();
      pool_.pop_back();
      return e;
    }
    hipEvent_t e;
    HIP_CHECK(hipEventCreate(&e));
    return e;
  }
  std::vector<Rec> pending_;
  std::vector<hipEvent_t> pool_;
};

// RAII bracket around one launch.
struct Timed {
  Profiler* p;
  hipStream_t s;
  int t;
  Timed(Profiler* prof, const char* name, double bytes, hipStream_t st)
      : p(prof), s(st), t(prof ? prof->begin(name, bytes, st) : -1) {}
  ~Timed() {
    if (p) p->end(t, s);
  }
};

}  // namespace mpa
