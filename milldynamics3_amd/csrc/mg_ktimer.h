// Opt-in per-kernel timing for the library's operators: when on, every launch made through
// MG_TIMED_LAUNCH is bracketed by a pair of HIP events on the launch's stream, and after the
// operator's own synchronisation the pairs' elapsed times are summed per kernel name. Used by the
// benchmark to measure a kernel's average launch time inside a timed run.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <map>
#include <string>
#include <vector>

namespace mgk {

struct Timer {
  bool on = false;
  std::vector<hipEvent_t> pool;
  struct Pending {
    const char* name;
    hipEvent_t a, b;
  };
  std::vector<Pending> pend;
  std::map<std::string, std::pair<double, unsigned long long>> acc;  // name -> (ms, launches)

  ~Timer() {
    for (auto& p : pend) {
      pool.push_back(p.a);
      pool.push_back(p.b);
    }
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
  }
  hipEvent_t take() {
    if (pool.empty()) {
      hipEvent_t e = nullptr;
      (void)hipEventCreate(&e);
      return e;
    }
    hipEvent_t e = pool.back();
    pool.pop_back();
    return e;
  }
  void set(bool v) {
    collect();
    on = v;
    acc.clear();
  }
  // Folds the finished pairs into the totals (call after the launches' stream was synchronised).
  void collect() {
    for (auto& p : pend) {
      float ms = 0.0f;
      if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
        auto& a = acc[p.name];
        a.first += ms;
        a.second++;
      }
      pool.push_back(p.a);
      pool.push_back(p.b);
    }
    pend.clear();
  }
  // {"name": [total_ms, launches], ...}
  std::string json() const {
    std::string s = "{";
    bool first = true;
    for (auto& kv : acc) {
      char b[256];
      std::snprintf(b, sizeof b, "%s\"%s\":[%.6f,%llu]", first ? "" : ",", kv.first.c_str(), kv.second.first,
                    kv.second.second);
      s += b;
      first = false;
    }
    return s + "}";
  }
};

}  // namespace mgk

// hipLaunchKernelGGL on `stream` with the launch timed when the timer is on.
#define MG_TIMED_LAUNCH_S(timer, stream, kernel, grid, block, shm, ...)           \
  do {                                                                            \
    hipEvent_t mg_e0_ = nullptr;                                                  \
    if ((timer).on) {                                                             \
      mg_e0_ = (timer).take();                                                    \
      (void)hipEventRecord(mg_e0_, (stream));                                     \
    }                                                                             \
    hipLaunchKernelGGL(kernel, grid, block, shm, (stream), __VA_ARGS__);          \
    if ((timer).on) {                                                             \
      hipEvent_t mg_e1_ = (timer).take();                                         \
      (void)hipEventRecord(mg_e1_, (stream));                                     \
      (timer).pend.push_back({#kernel, mg_e0_, mg_e1_});                          \
    }                                                                             \
  } while (0)

// The same on the null stream.
#define MG_TIMED_LAUNCH(timer, kernel, grid, block, shm, ...) \
  MG_TIMED_LAUNCH_S(timer, nullptr, kernel, grid, block, shm, __VA_ARGS__)
