// Kernel timing registry behind mk_profile_* (see profile.hpp).
#include <hip/hip_runtime.h>

#include <atomic>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "mk_internal.hpp"
#include "profile.hpp"

namespace mk {
namespace {

struct Pending {
  hipEvent_t a, b;
  const char *name;
  double bytes;
};
struct Total {
  uint64_t launches = 0;
  double ms = 0.0;
  double bytes = 0.0;
};

std::atomic<int> g_on{0};  // 0 off, 1 every kernel, 2 persistent loops only
std::mutex g_mu;
std::vector<Pending> g_pending;
std::vector<hipEvent_t> g_free;
std::map<std::string, Total> g_totals;
std::vector<std::string> g_names;  // stable storage for returned names
struct Step {
  const char *name;
  uint32_t step, small, hub_items;
  double us;
};
std::vector<Step> g_steps;
constexpr size_t kMaxSteps = 1 << 16;

hipEvent_t take_event() {
  if (!g_free.empty()) {
    hipEvent_t e = g_free.back();
    g_free.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

void resolve_locked() {
  for (auto &p : g_pending) {
    float ms = 0.f;
    if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      Total &t = g_totals[p.name];
      t.launches++;
      t.ms += ms;
      t.bytes += p.bytes;
    }
    g_free.push_back(p.a);
    g_free.push_back(p.b);
  }
  g_pending.clear();
}

}  // namespace

bool profiling_on() { return g_on.load(std::memory_order_relaxed) != 0; }
bool profiling_all() { return g_on.load(std::memory_order_relaxed) == 1; }

hipEvent_t prof_begin(hipStream_t s) {
  if (!profiling_all()) return nullptr;
  std::lock_guard<std::mutex> lk(g_mu);
  hipEvent_t e = take_event();
  if (e && hipEventRecord(e, s) != hipSuccess) {
    g_free.push_back(e);
    return nullptr;
  }
  return e;
}

void prof_end(hipEvent_t start, hipStream_t s, const char *name, double algo_bytes) {
  std::lock_guard<std::mutex> lk(g_mu);
  hipEvent_t e = take_event();
  if (!e || hipEventRecord(e, s) != hipSuccess) {
    g_free.push_back(start);
    if (e) g_free.push_back(e);
    return;
  }
  g_pending.push_back({start, e, name, algo_bytes});
  if (g_pending.size() > 4096) resolve_locked();
}

bool prof_events(hipEvent_t &start, hipEvent_t &stop) {
  std::lock_guard<std::mutex> lk(g_mu);
  start = take_event();
  stop = start ? take_event() : nullptr;
  if (start && stop) return true;
  if (start) g_free.push_back(start);
  start = stop = nullptr;
  return false;
}

void prof_commit(hipEvent_t start, hipEvent_t stop, const char *name, double algo_bytes) {
  std::lock_guard<std::mutex> lk(g_mu);
  g_pending.push_back({start, stop, name, algo_bytes});
  if (g_pending.size() > 4096) resolve_locked();
}

void prof_add_bytes(const char *name, double algo_bytes) {
  if (!profiling_on()) return;
  std::lock_guard<std::mutex> lk(g_mu);
  g_totals[name].bytes += algo_bytes;
}

void prof_time(const char *name, double ms) {
  if (!profiling_on()) return;
  std::lock_guard<std::mutex> lk(g_mu);
  Total &t = g_totals[name];
  t.launches++;
  t.ms += ms;
}

void prof_step(const char *name, uint32_t step, uint32_t small, uint32_t hub_items, double us) {
  if (!profiling_on()) return;
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_steps.size() < kMaxSteps) g_steps.push_back({name, step, small, hub_items, us});
}

}  // namespace mk

using namespace mk;

extern "C" {

int mk_profile_enable(int on) {
  if (on < 0 || on > 2) return fail(MK_ERR_BAD_ARG, "mk_profile_enable: 0, 1 or 2");
  g_on.store(on);
  return MK_OK;
}

int mk_profile_reset(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  resolve_locked();
  g_totals.clear();
  g_steps.clear();
  return MK_OK;
}

int mk_profile_step_count(int *count) {
  if (!count) return fail(MK_ERR_BAD_ARG, "mk_profile_step_count: NULL count");
  std::lock_guard<std::mutex> lk(g_mu);
  *count = int(g_steps.size());
  return MK_OK;
}

int mk_profile_step_get(int i, const char **name, uint32_t *step, uint32_t *small,
                        uint32_t *hub_items, double *us) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (i < 0 || size_t(i) >= g_steps.size())
    return fail(MK_ERR_BAD_ARG, "mk_profile_step_get: index out of range");
  const Step &s = g_steps[size_t(i)];
  if (name) *name = s.name;
  if (step) *step = s.step;
  if (small) *small = s.small;
  if (hub_items) *hub_items = s.hub_items;
  if (us) *us = s.us;
  return MK_OK;
}

int mk_profile_count(int *count) {
  if (!count) return fail(MK_ERR_BAD_ARG, "mk_profile_count: NULL count");
  std::lock_guard<std::mutex> lk(g_mu);
  resolve_locked();
  g_names.clear();
  for (auto &kv : g_totals) g_names.push_back(kv.first);
  *count = int(g_names.size());
  return MK_OK;
}

int mk_profile_get(int i, const char **name, uint64_t *launches, double *total_ms,
                   double *algo_bytes) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (i < 0 || size_t(i) >= g_names.size())
    return fail(MK_ERR_BAD_ARG, "mk_profile_get: index out of range (call mk_profile_count)");
  const Total &t = g_totals[g_names[i]];
  if (name) *name = g_names[i].c_str();
  if (launches) *launches = t.launches;
  if (total_ms) *total_ms = t.ms;
  if (algo_bytes) *algo_bytes = t.bytes;
  return MK_OK;
}

}  // extern "C"
