// Per-kernel timing with HIP events, enabled through mk_profile_enable().
//
// Each instrumented launch is timed by a pair of HIP events and tagged with
// its kernel name and its algorithmic byte count (the bytes the algorithm
// must move, not what the hardware moved).  prof_launch hands the events to
// hipExtLaunchKernelGGL, which records them at the dispatch's own start and
// end (the kernel's execution time, as a profiler's kernel trace reports
// it); ProfScope records them on the stream around a region of several
// launches.  Elapsed times are resolved lazily when the totals are read.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstdint>

namespace mk {

bool profiling_on();   // any level
bool profiling_all();  // every kernel timed (dispatch events)
// Returns the start event to pass to prof_end, or nullptr when disabled.
hipEvent_t prof_begin(hipStream_t s);
void prof_end(hipEvent_t start, hipStream_t s, const char *name, double algo_bytes);
// Add algorithmic bytes known only after a launch completed.
void prof_add_bytes(const char *name, double algo_bytes);
// Add one launch of `name` whose duration was measured another way (ms).
void prof_time(const char *name, double ms);
// Record one frontier step of a persistent loop: its frontier (small
// entries, hub bucket items) and duration.
void prof_step(const char *name, uint32_t step, uint32_t small, uint32_t hub_items, double us);

// A pair of pooled events for one timed launch; false when profiling is off
// or no event is available.
bool prof_events(hipEvent_t &start, hipEvent_t &stop);
// Register a launch timed with (start, stop).
void prof_commit(hipEvent_t start, hipEvent_t stop, const char *name, double algo_bytes);

// Launch `kernel` on stream s; when profiling, time the dispatch itself.
template <typename... KArgs, typename... Args>
void prof_launch(const char *name, double algo_bytes, void (*kernel)(KArgs...), dim3 grid,
                 dim3 block, uint32_t shmem, hipStream_t s, Args... args) {
  hipEvent_t a = nullptr, b = nullptr;
  if (profiling_all() && prof_events(a, b)) {
    hipExtLaunchKernelGGL(kernel, grid, block, shmem, s, a, b, 0u, KArgs(args)...);
    prof_commit(a, b, name, algo_bytes);
  } else {
    hipLaunchKernelGGL(kernel, grid, block, shmem, s, KArgs(args)...);
  }
}

struct ProfScope {
  hipEvent_t ev;
  hipStream_t s;
  const char *name;
  double bytes;
  ProfScope(hipStream_t stream, const char *kernel, double algo_bytes)
      : ev(prof_begin(stream)), s(stream), name(kernel), bytes(algo_bytes) {}
  ~ProfScope() {
    if (ev) prof_end(ev, s, name, bytes);
  }
};

}  // namespace mk
