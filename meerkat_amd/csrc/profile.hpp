// Per-kernel timing with HIP events, enabled through mk_profile_enable().
//
// Each instrumented launch is bracketed by two events on the stream it runs
// on and tagged with its kernel name and its algorithmic byte count (the
// bytes the algorithm must move, not what the hardware moved).  Elapsed
// times are resolved lazily when the totals are read.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mk {

bool profiling_on();
// Returns the start event to pass to prof_end, or nullptr when disabled.
hipEvent_t prof_begin(hipStream_t s);
void prof_end(hipEvent_t start, hipStream_t s, const char *name, double algo_bytes);
// Add algorithmic bytes known only after a launch completed.
void prof_add_bytes(const char *name, double algo_bytes);

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
