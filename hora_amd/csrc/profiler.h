// profiler.h -- optional per-kernel timing with HIP events on the launch stream.
#pragma once

#include "hm_common.h"

namespace hm {

bool profiling_enabled();

// Records a start event on construction and an end event on destruction,
// both on `s`, when profiling is enabled.
class ProfScope {
  public:
    ProfScope(const char *name, hipStream_t s);
    ~ProfScope();
    ProfScope(const ProfScope &) = delete;
    ProfScope &operator=(const ProfScope &) = delete;

  private:
    const char *name_;
    hipStream_t s_;
    hipEvent_t start_ = nullptr, stop_ = nullptr;
};

void profile_reset();
// Synchronises pending events and returns the totals for `name`.
bool profile_query(const char *name, double *total_ms, uint64_t *launches);

}  // namespace hm
