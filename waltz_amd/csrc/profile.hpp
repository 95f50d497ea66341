// Per-stage HIP-event timing (see profile.hip). Usage inside a pipeline:
//   { StageClock c("wzp_encode", stream); launch(...); }
// and profile_collect() after the pipeline's final stream synchronisation.
#pragma once

#include <hip/hip_runtime.h>

#include <cstring>

namespace wz {

class StageClock {
  public:
    StageClock(const char* name, hipStream_t s);
    ~StageClock();
    StageClock(const StageClock&) = delete;
    StageClock& operator=(const StageClock&) = delete;

  private:
    const char* name_;
    hipStream_t s_;
    void* a_ = nullptr;
    void* b_ = nullptr;
};

void profile_collect();

}  // namespace wz
