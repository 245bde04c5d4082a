// Per-kernel-class timing with HIP events on the library stream.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

namespace plm {

bool profiling_on();
// Switches the event pairs off for its lifetime (regions captured into a graph).
struct ProfPause {
    ProfPause();
    ~ProfPause();
    ProfPause(const ProfPause&) = delete;
    ProfPause& operator=(const ProfPause&) = delete;

  private:
    bool was_;
};
// Records an event pair around a region on `st` when profiling is enabled.
struct ProfScope {
    ProfScope(const char* name, hipStream_t st);
    ~ProfScope();
    ProfScope(const ProfScope&) = delete;
    ProfScope& operator=(const ProfScope&) = delete;

  private:
    std::string name_;
    hipStream_t st_;
    hipEvent_t a_ = nullptr, b_ = nullptr;
};

}  // namespace plm
