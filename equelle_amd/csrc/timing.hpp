// Live kernel timing with HIP events recorded on the runtime stream around
// selected launches (SpMV, assembly).  Off by default; when on, every timed
// launch is bracketed by two events from a reusable pool.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

namespace eqhip {

void timing_enable(bool on);
bool timing_enabled();
void timing_reset();
// Brackets: call begin() before the launch and end() after it, same name.
void timing_begin(const char* name);
void timing_end(const char* name);
// Synchronises, then sums the elapsed times of the recorded pairs.
void timing_query(const std::string& name, int* count, double* total_ms);

struct TimedLaunch {
    const char* name;
    bool on;
    explicit TimedLaunch(const char* n) : name(n), on(timing_enabled()) {
        if (on) timing_begin(name);
    }
    ~TimedLaunch() {
        if (on) timing_end(name);
    }
};

}  // namespace eqhip
