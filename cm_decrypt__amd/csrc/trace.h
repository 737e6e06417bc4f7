// Named ranges for rocprofv3 --marker-trace around the host-side steps of a launch (the
// stages of the plain pipeline, the fused launch and its KEM/positions follow-ups, key
// admission, graph replays, host-buffer calls and server launches).  Without a profiler
// attached a range costs two calls into the roctx library.
#pragma once

#include <rocprofiler-sdk-roctx/roctx.h>

namespace cmd {

class TraceRange {
public:
    explicit TraceRange(const char *name) { roctxRangePushA(name); }
    ~TraceRange() { end(); }
    // closes the range before the end of the scope (once)
    void end()
    {
        if (open_) roctxRangePop();
        open_ = false;
    }
    TraceRange(const TraceRange &) = delete;
    TraceRange &operator=(const TraceRange &) = delete;

private:
    bool open_ = true;
};

}  // namespace cmd
