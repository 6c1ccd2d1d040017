// Profiler ranges around the library's launch entry points, the counterpart of
// the reference's NVTX ranges "custom_ntt" / "custom_intt" around its timed
// transforms (src/v04_montgomery.cu:11-13, 827-839, 981-992, 1030-1043).
// roctx ranges are host-side and cost a call each; without a profiler
// attached (rocprofv3 --marker-trace) they record nothing.
#pragma once

#include <rocprofiler-sdk-roctx/roctx.h>

namespace ntt {

struct TraceRange {
    explicit TraceRange(const char* name) { roctxRangePushA(name); }
    ~TraceRange() { roctxRangePop(); }
    TraceRange(const TraceRange&) = delete;
    TraceRange& operator=(const TraceRange&) = delete;
};

}  // namespace ntt
