// Per-kernel timing with HIP events, recorded on the stream each kernel is
// launched on, for the benchmark's live roofline figures. Off unless started.
#pragma once

#include <hip/hip_runtime.h>

namespace nxs {

// PK_COST: census and costs, one kernel; PK_CHECK_MEDIAN: the left-right check and the median;
// the band search's costs, paths and selection; PK_SHRINK: coarse-to-fine's shrink and band
// bases.
enum ProfiledKernel {
    PK_COST,
    PK_PATH,
    PK_SELECT,
    PK_CHECK_MEDIAN,
    PK_BAND_COST,
    PK_BAND_PATH,
    PK_BAND_SELECT,
    PK_SHRINK,
    PK_COUNT
};

// Brackets one launch: begin() before it, end() after it, on the same stream.
// Both are no-ops while profiling is off.
int profile_begin(ProfiledKernel kernel, hipStream_t stream);
void profile_end(int slot, hipStream_t stream);

}  // namespace nxs
