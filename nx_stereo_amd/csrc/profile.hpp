// Per-kernel timing with HIP events, recorded on the stream each kernel is
// launched on, for the benchmark's live roofline figures. Off unless started.
#pragma once

#include <hip/hip_runtime.h>

namespace nxs {

enum ProfiledKernel { PK_CENSUS, PK_COST, PK_PATH, PK_SELECT, PK_CONSISTENCY, PK_MEDIAN, PK_COUNT };

// Brackets one launch: begin() before it, end() after it, on the same stream.
// Both are no-ops while profiling is off.
int profile_begin(ProfiledKernel kernel, hipStream_t stream);
void profile_end(int slot, hipStream_t stream);

}  // namespace nxs
