// Host-side mesh geometry shared by the regrid clustering (bsam_grid.cpp)
// and the ghost-fill planner (bsam_gpu.hip).  No device code.
#pragma once
#include <array>
#include <cstdint>
#include <functional>
#include <vector>

#include "../../include/bsam_gpu.h"

namespace bsam {

// PeriodicSetup (boundary.f90:805-872) + GetPeriodicOffsets (:875-954): the
// periodic image offsets of `level` (global cell index units of that level)
// for the root boundary codes `root_mthbc` (2 = periodic); `periodic` is
// false when no dimension is periodic.  No offsets below the root.
void periodic_offsets(int ndims, const bsam_patch_desc& root, int level,
                      std::vector<std::array<int, 3>>& offs, bool& periodic);

// fn(t) for t = 0 .. n-1 on the host threads of a process-wide pool (the
// caller runs t = 0 and waits for the rest).  A call made while the pool is
// busy (from inside fn, or from another thread), or after a fork, starts
// threads of its own instead; a thread that cannot be started leaves its
// share to the caller.
void parallel_for(int n, const std::function<void(int)>& fn);

// Record `msg` as the calling thread's bsam_last_error() and return code.
int32_t fail(int32_t code, const char* fmt, ...);

}  // namespace bsam
