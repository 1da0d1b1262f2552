// The step kernels' launches, shared by the host code and the translation units that compile
// the kernels (one per precision and coefficient mode, built in parallel).
#pragma once
#include <hip/hip_runtime.h>

#include "common.hpp"
#include "kernels.hpp"

namespace pxfdtd {

struct SideStream {
    hipStream_t s;
    hipEvent_t fork, join;
};

struct UnitCounts {
    int lean, x, slab;  // units in the lists, in this order
};

// One step's launches of kernel `kernel` (0 LDS tiles, 1-2 the 62-wide march, 3 the aligned
// march, 4 the 2D march) in coefficient mode MODE; defined in step_launch_impl.hpp and
// instantiated in the step_*.hip units.
template <typename T, int MODE>
int launch_mode(const StepArgs<T> &A, int kernel, int ty, hipStream_t s, size_t lds,
                const SideStream *side, const UnitCounts &n, int set);

}  // namespace pxfdtd
