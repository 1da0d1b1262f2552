// The step kernels' launches, shared by the host code and the translation units that compile
// the kernels (one per precision and coefficient mode, built in parallel).
#pragma once
#include <hip/hip_runtime.h>

#include <functional>

#include "common.hpp"
#include "kernels.hpp"

namespace pxfdtd {

struct SideStream {
    hipStream_t s;
    hipEvent_t fork, join;
    // work of the step before that only reads the fields this step reads (its monitor taps):
    // run on the side stream after the fork, in the slack beside the longer main-stream launches
    std::function<int(hipStream_t)> first = nullptr;
};
// the two-step march's CPML classes each on a stream of its own beside the lean launch, forked
// from and joined to the main stream by events (all at the default priority)
struct PairStreams {
    hipStream_t s[4];
    hipEvent_t fork, join[4];
};

struct UnitCounts {
    int lean, x, slab;  // units in the lists, in this order
    int slab_rows;      // rows per unit of the slab list (the aligned march: 1 below the tiling's 2)
};

// One step's launches of kernel `kernel` (0 LDS tiles, 1-2 the 62-wide march, 3 the aligned
// march, 4 the 2D march) in coefficient mode MODE; defined in step_launch_impl.hpp and
// instantiated in the step_*.hip units.
template <typename T, int MODE>
int launch_mode(const StepArgs<T> &A, int kernel, int ty, hipStream_t s, size_t lds,
                const SideStream *side, const UnitCounts &n, int set);

// the two-step march's units by CPML variant, in this order in the unit list: meeting no slab,
// only x slabs, only y slabs, only z slabs, and the rest
struct PairCounts {
    int lean, x, y, z, rest;
};
template <typename T>
struct PairTaps;
// Two steps in one pass of the two-step march (march2s.hpp) over A.units, each class of
// PairCounts in the variant compiled for its psi work (the CPML classes on the side stream where
// one is given), ry rows a strip.
template <typename T, int MODE>
int launch_pair(const StepArgs<T> &A, const PairTaps<T> &TP, int ry, hipStream_t s, size_t lds,
                const PairStreams *side, const PairCounts &n);

}  // namespace pxfdtd
