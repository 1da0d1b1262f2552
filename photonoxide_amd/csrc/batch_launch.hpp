// The server's batched step kernels: every admitted job's march units in one launch per kernel
// class. Declared here for the server (serving.hpp); the kernels are compiled in
// batch_launch.hip.
#pragma once
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace pxfdtd {

template <typename T>
struct JobStep {
    StepArgs<T> A[2];  // A[q]: reads copy q, writes copy 1 - q, the parameters of parity q
};
struct JobSeg {
    int32_t active;  // steps the job takes this segment (0: none)
    int32_t p0;      // the copy its first step reads
};
struct BatchBlock {
    int32_t job, first;  // a workgroup's job and its first of four units
};

// One step s of a segment for a kernel class: `kind` 3 (the aligned march, v rows per strip) or
// 4 (the 2D march, v its field set), coefficient mode `mode`, CPML variant `ax`. blocks[0,
// *count) (device memory: the active jobs' workgroups this segment) out of at most `capacity`.
void launch_batch_class(int kind, int mode, int v, int ax, const JobStep<float> *jobs, const JobSeg *seg,
                        const BatchBlock *blocks, const int *count, const int4 *units, int capacity, int s,
                        hipStream_t stream);

}  // namespace pxfdtd
