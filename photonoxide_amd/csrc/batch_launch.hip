// The server's batched step kernels (declared in batch_launch.hpp), compiled apart from the
// rest of the library so that the build runs in parallel.
#include "batch_launch.hpp"
#include "march64.hpp"
#include "march2d.hpp"

#include <algorithm>

namespace pxfdtd {

// The batched kernels are persistent: a fixed grid walks this segment's list of the active
// jobs' workgroups (the host compacts it each segment, so waiting and finished jobs cost
// nothing), each XCD's workgroups taking a contiguous run of it so that neighbouring units of
// a job share an L2. `n` is the list's length this segment, read from device memory.
__device__ __forceinline__ int batch_block(int it, int n) {
    const int g8 = (int)(gridDim.x / 8), per = (n + 7) / 8;
    const int x = (int)(blockIdx.x % 8), j = (int)(blockIdx.x / 8) + it * g8;
    return j < per ? x * per + j : n;  // n: none
}

// The tables the kernels read are never written while they run (__restrict__): their loads are
// scalar, and a job's parameters take SGPRs, not a VGPR each. The general slab variant at one row
// is held to 128 VGPRs, four waves per SIMD, where it fits without spilling (not with uint8 ids).
template <typename T, int MODE, int RY, int AX>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(RY == 1 && AX == 7 && MODE != 0 ? 4 : 1))) void march64_batch(const JobStep<T> *__restrict__ jobs, const JobSeg *__restrict__ seg,
                                                     const BatchBlock *__restrict__ blocks,
                                                     const int *__restrict__ count,
                                                     const int4 *__restrict__ units, int s) {
    __shared__ MarchShared<T> S;
    extern __shared__ unsigned char smem_raw[];
    T *tab = reinterpret_cast<T *>(smem_raw);
    int8_t *grp = reinterpret_cast<int8_t *>(tab + 256 * 6);
    const int n = uniform(*count);
    const int rounds = (((n + 7) / 8) + (int)(gridDim.x / 8) - 1) / (int)(gridDim.x / 8);
    for (int it = 0; it < rounds; it++) {
        const int b = batch_block(it, n);
        if (b >= n) break;  // uniform over the workgroup
        const BatchBlock B = blocks[b];
        const JobSeg g = seg[B.job];
        if (s >= g.active) continue;  // the whole workgroup: its job is done for this segment
        StepArgs<T> A = jobs[B.job].A[(g.p0 + s) & 1];
        A.wave_step = s;
        __syncthreads();  // the last block's waves are done with LDS
        march64_stage<T, MODE, AX != 0>(A, S, tab, grp);
        __syncthreads();
        const int4 u = units[uniform(B.first + (int)(threadIdx.x >> 6))];
        if (u.w > u.z) march64_unit<T, MODE, RY, AX>(A, u, S, tab, grp);  // else a padding unit
    }
}

template <typename T, int MODE, int AX, int SET>
__global__ __launch_bounds__(256) void march2d_batch(const JobStep<T> *__restrict__ jobs, const JobSeg *__restrict__ seg,
                                                     const BatchBlock *__restrict__ blocks,
                                                     const int *__restrict__ count,
                                                     const int4 *__restrict__ units, int s) {
    __shared__ MarchShared<T> S;
    extern __shared__ unsigned char smem_raw[];
    T *tab = reinterpret_cast<T *>(smem_raw);
    int8_t *grp = reinterpret_cast<int8_t *>(tab + 256 * 6);
    const int n = uniform(*count);
    const int rounds = (((n + 7) / 8) + (int)(gridDim.x / 8) - 1) / (int)(gridDim.x / 8);
    for (int it = 0; it < rounds; it++) {
        const int b = batch_block(it, n);
        if (b >= n) break;
        const BatchBlock B = blocks[b];
        const JobSeg g = seg[B.job];
        if (s >= g.active) continue;
        StepArgs<T> A = jobs[B.job].A[(g.p0 + s) & 1];
        A.wave_step = s;
        __syncthreads();
        march64_stage<T, MODE, AX != 0>(A, S, tab, grp);
        __syncthreads();
        const int4 u = units[uniform(B.first + (int)(threadIdx.x >> 6))];
        if (u.w > u.z) march2d_unit<T, MODE, AX, SET>(A, u, S, tab, grp);
    }
}

// the persistent grid: the class's workgroups when all its jobs step, at most 16384 (a multiple
// of 8, one run per XCD)
static int batch_grid(int capacity) { return 8 * std::max(1, std::min((capacity + 7) / 8, 2048)); }

template <int MODE, int RY, int AX>
static void launch_3d(const JobStep<float> *jobs, const JobSeg *seg, const BatchBlock *blocks, const int *count,
                      const int4 *units, int capacity, int s, size_t lds, hipStream_t stream) {
    hipLaunchKernelGGL((march64_batch<float, MODE, RY, AX>), dim3(batch_grid(capacity)), dim3(256), lds, stream,
                       jobs, seg, blocks, count, units, s);
}

template <int MODE, int AX, int SET>
static void launch_2d(const JobStep<float> *jobs, const JobSeg *seg, const BatchBlock *blocks, const int *count,
                      const int4 *units, int capacity, int s, size_t lds, hipStream_t stream) {
    hipLaunchKernelGGL((march2d_batch<float, MODE, AX, SET>), dim3(batch_grid(capacity)), dim3(256), lds, stream,
                       jobs, seg, blocks, count, units, s);
}

void launch_batch_class(int kind, int mode, int v, int ax, const JobStep<float> *jobs, const JobSeg *seg,
                        const BatchBlock *blocks, const int *count, const int4 *units, int capacity, int s,
                        hipStream_t stream) {
    const size_t lds = mode == 0 ? (size_t)256 * (6 * sizeof(float) + 3) : 0;
#define PX_CLASS(M, R, X) \
    if (kind == 3 && mode == M && v == R && ax == X) return launch_3d<M, R, X>(jobs, seg, blocks, count, units, capacity, s, lds, stream);
    PX_CLASS(0, 1, 0) PX_CLASS(0, 1, 1) PX_CLASS(0, 1, 7) PX_CLASS(0, 2, 0) PX_CLASS(0, 2, 1) PX_CLASS(0, 2, 7)
    PX_CLASS(1, 1, 0) PX_CLASS(1, 1, 1) PX_CLASS(1, 1, 7) PX_CLASS(1, 2, 0) PX_CLASS(1, 2, 1) PX_CLASS(1, 2, 7)
    PX_CLASS(2, 1, 0) PX_CLASS(2, 1, 1) PX_CLASS(2, 1, 7) PX_CLASS(2, 2, 0) PX_CLASS(2, 2, 1) PX_CLASS(2, 2, 7)
#undef PX_CLASS
#define PX_CLASS2(M, X, E) \
    if (kind == 4 && mode == M && ax == X && v == E) return launch_2d<M, X, E>(jobs, seg, blocks, count, units, capacity, s, lds, stream);
    PX_CLASS2(0, 0, 1) PX_CLASS2(0, 0, 2) PX_CLASS2(0, 0, 3) PX_CLASS2(0, 3, 1) PX_CLASS2(0, 3, 2) PX_CLASS2(0, 3, 3)
    PX_CLASS2(1, 0, 1) PX_CLASS2(1, 0, 2) PX_CLASS2(1, 0, 3) PX_CLASS2(1, 3, 1) PX_CLASS2(1, 3, 2) PX_CLASS2(1, 3, 3)
    PX_CLASS2(2, 0, 1) PX_CLASS2(2, 0, 2) PX_CLASS2(2, 0, 3) PX_CLASS2(2, 3, 1) PX_CLASS2(2, 3, 2) PX_CLASS2(2, 3, 3)
#undef PX_CLASS2
}

}  // namespace pxfdtd
