// The server's batched step kernels (declared in batch_launch.hpp), compiled apart from the
// rest of the library so that the build runs in parallel.
#include "batch_launch.hpp"
#include "march64.hpp"
#include "march2d.hpp"

namespace pxfdtd {

template <typename T, int MODE, int RY, int AX>
__global__ __launch_bounds__(256) void march64_batch(const JobStep<T> *jobs, const JobSeg *seg,
                                                     const BatchBlock *blocks, const int4 *units,
                                                     int nblocks, int per_xcd, int s) {
    __shared__ MarchShared<T> S;
    extern __shared__ unsigned char smem_raw[];
    T *tab = reinterpret_cast<T *>(smem_raw);
    int8_t *grp = reinterpret_cast<int8_t *>(tab + 256 * 6);
    const int b = (int)(blockIdx.x % 8) * per_xcd + (int)(blockIdx.x / 8);
    if (b >= nblocks) return;  // the whole workgroup
    const BatchBlock B = blocks[b];
    const JobSeg g = seg[B.job];
    if (s >= g.active) return;  // the whole workgroup: its job is done
    StepArgs<T> A = jobs[B.job].A[(g.p0 + s) & 1];
    A.wave_step = s;
    march64_stage<T, MODE, AX != 0>(A, S, tab, grp);
    __syncthreads();
    const int4 u = units[uniform(B.first + (int)(threadIdx.x >> 6))];
    if (u.w <= u.z) return;  // a padding unit
    march64_unit<T, MODE, RY, AX>(A, u, S, tab, grp);
}

template <typename T, int MODE, int AX, int SET>
__global__ __launch_bounds__(256) void march2d_batch(const JobStep<T> *jobs, const JobSeg *seg,
                                                     const BatchBlock *blocks, const int4 *units,
                                                     int nblocks, int per_xcd, int s) {
    __shared__ MarchShared<T> S;
    extern __shared__ unsigned char smem_raw[];
    T *tab = reinterpret_cast<T *>(smem_raw);
    int8_t *grp = reinterpret_cast<int8_t *>(tab + 256 * 6);
    const int b = (int)(blockIdx.x % 8) * per_xcd + (int)(blockIdx.x / 8);
    if (b >= nblocks) return;
    const BatchBlock B = blocks[b];
    const JobSeg g = seg[B.job];
    if (s >= g.active) return;
    StepArgs<T> A = jobs[B.job].A[(g.p0 + s) & 1];
    A.wave_step = s;
    march64_stage<T, MODE, AX != 0>(A, S, tab, grp);
    __syncthreads();
    const int4 u = units[uniform(B.first + (int)(threadIdx.x >> 6))];
    if (u.w <= u.z) return;
    march2d_unit<T, MODE, AX, SET>(A, u, S, tab, grp);
}

template <int MODE, int RY, int AX>
static void launch_3d(const JobStep<float> *jobs, const JobSeg *seg, const BatchBlock *blocks, const int4 *units,
                      int count, int s, size_t lds, hipStream_t stream) {
    const int per_xcd = (count + 7) / 8;
    hipLaunchKernelGGL((march64_batch<float, MODE, RY, AX>), dim3(8 * per_xcd), dim3(256), lds, stream, jobs,
                       seg, blocks, units, count, per_xcd, s);
}

template <int MODE, int AX, int SET>
static void launch_2d(const JobStep<float> *jobs, const JobSeg *seg, const BatchBlock *blocks, const int4 *units,
                      int count, int s, size_t lds, hipStream_t stream) {
    const int per_xcd = (count + 7) / 8;
    hipLaunchKernelGGL((march2d_batch<float, MODE, AX, SET>), dim3(8 * per_xcd), dim3(256), lds, stream, jobs,
                       seg, blocks, units, count, per_xcd, s);
}

void launch_batch_class(int kind, int mode, int v, int ax, const JobStep<float> *jobs, const JobSeg *seg,
                        const BatchBlock *blocks, const int4 *units, int count, int s, hipStream_t stream) {
    const size_t lds = mode == 0 ? (size_t)256 * (6 * sizeof(float) + 3) : 0;
#define PX_CLASS(M, R, X) \
    if (kind == 3 && mode == M && v == R && ax == X) return launch_3d<M, R, X>(jobs, seg, blocks, units, count, s, lds, stream);
    PX_CLASS(0, 1, 0) PX_CLASS(0, 1, 1) PX_CLASS(0, 1, 7) PX_CLASS(0, 2, 0) PX_CLASS(0, 2, 1) PX_CLASS(0, 2, 7)
    PX_CLASS(1, 1, 0) PX_CLASS(1, 1, 1) PX_CLASS(1, 1, 7) PX_CLASS(1, 2, 0) PX_CLASS(1, 2, 1) PX_CLASS(1, 2, 7)
    PX_CLASS(2, 1, 0) PX_CLASS(2, 1, 1) PX_CLASS(2, 1, 7) PX_CLASS(2, 2, 0) PX_CLASS(2, 2, 1) PX_CLASS(2, 2, 7)
#undef PX_CLASS
#define PX_CLASS2(M, X, E) \
    if (kind == 4 && mode == M && ax == X && v == E) return launch_2d<M, X, E>(jobs, seg, blocks, units, count, s, lds, stream);
    PX_CLASS2(0, 0, 1) PX_CLASS2(0, 0, 2) PX_CLASS2(0, 0, 3) PX_CLASS2(0, 3, 1) PX_CLASS2(0, 3, 2) PX_CLASS2(0, 3, 3)
    PX_CLASS2(1, 0, 1) PX_CLASS2(1, 0, 2) PX_CLASS2(1, 0, 3) PX_CLASS2(1, 3, 1) PX_CLASS2(1, 3, 2) PX_CLASS2(1, 3, 3)
    PX_CLASS2(2, 0, 1) PX_CLASS2(2, 0, 2) PX_CLASS2(2, 0, 3) PX_CLASS2(2, 3, 1) PX_CLASS2(2, 3, 2) PX_CLASS2(2, 3, 3)
#undef PX_CLASS2
}

}  // namespace pxfdtd
