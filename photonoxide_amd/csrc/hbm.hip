// The achievable HBM rate on this device, measured with streaming kernels of our own: read,
// write, copy and triad over buffers far larger than the caches (1 GiB each against 256 MB of
// Infinity Cache), 16-byte accesses, non-temporal stores, each at several grid sizes (4096 to
// 65536 workgroups of 256 in a grid-stride loop) keeping the best. It is the denominator the
// stepper's roofline fraction is also reported against, beside the nominal 8 TB/s, and the
// copy kernel of known size that scales rocprofv3's byte counters.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "../../include/pxfdtd.h"
#include "common.hpp"
#include "device.hpp"

namespace {

constexpr int kThreads = 256;

__global__ void __launch_bounds__(kThreads) hbm_read(const float4 *__restrict__ a, long long n, float *sink) {
    float acc = 0.f;
    for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kThreads) {
        const float4 v = a[i];
        acc += v.x + v.y + v.z + v.w;
    }
    // a store the compiler cannot drop, taken only when the sum is exactly -1 (never, for the
    // buffers filled with +1 here)
    if (acc == -1.f) sink[threadIdx.x] = acc;
}

typedef float f4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void put(float4 *p, float4 v) {
    __builtin_nontemporal_store(f4v{v.x, v.y, v.z, v.w}, (f4v *)p);
}

__global__ void __launch_bounds__(kThreads) hbm_write(float4 *__restrict__ a, long long n, float v) {
    for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kThreads)
        put(a + i, make_float4(v, v, v, v));
}

__global__ void __launch_bounds__(kThreads) hbm_copy(const float4 *__restrict__ a, float4 *__restrict__ b, long long n) {
    for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kThreads)
        put(b + i, a[i]);
}

__global__ void __launch_bounds__(kThreads) hbm_triad(float4 *__restrict__ a, const float4 *__restrict__ b,
                                                      const float4 *__restrict__ c, long long n, float s) {
    for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < n; i += (long long)gridDim.x * kThreads) {
        const float4 x = b[i], y = c[i];
        put(a + i, make_float4(x.x + s * y.x, x.y + s * y.y, x.z + s * y.z, x.w + s * y.w));
    }
}

struct DevBuf {
    void *p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

}  // namespace

extern "C" int pxfdtd_device_hbm_rates(pxfdtd_device *dev, int64_t bytes, int32_t reps, double out[4]) {
    if (bytes < (1 << 20) || bytes % 16 || reps < 1)
        return px_fail(PXFDTD_ERR_INVALID, "HBM rates need buffers of at least 1 MiB in 16-byte units and a repeat, "
                                           "got %lld bytes and %d", (long long)bytes, reps);
    HIPCHK(hipSetDevice(dev->ordinal));
    DevBuf a, b, c, sink;
    HIPCHK(hipMalloc(&a.p, bytes));
    HIPCHK(hipMalloc(&b.p, bytes));
    HIPCHK(hipMalloc(&c.p, bytes));
    HIPCHK(hipMalloc(&sink.p, kThreads * sizeof(float)));
    const long long n = bytes / 16;
    hipStream_t s = nullptr;
    HIPCHK(hipStreamCreate(&s));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = PXFDTD_OK;
    auto fail = [&](hipError_t e, const char *what) {
        if (rc == PXFDTD_OK && e != hipSuccess) rc = px_fail(PXFDTD_ERR_DEVICE, "%s: %s", what, hipGetErrorString(e));
    };
    fail(hipEventCreate(&e0), "hipEventCreate");
    fail(hipEventCreate(&e1), "hipEventCreate");
    const int grids[3] = {4096, 16384, 65536};
    int blocks = grids[0];
    const auto A = (float4 *)a.p, B = (float4 *)b.p, C = (float4 *)c.p;
    if (rc == PXFDTD_OK) {
        hbm_write<<<blocks, kThreads, 0, s>>>(A, n, 1.f);
        hbm_write<<<blocks, kThreads, 0, s>>>(B, n, 1.f);
        hbm_write<<<blocks, kThreads, 0, s>>>(C, n, 1.f);
        fail(hipGetLastError(), "hbm_write");
    }
    for (int kind = 0; kind < 4 && rc == PXFDTD_OK; kind++) {
        auto launch = [&]() {
            if (kind == 0) hbm_read<<<blocks, kThreads, 0, s>>>(A, n, (float *)sink.p);
            if (kind == 1) hbm_write<<<blocks, kThreads, 0, s>>>(B, n, 1.f);
            if (kind == 2) hbm_copy<<<blocks, kThreads, 0, s>>>(A, B, n);
            if (kind == 3) hbm_triad<<<blocks, kThreads, 0, s>>>(A, B, C, n, 1.f);
        };
        float best = 0.f;
        for (int g = 0; g < 3 && rc == PXFDTD_OK; g++) {
            blocks = grids[g];
            launch();  // warm
            fail(hipGetLastError(), "HBM kernel");
            for (int r = 0; r < reps && rc == PXFDTD_OK; r++) {
                fail(hipEventRecord(e0, s), "hipEventRecord");
                launch();
                fail(hipEventRecord(e1, s), "hipEventRecord");
                fail(hipEventSynchronize(e1), "hipEventSynchronize");
                float ms = 0.f;
                fail(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
                if (best == 0.f || ms < best) best = ms;
            }
        }
        const double moved = (double)bytes * (kind < 2 ? 1 : kind == 2 ? 2 : 3);
        out[kind] = rc == PXFDTD_OK ? moved / (best * 1e-3) / 1e9 : 0.0;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(s);
    (void)hipStreamDestroy(s);
    return rc;
}
