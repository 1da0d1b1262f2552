// Post-processing on the device: a flux monitor's flux from transform sums that stay in HBM
// (FluxPlane::flux, src/fdtd/monitors.rs:318-350, summed over faces as Simulation::flux,
// monitors.rs:530-547), so that an optimizer loop reads nf values instead of the planes' sums.
//
// Floating-point addition is not associative, so the sum is kept in the reference's order and
// only the terms are computed in parallel: one thread per term writes weight x sign x
// Re(E_mean conj(H~)) into a terms buffer (0 where the weight is 0, which the sequential sum
// skips: the running sum starts at +0.0 and can never become -0.0 under round-to-nearest, so
// adding +0.0 leaves it as it is), then one thread per frequency adds its terms in order. The
// result is the host's sequential sum bit for bit.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>

#include "../../include/pxfdtd.h"
#include "common.hpp"
#include "device.hpp"

namespace {

struct FaceK {
    const double *e[2];  // E along b, E along c
    const double *h[2];  // H along c, H along b (the partners of e[0], e[1])
    int axis, b, c, plane;
    int from[2], nu, nv;
    int windowed[2], wlo[2], whi[2];
    int lo[3], ex, ey, ez;
    double sign, step_b, step_c;
    long long base;  // the face's first term in the terms buffer
};

struct FacesK {
    FaceK f[PXFDTD_FLUX_MAX_FACES];
    int n, nf;
    long long total;  // terms of all faces
};

// FluxPlane::weight: a value's share of the window, halfway after node m (`half`) or on it
__device__ __forceinline__ double weight_of(const FaceK &F, int w, int m, bool half) {
    if (!F.windowed[w]) return 1.0;
    const int lo = F.wlo[w], hi = F.whi[w];
    if (half) return (lo < m && m < hi) ? 1.0 : ((m == lo || m == hi) ? 0.5 : 0.0);
    return (lo < m && m <= hi) ? 1.0 : 0.0;
}

// one thread per term: face q, frequency f, then the reference's loops (component pair, v, u)
__global__ void __launch_bounds__(256) flux_terms(FacesK P, double *__restrict__ terms) {
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= P.total) return;
    int q = 0;
    while (q + 1 < P.n && g >= P.f[q + 1].base) q++;
    const FaceK &F = P.f[q];
    const long long plane_terms = (long long)F.nu * F.nv;
    long long t = g - F.base;
    const int f = (int)(t / (2 * plane_terms));
    t -= (long long)f * 2 * plane_terms;
    const int pair = (int)(t / plane_terms);
    t -= pair * plane_terms;
    const int v = F.from[1] + (int)(t / F.nu), u = F.from[0] + (int)(t % F.nu);
    const int comp = pair == 0 ? F.b : F.c;
    const double sign = pair == 0 ? 0.5 : -0.5;
    const double weight = weight_of(F, 0, u, comp == F.b) * weight_of(F, 1, v, comp == F.c);
    double term = 0.0;
    if (weight != 0.0) {
        int at[3];
        at[F.axis] = F.plane;
        at[F.b] = u;
        at[F.c] = v;
        auto index = [&](const int *p) {
            return (((long long)f * F.ez + (p[2] - F.lo[2])) * F.ey + (p[1] - F.lo[1])) * F.ex + (p[0] - F.lo[0]);
        };
        const long long lo_at = index(at);
        at[F.axis] = F.plane + 1;
        const long long hi_at = index(at);
        const double *E = F.e[pair], *H = F.h[pair];
        const double lre = E[2 * lo_at], lim = E[2 * lo_at + 1];
        const double hre = E[2 * hi_at], him = E[2 * hi_at + 1];
        const double vre = H[2 * lo_at], vim = H[2 * lo_at + 1];
        const double mre = 0.5 * (lre + hre), mim = 0.5 * (lim + him);
        // Re(m conj(hv)) as complex multiplication forms it: m.re hv.re - m.im (-hv.im)
        term = weight * sign * (mre * vre - mim * -vim);
    }
    terms[g] = term;
}

// one thread per frequency: each face's terms in order, then the faces' signed fluxes from -0.0
__global__ void __launch_bounds__(64) flux_sum(FacesK P, const double *__restrict__ terms,
                                               double *__restrict__ out) {
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= P.nf) return;
    double total = -0.0;
    for (int q = 0; q < P.n; q++) {
        const FaceK &F = P.f[q];
        const long long len = 2LL * F.nu * F.nv;
        const double *t = terms + F.base + (long long)f * len;
        double face = 0.0;
        for (long long m = 0; m < len; m++) face += t[m];
        total += F.sign * (face * F.step_b * F.step_c);
    }
    out[f] = total;
}

}  // namespace

// The allocation-free flux's terms buffer: one array per GPU, held until the process ends.
// Host code uses the address hipGetSymbolAddress gives for the current device (scratch_of),
// never the array's name, which on the host is only a stand-in.
static constexpr long long kScratchValues = 1LL << 21;  // 16 MiB of FP64 terms
static __device__ __align__(256) double scratch[kScratchValues];

static std::mutex g_scratch_mutex;
static double *g_scratch_at[64] = {};

// Once per device, with that device current: the device's own address of `scratch`.
int flux_scratch_setup(int ordinal, double **out) {
    if (ordinal < 0 || ordinal >= 64) return px_fail(PXFDTD_ERR_INVALID, "no scratch for device %d", ordinal);
    std::lock_guard<std::mutex> g(g_scratch_mutex);
    if (!g_scratch_at[ordinal]) {
        int prev = 0;
        HIPCHK(hipGetDevice(&prev));
        HIPCHK(hipSetDevice(ordinal));
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        (void)hipSetDevice(prev);
        if (e != hipSuccess || !p)
            return px_fail(PXFDTD_ERR_DEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", ordinal,
                           hipGetErrorString(e));
        g_scratch_at[ordinal] = (double *)p;
    }
    *out = g_scratch_at[ordinal];
    return PXFDTD_OK;
}

// Checks the faces and lays out their kernel parameters and terms.
static int flux_plan(const pxfdtd_flux_face *faces, int32_t n, FacesK &P) {
    if (n < 1 || n > PXFDTD_FLUX_MAX_FACES)
        return px_fail(PXFDTD_ERR_INVALID, "a flux takes 1 to %d faces, not %d", PXFDTD_FLUX_MAX_FACES, n);
    P = FacesK{};
    P.n = n;
    P.nf = faces[0].nf;
    if (P.nf < 1) return px_fail(PXFDTD_ERR_INVALID, "a flux needs a frequency, got %d", P.nf);
    long long base = 0;
    for (int q = 0; q < n; q++) {
        const pxfdtd_flux_face &S = faces[q];
        FaceK &F = P.f[q];
        if (S.nf != P.nf)
            return px_fail(PXFDTD_ERR_INVALID, "face %d has %d frequencies where face 0 has %d", q, S.nf, P.nf);
        if (S.axis < 0 || S.axis > 2) return px_fail(PXFDTD_ERR_INVALID, "face %d: no axis %d", q, S.axis);
        F.axis = S.axis;
        F.b = (S.axis + 1) % 3;
        F.c = (S.axis + 2) % 3;
        F.plane = S.plane;
        for (int a = 0; a < 3; a++)
            if (S.lo[a] < 0 || S.hi[a] <= S.lo[a])
                return px_fail(PXFDTD_ERR_INVALID, "face %d: its transform's box is empty along axis %d", q, a);
        if (S.plane < S.lo[S.axis] || S.plane + 1 >= S.hi[S.axis])
            return px_fail(PXFDTD_ERR_INVALID, "face %d: nodes %d and %d along axis %d are not in its transform",
                           q, S.plane, S.plane + 1, S.axis);
        for (int w = 0; w < 2; w++) {
            const int along = w == 0 ? F.b : F.c;
            if (S.from[w] >= S.to[w] || S.from[w] < S.lo[along] || S.to[w] > S.hi[along])
                return px_fail(PXFDTD_ERR_INVALID, "face %d: values %d to %d along axis %d are not in its transform",
                               q, S.from[w], S.to[w], along);
            if (S.windowed[w] && !(S.wlo[w] < S.whi[w] && S.from[w] == S.wlo[w] && S.to[w] == S.whi[w] + 1))
                return px_fail(PXFDTD_ERR_INVALID,
                               "face %d: a window from %d to %d along axis %d runs over values %d to %d, not %d to %d",
                               q, S.wlo[w], S.whi[w], along, S.wlo[w], S.whi[w] + 1, S.from[w], S.to[w]);
            F.from[w] = S.from[w];
            F.windowed[w] = S.windowed[w] ? 1 : 0;
            F.wlo[w] = S.wlo[w];
            F.whi[w] = S.whi[w];
        }
        F.nu = S.to[0] - S.from[0];
        F.nv = S.to[1] - S.from[1];
        for (int a = 0; a < 3; a++) F.lo[a] = S.lo[a];
        F.ex = S.hi[0] - S.lo[0];
        F.ey = S.hi[1] - S.lo[1];
        F.ez = S.hi[2] - S.lo[2];
        F.e[0] = S.series[F.b];
        F.e[1] = S.series[F.c];
        F.h[0] = S.series[3 + F.c];
        F.h[1] = S.series[3 + F.b];
        for (int m = 0; m < 2; m++)
            if (!F.e[m] || !F.h[m]) return px_fail(PXFDTD_ERR_INVALID, "face %d: a series it reads is null", q);
        F.sign = S.sign;
        F.step_b = S.step_b;
        F.step_c = S.step_c;
        F.base = base;
        base += 2LL * P.nf * F.nu * F.nv;
    }
    P.total = base;
    return PXFDTD_OK;
}

static int flux_launch(const FacesK &P, double *terms, double *out, hipStream_t stream) {
    const long long blocks = (P.total + 255) / 256;
    if (blocks > 0x7fffffffLL) return px_fail(PXFDTD_ERR_INVALID, "a flux of %lld terms is too many", P.total);
    flux_terms<<<(unsigned)blocks, 256, 0, stream>>>(P, terms);
    HIPCHK(hipGetLastError());
    flux_sum<<<(P.nf + 63) / 64, 64, 0, stream>>>(P, terms, out);
    HIPCHK(hipGetLastError());
    return PXFDTD_OK;
}

extern "C" int pxfdtd_flux_device(pxfdtd_device *dev, const pxfdtd_flux_face *faces, int32_t n, double *out,
                                  void *stream) {
    FacesK P;
    TRY(flux_plan(faces, n, P));
    if (!out) return px_fail(PXFDTD_ERR_INVALID, "no output for the flux");
    HIPCHK(hipSetDevice(dev->ordinal));
    hipStream_t s = (hipStream_t)stream;
    double *terms = nullptr;
    HIPCHK(hipMalloc(&terms, P.total * sizeof(double)));
    int rc = flux_launch(P, terms, out, s);
    const hipError_t e = hipStreamSynchronize(s);
    (void)hipFree(terms);
    TRY(rc);
    if (e != hipSuccess) return px_fail(PXFDTD_ERR_DEVICE, "the flux: %s", hipGetErrorString(e));
    return PXFDTD_OK;
}

extern "C" int pxfdtd_flux_device_no_alloc(pxfdtd_device *dev, const pxfdtd_flux_face *faces, int32_t n,
                                           double *out, void *stream) {
    FacesK P;
    TRY(flux_plan(faces, n, P));
    if (!out) return px_fail(PXFDTD_ERR_INVALID, "no output for the flux");
    if (P.total > kScratchValues)
        return px_fail(PXFDTD_ERR_INVALID,
                       "the flux needs %lld terms; the allocation-free flux holds at most %lld (16 MiB)", P.total,
                       kScratchValues);
    if (!dev->scratch) return px_fail(PXFDTD_ERR_DEVICE, "device %d has no scratch address", dev->ordinal);
    HIPCHK(hipSetDevice(dev->ordinal));
    // every term is written before it is read: nothing in the scratch needs clearing
    return flux_launch(P, dev->scratch, out, (hipStream_t)stream);
}

// The host mirror's flux: on the device, nf values copied back to `out` (host memory).
int flux_to_host(pxfdtd_device *dev, const pxfdtd_flux_face *faces, int32_t n, double *out, void *stream) {
    FacesK P;
    TRY(flux_plan(faces, n, P));
    HIPCHK(hipSetDevice(dev->ordinal));
    double *dout = nullptr;
    HIPCHK(hipMalloc(&dout, P.nf * sizeof(double)));
    int rc = pxfdtd_flux_device(dev, faces, n, dout, stream);
    if (rc == PXFDTD_OK) {
        const hipError_t e = hipMemcpy(out, dout, P.nf * sizeof(double), hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = px_fail(PXFDTD_ERR_DEVICE, "reading the flux back: %s", hipGetErrorString(e));
    }
    (void)hipFree(dout);
    return rc;
}
