// Post-processing on the device: harmonic inversion (src/fdtd/harmonic.rs:62-184) of a signal
// held in device memory, so that a probe's record feeding an optimizer loop stays in HBM. The
// signal-sized work, the sums f, g, z^-M and the diagonal weighted sum at each of the filter
// diagonalization's points (harmonic.hpp sums_at and diagonal, O(N) each), runs one thread per
// point and shift p; only those values come back, and the small dense eigenproblem
// (harmonic_solve) runs on the host. Each thread adds its terms in the host's order with the
// host's complex products written out ((a c - b d) + (a d + b c) i, a real weight scaling
// both parts), and takes the points' reciprocals from the host, so the sums, and every result,
// are the host's bit for bit.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/pxfdtd.h"
#include "common.hpp"
#include "device.hpp"
#include "harmonic.hpp"

int px_fail(int code, const char *fmt, ...);  // resident.hip

namespace {

// pts[q] = (z, 1/z); out[q] = (f, g, z^-M, diagonal), complex each
__global__ void __launch_bounds__(256) harmonic_sums(const double *__restrict__ c, int m, int p,
                                                     const double *__restrict__ pts, int npts,
                                                     double *__restrict__ out) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= npts) return;
    const double zr = pts[4 * q], zi = pts[4 * q + 1], ir = pts[4 * q + 2], ii = pts[4 * q + 3];
    // f = sum_{l=0..M} c_(l+p) z^-l, the power by repeated products
    double fr = 0.0, fi = 0.0, pr = 1.0, pi = 0.0;
    for (int l = 0; l <= m; l++) {
        const double cr = c[2 * (l + p)], ci = c[2 * (l + p) + 1];
        fr += cr * pr - ci * pi;
        fi += cr * pi + ci * pr;
        const double nr = pr * ir - pi * ii, ni = pr * ii + pi * ir;
        pr = nr;
        pi = ni;
    }
    const double mr = pr * zr - pi * zi, mi = pr * zi + pi * zr;  // z^-M
    // g = sum_{l=M+1..2M} c_(l+p) z^-(l-M-1)
    double gr = 0.0, gi = 0.0, qr = 1.0, qi = 0.0;
    for (int l = m + 1; l <= 2 * m; l++) {
        const double cr = c[2 * (l + p)], ci = c[2 * (l + p) + 1];
        gr += cr * qr - ci * qi;
        gi += cr * qi + ci * qr;
        const double nr = qr * ir - qi * ii, ni = qr * ii + qi * ir;
        qr = nr;
        qi = ni;
    }
    // the diagonal: sum_{l=0..2M} (M - |M - l| + 1) c_(l+p) z^-l
    double dr = 0.0, di = 0.0, wr = 1.0, wi = 0.0;
    for (int l = 0; l <= 2 * m; l++) {
        const double weight = (double)(m + 1 - abs(l - m));
        const double ar = weight * c[2 * (l + p)], ai = weight * c[2 * (l + p) + 1];
        dr += ar * wr - ai * wi;
        di += ar * wi + ai * wr;
        const double nr = wr * ir - wi * ii, ni = wr * ii + wi * ir;
        wr = nr;
        wi = ni;
    }
    double *o = out + 8 * q;
    o[0] = fr;
    o[1] = fi;
    o[2] = gr;
    o[3] = gi;
    o[4] = mr;
    o[5] = mi;
    o[6] = dr;
    o[7] = di;
}

__global__ void __launch_bounds__(256) nonfinite(const double *__restrict__ c, long long n, int *__restrict__ flag) {
    const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= 2 * n) return;
    if (!isfinite(c[v])) atomicOr(flag, 1);
}

struct DeviceSignal : pxharm::Evaluator {
    const double *c;
    int m;
    hipStream_t stream;
    DeviceSignal(const double *signal, int m_, hipStream_t s) : c(signal), m(m_), stream(s) {}
    bool evaluate(int p, const std::vector<pxharm::cd> &z, std::vector<pxharm::Sums> &sums,
                  std::vector<pxharm::cd> &diag, std::string *why) const override {
        const int n = (int)z.size();
        std::vector<double> pts(4 * (size_t)n), out(8 * (size_t)n);
        for (int q = 0; q < n; q++) {
            const pxharm::cd zi = 1.0 / z[q];  // the host's quotient, as sums_at takes it
            pts[4 * q] = z[q].real();
            pts[4 * q + 1] = z[q].imag();
            pts[4 * q + 2] = zi.real();
            pts[4 * q + 3] = zi.imag();
        }
        double *d = nullptr;
        hipError_t e = hipMalloc(&d, sizeof(double) * 12 * (size_t)std::max(n, 1));
        if (e != hipSuccess) {
            *why = std::string("hipMalloc: ") + hipGetErrorString(e);
            return false;
        }
        double *dpts = d, *dout = d + 4 * (size_t)n;
        e = hipMemcpyAsync(dpts, pts.data(), sizeof(double) * pts.size(), hipMemcpyHostToDevice, stream);
        if (e == hipSuccess && n) {
            hipLaunchKernelGGL(harmonic_sums, dim3((n + 255) / 256), dim3(256), 0, stream, c, m, p, dpts, n, dout);
            e = hipGetLastError();
        }
        if (e == hipSuccess)
            e = hipMemcpyAsync(out.data(), dout, sizeof(double) * out.size(), hipMemcpyDeviceToHost, stream);
        if (e == hipSuccess) e = hipStreamSynchronize(stream);
        (void)hipFree(d);
        if (e != hipSuccess) {
            *why = std::string("the device's sums: ") + hipGetErrorString(e);
            return false;
        }
        sums.resize(n);
        diag.resize(n);
        for (int q = 0; q < n; q++) {
            const double *o = out.data() + 8 * q;
            sums[q].f = {o[0], o[1]};
            sums[q].g = {o[2], o[3]};
            sums[q].zm = {o[4], o[5]};
            diag[q] = {o[6], o[7]};
        }
        return true;
    }
};

}  // namespace

extern "C" int pxfdtd_harmonic_inversion_device(pxfdtd_device *dev, const double *signal, int64_t n, double dt,
                                                double low, double high, int32_t capacity, int32_t *found,
                                                double *out, void *stream) {
    *found = 0;
    if (!dev) return px_fail(PXFDTD_ERR_INVALID, "harmonic inversion on the device needs a device");
    if (n < 0 || (n > 0 && !signal)) return px_fail(PXFDTD_ERR_INVALID, "harmonic inversion needs a signal");
    if (n > (1LL << 30)) return px_fail(PXFDTD_ERR_INVALID, "harmonic inversion takes at most 2^30 samples");
    HIPCHK(hipSetDevice(dev->ordinal));
    hipStream_t s = (hipStream_t)stream;
    // the samples must be finite, as the host's check asks
    int finite = 1;
    if (n > 0) {
        int *flag = nullptr;
        HIPCHK(hipMalloc(&flag, sizeof(int)));
        hipError_t e = hipMemsetAsync(flag, 0, sizeof(int), s);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(nonfinite, dim3((unsigned)((2 * n + 255) / 256)), dim3(256), 0, s, signal, (long long)n,
                               flag);
            e = hipGetLastError();
        }
        int h = 0;
        if (e == hipSuccess) e = hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        (void)hipFree(flag);
        HIPCHK(e);
        finite = h == 0;
    }
    int m;
    std::vector<pxharm::cd> z;
    std::string why;
    if (!pxharm::harmonic_setup(n, finite != 0, dt, low, high, m, z, &why))
        return px_fail(PXFDTD_ERR_INVALID, "%s", why.c_str());
    std::vector<pxharm::Resonance> r;
    if (!pxharm::harmonic_solve(DeviceSignal(signal, m, s), m, z, dt, low, high, r, &why))
        return px_fail(PXFDTD_ERR_DEVICE, "%s", why.c_str());
    *found = (int32_t)r.size();
    for (int32_t q = 0; q < std::min<int32_t>(capacity, (int32_t)r.size()); q++) {
        out[5 * q] = r[q].frequency;
        out[5 * q + 1] = r[q].decay;
        out[5 * q + 2] = r[q].amplitude.real();
        out[5 * q + 3] = r[q].amplitude.imag();
        out[5 * q + 4] = r[q].error;
    }
    return PXFDTD_OK;
}
