// Coefficient construction on the device: the permittivity each value of E sees, averaged
// over 8^3 samples of its cell, harmonic along the component and arithmetic across it
// (Simulation::diagonal, src/fdtd.rs:619-653, over fdfd::averaged, src/fdfd/three.rs:563-588),
// the O(512 N x shapes) part of a 3D set-up. One thread per value computes the same sequence of
// IEEE 754 double operations as the host's restatement (simulation.cpp `averaged`), so the
// averages are its bits.
//
// Two exact shortcuts keep most values from sampling every shape:
//  - each shape's reciprocal e / (e e + 0 0) is computed once: the same operands give the same
//    bits in every sample;
//  - a shape whose region holds none of the cell's samples cannot decide any of them, and a
//    later shape that holds them all decides every one, so the samples are taken against the
//    last such covering shape (or the background) and the shapes after it that reach the
//    cell. When none reaches it, every sample has the same value and the average is that
//    value's uniform average, computed once on the host by the same function.
// The samples lie between the extreme ones (rounding is monotonic), which decide the box
// tests exactly; the cylinder tests are decided with a relative margin of 1e-9 on the squared
// radius, far beyond the rounding of the per-sample distance, and an undecided cylinder is
// sampled.
#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/pxfdtd.h"
#include "common.hpp"
#include "device.hpp"

namespace {

struct ShapeK {
    int kind;
    double lo[3], hi[3];
    double value, inv;  // the component's value and its reciprocal as the samples add it
    double uniform;     // the uniform average of `value`
};

struct GridK {
    int nx, ny, nz;
    double h[3], o[3];
};

constexpr int kSamples = 8;
constexpr int kList = 12;  // shapes after the covering one a value's samples test in registers

__device__ __forceinline__ bool inside(const ShapeK &s, const double p[3]) {
    if (s.kind == 0)
        return s.lo[0] < p[0] && p[0] < s.hi[0] && s.lo[1] < p[1] && p[1] < s.hi[1] && s.lo[2] < p[2] && p[2] < s.hi[2];
    const double dx = p[0] - s.lo[0], dy = p[1] - s.lo[1];
    return dx * dx + dy * dy < s.hi[0] * s.hi[0] && s.lo[2] < p[2] && p[2] < s.hi[2];
}

// whether the shape holds every sample of the box [a, b] of sample positions, and whether it
// may hold any
__device__ __forceinline__ void cover(const ShapeK &s, const double a[3], const double b[3], bool &all, bool &any) {
    if (s.kind == 0) {
        all = true;
        any = true;
        for (int d = 0; d < 3; d++) {
            all = all && s.lo[d] < a[d] && b[d] < s.hi[d];
            any = any && b[d] > s.lo[d] && a[d] < s.hi[d];
        }
        return;
    }
    const bool zall = s.lo[2] < a[2] && b[2] < s.hi[2];
    const bool zany = b[2] > s.lo[2] && a[2] < s.hi[2];
    const double r2 = s.hi[0] * s.hi[0];
    const double x0 = a[0] - s.lo[0], x1 = b[0] - s.lo[0], y0 = a[1] - s.lo[1], y1 = b[1] - s.lo[1];
    const double far = fmax(x0 * x0, x1 * x1) + fmax(y0 * y0, y1 * y1);
    const double nx = x0 > 0 ? x0 : (x1 < 0 ? x1 : 0.0), ny = y0 > 0 ? y0 : (y1 < 0 ? y1 : 0.0);
    const double near = nx * nx + ny * ny;
    all = zall && far < r2 * (1.0 - 1e-9);
    any = zany && near < r2 * (1.0 + 1e-9);
}

__global__ void __launch_bounds__(256) averaged_eps(int comp, GridK g, const ShapeK *__restrict__ shapes, int ns,
                                                    double bg_inv, double bg_uniform, double *__restrict__ out) {
    const long long N = (long long)g.nx * g.ny * g.nz;
    const long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= N) return;
    const int i = (int)(r % g.nx), j = (int)((r / g.nx) % g.ny), k = (int)(r / ((long long)g.nx * g.ny));
    double centre[3] = {g.o[0] + i * g.h[0], g.o[1] + j * g.h[1], g.o[2] + k * g.h[2]};
    centre[comp] += 0.5 * g.h[comp];
    const double n = kSamples;
    double a[3], b[3];
    for (int d = 0; d < 3; d++) {
        a[d] = centre[d] + ((0 + 0.5) / n - 0.5) * g.h[d];
        b[d] = centre[d] + ((kSamples - 1 + 0.5) / n - 0.5) * g.h[d];
    }
    // the last shape holding every sample, and the shapes after it that may hold some
    int base = -1, list[kList], nl = 0;
    bool overflow = false;
    for (int q = 0; q < ns; q++) {
        bool all, any;
        cover(shapes[q], a, b, all, any);
        if (all) {
            base = q;
            nl = 0;
            overflow = false;
        } else if (any) {
            if (nl < kList) list[nl++] = q;
            else overflow = true;
        }
    }
    if (nl == 0 && !overflow) {
        out[r] = base < 0 ? bg_uniform : shapes[base].uniform;
        return;
    }
    const double base_inv = base < 0 ? bg_inv : shapes[base].inv;
    const int ax = comp, bx = (comp + 1) % 3, cx = (comp + 2) % 3;
    double sum = 0.0, p[3] = {centre[0], centre[1], centre[2]};
    for (int sb = 0; sb < kSamples; sb++) {
        p[bx] = centre[bx] + ((sb + 0.5) / n - 0.5) * g.h[bx];
        for (int sc = 0; sc < kSamples; sc++) {
            p[cx] = centre[cx] + ((sc + 0.5) / n - 0.5) * g.h[cx];
            double inverse = 0.0;
            for (int sa = 0; sa < kSamples; sa++) {
                p[ax] = centre[ax] + ((sa + 0.5) / n - 0.5) * g.h[ax];
                double v = base_inv;
                if (!overflow) {
                    for (int m = 0; m < nl; m++)
                        if (inside(shapes[list[m]], p)) v = shapes[list[m]].inv;
                } else {
                    for (int q = base + 1; q < ns; q++)
                        if (inside(shapes[q], p)) v = shapes[q].inv;
                }
                inverse += v;
            }
            sum += n * inverse / (inverse * inverse + 0.0 * 0.0);
        }
    }
    out[r] = sum / (n * n);
}

struct DevMem {
    void *p = nullptr;
    ~DevMem() {
        if (p) (void)hipFree(p);
    }
};

}  // namespace

// simulation.cpp: the host's average of a scene's component at one cell (the restatement the
// device reproduces), used here for the uniform averages
namespace pxfdtd_host {
double averaged_at(const pxfdtd_scene *sc, int comp, const double centre[3], const double h[3]);
}

// The averaged permittivity of every value of E, [3][N], into host memory `eps3n`.
int averaged_on_device(pxfdtd_device *dev, const pxfdtd_grid *g, const pxfdtd_scene *sc, double *eps3n) {
    HIPCHK(hipSetDevice(dev->ordinal));
    const long long N = (long long)g->nx * g->ny * g->nz;
    const int ns = sc->n;
    GridK G = {g->nx, g->ny, g->nz, {g->dx, g->dy, g->dz}, {g->x0, g->y0, g->z0}};
    const double hs[3] = {g->dx, g->dy, g->dz}, o[3] = {0.0, 0.0, 0.0};
    DevMem shapes, out;
    HIPCHK(hipMalloc(&shapes.p, std::max(1, ns) * sizeof(ShapeK)));
    HIPCHK(hipMalloc(&out.p, N * sizeof(double)));
    hipStream_t st = nullptr;
    HIPCHK(hipStreamCreate(&st));
    int rc = PXFDTD_OK;
    for (int c = 0; c < 3 && rc == PXFDTD_OK; c++) {
        // each shape's value for this component, its reciprocal as a sample adds it, and the
        // uniform average of the value (the host's function over a scene of that value alone)
        std::vector<ShapeK> K(ns);
        for (int q = 0; q < ns; q++) {
            const pxfdtd_shape &s = sc->shapes[q];
            K[q].kind = s.kind;
            for (int d = 0; d < 3; d++) {
                K[q].lo[d] = s.lo[d];
                K[q].hi[d] = s.hi[d];
            }
            const double e = s.value[c];
            K[q].value = e;
            K[q].inv = 1.0 * e / (e * e + 0.0 * 0.0);
            pxfdtd_scene u = {{e, e, e}, 0, nullptr};
            K[q].uniform = pxfdtd_host::averaged_at(&u, c, o, hs);
        }
        const double e = sc->background[c];
        const double bg_inv = 1.0 * e / (e * e + 0.0 * 0.0);
        pxfdtd_scene u = {{e, e, e}, 0, nullptr};
        const double bg_uniform = pxfdtd_host::averaged_at(&u, c, o, hs);
        hipError_t err = hipSuccess;
        if (ns) err = hipMemcpyAsync(shapes.p, K.data(), ns * sizeof(ShapeK), hipMemcpyHostToDevice, st);
        if (err == hipSuccess) {
            const long long blocks = (N + 255) / 256;
            hipLaunchKernelGGL(averaged_eps, dim3((unsigned)blocks), dim3(256), 0, st, c, G, (const ShapeK *)shapes.p,
                               ns, bg_inv, bg_uniform, (double *)out.p);
            err = hipGetLastError();
        }
        if (err == hipSuccess)
            err = hipMemcpyAsync(eps3n + c * N, out.p, N * sizeof(double), hipMemcpyDeviceToHost, st);
        if (err == hipSuccess) err = hipStreamSynchronize(st);
        if (err != hipSuccess) rc = px_fail(PXFDTD_ERR_DEVICE, "averaging on the device: %s", hipGetErrorString(err));
    }
    (void)hipStreamSynchronize(st);
    (void)hipStreamDestroy(st);
    return rc;
}
