/* TEST INFRASTRUCTURE ONLY: part of the CPU oracle. Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, as the checker. The product never links it.
 *
 * Bit-portable maths for the oracle: the functions the reference computes its waveforms, DFT
 * phases and CPML profile with (src/portable/mod.rs:66-138 `exp`, :321-323 `sin_cos`,
 * :343-356 `powi`; src/portable/complex.rs:80-98 `exp`), restated from the published
 * algorithms they port:
 *   - exp: Sun's fdlibm e_exp.c (1993), argument reduction x = k ln2 + r and the Remez
 *     polynomial R(r^2) of degree 5, result scaled through the exponent bits;
 *   - sincos: musl libm sincos.c (the `libm` crate 0.2.16 is its Rust port), i.e. fdlibm's
 *     kernels __sin/__cos on [-pi/4, pi/4] after the medium-size Cody-Waite reduction
 *     __rem_pio2 in three rounds (arguments above 2^20 pi/2 are refused: no waveform or
 *     phase of a run reaches them);
 *   - powi: square and multiply over the binary digits of |n|, then 1/x^|n| for n < 0.
 * Compiled with -ffp-contract=off: every + - * / is one IEEE 754 rounding, as in Rust.
 * Pinned by: the reference's FDTD canary hash (tests/canary.rs:1007-1062), which runs every
 * waveform value of a 200-step run through exp and sincos, and by closed forms in the tests
 * (sin^2 + cos^2, exp(a+b) = exp(a) exp(b) to an ulp, known values).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static inline uint64_t bits_of(double x) { uint64_t u; memcpy(&u, &x, 8); return u; }
static inline double of_bits(uint64_t u) { double x; memcpy(&x, &u, 8); return x; }

double or_exp(double x) {
    const double o_threshold = of_bits(0x40862e42fefa39efULL);
    const double u_threshold = of_bits(0xc0874910d52d3051ULL);
    const double ln2_hi = of_bits(0x3fe62e42fee00000ULL);
    const double ln2_lo = of_bits(0x3dea39ef35793c76ULL);
    const double inv_ln2 = of_bits(0x3ff71547652b82feULL);
    const double two_m1000 = of_bits(0x0170000000000000ULL);
    const double p1 = of_bits(0x3fc555555555553eULL);
    const double p2 = of_bits(0xbf66c16c16bebd93ULL);
    const double p3 = of_bits(0x3f11566aaf25de2cULL);
    const double p4 = of_bits(0xbebbbd41c5d26bf1ULL);
    const double p5 = of_bits(0x3e66376972bea4d0ULL);
    uint32_t hx = (uint32_t)(bits_of(x) >> 32);
    int negative = (int)(hx >> 31);
    hx &= 0x7fffffffu;
    if (hx >= 0x40862e42u) {
        if (isnan(x)) return x;
        if (x > o_threshold) return INFINITY;
        if (x < u_threshold) return 0.0;
    }
    int k;
    double hi = 0.0, lo = 0.0;
    if (hx > 0x3fd62e42u) {
        if (hx < 0x3ff0a2b2u) {
            if (negative) { k = -1; hi = x + ln2_hi; lo = -ln2_lo; }
            else { k = 1; hi = x - ln2_hi; lo = ln2_lo; }
        } else {
            k = (int)(inv_ln2 * x + (negative ? -0.5 : 0.5));
            double t = (double)k;
            hi = x - t * ln2_hi;
            lo = t * ln2_lo;
        }
        x = hi - lo;
    } else if (hx < 0x3e300000u) {
        return 1.0 + x;
    } else {
        k = 0;
    }
    double t = x * x;
    double c = x - t * (p1 + t * (p2 + t * (p3 + t * (p4 + t * p5))));
    if (k == 0) return 1.0 - ((x * c) / (c - 2.0) - x);
    double y = 1.0 - ((lo - (x * c) / (2.0 - c)) - hi);
    if (k >= -1021) return of_bits(bits_of(y) + ((uint64_t)(int64_t)k << 52));
    return of_bits(bits_of(y) + ((uint64_t)(int64_t)(k + 1000) << 52)) * two_m1000;
}

/* fdlibm's kernels on [-pi/4, pi/4]; y is the tail of the reduced argument. */
static double k_sin(double x, double y, int iy) {
    const double s1 = -1.66666666666666324348e-01, s2 = 8.33333333332248946124e-03,
                 s3 = -1.98412698298579493134e-04, s4 = 2.75573137070700676789e-06,
                 s5 = -2.50507602534068634195e-08, s6 = 1.58969099521155010221e-10;
    double z = x * x, w = z * z;
    double r = s2 + z * (s3 + z * s4) + z * w * (s5 + z * s6);
    double v = z * x;
    if (iy == 0) return x + v * (s1 + z * r);
    return x - ((z * (0.5 * y - v * r) - y) - v * s1);
}

static double k_cos(double x, double y) {
    const double c1 = 4.16666666666666019037e-02, c2 = -1.38888888888741095749e-03,
                 c3 = 2.48015872894767294178e-05, c4 = -2.75573143513906633035e-07,
                 c5 = 2.08757232129817482790e-09, c6 = -1.13596475577881948265e-11;
    double z = x * x, w = z * z;
    double r = z * (c1 + z * (c2 + z * c3)) + w * w * (c4 + z * (c5 + z * c6));
    double hz = 0.5 * z;
    double ww = 1.0 - hz;
    return ww + (((1.0 - ww) - hz) + (z * r - x * y));
}

static int rem_pio2_medium(double x, uint32_t ix, double *y0, double *y1) {
    const double to_int = 6755399441055744.0; /* 1.5 / DBL_EPSILON */
    const double inv_pio2 = 6.36619772367581382433e-01;
    const double pio2_1 = 1.57079632673412561417e+00, pio2_1t = 6.07710050650619224932e-11;
    const double pio2_2 = 6.07710050630396597660e-11, pio2_2t = 2.02226624879595063154e-21;
    const double pio2_3 = 2.02226624871116645580e-21, pio2_3t = 8.47842766036889956997e-32;
    double fn = x * inv_pio2 + to_int;
    fn = fn - to_int;
    int n = (int)fn;
    double r = x - fn * pio2_1;
    double w = fn * pio2_1t;
    double y = r - w;
    int ey = (int)((bits_of(y) >> 52) & 0x7ff);
    int ex = (int)(ix >> 20);
    if (ex - ey > 16) {
        double t = r;
        w = fn * pio2_2;
        r = t - w;
        w = fn * pio2_2t - ((t - r) - w);
        y = r - w;
        ey = (int)((bits_of(y) >> 52) & 0x7ff);
        if (ex - ey > 49) {
            t = r;
            w = fn * pio2_3;
            r = t - w;
            w = fn * pio2_3t - ((t - r) - w);
            y = r - w;
        }
    }
    *y0 = y;
    *y1 = (r - y) - w;
    return n;
}

static int rem_pio2(double x, double *y0, double *y1) {
    const double pio2_1 = 1.57079632673412561417e+00, pio2_1t = 6.07710050650619224932e-11;
    uint64_t u = bits_of(x);
    int sign = (int)(u >> 63);
    uint32_t ix = (uint32_t)(u >> 32) & 0x7fffffffu;
    if (ix <= 0x400f6a7au) {
        if ((ix & 0xfffffu) == 0x921fbu) return rem_pio2_medium(x, ix, y0, y1);
        if (ix <= 0x4002d97cu) {
            if (!sign) {
                double z = x - pio2_1;
                *y0 = z - pio2_1t; *y1 = (z - *y0) - pio2_1t; return 1;
            } else {
                double z = x + pio2_1;
                *y0 = z + pio2_1t; *y1 = (z - *y0) + pio2_1t; return -1;
            }
        }
        if (!sign) {
            double z = x - 2.0 * pio2_1;
            *y0 = z - 2.0 * pio2_1t; *y1 = (z - *y0) - 2.0 * pio2_1t; return 2;
        } else {
            double z = x + 2.0 * pio2_1;
            *y0 = z + 2.0 * pio2_1t; *y1 = (z - *y0) + 2.0 * pio2_1t; return -2;
        }
    }
    if (ix <= 0x401c463bu) {
        if (ix <= 0x4015fdbcu) {
            if (ix == 0x4012d97cu) return rem_pio2_medium(x, ix, y0, y1);
            if (!sign) {
                double z = x - 3.0 * pio2_1;
                *y0 = z - 3.0 * pio2_1t; *y1 = (z - *y0) - 3.0 * pio2_1t; return 3;
            } else {
                double z = x + 3.0 * pio2_1;
                *y0 = z + 3.0 * pio2_1t; *y1 = (z - *y0) + 3.0 * pio2_1t; return -3;
            }
        }
        if (ix == 0x401921fbu) return rem_pio2_medium(x, ix, y0, y1);
        if (!sign) {
            double z = x - 4.0 * pio2_1;
            *y0 = z - 4.0 * pio2_1t; *y1 = (z - *y0) - 4.0 * pio2_1t; return 4;
        } else {
            double z = x + 4.0 * pio2_1;
            *y0 = z + 4.0 * pio2_1t; *y1 = (z - *y0) + 4.0 * pio2_1t; return -4;
        }
    }
    if (ix < 0x413921fbu) return rem_pio2_medium(x, ix, y0, y1);
    fprintf(stderr, "oracle: sincos argument %g beyond 2^20 pi/2 is not supported\n", x);
    abort();
}

void or_sincos(double x, double *s, double *c) {
    uint32_t ix = (uint32_t)(bits_of(x) >> 32) & 0x7fffffffu;
    if (ix <= 0x3fe921fbu) {
        if (ix < 0x3e46a09eu) { *s = x; *c = 1.0; return; }
        *s = k_sin(x, 0.0, 0);
        *c = k_cos(x, 0.0);
        return;
    }
    if (ix >= 0x7ff00000u) { *s = *c = x - x; return; }
    double y0, y1;
    int n = rem_pio2(x, &y0, &y1);
    double sv = k_sin(y0, y1, 1), cv = k_cos(y0, y1);
    switch (n & 3) {
    case 0: *s = sv; *c = cv; break;
    case 1: *s = cv; *c = -sv; break;
    case 2: *s = -sv; *c = -cv; break;
    default: *s = -cv; *c = sv; break;
    }
}

double or_sin(double x) { double s, c; or_sincos(x, &s, &c); return s; }
double or_cos(double x) { double s, c; or_sincos(x, &s, &c); return c; }

double or_powi(double x, int n) {
    double base = x, result = 1.0;
    unsigned e = n < 0 ? (unsigned)(-(long)n) : (unsigned)n;
    for (;;) {
        if (e & 1u) result *= base;
        e >>= 1;
        if (e == 0) break;
        base *= base;
    }
    return n < 0 ? 1.0 / result : result;
}

/* e^(a + ib) for finite arguments, as the reference's complex exp: a real z gives exp(a) with
 * the zero imaginary part kept. */
or_c64 or_cexp(or_c64 z) {
    double r = or_exp(z.re);
    if (z.im == 0.0) return (or_c64){r, z.im};
    double s, c;
    or_sincos(z.im, &s, &c);
    return (or_c64){r * c, r * s};
}
