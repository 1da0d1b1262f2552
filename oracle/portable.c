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
 *     __rem_pio2 in three rounds, and above 2^20 pi/2 fdlibm's k_rem_pio2 (Payne-Hanek with
 *     2/pi in 24-bit chunks);
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

/* The reduction of a large argument (fdlibm's k_rem_pio2.c, Sun Microsystems 1993, as musl's
 * __rem_pio2_large and Rust's libm rem_pio2_large call it with prec 1): x = sum tx[i] 2^(e0 -
 * 24 i) with 24-bit integer tx, times 2/pi in 24-bit chunks, the integer part mod 8 and the
 * fraction as y0 + y1 (times pi/2 in 24-bit pieces).
 * ipio2: 2/pi = sum ipio2[i] 2^(-24 (i + 1)); pio2_pieces: pi/2 = sum of 24-bit pieces at
 * 2^0, 2^-24, ... Both are the exact digits of the constants (written out with mpmath). */
static const int32_t ipio2[66] = {
    0xA2F983, 0x6E4E44, 0x1529FC, 0x2757D1, 0xF534DD, 0xC0DB62, 0x95993C, 0x439041, 0xFE5163,
    0xABDEBB, 0xC561B7, 0x246E3A, 0x424DD2, 0xE00649, 0x2EEA09, 0xD1921C, 0xFE1DEB, 0x1CB129,
    0xA73EE8, 0x8235F5, 0x2EBB44, 0x84E99C, 0x7026B4, 0x5F7E41, 0x3991D6, 0x398353, 0x39F49C,
    0x845F8B, 0xBDF928, 0x3B1FF8, 0x97FFDE, 0x05980F, 0xEF2F11, 0x8B5A0A, 0x6D1F6D, 0x367ECF,
    0x27CB09, 0xB74F46, 0x3F669E, 0x5FEA2D, 0x7527BA, 0xC7EBE5, 0xF17B3D, 0x0739F7, 0x8A5292,
    0xEA6BFB, 0x5FB11F, 0x8D5D08, 0x560330, 0x46FC7B, 0x6BABF0, 0xCFBC20, 0x9AF436, 0x1DA9E3,
    0x91615E, 0xE61B08, 0x659985, 0x5F14A0, 0x68408D, 0xFFD880, 0x4D7327, 0x310606, 0x1556CA,
    0x73A8C9, 0x60E27B, 0xC08C6B};
static const double pio2_pieces[8] = {
    1.57079625129699707031e+00, 7.54978941586159635335e-08, 5.39030252995776476554e-15,
    3.28200341580791294123e-22, 1.27065575308067607349e-29, 1.22933308981111328932e-36,
    2.73370053816464559624e-44, 2.16741683877804819444e-51};

static int rem_pio2_large(const double *x, double *y, int e0, int nx) {
    const double two24 = 16777216.0, twon24 = 5.9604644775390625e-08;
    const int jk = 4, jp = 4; /* prec 1: init_jk[1] */
    int32_t iq[20];
    double f[20], fq[20], q[20];
    int jx = nx - 1, jv = (e0 - 3) / 24, i, j, k, m, n, jz, carry, ih;
    double z, fw;
    if (jv < 0) jv = 0;
    int q0 = e0 - 24 * (jv + 1);
    j = jv - jx;
    m = jx + jk;
    for (i = 0; i <= m; i++, j++) f[i] = j < 0 ? 0.0 : (double)ipio2[j];
    for (i = 0; i <= jk; i++) {
        for (j = 0, fw = 0.0; j <= jx; j++) fw += x[j] * f[jx + i - j];
        q[i] = fw;
    }
    jz = jk;
    for (;;) {
        /* distill q[] into iq[] reversingly */
        for (i = 0, j = jz, z = q[jz]; j > 0; i++, j--) {
            fw = (double)(int32_t)(twon24 * z);
            iq[i] = (int32_t)(z - two24 * fw);
            z = q[j - 1] + fw;
        }
        z = scalbn(z, q0);
        z -= 8.0 * floor(z * 0.125);
        n = (int32_t)z;
        z -= (double)n;
        ih = 0;
        if (q0 > 0) {
            i = iq[jz - 1] >> (24 - q0);
            n += i;
            iq[jz - 1] -= i << (24 - q0);
            ih = iq[jz - 1] >> (23 - q0);
        } else if (q0 == 0) {
            ih = iq[jz - 1] >> 23;
        } else if (z >= 0.5) {
            ih = 2;
        }
        if (ih > 0) { /* q > 0.5 */
            n += 1;
            carry = 0;
            for (i = 0; i < jz; i++) {
                j = iq[i];
                if (carry == 0) {
                    if (j != 0) {
                        carry = 1;
                        iq[i] = 0x1000000 - j;
                    }
                } else {
                    iq[i] = 0xffffff - j;
                }
            }
            if (q0 == 1) iq[jz - 1] &= 0x7fffff;
            else if (q0 == 2) iq[jz - 1] &= 0x3fffff;
            if (ih == 2) {
                z = 1.0 - z;
                if (carry != 0) z -= scalbn(1.0, q0);
            }
        }
        /* a zero fraction: more terms of 2/pi, unless it is exact */
        if (z == 0.0) {
            j = 0;
            for (i = jz - 1; i >= jk; i--) j |= iq[i];
            if (j == 0) {
                for (k = 1; iq[jk - k] == 0; k++) {}
                for (i = jz + 1; i <= jz + k; i++) {
                    f[jx + i] = (double)ipio2[jv + i];
                    for (j = 0, fw = 0.0; j <= jx; j++) fw += x[j] * f[jx + i - j];
                    q[i] = fw;
                }
                jz += k;
                continue;
            }
        }
        break;
    }
    if (z == 0.0) {
        jz -= 1;
        q0 -= 24;
        while (iq[jz] == 0) {
            jz--;
            q0 -= 24;
        }
    } else {
        z = scalbn(z, -q0);
        if (z >= two24) {
            fw = (double)(int32_t)(twon24 * z);
            iq[jz] = (int32_t)(z - two24 * fw);
            jz += 1;
            q0 += 24;
            iq[jz] = (int32_t)fw;
        } else {
            iq[jz] = (int32_t)z;
        }
    }
    fw = scalbn(1.0, q0);
    for (i = jz; i >= 0; i--) {
        q[i] = fw * (double)iq[i];
        fw *= twon24;
    }
    for (i = jz; i >= 0; i--) {
        for (fw = 0.0, k = 0; k <= jp && k <= jz - i; k++) fw += pio2_pieces[k] * q[i + k];
        fq[jz - i] = fw;
    }
    fw = 0.0;
    for (i = jz; i >= 0; i--) fw += fq[i];
    y[0] = ih == 0 ? fw : -fw;
    fw = fq[0] - fw;
    for (i = 1; i <= jz; i++) fw += fq[i];
    y[1] = ih == 0 ? fw : -fw;
    return n & 7;
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
    /* |x| >= 2^20 pi/2: x's significand as three 24-bit pieces of an integer scaled by 2^e0 */
    double tx[3], ty[2];
    uint64_t v = (u & (((uint64_t)1 << 52) - 1)) | ((uint64_t)(0x3ff + 23) << 52);
    double z = of_bits(v);
    int i;
    for (i = 0; i < 2; i++) {
        tx[i] = (double)(int32_t)z;
        z = (z - tx[i]) * 16777216.0;
    }
    tx[i] = z;
    while (tx[i] == 0.0) i--;
    int n = rem_pio2_large(tx, ty, (int)(ix >> 20) - (0x3ff + 23), i + 1);
    if (sign) {
        *y0 = -ty[0];
        *y1 = -ty[1];
        return -n;
    }
    *y0 = ty[0];
    *y1 = ty[1];
    return n;
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
