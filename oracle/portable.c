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

/* The natural logarithm as the reference computes it (src/portable/mod.rs:141-220, its port of
 * fdlibm's e_log.c, Sun Microsystems 1993): x = 2^k (1 + f) with 1 + f in [sqrt(2)/2, sqrt(2)),
 * s = f / (2 + f), and log(1 + f) = f - s (f - R(s^2)) with fdlibm's degree-14 polynomial R and
 * the 0.5 f^2 correction above 0.7, ln 2 split into a 33-bit head and its tail. */
double or_ln(double x) {
    const double ln2_hi = of_bits(0x3fe62e42fee00000ULL), ln2_lo = of_bits(0x3dea39ef35793c76ULL);
    const double two54 = of_bits(0x4350000000000000ULL);
    const double lg1 = of_bits(0x3fe5555555555593ULL), lg2 = of_bits(0x3fd999999997fa04ULL);
    const double lg3 = of_bits(0x3fd2492494229359ULL), lg4 = of_bits(0x3fcc71c51d8e78afULL);
    const double lg5 = of_bits(0x3fc7466496cb03deULL), lg6 = of_bits(0x3fc39a09d078c69fULL);
    const double lg7 = of_bits(0x3fc2f112df3e5244ULL);
    int k = 0;
    uint64_t u = bits_of(x);
    if (u - 0x0010000000000000ULL >= 0x7ff0000000000000ULL - 0x0010000000000000ULL) {
        if (isnan(x) || x < 0.0) return NAN;
        if (x == 0.0) return -INFINITY;
        if (isinf(x)) return x;
        k -= 54; /* subnormal: scale up */
        x *= two54;
        u = bits_of(x);
    }
    int32_t high = (int32_t)(u >> 32);
    k += (high >> 20) - 1023;
    high &= 0x000fffff;
    const int32_t i = (high + 0x95f64) & 0x100000;
    /* x or x / 2 into [sqrt(2) / 2, sqrt(2)) */
    x = of_bits(((uint64_t)(uint32_t)(high | (i ^ 0x3ff00000)) << 32) | (u & 0xffffffffULL));
    k += i >> 20;
    const double f = x - 1.0, dk = (double)k;
    if ((0x000fffff & (2 + high)) < 3) { /* |f| < 2^-20 */
        if (f == 0.0) return dk * ln2_hi + dk * ln2_lo;
        const double r = f * f * (0.5 - (1.0 / 3.0) * f);
        return k == 0 ? f - r : dk * ln2_hi - ((r - dk * ln2_lo) - f);
    }
    const double s = f / (2.0 + f), z = s * s, w = z * z;
    const double t1 = w * (lg2 + w * (lg4 + w * lg6));
    const double t2 = z * (lg1 + w * (lg3 + w * (lg5 + w * lg7)));
    const double r = t2 + t1;
    if (((high - 0x6147a) | (0x6b851 - high)) > 0) {
        const double hfsq = 0.5 * f * f;
        return k == 0 ? f - (hfsq - s * (hfsq + r)) : dk * ln2_hi - ((hfsq - (s * (hfsq + r) + dk * ln2_lo)) - f);
    }
    return k == 0 ? f - s * (f - r) : dk * ln2_hi - ((s * (f - r) - dk * ln2_lo) - f);
}

/* x 2^n with one rounding (musl's scalbn): the factor applied in at most three exact steps. */
static double scale2(double x, int n) {
    double y = x;
    if (n > 1023) {
        y *= of_bits(0x7fe0000000000000ULL);
        n -= 1023;
        if (n > 1023) {
            y *= of_bits(0x7fe0000000000000ULL);
            n -= 1023;
            if (n > 1023) n = 1023;
        }
    } else if (n < -1022) {
        /* 2^-1022 2^53: keeps the significand exact until the last multiplication */
        y *= of_bits(0x0360000000000000ULL);
        n += 1022 - 53;
        if (n < -1022) {
            y *= of_bits(0x0360000000000000ULL);
            n += 1022 - 53;
            if (n < -1022) n = -1022;
        }
    }
    return y * of_bits((uint64_t)(0x3ff + n) << 52);
}

static inline uint32_t hi_word(double x) { return (uint32_t)(bits_of(x) >> 32); }
static inline uint32_t lo_word(double x) { return (uint32_t)bits_of(x); }
static inline double with_lo(double x, uint32_t lo) { return of_bits((bits_of(x) & 0xffffffff00000000ULL) | lo); }
static inline double with_hi(double x, uint32_t hi) { return of_bits(((uint64_t)hi << 32) | (bits_of(x) & 0xffffffffULL)); }

/* base^exponent as the reference computes it (src/portable/mod.rs:299-301 calls the `libm` crate
 * 0.2.16's pow, its port of fdlibm's e_pow.c, Sun Microsystems 2004): C's special cases, then
 * log2|x| = n + dp_h + z_h + z_l from (x - bp)/(x + bp) and a degree-12 polynomial, carried in
 * head + tail pairs whose heads have their low 32 bits cleared, times y split the same way, and
 * 2^(p_h + p_l) from fdlibm's exp kernel, scaled through the exponent bits. */
double or_pow(double x, double y) {
    static const double bp[] = {1.0, 1.5};
    static const double dp_h[] = {0.0, 5.84962487220764160156e-01}; /* 0x3fe2b803 40000000 */
    static const double dp_l[] = {0.0, 1.35003920212974897128e-08}; /* 0x3e4cfdeb 43cfd006 */
    const double two53 = 9007199254740992.0, huge = 1.0e300, tiny = 1.0e-300;
    /* (3/2)(log(x) - 2s - 2/3 s^3) */
    const double l1 = 5.99999999999994648725e-01, l2 = 4.28571428578550184252e-01;
    const double l3 = 3.33333329818377432918e-01, l4 = 2.72728123808534006489e-01;
    const double l5 = 2.30660745775561754067e-01, l6 = 2.06975017800338417784e-01;
    const double p1 = 1.66666666666666019037e-01, p2 = -2.77777777770155933842e-03;
    const double p3 = 6.61375632143793436117e-05, p4 = -1.65339022054652515390e-06;
    const double p5 = 4.13813679705723846039e-08;
    const double lg2 = 6.93147180559945286227e-01, lg2_h = 6.93147182464599609375e-01;
    const double lg2_l = -1.90465429995776804525e-09;
    const double ovt = 8.0085662595372944372e-17; /* -(1024 - log2(ovfl + .5ulp)) */
    const double cp = 9.61796693925975554329e-01, cp_h = 9.61796700954437255859e-01;
    const double cp_l = -7.02846165095275826516e-09;
    const double ivln2 = 1.44269504088896338700e+00, ivln2_h = 1.44269502162933349609e+00;
    const double ivln2_l = 1.92596299112661746887e-08;

    const int32_t hx = (int32_t)hi_word(x), hy = (int32_t)hi_word(y);
    const uint32_t lx = lo_word(x), ly = lo_word(y);
    int32_t ix = hx & 0x7fffffff;
    const int32_t iy = hy & 0x7fffffff;

    if ((iy | (int32_t)ly) == 0) return 1.0;           /* x^0 = 1 */
    if (hx == 0x3ff00000 && lx == 0) return 1.0;        /* 1^y = 1, even for a NaN y */
    if (ix > 0x7ff00000 || (ix == 0x7ff00000 && lx != 0) || iy > 0x7ff00000 || (iy == 0x7ff00000 && ly != 0))
        return x + y;                                   /* NaN */

    /* for x < 0: 0 if y is not an integer, 1 if it is odd, 2 if even */
    int yisint = 0;
    if (hx < 0) {
        if (iy >= 0x43400000) yisint = 2;
        else if (iy >= 0x3ff00000) {
            const int k = (iy >> 20) - 0x3ff;
            if (k > 20) {
                const uint32_t j = ly >> (52 - k);
                if ((j << (52 - k)) == ly) yisint = 2 - (int)(j & 1);
            } else if (ly == 0) {
                const int32_t j = iy >> (20 - k);
                if ((j << (20 - k)) == iy) yisint = 2 - (j & 1);
            }
        }
    }

    if (ly == 0) {
        if (iy == 0x7ff00000) { /* y is +-inf */
            if (((ix - 0x3ff00000) | (int32_t)lx) == 0) return 1.0;   /* (-1)^+-inf */
            if (ix >= 0x3ff00000) return hy >= 0 ? y : 0.0;           /* |x| > 1 */
            return hy >= 0 ? 0.0 : -y;                                /* |x| < 1 */
        }
        if (iy == 0x3ff00000) return hy >= 0 ? x : 1.0 / x; /* y is +-1 */
        if (hy == 0x40000000) return x * x;                 /* y is 2 */
        if (hy == 0x3fe00000 && hx >= 0) return sqrt(x);    /* y is 0.5, x >= +0 */
    }

    double ax = fabs(x);
    if (lx == 0 && (ix == 0x7ff00000 || ix == 0 || ix == 0x3ff00000)) { /* x is +-0, +-inf, +-1 */
        double z = ax;
        if (hy < 0) z = 1.0 / z;
        if (hx < 0) {
            if (((ix - 0x3ff00000) | yisint) == 0) z = (z - z) / (z - z); /* (-1)^non-int */
            else if (yisint == 1) z = -z;
        }
        return z;
    }

    double s = 1.0; /* the result's sign */
    if (hx < 0) {
        if (yisint == 0) return (x - x) / (x - x); /* (x < 0)^non-int is NaN */
        if (yisint == 1) s = -1.0;
    }

    double t1, t2;
    if (iy > 0x41e00000) { /* |y| > 2^31 */
        if (iy > 0x43f00000) { /* |y| > 2^64: over- or underflows */
            if (ix <= 0x3fefffff) return hy < 0 ? huge * huge : tiny * tiny;
            if (ix >= 0x3ff00000) return hy > 0 ? huge * huge : tiny * tiny;
        }
        if (ix < 0x3fefffff) return hy < 0 ? s * huge * huge : s * tiny * tiny;
        if (ix > 0x3ff00000) return hy > 0 ? s * huge * huge : s * tiny * tiny;
        /* |1 - x| <= 2^-20: log(x) by x - x^2/2 + x^3/3 - x^4/4 */
        const double t = ax - 1.0;
        const double w = (t * t) * (0.5 - t * (0.3333333333333333333333 - t * 0.25));
        const double u = ivln2_h * t;
        const double v = t * ivln2_l - w * ivln2;
        t1 = with_lo(u + v, 0);
        t2 = v - (t1 - u);
    } else {
        int n = 0;
        if (ix < 0x00100000) { /* subnormal x */
            ax *= two53;
            n -= 53;
            ix = (int32_t)hi_word(ax);
        }
        n += (ix >> 20) - 0x3ff;
        const int32_t j = ix & 0x000fffff;
        int k;
        ix = j | 0x3ff00000; /* normalize */
        if (j <= 0x3988E) k = 0;       /* |x| < sqrt(3/2) */
        else if (j < 0xBB67A) k = 1;   /* |x| < sqrt(3) */
        else { k = 0; n += 1; ix -= 0x00100000; }
        ax = with_hi(ax, (uint32_t)ix);

        /* ss = s_h + s_l = (x - 1)/(x + 1) or (x - 1.5)/(x + 1.5) */
        double u = ax - bp[k];
        double v = 1.0 / (ax + bp[k]);
        const double ss = u * v;
        const double s_h = with_lo(ss, 0);
        double t_h = of_bits((uint64_t)(uint32_t)(((ix >> 1) | 0x20000000) + 0x00080000 + (k << 18)) << 32);
        double t_l = ax - (t_h - bp[k]);
        const double s_l = v * ((u - s_h * t_h) - s_h * t_l);
        /* log(ax) */
        double s2 = ss * ss;
        double r = s2 * s2 * (l1 + s2 * (l2 + s2 * (l3 + s2 * (l4 + s2 * (l5 + s2 * l6)))));
        r += s_l * (s_h + ss);
        s2 = s_h * s_h;
        t_h = with_lo(3.0 + s2 + r, 0);
        t_l = r - ((t_h - 3.0) - s2);
        /* u + v = ss (1 + ...) */
        u = s_h * t_h;
        v = s_l * t_h + t_l * ss;
        /* 2/(3 log2) (ss + ...) */
        const double p_h = with_lo(u + v, 0);
        const double p_l = v - (p_h - u);
        const double z_h = cp_h * p_h;
        const double z_l = cp_l * p_h + p_l * cp + dp_l[k];
        /* log2(ax) = (ss + ..) 2/(3 log2) = n + dp_h + z_h + z_l */
        const double t = (double)n;
        t1 = with_lo(((z_h + z_l) + dp_h[k]) + t, 0);
        t2 = z_l - (((t1 - t) - dp_h[k]) - z_h);
    }

    /* (y1 + y2)(t1 + t2) with y1 y's head */
    const double y1 = with_lo(y, 0);
    const double pl = (y - y1) * t1 + y * t2;
    double ph = y1 * t1;
    double z = pl + ph;
    int32_t j = (int32_t)hi_word(z);
    const int32_t i = (int32_t)lo_word(z);
    if (j >= 0x40900000) { /* z >= 1024 */
        if (((j - 0x40900000) | i) != 0) return s * huge * huge;
        if (pl + ovt > z - ph) return s * huge * huge;
    } else if ((j & 0x7fffffff) >= 0x4090cc00) { /* z <= -1075 */
        if (((j - (int32_t)0xc090cc00) | i) != 0) return s * tiny * tiny;
        if (pl <= z - ph) return s * tiny * tiny;
    }

    /* 2^(ph + pl) */
    const int32_t ia = j & 0x7fffffff;
    int k = (ia >> 20) - 0x3ff;
    int32_t n = 0;
    if (ia > 0x3fe00000) { /* |z| > 0.5: n = [z + 0.5] */
        n = j + (0x00100000 >> (k + 1));
        k = ((n & 0x7fffffff) >> 20) - 0x3ff;
        const double t = of_bits((uint64_t)(uint32_t)(n & ~(0x000fffff >> k)) << 32);
        n = ((n & 0x000fffff) | 0x00100000) >> (20 - k);
        if (j < 0) n = -n;
        ph -= t;
    }
    const double t = with_lo(pl + ph, 0);
    const double u = t * lg2_h;
    const double v = (pl - (t - ph)) * lg2 + t * lg2_l;
    z = u + v;
    const double w = v - (z - u);
    const double tt = z * z;
    const double tt1 = z - tt * (p1 + tt * (p2 + tt * (p3 + tt * (p4 + tt * p5))));
    const double r = (z * tt1) / (tt1 - 2.0) - (w + z * w);
    z = 1.0 - (r - z);
    j = (int32_t)hi_word(z);
    j += (int32_t)((uint32_t)n << 20);
    if ((j >> 20) <= 0) z = scale2(z, n); /* subnormal result */
    else z = with_hi(z, (uint32_t)j);
    return s * z;
}
