// Bit-portable maths of the host layer (see portable.hpp).
#include "portable.hpp"

#include <cmath>
#include <cstring>
#include <string>

#include "../../include/pxfdtd.h"

int px_fail(int code, const char *fmt, ...);  // resident.hip

namespace pxfdtd {
namespace {

inline uint64_t bits(double x) {
    uint64_t u;
    std::memcpy(&u, &x, 8);
    return u;
}
inline double from_bits(uint64_t u) {
    double x;
    std::memcpy(&x, &u, 8);
    return x;
}

// fdlibm's constants, given by their bits
struct ExpConst {
    double o_thr = from_bits(0x40862e42fefa39efULL), u_thr = from_bits(0xc0874910d52d3051ULL);
    double ln2hi = from_bits(0x3fe62e42fee00000ULL), ln2lo = from_bits(0x3dea39ef35793c76ULL);
    double invln2 = from_bits(0x3ff71547652b82feULL), twom1000 = from_bits(0x0170000000000000ULL);
    double P[5] = {from_bits(0x3fc555555555553eULL), from_bits(0xbf66c16c16bebd93ULL),
                   from_bits(0x3f11566aaf25de2cULL), from_bits(0xbebbbd41c5d26bf1ULL),
                   from_bits(0x3e66376972bea4d0ULL)};
};
const ExpConst EC;

// minimax kernels of sin and cos on [-pi/4, pi/4] (fdlibm k_sin.c, k_cos.c)
constexpr double S[6] = {-1.66666666666666324348e-01, 8.33333333332248946124e-03,
                         -1.98412698298579493134e-04, 2.75573137070700676789e-06,
                         -2.50507602534068634195e-08, 1.58969099521155010221e-10};
constexpr double C[6] = {4.16666666666666019037e-02, -1.38888888888741095749e-03,
                         2.48015872894767294178e-05, -2.75573143513906633035e-07,
                         2.08757232129817482790e-09, -1.13596475577881948265e-11};

double ksin(double x, double y, bool tail) {
    const double z = x * x, w = z * z;
    const double r = S[1] + z * (S[2] + z * S[3]) + z * w * (S[4] + z * S[5]);
    const double v = z * x;
    if (!tail) return x + v * (S[0] + z * r);
    return x - ((z * (0.5 * y - v * r) - y) - v * S[0]);
}

double kcos(double x, double y) {
    const double z = x * x, w = z * z;
    const double r = z * (C[0] + z * (C[1] + z * C[2])) + w * w * (C[3] + z * (C[4] + z * C[5]));
    const double hz = 0.5 * z;
    const double one_minus = 1.0 - hz;
    return one_minus + (((1.0 - one_minus) - hz) + (z * r - x * y));
}

// pi/2 in three parts for Cody and Waite's reduction (fdlibm e_rem_pio2.c)
constexpr double PIO2[3] = {1.57079632673412561417e+00, 6.07710050630396597660e-11,
                            2.02226624871116645580e-21};
constexpr double PIO2T[3] = {6.07710050650619224932e-11, 2.02226624879595063154e-21,
                             8.47842766036889956997e-32};

int reduce_medium(double x, uint32_t hx, double &y0, double &y1) {
    const double to_int = 6755399441055744.0;  // 1.5 * 2^52
    const double inv_pio2 = 6.36619772367581382433e-01;
    double fn = x * inv_pio2 + to_int;
    fn = fn - to_int;
    const int n = (int)fn;
    double r = x - fn * PIO2[0];
    double w = fn * PIO2T[0];
    y0 = r - w;
    const int ex = (int)(hx >> 20);
    for (int round = 1; round < 3; round++) {
        const int ey = (int)((bits(y0) >> 52) & 0x7ff);
        if (ex - ey <= (round == 1 ? 16 : 49)) break;
        const double t = r;
        w = fn * PIO2[round];
        r = t - w;
        w = fn * PIO2T[round] - ((t - r) - w);
        y0 = r - w;
    }
    y1 = (r - y0) - w;
    return n;
}

// x - m pi/2 for |x| up to about 9 pi/4 with the nearest multiple m known
int reduce_small(double x, int m, double &y0, double &y1) {
    const double a = m * PIO2[0], at = m * PIO2T[0];
    const double z = x - a;
    y0 = z - at;
    y1 = (z - y0) - at;
    return m;
}

// x 2^n with one rounding (musl's scalbn): the factor applied in at most three exact steps. */
double scale2(double x, int n) {
    double y = x;
    if (n > 1023) {
        y *= from_bits(0x7fe0000000000000ULL);
        n -= 1023;
        if (n > 1023) {
            y *= from_bits(0x7fe0000000000000ULL);
            n -= 1023;
            if (n > 1023) n = 1023;
        }
    } else if (n < -1022) {
        /* 2^-1022 2^53: keeps the significand exact until the last multiplication */
        y *= from_bits(0x0360000000000000ULL);
        n += 1022 - 53;
        if (n < -1022) {
            y *= from_bits(0x0360000000000000ULL);
            n += 1022 - 53;
            if (n < -1022) n = -1022;
        }
    }
    return y * from_bits((uint64_t)(0x3ff + n) << 52);
}

inline uint32_t hi_word(double x) { return (uint32_t)(bits(x) >> 32); }
inline uint32_t lo_word(double x) { return (uint32_t)bits(x); }
inline double with_lo(double x, uint32_t lo) { return from_bits((bits(x) & 0xffffffff00000000ULL) | lo); }
inline double with_hi(double x, uint32_t hi) { return from_bits(((uint64_t)hi << 32) | (bits(x) & 0xffffffffULL)); }


}  // namespace

double pexp(double x) {
    uint32_t hx = (uint32_t)(bits(x) >> 32);
    const bool neg = (hx >> 31) != 0;
    hx &= 0x7fffffffu;
    if (hx >= 0x40862e42u) {
        if (std::isnan(x)) return x;
        if (x > EC.o_thr) return INFINITY;
        if (x < EC.u_thr) return 0.0;
    }
    int k = 0;
    double hi = 0.0, lo = 0.0;
    if (hx > 0x3fd62e42u) {
        if (hx < 0x3ff0a2b2u) {
            k = neg ? -1 : 1;
            hi = neg ? x + EC.ln2hi : x - EC.ln2hi;
            lo = neg ? -EC.ln2lo : EC.ln2lo;
        } else {
            k = (int)(EC.invln2 * x + (neg ? -0.5 : 0.5));
            const double t = (double)k;
            hi = x - t * EC.ln2hi;
            lo = t * EC.ln2lo;
        }
        x = hi - lo;
    } else if (hx < 0x3e300000u) {
        return 1.0 + x;
    }
    const double t = x * x;
    const double c = x - t * (EC.P[0] + t * (EC.P[1] + t * (EC.P[2] + t * (EC.P[3] + t * EC.P[4]))));
    if (k == 0) return 1.0 - ((x * c) / (c - 2.0) - x);
    const double y = 1.0 - ((lo - (x * c) / (2.0 - c)) - hi);
    if (k >= -1021) return from_bits(bits(y) + ((uint64_t)(int64_t)k << 52));
    return from_bits(bits(y) + ((uint64_t)(int64_t)(k + 1000) << 52)) * EC.twom1000;
}

bool psincos(double x, double &s, double &c) {
    const uint64_t u = bits(x);
    const uint32_t hx = (uint32_t)(u >> 32) & 0x7fffffffu;
    const bool neg = (u >> 63) != 0;
    if (hx <= 0x3fe921fbu) {
        if (hx < 0x3e46a09eu) {
            s = x;
            c = 1.0;
            return true;
        }
        s = ksin(x, 0.0, false);
        c = kcos(x, 0.0);
        return true;
    }
    if (hx >= 0x7ff00000u) {
        s = c = x - x;
        return true;
    }
    double y0, y1;
    int n;
    const int sg = neg ? -1 : 1;
    if (hx <= 0x400f6a7au && (hx & 0xfffffu) != 0x921fbu)
        n = reduce_small(x, sg * (hx <= 0x4002d97cu ? 1 : 2), y0, y1);
    else if (hx > 0x400f6a7au && hx <= 0x401c463bu && hx != 0x4012d97cu && hx != 0x401921fbu)
        n = reduce_small(x, sg * (hx <= 0x4015fdbcu ? 3 : 4), y0, y1);
    else if (hx < 0x413921fbu)
        n = reduce_medium(x, hx, y0, y1);
    else
        return false;
    const double sv = ksin(y0, y1, true), cv = kcos(y0, y1);
    switch (n & 3) {
    case 0: s = sv; c = cv; break;
    case 1: s = cv; c = -sv; break;
    case 2: s = -sv; c = -cv; break;
    default: s = -cv; c = sv; break;
    }
    return true;
}

double ppowi(double x, int n) {
    double base = x, result = 1.0;
    unsigned e = n < 0 ? 0u - (unsigned)n : (unsigned)n;
    while (true) {
        if (e & 1u) result *= base;
        e >>= 1;
        if (!e) break;
        base *= base;
    }
    return n < 0 ? 1.0 / result : result;
}

C64 pcexp(C64 z) {
    const double r = pexp(z.re);
    if (z.im == 0.0) return {r, z.im};
    double s, c;
    if (!psincos(z.im, s, c)) return {NAN, NAN};
    return {r * c, r * s};
}

/* The natural logarithm as the reference computes it (src/portable/mod.rs:141-220, its port of
 * fdlibm's e_log.c, Sun Microsystems 1993): x = 2^k (1 + f) with 1 + f in [sqrt(2)/2, sqrt(2)),
 * s = f / (2 + f), and log(1 + f) = f - s (f - R(s^2)) with fdlibm's degree-14 polynomial R and
 * the 0.5 f^2 correction above 0.7, ln 2 split into a 33-bit head and its tail. */
double pln(double x) {
    const double ln2_hi = from_bits(0x3fe62e42fee00000ULL), ln2_lo = from_bits(0x3dea39ef35793c76ULL);
    const double two54 = from_bits(0x4350000000000000ULL);
    const double lg1 = from_bits(0x3fe5555555555593ULL), lg2 = from_bits(0x3fd999999997fa04ULL);
    const double lg3 = from_bits(0x3fd2492494229359ULL), lg4 = from_bits(0x3fcc71c51d8e78afULL);
    const double lg5 = from_bits(0x3fc7466496cb03deULL), lg6 = from_bits(0x3fc39a09d078c69fULL);
    const double lg7 = from_bits(0x3fc2f112df3e5244ULL);
    int k = 0;
    uint64_t u = bits(x);
    if (u - 0x0010000000000000ULL >= 0x7ff0000000000000ULL - 0x0010000000000000ULL) {
        if (std::isnan(x) || x < 0.0) return std::nan("");
        if (x == 0.0) return -HUGE_VAL;
        if (std::isinf(x)) return x;
        k -= 54; /* subnormal: scale up */
        x *= two54;
        u = bits(x);
    }
    int32_t high = (int32_t)(u >> 32);
    k += (high >> 20) - 1023;
    high &= 0x000fffff;
    const int32_t i = (high + 0x95f64) & 0x100000;
    /* x or x / 2 into [sqrt(2) / 2, sqrt(2)) */
    x = from_bits(((uint64_t)(uint32_t)(high | (i ^ 0x3ff00000)) << 32) | (u & 0xffffffffULL));
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

/* base^exponent as the reference computes it (src/portable/mod.rs:299-301 calls the `libm` crate
 * 0.2.16's pow, its port of fdlibm's e_pow.c, Sun Microsystems 2004): C's special cases, then
 * log2|x| = n + dp_h + z_h + z_l from (x - bp)/(x + bp) and a degree-12 polynomial, carried in
 * head + tail pairs whose heads have their low 32 bits cleared, times y split the same way, and
 * 2^(p_h + p_l) from fdlibm's exp kernel, scaled through the exponent bits. */
double ppow(double x, double y) {
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
        if (hy == 0x3fe00000 && hx >= 0) return std::sqrt(x);    /* y is 0.5, x >= +0 */
    }

    double ax = std::fabs(x);
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
        double t_h = from_bits((uint64_t)(uint32_t)(((ix >> 1) | 0x20000000) + 0x00080000 + (k << 18)) << 32);
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
        const double t = from_bits((uint64_t)(uint32_t)(n & ~(0x000fffff >> k)) << 32);
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

/* sqrt(x^2 + y^2) with no overflow or underflow in between, as the reference computes it
 * (src/portable/mod.rs:246-248 takes the `libm` crate 0.2.16's hypot, its port of musl's
 * hypot.c): |x| >= |y| ordered by their bits, inf before NaN, the larger alone where the
 * exponents are more than 64 apart, a scale of 2^-+700 where a square could leave the range, and
 * each square as an exact head and tail by Dekker's split at 2^27 + 1. */
double phypot(double x, double y) {
    uint64_t ux = bits(x) & 0x7fffffffffffffffULL, uy = bits(y) & 0x7fffffffffffffffULL;
    if (ux < uy) {
        const uint64_t t = ux;
        ux = uy;
        uy = t;
    }
    const int ex = (int)(ux >> 52), ey = (int)(uy >> 52);
    x = from_bits(ux);
    y = from_bits(uy);
    if (ey == 0x7ff) return y;
    if (ex == 0x7ff || uy == 0) return x;
    if (ex - ey > 64) return x + y;
    double z = 1.0;
    if (ex > 0x3ff + 510) {
        z = 0x1p700;
        x *= 0x1p-700;
        y *= 0x1p-700;
    } else if (ey < 0x3ff - 450) {
        z = 0x1p-700;
        x *= 0x1p700;
        y *= 0x1p700;
    }
    auto square = [](double v, double &hi, double &lo) {
        const double vc = v * 134217729.0;  // 2^27 + 1
        const double vh = v - vc + vc, vl = v - vh;
        hi = v * v;
        lo = vh * vh - hi + 2.0 * vh * vl + vl * vl;
    };
    double hx, lx, hy, ly;
    square(x, hx, lx);
    square(y, hy, ly);
    return z * std::sqrt(ly + lx + hy + hx);
}

/* The arctangent (src/portable/mod.rs:253-255, the `libm` crate's port of fdlibm's s_atan.c, Sun
 * Microsystems 1993): |x| reduced to [-7/16, 7/16] about 0, 0.5, 1, 1.5 or inf, atan of the
 * reduced argument by fdlibm's odd degree-23 polynomial in two halves, plus atan of the centre as
 * a head and tail. */
double patan(double x) {
    static const uint64_t hi_bits[4] = {0x3fddac670561bb4fULL, 0x3fe921fb54442d18ULL, 0x3fef730bd281f69bULL,
                                        0x3ff921fb54442d18ULL};
    static const uint64_t lo_bits[4] = {0x3c7a2b7f222f65e2ULL, 0x3c81a62633145c07ULL, 0x3c7007887af0cbbdULL,
                                        0x3c91a62633145c07ULL};
    static const uint64_t at_bits[11] = {0x3fd555555555550dULL, 0xbfc999999998ebc4ULL, 0x3fc24924920083ffULL,
                                         0xbfbc71c6fe231671ULL, 0x3fb745cdc54c206eULL, 0xbfb3b0f2af749a6dULL,
                                         0x3fb10d66a0d03d51ULL, 0xbfadde2d52defd9aULL, 0x3fa97b4b24760debULL,
                                         0xbfa2b4442c6a6c2fULL, 0x3f90ad3ae322da11ULL};
    double at[11];
    for (int q = 0; q < 11; q++) at[q] = from_bits(at_bits[q]);
    uint32_t ix = hi_word(x);
    const bool negative = (ix >> 31) != 0;
    ix &= 0x7fffffffu;
    int id;
    if (ix >= 0x44100000u) {  // |x| >= 2^66 or NaN
        if (std::isnan(x)) return x;
        const double z = from_bits(hi_bits[3]);
        return negative ? -z : z;
    }
    if (ix < 0x3fdc0000u) {  // |x| < 7/16
        if (ix < 0x3e400000u) return x;
        id = -1;
    } else {
        x = std::fabs(x);
        if (ix < 0x3ff30000u) {
            if (ix < 0x3fe60000u) {
                id = 0;
                x = (2.0 * x - 1.0) / (2.0 + x);
            } else {
                id = 1;
                x = (x - 1.0) / (x + 1.0);
            }
        } else if (ix < 0x40038000u) {
            id = 2;
            x = (x - 1.5) / (1.0 + 1.5 * x);
        } else {
            id = 3;
            x = -1.0 / x;
        }
    }
    const double z = x * x, w = z * z;
    const double s1 = z * (at[0] + w * (at[2] + w * (at[4] + w * (at[6] + w * (at[8] + w * at[10])))));
    const double s2 = w * (at[1] + w * (at[3] + w * (at[5] + w * (at[7] + w * at[9]))));
    if (id < 0) return x - x * (s1 + s2);
    const double r = from_bits(hi_bits[id]) - (x * (s1 + s2) - from_bits(lo_bits[id]) - x);
    return negative ? -r : r;
}

/* The angle of (x, y) in [-pi, pi] (src/portable/mod.rs:265-266, the `libm` crate's port of
 * fdlibm's e_atan2.c): C's cases for zeros and infinities by the signs, pi/2 where |y/x| > 2^64,
 * else atan|y/x| moved to the quadrant with pi as a head and tail. */
double patan2(double y, double x) {
    const double pi = from_bits(0x400921fb54442d18ULL), pi_lo = from_bits(0x3ca1a62633145c07ULL);
    if (std::isnan(x) || std::isnan(y)) return x + y;
    uint32_t ix = hi_word(x), iy = hi_word(y);
    const uint32_t lx = lo_word(x), ly = lo_word(y);
    if (((ix - 0x3ff00000u) | lx) == 0) return patan(y);  // x = 1
    const uint32_t m = ((iy >> 31) & 1u) | ((ix >> 30) & 2u);  // 2 sign(x) + sign(y)
    ix &= 0x7fffffffu;
    iy &= 0x7fffffffu;
    if ((iy | ly) == 0) {
        switch (m) {
        case 0:
        case 1: return y;
        case 2: return pi;
        default: return -pi;
        }
    }
    if ((ix | lx) == 0) return (m & 1u) ? -pi / 2.0 : pi / 2.0;
    if (ix == 0x7ff00000u) {
        if (iy == 0x7ff00000u) {
            switch (m) {
            case 0: return pi / 4.0;
            case 1: return -pi / 4.0;
            case 2: return 3.0 * pi / 4.0;
            default: return -3.0 * pi / 4.0;
            }
        }
        switch (m) {
        case 0: return 0.0;
        case 1: return -0.0;
        case 2: return pi;
        default: return -pi;
        }
    }
    if (ix + (64u << 20) < iy || iy == 0x7ff00000u) return (m & 1u) ? -pi / 2.0 : pi / 2.0;
    const double z = ((m & 2u) && iy + (64u << 20) < ix) ? 0.0 : patan(std::fabs(y / x));
    switch (m) {
    case 0: return z;
    case 1: return -z;
    case 2: return pi - (z - pi_lo);
    default: return (z - pi_lo) - pi;
    }
}

/* The arcsine (src/portable/mod.rs:249-251, the `libm` crate's port of fdlibm's e_asin.c):
 * x + x R(x^2) below 0.5 with fdlibm's rational R, pi/2 - 2 asin(sqrt((1 - |x|)/2)) above, the
 * square root split into a 32-bit head and its remainder between 0.5 and 0.975. */
double pasin(double x) {
    const double pio2_hi = from_bits(0x3ff921fb54442d18ULL), pio2_lo = from_bits(0x3c91a62633145c07ULL);
    auto rational = [](double z) {
        const double p0 = from_bits(0x3fc5555555555555ULL), p1 = from_bits(0xbfd4d61203eb6f7dULL),
                     p2 = from_bits(0x3fc9c1550e884455ULL), p3 = from_bits(0xbfa48228b5688f3bULL),
                     p4 = from_bits(0x3f49efe07501b288ULL), p5 = from_bits(0x3f023de10dfdf709ULL);
        const double q1 = from_bits(0xc0033a271c8a2d4bULL), q2 = from_bits(0x40002ae59c598ac8ULL),
                     q3 = from_bits(0xbfe6066c1b8d0159ULL), q4 = from_bits(0x3fb3b8c5b12e9282ULL);
        const double p = z * (p0 + z * (p1 + z * (p2 + z * (p3 + z * (p4 + z * p5)))));
        const double q = 1.0 + z * (q1 + z * (q2 + z * (q3 + z * q4)));
        return p / q;
    };
    const uint32_t hx = hi_word(x), ix = hx & 0x7fffffffu;
    if (ix >= 0x3ff00000u) {  // |x| >= 1 or NaN
        if (((ix - 0x3ff00000u) | lo_word(x)) == 0) return x * pio2_hi;
        return 0.0 / (x - x);
    }
    if (ix < 0x3fe00000u) {  // |x| < 0.5
        if (ix < 0x3e500000u && ix >= 0x00100000u) return x;
        return x + x * rational(x * x);
    }
    const double z = (1.0 - std::fabs(x)) * 0.5, s = std::sqrt(z), r = rational(z);
    double v;
    if (ix >= 0x3fef3333u) {  // |x| > 0.975
        v = pio2_hi - (2.0 * (s + s * r) - pio2_lo);
    } else {
        const double f = with_lo(s, 0), c = (z - f * f) / (s + f);
        v = 0.5 * pio2_hi - (2.0 * s * r - (pio2_lo - 2.0 * c) - (0.5 * pio2_hi - 2.0 * f));
    }
    return (hx >> 31) ? -v : v;
}

namespace {
// log(1 + f) - f + f^2/2 for 1 + f in [sqrt(2)/2, sqrt(2)): fdlibm's polynomial in s^2,
// s = f / (2 + f), the part musl's log and log1p share
struct LogPoly {
    double s, hfsq, r;
};
LogPoly log_poly(double f) {
    const double lg1 = from_bits(0x3fe5555555555593ULL), lg2 = from_bits(0x3fd999999997fa04ULL);
    const double lg3 = from_bits(0x3fd2492494229359ULL), lg4 = from_bits(0x3fcc71c51d8e78afULL);
    const double lg5 = from_bits(0x3fc7466496cb03deULL), lg6 = from_bits(0x3fc39a09d078c69fULL);
    const double lg7 = from_bits(0x3fc2f112df3e5244ULL);
    LogPoly q;
    q.hfsq = 0.5 * f * f;
    q.s = f / (2.0 + f);
    const double z = q.s * q.s, w = z * z;
    const double t1 = w * (lg2 + w * (lg4 + w * lg6));
    const double t2 = z * (lg1 + w * (lg3 + w * (lg5 + w * lg7)));
    q.r = t2 + t1;
    return q;
}
const double LN2_HI = from_bits(0x3fe62e42fee00000ULL), LN2_LO = from_bits(0x3dea39ef35793c76ULL);

/* The natural logarithm as the `libm` crate 0.2.16 computes it (musl's log.c, its shorter
 * rewrite of fdlibm's e_log.c, whose bits differ at times from the fdlibm port `pln`): the
 * reference's acosh takes it through `libm::acosh` (src/portable/mod.rs:310-315). */
double musl_log(double x) {
    uint64_t u = bits(x);
    uint32_t hx = (uint32_t)(u >> 32);
    int k = 0;
    if (hx < 0x00100000u || (hx >> 31)) {
        if ((u << 1) == 0) return -1.0 / (x * x);
        if (hx >> 31) return (x - x) / 0.0;
        k -= 54;  // subnormal: scale up
        x *= 0x1p54;
        u = bits(x);
        hx = (uint32_t)(u >> 32);
    } else if (hx >= 0x7ff00000u) {
        return x;
    } else if (hx == 0x3ff00000u && (u << 32) == 0) {
        return 0.0;
    }
    hx += 0x3ff00000u - 0x3fe6a09eu;
    k += (int)(hx >> 20) - 0x3ff;
    hx = (hx & 0x000fffffu) + 0x3fe6a09eu;
    x = from_bits((uint64_t)hx << 32 | (u & 0xffffffffULL));
    const double f = x - 1.0;
    const LogPoly q = log_poly(f);
    const double dk = (double)k;
    return q.s * (q.hfsq + q.r) + dk * LN2_LO - q.hfsq + f + dk * LN2_HI;
}
}  // namespace

/* ln(1 + x), accurate near 0 (src/portable/mod.rs:236-238, the `libm` crate's port of musl's
 * log1p.c after fdlibm's s_log1p.c): 1 + x = 2^k (1 + f) with a correction c for the rounding of
 * 1 + x, then the logarithm's polynomial. */
double plog1p(double x) {
    const uint32_t hx = hi_word(x);
    int k = 1;
    double c = 0.0, f = 0.0;
    if (hx < 0x3fda827au || (hx >> 31)) {  // 1 + x < sqrt(2)
        if (hx >= 0xbff00000u) {            // x <= -1
            if (x == -1.0) return x / 0.0;
            return (x - x) / 0.0;
        }
        if ((hx << 1) < (0x3ca00000u << 1)) return x;  // |x| < 2^-53
        if (hx <= 0xbfd2bec4u) {                        // sqrt(2)/2 <= 1 + x < sqrt(2)
            k = 0;
            c = 0.0;
            f = x;
        }
    } else if (hx >= 0x7ff00000u) {
        return x;
    }
    if (k) {
        const double u = 1.0 + x;
        uint64_t ui = bits(u);
        uint32_t hu = (uint32_t)(ui >> 32);
        hu += 0x3ff00000u - 0x3fe6a09eu;
        k = (int)(hu >> 20) - 0x3ff;
        if (k < 54) {
            c = k >= 2 ? 1.0 - (u - x) : x - (u - 1.0);
            c /= u;
        } else {
            c = 0.0;
        }
        hu = (hu & 0x000fffffu) + 0x3fe6a09eu;
        ui = (uint64_t)hu << 32 | (ui & 0xffffffffULL);
        f = from_bits(ui) - 1.0;
    }
    const LogPoly q = log_poly(f);
    const double dk = (double)k;
    return q.s * (q.hfsq + q.r) + (dk * LN2_LO + c) - q.hfsq + f + dk * LN2_HI;
}

/* The inverse hyperbolic cosine (src/portable/mod.rs:310-315): NaN below 1, else the `libm`
 * crate's port of musl's acosh.c, log1p(t + sqrt(t^2 + 2t)) with t = x - 1 below 2,
 * log(2x - 1/(x + sqrt(x^2 - 1))) below 2^26 and log x + ln 2 above, with musl's log. */
double pacosh(double x) {
    if (x < 1.0) return NAN;
    const unsigned e = (unsigned)(bits(x) >> 52) & 0x7ffu;
    if (e < 0x3ff + 1) return plog1p(x - 1.0 + std::sqrt((x - 1.0) * (x - 1.0) + 2.0 * (x - 1.0)));
    if (e < 0x3ff + 26) return musl_log(2.0 * x - 1.0 / (x + std::sqrt(x * x - 1.0)));
    return musl_log(x) + from_bits(0x3fe62e42fefa39efULL);
}

double pcabs(C64 z) { return phypot(z.re, z.im); }
double pcarg(C64 z) { return patan2(z.im, z.re); }

/* The principal square root (src/portable/complex.rs:132-180): Kahan's formulas, t =
 * sqrt((|a| + |z|)/2), then (t, b/2t) for a >= 0 and (|b|/2t, +-t) below, scaled by 1/4 (and 2
 * after) near overflow and by 2^54 (and 2^-27) near underflow; C99's special values, a zero
 * imaginary part's sign kept on the cut. */
C64 pcsqrt(C64 z) {
    const double big = from_bits(0x7fc0000000000000ULL), small = from_bits(0x0170000000000000ULL);
    const double up = from_bits(0x4350000000000000ULL), down = from_bits(0x3e40000000000000ULL);
    double x = z.re, y = z.im;
    if (y == 0.0) {
        if (!std::signbit(x)) return {std::sqrt(x), y};
        return {0.0, std::copysign(std::sqrt(-x), y)};
    }
    if (std::isinf(y)) return {INFINITY, y};
    if (std::isnan(x) || std::isnan(y)) return {NAN, NAN};
    if (std::isinf(x)) return x > 0.0 ? C64{x, std::copysign(0.0, y)} : C64{0.0, std::copysign(INFINITY, y)};
    const double m = std::fmax(std::fabs(x), std::fabs(y));
    double scale = 1.0;
    if (m >= big) {
        x *= 0.25;
        y *= 0.25;
        scale = 2.0;
    } else if (m < small) {
        x *= up;
        y *= up;
        scale = down;
    }
    const double t = std::sqrt(0.5 * (std::fabs(x) + phypot(x, y)));
    double re, im;
    if (x >= 0.0) {
        re = t;
        im = y / (2.0 * t);
    } else {
        re = std::fabs(y) / (2.0 * t);
        im = std::copysign(t, y);
    }
    return {re * scale, im * scale};
}

/* The principal logarithm (src/portable/complex.rs:106-116): ln|z| + i arg z, the real part as
 * ln_1p((a - 1)(a + 1) + b^2)/2 where |z| is within [1/sqrt(2), sqrt(2)] (a >= b the parts'
 * sizes). */
C64 pcln(C64 z) {
    const double r = pcabs(z), theta = pcarg(z);
    if (r >= from_bits(0x3fe6a09e667f3bcdULL) && r <= from_bits(0x3ff6a09e667f3bcdULL)) {
        const double x = std::fabs(z.re), y = std::fabs(z.im);
        const double a = x >= y ? x : y, b = x >= y ? y : x;
        const double t = (a - 1.0) * (a + 1.0) + b * b;
        return {0.5 * plog1p(t), theta};
    }
    return {pln(r), theta};
}

/* The principal arcsine (src/portable/complex.rs:344-407): Hull, Fairgrieve and Tang's algorithm
 * (ACM TOMS 23, 299, 1997) on the first quadrant, r = |z + 1|, s = |z - 1|, A = (r + s)/2,
 * B = x/A; the real part asin B up to B = 0.6417 and an arctangent of positive sums above, the
 * imaginary ln_1p of positive sums up to A = 1.5 and ln(A + sqrt(A^2 - 1)) above; z itself below
 * 2^-27, atan2 and ln|z| + ln 2 beyond 2^500, the limits below an imaginary part of 2^-500; the
 * signs put back by asin(-z) = -asin z and asin(conj z) = conj asin z. */
C64 pcasin(C64 z) {
    const double large = from_bits(0x5f30000000000000ULL), tiny = from_bits(0x20b0000000000000ULL);
    const double small = from_bits(0x3e40000000000000ULL), half_pi = from_bits(0x3ff921fb54442d18ULL);
    const double x = std::fabs(z.re), y = std::fabs(z.im);
    if (std::isnan(x) || std::isnan(y)) {
        if (std::isinf(y)) return {NAN, z.im};
        if (x == 0.0) return {z.re, NAN};
        if (std::isinf(x)) return {NAN, INFINITY};
        return {NAN, NAN};
    }
    double re, im;
    const double big = std::fmax(x, y);
    if (big < small) {
        re = x;
        im = y;
    } else if (big >= large) {
        re = patan2(x, y);
        im = pln(phypot(x, y)) + from_bits(0x3fe62e42fefa39efULL);
    } else if (y < tiny) {
        if (x < 1.0) {
            re = pasin(x);
            im = y / std::sqrt((1.0 - x) * (1.0 + x));
        } else if (x == 1.0) {
            re = half_pi;
            im = std::sqrt(y);
        } else {
            re = half_pi;
            im = pacosh(x);
        }
    } else {
        const double xp1 = x + 1.0;
        const double r = phypot(xp1, y), s = phypot(x - 1.0, y);
        const double a = 0.5 * (r + s), b = x / a, yy = y * y;
        if (b <= 0.6417) {
            re = pasin(b);
        } else {
            const double apx = a + x;
            if (x <= 1.0)
                re = patan(x / std::sqrt(0.5 * apx * (yy / (r + xp1) + (s + (1.0 - x)))));
            else
                re = patan(x / (y * std::sqrt(0.5 * (apx / (r + xp1) + apx / (s + (x - 1.0))))));
        }
        if (a <= 1.5) {
            const double am1 = x < 1.0 ? 0.5 * (yy / (r + xp1) + yy / (s + (1.0 - x)))
                                       : 0.5 * (yy / (r + xp1) + (s + (x - 1.0)));
            im = plog1p(am1 + std::sqrt(am1 * (a + 1.0)));
        } else {
            im = pln(a + std::sqrt(a * a - 1.0));
        }
    }
    return {std::copysign(re, z.re), std::copysign(im, z.im)};
}

}  // namespace pxfdtd

extern "C" int pxfdtd_portable(const char *name, double a, double b, double *re, double *im) {
    using namespace pxfdtd;
    if (!name || !re || !im) return px_fail(PXFDTD_ERR_INVALID, "pxfdtd_portable needs a name and two outputs");
    const std::string f(name);
    double r = 0.0, i = 0.0;
    C64 w{0.0, 0.0};
    bool complex = false;
    if (f == "exp") r = pexp(a);
    else if (f == "ln") r = pln(a);
    else if (f == "ln_1p") r = plog1p(a);
    else if (f == "asin") r = pasin(a);
    else if (f == "atan") r = patan(a);
    else if (f == "acosh") r = pacosh(a);
    else if (f == "atan2") r = patan2(a, b);
    else if (f == "hypot") r = phypot(a, b);
    else if (f == "powf") r = ppow(a, b);
    else if (f == "powi") r = ppowi(a, (int)b);
    else {
        complex = true;
        const C64 z{a, b};
        if (f == "c_sqrt") w = pcsqrt(z);
        else if (f == "c_asin") w = pcasin(z);
        else if (f == "c_ln") w = pcln(z);
        else if (f == "c_polar") w = {pcabs(z), pcarg(z)};
        else return px_fail(PXFDTD_ERR_INVALID, "no portable function named '%s'", name);
    }
    *re = complex ? w.re : r;
    *im = complex ? w.im : i;
    return PXFDTD_OK;
}
