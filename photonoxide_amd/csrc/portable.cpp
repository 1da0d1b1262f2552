// Bit-portable maths of the host layer (see portable.hpp).
#include "portable.hpp"

#include <cmath>
#include <cstring>

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

}  // namespace pxfdtd
