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

}  // namespace pxfdtd
