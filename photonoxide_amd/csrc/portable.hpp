// Bit-portable maths for the host's waveforms, transform phases, CPML coefficients, beams, port
// modes and harmonic inversion: the functions photonoxide computes them with
// (src/portable/mod.rs:66-138 exp, :321-323 sin_cos, :343-356 powi, :141-220 ln, :299-301 powf,
// :230-315 the `libm` crate's; src/portable/complex.rs), from the published algorithms they port:
// fdlibm's e_exp.c, e_log.c, e_pow.c, s_atan.c, e_atan2.c and e_asin.c (Sun Microsystems, 1993
// and 2004), musl's sincos (fdlibm kernels after the medium-size Cody-Waite reduction), hypot,
// log1p, log and acosh, Kahan's complex square root and Hull, Fairgrieve and Tang's complex
// arcsine. Only + - * /, sqrt and exact bit operations, compiled without contraction, so each
// value has the same bits as the reference's.
#pragma once
#include <cstdint>

namespace pxfdtd {
struct C64 {
    double re, im;
};
double pexp(double x);
// false for |x| >= 2^20 pi/2, beyond the medium-size reduction (no run reaches it)
bool psincos(double x, double &s, double &c);
double ppowi(double x, int n);
// fdlibm e_log.c as src/portable/mod.rs:141-220 ports it; fdlibm e_pow.c as the `libm` crate
// 0.2.16 ports it (src/portable/mod.rs:299-301)
double pln(double x);
double ppow(double x, double y);
C64 pcexp(C64 z);
// the `libm` crate 0.2.16's hypot (musl), atan, atan2, asin (fdlibm), log1p (musl) and acosh
// (musl, NaN below 1) as src/portable/mod.rs:230-315 takes them
double phypot(double x, double y);
double patan(double x);
double patan2(double y, double x);
double pasin(double x);
double plog1p(double x);
double pacosh(double x);
// src/portable/complex.rs:26-39 abs and arg, :106-116 ln, :132-180 sqrt, :344-407 asin
double pcabs(C64 z);
double pcarg(C64 z);
C64 pcln(C64 z);
C64 pcsqrt(C64 z);
C64 pcasin(C64 z);
}  // namespace pxfdtd
