// Bit-portable maths for the host's waveforms, transform phases and CPML coefficients: the
// functions photonoxide computes them with (src/portable/mod.rs:66-138 exp, :321-323 sin_cos,
// :343-356 powi, :141-220 ln, :299-301 powf; src/portable/complex.rs:80-98 complex exp), from the
// published algorithms they port: fdlibm's e_exp.c, e_log.c and e_pow.c (Sun Microsystems, 1993
// and 2004) and musl's sincos (fdlibm kernels after the medium-size Cody-Waite reduction). Only
// + - * /, sqrt and exact bit operations, compiled without contraction, so a waveform's value
// and a CPML coefficient have the same bits as the reference's.
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
}  // namespace pxfdtd
