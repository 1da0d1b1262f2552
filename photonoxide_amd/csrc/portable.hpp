// Bit-portable maths for the host's waveforms, transform phases and CPML coefficients: the
// functions photonoxide computes them with (src/portable/mod.rs:66-138 exp, :321-323 sin_cos,
// :343-356 powi; src/portable/complex.rs:80-98 complex exp), from the published algorithms they
// port: fdlibm's e_exp.c (Sun Microsystems, 1993) and musl's sincos (fdlibm kernels after the
// medium-size Cody-Waite reduction). Only + - * / and exact bit operations, compiled without
// contraction, so a waveform's value has the same bits as the reference's.
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
C64 pcexp(C64 z);
}  // namespace pxfdtd
