// Harmonic inversion (harmonic_inversion, src/fdtd/harmonic.rs:62-184): a signal's frequencies,
// decay rates and amplitudes by the filter diagonalization of V. A. Mandelshtam and H. S. Taylor,
// J. Chem. Phys. 107, 6756 (1997), Section II and its summary of the algorithm, as the reference
// applies it:
// - the signal c_n = sum_k d_k e^(-i n tau w_k), sampled every tau, on J points z_j = e^(-i phi_j)
//   across the window (J = ceil(N tau (w_max - w_min) / 4 pi), at least 8), M = (N - 3) / 2;
// - U^(p)(z, z') = sum_{n, n' = 0..M} z^-n z'^-n' c_(n + n' + p) by their Eq. 25's closed form,
//   [z f(z') - z' f(z) - z^-M g(z') + z'^-M g(z)] / (z - z'), and on the diagonal the weighted
//   sum (harmonic.rs:186-220); f and g are taken once per point here, each as the reference
//   sums it;
// - U^(1) B = u U^(0) B on U^(0)'s singular vectors above 1e-11 of its largest (a one-sided
//   Jacobi SVD here, faer's there), the reduced matrix's eigenvalues by Householder reduction and
//   the Hessenberg QR of eigen.hpp;
// - each u_k in the window with B normalized as B^T U^(0) B = 1, its error
//   |(U^(2) - u^2 U^(0)) B| / |u^2 U^(0) B| and its amplitude ((1/(M + 1)) sum_j B_j
//   U^(0)(z_j, u_k))^2 (Eq. 27); sorted by frequency.
// The eigensolvers differ from the reference's, so results agree to round-off, not bit for bit.
#pragma once
#include <algorithm>
#include <cmath>
#include <complex>
#include <string>
#include <vector>

#include "eigen.hpp"
#include "portable.hpp"

namespace pxharm {

using cd = std::complex<double>;

struct Resonance {
    double frequency, decay;
    cd amplitude;
    double error;
};

// f(z) = sum_{l=0..M} c_(l+p) z^-l and g(z) = sum_{l=M+1..2M} c_(l+p) z^-(l-M-1), each by
// repeated products as the reference sums them, and z^-M
struct Sums {
    cd f, g, zm;
};

inline Sums sums_at(const std::vector<cd> &c, int m, int p, cd z) {
    const cd zi = 1.0 / z;
    Sums s{0.0, 0.0, 0.0};
    cd pz = 1.0;
    for (int l = 0; l <= m; l++) {
        s.f += c[l + p] * pz;
        pz *= zi;
    }
    s.zm = pz * z;
    cd qz = 1.0;
    for (int l = m + 1; l <= 2 * m; l++) {
        s.g += c[l + p] * qz;
        qz *= zi;
    }
    return s;
}

// the diagonal: sum_{l=0..2M} (M - |M - l| + 1) c_(l+p) z^-l
inline cd diagonal(const std::vector<cd> &c, int m, int p, cd z) {
    const cd zi = 1.0 / z;
    cd total = 0.0, power = 1.0;
    for (int l = 0; l <= 2 * m; l++) {
        const double weight = (double)(m + 1 - std::abs(l - m));
        total += weight * c[l + p] * power;
        power *= zi;
    }
    return total;
}

// A = W S V^H by one-sided Jacobi on A's columns (n x n, row-major): the singular values in
// descending order with W's and V's columns in the same order
inline void svd(const std::vector<cd> &A, int n, std::vector<cd> &W, std::vector<double> &S, std::vector<cd> &V) {
    std::vector<cd> a(A);
    V.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) V[(size_t)i * n + i] = 1.0;
    for (int sweep = 0; sweep < 60; sweep++) {
        double off = 0.0;
        for (int p = 0; p < n - 1; p++)
            for (int q = p + 1; q < n; q++) {
                double alpha = 0.0, beta = 0.0;
                cd gamma = 0.0;
                for (int r = 0; r < n; r++) {
                    const cd ap = a[(size_t)r * n + p], aq = a[(size_t)r * n + q];
                    alpha += std::norm(ap);
                    beta += std::norm(aq);
                    gamma += std::conj(ap) * aq;
                }
                const double g = std::abs(gamma);
                if (g <= 1e-15 * std::sqrt(alpha * beta) || g == 0.0) continue;
                off = std::max(off, g / std::sqrt(alpha * beta));
                const cd phase = gamma / g;
                const double zeta = (beta - alpha) / (2.0 * g);
                const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
                const double cs = 1.0 / std::sqrt(1.0 + t * t), sn = cs * t;
                // [a_p, a_q] <- [a_p, a_q] [[c, s e^(i phi)], [-s e^(-i phi), c]]
                for (int r = 0; r < n; r++) {
                    const cd ap = a[(size_t)r * n + p], aq = a[(size_t)r * n + q];
                    a[(size_t)r * n + p] = cs * ap - sn * std::conj(phase) * aq;
                    a[(size_t)r * n + q] = sn * phase * ap + cs * aq;
                    const cd vp = V[(size_t)r * n + p], vq = V[(size_t)r * n + q];
                    V[(size_t)r * n + p] = cs * vp - sn * std::conj(phase) * vq;
                    V[(size_t)r * n + q] = sn * phase * vp + cs * vq;
                }
            }
        if (off <= 1e-15) break;
    }
    std::vector<double> s(n);
    for (int k = 0; k < n; k++) {
        double v = 0.0;
        for (int r = 0; r < n; r++) v += std::norm(a[(size_t)r * n + k]);
        s[k] = std::sqrt(v);
    }
    std::vector<int> order(n);
    for (int k = 0; k < n; k++) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return s[x] > s[y]; });
    W.assign((size_t)n * n, 0.0);
    S.assign(n, 0.0);
    std::vector<cd> v2((size_t)n * n);
    for (int k = 0; k < n; k++) {
        const int o = order[k];
        S[k] = s[o];
        for (int r = 0; r < n; r++) {
            W[(size_t)r * n + k] = s[o] > 0.0 ? a[(size_t)r * n + o] / s[o] : cd(0.0, 0.0);
            v2[(size_t)r * n + k] = V[(size_t)r * n + o];
        }
    }
    V.swap(v2);
}

// The eigenpairs of a general complex n x n matrix (row-major): Householder reduction to
// Hessenberg form, then pxeig's Schur form; eigenvectors of unit length. False if the QR fails.
inline bool eigen(const std::vector<cd> &A, int n, std::vector<cd> &values, std::vector<cd> &vectors) {
    std::vector<cd> H(A), Qh((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) Qh[(size_t)i * n + i] = 1.0;
    for (int k = 0; k + 2 < n; k++) {
        double xn = 0.0;
        for (int r = k + 1; r < n; r++) xn += std::norm(H[(size_t)r * n + k]);
        xn = std::sqrt(xn);
        if (xn == 0.0) continue;
        const cd x0 = H[(size_t)(k + 1) * n + k];
        const cd alpha = (std::abs(x0) > 0.0 ? -x0 / std::abs(x0) : cd(-1.0, 0.0)) * xn;
        std::vector<cd> v((size_t)n, 0.0);
        for (int r = k + 1; r < n; r++) v[r] = H[(size_t)r * n + k];
        v[k + 1] -= alpha;
        double vn = 0.0;
        for (int r = k + 1; r < n; r++) vn += std::norm(v[r]);
        vn = std::sqrt(vn);
        if (vn == 0.0) continue;
        for (int r = k + 1; r < n; r++) v[r] /= vn;
        // H <- (I - 2 v v^H) H (I - 2 v v^H), Qh <- Qh (I - 2 v v^H)
        for (int c = 0; c < n; c++) {
            cd d = 0.0;
            for (int r = k + 1; r < n; r++) d += std::conj(v[r]) * H[(size_t)r * n + c];
            for (int r = k + 1; r < n; r++) H[(size_t)r * n + c] -= 2.0 * v[r] * d;
        }
        for (int r = 0; r < n; r++) {
            cd d = 0.0;
            for (int c = k + 1; c < n; c++) d += H[(size_t)r * n + c] * v[c];
            for (int c = k + 1; c < n; c++) H[(size_t)r * n + c] -= 2.0 * d * std::conj(v[c]);
            cd e = 0.0;
            for (int c = k + 1; c < n; c++) e += Qh[(size_t)r * n + c] * v[c];
            for (int c = k + 1; c < n; c++) Qh[(size_t)r * n + c] -= 2.0 * e * std::conj(v[c]);
        }
        for (int r = k + 2; r < n; r++) H[(size_t)r * n + k] = 0.0;
    }
    std::vector<cd> Qs;
    if (!pxeig::hessenberg_schur(H, Qs, n)) return false;
    values.assign(n, 0.0);
    vectors.assign((size_t)n * n, 0.0);  // column k: eigenvector k
    for (int k = 0; k < n; k++) {
        values[k] = H[(size_t)k * n + k];
        const std::vector<cd> z = pxeig::triangular_vector(H, n, k);
        std::vector<cd> y((size_t)n, 0.0), x((size_t)n, 0.0);
        for (int r = 0; r < n; r++)
            for (int c = 0; c < n; c++) y[r] += Qs[(size_t)r * n + c] * z[c];
        double xn = 0.0;
        for (int r = 0; r < n; r++) {
            for (int c = 0; c < n; c++) x[r] += Qh[(size_t)r * n + c] * y[c];
            xn += std::norm(x[r]);
        }
        xn = std::sqrt(xn);
        for (int r = 0; r < n; r++) vectors[(size_t)r * n + k] = x[r] / xn;
    }
    return true;
}

// The signal's sums at points z for shift p (sums_at) and the diagonal weighted sum (diagonal),
// each point's from the same operations in the same order wherever they are computed: the host
// below, or the device (harmonic_device.hip) from a signal held there.
struct Evaluator {
    virtual ~Evaluator() = default;
    // for each point: its sums and its diagonal; false with the reason in *why on failure
    virtual bool evaluate(int p, const std::vector<cd> &z, std::vector<Sums> &sums, std::vector<cd> &diag,
                          std::string *why) const = 0;
};

struct HostSignal : Evaluator {
    const std::vector<cd> &c;
    int m;
    HostSignal(const std::vector<cd> &signal, int m_) : c(signal), m(m_) {}
    bool evaluate(int p, const std::vector<cd> &z, std::vector<Sums> &sums, std::vector<cd> &diag,
                  std::string *) const override {
        const int n = (int)z.size();
        sums.resize(n);
        diag.resize(n);
#pragma omp parallel for schedule(static) if (n * (long long)m > 200000)
        for (int q = 0; q < n; q++) {
            sums[q] = sums_at(c, m, p, z[q]);
            diag[q] = diagonal(c, m, p, z[q]);
        }
        return true;
    }
};

// element of U^(p) at (z, w) from the points' sums, z's diagonal where the points meet
inline cd element_of(cd z, const Sums &sz, cd dz, cd w, const Sums &sw) {
    if (std::abs(z - w) <= 1e-13 * std::max(std::abs(z), std::abs(w))) return dz;
    return (z * sw.f - w * sz.f - sz.zm * sw.g + sw.zm * sz.g) / (z - w);
}

// harmonic inversion's checks, and its parameters for n samples every dt over [f_low, f_high]:
// M and the points z_j; false with the reason in *why
inline bool harmonic_setup(long long n, bool finite, double dt, double f_low, double f_high, int &m,
                           std::vector<cd> &z, std::string *why) {
    if (!(std::isfinite(dt) && dt > 0.0)) {
        *why = "harmonic inversion needs a positive, finite time step";
        return false;
    }
    if (!(f_high > f_low) || f_high * dt >= 0.5) {
        *why = "harmonic inversion's window must be from low to high below the Nyquist frequency " +
               std::to_string(0.5 / dt) + " c/um";
        return false;
    }
    if (n < 8 || !finite) {
        *why = "harmonic inversion needs at least 8 finite samples";
        return false;
    }
    const double TAU_ = 6.283185307179586, PI_ = 3.141592653589793;
    m = (int)((n - 3) / 2);
    const double phi_low = TAU_ * f_low * dt, phi_high = TAU_ * f_high * dt;
    const int j = std::max(8, (int)std::ceil((double)n * (phi_high - phi_low) / (4.0 * PI_)));
    z.resize(j);
    for (int q = 0; q < j; q++) {
        const double phi = phi_low + ((double)q + 0.5) * (phi_high - phi_low) / (double)j;
        const pxfdtd::C64 p = pxfdtd::pcexp({0.0, -phi});  // portable::complex::exp
        z[q] = cd(p.re, p.im);
    }
    return true;
}

// w = (-arg u, ln |u|) / tau, with the portable modulus, argument and logarithm
inline cd angular(const cd &u, double tau) {
    const pxfdtd::C64 v = {u.real(), u.imag()};
    const double re = -pxfdtd::pcarg(v), im = pxfdtd::pln(pxfdtd::pcabs(v));
    return cd(re / tau, im / tau);
}

// the filter diagonalization from the signal's sums (taken by `ev`) at the points z
inline bool harmonic_solve(const Evaluator &ev, int m, const std::vector<cd> &z, double dt, double f_low,
                           double f_high, std::vector<Resonance> &found, std::string *why) {
    found.clear();
    const double TAU_ = 6.283185307179586;
    const double tau = dt;
    const int j = (int)z.size();
    std::vector<Sums> sums[3];
    std::vector<cd> diag[3];
    for (int p = 0; p < 3; p++)
        if (!ev.evaluate(p, z, sums[p], diag[p], why)) return false;
    std::vector<cd> U[3];
    for (int p = 0; p < 3; p++) {
        U[p].assign((size_t)j * j, 0.0);
#pragma omp parallel for schedule(dynamic) if (j * (long long)j > 40000)
        for (int a = 0; a < j; a++)
            for (int b = 0; b < j; b++)
                U[p][(size_t)a * j + b] = element_of(z[a], sums[p][a], diag[p][a], z[b], sums[p][b]);
    }
    std::vector<cd> W, Vs;
    std::vector<double> S;
    svd(U[0], j, W, S, Vs);
    int rank = 0;
    for (int k = 0; k < j; k++)
        if (S[k] > 1e-11 * S[0]) rank++;
    if (rank == 0) return true;
    // S^-1 W^H U1 V y = u y on the kept singular vectors
    std::vector<cd> red((size_t)rank * rank, 0.0), t1((size_t)j * rank, 0.0);
    for (int r = 0; r < j; r++)
        for (int b = 0; b < rank; b++) {
            cd s = 0.0;
            for (int c = 0; c < j; c++) s += U[1][(size_t)r * j + c] * Vs[(size_t)c * j + b];
            t1[(size_t)r * rank + b] = s;
        }
    for (int a = 0; a < rank; a++)
        for (int b = 0; b < rank; b++) {
            cd s = 0.0;
            for (int r = 0; r < j; r++) s += std::conj(W[(size_t)r * j + a]) * t1[(size_t)r * rank + b];
            red[(size_t)a * rank + b] = s / S[a];
        }
    std::vector<cd> values, vectors;
    if (!eigen(red, rank, values, vectors)) {
        *why = "harmonic inversion's eigenproblem failed";
        return false;
    }
    // the eigenvalues in the window, then the signal's sums at each at once
    std::vector<int> kept;
    std::vector<cd> uks;
    for (int k = 0; k < rank; k++) {
        const cd uk = values[k];
        if (pxfdtd::pcabs({uk.real(), uk.imag()}) == 0.0) continue;
        const cd omega = angular(uk, tau);
        const double frequency = omega.real() / TAU_;
        if (!(frequency >= f_low && frequency <= f_high)) continue;
        kept.push_back(k);
        uks.push_back(uk);
    }
    std::vector<Sums> su;
    std::vector<cd> du;
    if (!uks.empty() && !ev.evaluate(0, uks, su, du, why)) return false;
    for (size_t e = 0; e < kept.size(); e++) {
        const int k = kept[e];
        const cd uk = values[k];
        const cd omega = angular(uk, tau);
        const double frequency = omega.real() / TAU_;
        std::vector<cd> b((size_t)j, 0.0);
        for (int r = 0; r < j; r++)
            for (int c = 0; c < rank; c++) b[r] += Vs[(size_t)r * j + c] * vectors[(size_t)c * rank + k];
        auto times = [&](int p, const std::vector<cd> &x) {
            std::vector<cd> y((size_t)j, 0.0);
            for (int r = 0; r < j; r++)
                for (int c = 0; c < j; c++) y[r] += U[p][(size_t)r * j + c] * x[c];
            return y;
        };
        std::vector<cd> u0b = times(0, b);
        cd norm = 0.0;
        for (int r = 0; r < j; r++) norm += b[r] * u0b[r];
        const pxfdtd::C64 sq = pxfdtd::pcsqrt({norm.real(), norm.imag()});
        const double sn = sq.re * sq.re + sq.im * sq.im;  // num-complex's inv
        const cd scale = cd(sq.re / sn, 0.0 - sq.im / sn);
        for (cd &x : b) x *= scale;
        u0b = times(0, b);
        const std::vector<cd> u2b = times(2, b);
        const cd uk2 = uk * uk;
        double residual = 0.0, size = 0.0;
        for (int r = 0; r < j; r++) {
            residual += std::norm(u2b[r] - uk2 * u0b[r]);
            size += std::norm(uk2 * u0b[r]);
        }
        cd sum = 0.0;
        for (int a = 0; a < j; a++) sum += b[a] * element_of(z[a], sums[0][a], diag[0][a], uk, su[e]);
        const cd root = sum / (double)(m + 1);
        found.push_back({frequency, -omega.imag(), root * root, std::sqrt(residual) / std::sqrt(size)});
    }
    std::sort(found.begin(), found.end(), [](const Resonance &x, const Resonance &y) { return x.frequency < y.frequency; });
    return true;
}

// harmonic_inversion: false with the reason in *why for a refused input or a failed eigenproblem
inline bool harmonic_inversion(const std::vector<cd> &signal, double dt, double f_low, double f_high,
                               std::vector<Resonance> &found, std::string *why) {
    found.clear();
    bool finite = true;
    for (const cd &v : signal) finite = finite && std::isfinite(v.real()) && std::isfinite(v.imag());
    int m;
    std::vector<cd> z;
    if (!harmonic_setup((long long)signal.size(), finite, dt, f_low, f_high, m, z, why)) return false;
    return harmonic_solve(HostSignal(signal, m), m, z, dt, f_low, f_high, found, why);
}

}  // namespace pxharm
