// The port-mode solver's linear algebra (host): the eigenpairs of a sparse complex matrix
// nearest a shift, by Arnoldi on the shifted inverse (A - s)^-1, whose solves use a banded LU
// with partial pivoting, and a complex Hessenberg QR for the small projected problem.
//
// The reference takes the same eigenpairs from its own Krylov-Schur (src/eigen.rs:147-181,
// crate::eigen::nearest); a mode is then fixed by its normalization, so the two agree to the
// solvers' tolerance, not bit for bit.
#pragma once
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <string>
#include <tuple>
#include <vector>

namespace pxeig {

using cd = std::complex<double>;

// A banded LU of an n x n matrix with kl sub- and ku super-diagonals: row i keeps columns
// i - kl to i + kl + ku (the extra kl for the pivots' fill), the multipliers of each step kept
// apart with the step's pivot row.
struct BandLU {
    long long n = 0;
    int kl = 0, ku = 0, w = 0;
    std::vector<cd> a;      // [n][w]
    std::vector<cd> l;      // [n][kl]
    std::vector<long long> piv;

    cd &at(long long i, long long c) { return a[(size_t)i * w + (size_t)(c - (i - kl))]; }

    // false if a pivot vanishes (the shifted matrix singular)
    bool factor(long long n_, int kl_, int ku_, const std::vector<std::tuple<long long, long long, cd>> &entries) {
        n = n_;
        kl = kl_;
        ku = ku_;
        w = 2 * kl + ku + 1;
        a.assign((size_t)n * w, cd(0.0, 0.0));
        l.assign((size_t)n * std::max(kl, 1), cd(0.0, 0.0));
        piv.assign((size_t)n, 0);
        for (const auto &e : entries) at(std::get<0>(e), std::get<1>(e)) += std::get<2>(e);
        for (long long i = 0; i < n; i++) {
            const long long last = std::min(n - 1, i + kl);
            long long p = i;
            double best = std::abs(at(i, i));
            for (long long r = i + 1; r <= last; r++)
                if (std::abs(at(r, i)) > best) {
                    best = std::abs(at(r, i));
                    p = r;
                }
            if (best == 0.0) return false;
            piv[i] = p;
            const long long cend = std::min(n - 1, i + kl + ku);
            if (p != i)
                for (long long c = i; c <= cend; c++) std::swap(at(i, c), at(p, c));
            const cd d = at(i, i);
            const long long len = cend - i;  // columns i + 1 .. cend
            const double *pivot_row = reinterpret_cast<const double *>(&at(i, i + 1));
            // the rows below the pivot are independent: in parallel when there is enough of them
#pragma omp parallel for schedule(static) if ((last - i) * len > 65536)
            for (long long r = i + 1; r <= last; r++) {
                const cd m = at(r, i) / d;
                l[(size_t)i * kl + (r - i - 1)] = m;
                at(r, i) = 0.0;
                if (m == cd(0.0, 0.0)) continue;
                // row r -= m row i, the complex products written out (no library call per value)
                const double mr = m.real(), mi = m.imag();
                double *row = reinterpret_cast<double *>(&at(r, i + 1));
                for (long long c = 0; c < len; c++) {
                    const double ar = pivot_row[2 * c], ai = pivot_row[2 * c + 1];
                    row[2 * c] -= mr * ar - mi * ai;
                    row[2 * c + 1] -= mr * ai + mi * ar;
                }
            }
        }
        return true;
    }

    void solve(std::vector<cd> &b) {
        double *x = reinterpret_cast<double *>(b.data());
        for (long long i = 0; i < n; i++) {
            if (piv[i] != i) std::swap(b[i], b[piv[i]]);
            const long long last = std::min(n - 1, i + kl);
            const double br = x[2 * i], bi = x[2 * i + 1];
            const double *li = reinterpret_cast<const double *>(&l[(size_t)i * kl]);
            for (long long r = i + 1; r <= last; r++) {
                const double mr = li[2 * (r - i - 1)], mi = li[2 * (r - i - 1) + 1];
                x[2 * r] -= mr * br - mi * bi;
                x[2 * r + 1] -= mr * bi + mi * br;
            }
        }
        for (long long i = n - 1; i >= 0; i--) {
            const long long cend = std::min(n - 1, i + kl + ku);
            double sr = x[2 * i], si = x[2 * i + 1];
            const double *row = reinterpret_cast<const double *>(&at(i, i));
            for (long long c = 1; c <= cend - i; c++) {
                const double ar = row[2 * c], ai = row[2 * c + 1], vr = x[2 * (i + c)], vi = x[2 * (i + c) + 1];
                sr -= ar * vr - ai * vi;
                si -= ar * vi + ai * vr;
            }
            b[i] = cd(sr, si) / at(i, i);
        }
    }
};

// The Schur form of a k x k upper Hessenberg H (row-major), in place by single-shift QR with
// Wilkinson's shift and deflation, Q accumulating the rotations (H_in = Q T Q^H); false if it
// does not converge.
inline bool hessenberg_schur(std::vector<cd> &H, std::vector<cd> &Q, int k) {
    auto h = [&](int i, int j) -> cd & { return H[(size_t)i * k + j]; };
    Q.assign((size_t)k * k, cd(0.0, 0.0));
    for (int i = 0; i < k; i++) Q[(size_t)i * k + i] = 1.0;
    const double eps = 2.220446049250313e-16;
    int hi = k - 1, iter = 0, total = 0;
    while (hi > 0) {
        int lo = hi;
        while (lo > 0 && std::abs(h(lo, lo - 1)) > eps * (std::abs(h(lo - 1, lo - 1)) + std::abs(h(lo, lo)))) lo--;
        if (lo > 0) h(lo, lo - 1) = 0.0;
        if (lo == hi) {
            hi--;
            iter = 0;
            continue;
        }
        if (++total > 60 * k) return false;
        iter++;
        cd mu;
        if (iter % 11 == 10) {
            mu = h(hi, hi) + std::abs(h(hi, hi - 1));  // an exceptional shift
        } else {
            const cd a = h(hi - 1, hi - 1), b = h(hi - 1, hi), c = h(hi, hi - 1), d = h(hi, hi);
            const cd half = 0.5 * (a + d), disc = std::sqrt(0.25 * (a - d) * (a - d) + b * c);
            const cd m1 = half + disc, m2 = half - disc;
            mu = std::abs(m1 - d) < std::abs(m2 - d) ? m1 : m2;
        }
        cd x = h(lo, lo) - mu, y = h(lo + 1, lo);
        for (int m = lo; m < hi; m++) {
            // G [x; y] = [r; 0], G = [c s; -conj(s) c], c real
            double c;
            cd s;
            const double ax = std::abs(x), ay = std::abs(y);
            if (ay == 0.0) {
                c = 1.0;
                s = 0.0;
            } else if (ax == 0.0) {
                c = 0.0;
                s = std::conj(y) / ay;
            } else {
                const double norm = std::hypot(ax, ay);
                c = ax / norm;
                s = (x / ax) * std::conj(y) / norm;
            }
            for (int j = std::max(lo, m - 1); j < k; j++) {
                const cd h1 = h(m, j), h2 = h(m + 1, j);
                h(m, j) = c * h1 + s * h2;
                h(m + 1, j) = -std::conj(s) * h1 + c * h2;
            }
            const int rend = std::min(hi, m + 2);
            for (int i = 0; i <= rend; i++) {
                const cd h1 = h(i, m), h2 = h(i, m + 1);
                h(i, m) = h1 * c + h2 * std::conj(s);
                h(i, m + 1) = -h1 * s + h2 * c;
            }
            for (int i = 0; i < k; i++) {
                const cd q1 = Q[(size_t)i * k + m], q2 = Q[(size_t)i * k + m + 1];
                Q[(size_t)i * k + m] = q1 * c + q2 * std::conj(s);
                Q[(size_t)i * k + m + 1] = -q1 * s + q2 * c;
            }
            if (m < hi - 1) {
                x = h(m + 1, m);
                y = h(m + 2, m);
            }
        }
    }
    return true;
}

// The eigenvector of the upper triangular T (row-major k x k) for its i-th diagonal value
inline std::vector<cd> triangular_vector(const std::vector<cd> &T, int k, int i) {
    std::vector<cd> z((size_t)k, cd(0.0, 0.0));
    z[i] = 1.0;
    const cd t = T[(size_t)i * k + i];
    double scale = 0.0;
    for (int j = 0; j < k; j++) scale = std::max(scale, std::abs(T[(size_t)j * k + j]));
    const double small = 1e-14 * std::max(scale, 1e-300);
    for (int j = i - 1; j >= 0; j--) {
        cd s = 0.0;
        for (int q = j + 1; q <= i; q++) s += T[(size_t)j * k + q] * z[q];
        cd d = T[(size_t)j * k + j] - t;
        if (std::abs(d) < small) d = small;
        z[j] = -s / d;
    }
    return z;
}

struct Pair {
    cd value;
    std::vector<cd> vector;
};

// The `count` eigenvalues of the n x n matrix (entries summed where repeated) nearest `shift`,
// nearest first, with unit eigenvectors: each accepted when its residual in the shifted inverse,
// as the Arnoldi decomposition gives it, is below `tolerance` of its Ritz value. An empty
// result and a message in *why if the shifted matrix is singular or the pairs do not converge.
inline std::vector<Pair> nearest(long long n, const std::vector<std::tuple<long long, long long, cd>> &entries,
                                 cd shift, int count, double tolerance, std::string *why) {
    std::vector<Pair> out;
    if (n <= 0 || count <= 0) return out;
    count = (int)std::min<long long>(count, n);
    int kl = 0, ku = 0;
    std::vector<std::tuple<long long, long long, cd>> shifted(entries);
    for (const auto &e : entries) {
        const long long d = std::get<0>(e) - std::get<1>(e);
        kl = std::max<long long>(kl, d);
        ku = std::max<long long>(ku, -d);
    }
    for (long long i = 0; i < n; i++) shifted.emplace_back(i, i, -shift);
    // the band's factorization takes about n kl (kl + ku) complex products and n (2 kl + ku)
    // values: past about a minute of host work (1e11 products; 5e10 take 32 s on 8 threads) or
    // 4 GB, refuse rather than grind
    if ((double)n * kl * (double)(kl + ku) > 1e11 || (double)n * (2.0 * kl + ku + 1) * 16.0 > 4e9) {
        *why = "the plane's problem is too large for the banded solve (n = " + std::to_string(n) + ", band " +
               std::to_string(kl) + "): give a window around the guide";
        return out;
    }
    BandLU lu;
    if (!lu.factor(n, kl, ku, shifted)) {
        *why = "the shifted matrix is singular";
        return out;
    }
    const int cap = (int)std::min<long long>(n, 400);
    std::vector<std::vector<cd>> V;
    std::vector<cd> Hm((size_t)(cap + 1) * cap, cd(0.0, 0.0));  // (cap + 1) x cap, row-major
    auto H = [&](int i, int j) -> cd & { return Hm[(size_t)i * cap + j]; };
    std::vector<cd> v((size_t)n);
    // a fixed start with every value in play
    double nrm = 0.0;
    for (long long i = 0; i < n; i++) {
        v[i] = cd(1.0 + 0.37 * std::sin(1.3 * (double)i), 0.21 * std::cos(0.7 * (double)i));
        nrm += std::norm(v[i]);
    }
    nrm = std::sqrt(nrm);
    for (auto &x : v) x /= nrm;
    V.push_back(v);
    int k = 0;
    bool breakdown = false;
    for (int j = 0; j < cap; j++) {
        std::vector<cd> w = V[j];
        lu.solve(w);
        for (int pass = 0; pass < 2; pass++)
            for (int i = 0; i <= j; i++) {
                const double *vi = reinterpret_cast<const double *>(V[i].data());
                double *wv = reinterpret_cast<double *>(w.data());
                double dr = 0.0, di = 0.0;  // conj(v) . w
                for (long long q = 0; q < n; q++) {
                    dr += vi[2 * q] * wv[2 * q] + vi[2 * q + 1] * wv[2 * q + 1];
                    di += vi[2 * q] * wv[2 * q + 1] - vi[2 * q + 1] * wv[2 * q];
                }
                H(i, j) += cd(dr, di);
                for (long long q = 0; q < n; q++) {
                    const double ar = vi[2 * q], ai = vi[2 * q + 1];
                    wv[2 * q] -= dr * ar - di * ai;
                    wv[2 * q + 1] -= dr * ai + di * ar;
                }
            }
        double wn = 0.0;
        for (const cd &x : w) wn += std::norm(x);
        wn = std::sqrt(wn);
        H(j + 1, j) = wn;
        k = j + 1;
        double hscale = 0.0;
        for (int i = 0; i <= j; i++) hscale = std::max(hscale, std::abs(H(i, j)));
        breakdown = wn <= 1e-14 * std::max(hscale, 1e-300);
        const bool check = breakdown || k == cap || (k >= count + 8 && (k - count) % 8 == 0);
        if (check) {
            std::vector<cd> T((size_t)k * k), Q;
            for (int r = 0; r < k; r++)
                for (int c = 0; c < k; c++) T[(size_t)r * k + c] = H(r, c);
            if (!hessenberg_schur(T, Q, k)) {
                *why = "the projected eigenproblem did not converge";
                return {};
            }
            std::vector<int> order(k);
            for (int i = 0; i < k; i++) order[i] = i;
            std::sort(order.begin(), order.end(), [&](int p, int q) {
                return std::abs(T[(size_t)p * k + p]) > std::abs(T[(size_t)q * k + q]);
            });
            bool done = true;
            std::vector<Pair> found;
            for (int t = 0; t < count; t++) {
                const int i = order[t];
                const cd theta = T[(size_t)i * k + i];
                const std::vector<cd> z = triangular_vector(T, k, i);
                std::vector<cd> y((size_t)k, cd(0.0, 0.0));
                for (int r = 0; r < k; r++)
                    for (int c = 0; c < k; c++) y[r] += Q[(size_t)r * k + c] * z[c];
                double yn = 0.0;
                for (const cd &x : y) yn += std::norm(x);
                yn = std::sqrt(yn);
                const double residual = wn * std::abs(y[k - 1]) / yn;
                if (!breakdown && !(residual <= tolerance * std::abs(theta))) done = false;
                Pair p;
                p.value = shift + 1.0 / theta;
                p.vector.assign((size_t)n, cd(0.0, 0.0));
                for (int c = 0; c < k; c++) {
                    const cd yc = y[c] / yn;
                    for (long long q = 0; q < n; q++) p.vector[q] += yc * V[c][q];
                }
                double pn = 0.0;
                for (const cd &x : p.vector) pn += std::norm(x);
                pn = std::sqrt(pn);
                for (cd &x : p.vector) x /= pn;
                found.push_back(std::move(p));
            }
            if (done) {
                // one step of inverse iteration at each pair's own eigenvalue: the Arnoldi
                // residual bounds A's only to |A - shift| times the tolerance
                for (Pair &p : found) {
                    std::vector<std::tuple<long long, long long, cd>> near(entries);
                    for (long long i = 0; i < n; i++) near.emplace_back(i, i, -p.value);
                    BandLU polish;
                    if (!polish.factor(n, kl, ku, near)) continue;  // exactly singular: x is exact
                    std::vector<cd> x = p.vector;
                    polish.solve(x);
                    double xn = 0.0;
                    for (const cd &v : x) xn += std::norm(v);
                    xn = std::sqrt(xn);
                    if (!(xn > 0.0) || !std::isfinite(xn)) continue;
                    for (long long q = 0; q < n; q++) p.vector[q] = x[q] / xn;
                    // the eigenvalue from the polished vector: x^H A x (x of unit length)
                    std::vector<cd> ax((size_t)n, cd(0.0, 0.0));
                    for (const auto &e : entries)
                        ax[std::get<0>(e)] += std::get<2>(e) * p.vector[std::get<1>(e)];
                    cd rq = 0.0;
                    for (long long q = 0; q < n; q++) rq += std::conj(p.vector[q]) * ax[q];
                    p.value = rq;
                }
                return found;
            }
            if (breakdown || k == cap) {
                *why = "the eigenpairs did not converge";
                return {};
            }
        }
        for (cd &x : w) x /= wn;
        V.push_back(std::move(w));
    }
    *why = "the eigenpairs did not converge";
    return {};
}

}  // namespace pxeig
