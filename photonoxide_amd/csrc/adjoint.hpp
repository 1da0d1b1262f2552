// The reverse-time adjoint of the fused Yee + CPML step (pxfdtd_adjoint_*).
//
// photonoxide's own adjoint is frequency-domain, from two forward runs' transforms, with no
// checkpointing (src/fdtd/adjoint.rs:43-46, 496-569); this is the exact discrete transpose of
// the time stepping instead. One step maps (E, H, psi_E, psi_H) to new values by four linear
// stages (the reference's order, src/fdtd/kernel/reference.rs:14-180):
//   F1  psi_H[c,w] = bH psi_H[c,w] + aH D+_w(E_o) / step_w
//   F2  H_c -= dt (sum_w s'_cw fh_w D+_w(E_o) + sum_w s'_cw psi_H[c,w])
//   F3  psi_E[c,w] = bE psi_E[c,w] + aE D-_w(H_o) / step_w
//   F4  E_c = ca_c E_c + cb_c (sum_w s_cw fe_w D-_w(H_o) + sum_w s_cw psi_E[c,w])
// (w the two axes other than c, o the field component its difference reads, s = +1 for the
// first of c's others and -1 for the second; D+ a forward difference, zero past the high wall,
// D- a backward one, zero before the low wall). Its transpose runs the stages backwards on the
// adjoint variables (written with a bar in the comments):
//   R1a psi~E = psibar_E[c,w] + s_cw cb_c Ebar_c;  q = aE psi~E / step_w;  psibar_E = bE psi~E
//   R1b Hbar_o += -s_cw D+_w(cb_c fe_w Ebar_c) - D+_w(q)          (the transpose of D- is -D+)
//   R2a psi~H = psibar_H[c,w] - dt s'_cw Hbar_c;  q' = aH psi~H / step_w;  psibar_H = bH psi~H
//   R2b Ebar_o = ca_o Ebar_o + dt s'_cw D-_w(fh_w Hbar_c) - D-_w(q')
// with the same walls and periodic wraps as the forward. The gradient with respect to cb of a
// value is the sum over steps of its Ebar after the step times the term cb multiplies in F4
// (recomputed from checkpoints), and the objective's derivative enters Ebar after each step at
// the monitor's values.
#pragma once
#include "kernels.hpp"

namespace pxfdtd {

struct AdjGrid {
    int nx, ny, nz, pitch;
    long long plane, nval;
    int per[3];
    int low[3], high[3];
    int psi_off[2][3][3][2];  // [field][component][axis][side], as the forward's
    int cpml_off[3][2][2];    // b of [axis][side][field] in coef, a after count values
    int fe_off[3], fh_off[3];
};

template <typename T>
struct AdjArgs {
    AdjGrid g;
    T dt;
    T step[3];
    const T *coef;  // the forward's factors and CPML b, a
    const T *ca, *cb;  // [3][nval], pitched
    T *eb, *hb;        // Ebar, Hbar [3][nval]
    T *pb;             // psibar, the forward's psi layout
    T *q;              // R1a / R2a's q, the same layout
};

__device__ __forceinline__ int adj_wrap(int m, int n, int per) {
    if (m >= 0 && m < n) return m;
    if (!per) return -1;
    return m < 0 ? m + n : m - n;
}

// the index of q (field F, component c, axis w) at cell (i, j, k), or -1 outside its slab
template <int F, typename T>
__device__ __forceinline__ long long adj_qidx(const AdjArgs<T> &A, int c, int w, int i, int j, int k) {
    const AdjGrid &g = A.g;
    const int n[3] = {g.nx, g.ny, g.nz};
    const int m[3] = {i, j, k};
    if (i < 0 || j < 0 || k < 0) return -1;
    const int side = side_of(m[w], n[w], g.low[w], g.high[w]);
    if (side < 0) return -1;
    const int base = g.psi_off[F][c][w][side];
    if (base < 0) return -1;
    const int cnt = side == 1 ? g.high[w] : g.low[w];
    const int s = m[w] - (side == 1 ? n[w] - g.high[w] : 0);
    long long q;
    if (w == 0)
        q = ((long long)k * g.ny + j) * cnt + s;
    else if (w == 1)
        q = ((long long)k * cnt + s) * g.nx + i;
    else
        q = ((long long)s * g.ny + j) * g.nx + i;
    return base + q;
}

// R1a and R2a: one thread per psi value of the slabs of field F (table of slabs by offset)
struct AdjSlab {
    int field, comp, axis, side, from, count;
    long long offset, len;
};

template <typename T, int F>
__global__ void adj_psi(AdjArgs<T> A, const AdjSlab *slabs, int nslabs, long long total) {
    const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= total) return;
    int sl = -1;
    for (int q = 0; q < nslabs; q++)
        if (v >= slabs[q].offset && v < slabs[q].offset + slabs[q].len) sl = q;
    if (sl < 0 || slabs[sl].field != F) return;
    const AdjSlab S = slabs[sl];
    const AdjGrid &g = A.g;
    const long long lv = v - S.offset;
    // the slab's box: the grid with axis w reduced to count, i fastest
    const int bn[3] = {S.axis == 0 ? S.count : g.nx, S.axis == 1 ? S.count : g.ny, S.axis == 2 ? S.count : g.nz};
    int m[3];
    m[0] = (int)(lv % bn[0]);
    m[1] = (int)((lv / bn[0]) % bn[1]);
    m[2] = (int)(lv / ((long long)bn[0] * bn[1]));
    const int s = m[S.axis];
    m[S.axis] += S.from;
    const long long cell = (long long)m[2] * g.plane + (long long)m[1] * g.pitch + m[0];
    const int c = S.comp, w = S.axis;
    const int a = (c + 1) % 3;
    const T sign = w == a ? T(1) : T(-1);
    const int co = g.cpml_off[w][S.side][F];
    const T b = A.coef[co + s], aa = A.coef[co + S.count + s];
    T tot;
    if (F == 0)
        tot = A.pb[v] + sign * A.cb[c * g.nval + cell] * A.eb[c * g.nval + cell];
    else
        tot = A.pb[v] - A.dt * sign * A.hb[c * g.nval + cell];
    A.q[v] = aa * tot / A.step[w];
    A.pb[v] = b * tot;
}

// R1b: Hbar += the transposes of F4's differences and psi_E's, one thread per cell
template <typename T>
__global__ void adj_h(AdjArgs<T> A) {
    const AdjGrid &g = A.g;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.plane * g.nz) return;
    const int i = (int)(t % g.pitch), j = (int)((t / g.pitch) % g.ny), k = (int)(t / g.plane);
    if (i >= g.nx) return;
    const int m[3] = {i, j, k};
    const int n[3] = {g.nx, g.ny, g.nz};
    // g_cw at a cell: cb_c fe_w Ebar_c, zero past the high wall
    auto gv = [&](int c, int w, int d) -> T {  // at the cell + d e_w (d = 0 or 1)
        int mm[3] = {m[0], m[1], m[2]};
        mm[w] = adj_wrap(mm[w] + d, n[w], g.per[w]);
        if (mm[w] < 0) return T(0);
        const long long cell = (long long)mm[2] * g.plane + (long long)mm[1] * g.pitch + mm[0];
        return A.cb[c * g.nval + cell] * A.coef[g.fe_off[w] + mm[w]] * A.eb[c * g.nval + cell];
    };
    auto qv = [&](int c, int w, int d) -> T {
        int mm[3] = {m[0], m[1], m[2]};
        mm[w] = adj_wrap(mm[w] + d, n[w], g.per[w]);
        if (mm[w] < 0) return T(0);
        const long long q = adj_qidx<0>(A, c, w, mm[0], mm[1], mm[2]);
        return q >= 0 ? A.q[q] : T(0);
    };
    const long long own = (long long)k * g.plane + (long long)j * g.pitch + i;
    // Hbar_o for o = x, y, z: the E components and axes that read H_o (c's others (a, b):
    // E_c reads H_b along a with +1 and H_a along b with -1)
    for (int o = 0; o < 3; o++) {
        T acc = T(0);
        for (int c = 0; c < 3; c++) {
            if (c == o) continue;
            const int a = (c + 1) % 3, b = (c + 2) % 3;
            // E_c reads H_o along w: o == b -> w = a, sign +1; o == a -> w = b, sign -1
            const int w = o == b ? a : b;
            const T s = o == b ? T(1) : T(-1);
            acc = acc - s * (gv(c, w, 1) - gv(c, w, 0));
            acc = acc - (qv(c, w, 1) - qv(c, w, 0));
        }
        A.hb[o * g.nval + own] += acc;
    }
}

// R2b: Ebar = ca Ebar + the transposes of F2's differences and psi_H's, one thread per cell
template <typename T>
__global__ void adj_e(AdjArgs<T> A) {
    const AdjGrid &g = A.g;
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.plane * g.nz) return;
    const int i = (int)(t % g.pitch), j = (int)((t / g.pitch) % g.ny), k = (int)(t / g.plane);
    if (i >= g.nx) return;
    const int m[3] = {i, j, k};
    const int n[3] = {g.nx, g.ny, g.nz};
    // h_cw at a cell: fh_w Hbar_c, zero before the low wall
    auto hv = [&](int c, int w, int d) -> T {  // at the cell - d e_w
        int mm[3] = {m[0], m[1], m[2]};
        mm[w] = adj_wrap(mm[w] - d, n[w], g.per[w]);
        if (mm[w] < 0) return T(0);
        const long long cell = (long long)mm[2] * g.plane + (long long)mm[1] * g.pitch + mm[0];
        return A.coef[g.fh_off[w] + mm[w]] * A.hb[c * g.nval + cell];
    };
    auto qv = [&](int c, int w, int d) -> T {
        int mm[3] = {m[0], m[1], m[2]};
        mm[w] = adj_wrap(mm[w] - d, n[w], g.per[w]);
        if (mm[w] < 0) return T(0);
        const long long q = adj_qidx<1>(A, c, w, mm[0], mm[1], mm[2]);
        return q >= 0 ? A.q[q] : T(0);
    };
    const long long own = (long long)k * g.plane + (long long)j * g.pitch + i;
    for (int o = 0; o < 3; o++) {
        T acc = A.ca[o * g.nval + own] * A.eb[o * g.nval + own];
        for (int c = 0; c < 3; c++) {
            if (c == o) continue;
            const int a = (c + 1) % 3, b = (c + 2) % 3;
            // H_c reads E_b along a with +1 and E_a along b with -1
            const int w = o == b ? a : b;
            const T s = o == b ? T(1) : T(-1);
            acc = acc + A.dt * s * (hv(c, w, 0) - hv(c, w, 1));
            acc = acc - (qv(c, w, 0) - qv(c, w, 1));
        }
        A.eb[o * g.nval + own] = acc;
    }
}

// the objective's derivative at the monitor's values: Ebar += 2 Re(w p), w = conj(S) m
template <typename T>
__global__ void adj_seed(T *eb, const int64_t *index, const double *w, long long n, double pr, double pi) {
    const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    eb[index[v]] += (T)(2.0 * (w[2 * v] * pr - w[2 * v + 1] * pi));
}

// grad[d] += Ebar at design value d times the term cb multiplied there, in FP64
template <typename T>
__global__ void adj_grad(const T *eb, const int64_t *index, const T *term, long long n, double *grad) {
    const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    grad[v] += (double)eb[index[v]] * (double)term[v];
}

// the term cb multiplies in F4 at each design value, from the step's new H (no psi and no
// current there: the design lies outside the CPML slabs and the sources)
template <typename T>
__global__ void adj_term(AdjGrid g, const T *coef, const T *h, const int64_t *index, const int8_t *comp,
                         long long n, T *term) {
    const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const long long idx = index[v];
    const int c = comp[v];
    const int i = (int)(idx % g.pitch), j = (int)((idx / g.pitch) % g.ny), k = (int)(idx / g.plane);
    const int m[3] = {i, j, k};
    const int n3[3] = {g.nx, g.ny, g.nz};
    const int a = (c + 1) % 3, b = (c + 2) % 3;
    auto back = [&](int comp_h, int w) -> T {  // H_comp(cell) - H_comp(cell - e_w), zero before a wall
        int mm[3] = {m[0], m[1], m[2]};
        mm[w] = adj_wrap(mm[w] - 1, n3[w], g.per[w]);
        const T here = h[comp_h * g.nval + idx];
        if (mm[w] < 0) return here;
        const long long cell = (long long)mm[2] * g.plane + (long long)mm[1] * g.pitch + mm[0];
        return here - h[comp_h * g.nval + cell];
    };
    // curl_c = D-_a(H_b) fe_a - D-_b(H_a) fe_b, as F4 computes it
    const T da = back(b, a), db = back(a, b);
    term[v] = da * coef[g.fe_off[a] + m[a]] - db * coef[g.fe_off[b] + m[b]];
}

}  // namespace pxfdtd
