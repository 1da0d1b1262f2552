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
    T step[3];  // f64: dx, dy, dz; f32: their reciprocals (see per_step)
    const T *coef;  // the forward's factors and CPML b, a
    const T *ca, *cb;  // [3][nval], pitched
    T *eb, *hb;        // Ebar, Hbar [3][nval]
    const T *pb_in;    // psibar before the step, the forward's psi layout
    T *pb_out;         // psibar after it (the other of two buffers)
};

__device__ __forceinline__ int adj_wrap(int m, int n, int per) {
    if (m >= 0 && m < n) return m;
    if (!per) return -1;
    return m < 0 ? m + n : m - n;
}

// One of two launch-wide values by a lane's side. Both are taken into scalar registers first:
// a plain select of two argument fields compiles into a per-lane load from the argument segment
// at a selected address, which adds a memory round trip in front of the load it indexes.
__device__ __forceinline__ int adj_by_side(int side, int low, int high) {
    const int l = __builtin_amdgcn_readfirstlane(low), h = __builtin_amdgcn_readfirstlane(high);
    return side == 1 ? h : l;
}

// Where a point sits among axis w's CPML slabs: the side (-1 outside both, or past a wall), its
// coordinate within the slab, and its index within one component's slab (the forward's layout:
// along x [k][j][s], along y [k][s][i], along z [s][j][i]). m is the point, m[w] < 0 past a wall.
struct AdjPt {
    int side, s, q;
};

__device__ __forceinline__ AdjPt adj_pt(const AdjGrid &g, int w, const int m[3]) {
    AdjPt p;
    const int n = w == 0 ? g.nx : (w == 1 ? g.ny : g.nz);
    p.side = m[w] < 0 ? -1 : side_of(m[w], n, g.low[w], g.high[w]);
    const int cnt = adj_by_side(p.side, g.low[w], g.high[w]);
    p.s = m[w] - (p.side == 1 ? n - g.high[w] : 0);
    if (w == 0)
        p.q = (m[2] * g.ny + m[1]) * cnt + p.s;
    else if (w == 1)
        p.q = (m[2] * cnt + p.s) * g.nx + m[0];
    else
        p.q = (p.s * g.ny + m[1]) * g.nx + m[0];
    return p;
}

// The psibar index of component c's value along axis w at point p, -1 where it has none.
template <int F>
__device__ __forceinline__ int adj_pidx(const AdjGrid &g, int c, int w, const AdjPt &p) {
    if (p.side < 0) return -1;
    const int base = adj_by_side(p.side, g.psi_off[F][c][w][0], g.psi_off[F][c][w][1]);
    return base < 0 ? -1 : base + p.q;
}

// Axis w's CPML b and a of field F at point p (the table's first entry, unused, outside the
// slabs, so that the load needs no branch).
template <int F, typename T>
__device__ __forceinline__ void adj_ba(const AdjArgs<T> &A, int w, const AdjPt &p, T &b, T &a) {
    const int co = p.side < 0 ? 0 : adj_by_side(p.side, A.g.cpml_off[w][0][F], A.g.cpml_off[w][1][F]) + p.s;
    const int cnt = p.side < 0 ? 0 : adj_by_side(p.side, A.g.low[w], A.g.high[w]);
    b = A.coef[co];
    a = A.coef[co + cnt];
}

// The launch shape of adj_h and adj_e: 64 x 4 threads over (i, j), one block row per k.
constexpr int ADJ_BX = 64, ADJ_BY = 4;

// R1a + R1b: Hbar += the transposes of F4's differences and psi_E's, one thread per cell. The
// thread updates its own cell's psibar_E values (R1a: psi~ = psibar + s cb Ebar, psibar <- b
// psi~ into the step's output buffer, q = a psi~ / step_w) and recomputes the q of the cell +
// e_w from the step's input buffer, so no other thread's output is read. Every load is issued
// before the arithmetic, with indices that stay in bounds where a term is zero, so that a wave
// waits on memory once rather than once per branch.
template <typename T>
__global__ __launch_bounds__(ADJ_BX *ADJ_BY) void adj_h(AdjArgs<T> A) {
    const AdjGrid &g = A.g;
    const int i = blockIdx.x * ADJ_BX + threadIdx.x, j = blockIdx.y * ADJ_BY + threadIdx.y, k = blockIdx.z;
    if (i >= g.nx || j >= g.ny) return;
    const int m[3] = {i, j, k};
    const int n[3] = {g.nx, g.ny, g.nz};
    const long long own = (long long)k * g.plane + (long long)j * g.pitch + i;
    const long long stride[3] = {1, g.pitch, g.plane};
    // the cell + e_w: its coordinate along w (-1 past a high wall, where its loads read the cell
    // itself and its terms are zero), index and slab positions
    int up[3], upc[3];
    long long nb[3];
    AdjPt p[2][3];
#pragma unroll
    for (int w = 0; w < 3; w++) {
        up[w] = adj_wrap(m[w] + 1, n[w], g.per[w]);
        upc[w] = up[w] < 0 ? m[w] : up[w];
        nb[w] = own + (long long)(upc[w] - m[w]) * stride[w];
        int mm[3] = {i, j, k};
        p[0][w] = adj_pt(g, w, mm);
        mm[w] = up[w];
        p[1][w] = adj_pt(g, w, mm);
    }
    // E_c reads H_b along a (t = 0, s = +1) and H_a along b (t = 1, s = -1): w = (c + 1 + t) % 3
    T cb0[3], eb0[3], cb1[3][2], eb1[3][2], pb[2][3][2], fe[2][3], bq[2][3], aq[2][3];
    int pi[2][3][2];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        cb0[c] = A.cb[c * g.nval + own];
        eb0[c] = A.eb[c * g.nval + own];
#pragma unroll
        for (int t = 0; t < 2; t++) {
            const int w = (c + 1 + t) % 3;
            cb1[c][t] = A.cb[c * g.nval + nb[w]];
            eb1[c][t] = A.eb[c * g.nval + nb[w]];
#pragma unroll
            for (int d = 0; d < 2; d++) {
                pi[d][c][t] = adj_pidx<0>(g, c, w, p[d][w]);
                pb[d][c][t] = A.pb_in[pi[d][c][t] < 0 ? 0 : pi[d][c][t]];
            }
        }
    }
#pragma unroll
    for (int w = 0; w < 3; w++) {
        fe[0][w] = A.coef[g.fe_off[w] + m[w]];
        fe[1][w] = A.coef[g.fe_off[w] + upc[w]];
        adj_ba<0>(A, w, p[0][w], bq[0][w], aq[0][w]);
        adj_ba<0>(A, w, p[1][w], bq[1][w], aq[1][w]);
    }
    // g_cw = cb_c fe_w Ebar_c; Hbar_o gets -s (g_cw(+e_w) - g_cw) - (q_cw(+e_w) - q_cw)
    T acc[3] = {T(0), T(0), T(0)};
#pragma unroll
    for (int c = 0; c < 3; c++) {
#pragma unroll
        for (int t = 0; t < 2; t++) {
            const int w = (c + 1 + t) % 3, o = (c + 2 - t) % 3;
            const T s = t == 0 ? T(1) : T(-1);
            const T g1 = up[w] >= 0 ? cb1[c][t] * fe[1][w] * eb1[c][t] : T(0);
            T q1 = T(0);
            if (pi[1][c][t] >= 0) q1 = per_step(aq[1][w] * (pb[1][c][t] + s * cb1[c][t] * eb1[c][t]), A.step[w]);
            const T g0 = cb0[c] * fe[0][w] * eb0[c];
            T q0 = T(0);
            if (pi[0][c][t] >= 0) {
                const T tot = pb[0][c][t] + s * cb0[c] * eb0[c];
                A.pb_out[pi[0][c][t]] = bq[0][w] * tot;
                q0 = per_step(aq[0][w] * tot, A.step[w]);
            }
            acc[o] = acc[o] - s * (g1 - g0);
            acc[o] = acc[o] - (q1 - q0);
        }
    }
#pragma unroll
    for (int o = 0; o < 3; o++) A.hb[o * g.nval + own] += acc[o];
}

// R2a + R2b: Ebar = ca Ebar + the transposes of F2's differences and psi_H's, one thread per
// cell; psibar_H as psibar_E in adj_h (R2a: psi~ = psibar - dt s Hbar), with the cell - e_w
// for the neighbour.
template <typename T>
__global__ __launch_bounds__(ADJ_BX *ADJ_BY) void adj_e(AdjArgs<T> A) {
    const AdjGrid &g = A.g;
    const int i = blockIdx.x * ADJ_BX + threadIdx.x, j = blockIdx.y * ADJ_BY + threadIdx.y, k = blockIdx.z;
    if (i >= g.nx || j >= g.ny) return;
    const int m[3] = {i, j, k};
    const int n[3] = {g.nx, g.ny, g.nz};
    const long long own = (long long)k * g.plane + (long long)j * g.pitch + i;
    const long long stride[3] = {1, g.pitch, g.plane};
    int down[3], downc[3];
    long long nb[3];
    AdjPt p[2][3];
#pragma unroll
    for (int w = 0; w < 3; w++) {
        down[w] = adj_wrap(m[w] - 1, n[w], g.per[w]);
        downc[w] = down[w] < 0 ? m[w] : down[w];
        nb[w] = own + (long long)(downc[w] - m[w]) * stride[w];
        int mm[3] = {i, j, k};
        p[0][w] = adj_pt(g, w, mm);
        mm[w] = down[w];
        p[1][w] = adj_pt(g, w, mm);
    }
    // H_c reads E_b along a (t = 0, s = +1) and E_a along b (t = 1, s = -1)
    T acc[3], hb0[3], hb1[3][2], pb[2][3][2], fh[2][3], bq[2][3], aq[2][3];
    int pi[2][3][2];
#pragma unroll
    for (int o = 0; o < 3; o++) acc[o] = A.ca[o * g.nval + own] * A.eb[o * g.nval + own];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        hb0[c] = A.hb[c * g.nval + own];
#pragma unroll
        for (int t = 0; t < 2; t++) {
            const int w = (c + 1 + t) % 3;
            hb1[c][t] = A.hb[c * g.nval + nb[w]];
#pragma unroll
            for (int d = 0; d < 2; d++) {
                pi[d][c][t] = adj_pidx<1>(g, c, w, p[d][w]);
                pb[d][c][t] = A.pb_in[pi[d][c][t] < 0 ? 0 : pi[d][c][t]];
            }
        }
    }
#pragma unroll
    for (int w = 0; w < 3; w++) {
        fh[0][w] = A.coef[g.fh_off[w] + m[w]];
        fh[1][w] = A.coef[g.fh_off[w] + downc[w]];
        adj_ba<1>(A, w, p[0][w], bq[0][w], aq[0][w]);
        adj_ba<1>(A, w, p[1][w], bq[1][w], aq[1][w]);
    }
    // h_cw = fh_w Hbar_c, zero before a low wall; Ebar_o gets dt s (h_cw - h_cw(-e_w)) -
    // (q_cw - q_cw(-e_w))
#pragma unroll
    for (int c = 0; c < 3; c++) {
#pragma unroll
        for (int t = 0; t < 2; t++) {
            const int w = (c + 1 + t) % 3, o = (c + 2 - t) % 3;
            const T s = t == 0 ? T(1) : T(-1);
            const T h0 = fh[0][w] * hb0[c];
            T q0 = T(0);
            if (pi[0][c][t] >= 0) {
                const T tot = pb[0][c][t] - A.dt * s * hb0[c];
                A.pb_out[pi[0][c][t]] = bq[0][w] * tot;
                q0 = per_step(aq[0][w] * tot, A.step[w]);
            }
            const T h1 = down[w] >= 0 ? fh[1][w] * hb1[c][t] : T(0);
            T q1 = T(0);
            if (pi[1][c][t] >= 0) q1 = per_step(aq[1][w] * (pb[1][c][t] - A.dt * s * hb1[c][t]), A.step[w]);
            acc[o] = acc[o] + A.dt * s * (h0 - h1);
            acc[o] = acc[o] - (q0 - q1);
        }
    }
#pragma unroll
    for (int o = 0; o < 3; o++) A.eb[o * g.nval + own] = acc[o];
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
