// The 2D march: the fused Yee + CPML step of a grid one cell thick along z with z periodic
// (photonoxide's 2D simulations, src/fdtd.rs:680-796 with nz = 1).
//
// E and H at z + 1 and z - 1 are then the values themselves, so every difference along z is
// an exact zero and the update splits into two independent sets: TE (Ex, Ey, Hz) and TM (Ez,
// Hx, Hy). The variant SET steps TE (1), TM (2) or both (3); a set whose fields, psi and
// sources are all zero stays zero, so stepping only the other one gives the same bits as the
// 3D update. The zero terms are kept where they decide a sign (0 - x is not -x at x = 0), so
// each value is the 3D kernel's bits.
//
// Each wavefront owns a strip of 64 x-values (lane L at x = i0 + L, i0 a multiple of 64) and
// marches along y over a chunk of rows, keeping the row below's new H in registers (E's back
// difference along y) and loading E one row ahead. The columns just outside the strip come from
// a two-lane load per array: lane 0 holds x = i0 - 1 and recomputes H there (its own i - 1
// neighbour for E), lane 63 holds E at x = i0 + 64 (its i + 1 neighbour for H); DPP wave shifts
// move the rest between lanes. The injection lists are per strip and row.
#pragma once
#include "march64.hpp"

namespace pxfdtd {

// the injections of row j on the value (field, i) in `mask`, in the strip's list order
template <typename T, int FIELD>
__device__ __forceinline__ void inject2d(const StepArgs<T> &A, const StepCold<T> &K, int strip, int i,
                                         int j, int mask, T &v0, T &v1, T &v2, T cb0, T cb1, T cb2) {
    const int32_t *off = K.inj_off + (long long)strip * (A.ny + 1);
    const int lo = off[j], hi = off[j + 1];
    for (int q = lo; q < hi; q++) {
        const Injection in = K.inj[q];
        const int c = in.code & 3;
        if ((in.code >> 2) != FIELD || in.i != i || !((mask >> c) & 1)) continue;
        apply_injection<T, FIELD>(A, K, in, c, v0, v1, v2, cb0, cb1, cb2, A.wave_step);
    }
}

// One wavefront's unit: strip unit.x, rows [unit.z, unit.w).
template <typename T, int MODE, int AX, int SET>
__device__ __forceinline__ void march2d_unit(const StepArgs<T> &A, const int4 unit,
                                             const MarchShared<T> &S, const T *tab,
                                             const int8_t *grp) {
    constexpr bool SLABS = AX != 0, TE = SET & 1, TM = SET & 2;
    const StepCold<T> &K = S.K;
    const T *cp = S.cp;
    const int lane = threadIdx.x & 63;
    const int nx = A.nx, ny = A.ny, pitch = A.pitch;
    const int cbase = 2 * (nx + ny + 1);
    const int bx = unit.x, j0 = unit.z, j1 = unit.w;
    const int i0 = bx * SW;
    const int perx = A.flags & 1, pery = (A.flags >> 1) & 1;
    bool injects = false;
    if (A.flags & 16) {
        const int32_t *io = K.inj_off + (long long)bx * (ny + 1);
        injects = io[ny] > io[0];
    }
    const int i = i0 + lane;
    const bool own_x = i < nx;
    const int xi = own_x ? i : (i == nx && perx ? 0 : -1);
    const int xl = wrap_any(i0 - 1, nx, perx);
    const int xr = i0 + SW < nx ? i0 + SW : (i0 + SW == nx && perx ? 0 : -1);
    // the column this lane holds at the edges: lane 0 the left one, lane 63 the right one
    const int xe = lane == 0 ? xl : (lane == 63 ? xr : -1);
    const uint32_t cbytes = (uint32_t)A.nval * sizeof(T);
    const __amdgpu_buffer_rsrc_t rE = rsrc(A.ein, 3 * cbytes), rH = rsrc(A.hin, 3 * cbytes),
                                 wE = rsrc(A.eout, 3 * cbytes), wH = rsrc(A.hout, 3 * cbytes),
                                 rid = rsrc(A.ids, MODE == 0 ? (uint32_t)A.nval : 0u);
    const T dt = A.dt;
    const T stx = A.step[0], sty = A.step[1];
    const T *coef = A.coef;
    const int fex = 0, fhx = nx, fey = 2 * nx, fhy = 2 * nx + ny;
    const T fxh = coef[fhx + max(xi, 0)], fxe = coef[fex + max(xi, 0)];
    const T fxh_l = coef[fhx + max(xl, 0)];
    const uint32_t vown = own_x ? (uint32_t)xi * sizeof(T) : OUTSIDE;

    // CPML: the x slab per lane (and at the left column, lane 0's), the y slab per row
    int lx = 0, hx = 0, ly = 0, hy = 0;
    int xs = -1, xcnt = 0, xsx = 0, xsl = -1, xlcnt = 0, xls = 0;
    T bxh = 0, axh = 0, bxe = 0, axe = 0, bxl = 0, axl = 0;
    __amdgpu_buffer_rsrc_t rPi = rE, rPo = wE;
    if (SLABS) {
        lx = uniform(K.low[0]); hx = uniform(K.high[0]);
        ly = uniform(K.low[1]); hy = uniform(K.high[1]);
        const uint32_t pb = (uint32_t)uniform(K.psi_len) * sizeof(T);
        rPi = rsrc(uniform_ptr(K.psi_in), pb);
        rPo = rsrc(uniform_ptr(K.psi_out), pb);
        xs = own_x ? side_of(xi, nx, lx, hx) : -1;
        xcnt = xs == 1 ? hx : lx;
        xsx = xi - (xs == 1 ? nx - hx : 0);
        if (xs >= 0) {
            const int ch = K.cpml_off[0][xs][1] - cbase, ce = K.cpml_off[0][xs][0] - cbase;
            bxh = cp[ch + xsx]; axh = cp[ch + xcnt + xsx];
            bxe = cp[ce + xsx]; axe = cp[ce + xcnt + xsx];
        }
        xsl = xl >= 0 ? side_of(xl, nx, lx, hx) : -1;
        xlcnt = xsl == 1 ? hx : lx;
        xls = xl - (xsl == 1 ? nx - hx : 0);
        if (xsl >= 0) {
            const int ch = K.cpml_off[0][xsl][1] - cbase;
            bxl = cp[ch + xls]; axl = cp[ch + xlcnt + xls];
        }
    }
    // x slab psi byte offset of (field F, component c) on row j at this lane; none: out of range
    auto xo = [&](int F, int c, int j) -> uint32_t {
        return xs >= 0 ? (uint32_t)(K.psi_off[F][c][0][xs] + j * xcnt + xsx) * sizeof(T) : OUTSIDE;
    };
    // y slab: the row's first value (soffset), the lane's value by vown
    auto yo = [&](int F, int c, int sd, int s) -> uint32_t {
        return (uint32_t)(uniform(K.psi_off[F][c][1][sd]) + s * nx) * sizeof(T);
    };

    auto ywrap = [&](int jj) { return wrap_any(jj, ny, pery); };
    const bool has_pro = j0 > 0 || pery;
    const int jfirst = has_pro ? j0 - 1 : j0;

    // a row's E: own values and the edge column values (lane 0 left, lane 63 right)
    struct ERow {
        T ex, ey, ez, gx, gy, gz;
    };
    auto load_e = [&](ERow &R, int jj) {
        const int j = ywrap(jj);
        const uint32_t o = (xi >= 0 && j >= 0) ? (uint32_t)(j * pitch + xi) * sizeof(T) : OUTSIDE;
        const uint32_t g = (xe >= 0 && j >= 0) ? (uint32_t)(j * pitch + xe) * sizeof(T) : OUTSIDE;
        R.ex = R.ey = R.ez = R.gx = R.gy = R.gz = T(0);
        if (TE) {
            R.ex = bload<T>(rE, o, 0);
            R.ey = bload<T>(rE, o, cbytes);
            R.gx = bload<T>(rE, lane == 0 ? g : OUTSIDE, 0);
            R.gy = bload<T>(rE, g, cbytes);
        }
        if (TM) {
            R.ez = bload<T>(rE, o, 2 * cbytes);
            R.gz = bload<T>(rE, g, 2 * cbytes);
        }
    };

    ERow cur;
    load_e(cur, jfirst);
    T hzp = T(0), hxp = T(0);  // new Hz, Hx of the row below
    for (int jj = jfirst; jj < j1; jj++) {
        const bool pro = jj < j0;
        const int j = ywrap(jj);
        ERow nxt;
        load_e(nxt, jj + 1);
        const uint32_t o = xi >= 0 ? (uint32_t)(j * pitch + xi) * sizeof(T) : OUTSIDE;
        const uint32_t gl = (lane == 0 && xl >= 0) ? (uint32_t)(j * pitch + xl) * sizeof(T) : OUTSIDE;
        const uint32_t st = (own_x && !pro) ? (uint32_t)(j * pitch + i) * sizeof(T) : OUTSIDE;
        T hx_o = T(0), hy_o = T(0), hz_o = T(0), ghy = T(0), ghz = T(0);
        if (TE) {
            hz_o = bload<T>(rH, o, 2 * cbytes);
            ghz = bload<T>(rH, gl, 2 * cbytes);
        }
        if (TM) {
            hx_o = bload<T>(rH, o, 0);
            hy_o = bload<T>(rH, o, cbytes);
            ghy = bload<T>(rH, gl, cbytes);
        }
        uint8_t id = 0;
        if (MODE == 0)
            id = __builtin_amdgcn_raw_buffer_load_b8(rid, (int)(st == OUTSIDE ? OUTSIDE : st / (uint32_t)sizeof(T)), 0, 0);
        const T fyh = coef[fhy + j], fye = coef[fey + j];
        // psi of this row, loaded with the fields
        const int ys = SLABS ? side_of(j, ny, ly, hy) : -1;
        const int ycnt = ys == 1 ? hy : ly, yl = j - (ys == 1 ? ny - hy : 0);
        T qxh[2] = {T(0), T(0)}, qxe[2] = {T(0), T(0)}, qyh[2] = {T(0), T(0)}, qye[2] = {T(0), T(0)};
        T hq[3] = {T(0), T(0), T(0)};
        if (SLABS) {
            if (xs >= 0 || xsl >= 0) {  // (Cy, x) and (Cz, x)
                if (TM) qxh[0] = bload_once<T>(rPi, xo(1, 1, j), 0);
                if (TE) qxh[1] = bload_once<T>(rPi, xo(1, 2, j), 0);
                if (TE) qxe[0] = bload_once<T>(rPi, xo(0, 1, j), 0);
                if (TM) qxe[1] = bload_once<T>(rPi, xo(0, 2, j), 0);
                if (xsl >= 0 && !pro) {
                    const uint32_t q = lane == 0 ? (uint32_t)(j * xlcnt + xls) : OUTSIDE;
                    if (TM) hq[0] = bload_once<T>(rPi, q == OUTSIDE ? OUTSIDE : (uint32_t)(uniform(K.psi_off[1][1][0][xsl]) + q) * sizeof(T), 0);
                    if (TE) hq[1] = bload_once<T>(rPi, q == OUTSIDE ? OUTSIDE : (uint32_t)(uniform(K.psi_off[1][2][0][xsl]) + q) * sizeof(T), 0);
                }
            }
            if (ys >= 0) {  // (Cx, y) and (Cz, y)
                if (TM) qyh[0] = bload_once<T>(rPi, vown, yo(1, 0, ys, yl));
                if (TE) qyh[1] = bload_once<T>(rPi, vown, yo(1, 2, ys, yl));
                if (TE) qye[0] = bload_once<T>(rPi, vown, yo(0, 0, ys, yl));
                if (TM) qye[1] = bload_once<T>(rPi, vown, yo(0, 2, ys, yl));
                if (TE && !pro && xl >= 0)
                    hq[2] = bload_once<T>(rPi, lane == 0 ? (uint32_t)xl * sizeof(T) : OUTSIDE, yo(1, 2, ys, yl));
            }
        }
        const bool ok = xi >= 0;
        const bool wr = !pro && own_x;
        // new H on the row
        T hz = T(0), hx = T(0), hy = T(0);
        if (TE) {
            const T ey_n = next_lane(cur.ey, cur.gy);  // lane 63: Ey at x = i0 + 64
            const T dxz = ey_n - cur.ey, dyz = nxt.ex - cur.ex;
            hz = hz_o - dt * (dxz * fxh - dyz * fyh);
            if (SLABS) {
                if (xs >= 0 || xsl >= 0) {
                    const T u = bxh * qxh[1] + axh * (per_step(dxz, stx));
                    hz = xs >= 0 ? hz - dt * (T(1) * u) : hz;
                    if (wr) bstore<T>(u, rPo, xo(1, 2, j), 0);
                }
                if (ys >= 0) {
                    const int co = K.cpml_off[1][ys][1] - cbase;
                    const T u = cp[co + yl] * qyh[1] + cp[co + ycnt + yl] * (per_step(dyz, sty));
                    hz = hz - dt * (T(-1) * u);
                    if (wr) bstore<T>(u, rPo, vown, yo(1, 2, ys, yl));
                }
            }
        }
        if (TM) {
            const T ez_n = next_lane(cur.ez, cur.gz);
            const T dyx = nxt.ez - cur.ez, dxy = ez_n - cur.ez;
            hx = hx_o - dt * (dyx * fyh);
            hy = hy_o - dt * (T(0) - dxy * fxh);
            if (SLABS) {
                if (xs >= 0 || xsl >= 0) {
                    const T u = bxh * qxh[0] + axh * (per_step(dxy, stx));
                    hy = xs >= 0 ? hy - dt * (T(-1) * u) : hy;
                    if (wr) bstore<T>(u, rPo, xo(1, 1, j), 0);
                }
                if (ys >= 0) {
                    const int co = K.cpml_off[1][ys][1] - cbase;
                    const T u = cp[co + yl] * qyh[0] + cp[co + ycnt + yl] * (per_step(dyx, sty));
                    hx = hx - dt * (T(1) * u);
                    if (wr) bstore<T>(u, rPo, vown, yo(1, 0, ys, yl));
                }
            }
        }
        if (injects && ok) inject2d<T, 1>(A, K, bx, xi, j, 7, hx, hy, hz, T(0), T(0), T(0));
        hx = ok ? hx : T(0);
        hy = ok ? hy : T(0);
        hz = ok ? hz : T(0);
        if (!pro) {
            // H at x = i0 - 1 on lane 0, from its own gathered values and its own E at i0
            T hz_h = T(0), hy_h = T(0);
            const bool hok = lane == 0 && xl >= 0;
            if (TE) {
                const T dxz = cur.ey - cur.gy, dyz = nxt.gx - cur.gx;
                T v = ghz - dt * (dxz * fxh_l - dyz * fyh);
                if (SLABS) {
                    if (xsl >= 0) v = v - dt * (T(1) * (bxl * hq[1] + axl * (per_step(dxz, stx))));
                    if (ys >= 0) {
                        const int co = K.cpml_off[1][ys][1] - cbase;
                        v = v - dt * (T(-1) * (cp[co + yl] * hq[2] + cp[co + ycnt + yl] * (per_step(dyz, sty))));
                    }
                }
                hz_h = v;
            }
            if (TM) {
                const T dxy = cur.ez - cur.gz;
                T v = ghy - dt * (T(0) - dxy * fxh_l);
                if (SLABS && xsl >= 0) v = v - dt * (T(-1) * (bxl * hq[0] + axl * (per_step(dxy, stx))));
                hy_h = v;
            }
            if (injects && hok) {
                T d0 = T(0);
                inject2d<T, 1>(A, K, bx, xl, j, 6, d0, hy_h, hz_h, T(0), T(0), T(0));
            }
            hz_h = hok ? hz_h : T(0);
            hy_h = hok ? hy_h : T(0);
            // stores of H, then new E on the row
            if (TE) bstore<T>(hz, wH, st, 2 * cbytes);
            if (TM) {
                bstore<T>(hx, wH, st, 0);
                bstore<T>(hy, wH, st, cbytes);
            }
            const T hz_p = TE ? prev_lane(hz, hz_h) : T(0);  // lane 0: Hz at x = i0 - 1
            const T hy_p = TM ? prev_lane(hy, hy_h) : T(0);
            const bool own = st != OUTSIDE;
            T ca0, ca1, ca2, cb0, cb1, cb2;
            int g0, g1, g2;
            const long long r = (long long)(st / (uint32_t)sizeof(T));
            if (MODE == 0) {
                const T *e = tab + 6 * id;
                ca0 = e[0]; cb0 = e[1]; ca1 = e[2]; cb1 = e[3]; ca2 = e[4]; cb2 = e[5];
                g0 = grp[3 * id]; g1 = grp[3 * id + 1]; g2 = grp[3 * id + 2];
            } else if (MODE == 2) {
                const long long nv = A.nval;
                cb0 = own ? A.cacb[r] : T(0);
                cb1 = own ? A.cacb[nv + r] : T(0);
                cb2 = own ? A.cacb[2 * nv + r] : T(0);
                ca0 = lossless_ca(cb0); ca1 = lossless_ca(cb1); ca2 = lossless_ca(cb2);
                g0 = g1 = g2 = -1;
            } else {
                const long long nv = A.nval;
                ca0 = own ? A.cacb[r] : T(0);
                ca1 = own ? A.cacb[nv + r] : T(0);
                ca2 = own ? A.cacb[2 * nv + r] : T(0);
                cb0 = own ? A.cacb[3 * nv + r] : T(0);
                cb1 = own ? A.cacb[4 * nv + r] : T(0);
                cb2 = own ? A.cacb[5 * nv + r] : T(0);
                g0 = own && A.grp ? A.grp[r] : -1;
                g1 = own && A.grp ? A.grp[nv + r] : -1;
                g2 = own && A.grp ? A.grp[2 * nv + r] : -1;
            }
            T ex = T(0), ey = T(0), ez = T(0);
            T dyx = T(0), dxy2 = T(0), dxz2 = T(0), dyz2 = T(0);
            if (TE) {
                dyx = hz - hzp;
                dxy2 = hz - hz_p;
                ex = ca0 * cur.ex + cb0 * (dyx * fye);
                ey = ca1 * cur.ey + cb1 * (T(0) - dxy2 * fxe);
            }
            if (TM) {
                dxz2 = hy - hy_p;
                dyz2 = hx - hxp;
                ez = ca2 * cur.ez + cb2 * (dxz2 * fxe - dyz2 * fye);
            }
            if (SLABS && own) {
                if (xs >= 0) {
                    if (TE) {
                        const T u = bxe * qxe[0] + axe * (per_step(dxy2, stx));
                        ey = ey + cb1 * (T(-1) * u);
                        bstore<T>(u, rPo, xo(0, 1, j), 0);
                    }
                    if (TM) {
                        const T u = bxe * qxe[1] + axe * (per_step(dxz2, stx));
                        ez = ez + cb2 * (T(1) * u);
                        bstore<T>(u, rPo, xo(0, 2, j), 0);
                    }
                }
                if (ys >= 0) {
                    const int co = K.cpml_off[1][ys][0] - cbase;
                    const T bb = cp[co + yl], aa = cp[co + ycnt + yl];
                    if (TE) {
                        const T u = bb * qye[0] + aa * (per_step(dyx, sty));
                        ex = ex + cb0 * (T(1) * u);
                        bstore<T>(u, rPo, vown, yo(0, 0, ys, yl));
                    }
                    if (TM) {
                        const T u = bb * qye[1] + aa * (per_step(dyz2, sty));
                        ez = ez + cb2 * (T(-1) * u);
                        bstore<T>(u, rPo, vown, yo(0, 2, ys, yl));
                    }
                }
            }
            if (injects && own) inject2d<T, 0>(A, K, bx, xi, j, SET == 3 ? 7 : (TE ? 3 : 4), ex, ey, ez, cb0, cb1, cb2);
            if (TE && own && g0 >= 0) ex = media_update(A, K, g0, 0, (int)r, ex, cb0, cur.ex);
            if (TE && own && g1 >= 0) ey = media_update(A, K, g1, 1, (int)r, ey, cb1, cur.ey);
            if (TM && own && g2 >= 0) ez = media_update(A, K, g2, 2, (int)r, ez, cb2, cur.ez);
            if (TE) {
                bstore<T>(ex, wE, st, 0);
                bstore<T>(ey, wE, st, cbytes);
            }
            if (TM) bstore<T>(ez, wE, st, 2 * cbytes);
        }
        hzp = hz;
        hxp = hx;
        cur = nxt;
    }
}

template <typename T, int MODE, int AX, int SET>
__global__ __launch_bounds__(256) void march2d_step(StepArgs<T> A) {
    __shared__ MarchShared<T> S;
    extern __shared__ unsigned char smem_raw[];
    T *tab = reinterpret_cast<T *>(smem_raw);
    int8_t *grp = reinterpret_cast<int8_t *>(tab + 256 * 6);
    march64_stage<T, MODE, AX != 0>(A, S, tab, grp);
    __syncthreads();
    const int b = (int)(blockIdx.x % 8) * A.per_xcd + (int)(blockIdx.x / 8);
    const int w = uniform(b * 4 + (int)(threadIdx.x >> 6));
    if (w >= A.blocks) return;
    march2d_unit<T, MODE, AX, SET>(A, A.units[w], S, tab, grp);
}

}  // namespace pxfdtd
