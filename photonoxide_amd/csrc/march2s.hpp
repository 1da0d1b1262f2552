// The two-step march: two time steps of the fused Yee + CPML update in one pass over HBM.
//
// A step reads E, H and psi and writes them back; two steps in one pass read and write them
// once, so the bytes per cell-update halve. Each wavefront owns a strip of SW2 = 60 values in x
// by RY rows in y and marches along z over a chunk of planes, and it computes the first step's
// values on the margin around its strip that the second step reads:
//   - lanes: lane L holds x = i0 - 2 + L (i0 = 60 bx), so the strip's own columns are lanes
//     2 .. 61 and every neighbour either step needs along x is another lane's (DPP shifts, no
//     gathers): H1 (the first step's H) is valid on lanes 0 .. 62, E1 on 1 .. 62, H2 on 1 .. 61
//     and E2 on 2 .. 61;
//   - rows: E0 (the pass's input E) on rows j0 - 2 .. j0 + RY + 1, H1 on j0 - 2 .. j0 + RY,
//     E1 on j0 - 1 .. j0 + RY, H2 on j0 - 1 .. j0 + RY - 1 and E2 on the own rows;
//   - planes: the second step runs a plane behind the first (E2 at plane k needs H2 at k - 1,
//     H2 at k needs E1 at k + 1), so a chunk [kz0, kz1) computes H1 from kz0 - 2 and stores E2,
//     H2 of plane k when the march is at plane k + 1.
// The CPML's psi of the first step stays in registers for the second (the values of H2, E2's
// rows, one plane); only the second step's psi of own values is stored. Sources and currents
// apply at every value either step computes (the tile's list covers its margin), with the
// waves of each step. The monitored values the first step leaves are written to the staging
// rows from here (each own value by its one owner); the second step's are the taps' after the
// pass, from the fields stored.
//
// Every value is computed with the same operations in the same order as one step of the
// aligned march (march64.hpp) and of the reference's loops (src/fdtd/kernel/reference.rs:14-180),
// and only own values of the second step are stored, so one pass is the bits of two steps of
// the other kernels. Values beyond a wall are zero, as the others hold them; across a periodic
// side the lanes and rows wrap.
#pragma once
#include <type_traits>

#include "march64.hpp"

namespace pxfdtd {

constexpr int SW2 = 60;  // own columns per strip of the two-step march

// A monitored value the first step leaves: its column and its slot in a staging row. A tile's
// values are listed per (plane, own row, component Ex Ey Ez Hx Hy Hz).
struct TapAt {
    int32_t i, pad;
    int64_t slot;
};
template <typename T>
struct PairTaps {
    const int32_t *off;  // [tile][nz][RY][6] + 1 into at, or null for no taps
    const TapAt *at;
    T *stage;            // the first step's staging row
};

template <typename T, int MODE, int RY, int AX>
__device__ __forceinline__ void march2s_unit(const StepArgs<T> &A, const PairTaps<T> &TP, const int4 unit,
                                             const MarchShared<T> &S, const T *tab) {
    constexpr int NE = RY + 4;  // rows of E0: j0 - 2 .. j0 + RY + 1
    constexpr int NH = RY + 3;  // rows of H0 and H1: j0 - 2 .. j0 + RY
    constexpr int N1 = RY + 2;  // rows of E1: j0 - 1 .. j0 + RY (E0 rows 1 .. RY + 2)
    constexpr int N2 = RY + 1;  // rows of H2: j0 - 1 .. j0 + RY - 1 (E0 rows 1 .. RY + 1)
    constexpr bool SLABS = AX != 0, XS = AX & 1, YS = AX & 2, ZS = AX & 4;
    const StepCold<T> &K = S.K;
    const T *cp = S.cp;
    const int lane = threadIdx.x & 63;
    const int bx = unit.x, by = unit.y, kz0 = unit.z, kz1 = unit.w;
    const int t = by * A.tiles_x + bx;
    const int i0 = bx * SW2, j0 = by * RY;
    const int nx = A.nx, ny = A.ny, nz = A.nz, pitch = A.pitch;
    const int perx = A.flags & 1, pery = (A.flags >> 1) & 1, perz = (A.flags >> 2) & 1;
    const int cbase = 2 * (nx + ny + nz);
    // the sources and currents are listed per (tile, plane, E0 row, field), the monitored values
    // per (tile, plane, own row, component): a value's update reads its own short list
    const long long per_tile_inj = (long long)nz * NE * 2, per_tile_tap = (long long)nz * RY * 6;
    bool injects = false;
    if (A.flags & 16) {
        const int32_t *io = K.inj_off + (long long)t * per_tile_inj;
        injects = io[per_tile_inj] > io[0];
    }
    bool taps = false;
    if (TP.off) {
        const int32_t *to = TP.off + (long long)t * per_tile_tap;
        taps = to[per_tile_tap] > to[0];
    }
    const int ws1 = A.wave_step, ws2 = A.wave_step + 1;

    const int x = i0 - 2 + lane;
    const int xi = wrap_any(x, nx, perx);  // -1 beyond a wall
    const bool xok = xi >= 0;
    const bool own_x = lane >= 2 && lane < 2 + SW2 && x < nx;
    int yj[NE];
    uint32_t vo[NE];
#pragma unroll
    for (int r = 0; r < NE; r++) {
        yj[r] = wrap_any(j0 - 2 + r, ny, pery);
        vo[r] = (xok && yj[r] >= 0) ? (uint32_t)(yj[r] * pitch + xi) * sizeof(T) : OUTSIDE;
    }
    uint32_t so[RY];  // stores of the own rows (E0 rows 2 .. RY + 1)
#pragma unroll
    for (int q = 0; q < RY; q++) {
        const int jr = j0 + q;
        so[q] = (own_x && jr < ny) ? (uint32_t)(jr * pitch + x) * sizeof(T) : OUTSIDE;
    }

    const uint32_t cbytes = (uint32_t)A.nval * sizeof(T), pbytes = (uint32_t)A.plane * sizeof(T);
    const __amdgpu_buffer_rsrc_t rE = rsrc(A.ein, 3 * cbytes), rH = rsrc(A.hin, 3 * cbytes),
                                 wE = rsrc(A.eout, 3 * cbytes), wH = rsrc(A.hout, 3 * cbytes),
                                 rid = rsrc(A.ids, MODE == 0 ? (uint32_t)A.nval : 0u),
                                 rnone = rsrc(A.ein, 0u);
    const T dt = A.dt;
    const T stx = A.step[0], sty = A.step[1], stz = A.step[2];
    const T *coef = A.coef;
    const int fex = 0, fhx = nx, fey = 2 * nx, fhy = 2 * nx + ny, fez = 2 * (nx + ny), fhz = 2 * (nx + ny) + nz;
    const T fxh = coef[fhx + max(xi, 0)], fxe = coef[fex + max(xi, 0)];
    T fyh[NE], fye[NE];
#pragma unroll
    for (int r = 0; r < NE; r++) {
        fyh[r] = coef[fhy + max(yj[r], 0)];
        fye[r] = coef[fey + max(yj[r], 0)];
    }
    auto zwrap = [&](int kk) __attribute__((always_inline)) { return wrap_any(kk, nz, perz); };

    // CPML geometry fixed for the unit: the x slab per lane, the y slab per row
    int lx = 0, hx = 0, ly = 0, hy = 0, lz = 0, hz = 0;
    int xs = -1, xcnt = 0;
    int xb[2][2] = {{0, 0}, {0, 0}};  // x slab psi base per [field][component y, z] at this lane
    T bxh = 0, axh = 0, bxe = 0, axe = 0;
    int ys[NE], yl[NE];
    bool xany = false;
    __amdgpu_buffer_rsrc_t rPi = rnone, rPo = rnone;
    const uint32_t vcol = xok ? (uint32_t)xi * sizeof(T) : OUTSIDE;      // a lane's column, any lane
    const uint32_t vown = own_x ? (uint32_t)xi * sizeof(T) : OUTSIDE;    // the own lanes'
    if (SLABS) {
        lx = uniform(K.low[0]); hx = uniform(K.high[0]);
        ly = uniform(K.low[1]); hy = uniform(K.high[1]);
        lz = uniform(K.low[2]); hz = uniform(K.high[2]);
        const uint32_t pb = (uint32_t)uniform(K.psi_len) * sizeof(T);
        rPi = rsrc(uniform_ptr(K.psi_in), pb);
        rPo = rsrc(uniform_ptr(K.psi_out), pb);
        xs = xok ? side_of(xi, nx, lx, hx) : -1;
        xcnt = xs == 1 ? hx : lx;
        const int xloc = xi - (xs == 1 ? nx - hx : 0);
        if (XS && xs >= 0) {
            const int ch = K.cpml_off[0][xs][1] - cbase, ce = K.cpml_off[0][xs][0] - cbase;
            bxh = cp[ch + xloc]; axh = cp[ch + xcnt + xloc];
            bxe = cp[ce + xloc]; axe = cp[ce + xcnt + xloc];
#pragma unroll
            for (int F = 0; F < 2; F++) {
                xb[F][0] = K.psi_off[F][1][0][xs] + xloc;
                xb[F][1] = K.psi_off[F][2][0][xs] + xloc;
            }
        }
        // a lane of the strip, margin included, in an x slab
        xany = (lx > 0 && i0 - 2 < lx) || (hx > 0 && min(i0 + SW2 + 2, nx) > nx - hx);
#pragma unroll
        for (int r = 0; r < NE; r++) {
            ys[r] = yj[r] >= 0 ? side_of(yj[r], ny, ly, hy) : -1;
            yl[r] = yj[r] - (ys[r] == 1 ? ny - hy : 0);
        }
    }
    // psi byte offsets: x slab of (field F, component slot s = 0 y / 1 z) at row r, plane k, the
    // lanes in `lanes`; y and z slabs: the row's first value (scalar), the lane's column apart
    // x slab: each lane's base (OUTSIDE off the slab; for stores, off the own lanes too) plus the
    // row and plane's (k ny + j) times the slab's width
    uint32_t xvl[2][2], xvs[2][2];
#pragma unroll
    for (int F = 0; F < 2; F++)
#pragma unroll
        for (int sl = 0; sl < 2; sl++) {
            xvl[F][sl] = xs >= 0 ? (uint32_t)xb[F][sl] * sizeof(T) : OUTSIDE;
            xvs[F][sl] = xs >= 0 && own_x ? (uint32_t)xb[F][sl] * sizeof(T) : OUTSIDE;
        }
    // the two sides' widths are scalars, so the row's offset is a scalar product and one select
    const uint32_t xlb = (uint32_t)lx * sizeof(T), xhb = (uint32_t)hx * sizeof(T);
    auto xrow = [&](int r, int k) __attribute__((always_inline)) -> uint32_t {
        const uint32_t kj = (uint32_t)uniform(k * ny + yj[r]);
        return xs == 1 ? kj * xhb : kj * xlb;
    };
    auto yo = [&](int F, int c, int r, int k) __attribute__((always_inline)) -> uint32_t {
        const int cnt = ys[r] == 1 ? hy : ly;
        return (uint32_t)(uniform(K.psi_off[F][c][1][ys[r]]) + (k * cnt + yl[r]) * nx) * sizeof(T);
    };
    auto zo = [&](int F, int c, int r, int zs, int zl) __attribute__((always_inline)) -> uint32_t {
        return (uint32_t)(uniform(K.psi_off[F][c][2][zs]) + (zl * ny + yj[r]) * nx) * sizeof(T);
    };

    struct EPlane {
        T e[NE][3];
    };
    struct HPlane {
        T h[NH][3];
        uint8_t id[N1];  // E1's rows
    };
    auto load_e = [&](EPlane &P, int kk) __attribute__((always_inline)) {
        const int kp = zwrap(kk);
        const __amdgpu_buffer_rsrc_t rr = kp >= 0 ? rE : rnone;
        const uint32_t s = kp >= 0 ? (uint32_t)kp * pbytes : 0u;
#pragma unroll
        for (int r = 0; r < NE; r++) {
            P.e[r][0] = bload<T>(rr, vo[r], s);
            P.e[r][1] = r == NE - 1 ? T(0) : bload<T>(rr, vo[r], s + cbytes);  // the top row: Ex, Ez
            P.e[r][2] = bload<T>(rr, vo[r], s + 2 * cbytes);
        }
    };
    auto load_h = [&](HPlane &P, int kk) __attribute__((always_inline)) {
        const int kp = zwrap(kk);
        const __amdgpu_buffer_rsrc_t rr = kp >= 0 ? rH : rnone;
        const __amdgpu_buffer_rsrc_t ri = kp >= 0 ? rid : rnone;
        const uint32_t s = kp >= 0 ? (uint32_t)kp * pbytes : 0u;
#pragma unroll
        for (int r = 0; r < NH; r++) {
            P.h[r][0] = bload<T>(rr, vo[r], s);
            P.h[r][1] = r == 0 ? T(0) : bload<T>(rr, vo[r], s + cbytes);  // the bottom row: Hx, Hz
            P.h[r][2] = bload<T>(rr, vo[r], s + 2 * cbytes);
        }
#pragma unroll
        for (int q = 0; q < N1; q++) {
            if (MODE == 0) {
                const uint32_t o = vo[q + 1] == OUTSIDE ? OUTSIDE : vo[q + 1] / (uint32_t)sizeof(T);
                P.id[q] = __builtin_amdgcn_raw_buffer_load_b8(ri, (int)o, (int)(kp >= 0 ? (uint32_t)kp * (uint32_t)A.plane : 0u), 0);
            } else {
                P.id[q] = 0;
            }
        }
    };
    struct Coef {
        T ca[3], cb[3];
    };
    // ca, cb of a value at E0 row r of plane k (mode 0: from its id)
    auto coef_at = [&](int r, int k, uint8_t id, bool ok) __attribute__((always_inline)) -> Coef {
        Coef c;
        if (MODE == 0) {
            const T *e = tab + 6 * id;
            c.ca[0] = e[0]; c.cb[0] = e[1]; c.ca[1] = e[2]; c.cb[1] = e[3]; c.ca[2] = e[4]; c.cb[2] = e[5];
        } else {
            const long long nv = A.nval;
            const long long o = ok ? (long long)k * A.plane + (long long)yj[r] * pitch + xi : 0;
#pragma unroll
            for (int c3 = 0; c3 < 3; c3++) {
                if (MODE == 2) {
                    c.cb[c3] = ok ? A.cacb[c3 * nv + o] : T(0);
                    c.ca[c3] = lossless_ca(c.cb[c3]);
                } else {
                    c.ca[c3] = ok ? A.cacb[c3 * nv + o] : T(0);
                    c.cb[c3] = ok ? A.cacb[(3 + c3) * nv + o] : T(0);
                }
            }
        }
        return c;
    };

    // psi of one H value by axis: x (Hy, Hz), y (Hx, Hz), z (Hx, Hy); of one E value: x (Ey, Ez),
    // y (Ex, Ez), z (Ex, Ey)
    struct Psi {
        T x[2], y[2], z[2];
    };
    // A slab plane's coefficients: the z slab side, local plane and b, a for E and H
    struct ZSlab {
        int side, loc;
        T bh, ah, be, ae;
    };
    auto zslab = [&](int k) __attribute__((always_inline)) -> ZSlab {
        ZSlab z = {-1, 0, T(0), T(0), T(0), T(0)};
        if (ZS && k >= 0) {
            z.side = side_of(k, nz, lz, hz);
            if (z.side >= 0) {
                z.loc = k - (z.side == 1 ? nz - hz : 0);
                const int cnt = z.side == 1 ? hz : lz;
                const int ch = K.cpml_off[2][z.side][1] - cbase, ce = K.cpml_off[2][z.side][0] - cbase;
                z.bh = cp[ch + z.loc]; z.ah = cp[ch + cnt + z.loc];
                z.be = cp[ce + z.loc]; z.ae = cp[ce + cnt + z.loc];
            }
        }
        return z;
    };
    // the y slab's b, a at E0 row r for field F (0 E, 1 H)
    auto yba = [&](int F, int r, T &b, T &a) __attribute__((always_inline)) {
        const int cnt = ys[r] == 1 ? hy : ly;
        const int co = K.cpml_off[1][ys[r]][F] - cbase;
        b = cp[co + yl[r]];
        a = cp[co + cnt + yl[r]];
    };
    // psi of field F at E0 row r of plane k (in), loaded for the lanes in `lanes`
    auto load_psi = [&](int F, int r, int k, const ZSlab &zs, bool lanes) __attribute__((always_inline)) -> Psi {
        Psi p = {{T(0), T(0)}, {T(0), T(0)}, {T(0), T(0)}};
        if (XS && xany && yj[r] >= 0) {
            const uint32_t ro = xrow(r, k);
            p.x[0] = bload_once<T>(rPi, (lanes ? xvl[F][0] : xvs[F][0]) + ro, 0);
            p.x[1] = bload_once<T>(rPi, (lanes ? xvl[F][1] : xvs[F][1]) + ro, 0);
        }
        const uint32_t vl = lanes ? vcol : OUTSIDE;
        if (YS && ys[r] >= 0) {
            p.y[0] = bload_once<T>(rPi, vl, yo(F, 0, r, k));
            p.y[1] = bload_once<T>(rPi, vl, yo(F, 2, r, k));
        }
        if (ZS && zs.side >= 0 && yj[r] >= 0) {
            p.z[0] = bload_once<T>(rPi, vl, zo(F, 0, r, zs.side, zs.loc));
            p.z[1] = bload_once<T>(rPi, vl, zo(F, 1, r, zs.side, zs.loc));
        }
        return p;
    };
    auto store_psi = [&](int F, int r, int k, const ZSlab &zs, const Psi &p) __attribute__((always_inline)) {
        if (XS && xany && yj[r] >= 0) {
            const uint32_t ro = xrow(r, k);
            bstore<T>(p.x[0], rPo, xvs[F][0] + ro, 0);
            bstore<T>(p.x[1], rPo, xvs[F][1] + ro, 0);
        }
        if (YS && ys[r] >= 0) {
            bstore<T>(p.y[0], rPo, vown, yo(F, 0, r, k));
            bstore<T>(p.y[1], rPo, vown, yo(F, 2, r, k));
        }
        if (ZS && zs.side >= 0 && yj[r] >= 0) {
            bstore<T>(p.z[0], rPo, vown, zo(F, 0, r, zs.side, zs.loc));
            bstore<T>(p.z[1], rPo, vown, zo(F, 1, r, zs.side, zs.loc));
        }
    };

    // One H value: H - dt curl E, then the slabs' terms (per component in axis order), from the
    // differences d = (dyx, dzx, dxy, dzy, dxz, dyz) and psi q; returns the new psi in q
    auto h_value = [&](int r, const ZSlab &zs, const T ho[3], const T d[6], T fzh, Psi &q, T v[3]) __attribute__((always_inline)) {
        v[0] = ho[0] - dt * (d[0] * fyh[r] - d[1] * fzh);
        v[1] = ho[1] - dt * (d[3] * fzh - d[2] * fxh);
        v[2] = ho[2] - dt * (d[4] * fxh - d[5] * fyh[r]);
        if (SLABS) {
            if (XS && xany) {
                const bool in = xs >= 0;
                const T u1 = bxh * q.x[0] + axh * per_step(d[2], stx);
                const T u2 = bxh * q.x[1] + axh * per_step(d[4], stx);
                v[1] = in ? v[1] - dt * (T(-1) * u1) : v[1];
                v[2] = in ? v[2] - dt * (T(1) * u2) : v[2];
                q.x[0] = u1;
                q.x[1] = u2;
            }
            if (YS && ys[r] >= 0) {
                T bb, aa;
                yba(1, r, bb, aa);
                const T u0 = bb * q.y[0] + aa * per_step(d[0], sty);
                const T u2 = bb * q.y[1] + aa * per_step(d[5], sty);
                v[0] = v[0] - dt * (T(1) * u0);
                v[2] = v[2] - dt * (T(-1) * u2);
                q.y[0] = u0;
                q.y[1] = u2;
            }
            if (ZS && zs.side >= 0) {
                const T u0 = zs.bh * q.z[0] + zs.ah * per_step(d[1], stz);
                const T u1 = zs.bh * q.z[1] + zs.ah * per_step(d[3], stz);
                v[0] = v[0] - dt * (T(-1) * u0);
                v[1] = v[1] - dt * (T(1) * u1);
                q.z[0] = u0;
                q.z[1] = u1;
            }
        }
    };
    // One E value: ca E + cb curl H, then the slabs' terms, from d = (dyx, dzx, dxy, dzy, dxz,
    // dyz) of H and psi q; returns the new psi in q
    auto e_value = [&](int r, const ZSlab &zs, const T eo[3], const Coef &c, const T d[6], T fze, Psi &q, T v[3]) __attribute__((always_inline)) {
        v[0] = c.ca[0] * eo[0] + c.cb[0] * (d[0] * fye[r] - d[1] * fze);
        v[1] = c.ca[1] * eo[1] + c.cb[1] * (d[3] * fze - d[2] * fxe);
        v[2] = c.ca[2] * eo[2] + c.cb[2] * (d[4] * fxe - d[5] * fye[r]);
        if (SLABS) {
            if (XS && xany) {
                const bool in = xs >= 0;
                const T u1 = bxe * q.x[0] + axe * per_step(d[2], stx);
                const T u2 = bxe * q.x[1] + axe * per_step(d[4], stx);
                v[1] = in ? v[1] + c.cb[1] * (T(-1) * u1) : v[1];
                v[2] = in ? v[2] + c.cb[2] * (T(1) * u2) : v[2];
                q.x[0] = u1;
                q.x[1] = u2;
            }
            if (YS && ys[r] >= 0) {
                T bb, aa;
                yba(0, r, bb, aa);
                const T u0 = bb * q.y[0] + aa * per_step(d[0], sty);
                const T u2 = bb * q.y[1] + aa * per_step(d[5], sty);
                v[0] = v[0] + c.cb[0] * (T(1) * u0);
                v[2] = v[2] + c.cb[2] * (T(-1) * u2);
                q.y[0] = u0;
                q.y[1] = u2;
            }
            if (ZS && zs.side >= 0) {
                const T u0 = zs.be * q.z[0] + zs.ae * per_step(d[1], stz);
                const T u1 = zs.be * q.z[1] + zs.ae * per_step(d[3], stz);
                v[0] = v[0] + c.cb[0] * (T(-1) * u0);
                v[1] = v[1] + c.cb[1] * (T(1) * u1);
                q.z[0] = u0;
                q.z[1] = u1;
            }
        }
    };
    // the sources, currents of field FIELD at E0 row r of plane k, with the waves of row ws
    auto inj = [&](auto field, int r, int k, int ws, T (&v)[3], T cb0, T cb1, T cb2) __attribute__((always_inline)) {
        constexpr int FIELD = decltype(field)::value;
        const int32_t *io = K.inj_off + (long long)t * per_tile_inj + ((long long)k * NE + r) * 2 + FIELD;
        const int lo = io[0], hi = io[1];
        for (int m = lo; m < hi; m++) {
            const Injection in = K.inj[m];
            if (in.i != xi) continue;
            apply_injection<T, FIELD>(A, K, in, in.code & 3, v[0], v[1], v[2], cb0, cb1, cb2, ws);
        }
    };
    // the first step's monitored values at plane k (own rows): E1 row q + 1 and H1 row q + 2
    auto tap = [&](int k, const T (&e1)[N1][3], const T (&h1)[NH][3]) __attribute__((always_inline)) {
        const int32_t *to = TP.off + (long long)t * per_tile_tap + (long long)k * RY * 6;
#pragma unroll
        for (int q = 0; q < RY; q++) {
#pragma unroll
            for (int c = 0; c < 6; c++) {
                const int lo = to[q * 6 + c], hi = to[q * 6 + c + 1];
                const T v = c < 3 ? e1[q + 1][c] : h1[q + 2][c - 3];
                for (int m = lo; m < hi; m++) {
                    const TapAt a = TP.at[m];
                    if (own_x && xi == a.i) TP.stage[a.slot] = v;
                }
            }
        }
    };

    // carried between planes: H1 of the plane below (E0 rows 1 .. RY + 2; all components on
    // rows 1 .. RY + 1, H2's old H) and its psi on H2's rows, E1 of the plane below and its psi
    // on E2's rows, H2's Hx, Hy of the plane two below (E2's rows), and the plane below's
    // coefficients of E2's rows
    T h1p[N1][3], e1p[N1][3], h2p[RY][2];
    Psi ph1[N2], pe1[RY];
    Coef cfp[RY];
    uint8_t idp[RY];
#pragma unroll
    for (int q = 0; q < N1; q++) h1p[q][0] = h1p[q][1] = h1p[q][2] = e1p[q][0] = e1p[q][1] = e1p[q][2] = T(0);
#pragma unroll
    for (int q = 0; q < N2; q++) ph1[q] = Psi{{T(0), T(0)}, {T(0), T(0)}, {T(0), T(0)}};
#pragma unroll
    for (int q = 0; q < RY; q++) {
        h2p[q][0] = h2p[q][1] = T(0);
        pe1[q] = Psi{{T(0), T(0)}, {T(0), T(0)}, {T(0), T(0)}};
        cfp[q] = Coef{};
        idp[q] = 0;
    }

    auto step_plane = [&](const int kk, const EPlane &cur, const EPlane &nxt, const HPlane &hc) __attribute__((always_inline)) {
        const int k = zwrap(kk);       // H1, E1
        const int kb = zwrap(kk - 1);  // H2, E2
        const bool kok = k >= 0, kbok = kb >= 0;
        const T fzh = coef[fhz + max(k, 0)], fze = coef[fez + max(k, 0)];
        const T fzhb = coef[fhz + max(kb, 0)], fzeb = coef[fez + max(kb, 0)];
        const ZSlab zs = zslab(k), zsb = zslab(kb);

        // the first step's psi at plane k: H1's rows and E1's, the margin's lanes included
        Psi qh[NH], qe[N1];
#pragma unroll
        for (int r = 0; r < NH; r++) qh[r] = SLABS ? load_psi(1, r, k, zs, true) : Psi{};
#pragma unroll
        for (int q = 0; q < N1; q++) qe[q] = SLABS ? load_psi(0, q + 1, k, zs, true) : Psi{};

        // H1 at plane kk, E0 rows 0 .. RY + 2
        T h1[NH][3];
#pragma unroll
        for (int r = 0; r < NH; r++) {
            const T ey_n = next_lane(cur.e[r][1], T(0));
            const T ez_n = next_lane(cur.e[r][2], T(0));
            const T d[6] = {cur.e[r + 1][2] - cur.e[r][2], nxt.e[r][1] - cur.e[r][1], ez_n - cur.e[r][2],
                            nxt.e[r][0] - cur.e[r][0], ey_n - cur.e[r][1], cur.e[r + 1][0] - cur.e[r][0]};
            T v[3];
            h_value(r, zs, hc.h[r], d, fzh, qh[r], v);
            const bool ok = kok && xok && yj[r] >= 0;
            if (injects && ok) inj(std::integral_constant<int, 1>{}, r, k, ws1, v, T(0), T(0), T(0));
            h1[r][0] = ok ? v[0] : T(0);
            h1[r][1] = ok ? v[1] : T(0);
            h1[r][2] = ok ? v[2] : T(0);
        }
        // E1 at plane kk, E0 rows 1 .. RY + 2
        T e1[N1][3];
        Coef cf[N1];
#pragma unroll
        for (int q = 0; q < N1; q++) {
            const int r = q + 1;
            const bool ok = kok && xok && yj[r] >= 0;
            cf[q] = coef_at(r, k, hc.id[q], ok);
            const T hy_p = prev_lane(h1[r][1], T(0));
            const T hz_p = prev_lane(h1[r][2], T(0));
            const T d[6] = {h1[r][2] - h1[r - 1][2], h1[r][1] - h1p[q][1], h1[r][2] - hz_p,
                            h1[r][0] - h1p[q][0], h1[r][1] - hy_p, h1[r][0] - h1[r - 1][0]};
            T v[3];
            e_value(r, zs, cur.e[r], cf[q], d, fze, qe[q], v);
            if (injects && ok) inj(std::integral_constant<int, 0>{}, r, k, ws1, v, cf[q].cb[0], cf[q].cb[1], cf[q].cb[2]);
            e1[q][0] = ok ? v[0] : T(0);
            e1[q][1] = ok ? v[1] : T(0);
            e1[q][2] = ok ? v[2] : T(0);
        }
        if (taps && kk >= kz0 && kk < kz1) tap(k, e1, h1);

        // H2 at plane kk - 1, E0 rows 1 .. RY + 1, from E1 there and above
        T h2[N2][3];
#pragma unroll
        for (int q = 0; q < N2; q++) {
            const int r = q + 1;
            const T ey_n = next_lane(e1p[q][1], T(0));
            const T ez_n = next_lane(e1p[q][2], T(0));
            const T d[6] = {e1p[q + 1][2] - e1p[q][2], e1[q][1] - e1p[q][1], ez_n - e1p[q][2],
                            e1[q][0] - e1p[q][0], ey_n - e1p[q][1], e1p[q + 1][0] - e1p[q][0]};
            T v[3];
            Psi p = ph1[q];
            h_value(r, zsb, h1p[q], d, fzhb, p, v);
            const bool ok = kbok && xok && yj[r] >= 0;
            if (injects && ok) inj(std::integral_constant<int, 1>{}, r, kb, ws2, v, T(0), T(0), T(0));
            h2[q][0] = ok ? v[0] : T(0);
            h2[q][1] = ok ? v[1] : T(0);
            h2[q][2] = ok ? v[2] : T(0);
            ph1[q] = p;  // psi^(n+2) on H2's rows, stored below for the own ones
        }
        // E2 at plane kk - 1, the own rows (E0 rows 2 .. RY + 1), and the stores
        const bool store = kk - 1 >= kz0 && kk - 1 < kz1 && kbok;
        const uint32_t sk = (uint32_t)max(kb, 0) * pbytes;
#pragma unroll
        for (int q = 0; q < RY; q++) {
            const int r = q + 2;
            const T hy_p = prev_lane(h2[q + 1][1], T(0));
            const T hz_p = prev_lane(h2[q + 1][2], T(0));
            const T d[6] = {h2[q + 1][2] - h2[q][2], h2[q + 1][1] - h2p[q][1], h2[q + 1][2] - hz_p,
                            h2[q + 1][0] - h2p[q][0], h2[q + 1][1] - hy_p, h2[q + 1][0] - h2[q][0]};
            const Coef c = MODE == 0 ? coef_at(r, kb, idp[q], true) : cfp[q];
            T v[3];
            Psi p = pe1[q];
            e_value(r, zsb, e1p[q + 1], c, d, fzeb, p, v);
            const bool own = so[q] != OUTSIDE;
            if (injects && own && kbok) inj(std::integral_constant<int, 0>{}, r, kb, ws2, v, c.cb[0], c.cb[1], c.cb[2]);
            if (store) {
                const uint32_t st = so[q];
                bstore<T>(h2[q + 1][0], wH, st, sk);
                bstore<T>(h2[q + 1][1], wH, st, sk + cbytes);
                bstore<T>(h2[q + 1][2], wH, st, sk + 2 * cbytes);
                bstore<T>(v[0], wE, st, sk);
                bstore<T>(v[1], wE, st, sk + cbytes);
                bstore<T>(v[2], wE, st, sk + 2 * cbytes);
                if (SLABS && j0 + q < ny) {
                    store_psi(1, r, kb, zsb, ph1[q + 1]);
                    store_psi(0, r, kb, zsb, p);
                }
            }
        }
        // what the next plane needs from this one
#pragma unroll
        for (int q = 0; q < N1; q++) {
            h1p[q][0] = h1[q + 1][0];
            h1p[q][1] = h1[q + 1][1];
            h1p[q][2] = h1[q + 1][2];
            e1p[q][0] = e1[q][0];
            e1p[q][1] = e1[q][1];
            e1p[q][2] = e1[q][2];
        }
#pragma unroll
        for (int q = 0; q < N2; q++) ph1[q] = qh[q + 1];
#pragma unroll
        for (int q = 0; q < RY; q++) {
            h2p[q][0] = h2[q + 1][0];
            h2p[q][1] = h2[q + 1][1];
            pe1[q] = qe[q + 1];
            if (MODE == 0) idp[q] = hc.id[q + 1];
            else cfp[q] = cf[q + 1];
        }
    };

    const int kfirst = kz0 - 2;
    EPlane cur, nxt;
    HPlane hcur;
    load_e(cur, kfirst);
    load_e(nxt, kfirst + 1);
    load_h(hcur, kfirst);
    for (int kk = kfirst; kk <= kz1; kk++) {
        EPlane nxt2;
        HPlane hnxt;
        load_e(nxt2, kk + 2);
        load_h(hnxt, kk + 1);
        step_plane(kk, cur, nxt, hcur);
        cur = nxt;
        nxt = nxt2;
        hcur = hnxt;
    }
}

template <typename T, int MODE, int RY, int AX>
__global__ __launch_bounds__(256) void march2s_step(StepArgs<T> A, PairTaps<T> TP) {
    __shared__ MarchShared<T> S;
    extern __shared__ unsigned char smem_raw[];
    T *tab = reinterpret_cast<T *>(smem_raw);                 // [256][3][2]
    int8_t *grp = reinterpret_cast<int8_t *>(tab + 256 * 6);  // [256][3]
    march64_stage<T, MODE, AX != 0>(A, S, tab, grp);
    __syncthreads();  // the parameters in LDS
    const int b = (int)(blockIdx.x % 8) * A.per_xcd + (int)(blockIdx.x / 8);
    const int w = uniform(b * 4 + (int)(threadIdx.x >> 6));
    if (w >= A.blocks) return;
    march2s_unit<T, MODE, RY, AX>(A, TP, A.units[w], S, tab);
}

}  // namespace pxfdtd
