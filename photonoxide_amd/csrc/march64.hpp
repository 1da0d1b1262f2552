// The aligned register march: the fused Yee + CPML step the time-stepper runs by default.
//
// Each wavefront owns a strip of 64 values in x (lane L holds x = i0 + L with i0 a multiple
// of 64, so each row it loads or stores is two whole aligned 128-byte lines) by RY rows in y,
// and marches along z over a chunk of planes. At each plane it computes H's new values on the
// strip and on the row below it (the row's H is what E's back difference along y reads, and
// is recomputed here rather than exchanged), then E's new values on the strip. Neighbours
// along x move between lanes by DPP wave shifts; the two columns just outside the strip come
// from gathers, a handful of lanes loading one value each:
//   - E at x = i0 + 64 (Ey, Ez on the rows H is computed on): lane 63's i + 1 neighbour;
//   - old H and E at x = i0 - 1: lane m < RY recomputes Hy, Hz at (i0 - 1, j0 + m), lane 0's
//     i - 1 neighbour for E.
// Along z each wave keeps the plane below's new H in registers. Only the strip's own values
// are stored (through buffer stores whose offset is out of range elsewhere, so no lane
// branches around them), and every value is computed with the same operations in the same
// order as the reference's loops (src/fdtd/kernel/reference.rs:14-180), so any tiling gives
// the same bits.
//
// CPML: the variant AX compiles in the psi work of the axes in its bit mask (1 x, 2 y, 4 z); a
// strip and chunk runs in the variant of the slabs it meets. Every psi value a plane needs is
// loaded at the top of the plane with the fields, behind branches that are uniform over the
// wave (per plane for the z slabs, per row for the y slabs, per unit for the x slabs, whose
// lanes are fixed), so the slab work adds no round trip to memory.
#pragma once
#include "kernels.hpp"

namespace pxfdtd {

constexpr int SW = 64;  // strip width: one value per lane

template <int CTRL>
__device__ __forceinline__ int dpp_i(int old, int v) {
    return __builtin_amdgcn_update_dpp(old, v, CTRL, 0xf, 0xf, false);
}
template <int CTRL>
__device__ __forceinline__ float dpp(float old, float v) {
    return __builtin_bit_cast(float, dpp_i<CTRL>(__builtin_bit_cast(int, old), __builtin_bit_cast(int, v)));
}
template <int CTRL>
__device__ __forceinline__ double dpp(double old, double v) {
    const unsigned long long o = __builtin_bit_cast(unsigned long long, old);
    const unsigned long long x = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)dpp_i<CTRL>((int)(unsigned)o, (int)(unsigned)x);
    const unsigned hi = (unsigned)dpp_i<CTRL>((int)(unsigned)(o >> 32), (int)(unsigned)(x >> 32));
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
// lane L takes lane L + 1's v, lane 63 takes `fill` (DPP wave_shl:1, no write from beyond)
template <typename T>
__device__ __forceinline__ T next_lane(T v, T fill) { return dpp<0x130>(fill, v); }
// lane L takes lane L - 1's v, lane 0 takes `fill` (DPP wave_shr:1)
template <typename T>
__device__ __forceinline__ T prev_lane(T v, T fill) { return dpp<0x138>(fill, v); }

// lane l's value, the same for every lane
__device__ __forceinline__ float lane_value(float v, int l) {
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}
__device__ __forceinline__ double lane_value(double v, int l) {
    const unsigned long long x = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)x, l);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(x >> 32), l);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
template <typename P>
__device__ __forceinline__ P *uniform_ptr(P *p) {
    const unsigned long long x = (unsigned long long)p;
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)x);
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(x >> 32));
    return (P *)(((unsigned long long)hi << 32) | lo);
}

typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
// stores of the step's new fields and psi, which the next step reads a whole grid later:
// non-temporal (cache policy NT), so they do not displace rows still to be read this step
template <typename T>
__device__ __forceinline__ void bstore(T v, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff);
template <>
__device__ __forceinline__ void bstore<float>(float v, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)voff, (int)soff, 2);
}
template <>
__device__ __forceinline__ void bstore<double>(double v, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, (int)voff, (int)soff, 2);
}

// m on the grid: wrapped across a periodic axis (any distance), -1 beyond a wall
__device__ __forceinline__ int wrap_any(int m, int n, int per) {
    if (m >= 0 && m < n) return m;
    if (!per) return -1;
    const int q = m % n;
    return q < 0 ? q + n : q;
}

// What a workgroup stages in LDS before its wavefronts march: the cold parameters, the
// material table (uint8 ids) and the CPML's b and a. All four waves of a workgroup belong to
// one problem.
template <typename T>
struct MarchShared {
    StepCold<T> K;
    T cp[2048];  // the CPML's b and a (the SLABS variants use it)
};

template <typename T, int MODE, bool SLABS>
__device__ __forceinline__ void march64_stage(const StepArgs<T> &A, MarchShared<T> &S, T *tab,
                                              int8_t *grp) {
    const int tid = threadIdx.x;
    if (MODE == 0) {
        for (int q = tid; q < 256 * 6; q += 256) tab[q] = A.table[q];
        const int8_t *g = reinterpret_cast<const int8_t *>(A.table + 256 * 6);
        for (int q = tid; q < 256 * 3; q += 256) grp[q] = (A.flags & 32) ? int8_t(-1) : g[q];
    }
    {
        const int *src = reinterpret_cast<const int *>(A.cold);
        int *dst = reinterpret_cast<int *>(&S.K);
        for (int q = tid; q < (int)(sizeof(StepCold<T>) / 4); q += 256) dst[q] = src[q];
    }
    const int cbase = 2 * (A.nx + A.ny + A.nz);
    if (SLABS)
        for (int q = tid; q < A.ncpml && q < 2048; q += 256) S.cp[q] = A.coef[cbase + q];
}

// One wavefront's unit of the march: strip bx, row strip by, planes [kz0, kz1).
template <typename T, int MODE, int RY, int AX>
__device__ __forceinline__ void march64_unit(const StepArgs<T> &A, const int4 unit,
                                             const MarchShared<T> &S, const T *tab,
                                             const int8_t *grp) {
    constexpr int NR = RY + 2;  // rows of E held: j0 - 1 .. j0 + RY
    constexpr bool SLABS = AX != 0, XS = AX & 1, YS = AX & 2, ZS = AX & 4;
    const StepCold<T> &K = S.K;
    const T *cp = S.cp;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int cbase = 2 * (A.nx + A.ny + A.nz);
    const int bx = unit.x, by = unit.y, kz0 = unit.z, kz1 = unit.w;
    const int i0 = bx * SW, j0 = by * RY;
    const int t = (RY == A.tile_rows ? by : j0 / A.tile_rows) * A.tiles_x + bx;  // its injections' tile
    const int nx = A.nx, ny = A.ny, nz = A.nz, pitch = A.pitch;
    const int perx = A.flags & 1, pery = (A.flags >> 1) & 1, perz = (A.flags >> 2) & 1;
    bool injects = false;
    if (A.flags & 16) {
        const int32_t *io = K.inj_off + (long long)t * (nz + 1);
        injects = io[nz] > io[0];
    }

    // this lane's column, and the columns left and right of the strip (-1: beyond a wall)
    const int i = i0 + lane;
    const bool own_x = i < nx;
    const int xi = own_x ? i : (i == nx && perx ? 0 : -1);
    const int xl = wrap_any(i0 - 1, nx, perx);
    const int xr = i0 + SW < nx ? i0 + SW : (i0 + SW == nx && perx ? 0 : -1);
    int yj[NR];
    uint32_t vo[NR], so[NR];  // byte offsets in a plane: loads (any cell) and stores (own)
#pragma unroll
    for (int r = 0; r < NR; r++) {
        const int jr = j0 - 1 + r;
        yj[r] = wrap_any(jr, ny, pery);
        vo[r] = (xi < 0 || yj[r] < 0) ? OUTSIDE : (uint32_t)(yj[r] * pitch + xi) * sizeof(T);
        so[r] = (own_x && jr >= 0 && jr < ny) ? (uint32_t)(jr * pitch + i) * sizeof(T) : OUTSIDE;
    }
    // the gathers: lane m (< 32) takes (i0 - 1, j0 + m), lane 32 + m takes (i0 + 64, j0 - 1 + m)
    const int m = lane & 31;
    const int gyl = wrap_any(j0 + m, ny, pery), gyr = wrap_any(j0 - 1 + m, ny, pery);
    const bool lok = lane < 32 && m <= RY && xl >= 0 && gyl >= 0;
    const bool rok = lane >= 32 && m <= RY && xr >= 0 && gyr >= 0;
    const uint32_t lof = lok ? (uint32_t)(gyl * pitch + xl) * sizeof(T) : OUTSIDE;
    const uint32_t go_x = lof;                                        // Ex, m = 0..RY
    const uint32_t go_yz = (lok && m < RY) ? lof
                           : (rok ? (uint32_t)(gyr * pitch + xr) * sizeof(T) : OUTSIDE);
    const uint32_t go_h = (lok && m < RY) ? lof : OUTSIDE;           // old Hy, Hz
    const bool hok = lane < RY && xl >= 0 && gyl >= 0 && j0 + m < ny;  // a halo H this lane makes

    const uint32_t cbytes = (uint32_t)A.nval * sizeof(T), pbytes = (uint32_t)A.plane * sizeof(T);
    const __amdgpu_buffer_rsrc_t rE = rsrc(A.ein, 3 * cbytes), rH = rsrc(A.hin, 3 * cbytes),
                                 wE = rsrc(A.eout, 3 * cbytes), wH = rsrc(A.hout, 3 * cbytes),
                                 rid = rsrc(A.ids, MODE == 0 ? (uint32_t)A.nval : 0u),
                                 rnone = rsrc(A.ein, 0u);
    const T dt = A.dt;
    const T stx = A.step[0], sty = A.step[1], stz = A.step[2];
    const T *coef = A.coef;
    const int fex = 0, fhx = nx, fey = 2 * nx, fhy = 2 * nx + ny, fez = 2 * (nx + ny),
              fhz = 2 * (nx + ny) + nz;
    const T fxh = coef[fhx + max(xi, 0)], fxe = coef[fex + max(xi, 0)];
    const T fxh_l = coef[fhx + max(xl, 0)], fyh_l = coef[fhy + max(gyl, 0)];
    // the rows' difference factors, fixed for the unit: in registers, not reloaded each plane
    T fyh_r[RY + 1], fye_r[RY + 1];
#pragma unroll
    for (int r = 0; r <= RY; r++) {
        const int jr = wrap_any(j0 - 1 + r, ny, pery);
        fyh_r[r] = coef[fhy + max(jr, 0)];
        fye_r[r] = coef[fey + max(jr, 0)];
    }

    // CPML geometry fixed for the unit: the x slab per lane (and at the column left of the
    // strip), the y slab per row. psi moves through buffer loads and stores whose offset is
    // out of range where a lane has no psi value, so no lane branches around them.
    int lx = 0, hx = 0, ly = 0, hy = 0, lz = 0, hz = 0;
    int xs = -1, xcnt = 0, xsx = 0, xsl = -1, xlcnt = 0, xls = 0;
    int xb[2][2] = {{0, 0}, {0, 0}};  // x slab psi base per [field][component y, z] at this lane
    T bxh = 0, axh = 0, bxe = 0, axe = 0, bxl = 0, axl = 0;
    int ys[NR], yl[NR];
    bool xany = false;
    __amdgpu_buffer_rsrc_t rPi = rnone, rPo = rnone;
    const uint32_t vown = own_x ? (uint32_t)xi * sizeof(T) : OUTSIDE;
    if (SLABS) {
        lx = uniform(K.low[0]); hx = uniform(K.high[0]);
        ly = uniform(K.low[1]); hy = uniform(K.high[1]);
        lz = uniform(K.low[2]); hz = uniform(K.high[2]);
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
#pragma unroll
            for (int F = 0; F < 2; F++) {
                xb[F][0] = K.psi_off[F][1][0][xs] + xsx;
                xb[F][1] = K.psi_off[F][2][0][xs] + xsx;
            }
        }
        xany = (lx > 0 && i0 < lx) || (hx > 0 && min(i0 + SW, nx) > nx - hx);
        xsl = xl >= 0 ? side_of(xl, nx, lx, hx) : -1;
        xlcnt = xsl == 1 ? hx : lx;
        xls = xl - (xsl == 1 ? nx - hx : 0);
        if (xsl >= 0) {
            const int ch = K.cpml_off[0][xsl][1] - cbase;
            bxl = cp[ch + xls]; axl = cp[ch + xlcnt + xls];
        }
#pragma unroll
        for (int r = 0; r < NR; r++) {
            ys[r] = yj[r] >= 0 ? side_of(yj[r], ny, ly, hy) : -1;
            yl[r] = yj[r] - (ys[r] == 1 ? ny - hy : 0);
        }
    }

    auto zwrap = [&](int kk) { return wrap_any(kk, nz, perz); };
    const bool has_pro = kz0 > 0 || perz;
    const int kfirst = has_pro ? kz0 - 1 : kz0;

    struct EPlane {
        T e[NR][3];
        T gx, gy, gz;  // gathered: left column Ex (m = 0..RY), Ey, Ez (m < RY); right Ey, Ez
    };
    auto load_e = [&](EPlane &P, int kk) {
        const int kp = zwrap(kk);
        const __amdgpu_buffer_rsrc_t rr = kp >= 0 ? rE : rnone;
        const uint32_t s = kp >= 0 ? (uint32_t)kp * pbytes : 0u;
#pragma unroll
        for (int r = 0; r < NR; r++) {
            P.e[r][0] = bload<T>(rr, vo[r], s);
            P.e[r][1] = bload<T>(rr, vo[r], s + cbytes);
            P.e[r][2] = bload<T>(rr, vo[r], s + 2 * cbytes);
        }
        P.gx = bload<T>(rr, go_x, s);
        P.gy = bload<T>(rr, go_yz, s + cbytes);
        P.gz = bload<T>(rr, go_yz, s + 2 * cbytes);
    };

    // old H of a plane (own rows j0 - 1 .. j0 + RY - 1, the left column's Hy, Hz) and the
    // material ids of the strip's rows
    struct HPlane {
        T ho[RY + 1][3];
        T ghy, ghz;
        T fzh, fze;  // the plane's difference factors along z
        uint8_t id[RY];
    };
    auto load_h = [&](HPlane &P, int kk) {
        const int kp = zwrap(kk);
        const __amdgpu_buffer_rsrc_t rr = kp >= 0 ? rH : rnone;
        const __amdgpu_buffer_rsrc_t ri = kp >= 0 ? rid : rnone;
        const uint32_t s = kp >= 0 ? (uint32_t)kp * pbytes : 0u;
#pragma unroll
        for (int r = 0; r <= RY; r++) {
            P.ho[r][0] = bload<T>(rr, vo[r], s);
            P.ho[r][1] = r == 0 ? T(0) : bload<T>(rr, vo[r], s + cbytes);
            P.ho[r][2] = bload<T>(rr, vo[r], s + 2 * cbytes);
        }
        P.ghy = bload<T>(rr, go_h, s + cbytes);
        P.ghz = bload<T>(rr, go_h, s + 2 * cbytes);
        P.fzh = coef[fhz + max(kp, 0)];
        P.fze = coef[fez + max(kp, 0)];
#pragma unroll
        for (int r = 0; r < RY; r++) {
            if (MODE == 0) {
                const uint32_t o = so[r + 1] == OUTSIDE ? OUTSIDE : so[r + 1] / (uint32_t)sizeof(T);
                P.id[r] = __builtin_amdgcn_raw_buffer_load_b8(ri, (int)o, (int)(kp >= 0 ? (uint32_t)kp * (uint32_t)A.plane : 0u), 0);
            } else {
                P.id[r] = 0;
            }
        }
    };

    // The lean variant is software pipelined by a plane: the loads of E two planes up and of
    // old H one plane up are issued before the arithmetic of this plane, whose own were issued
    // a plane before. The slab variants, whose psi values take the registers that would need,
    // issue E one plane up and old H of this plane at the top of the plane.
    constexpr bool DEEP = AX == 0;
    T hp[RY][2];  // new Hx, Hy of the plane below, on the strip's rows
#pragma unroll
    for (int r = 0; r < RY; r++) hp[r][0] = hp[r][1] = T(0);

    // the arithmetic of plane kk from E of it and the plane above and old H of it
    // the arithmetic of plane kk; `ahead` issues loads for a later plane once this plane's psi
    // loads are issued, so that waiting for these does not wait for those
    auto step_plane = [&](const int kk, const EPlane &cur, const EPlane &nxt, const HPlane &hcur,
                          auto &&ahead) {
        const bool pro = kk < kz0;
        const int k = zwrap(kk);
        const uint32_t sk = (uint32_t)k * pbytes;
        const T (&ho)[RY + 1][3] = hcur.ho;
        const T ghy = hcur.ghy, ghz = hcur.ghz;
        const uint8_t (&id)[RY] = hcur.id;
        const T fzh = hcur.fzh, fze = hcur.fze;

        // this plane's psi values, loaded with the fields: per [row][component] for H (rows
        // j0 - 1 .. j0 + RY - 1) and E (the strip's rows), per axis; 0 where none
        int zs = -1, zsl = 0, zcnt = 0;
        T qx[RY + 1][2], qxe[RY][2], qy[RY + 1][2], qye[RY][2], qz[RY + 1][2], qze[RY][2], hq[4];
        // x slab psi byte offset of (field F, component y or z) on row r; out of range if none
        auto xo = [&](int F, int c, int r) -> uint32_t {
            return (xs >= 0 && yj[r] >= 0) ? (uint32_t)(xb[F][c] + (k * ny + yj[r]) * xcnt) * sizeof(T) : OUTSIDE;
        };
        auto yo = [&](int F, int c, int r) -> uint32_t {  // y slab: the row's first value
            const int cnt = ys[r] == 1 ? hy : ly;
            return (uint32_t)(uniform(K.psi_off[F][c][1][ys[r]]) + (k * cnt + yl[r]) * nx) * sizeof(T);
        };
        auto zo = [&](int F, int c, int r) -> uint32_t {  // z slab: the row's first value
            return (uint32_t)(uniform(K.psi_off[F][c][2][zs]) + (zsl * ny + yj[r]) * nx) * sizeof(T);
        };
        if (SLABS) {
            zs = side_of(k, nz, lz, hz);
            zsl = k - (zs == 1 ? nz - hz : 0);
            zcnt = zs == 1 ? hz : lz;
#pragma unroll
            for (int r = 0; r <= RY; r++) {
                qx[r][0] = qx[r][1] = qy[r][0] = qy[r][1] = qz[r][0] = qz[r][1] = T(0);
                if (r > 0) qxe[r - 1][0] = qxe[r - 1][1] = qye[r - 1][0] = qye[r - 1][1] = qze[r - 1][0] = qze[r - 1][1] = T(0);
            }
            hq[0] = hq[1] = hq[2] = hq[3] = T(0);
            if (XS && xany) {
#pragma unroll
                for (int r = 0; r <= RY; r++) {
                    if (r > 0) qx[r][0] = bload_once<T>(rPi, xo(1, 0, r), 0);
                    qx[r][1] = bload_once<T>(rPi, xo(1, 1, r), 0);
                    if (r > 0) {
                        qxe[r - 1][0] = bload_once<T>(rPi, xo(0, 0, r), 0);
                        qxe[r - 1][1] = bload_once<T>(rPi, xo(0, 1, r), 0);
                    }
                }
            }
#pragma unroll
            for (int r = 0; r <= RY; r++) {
                if (YS && ys[r] >= 0) {
                    qy[r][0] = bload_once<T>(rPi, vown, yo(1, 0, r));
                    qy[r][1] = bload_once<T>(rPi, vown, yo(1, 2, r));
                    if (r > 0) {
                        qye[r - 1][0] = bload_once<T>(rPi, vown, yo(0, 0, r));
                        qye[r - 1][1] = bload_once<T>(rPi, vown, yo(0, 2, r));
                    }
                }
                if (ZS && zs >= 0 && yj[r] >= 0) {
                    qz[r][0] = bload_once<T>(rPi, vown, zo(1, 0, r));
                    if (r > 0) {
                        qz[r][1] = bload_once<T>(rPi, vown, zo(1, 1, r));
                        qze[r - 1][0] = bload_once<T>(rPi, vown, zo(0, 0, r));
                        qze[r - 1][1] = bload_once<T>(rPi, vown, zo(0, 1, r));
                    }
                }
            }
            if (!pro) {  // the column left of the strip: Hy, Hz on lane m < RY (read only)
                if (XS && xsl >= 0) {
                    const int q = (k * ny + gyl) * xlcnt + xls;
                    hq[0] = bload_once<T>(rPi, hok ? (uint32_t)(uniform(K.psi_off[1][1][0][xsl]) + q) * sizeof(T) : OUTSIDE, 0);
                    hq[1] = bload_once<T>(rPi, hok ? (uint32_t)(uniform(K.psi_off[1][2][0][xsl]) + q) * sizeof(T) : OUTSIDE, 0);
                }
                const int sd = YS && hok ? side_of(gyl, ny, ly, hy) : -1;
                if (sd >= 0) {
                    const int cnt = sd == 1 ? hy : ly, s = gyl - (sd == 1 ? ny - hy : 0);
                    hq[2] = bload_once<T>(rPi, (uint32_t)(K.psi_off[1][2][1][sd] + (k * cnt + s) * nx + xl) * sizeof(T), 0);
                }
                if (ZS && zs >= 0)
                    hq[3] = bload_once<T>(rPi, hok ? (uint32_t)(uniform(K.psi_off[1][1][2][zs]) + (zsl * ny + gyl) * nx + xl) * sizeof(T) : OUTSIDE, 0);
            }
        }

        ahead();

        // new H on rows j0 - 1 .. j0 + RY - 1 (the first only Hx, Hz)
        T hn[RY + 1][3];
#pragma unroll
        for (int r = 0; r <= RY; r++) {
            const T ey_n = next_lane(cur.e[r][1], lane_value(cur.gy, 32 + r));
            const T ez_n = next_lane(cur.e[r][2], lane_value(cur.gz, 32 + r));
            const bool ok = xi >= 0 && yj[r] >= 0;
            const T fyh = fyh_r[r];
            const T dyx = cur.e[r + 1][2] - cur.e[r][2], dzx = nxt.e[r][1] - cur.e[r][1];  // Hx
            const T dzy = nxt.e[r][0] - cur.e[r][0], dxy = ez_n - cur.e[r][2];            // Hy
            const T dxz = ey_n - cur.e[r][1], dyz = cur.e[r + 1][0] - cur.e[r][0];        // Hz
            T v0 = ho[r][0] - dt * (dyx * fyh - dzx * fzh);
            T v1 = ho[r][1] - dt * (dzy * fzh - dxy * fxh);
            T v2 = ho[r][2] - dt * (dxz * fxh - dyz * fyh);
            if (SLABS) {
                // the strip's own rows write their psi; the row below's is only read
                const bool wr = !pro && r > 0 && j0 - 1 + r < ny;
                if (XS && xany) {
                    const bool in = xs >= 0;
                    if (r > 0) {
                        const T u = bxh * qx[r][0] + axh * (per_step(dxy, stx));
                        v1 = in ? v1 - dt * (T(-1) * u) : v1;
                        if (wr) bstore<T>(u, rPo, xo(1, 0, r), 0);
                    }
                    const T u = bxh * qx[r][1] + axh * (per_step(dxz, stx));
                    v2 = in ? v2 - dt * (T(1) * u) : v2;
                    if (wr) bstore<T>(u, rPo, xo(1, 1, r), 0);
                }
                if (YS && ys[r] >= 0) {
                    const int cnt = ys[r] == 1 ? hy : ly;
                    const int co = K.cpml_off[1][ys[r]][1] - cbase;
                    const T bb = cp[co + yl[r]], aa = cp[co + cnt + yl[r]];
                    const T u0 = bb * qy[r][0] + aa * (per_step(dyx, sty));
                    const T u2 = bb * qy[r][1] + aa * (per_step(dyz, sty));
                    v0 = v0 - dt * (T(1) * u0);
                    v2 = v2 - dt * (T(-1) * u2);
                    if (wr) {
                        bstore<T>(u0, rPo, vown, yo(1, 0, r));
                        bstore<T>(u2, rPo, vown, yo(1, 2, r));
                    }
                }
                if (ZS && zs >= 0 && yj[r] >= 0) {
                    const int co = K.cpml_off[2][zs][1] - cbase;
                    const T bb = cp[co + zsl], aa = cp[co + zcnt + zsl];
                    const T u0 = bb * qz[r][0] + aa * (per_step(dzx, stz));
                    v0 = v0 - dt * (T(-1) * u0);
                    if (wr) bstore<T>(u0, rPo, vown, zo(1, 0, r));
                    if (r > 0) {
                        const T u1 = bb * qz[r][1] + aa * (per_step(dzy, stz));
                        v1 = v1 - dt * (T(1) * u1);
                        if (wr) bstore<T>(u1, rPo, vown, zo(1, 1, r));
                    }
                }
            }
            if (injects && ok && !(pro && r == 0))
                inject<T, 1>(A, K, t, xi, yj[r], k, r == 0 ? 5 : 7, v0, v1, v2, T(0), T(0), T(0));
            hn[r][0] = ok ? v0 : T(0);
            hn[r][1] = ok ? v1 : T(0);
            hn[r][2] = ok ? v2 : T(0);
        }

        if (!pro) {
            // Hy, Hz at (i0 - 1, j0 + m) on lane m < RY, from the gathered values and lane 0's
            // own Ey, Ez of that row
            T ey0 = T(0), ez0 = T(0);
#pragma unroll
            for (int q = 0; q < RY; q++) {
                const T a = lane_value(cur.e[q + 1][1], 0), c = lane_value(cur.e[q + 1][2], 0);
                ey0 = m == q ? a : ey0;
                ez0 = m == q ? c : ez0;
            }
            const T exn = next_lane(cur.gx, T(0));  // Ex at (i0 - 1, j0 + m + 1)
            const T dzy = nxt.gx - cur.gx, dxy = ez0 - cur.gz;
            const T dxz = ey0 - cur.gy, dyz = exn - cur.gx;
            T v1 = ghy - dt * (dzy * fzh - dxy * fxh_l);
            T v2 = ghz - dt * (dxz * fxh_l - dyz * fyh_l);
            if (SLABS) {
                if (XS && xsl >= 0) {
                    v1 = v1 - dt * (T(-1) * (bxl * hq[0] + axl * (per_step(dxy, stx))));
                    v2 = v2 - dt * (T(1) * (bxl * hq[1] + axl * (per_step(dxz, stx))));
                }
                const int sd = YS && hok ? side_of(gyl, ny, ly, hy) : -1;
                if (sd >= 0) {
                    const int cnt = sd == 1 ? hy : ly, s = gyl - (sd == 1 ? ny - hy : 0);
                    const int co = K.cpml_off[1][sd][1] - cbase;
                    v2 = v2 - dt * (T(-1) * (cp[co + s] * hq[2] + cp[co + cnt + s] * (per_step(dyz, sty))));
                }
                if (ZS && zs >= 0) {
                    const int co = K.cpml_off[2][zs][1] - cbase;
                    v1 = v1 - dt * (T(1) * (cp[co + zsl] * hq[3] + cp[co + zcnt + zsl] * (per_step(dzy, stz))));
                }
            }
            T v0 = T(0);
            if (injects && hok) inject<T, 1>(A, K, t, xl, gyl, k, 6, v0, v1, v2, T(0), T(0), T(0));
            const T hy_h = hok ? v1 : T(0), hz_h = hok ? v2 : T(0);

            // new E on the strip
#pragma unroll
            for (int r = 1; r <= RY; r++) {
                const T hy_p = prev_lane(hn[r][1], lane_value(hy_h, r - 1));
                const T hz_p = prev_lane(hn[r][2], lane_value(hz_h, r - 1));
                const uint32_t st = so[r];
                const bool own = st != OUTSIDE;
                bstore<T>(hn[r][0], wH, st, sk);
                bstore<T>(hn[r][1], wH, st, sk + cbytes);
                bstore<T>(hn[r][2], wH, st, sk + 2 * cbytes);
                const int j = yj[r];
                const T dyx = hn[r][2] - hn[r - 1][2], dzx = hn[r][1] - hp[r - 1][1];  // Ex
                const T dzy2 = hn[r][0] - hp[r - 1][0], dxy2 = hn[r][2] - hz_p;        // Ey
                const T dxz2 = hn[r][1] - hy_p, dyz2 = hn[r][0] - hn[r - 1][0];        // Ez
                T ca0, ca1, ca2, cb0, cb1, cb2;
                int g0, g1, g2;
                const long long o = (long long)k * A.plane + (long long)(st / (uint32_t)sizeof(T));
                if (MODE == 0) {
                    const int q = id[r - 1];
                    const T *e = tab + 6 * q;
                    ca0 = e[0]; cb0 = e[1]; ca1 = e[2]; cb1 = e[3]; ca2 = e[4]; cb2 = e[5];
                    g0 = grp[3 * q]; g1 = grp[3 * q + 1]; g2 = grp[3 * q + 2];
                } else if (MODE == 2) {
                    const long long nv = A.nval;
                    cb0 = own ? A.cacb[o] : T(0);
                    cb1 = own ? A.cacb[nv + o] : T(0);
                    cb2 = own ? A.cacb[2 * nv + o] : T(0);
                    ca0 = lossless_ca(cb0); ca1 = lossless_ca(cb1); ca2 = lossless_ca(cb2);
                    g0 = g1 = g2 = -1;
                } else {
                    const long long nv = A.nval;
                    ca0 = own ? A.cacb[o] : T(0);
                    ca1 = own ? A.cacb[nv + o] : T(0);
                    ca2 = own ? A.cacb[2 * nv + o] : T(0);
                    cb0 = own ? A.cacb[3 * nv + o] : T(0);
                    cb1 = own ? A.cacb[4 * nv + o] : T(0);
                    cb2 = own ? A.cacb[5 * nv + o] : T(0);
                    g0 = own && A.grp ? A.grp[o] : -1;
                    g1 = own && A.grp ? A.grp[nv + o] : -1;
                    g2 = own && A.grp ? A.grp[2 * nv + o] : -1;
                }
                const T fye = fye_r[r];
                T v0 = ca0 * cur.e[r][0] + cb0 * (dyx * fye - dzx * fze);
                T w1 = ca1 * cur.e[r][1] + cb1 * (dzy2 * fze - dxy2 * fxe);
                T w2 = ca2 * cur.e[r][2] + cb2 * (dxz2 * fxe - dyz2 * fye);
                if (SLABS) {
                    const bool wr = j0 - 1 + r < ny;
                    if (XS && xany) {
                        const bool in = xs >= 0;
                        const T u1 = bxe * qxe[r - 1][0] + axe * (per_step(dxy2, stx));
                        const T u2 = bxe * qxe[r - 1][1] + axe * (per_step(dxz2, stx));
                        w1 = in ? w1 + cb1 * (T(-1) * u1) : w1;
                        w2 = in ? w2 + cb2 * (T(1) * u2) : w2;
                        if (wr) {
                            bstore<T>(u1, rPo, xo(0, 0, r), 0);
                            bstore<T>(u2, rPo, xo(0, 1, r), 0);
                        }
                    }
                    if (YS && ys[r] >= 0) {
                        const int cnt = ys[r] == 1 ? hy : ly;
                        const int co = K.cpml_off[1][ys[r]][0] - cbase;
                        const T bb = cp[co + yl[r]], aa = cp[co + cnt + yl[r]];
                        const T u0 = bb * qye[r - 1][0] + aa * (per_step(dyx, sty));
                        const T u2 = bb * qye[r - 1][1] + aa * (per_step(dyz2, sty));
                        v0 = v0 + cb0 * (T(1) * u0);
                        w2 = w2 + cb2 * (T(-1) * u2);
                        if (wr) {
                            bstore<T>(u0, rPo, vown, yo(0, 0, r));
                            bstore<T>(u2, rPo, vown, yo(0, 2, r));
                        }
                    }
                    if (ZS && zs >= 0 && j >= 0) {
                        const int co = K.cpml_off[2][zs][0] - cbase;
                        const T bb = cp[co + zsl], aa = cp[co + zcnt + zsl];
                        const T u0 = bb * qze[r - 1][0] + aa * (per_step(dzx, stz));
                        const T u1 = bb * qze[r - 1][1] + aa * (per_step(dzy2, stz));
                        v0 = v0 + cb0 * (T(-1) * u0);
                        w1 = w1 + cb1 * (T(1) * u1);
                        if (wr) {
                            bstore<T>(u0, rPo, vown, zo(0, 0, r));
                            bstore<T>(u1, rPo, vown, zo(0, 1, r));
                        }
                    }
                }
                if (injects && own) inject<T, 0>(A, K, t, xi, j, k, 7, v0, w1, w2, cb0, cb1, cb2);
                if (own && g0 >= 0) v0 = media_update(A, K, g0, 0, (int)o, v0, cb0, cur.e[r][0]);
                if (own && g1 >= 0) w1 = media_update(A, K, g1, 1, (int)o, w1, cb1, cur.e[r][1]);
                if (own && g2 >= 0) w2 = media_update(A, K, g2, 2, (int)o, w2, cb2, cur.e[r][2]);
                bstore<T>(v0, wE, st, sk);
                bstore<T>(w1, wE, st, sk + cbytes);
                bstore<T>(w2, wE, st, sk + 2 * cbytes);
            }
        }
#pragma unroll
        for (int r = 0; r < RY; r++) {
            hp[r][0] = hn[r + 1][0];
            hp[r][1] = hn[r + 1][1];
        }
    };

    if (DEEP) {
        EPlane cur, nxt;
        HPlane hcur;
        load_e(cur, kfirst);
        load_e(nxt, kfirst + 1);
        load_h(hcur, kfirst);
        for (int kk = kfirst; kk < kz1; kk++) {
            EPlane nxt2;
            HPlane hnxt;
            load_e(nxt2, kk + 2);
            load_h(hnxt, kk + 1);
            step_plane(kk, cur, nxt, hcur, [] {});
            cur = nxt;
            nxt = nxt2;
            hcur = hnxt;
        }
    } else {
        // old H a plane ahead, issued after the plane's psi loads
        EPlane cur, nxt;
        HPlane hcur;
        load_e(cur, kfirst);
        load_h(hcur, kfirst);
        for (int kk = kfirst; kk < kz1; kk++) {
            load_e(nxt, kk + 1);
            HPlane hnxt;
            step_plane(kk, cur, nxt, hcur, [&] { load_h(hnxt, kk + 1); });
            cur = nxt;
            hcur = hnxt;
        }
    }
}

template <typename T, int MODE, int RY, int AX>
__device__ __forceinline__ void march64_kernel(const StepArgs<T> &A) {
    __shared__ MarchShared<T> S;
    extern __shared__ unsigned char smem_raw[];
    T *tab = reinterpret_cast<T *>(smem_raw);                 // [256][3][2]
    int8_t *grp = reinterpret_cast<int8_t *>(tab + 256 * 6);  // [256][3]
    march64_stage<T, MODE, AX != 0>(A, S, tab, grp);
    __syncthreads();  // the only barrier: the parameters in LDS
    // workgroups are dealt round-robin to the 8 XCDs: each XCD takes a contiguous run of
    // units, so neighbouring strips (which read each other's edges) share an L2
    const int b = (int)(blockIdx.x % 8) * A.per_xcd + (int)(blockIdx.x / 8);
    const int w = uniform(b * 4 + (int)(threadIdx.x >> 6));
    if (w >= A.blocks) return;
    march64_unit<T, MODE, RY, AX>(A, A.units[w], S, tab, grp);
}

template <typename T, int MODE, int RY, int AX>
__global__ __launch_bounds__(256) void march64_step(StepArgs<T> A) {
    march64_kernel<T, MODE, RY, AX>(A);
}

// The same held to 128 VGPRs in f32, four waves per SIMD: the general slab variant at one row per
// strip, which needs 130 unbounded (three waves) and fits 128 without spilling
template <typename T, int MODE, int RY, int AX>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : 1))) void march64_step_4w(StepArgs<T> A) {
    march64_kernel<T, MODE, RY, AX>(A);
}


}  // namespace pxfdtd
