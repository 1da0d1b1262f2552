// Device code of the MI355X (gfx950) FDTD time-stepper: the fused Yee + CPML step, the taps
// that copy probe and transform values out, and the FP64 transform accumulation.
//
// The arithmetic of every value follows the reference's plain loops operation for operation
// (photonoxide src/fdtd/kernel/reference.rs:14-180: the curl as
// D_a(F_b) f_a - D_b(F_a) f_b, then each CPML slab in axis order x -> y -> z,
// psi <- b psi + a (difference / step), H -= dt (sign psi), E += cb (sign psi)), and it is
// compiled with -ffp-contract=off, so the f64 build gives the reference's bits and the f32
// build the same operations rounded to f32.
//
// Layout in HBM: each component of E and H is its own array [nz][ny][pitch] (i fastest,
// pitch = nx rounded up to 64 values, so a wave's 64 lanes read whole aligned rows), in two
// copies: a step reads copy p and writes copy 1 - p. The CPML's psi lives only in the slabs,
// one array per (field, component, axis, side) as the reference keeps it ([k][j][i] over the
// slab's box), also in two copies. E's coefficients are, where at most 256 distinct
// (ca, cb, medium) triples over the three components occur, one uint8 material id per cell
// into a table staged in LDS (1 byte per cell instead of 24); otherwise whole ca, cb arrays.
//
// One workgroup owns a tile of 64 x TY values in x and y (a wavefront per row, plus one
// wavefront for the halo) and marches along z over a chunk of planes. At each plane:
//   1. E of the next plane goes into an LDS ring of two planes over the tile and its halo
//      (one value around it), from registers loaded one plane ahead;
//   2. the loads for the plane after are issued, so HBM traffic overlaps the arithmetic;
//   3. H's new values over the tile, and over the row below and the column left of it (the
//      neighbours E's back differences read, recomputed here rather than exchanged), from the
//      E planes in LDS; the tile's own go to HBM, all to LDS;
//   4. E's new values over the tile from the new H in LDS and, along z, the new H of the
//      plane below kept in registers.
// The chunk's first plane computes H of the plane below it once (a prologue). Only the
// tile's own values are written, so the redundant halo work never races, and every value is
// the same bits whatever the tiling.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pxfdtd {

constexpr int TX = 64;  // tile width in x: one wavefront per row

struct Injection {
    int32_t k;          // plane
    int32_t i, j;       // position (wrapped into the grid)
    int32_t code;       // field * 4 + component
    int32_t emitter;
    int32_t pad;
    double a_re, a_im;  // complex amplitude: amount = Re(a w)
};

struct MediumGroup {
    int32_t poles;
    int32_t pad;
    double chi;
    double coef[4][3];  // alpha, xi, gamma per pole
};

// What a slab cell, an injection or a dispersive value needs: kept in device memory and read
// where used, so that it holds no scalar registers across the plane loop.
template <typename T>
struct StepCold {
    int low[3], high[3];      // CPML cells per axis side
    int cpml_off[3][2][2];    // b of [axis][side][field E=0,H=1] in coef; a follows, count each
    int psi_off[2][3][3][2];  // [field][component][axis][side], -1 if none
    const T *psi_in;
    T *psi_out;
    // injections: per tile a list sorted by plane, inj_off[tile * (nz + 1) + k]
    const int32_t *inj_off;
    const Injection *inj;
    const double *waves;      // the call's emitters' w, [steps][emitters] complex
    int emitters;
    // media (dense per-value state, read only where a value's group is set)
    const MediumGroup *groups;
    int max_poles;
    T *m_before;              // [3][nval]
    T *m_cur, *m_prev;        // [3][nval][max_poles]
    int psi_len;              // psi values per copy
    // plane waves' auxiliary fields, FP64: an injection with emitter < 0 is a coupling of a
    // plane wave's box and reads its incident field at aux[-1 - emitter]
    const double *aux;
};

// The kernel's arguments: what every plane reads.
template <typename T>
struct StepArgs {
    int nx, ny, nz, pitch;
    int plane;                // pitch * ny
    int nval;                 // values per component: a field's components are nval apart
    int flags;                // bit a: axis a periodic; bit 3: CPML slabs; bit 4: injections;
                              // bit 5: media in media_pass after the step kernel
    int tiles_x, tiles_y, kz;
    int tile_rows;            // rows of the tiles the injection lists are kept by (a march unit
                              // may be fewer rows: its tile is the one holding its first row)
    int blocks;               // tiles_x * tiles_y * chunks
    int per_xcd;              // blocks dealt to each XCD, contiguous in the tile order
    const T *ein, *hin;       // copy p, [3][nval] each
    T *eout, *hout;           // copy 1 - p
    // coefficients, mode 0: a uint8 material id per cell into a table staged in LDS,
    // [256][3][2] (ca, cb per component) then the media groups [256][3] as int8
    const uint8_t *ids;
    const T *table;
    // mode 1: ca then cb, [2][3][nval], and the media group per value [3][nval] (or null);
    // mode 2 (a lossless problem without media): cb alone, [3][nval], ca being 0 where cb is 0
    // (walls, conductors) and 1 elsewhere
    const T *cacb;
    const int8_t *grp;
    T dt;
    T step[3];                // f64: dx, dy, dz; f32: 1/dx, 1/dy, 1/dz (see per_step)
    // kappa/step at the nodes (fe) and halves (fh) along x, y, z: fe_x, fh_x, fe_y, fh_y,
    // fe_z, fh_z, n_axis values each; then the CPML's b and a
    const T *coef;
    int ncpml;                // CPML coefficients after the difference factors
    const int4 *units;        // register march: the strips and plane chunks of this launch
    const StepCold<T> *cold;  // the one for this step's parity
    int wave_step;            // this step's row of cold->waves
};

// The CPML's psi takes a field difference over the step. The reference's CPU loops divide
// (src/fdtd/kernel/reference.rs:102-180); its GPU kernel multiplies by the host's 1/step rounded
// to the working precision (src/fdtd/gpu/yee.wgsl update_h/update_e, with reals from
// src/fdtd/gpu/resident.rs:170). f64 here does the first (the CPU path's bits), f32 the second
// (one multiplication in place of an IEEE division): StepArgs::step holds the steps in f64 and
// their reciprocals in f32.
template <typename T>
__device__ __forceinline__ T per_step(T d, T s);
template <>
__device__ __forceinline__ float per_step<float>(float d, float s) { return d * s; }
template <>
__device__ __forceinline__ double per_step<double>(double d, double s) { return d / s; }

// A workgroup barrier that waits only for this wave's LDS operations, not for its global
// loads: the loads issued for the next plane stay in flight across it. (__syncthreads also
// waits for every outstanding global load, which would drain that prefetch.)
__device__ __forceinline__ void lds_barrier() {
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// psi's slab side for coordinate m along an axis: 0 low, 1 high, -1 none
__device__ __forceinline__ int side_of(int m, int n, int low, int high) {
    return m < low ? 0 : (m >= n - high ? 1 : -1);
}

// a coordinate wrapped onto the grid across a periodic side, or -1 beyond a wall
__device__ __forceinline__ int wrap(int m, int n, int per) {
    if (m >= 0 && m < n) return m;
    if (!per) return -1;
    return m < 0 ? m + n : m - n;
}

// Where a value (i, j, k) sits in the CPML slabs: per axis its side (-1: none), its index
// along the axis within the slab, and its index within the slab's psi box.
struct SlabAt {
    int side[3], s[3], q[3];
};

template <typename T>
__device__ __forceinline__ SlabAt slab_at(const StepArgs<T> &A, const StepCold<T> &K, int i,
                                          int j, int k) {
    SlabAt P;
    const int n[3] = {A.nx, A.ny, A.nz};
    const int m[3] = {i, j, k};
#pragma unroll
    for (int w = 0; w < 3; w++) {
        P.side[w] = side_of(m[w], n[w], K.low[w], K.high[w]);
        const int from = P.side[w] == 1 ? n[w] - K.high[w] : 0;
        const int cnt = P.side[w] == 1 ? K.high[w] : K.low[w];
        P.s[w] = m[w] - from;
    }
    const int cx = P.side[0] == 1 ? K.high[0] : K.low[0];
    const int cy = P.side[1] == 1 ? K.high[1] : K.low[1];
    P.q[0] = (k * A.ny + j) * cx + P.s[0];
    P.q[1] = (k * cy + P.s[1]) * A.nx + i;
    P.q[2] = (P.s[2] * A.ny + j) * A.nx + i;
    return P;
}

// psi <- b psi + a d for (field F, component C, axis W) at a slab value; written if asked
template <int F, int C, int W, typename T>
__device__ __forceinline__ T psi_at(const StepArgs<T> &A, const StepCold<T> &K, const SlabAt &P,
                                    T d, bool write) {
    const int side = P.side[W];
    const int off = (side == 1 ? K.psi_off[F][C][W][1] : K.psi_off[F][C][W][0]) + P.q[W];
    const int co = side == 1 ? K.cpml_off[W][1][F] : K.cpml_off[W][0][F];
    const int cnt = side == 1 ? K.high[W] : K.low[W];
    const T v = A.coef[co + P.s[W]] * K.psi_in[off] + A.coef[co + cnt + P.s[W]] * d;
    if (write) K.psi_out[off] = v;
    return v;
}

// One injection on component c of a value: a source value, H -= dt Re(a w) or
// E -= cb Re(a w) (src/fdtd.rs:1346-1395), or a plane wave's coupling, H -= (dt weight) F or
// E -= (cb weight) F with F its incident field (src/fdtd/sources.rs:971-990), both in the
// reference's association. Selects, not an indexed array: the three values stay in registers.
template <typename T, int FIELD>
__device__ __forceinline__ void apply_injection(const StepArgs<T> &A, const StepCold<T> &K,
                                                const Injection &in, int c, T &v0, T &v1, T &v2,
                                                T cb0, T cb1, T cb2, int ws) {
    if (in.emitter < 0) {
        const T f = T(K.aux[-1LL - in.emitter]);
        const T wt = T(in.a_re);
        if (FIELD == 1) {
            const T d = (A.dt * wt) * f;
            v0 = c == 0 ? v0 - d : v0;
            v1 = c == 1 ? v1 - d : v1;
            v2 = c == 2 ? v2 - d : v2;
        } else {
            v0 = c == 0 ? v0 - (cb0 * wt) * f : v0;
            v1 = c == 1 ? v1 - (cb1 * wt) * f : v1;
            v2 = c == 2 ? v2 - (cb2 * wt) * f : v2;
        }
        return;
    }
    const double *w = K.waves + 2 * ((long long)ws * K.emitters + in.emitter);
    const T amount = T(in.a_re * w[0] - in.a_im * w[1]);
    if (FIELD == 1) {
        const T d = A.dt * amount;
        v0 = c == 0 ? v0 - d : v0;
        v1 = c == 1 ? v1 - d : v1;
        v2 = c == 2 ? v2 - d : v2;
    } else {
        v0 = c == 0 ? v0 - cb0 * amount : v0;
        v1 = c == 1 ? v1 - cb1 * amount : v1;
        v2 = c == 2 ? v2 - cb2 * amount : v2;
    }
}

// the injections of plane k on the value (field, i, j) in `mask`, in the tile's list order
// the sources, currents and couplings of tile `tile`'s list at value (i, j, k), wave row ws
template <typename T, int FIELD>
__device__ __forceinline__ void inject_ws(const StepArgs<T> &A, const StepCold<T> &K, int tile, int ws, int i,
                                          int j, int k, int mask, T &v0, T &v1, T &v2, T cb0, T cb1,
                                          T cb2) {
    const int32_t *off = K.inj_off + (long long)tile * (A.nz + 1);
    const int lo = off[k], hi = off[k + 1];
    for (int q = lo; q < hi; q++) {
        const Injection in = K.inj[q];
        const int c = in.code & 3;
        if ((in.code >> 2) != FIELD || in.i != i || in.j != j || !((mask >> c) & 1)) continue;
        apply_injection<T, FIELD>(A, K, in, c, v0, v1, v2, cb0, cb1, cb2, ws);
    }
}
template <typename T, int FIELD>
__device__ __forceinline__ void inject(const StepArgs<T> &A, const StepCold<T> &K, int tile, int i,
                                       int j, int k, int mask, T &v0, T &v1, T &v2, T cb0, T cb1,
                                       T cb2) {
    inject_ws<T, FIELD>(A, K, tile, A.wave_step, i, j, k, mask, v0, v1, v2, cb0, cb1, cb2);
}

// the ADE rest of E's update and the polarization currents' step at one value
// (src/fdtd/media.rs:670-705); eold is E^n, the value saved before the update
template <typename T>
__device__ __forceinline__ T media_update(const StepArgs<T> &A, const StepCold<T> &K, int g, int c,
                                          int r, T v, T cb, T eold) {
    const MediumGroup &G = K.groups[g];
    const long long s = (long long)c * A.nval + r;
    const int mp = K.max_poles;
    T *m_cur = K.m_cur, *m_prev = K.m_prev;
    // the state's loads in a loop of their own, so that they issue together
    T cur[4], prv[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        if (q >= G.poles) break;
        cur[q] = m_cur[s * mp + q];
        prv[q] = m_prev[s * mp + q];
    }
    const T before = K.m_before[s];
    T rest = T(G.chi) * before;
#pragma unroll
    for (int q = 0; q < 4; q++) {
        if (q >= G.poles) break;
        const T alpha = T(G.coef[q][0]), xi = T(G.coef[q][1]);
        rest = rest - T(0.5) * ((alpha + T(1)) * cur[q] + xi * prv[q]);
    }
    const T next = v + cb * rest;
    const T slope = (next - before) / (T(2) * A.dt);
#pragma unroll
    for (int q = 0; q < 4; q++) {
        if (q >= G.poles) break;
        const T alpha = T(G.coef[q][0]), xi = T(G.coef[q][1]), gamma = T(G.coef[q][2]);
        const T jn = alpha * cur[q] + xi * prv[q] + gamma * slope;
        m_prev[s * mp + q] = cur[q];
        m_cur[s * mp + q] = jn;
    }
    K.m_before[s] = eold;
    return next;
}

// ca of a lossless problem's value from its cb (coefficient mode 2): 0 where cb is 0 (walls and
// conductors zero both, src/fdtd.rs:738-756, 995-1010), 1 elsewhere
template <typename T>
__device__ __forceinline__ T lossless_ca(T cb) { return cb == T(0) ? T(0) : T(1); }

// The dispersive values' update as a pass of its own after a step kernel that left it out
// (flags bit 5): the step kernel stored each value's curl update with its sources, E^n is
// still in the input copy, and this adds the media term and advances the state with
// media_update, the arithmetic the step kernels apply inline (src/fdtd/media.rs:670-705). One
// thread per value in a medium: `vals` holds c * nval + r, `vgrp` its group, `vcb` its cb.
template <typename T>
__global__ __launch_bounds__(256) void media_pass(StepArgs<T> A, const uint32_t *vals,
                                                  const int8_t *vgrp, const T *vcb, long long count) {
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= count) return;
    const StepCold<T> &K = *A.cold;
    const uint32_t s = vals[e];
    const uint32_t nv = (uint32_t)A.nval;
    const int c = s >= 2 * nv ? 2 : (s >= nv ? 1 : 0);
    const int r = (int)(s - (uint32_t)c * nv);
    A.eout[s] = media_update(A, K, vgrp[e], c, r, A.eout[s], vcb[e], A.ein[s]);
}

// What a thread loads one or two planes ahead: E of a plane (its own value and a halo
// value), old H of another (own, and the halo row's Hx, Hz and column's Hy, Hz) and the
// material id.
template <typename T>
struct Ring {
    T e0, e1, e2;       // own E
    T x0, x1, x2;       // halo E
    T h0, h1, h2;       // own old H
    T r0, r2, l1, l2;   // halo row Hx, Hz; halo column Hy, Hz
    uint8_t id;
};

template <typename T, int MODE, int TY>
__global__ __launch_bounds__(TX *(TY + 1)) void fused_step(StepArgs<T> A) {
    constexpr int W = TX + 2;   // LDS columns: i0 - 1 .. i0 + 64
    constexpr int HR = TY + 2;  // LDS rows: j0 - 1 .. j0 + TY
    constexpr int NT = TX * (TY + 1);
    constexpr int NHALO = 2 * W + 2 * TY;  // E positions around the tile
    static_assert(NHALO <= NT, "the halo has one thread per position");
    __shared__ T es[2][3][HR][W];        // E, two planes
    __shared__ T hs[3][TY + 1][TX + 1];  // new H: rows j0 - 1 .., columns i0 - 1 ..
    __shared__ StepCold<T> K;            // the slab, injection and media parameters
    extern __shared__ unsigned char smem_raw[];
    T *tab = reinterpret_cast<T *>(smem_raw);                 // [256][3][2]
    int8_t *grp = reinterpret_cast<int8_t *>(tab + 256 * 6);  // [256][3]

    const int lane = threadIdx.x, wave = threadIdx.y;
    const int tid = wave * TX + lane;
    if (MODE == 0) {
        for (int q = tid; q < 256 * 6; q += NT) tab[q] = A.table[q];
        const int8_t *g = reinterpret_cast<const int8_t *>(A.table + 256 * 6);
        for (int q = tid; q < 256 * 3; q += NT) grp[q] = (A.flags & 32) ? int8_t(-1) : g[q];
    }
    {
        const int *src = reinterpret_cast<const int *>(A.cold);
        int *dst = reinterpret_cast<int *>(&K);
        for (int q = tid; q < (int)(sizeof(StepCold<T>) / 4); q += NT) dst[q] = src[q];
    }

    // blocks are dealt round-robin to the 8 XCDs: give each XCD a contiguous run of tiles,
    // so neighbouring tiles (which read each other's edges) share an L2
    const int b = (int)(blockIdx.x % 8) * A.per_xcd + (int)(blockIdx.x / 8);
    if (b >= A.blocks) return;
    const int tiles = A.tiles_x * A.tiles_y;
    const int bz = b / tiles, t = b % tiles;
    const int by = t / A.tiles_x, bx = t % A.tiles_x;
    const int i0 = bx * TX, j0 = by * TY;
    const int kz0 = bz * A.kz;
    const int kz1 = min(A.nz, kz0 + A.kz);
    const int nx = A.nx, ny = A.ny, nz = A.nz, pitch = A.pitch, plane = A.plane;
    const int perx = A.flags & 1, pery = (A.flags >> 1) & 1, perz = (A.flags >> 2) & 1;
    const bool slabs = (A.flags & 8) != 0, injects = (A.flags & 16) != 0;

    // own value
    const bool own = wave < TY;
    const int oi = i0 + lane, oj = j0 + wave;
    const bool own_ok = own && oi < nx && oj < ny;
    const int orow = oj * pitch + oi;  // offset within a plane
    // halo E position this thread loads: row j0-1, row j0+TY, column i0-1, column i0+64
    int hrho = -1, hkap = -1, hoff = -1;
    if (tid < NHALO) {
        if (tid < W) { hrho = 0; hkap = tid; }
        else if (tid < 2 * W) { hrho = HR - 1; hkap = tid - W; }
        else if (tid < 2 * W + TY) { hrho = 1 + tid - 2 * W; hkap = 0; }
        else { hrho = 1 + tid - 2 * W - TY; hkap = W - 1; }
        const int wi = wrap(i0 - 1 + hkap, nx, perx);
        const int wj = wrap(j0 - 1 + hrho, ny, pery);
        hoff = (wi < 0 || wj < 0) ? -1 : wj * pitch + wi;
    }
    // own lanes beyond the grid hold the wrapped value (periodic sides) or zero
    int eoff = -1;  // where this own lane's LDS E value comes from
    if (own) {
        const int wi = wrap(oi, nx, perx), wj = wrap(oj, ny, pery);
        eoff = (wi < 0 || wj < 0) ? -1 : wj * pitch + wi;
    }
    // halo H positions this thread computes (halo wave): row j0-1 at i0+lane (Hz, Hx);
    // column i0-1 at j0+lane (Hz, Hy) for lane < TY
    const bool hw = wave == TY;
    const int hrj = wrap(j0 - 1, ny, pery);
    const bool hrow_ok = hw && oi < nx && hrj >= 0;
    const int hci = wrap(i0 - 1, nx, perx);
    const int hcj = j0 + lane;
    const bool hcol_ok = hw && lane < TY && hcj < ny && hci >= 0;
    const int hrow_off = hrow_ok ? hrj * pitch + oi : -1, hcol_off = hcol_ok ? hcj * pitch + hci : -1;
    const int own_off = own_ok ? orow : -1;

    auto zplane = [&](int kk) -> int { return wrap(kk, nz, perz); };
    // loads from a valid address always, the value zeroed beyond a wall: no branch per load
    auto ld = [&](const T *a, int kp, int off) -> T {
        const bool ok = kp >= 0 && off >= 0;
        const T v = a[ok ? kp * plane + off : 0];
        return ok ? v : T(0);
    };
    const int nv = A.nval;
    const T *ein = A.ein, *hin = A.hin;
    auto load_e = [&](Ring<T> &R, int kk) {
        const int kp = zplane(kk);
        R.e0 = ld(ein, kp, eoff); R.e1 = ld(ein + nv, kp, eoff); R.e2 = ld(ein + 2 * nv, kp, eoff);
        R.x0 = ld(ein, kp, hoff); R.x1 = ld(ein + nv, kp, hoff); R.x2 = ld(ein + 2 * nv, kp, hoff);
    };
    auto load_h = [&](Ring<T> &R, int kk) {
        const int kp = zplane(kk);
        if (own) {
            R.h0 = ld(hin, kp, own_off); R.h1 = ld(hin + nv, kp, own_off); R.h2 = ld(hin + 2 * nv, kp, own_off);
        } else {
            R.r0 = ld(hin, kp, hrow_off); R.r2 = ld(hin + 2 * nv, kp, hrow_off);
            R.l1 = ld(hin + nv, kp, hcol_off); R.l2 = ld(hin + 2 * nv, kp, hcol_off);
        }
    };
    auto load_id = [&](Ring<T> &R, int kk) {
        if (MODE == 0) {
            const int kp = zplane(kk);
            R.id = A.ids[own_ok && kp >= 0 ? kp * plane + orow : 0];
        }
    };

    const bool has_pro = kz0 > 0 || perz;
    const int kfirst = has_pro ? kz0 - 1 : kz0;

    // E of the first plane into LDS; the two rings take the planes after
    {
        Ring<T> R0;
        load_e(R0, kfirst);
        if (own) {
            es[kfirst & 1][0][wave + 1][lane + 1] = R0.e0;
            es[kfirst & 1][1][wave + 1][lane + 1] = R0.e1;
            es[kfirst & 1][2][wave + 1][lane + 1] = R0.e2;
        }
        if (tid < NHALO) {
            es[kfirst & 1][0][hrho][hkap] = R0.x0;
            es[kfirst & 1][1][hrho][hkap] = R0.x1;
            es[kfirst & 1][2][hrho][hkap] = R0.x2;
        }
    }
    Ring<T> RA, RB;  // RA: E(kk+1), H(kk), id(kk) for kk - kfirst even; RB for odd
    RA.e0 = RA.e1 = RA.e2 = RA.x0 = RA.x1 = RA.x2 = T(0);
    RA.h0 = RA.h1 = RA.h2 = RA.r0 = RA.r2 = RA.l1 = RA.l2 = T(0);
    RA.id = 0;
    RB = RA;
    load_e(RA, kfirst + 1);
    load_h(RA, kfirst);
    load_id(RA, kfirst);
    if (kfirst + 1 < kz1) {
        load_e(RB, kfirst + 2);
        load_h(RB, kfirst + 1);
        load_id(RB, kfirst + 1);
    }
    T ec0 = T(0), ec1 = T(0), ec2 = T(0);  // E of this plane, own (E^n)
    T hp0 = T(0), hp1 = T(0);               // new H of the plane below, own (x, y)
    const T dt = A.dt;
    const T stx = A.step[0], sty = A.step[1], stz = A.step[2];
    const T *coef = A.coef;
    const int fex = 0, fhx = nx, fey = 2 * nx, fhy = 2 * nx + ny, fez = 2 * (nx + ny),
              fhz = 2 * (nx + ny) + nz;
    __syncthreads();  // the parameters and the first plane in LDS
    if (own) {
        ec0 = es[kfirst & 1][0][wave + 1][lane + 1];
        ec1 = es[kfirst & 1][1][wave + 1][lane + 1];
        ec2 = es[kfirst & 1][2][wave + 1][lane + 1];
    }

    auto iteration = [&](const int kk, Ring<T> &R) {
        const bool pro = kk < kz0;
        const int k = zplane(kk);
        const int s0 = kk & 1, s1 = (kk + 1) & 1;
        // 1. E of the next plane into LDS, then the ring's E takes the plane 3 ahead
        if (own) {
            es[s1][0][wave + 1][lane + 1] = R.e0;
            es[s1][1][wave + 1][lane + 1] = R.e1;
            es[s1][2][wave + 1][lane + 1] = R.e2;
        }
        if (tid < NHALO) {
            es[s1][0][hrho][hkap] = R.x0;
            es[s1][1][hrho][hkap] = R.x1;
            es[s1][2][hrho][hkap] = R.x2;
        }
        if (kk + 2 < kz1) load_e(R, kk + 3);
        lds_barrier();
        // 2. new H over the tile and its halo
        const T fzh = coef[fhz + k];
        if (own_ok) {
            const int ka = lane + 1, ra = wave + 1;
            const T ex = es[s0][0][ra][ka], ey = es[s0][1][ra][ka], ez = es[s0][2][ra][ka];
            const T dyx = es[s0][2][ra + 1][ka] - ez, dzx = es[s1][1][ra][ka] - ey;  // Hx
            const T dzy = es[s1][0][ra][ka] - ex, dxy = es[s0][2][ra][ka + 1] - ez;  // Hy
            const T dxz = es[s0][1][ra][ka + 1] - ey, dyz = es[s0][0][ra + 1][ka] - ex;  // Hz
            const T fxh = coef[fhx + oi], fyh = coef[fhy + oj];
            T v0 = R.h0 - dt * (dyx * fyh - dzx * fzh);
            T v1 = R.h1 - dt * (dzy * fzh - dxy * fxh);
            T v2 = R.h2 - dt * (dxz * fxh - dyz * fyh);
            if (slabs) {
                const SlabAt P = slab_at(A, K, oi, oj, k);
                const bool wr = !pro;
                if (P.side[1] >= 0) v0 = v0 - dt * (T(1) * psi_at<1, 0, 1>(A, K, P, per_step(dyx, sty), wr));
                if (P.side[2] >= 0) v0 = v0 - dt * (T(-1) * psi_at<1, 0, 2>(A, K, P, per_step(dzx, stz), wr));
                if (P.side[0] >= 0) v1 = v1 - dt * (T(-1) * psi_at<1, 1, 0>(A, K, P, per_step(dxy, stx), wr));
                if (P.side[2] >= 0) v1 = v1 - dt * (T(1) * psi_at<1, 1, 2>(A, K, P, per_step(dzy, stz), wr));
                if (P.side[0] >= 0) v2 = v2 - dt * (T(1) * psi_at<1, 2, 0>(A, K, P, per_step(dxz, stx), wr));
                if (P.side[1] >= 0) v2 = v2 - dt * (T(-1) * psi_at<1, 2, 1>(A, K, P, per_step(dyz, sty), wr));
            }
            if (injects) inject<T, 1>(A, K, t, oi, oj, k, 7, v0, v1, v2, T(0), T(0), T(0));
            if (!pro) {
                const int r = k * plane + orow;
                A.hout[r] = v0;
                A.hout[nv + r] = v1;
                A.hout[2 * nv + r] = v2;
            }
            hs[0][wave + 1][lane + 1] = v0;
            hs[1][wave + 1][lane + 1] = v1;
            hs[2][wave + 1][lane + 1] = v2;
        }
        if (hw && !pro) {
            // the row below the tile: Hz and Hx at (i0 + lane, j0 - 1)
            if (hrow_ok) {
                const int ka = lane + 1, ra = 0;
                const T ex = es[s0][0][ra][ka], ey = es[s0][1][ra][ka], ez = es[s0][2][ra][ka];
                const T dyx = es[s0][2][ra + 1][ka] - ez, dzx = es[s1][1][ra][ka] - ey;
                const T dxz = es[s0][1][ra][ka + 1] - ey, dyz = es[s0][0][ra + 1][ka] - ex;
                const T fxh = coef[fhx + oi], fyh = coef[fhy + hrj];
                T v0 = R.r0 - dt * (dyx * fyh - dzx * fzh);
                T v2 = R.r2 - dt * (dxz * fxh - dyz * fyh);
                if (slabs) {
                    const SlabAt P = slab_at(A, K, oi, hrj, k);
                    if (P.side[1] >= 0) v0 = v0 - dt * (T(1) * psi_at<1, 0, 1>(A, K, P, per_step(dyx, sty), false));
                    if (P.side[2] >= 0) v0 = v0 - dt * (T(-1) * psi_at<1, 0, 2>(A, K, P, per_step(dzx, stz), false));
                    if (P.side[0] >= 0) v2 = v2 - dt * (T(1) * psi_at<1, 2, 0>(A, K, P, per_step(dxz, stx), false));
                    if (P.side[1] >= 0) v2 = v2 - dt * (T(-1) * psi_at<1, 2, 1>(A, K, P, per_step(dyz, sty), false));
                }
                T v1 = T(0);
                if (injects) inject<T, 1>(A, K, t, oi, hrj, k, 5, v0, v1, v2, T(0), T(0), T(0));
                hs[0][0][lane + 1] = v0;
                hs[2][0][lane + 1] = v2;
            } else {
                hs[0][0][lane + 1] = T(0);
                hs[2][0][lane + 1] = T(0);
            }
            // the column left of the tile: Hz and Hy at (i0 - 1, j0 + lane)
            if (hcol_ok) {
                const int ka = 0, ra = lane + 1;
                const T ex = es[s0][0][ra][ka], ey = es[s0][1][ra][ka], ez = es[s0][2][ra][ka];
                const T dzy = es[s1][0][ra][ka] - ex, dxy = es[s0][2][ra][ka + 1] - ez;
                const T dxz = es[s0][1][ra][ka + 1] - ey, dyz = es[s0][0][ra + 1][ka] - ex;
                const T fxh = coef[fhx + hci], fyh = coef[fhy + hcj];
                T v1 = R.l1 - dt * (dzy * fzh - dxy * fxh);
                T v2 = R.l2 - dt * (dxz * fxh - dyz * fyh);
                if (slabs) {
                    const SlabAt P = slab_at(A, K, hci, hcj, k);
                    if (P.side[0] >= 0) v1 = v1 - dt * (T(-1) * psi_at<1, 1, 0>(A, K, P, per_step(dxy, stx), false));
                    if (P.side[2] >= 0) v1 = v1 - dt * (T(1) * psi_at<1, 1, 2>(A, K, P, per_step(dzy, stz), false));
                    if (P.side[0] >= 0) v2 = v2 - dt * (T(1) * psi_at<1, 2, 0>(A, K, P, per_step(dxz, stx), false));
                    if (P.side[1] >= 0) v2 = v2 - dt * (T(-1) * psi_at<1, 2, 1>(A, K, P, per_step(dyz, sty), false));
                }
                T v0 = T(0);
                if (injects) inject<T, 1>(A, K, t, hci, hcj, k, 6, v0, v1, v2, T(0), T(0), T(0));
                hs[1][lane + 1][0] = v1;
                hs[2][lane + 1][0] = v2;
            } else if (lane < TY) {
                hs[1][lane + 1][0] = T(0);
                hs[2][lane + 1][0] = T(0);
            }
        }
        const uint8_t idc = R.id;
        if (kk + 2 < kz1) load_h(R, kk + 2);
        lds_barrier();
        // 3. new E over the tile
        if (own_ok && !pro) {
            const int ka = lane + 1, ra = wave + 1;
            const T hx = hs[0][ra][ka], hy = hs[1][ra][ka], hz = hs[2][ra][ka];
            const T dyx = hz - hs[2][ra - 1][ka], dzx = hy - hp1;                 // Ex
            const T dzy = hx - hp0, dxy = hz - hs[2][ra][ka - 1];                 // Ey
            const T dxz = hy - hs[1][ra][ka - 1], dyz = hx - hs[0][ra - 1][ka];   // Ez
            T ca0, ca1, ca2, cb0, cb1, cb2;
            int g0, g1, g2;
            const int r = k * plane + orow;
            if (MODE == 0) {
                const T *e = tab + 6 * idc;
                ca0 = e[0]; cb0 = e[1]; ca1 = e[2]; cb1 = e[3]; ca2 = e[4]; cb2 = e[5];
                g0 = grp[3 * idc]; g1 = grp[3 * idc + 1]; g2 = grp[3 * idc + 2];
            } else if (MODE == 2) {
                cb0 = A.cacb[r]; cb1 = A.cacb[nv + r]; cb2 = A.cacb[2 * nv + r];
                ca0 = lossless_ca(cb0); ca1 = lossless_ca(cb1); ca2 = lossless_ca(cb2);
                g0 = g1 = g2 = -1;
            } else {
                ca0 = A.cacb[r]; ca1 = A.cacb[nv + r]; ca2 = A.cacb[2 * nv + r];
                cb0 = A.cacb[3 * nv + r]; cb1 = A.cacb[4 * nv + r]; cb2 = A.cacb[5 * nv + r];
                g0 = A.grp ? A.grp[r] : -1;
                g1 = A.grp ? A.grp[nv + r] : -1;
                g2 = A.grp ? A.grp[2 * nv + r] : -1;
            }
            const T fxe = coef[fex + oi], fye = coef[fey + oj], fze = coef[fez + k];
            T v0 = ca0 * ec0 + cb0 * (dyx * fye - dzx * fze);
            T v1 = ca1 * ec1 + cb1 * (dzy * fze - dxy * fxe);
            T v2 = ca2 * ec2 + cb2 * (dxz * fxe - dyz * fye);
            if (slabs) {
                const SlabAt P = slab_at(A, K, oi, oj, k);
                if (P.side[1] >= 0) v0 = v0 + cb0 * (T(1) * psi_at<0, 0, 1>(A, K, P, per_step(dyx, sty), true));
                if (P.side[2] >= 0) v0 = v0 + cb0 * (T(-1) * psi_at<0, 0, 2>(A, K, P, per_step(dzx, stz), true));
                if (P.side[0] >= 0) v1 = v1 + cb1 * (T(-1) * psi_at<0, 1, 0>(A, K, P, per_step(dxy, stx), true));
                if (P.side[2] >= 0) v1 = v1 + cb1 * (T(1) * psi_at<0, 1, 2>(A, K, P, per_step(dzy, stz), true));
                if (P.side[0] >= 0) v2 = v2 + cb2 * (T(1) * psi_at<0, 2, 0>(A, K, P, per_step(dxz, stx), true));
                if (P.side[1] >= 0) v2 = v2 + cb2 * (T(-1) * psi_at<0, 2, 1>(A, K, P, per_step(dyz, sty), true));
            }
            if (injects) inject<T, 0>(A, K, t, oi, oj, k, 7, v0, v1, v2, cb0, cb1, cb2);
            if (g0 >= 0) v0 = media_update(A, K, g0, 0, r, v0, cb0, ec0);
            if (g1 >= 0) v1 = media_update(A, K, g1, 1, r, v1, cb1, ec1);
            if (g2 >= 0) v2 = media_update(A, K, g2, 2, r, v2, cb2, ec2);
            A.eout[r] = v0;
            A.eout[nv + r] = v1;
            A.eout[2 * nv + r] = v2;
        }
        if (kk + 2 < kz1) load_id(R, kk + 2);
        if (own) {
            hp0 = hs[0][wave + 1][lane + 1];
            hp1 = hs[1][wave + 1][lane + 1];
            ec0 = es[s1][0][wave + 1][lane + 1];
            ec1 = es[s1][1][wave + 1][lane + 1];
            ec2 = es[s1][2][wave + 1][lane + 1];
        }
    };

    for (int kk = kfirst; kk < kz1; kk += 2) {
        iteration(kk, RA);
        if (kk + 1 < kz1) iteration(kk + 1, RB);
    }
}

// ---------------------------------------------------------------------------------------
// The register-march kernel: no LDS exchange and no barrier in the plane loop.
//
// Each wavefront owns a strip of RW = 62 values in x by RY rows in y and marches along z.
// Lane L holds x = i0 - 1 + L: lane 0 is the column left of the strip (its H is the i - 1
// neighbour of lane 1's E update), lane 63 the column right of it (its E is the i + 1
// neighbour of lane 62's H update); lanes 1..62 are the strip's own values. Each lane holds
// RY + 2 rows of E (the row below the strip, the strip, the row above), computes H on RY + 1
// rows (the row below and the strip) and E on the strip. Neighbours along x come from the
// adjacent lane (a cross-lane shuffle), along y from the thread's own registers, along z
// from the plane before (kept) and after (loaded). The waves of a workgroup are independent:
// the chip stays busy from many waves in flight rather than from a synchronized pipeline.

// The CPML updates of one value's components in one half step, split in two: psi_load issues
// every psi load of the value (from the top of the plane, with the field loads, so their
// latencies overlap), psi_finish updates and stores them. Combos in the reference's order per
// component (axis x -> y -> z): (Cx, y) (Cx, z) (Cy, x) (Cy, z) (Cz, x) (Cz, y). `cp` is the
// CPML coefficient table staged in LDS (b then a, count each, per [axis][side][field] at
// K.cpml_off minus `cbase`).
template <int F, typename T>
__device__ __forceinline__ void psi_load(const StepCold<T> &K, const SlabAt &P, int mask, T old[6]) {
    constexpr int CC[6] = {0, 0, 1, 1, 2, 2}, WW[6] = {1, 2, 0, 2, 0, 1};
#pragma unroll
    for (int n = 0; n < 6; n++) {
        const int c = CC[n], w = WW[n];
        const int side = P.side[w];
        const bool on = side >= 0 && ((mask >> c) & 1);
        const int off = (side == 1 ? K.psi_off[F][c][w][1] : K.psi_off[F][c][w][0]) + P.q[w];
        old[n] = on ? K.psi_in[off] : T(0);
    }
}

template <int F, typename T>
__device__ __forceinline__ void psi_finish(const StepCold<T> &K, const T *cp, int cbase,
                                           const SlabAt &P, int mask, const T old[6], T d_yx,
                                           T d_zx, T d_xy, T d_zy, T d_xz, T d_yz, T stx, T sty,
                                           T stz, bool write, T u[6]) {
    constexpr int CC[6] = {0, 0, 1, 1, 2, 2}, WW[6] = {1, 2, 0, 2, 0, 1};
    const T d[6] = {per_step(d_yx, sty), per_step(d_zx, stz), per_step(d_xy, stx), per_step(d_zy, stz), per_step(d_xz, stx), per_step(d_yz, sty)};
#pragma unroll
    for (int n = 0; n < 6; n++) {
        const int c = CC[n], w = WW[n];
        const int side = P.side[w];
        const bool on = side >= 0 && ((mask >> c) & 1);
        const int co = (side == 1 ? K.cpml_off[w][1][F] : K.cpml_off[w][0][F]) - cbase;
        const int cnt = side == 1 ? K.high[w] : K.low[w];
        u[n] = on ? cp[co + P.s[w]] * old[n] + cp[co + cnt + P.s[w]] * d[n] : T(0);
        if (write && on) {
            const int off = (side == 1 ? K.psi_off[F][c][w][1] : K.psi_off[F][c][w][0]) + P.q[w];
            K.psi_out[off] = u[n];
        }
    }
}

// Raw buffer access (CDNA's MUBUF): a 128-bit resource in scalar registers holds the base and
// the size in bytes; a lane's 32-bit byte offset at or beyond the size reads 0. Wall
// neighbours and planes beyond a wall then load as zero with no branch and no 64-bit address
// arithmetic per load.
constexpr uint32_t OUTSIDE = 0x80000000u;  // an offset beyond every array (each < 2 GiB)

__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc(const void *base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), 0, (int)bytes, 0x00020000);
}
template <typename T>
__device__ __forceinline__ T bload(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff);
template <>
__device__ __forceinline__ float bload<float>(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}
template <>
__device__ __forceinline__ double bload<double>(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0));
}
// a load of a value read once per step (CPML psi): non-temporal (cache policy NT), so it does
// not displace the field rows the neighbouring strips read again
template <typename T>
__device__ __forceinline__ T bload_once(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff);
template <>
__device__ __forceinline__ float bload_once<float>(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 2));
}
template <>
__device__ __forceinline__ double bload_once<double>(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 2));
}

constexpr int RW = 62;

template <typename T>
__device__ __forceinline__ T from_next_lane(T v) { return __shfl_down(v, 1, 64); }
template <typename T>
__device__ __forceinline__ T from_prev_lane(T v) { return __shfl_up(v, 1, 64); }

// SLABS: the strips and plane chunks that touch a CPML slab run in this kernel with the psi
// updates compiled in; the others in one without them, lean in registers. PF: the loads of
// the next plane are issued before this plane's arithmetic (more bytes in flight per wave,
// at the cost of registers).
template <typename T, int MODE, int RY, bool SLABS, bool PF>
__global__ __launch_bounds__(256) void march_step(StepArgs<T> A) {
    constexpr int NR = RY + 2;  // rows of E held: j0 - 1 .. j0 + RY
    constexpr int CPML_MAX = 2048;
    __shared__ StepCold<T> K;
    __shared__ T cp[SLABS ? CPML_MAX : 1];  // the CPML's b and a
    extern __shared__ unsigned char smem_raw[];
    T *tab = reinterpret_cast<T *>(smem_raw);                 // [256][3][2]
    int8_t *grp = reinterpret_cast<int8_t *>(tab + 256 * 6);  // [256][3]
    const int tid = threadIdx.x;
    if (MODE == 0) {
        for (int q = tid; q < 256 * 6; q += 256) tab[q] = A.table[q];
        const int8_t *g = reinterpret_cast<const int8_t *>(A.table + 256 * 6);
        for (int q = tid; q < 256 * 3; q += 256) grp[q] = (A.flags & 32) ? int8_t(-1) : g[q];
    }
    {
        const int *src = reinterpret_cast<const int *>(A.cold);
        int *dst = reinterpret_cast<int *>(&K);
        for (int q = tid; q < (int)(sizeof(StepCold<T>) / 4); q += 256) dst[q] = src[q];
    }
    const int cbase = 2 * (A.nx + A.ny + A.nz);
    if (SLABS)
        for (int q = tid; q < A.ncpml && q < CPML_MAX; q += 256) cp[q] = A.coef[cbase + q];
    __syncthreads();  // the only barrier: the parameters in LDS

    const int lane = tid & 63;
    // workgroups are dealt round-robin to the 8 XCDs: each XCD takes a contiguous run of
    // units, so neighbouring strips (which read each other's edges) share an L2
    const int b = (int)(blockIdx.x % 8) * A.per_xcd + (int)(blockIdx.x / 8);
    const int w = b * 4 + (tid >> 6);
    if (w >= A.blocks) return;
    const int4 unit = A.units[w];  // strip x, strip y, first plane, end plane
    const int bx = unit.x, by = unit.y, kz0 = unit.z, kz1 = unit.w;
    const int t = by * A.tiles_x + bx;
    const int i0 = bx * RW, j0 = by * RY;
    const int nx = A.nx, ny = A.ny, nz = A.nz, pitch = A.pitch, plane = A.plane, nv = A.nval;
    const int perx = A.flags & 1, pery = (A.flags >> 1) & 1, perz = (A.flags >> 2) & 1;
    // injections: only strips whose list holds any
    bool injects = false;
    if (A.flags & 16) {
        const int32_t *io = K.inj_off + (long long)t * (nz + 1);
        injects = io[nz] > io[0];
    }

    const int i = i0 - 1 + lane;
    const int xi = wrap(i, nx, perx);  // -1 beyond a wall
    const bool own_x = lane >= 1 && lane <= RW && i < nx;
    int yj[NR], off[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) {
        yj[r] = wrap(j0 - 1 + r, ny, pery);
        off[r] = (xi < 0 || yj[r] < 0) ? -1 : yj[r] * pitch + xi;
    }
    auto zplane = [&](int kk) -> int { return wrap(kk, nz, perz); };
    // byte offsets within a plane, OUTSIDE beyond a wall; the plane's offset rides in soffset
    uint32_t vo[NR];
#pragma unroll
    for (int r = 0; r < NR; r++) vo[r] = off[r] >= 0 ? (uint32_t)off[r] * sizeof(T) : OUTSIDE;
    const uint32_t cbytes = (uint32_t)nv * sizeof(T);
    const __amdgpu_buffer_rsrc_t rex = rsrc(A.ein, cbytes), rey = rsrc(A.ein + nv, cbytes),
                                 rez = rsrc(A.ein + 2 * nv, cbytes), rhx = rsrc(A.hin, cbytes),
                                 rhy = rsrc(A.hin + nv, cbytes), rhz = rsrc(A.hin + 2 * nv, cbytes),
                                 rid = rsrc(A.ids, MODE == 0 ? (uint32_t)nv : 0u),
                                 rnone = rsrc(A.ein, 0u);
    const uint32_t pbytes = (uint32_t)plane * sizeof(T);
    // a plane beyond a wall reads through an empty resource: zeros
    auto pick = [&](__amdgpu_buffer_rsrc_t r, int kp) { return kp >= 0 ? r : rnone; };
    auto soff = [&](int kp) { return kp >= 0 ? (uint32_t)kp * pbytes : 0u; };
    const T dt = A.dt;
    const T stx = A.step[0], sty = A.step[1], stz = A.step[2];
    const T *coef = A.coef;
    const int fex = 0, fhx = nx, fey = 2 * nx, fhy = 2 * nx + ny, fez = 2 * (nx + ny),
              fhz = 2 * (nx + ny) + nz;
    const T fxh = coef[fhx + max(xi, 0)], fxe = coef[fex + max(xi, 0)];

    const bool has_pro = kz0 > 0 || perz;
    const int kfirst = has_pro ? kz0 - 1 : kz0;

    // what a plane kk loads: E of plane kk + 1, old H and the ids of plane kk
    struct Loads {
        T en[NR][3];
        T ho[RY + 1][3];
        uint8_t id[RY];
    };
    auto load = [&](Loads &L, int kk) {
        const int k = zplane(kk), k1 = zplane(kk + 1);
        const uint32_t s1 = soff(k1), s0 = soff(k);
        const __amdgpu_buffer_rsrc_t ax = pick(rex, k1), ay = pick(rey, k1), az = pick(rez, k1);
        const __amdgpu_buffer_rsrc_t hx = pick(rhx, k), hy = pick(rhy, k), hz = pick(rhz, k);
#pragma unroll
        for (int r = 0; r < NR; r++) {
            L.en[r][0] = bload<T>(ax, vo[r], s1);
            L.en[r][1] = bload<T>(ay, vo[r], s1);
            L.en[r][2] = bload<T>(az, vo[r], s1);
        }
#pragma unroll
        for (int r = 0; r <= RY; r++) {
            L.ho[r][0] = bload<T>(hx, vo[r], s0);
            L.ho[r][1] = r == 0 ? T(0) : bload<T>(hy, vo[r], s0);
            L.ho[r][2] = bload<T>(hz, vo[r], s0);
        }
        if (MODE == 0) {
            const __amdgpu_buffer_rsrc_t ri = pick(rid, k);
            const uint32_t si = k >= 0 ? (uint32_t)k * (uint32_t)plane : 0u;
#pragma unroll
            for (int r = 0; r < RY; r++)
                L.id[r] = __builtin_amdgcn_raw_buffer_load_b8(
                    ri, (int)(off[r + 1] >= 0 ? (uint32_t)off[r + 1] : OUTSIDE), (int)si, 0);
        } else {
#pragma unroll
            for (int r = 0; r < RY; r++) L.id[r] = 0;
        }
    };

    T ek[NR][3];  // E of this plane
    {
        const int kp = zplane(kfirst);
        const uint32_t s0 = soff(kp);
        const __amdgpu_buffer_rsrc_t ax = pick(rex, kp), ay = pick(rey, kp), az = pick(rez, kp);
#pragma unroll
        for (int r = 0; r < NR; r++) {
            ek[r][0] = bload<T>(ax, vo[r], s0);
            ek[r][1] = bload<T>(ay, vo[r], s0);
            ek[r][2] = bload<T>(az, vo[r], s0);
        }
    }
    T hp[RY][2];  // new Hx, Hy of the plane below, on the strip's rows
#pragma unroll
    for (int r = 0; r < RY; r++) hp[r][0] = hp[r][1] = T(0);
    Loads nxt;
    if (PF) load(nxt, kfirst);

    for (int kk = kfirst; kk < kz1; kk++) {
        const bool pro = kk < kz0;
        const int k = zplane(kk);
        Loads L;
        if (PF) {
            L = nxt;
            if (kk + 1 < kz1) load(nxt, kk + 1);
        } else {
            load(L, kk);
        }
        // the psi values this plane updates, loaded with the fields
        SlabAt PH[RY + 1];
        T psh[RY + 1][6], pse[RY][6];
        if (SLABS) {
#pragma unroll
            for (int r = 0; r <= RY; r++) {
                PH[r] = slab_at(A, K, max(xi, 0), max(yj[r], 0), k);
                const bool ok = off[r] >= 0 && lane < 63;
                psi_load<1>(K, PH[r], ok ? (r == 0 ? 5 : 7) : 0, psh[r]);
                if (r > 0) psi_load<0>(K, PH[r], own_x && j0 - 1 + r < ny ? 7 : 0, pse[r - 1]);
            }
        }
        const T fzh = coef[fhz + k], fze = coef[fez + k];
        // new H on rows j0 - 1 .. j0 + RY - 1
        T hn[RY + 1][3];
#pragma unroll
        for (int r = 0; r <= RY; r++) {
            const T ey_n = from_next_lane(ek[r][1]);
            const T ez_n = from_next_lane(ek[r][2]);
            const int j = yj[r];
            const bool ok = off[r] >= 0 && lane < 63;
            const T fyh = coef[fhy + max(j, 0)];
            const T dyx = ek[r + 1][2] - ek[r][2], dzx = L.en[r][1] - ek[r][1];  // Hx
            const T dzy = L.en[r][0] - ek[r][0], dxy = ez_n - ek[r][2];          // Hy
            const T dxz = ey_n - ek[r][1], dyz = ek[r + 1][0] - ek[r][0];        // Hz
            T v0 = L.ho[r][0] - dt * (dyx * fyh - dzx * fzh);
            T v1 = L.ho[r][1] - dt * (dzy * fzh - dxy * fxh);
            T v2 = L.ho[r][2] - dt * (dxz * fxh - dyz * fyh);
            if (SLABS) {
                const SlabAt &P = PH[r];
                const int mask = ok ? (r == 0 ? 5 : 7) : 0;
                const bool wr = !pro && r > 0 && own_x && j0 - 1 + r < ny;
                T u[6];
                psi_finish<1>(K, cp, cbase, P, mask, psh[r], dyx, dzx, dxy, dzy, dxz, dyz, stx, sty,
                              stz, wr, u);
                if (P.side[1] >= 0) v0 = v0 - dt * (T(1) * u[0]);
                if (P.side[2] >= 0) v0 = v0 - dt * (T(-1) * u[1]);
                if (r > 0 && P.side[0] >= 0) v1 = v1 - dt * (T(-1) * u[2]);
                if (r > 0 && P.side[2] >= 0) v1 = v1 - dt * (T(1) * u[3]);
                if (P.side[0] >= 0) v2 = v2 - dt * (T(1) * u[4]);
                if (P.side[1] >= 0) v2 = v2 - dt * (T(-1) * u[5]);
            }
            if (injects && ok) inject<T, 1>(A, K, t, xi, j, k, r == 0 ? 5 : 7, v0, v1, v2, T(0), T(0), T(0));
            hn[r][0] = ok ? v0 : T(0);
            hn[r][1] = ok ? v1 : T(0);
            hn[r][2] = ok ? v2 : T(0);
        }
        // new E on the strip
#pragma unroll
        for (int r = 1; r <= RY; r++) {
            const T hy_p = from_prev_lane(hn[r][1]);
            const T hz_p = from_prev_lane(hn[r][2]);
            const bool own = own_x && j0 - 1 + r < ny;
            if (!pro && own) {
                const int o = k * plane + off[r];
                const uint32_t uo = (uint32_t)o;
                A.hout[uo] = hn[r][0];
                A.hout[(uint32_t)nv + uo] = hn[r][1];
                A.hout[2u * (uint32_t)nv + uo] = hn[r][2];
                const int j = yj[r];
                const T dyx = hn[r][2] - hn[r - 1][2], dzx = hn[r][1] - hp[r - 1][1];  // Ex
                const T dzy = hn[r][0] - hp[r - 1][0], dxy = hn[r][2] - hz_p;          // Ey
                const T dxz = hn[r][1] - hy_p, dyz = hn[r][0] - hn[r - 1][0];          // Ez
                T ca0, ca1, ca2, cb0, cb1, cb2;
                int g0, g1, g2;
                if (MODE == 0) {
                    const int q = L.id[r - 1];
                    const T *e = tab + 6 * q;
                    ca0 = e[0]; cb0 = e[1]; ca1 = e[2]; cb1 = e[3]; ca2 = e[4]; cb2 = e[5];
                    g0 = grp[3 * q]; g1 = grp[3 * q + 1]; g2 = grp[3 * q + 2];
                } else if (MODE == 2) {
                    cb0 = A.cacb[o]; cb1 = A.cacb[nv + o]; cb2 = A.cacb[2 * nv + o];
                    ca0 = lossless_ca(cb0); ca1 = lossless_ca(cb1); ca2 = lossless_ca(cb2);
                    g0 = g1 = g2 = -1;
                } else {
                    ca0 = A.cacb[o]; ca1 = A.cacb[nv + o]; ca2 = A.cacb[2 * nv + o];
                    cb0 = A.cacb[3 * nv + o]; cb1 = A.cacb[4 * nv + o]; cb2 = A.cacb[5 * nv + o];
                    g0 = A.grp ? A.grp[o] : -1;
                    g1 = A.grp ? A.grp[nv + o] : -1;
                    g2 = A.grp ? A.grp[2 * nv + o] : -1;
                }
                const T fye = coef[fey + j];
                T v0 = ca0 * ek[r][0] + cb0 * (dyx * fye - dzx * fze);
                T v1 = ca1 * ek[r][1] + cb1 * (dzy * fze - dxy * fxe);
                T v2 = ca2 * ek[r][2] + cb2 * (dxz * fxe - dyz * fye);
                if (SLABS) {
                    const SlabAt &P = PH[r];
                    T u[6];
                    psi_finish<0>(K, cp, cbase, P, 7, pse[r - 1], dyx, dzx, dxy, dzy, dxz, dyz, stx,
                                  sty, stz, true, u);
                    if (P.side[1] >= 0) v0 = v0 + cb0 * (T(1) * u[0]);
                    if (P.side[2] >= 0) v0 = v0 + cb0 * (T(-1) * u[1]);
                    if (P.side[0] >= 0) v1 = v1 + cb1 * (T(-1) * u[2]);
                    if (P.side[2] >= 0) v1 = v1 + cb1 * (T(1) * u[3]);
                    if (P.side[0] >= 0) v2 = v2 + cb2 * (T(1) * u[4]);
                    if (P.side[1] >= 0) v2 = v2 + cb2 * (T(-1) * u[5]);
                }
                if (injects) inject<T, 0>(A, K, t, xi, j, k, 7, v0, v1, v2, cb0, cb1, cb2);
                if (g0 >= 0) v0 = media_update(A, K, g0, 0, o, v0, cb0, ek[r][0]);
                if (g1 >= 0) v1 = media_update(A, K, g1, 1, o, v1, cb1, ek[r][1]);
                if (g2 >= 0) v2 = media_update(A, K, g2, 2, o, v2, cb2, ek[r][2]);
                A.eout[uo] = v0;
                A.eout[(uint32_t)nv + uo] = v1;
                A.eout[2u * (uint32_t)nv + uo] = v2;
            }
        }
#pragma unroll
        for (int r = 0; r < RY; r++) {
            hp[r][0] = hn[r + 1][0];
            hp[r][1] = hn[r + 1][1];
        }
#pragma unroll
        for (int r = 0; r < NR; r++) {
            ek[r][0] = L.en[r][0];
            ek[r][1] = L.en[r][1];
            ek[r][2] = L.en[r][2];
        }
    }
}

// Copies this step's probe and transform values out of copy `src`: out[v] = field[index[v]].
// ------------------------------------------------------------------ plane waves
// A plane wave's auxiliary 1D grid on the device: fields at aux + field_off, E's three
// components then H~'s, len each; its ca, cb, da, db at coef + coef_off, len each.
struct PlaneDesc {
    int g[3];
    int emitter;
    long long len, pad, field_off, coef_off;
    double step[3];
    int nsheet;
    int sheet_c[6];
    long long sheet_i[6];
    double sheet_j[6];
};

// (f[i + g_a] - f[i - g_a]) / step_a, exactly 0 where g_a = 0 (src/fdtd/sources.rs:993-1003)
__device__ __forceinline__ double aux_difference(const PlaneDesc &d, const double *f, int a,
                                                 long long i) {
    const int o = d.g[a];
    if (o == 0) return 0.0;
    return (f[i + o] - f[i - o]) / d.step[a];
}

// H~ <- da H~ - db curl E over every plane wave's samples pad .. len - pad - 1
// (sources.rs:1011-1021); one thread per sample, blockIdx.y the plane wave
static __global__ __launch_bounds__(256) void aux_step_h(const PlaneDesc *planes, double *aux,
                                                  const double *coef) {
    const PlaneDesc &d = planes[blockIdx.y];
    const long long i = d.pad + (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.len - d.pad) return;
    const double *e = aux + d.field_off;
    double *h = aux + d.field_off + 3 * d.len;
    const double *da = coef + d.coef_off + 2 * d.len, *db = coef + d.coef_off + 3 * d.len;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const int a = (c + 1) % 3, b = (c + 2) % 3;
        const double curl = aux_difference(d, e + b * d.len, a, i) - aux_difference(d, e + a * d.len, b, i);
        h[c * d.len + i] = da[i] * h[c * d.len + i] - db[i] * curl;
    }
}

// E <- ca E + cb curl H~, then the sheet E_c[i] -= cb[i] j w (sources.rs:1024-1039), w the
// plane wave's emitter value this step
static __global__ __launch_bounds__(256) void aux_step_e(const PlaneDesc *planes, double *aux,
                                                  const double *coef, const double *waves,
                                                  int emitters, int wave_step) {
    const PlaneDesc &d = planes[blockIdx.y];
    const long long i = d.pad + (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= d.len - d.pad) return;
    double *e = aux + d.field_off;
    const double *h = aux + d.field_off + 3 * d.len;
    const double *ca = coef + d.coef_off, *cb = coef + d.coef_off + d.len;
    const double w = waves[2 * ((long long)wave_step * emitters + d.emitter)];
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const int a = (c + 1) % 3, b = (c + 2) % 3;
        const double curl = aux_difference(d, h + b * d.len, a, i) - aux_difference(d, h + a * d.len, b, i);
        double v = ca[i] * e[c * d.len + i] + cb[i] * curl;
        for (int q = 0; q < d.nsheet; q++)
            if (d.sheet_c[q] == c && d.sheet_i[q] == i) v -= cb[i] * d.sheet_j[q] * w;
        e[c * d.len + i] = v;
    }
}

template <typename T>
__global__ void tap_values(const T *const *fields /* Ex Ey Ez Hx Hy Hz */, const int64_t *index,
                           const int8_t *which, long long n, T *out) {
    const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    out[v] = fields[which[v]][index[v]];
}

// sums += sum over the staged steps of F[s][v] e^(i w t_s) dt, in step order, in FP64: F (real)
// times the phase as num-complex multiplies (F + 0i)(p) (src/fdtd/monitors.rs:185-206). Each
// value's sum is alone and its terms are added in step order: no atomics, the same bits as
// adding one step at a time.
struct SeriesInfo {
    int64_t base_v;    // first value in the staged layout
    int64_t volume;
    int64_t sums_off;  // first complex sum: [f][volume]
    int32_t nf;
    int32_t slot;      // first phase slot
};

template <typename T>
__global__ void dft_accumulate(const T *staged, long long row, long long nvals, int steps,
                               const int32_t *series_of, const SeriesInfo *series,
                               const double *phases, int slots, double *sums) {
    const long long v = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nvals) return;
    const SeriesInfo S = series[series_of[v]];
    const long long local = v - S.base_v;
    for (int f = 0; f < S.nf; f++) {
        const long long o = S.sums_off + (long long)f * S.volume + local;
        double re = sums[2 * o], im = sums[2 * o + 1];
        for (int s = 0; s < steps; s++) {
            const double F = (double)staged[(long long)s * row + v];
            const double pr = phases[2 * ((long long)s * slots + S.slot + f)];
            const double pi = phases[2 * ((long long)s * slots + S.slot + f) + 1];
            re += F * pr - 0.0 * pi;
            im += F * pi + 0.0 * pr;
        }
        sums[2 * o] = re;
        sums[2 * o + 1] = im;
    }
}

}  // namespace pxfdtd
