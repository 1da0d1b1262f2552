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
// slab's box), also in two copies. E's coefficients are a per-component material id (uint8,
// or uint16 when a component has more than 256 distinct (ca, cb) pairs) into a table staged
// in LDS: 1 byte per value instead of 8.
//
// One workgroup owns a tile of 64 x TY values in x and y and marches along z over a chunk of
// planes. At each plane it computes H's new values over the tile and the halo one value
// below in x and y (the neighbours E's back differences read), exchanges them through LDS,
// then updates E over the tile. H at the plane below is kept in registers from the previous
// plane; the chunk's first plane computes it once (a prologue). Only the tile's own values
// are written, so the redundant halo work never races.
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

template <typename T>
struct StepArgs {
    int nx, ny, nz, pitch;
    long long plane;          // pitch * ny
    int periodic[3];
    int tiles_x, tiles_y, chunks, kz;
    int low[3], high[3];      // CPML cells per axis side
    const T *e_in[3], *h_in[3];
    T *e_out[3], *h_out[3];
    // coefficients, mode 0: a uint8 id per value into a table staged in LDS
    const uint8_t *ids[3];
    const T *table;           // [3][table_stride][2]: ca, cb
    const int8_t *group_of;   // [3][table_stride]: media group or -1
    int table_stride;
    // mode 1: ca, cb and the media group per value
    const T *ca[3], *cb[3];
    const int8_t *grp[3];     // may be null: no media
    T dt;
    T step[3];
    const T *fe[3], *fh[3];   // kappa/step along each axis, nodes and halves
    const T *pb[3][2][2];     // CPML b per [axis][side][field E=0,H=1], count values
    const T *pa[3][2][2];
    const T *psi_in;
    T *psi_out;
    long long psi_off[2][3][3][2];  // [field][component][axis][side], -1 if none
    // injections: per tile a list sorted by plane, inj_off[tile * (nz + 1) + k]
    const int32_t *inj_off;
    const Injection *inj;
    const double *waves;      // this step's emitters' w, complex
    // media (dense per-value state, read only where a value's group is set)
    const MediumGroup *groups;
    int ngroups;
    int max_poles;
    T *m_before;              // [3][N]
    T *m_cur, *m_prev;        // [3][N][max_poles]
    long long mstride;        // N with pitch (values per component)
};

template <typename T>
__device__ __forceinline__ T ld(const T *a, long long idx) {
    return idx >= 0 ? a[idx] : T(0);
}

// index of the neighbour one step forward/back along an axis, -1 beyond a wall
__device__ __forceinline__ long long fwd(long long r, int m, int n, long long stride, int per) {
    return m + 1 < n ? r + stride : (per ? r - (long long)(n - 1) * stride : -1);
}
__device__ __forceinline__ long long bwd(long long r, int m, int n, long long stride, int per) {
    return m > 0 ? r - stride : (per ? r + (long long)(n - 1) * stride : -1);
}

// psi's slab side for coordinate m along an axis: 0 low, 1 high, -1 none
__device__ __forceinline__ int side_of(int m, int n, int low, int high) {
    return m < low ? 0 : (m >= n - high ? 1 : -1);
}

template <typename T>
__device__ __forceinline__ T psi_update(const StepArgs<T> &A, int F, int c, int w, int side,
                                        int i, int j, int k, T d, bool write) {
    const int m[3] = {i, j, k};
    const int n[3] = {A.nx, A.ny, A.nz};
    int from[3] = {0, 0, 0}, cnt[3] = {n[0], n[1], n[2]};
    from[w] = side == 0 ? 0 : n[w] - A.high[w];
    cnt[w] = side == 0 ? A.low[w] : A.high[w];
    const long long q = ((long long)(m[2] - from[2]) * cnt[1] + (m[1] - from[1])) * cnt[0] +
                        (m[0] - from[0]);
    const long long off = A.psi_off[F][c][w][side] + q;
    const int s = m[w] - from[w];
    const T b = A.pb[w][side][F][s], a = A.pa[w][side][F][s];
    const T v = b * A.psi_in[off] + a * d;
    if (write) A.psi_out[off] = v;
    return v;
}

// H's new components at (i, j, k) in `mask` (bit c), from copy p; psi written if `write`.
template <typename T>
__device__ __forceinline__ void h_new(const StepArgs<T> &A, int i, int j, int k, int mask,
                                      bool write, T out[3]) {
    const long long r = (long long)k * A.plane + (long long)j * A.pitch + i;
    const long long rx = fwd(r, i, A.nx, 1, A.periodic[0]);
    const long long ry = fwd(r, j, A.ny, A.pitch, A.periodic[1]);
    const long long rz = fwd(r, k, A.nz, A.plane, A.periodic[2]);
    const T ex = A.e_in[0][r], ey = A.e_in[1][r], ez = A.e_in[2][r];
    const T dt = A.dt;
    const int sx = side_of(i, A.nx, A.low[0], A.high[0]);
    const int sy = side_of(j, A.ny, A.low[1], A.high[1]);
    const int sz = side_of(k, A.nz, A.low[2], A.high[2]);
    if (mask & 1) {  // Hx: D_y(Ez) fy - D_z(Ey) fz
        const T dy = ld(A.e_in[2], ry) - ez, dz = ld(A.e_in[1], rz) - ey;
        T v = A.h_in[0][r];
        v = v - dt * (dy * A.fh[1][j] - dz * A.fh[2][k]);
        if (sy >= 0) v = v - dt * (T(1) * psi_update(A, 1, 0, 1, sy, i, j, k, dy / A.step[1], write));
        if (sz >= 0) v = v - dt * (T(-1) * psi_update(A, 1, 0, 2, sz, i, j, k, dz / A.step[2], write));
        out[0] = v;
    }
    if (mask & 2) {  // Hy: D_z(Ex) fz - D_x(Ez) fx
        const T dz = ld(A.e_in[0], rz) - ex, dx = ld(A.e_in[2], rx) - ez;
        T v = A.h_in[1][r];
        v = v - dt * (dz * A.fh[2][k] - dx * A.fh[0][i]);
        if (sx >= 0) v = v - dt * (T(-1) * psi_update(A, 1, 1, 0, sx, i, j, k, dx / A.step[0], write));
        if (sz >= 0) v = v - dt * (T(1) * psi_update(A, 1, 1, 2, sz, i, j, k, dz / A.step[2], write));
        out[1] = v;
    }
    if (mask & 4) {  // Hz: D_x(Ey) fx - D_y(Ex) fy
        const T dx = ld(A.e_in[1], rx) - ey, dy = ld(A.e_in[0], ry) - ex;
        T v = A.h_in[2][r];
        v = v - dt * (dx * A.fh[0][i] - dy * A.fh[1][j]);
        if (sx >= 0) v = v - dt * (T(1) * psi_update(A, 1, 2, 0, sx, i, j, k, dx / A.step[0], write));
        if (sy >= 0) v = v - dt * (T(-1) * psi_update(A, 1, 2, 1, sy, i, j, k, dy / A.step[1], write));
        out[2] = v;
    }
}

// the injections of plane k on the value (field, i, j) in `mask`, in the tile's list order
template <typename T, int FIELD>
__device__ __forceinline__ void inject(const StepArgs<T> &A, int tile, int i, int j, int k,
                                       int mask, T v[3], const T cb[3]) {
    if (!A.inj_off) return;
    const int32_t *off = A.inj_off + (long long)tile * (A.nz + 1);
    const int lo = off[k], hi = off[k + 1];
    for (int q = lo; q < hi; q++) {
        const Injection in = A.inj[q];
        const int c = in.code & 3;
        if ((in.code >> 2) != FIELD || in.i != i || in.j != j || !((mask >> c) & 1)) continue;
        const double wr = A.waves[2 * in.emitter], wi = A.waves[2 * in.emitter + 1];
        const T amount = T(in.a_re * wr - in.a_im * wi);
        if (FIELD == 1)
            v[c] = v[c] - A.dt * amount;
        else
            v[c] = v[c] - cb[c] * amount;
    }
}

template <typename T, int MODE, int TY>
__global__ __launch_bounds__(TX *(TY + 2)) void fused_step(StepArgs<T> A) {
    __shared__ T hs[3][TY + 1][TX + 1];
    extern __shared__ unsigned char smem_raw[];
    T *tab = reinterpret_cast<T *>(smem_raw);  // [3][stride][2]
    int8_t *grp = reinterpret_cast<int8_t *>(tab + 6 * A.table_stride);
    const int lane = threadIdx.x, wave = threadIdx.y;
    if (MODE == 0) {
        for (int q = wave * TX + lane; q < 6 * A.table_stride; q += TX * (TY + 2)) tab[q] = A.table[q];
        for (int q = wave * TX + lane; q < 3 * A.table_stride; q += TX * (TY + 2)) grp[q] = A.group_of[q];
    }

    const int b = blockIdx.x;
    const int tiles = A.tiles_x * A.tiles_y;
    const int bz = b / tiles, t = b % tiles;
    const int by = t / A.tiles_x, bx = t % A.tiles_x;
    const int i0 = bx * TX, j0 = by * TY;
    const int kz0 = bz * A.kz;
    const int kz1 = min(A.nz, kz0 + A.kz);

    // what this thread computes: own value, a value of the halo row (j0 - 1) or of the halo
    // column (i0 - 1)
    int role, hi_ = -1, hj = -1, mask = 0;
    if (wave < TY) {
        role = 0;
        hi_ = i0 + lane;
        hj = j0 + wave;
        mask = 7;
        if (hi_ >= A.nx || hj >= A.ny) role = -1;
    } else if (wave == TY) {
        role = 1;  // halo row: Hz, Hx
        hi_ = i0 + lane;
        hj = j0 - 1;
        mask = 5;
        if (hj < 0) hj = A.periodic[1] ? A.ny - 1 : -2;
        if (hi_ >= A.nx || hj == -2) role = -1;
    } else {
        role = 2;  // halo column: Hz, Hy
        hi_ = i0 - 1;
        hj = j0 + lane - 1;
        mask = 6;
        if (hi_ < 0) hi_ = A.periodic[0] ? A.nx - 1 : -2;
        if (lane == 0 || lane > TY || hj >= A.ny || hi_ == -2) role = -1;
    }
    // LDS slot
    const int li = role == 2 ? 0 : lane + 1;
    const int lj = role == 0 ? wave + 1 : (role == 1 ? 0 : lane);
    __syncthreads();

    T hprev[3] = {T(0), T(0), T(0)};
    if (role == 0) {
        if (kz0 > 0 || A.periodic[2]) {
            const int kp = kz0 > 0 ? kz0 - 1 : A.nz - 1;
            T v[3];
            h_new(A, hi_, hj, kp, 3, false, v);
            const T nocb[3] = {T(0), T(0), T(0)};
            inject<T, 1>(A, t, hi_, hj, kp, 3, v, nocb);
            hprev[0] = v[0];
            hprev[1] = v[1];
        }
    }

    for (int k = kz0; k < kz1; k++) {
        // phase 1: H at the plane over the tile and its halo
        if (role >= 0) {
            T v[3] = {T(0), T(0), T(0)};
            h_new(A, hi_, hj, k, mask, role == 0, v);
            const T nocb[3] = {T(0), T(0), T(0)};
            inject<T, 1>(A, t, hi_, hj, k, mask, v, nocb);
            if (role == 0) {
                const long long r = (long long)k * A.plane + (long long)hj * A.pitch + hi_;
                A.h_out[0][r] = v[0];
                A.h_out[1][r] = v[1];
                A.h_out[2][r] = v[2];
            }
            if (mask & 1) hs[0][lj][li] = v[0];
            if (mask & 2) hs[1][lj][li] = v[1];
            if (mask & 4) hs[2][lj][li] = v[2];
        } else {
            // beyond a wall or the grid: zero neighbours for the back differences
            if (wave == TY && lane < TX) { hs[0][0][lane + 1] = T(0); hs[2][0][lane + 1] = T(0); }
            if (wave == TY + 1 && lane <= TY) { hs[1][lane][0] = T(0); hs[2][lane][0] = T(0); }
        }
        __syncthreads();
        // phase 2: E over the tile
        if (role == 0) {
            const int i = hi_, j = hj;
            const long long r = (long long)k * A.plane + (long long)j * A.pitch + i;
            const T hx = hs[0][lj][li], hy = hs[1][lj][li], hz = hs[2][lj][li];
            const T hz_jm = hs[2][lj - 1][li], hx_jm = hs[0][lj - 1][li];
            const T hz_im = hs[2][lj][li - 1], hy_im = hs[1][lj][li - 1];
            const int sx = side_of(i, A.nx, A.low[0], A.high[0]);
            const int sy = side_of(j, A.ny, A.low[1], A.high[1]);
            const int sz = side_of(k, A.nz, A.low[2], A.high[2]);
            T ca[3], cb[3];
            int grp_c[3];
            for (int c = 0; c < 3; c++) {
                if (MODE == 0) {
                    const int id = (int)A.ids[c][r];
                    ca[c] = tab[(c * A.table_stride + id) * 2];
                    cb[c] = tab[(c * A.table_stride + id) * 2 + 1];
                    grp_c[c] = grp[c * A.table_stride + id];
                } else {
                    ca[c] = A.ca[c][r];
                    cb[c] = A.cb[c][r];
                    grp_c[c] = A.grp[c] ? (int)A.grp[c][r] : -1;
                }
            }
            T v[3];
            const T eold[3] = {A.e_in[0][r], A.e_in[1][r], A.e_in[2][r]};
            {  // Ex: D_y(Hz) fy - D_z(Hy) fz
                const T dy = hz - hz_jm, dz = hy - hprev[1];
                T e = ca[0] * eold[0] + cb[0] * (dy * A.fe[1][j] - dz * A.fe[2][k]);
                if (sy >= 0) e = e + cb[0] * (T(1) * psi_update(A, 0, 0, 1, sy, i, j, k, dy / A.step[1], true));
                if (sz >= 0) e = e + cb[0] * (T(-1) * psi_update(A, 0, 0, 2, sz, i, j, k, dz / A.step[2], true));
                v[0] = e;
            }
            {  // Ey: D_z(Hx) fz - D_x(Hz) fx
                const T dz = hx - hprev[0], dx = hz - hz_im;
                T e = ca[1] * eold[1] + cb[1] * (dz * A.fe[2][k] - dx * A.fe[0][i]);
                if (sx >= 0) e = e + cb[1] * (T(-1) * psi_update(A, 0, 1, 0, sx, i, j, k, dx / A.step[0], true));
                if (sz >= 0) e = e + cb[1] * (T(1) * psi_update(A, 0, 1, 2, sz, i, j, k, dz / A.step[2], true));
                v[1] = e;
            }
            {  // Ez: D_x(Hy) fx - D_y(Hx) fy
                const T dx = hy - hy_im, dy = hx - hx_jm;
                T e = ca[2] * eold[2] + cb[2] * (dx * A.fe[0][i] - dy * A.fe[1][j]);
                if (sx >= 0) e = e + cb[2] * (T(1) * psi_update(A, 0, 2, 0, sx, i, j, k, dx / A.step[0], true));
                if (sy >= 0) e = e + cb[2] * (T(-1) * psi_update(A, 0, 2, 1, sy, i, j, k, dy / A.step[1], true));
                v[2] = e;
            }
            inject<T, 0>(A, t, i, j, k, 7, v, cb);
            // dispersive media: the rest of E's update and the currents' step
            // (src/fdtd/media.rs:670-705); E's old value is the saved E^n
            for (int c = 0; c < 3; c++) {
                const int g = grp_c[c];
                if (g < 0) continue;
                const MediumGroup &G = A.groups[g];
                const long long s = (long long)c * A.mstride + r;
                const T before = A.m_before[s];
                const T chi = T(G.chi);
                T rest = chi * before;
                T cur[4], prv[4];
                for (int q = 0; q < G.poles; q++) {
                    cur[q] = A.m_cur[s * A.max_poles + q];
                    prv[q] = A.m_prev[s * A.max_poles + q];
                    const T alpha = T(G.coef[q][0]), xi = T(G.coef[q][1]);
                    rest = rest - T(0.5) * ((alpha + T(1)) * cur[q] + xi * prv[q]);
                }
                const T next = v[c] + cb[c] * rest;
                const T slope = (next - before) / (T(2) * A.dt);
                for (int q = 0; q < G.poles; q++) {
                    const T alpha = T(G.coef[q][0]), xi = T(G.coef[q][1]), gamma = T(G.coef[q][2]);
                    const T jn = alpha * cur[q] + xi * prv[q] + gamma * slope;
                    A.m_prev[s * A.max_poles + q] = cur[q];
                    A.m_cur[s * A.max_poles + q] = jn;
                }
                A.m_before[s] = eold[c];
                v[c] = next;
            }
            A.e_out[0][r] = v[0];
            A.e_out[1][r] = v[1];
            A.e_out[2][r] = v[2];
            hprev[0] = hx;
            hprev[1] = hy;
        }
        __syncthreads();
    }
}

// Copies this step's probe and transform values out of copy `src`: out[v] = field[index[v]].
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
