/* TEST INFRASTRUCTURE ONLY: the CPU speed baseline bench.py times beside the device (its
 * cpu_baseline leg), never the thing measured or shipped.
 *
 * A restatement of the reference's CPU kernel for one Yee + CPML step in f32
 * (src/fdtd/kernel.rs:248-372 curl_update over rows, :461-579 update_slabs), stepped as the
 * reference's CPU path chooses (src/fdtd/kernel/blocked.rs):
 *   - the whole grid each step: the rows, each a plain loop over i that the compiler
 *     vectorizes, shared among threads in chunks of ceil(4096 / nx) rows (kernel.rs:31, rayon
 *     there, OpenMP here), then each CPML slab's rows;
 *   - or, for a grid beyond the caches, by Malas et al.'s diamonds (blocked.rs:1-40, 545-661):
 *     tiles of whole rows along y take s steps at a time, shrinking by a row at each inner side
 *     each step, the gaps around their boundaries filled at the next block by the tiles centred
 *     on them, each task a wavefront along z (plane k at step m right after plane k + 1 at step
 *     m - 1, each plane's H then its E). Each row takes its curl and then its slabs' psi in the
 *     slabs' order, the same operations on every value as the whole-grid path, so both give the
 *     same bits (or_cpu_blocked_differs counts the values that differ: none).
 * The tiles are Blocking::auto's (blocked.rs:76-114): from 40 MiB of E, H and coefficients on,
 * the most steps up to 5 whose tiles of 2s + 2 rows keep at most 4.5 MiB of planes in flight
 * and number at least min(threads, 16).
 * A uniform medium (ca = 1, cb = dt/eps) with a CPML on every face: the kernel's work per
 * cell-update, whatever the geometry. Timed as the reference times its kernels
 * (src/fdtd/gpu/rates.rs:49-100): a warm run, then the step count doubled from 4 until one run
 * takes more than 0.3 s, then the best of 3 runs of that many steps. The fields are first
 * touched by the threads, in the chunks, that step them, so their pages sit on those threads'
 * memory nodes.
 *
 * or_cpu_triad is the STREAM triad the reference measures its machine's roof with
 * (src/traffic.rs:59-80): a = b + s c over doubles in chunks of 2^14 on all threads, 24 bytes an
 * element, the best of `repeats` runs. */
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

typedef struct {
    int axis, side, field, comp, from, count;
    float *b, *a, *psi;
} cslab;

typedef struct {
    int nx, ny, nz;
    long plane;
    float *f[6], *ca[3], *cb[3];
    const float *zeros;
    float fac[3][4096];
    cslab slabs[24];
    int ns;
    float h, dt;
} ctx_t;

static double now(void) { return omp_get_wtime(); }

/* the curl of H (forward) or E on row (j, k) */
static void curl_row(const ctx_t *c, int forward, int j, int k) {
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    const long plane = c->plane, base = ((long)k * ny + j) * nx;
    const float fy = c->fac[1][j], fz = c->fac[2][k], dt = c->dt;
    const float *fx = c->fac[0], *zeros = c->zeros;
    float *const *f = c->f;
    if (forward) {
        /* H -= dt curl E, forward differences; rows j+1, k+1 (zeros beyond) */
        const float *ex = f[0] + base, *ey = f[1] + base, *ez = f[2] + base;
        const float *exy = j + 1 < ny ? ex + nx : zeros, *ezy = j + 1 < ny ? ez + nx : zeros;
        const float *exz = k + 1 < nz ? ex + plane : zeros, *eyz = k + 1 < nz ? ey + plane : zeros;
        float *hx = f[3] + base, *hy = f[4] + base, *hz = f[5] + base;
        for (int i = 0; i < nx; i++) hx[i] = hx[i] - dt * ((ezy[i] - ez[i]) * fy - (eyz[i] - ey[i]) * fz);
        for (int i = 0; i < nx - 1; i++) hy[i] = hy[i] - dt * ((exz[i] - ex[i]) * fz - (ez[i + 1] - ez[i]) * fx[i]);
        hy[nx - 1] = hy[nx - 1] - dt * ((exz[nx - 1] - ex[nx - 1]) * fz - (0.f - ez[nx - 1]) * fx[nx - 1]);
        for (int i = 0; i < nx - 1; i++) hz[i] = hz[i] - dt * ((ey[i + 1] - ey[i]) * fx[i] - (exy[i] - ex[i]) * fy);
        hz[nx - 1] = hz[nx - 1] - dt * ((0.f - ey[nx - 1]) * fx[nx - 1] - (exy[nx - 1] - ex[nx - 1]) * fy);
    } else {
        /* E = ca E + cb curl H, back differences; rows j-1, k-1 */
        const float *hx = f[3] + base, *hy = f[4] + base, *hz = f[5] + base;
        const float *hxy = j > 0 ? hx - nx : zeros, *hzy = j > 0 ? hz - nx : zeros;
        const float *hxz = k > 0 ? hx - plane : zeros, *hyz = k > 0 ? hy - plane : zeros;
        float *ex = f[0] + base, *ey = f[1] + base, *ez = f[2] + base;
        const float *ca0 = c->ca[0] + base, *ca1 = c->ca[1] + base, *ca2 = c->ca[2] + base;
        const float *cb0 = c->cb[0] + base, *cb1 = c->cb[1] + base, *cb2 = c->cb[2] + base;
        for (int i = 0; i < nx; i++) ex[i] = ca0[i] * ex[i] + cb0[i] * ((hz[i] - hzy[i]) * fy - (hy[i] - hyz[i]) * fz);
        ey[0] = ca1[0] * ey[0] + cb1[0] * ((hx[0] - hxz[0]) * fz - (hz[0] - 0.f) * fx[0]);
        for (int i = 1; i < nx; i++)
            ey[i] = ca1[i] * ey[i] + cb1[i] * ((hx[i] - hxz[i]) * fz - (hz[i] - hz[i - 1]) * fx[i]);
        ez[0] = ca2[0] * ez[0] + cb2[0] * ((hy[0] - 0.f) * fx[0] - (hx[0] - hxy[0]) * fy);
        for (int i = 1; i < nx; i++)
            ez[i] = ca2[i] * ez[i] + cb2[i] * ((hy[i] - hy[i - 1]) * fx[i] - (hx[i] - hxy[i]) * fy);
    }
}

/* one value of a slab: psi <- b psi + a dF/dw, the field corrected by sign psi */
static inline void slab_value(const ctx_t *c, const cslab *sl, int forward, const int m[3], long r, long q) {
    const int w = sl->axis, cc = sl->comp, n[3] = {c->nx, c->ny, c->nz};
    const int first = (cc + 1) % 3;
    const float sign = w == first ? 1.0f : -1.0f;
    const int diff = w == first ? (cc + 2) % 3 : first;
    const float *src = forward ? c->f[diff] : c->f[3 + diff];
    float *tgt = forward ? c->f[3 + cc] : c->f[cc];
    const long stride = w == 0 ? 1 : (w == 1 ? c->nx : c->plane);
    const int mw = m[w];
    float other;
    if (forward) other = mw + 1 < n[w] ? src[r + stride] : 0.f;
    else other = mw > 0 ? src[r - stride] : 0.f;
    const float d = (forward ? other - src[r] : src[r] - other) / c->h;
    const int s = mw - sl->from;
    sl->psi[q] = sl->b[s] * sl->psi[q] + sl->a[s] * d;
    const float u = sign * sl->psi[q];
    tgt[r] = forward ? tgt[r] - c->dt * u : tgt[r] + c->cb[cc][r] * u;
}

/* the slabs of H (forward) or E over row (j, k), in their order */
static void slab_rows(const ctx_t *c, int forward, int j, int k) {
    const int nx = c->nx, ny = c->ny;
    for (int q = 0; q < c->ns; q++) {
        const cslab *sl = &c->slabs[q];
        if (sl->field != (forward ? 1 : 0)) continue;
        const int w = sl->axis;
        int m[3] = {0, j, k};
        if (w == 0) {
            for (int li = 0; li < sl->count; li++) {
                m[0] = sl->from + li;
                const long r = ((long)k * ny + j) * nx + m[0];
                slab_value(c, sl, forward, m, r, ((long)k * ny + j) * sl->count + li);
            }
        } else if (w == 1 && j >= sl->from && j < sl->from + sl->count) {
            for (int i = 0; i < nx; i++) {
                m[0] = i;
                slab_value(c, sl, forward, m, ((long)k * ny + j) * nx + i,
                           ((long)k * sl->count + (j - sl->from)) * nx + i);
            }
        } else if (w == 2 && k >= sl->from && k < sl->from + sl->count) {
            for (int i = 0; i < nx; i++) {
                m[0] = i;
                slab_value(c, sl, forward, m, ((long)k * ny + j) * nx + i,
                           ((long)(k - sl->from) * ny + j) * nx + i);
            }
        }
    }
}

static void row(const ctx_t *c, int forward, int j, int k) {
    curl_row(c, forward, j, k);
    slab_rows(c, forward, j, k);
}

/* the whole grid, one step at a time: each field's curl over every row, then each slab */
static void run_whole(ctx_t *c, int steps) {
    const int rows_per = (4096 + c->nx - 1) / c->nx;
    const long nrows = (long)c->ny * c->nz, nchunks = (nrows + rows_per - 1) / rows_per;
    const int n[3] = {c->nx, c->ny, c->nz};
    for (int st = 0; st < steps; st++)
        for (int half = 0; half < 2; half++) {
            const int forward = half == 0; /* H first, then E */
#pragma omp parallel for schedule(static)
            for (long chunk = 0; chunk < nchunks; chunk++)
                for (long rw = chunk * rows_per; rw < (chunk + 1) * rows_per && rw < nrows; rw++)
                    curl_row(c, forward, (int)(rw % c->ny), (int)(rw / c->ny));
            for (int q = 0; q < c->ns; q++) {
                const cslab *sl = &c->slabs[q];
                if (sl->field != (forward ? 1 : 0)) continue;
                const int w = sl->axis;
                int from[3] = {0, 0, 0}, cnt[3] = {n[0], n[1], n[2]};
                from[w] = sl->from;
                cnt[w] = sl->count;
#pragma omp parallel for schedule(static) collapse(2)
                for (int lk = 0; lk < cnt[2]; lk++)
                    for (int lj = 0; lj < cnt[1]; lj++)
                        for (int li = 0; li < cnt[0]; li++) {
                            const int m[3] = {from[0] + li, from[1] + lj, from[2] + lk};
                            const long r = (long)m[2] * c->plane + (long)m[1] * c->nx + m[0];
                            slab_value(c, sl, forward, m, r, ((long)lk * cnt[1] + lj) * cnt[0] + li);
                        }
            }
        }
}

/* ---------------------------------------------------------------- diamonds (blocked.rs) */
typedef struct {
    int gap;        /* 1: the rows around boundary b, growing; 0: the tile [a, b), shrinking */
    int a, b;
    int first, steps;
} segment;

/* Kind::rows (blocked.rs:222-234): the rows [lo, hi) of H (e = 0) or E (e = 1) at step m */
static void seg_rows(const segment *s, int e, int m, int ny, int *lo, int *hi) {
    if (s->gap) {
        *lo = s->b - 1 - m;
        *hi = s->b + m + e;
        return;
    }
    const int l = s->a == 0 ? 0 : s->a + m + e;
    const int h = s->b == ny ? ny : s->b - 1 - m;
    *lo = l;
    *hi = h > l ? h : l;
}

/* Task::run (blocked.rs:303-330): the wavefront along z over the task's steps */
static void run_task(const ctx_t *c, const segment *segs, int nseg) {
    const int first = segs[0].first;
    int total = 0;
    for (int q = 0; q < nseg; q++) total += segs[q].steps;
    for (int p = 0; p < c->nz + total - 1; p++)
        for (int m = 0; m < total && m <= p; m++) {
            const int k = p - m;
            if (k >= c->nz) continue;
            const int g = first + m;
            const segment *s = &segs[0];
            for (int q = 0; q < nseg; q++)
                if (g < segs[q].first + segs[q].steps) {
                    s = &segs[q];
                    break;
                }
            for (int e = 0; e < 2; e++) {
                int lo, hi;
                seg_rows(s, e, g - s->first, c->ny, &lo, &hi);
                for (int j = lo; j < hi; j++) row(c, e == 0, j, k);
            }
        }
}

/* step_blocked (blocked.rs:545-661) with diamonds: block p's tiles have their boundaries at x
 * (even p) or at x's midpoints (odd p); each tile of block p takes block p - 1's gap inside it
 * first, then its own steps, shrinking */
static void run_blocked(ctx_t *c, int steps, int s, int rows) {
    if (steps <= 0) return;
    if (s > steps) s = steps;
    if (s < 1) s = 1;
    const int ny = c->ny;
    int tiles = ny / (rows > 0 ? rows : 1);
    if (ny / (2 * s + 2) < tiles) tiles = ny / (2 * s + 2);
    if (tiles < 1) tiles = 1;
    int *bx = malloc(sizeof(int) * (tiles + 1)), *by = malloc(sizeof(int) * (tiles + 1));
    const int nbx = tiles - 1, nby = tiles > 1 ? tiles - 2 : 0;
    for (int t = 1; t < tiles; t++) bx[t - 1] = (int)((long)t * ny / tiles);
    for (int q = 0; q + 1 < nbx; q++) by[q] = (bx[q] + bx[q + 1]) / 2;
    const int nblocks = (steps + s - 1) / s;
    segment *tasks = malloc(sizeof(segment) * 2 * (tiles + 2));
    int *counts = malloc(sizeof(int) * (tiles + 2));
    for (int p = 0; p <= nblocks; p++) {
        const int *bp = (p % 2 == 0) ? bx : by, nbp = (p % 2 == 0) ? nbx : nby;
        const int *bq = (p % 2 == 1) ? bx : by, nbq = (p % 2 == 1) ? nbx : nby;
        const int pfirst = (p - 1) * s, psteps = p >= 1 ? (steps - pfirst < s ? steps - pfirst : s) : 0;
        int ntasks = 0;
        if (p < nblocks) {
            const int first = p * s, n = steps - first < s ? steps - first : s;
            for (int t = 0; t <= nbp; t++) {
                const int a = t == 0 ? 0 : bp[t - 1], b = t == nbp ? ny : bp[t];
                int ns = 0;
                segment *seg = &tasks[2 * ntasks];
                if (p >= 1)
                    for (int q = 0; q < nbq; q++)
                        if (a < bq[q] && bq[q] < b) {
                            seg[ns++] = (segment){1, 0, bq[q], pfirst, psteps};
                            break;
                        }
                seg[ns++] = (segment){0, a, b, first, n};
                counts[ntasks++] = ns;
            }
        } else {
            for (int q = 0; q < nbq; q++) {
                tasks[2 * ntasks] = (segment){1, 0, bq[q], pfirst, psteps};
                counts[ntasks++] = 1;
            }
        }
#pragma omp parallel for schedule(dynamic, 1)
        for (int t = 0; t < ntasks; t++) run_task(c, &tasks[2 * t], counts[t]);
    }
    free(bx);
    free(by);
    free(tasks);
    free(counts);
}

/* Blocking::auto (blocked.rs:76-114) for f32 on `threads`: s steps at a time in tiles of about
 * `rows` rows, or s = 0 for the whole grid */
static void auto_blocking(int nx, int ny, int nz, int threads, int *s, int *rows) {
    const long cell = 12 * 4, bytes = (long)nx * ny * nz * cell;
    *s = 0;
    *rows = 0;
    if (bytes < (40L << 20)) return;
    const int enough = threads < 1 ? 1 : (threads > 16 ? 16 : threads);
    for (int st = 5; st >= 1; st--) {
        const int r = 2 * st + 2;
        if ((long)(r < nz ? r : nz) * r * nx * cell <= (9L << 19) && ny / r >= enough) {
            *s = st;
            *rows = r;
            return;
        }
    }
}

static void setup(ctx_t *c, int nx, int ny, int nz, int pml) {
    memset(c, 0, sizeof *c);
    c->nx = nx;
    c->ny = ny;
    c->nz = nz;
    c->plane = (long)nx * ny;
    const long N = c->plane * nz;
    const int n[3] = {nx, ny, nz};
    c->h = 0.02f;
    c->dt = (float)(0.9 * 0.02 / sqrt(3.0));
    const int rows_per = (4096 + nx - 1) / nx;
    const long nrows = (long)ny * nz, nchunks = (nrows + rows_per - 1) / rows_per;
    for (int q = 0; q < 6; q++) c->f[q] = aligned_alloc(64, ((N * 4 + 63) / 64) * 64);
    for (int q = 0; q < 3; q++) {
        c->cb[q] = aligned_alloc(64, ((N * 4 + 63) / 64) * 64);
        c->ca[q] = aligned_alloc(64, ((N * 4 + 63) / 64) * 64);
    }
    /* first touch by the threads that step each chunk of rows, in the same static schedule */
#pragma omp parallel for schedule(static)
    for (long chunk = 0; chunk < nchunks; chunk++) {
        const long r0 = chunk * rows_per * nx, r1 = (chunk + 1) * rows_per * nx < N ? (chunk + 1) * rows_per * nx : N;
        for (int q = 0; q < 6; q++)
            for (long r = r0; r < r1; r++) c->f[q][r] = 1e-3f * (float)((r * 2654435761u + q) % 1000) - 0.5f;
        for (int q = 0; q < 3; q++)
            for (long r = r0; r < r1; r++) {
                c->cb[q][r] = c->dt / 2.1f;
                c->ca[q][r] = 1.0f;
            }
    }
    c->zeros = calloc(nx, sizeof(float));
    for (int a = 0; a < 3; a++)
        for (int m = 0; m < n[a] && m < 4096; m++) c->fac[a][m] = 1.0f / c->h;
    c->ns = 0;
    for (int a = 0; a < 3; a++)
        for (int s = 0; s < 2; s++)
            for (int F = 0; F < 2; F++)
                for (int cc = 1; cc <= 2; cc++) {
                    cslab *sl = &c->slabs[c->ns++];
                    sl->axis = a;
                    sl->side = s;
                    sl->field = F;
                    sl->comp = (a + cc) % 3;
                    sl->from = s == 0 ? 0 : n[a] - pml;
                    sl->count = pml;
                    sl->b = malloc(pml * sizeof(float));
                    sl->a = malloc(pml * sizeof(float));
                    for (int m = 0; m < pml; m++) {
                        sl->b[m] = 0.9f - 0.01f * m;
                        sl->a[m] = -0.05f + 0.002f * m;
                    }
                    sl->psi = calloc(N / n[a] * pml, sizeof(float));
                }
}

static void teardown(ctx_t *c) {
    for (int q = 0; q < 6; q++) free(c->f[q]);
    for (int q = 0; q < 3; q++) {
        free(c->cb[q]);
        free(c->ca[q]);
    }
    free((void *)c->zeros);
    for (int q = 0; q < c->ns; q++) {
        free(c->slabs[q].b);
        free(c->slabs[q].a);
        free(c->slabs[q].psi);
    }
}

static void run_steps(ctx_t *c, int steps, int s, int rows) {
    if (s > 0) run_blocked(c, steps, s, rows);
    else run_whole(c, steps);
}

double or_cpu_rate_f32(int nx, int ny, int nz, int pml, int threads, long *steps_out) {
    return or_cpu_rate_blocked_f32(nx, ny, nz, pml, threads, -1, 0, steps_out);
}

double or_cpu_rate_blocked_f32(int nx, int ny, int nz, int pml, int threads, int s, int rows, long *steps_out) {
    if (threads > 0) omp_set_num_threads(threads);
    if (s < 0) auto_blocking(nx, ny, nz, threads > 0 ? threads : omp_get_max_threads(), &s, &rows);
    ctx_t *c = malloc(sizeof(ctx_t));
    setup(c, nx, ny, nz, pml);
    const long N = (long)nx * ny * nz;
    long steps = 4;
    run_steps(c, (int)steps, s, rows);
    for (;;) {
        const double t0 = now();
        run_steps(c, (int)steps, s, rows);
        if (now() - t0 > 0.3) break;
        steps *= 2;
    }
    double best = 0.0;
    for (int rep = 0; rep < 3; rep++) {
        const double t0 = now();
        run_steps(c, (int)steps, s, rows);
        const double rate = (double)N * steps / (now() - t0);
        if (rate > best) best = rate;
    }
    teardown(c);
    free(c);
    if (steps_out) *steps_out = steps;
    return best;
}

void or_cpu_auto_blocking(int nx, int ny, int nz, int threads, int *s, int *rows) {
    auto_blocking(nx, ny, nz, threads, s, rows);
}

long or_cpu_blocked_differs(int nx, int ny, int nz, int pml, int steps, int s, int rows, int threads) {
    if (threads > 0) omp_set_num_threads(threads);
    ctx_t *a = malloc(sizeof(ctx_t)), *b = malloc(sizeof(ctx_t));
    setup(a, nx, ny, nz, pml);
    setup(b, nx, ny, nz, pml);
    run_whole(a, steps);
    run_blocked(b, steps, s, rows);
    const long N = (long)nx * ny * nz;
    long differ = 0;
    for (int q = 0; q < 6; q++)
        for (long r = 0; r < N; r++) differ += a->f[q][r] != b->f[q][r];
    for (int q = 0; q < a->ns; q++) {
        const long vol = N / (q < 8 ? nx : (q < 16 ? ny : nz)) * pml;
        for (long r = 0; r < vol; r++) differ += a->slabs[q].psi[r] != b->slabs[q].psi[r];
    }
    /* a run that moved nothing would match anything: the fields must have changed */
    double moved = 0.0;
    for (long r = 0; r < N; r++) moved += fabs((double)a->f[3][r] - (double)(1e-3f * (float)((r * 2654435761u + 3) % 1000) - 0.5f));
    teardown(a);
    teardown(b);
    free(a);
    free(b);
    return moved > 0.0 ? differ : -1;
}

double or_cpu_triad(long len, int repeats, int threads) {
    if (threads > 0) omp_set_num_threads(threads);
    const long chunk = 1L << 14, nchunks = (len + chunk - 1) / chunk;
    double *a = aligned_alloc(64, ((len * 8 + 63) / 64) * 64);
    double *b = aligned_alloc(64, ((len * 8 + 63) / 64) * 64);
    double *c = aligned_alloc(64, ((len * 8 + 63) / 64) * 64);
#pragma omp parallel for schedule(static)
    for (long q = 0; q < nchunks; q++)
        for (long i = q * chunk; i < (q + 1) * chunk && i < len; i++) {
            a[i] = 0.0;
            b[i] = 1.0;
            c[i] = 2.0;
        }
    volatile double s_ = 3.0;
    const double s = s_;
    double best = 1e30;
    for (int r = 0; r < (repeats > 0 ? repeats : 1); r++) {
        const double t0 = now();
#pragma omp parallel for schedule(static)
        for (long q = 0; q < nchunks; q++)
            for (long i = q * chunk; i < (q + 1) * chunk && i < len; i++) a[i] = b[i] + s * c[i];
        const double el = now() - t0;
        if (el < best) best = el;
    }
    volatile double keep = a[len / 2];
    (void)keep;
    free(a);
    free(b);
    free(c);
    return 24.0 * (double)len / best / 1e9;
}
