/* TEST INFRASTRUCTURE ONLY: the CPU speed baseline bench.py times beside the device (its
 * cpu_baseline leg), never the thing measured or shipped.
 *
 * A restatement of the reference's CPU kernel for one Yee + CPML step in f32
 * (src/fdtd/kernel.rs:248-372 curl_update over rows, :461-579 update_slabs): the rows of the
 * grid, each a plain loop over i that the compiler vectorizes, shared among threads in chunks of
 * ceil(4096 / nx) rows (kernel.rs:31, rayon there, OpenMP here), then each CPML slab's rows.
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

static double now(void) { return omp_get_wtime(); }

/* `steps` steps of the kernel */
static void run_steps(int steps, int nx, int ny, int nz, float **f, float **ca, float **cb, const float *zeros,
                      float fac[3][4096], cslab *slabs, int ns, float h, float dt);

double or_cpu_rate_f32(int nx, int ny, int nz, int pml, int threads, long *steps_out) {
    if (threads > 0) omp_set_num_threads(threads);
    const long N = (long)nx * ny * nz;
    const int n[3] = {nx, ny, nz};
    const float h = 0.02f, dt = (float)(0.9 * 0.02 / sqrt(3.0));
    const int rows_per = (4096 + nx - 1) / nx;
    const long nrows = (long)ny * nz, nchunks = (nrows + rows_per - 1) / rows_per;
    float *f[6], *cb[3], *ca[3];
    for (int q = 0; q < 6; q++) f[q] = aligned_alloc(64, ((N * 4 + 63) / 64) * 64);
    for (int c = 0; c < 3; c++) {
        cb[c] = aligned_alloc(64, ((N * 4 + 63) / 64) * 64);
        ca[c] = aligned_alloc(64, ((N * 4 + 63) / 64) * 64);
    }
    /* first touch by the threads that step each chunk of rows, in the same static schedule */
#pragma omp parallel for schedule(static)
    for (long chunk = 0; chunk < nchunks; chunk++) {
        const long r0 = chunk * rows_per * nx, r1 = (chunk + 1) * rows_per * nx < N ? (chunk + 1) * rows_per * nx : N;
        for (int q = 0; q < 6; q++)
            for (long r = r0; r < r1; r++) f[q][r] = 1e-3f * (float)((r * 2654435761u + q) % 1000) - 0.5f;
        for (int c = 0; c < 3; c++)
            for (long r = r0; r < r1; r++) { cb[c][r] = dt / 2.1f; ca[c][r] = 1.0f; }
    }
    float *zeros = calloc(nx, sizeof(float));
    static float fac[3][4096];
    for (int a = 0; a < 3; a++)
        for (int m = 0; m < n[a] && m < 4096; m++) fac[a][m] = 1.0f / h;
    cslab slabs[24];
    int ns = 0;
    for (int a = 0; a < 3; a++)
        for (int s = 0; s < 2; s++)
            for (int F = 0; F < 2; F++)
                for (int cc = 1; cc <= 2; cc++) {
                    cslab *sl = &slabs[ns++];
                    sl->axis = a; sl->side = s; sl->field = F; sl->comp = (a + cc) % 3;
                    sl->from = s == 0 ? 0 : n[a] - pml;
                    sl->count = pml;
                    sl->b = malloc(pml * sizeof(float));
                    sl->a = malloc(pml * sizeof(float));
                    for (int m = 0; m < pml; m++) { sl->b[m] = 0.9f; sl->a[m] = -0.05f; }
                    const long vol = N / n[a] * pml;
                    sl->psi = calloc(vol, sizeof(float));
                }
    long steps = 4;
    run_steps(steps, nx, ny, nz, f, ca, cb, zeros, fac, slabs, ns, h, dt);
    for (;;) {
        const double t0 = now();
        run_steps(steps, nx, ny, nz, f, ca, cb, zeros, fac, slabs, ns, h, dt);
        if (now() - t0 > 0.3) break;
        steps *= 2;
    }
    double best = 0.0;
    for (int rep = 0; rep < 3; rep++) {
        const double t0 = now();
        run_steps(steps, nx, ny, nz, f, ca, cb, zeros, fac, slabs, ns, h, dt);
        const double rate = (double)N * steps / (now() - t0);
        if (rate > best) best = rate;
    }
    for (int q = 0; q < 6; q++) free(f[q]);
    for (int c = 0; c < 3; c++) { free(cb[c]); free(ca[c]); }
    free(zeros);
    for (int q = 0; q < ns; q++) { free(slabs[q].b); free(slabs[q].a); free(slabs[q].psi); }
    if (steps_out) *steps_out = steps;
    return best;
}

double or_cpu_triad(long len, int repeats, int threads) {
    if (threads > 0) omp_set_num_threads(threads);
    const long chunk = 1L << 14, nchunks = (len + chunk - 1) / chunk;
    double *a = aligned_alloc(64, ((len * 8 + 63) / 64) * 64);
    double *b = aligned_alloc(64, ((len * 8 + 63) / 64) * 64);
    double *c = aligned_alloc(64, ((len * 8 + 63) / 64) * 64);
#pragma omp parallel for schedule(static)
    for (long q = 0; q < nchunks; q++)
        for (long i = q * chunk; i < (q + 1) * chunk && i < len; i++) { a[i] = 0.0; b[i] = 1.0; c[i] = 2.0; }
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

static void run_steps(int steps, int nx, int ny, int nz, float **f, float **ca, float **cb, const float *zeros,
                      float fac[3][4096], cslab *slabs, int ns, float h, float dt) {
    const long plane = (long)nx * ny;
    const int n[3] = {nx, ny, nz};
    const int rows_per = (4096 + nx - 1) / nx;
    const long nrows = (long)ny * nz;
    for (int st = 0; st < steps; st++) {
        for (int half = 0; half < 2; half++) {
            const int forward = half == 0; /* H first, then E */
            float **out = forward ? f + 3 : f;
            float **in = forward ? f : f + 3;
#pragma omp parallel for schedule(static)
            for (long chunk = 0; chunk < (nrows + rows_per - 1) / rows_per; chunk++) {
                for (long row = chunk * rows_per; row < (chunk + 1) * rows_per && row < nrows; row++) {
                    const int j = (int)(row % ny), k = (int)(row / ny);
                    const long base = row * nx;
                    const float fy = fac[1][j], fz = fac[2][k];
                    const float *fx = fac[0];
                    if (forward) {
                        /* H -= dt curl E, forward differences; rows j+1, k+1 (zeros beyond) */
                        const float *ex = f[0] + base, *ey = f[1] + base, *ez = f[2] + base;
                        const float *exy = j + 1 < ny ? ex + nx : zeros, *ezy = j + 1 < ny ? ez + nx : zeros;
                        const float *exz = k + 1 < nz ? ex + plane : zeros, *eyz = k + 1 < nz ? ey + plane : zeros;
                        float *hx = f[3] + base, *hy = f[4] + base, *hz = f[5] + base;
                        for (int i = 0; i < nx; i++)
                            hx[i] = hx[i] - dt * ((ezy[i] - ez[i]) * fy - (eyz[i] - ey[i]) * fz);
                        for (int i = 0; i < nx - 1; i++)
                            hy[i] = hy[i] - dt * ((exz[i] - ex[i]) * fz - (ez[i + 1] - ez[i]) * fx[i]);
                        hy[nx - 1] = hy[nx - 1] - dt * ((exz[nx - 1] - ex[nx - 1]) * fz - (0.f - ez[nx - 1]) * fx[nx - 1]);
                        for (int i = 0; i < nx - 1; i++)
                            hz[i] = hz[i] - dt * ((ey[i + 1] - ey[i]) * fx[i] - (exy[i] - ex[i]) * fy);
                        hz[nx - 1] = hz[nx - 1] - dt * ((0.f - ey[nx - 1]) * fx[nx - 1] - (exy[nx - 1] - ex[nx - 1]) * fy);
                    } else {
                        /* E = ca E + cb curl H, back differences; rows j-1, k-1 */
                        const float *hx = f[3] + base, *hy = f[4] + base, *hz = f[5] + base;
                        const float *hxy = j > 0 ? hx - nx : zeros, *hzy = j > 0 ? hz - nx : zeros;
                        const float *hxz = k > 0 ? hx - plane : zeros, *hyz = k > 0 ? hy - plane : zeros;
                        float *ex = f[0] + base, *ey = f[1] + base, *ez = f[2] + base;
                        const float *ca0 = ca[0] + base, *ca1 = ca[1] + base, *ca2 = ca[2] + base;
                        const float *cb0 = cb[0] + base, *cb1 = cb[1] + base, *cb2 = cb[2] + base;
                        for (int i = 0; i < nx; i++)
                            ex[i] = ca0[i] * ex[i] + cb0[i] * ((hz[i] - hzy[i]) * fy - (hy[i] - hyz[i]) * fz);
                        ey[0] = ca1[0] * ey[0] + cb1[0] * ((hx[0] - hxz[0]) * fz - (hz[0] - 0.f) * fx[0]);
                        for (int i = 1; i < nx; i++)
                            ey[i] = ca1[i] * ey[i] + cb1[i] * ((hx[i] - hxz[i]) * fz - (hz[i] - hz[i - 1]) * fx[i]);
                        ez[0] = ca2[0] * ez[0] + cb2[0] * ((hy[0] - 0.f) * fx[0] - (hx[0] - hxy[0]) * fy);
                        for (int i = 1; i < nx; i++)
                            ez[i] = ca2[i] * ez[i] + cb2[i] * ((hy[i] - hy[i - 1]) * fx[i] - (hx[i] - hxy[i]) * fy);
                    }
                }
            }
            /* the slabs of this field, each a parallel loop over its rows */
            for (int q = 0; q < ns; q++) {
                cslab *sl = &slabs[q];
                if (sl->field != (forward ? 1 : 0)) continue;
                const int w = sl->axis, c = sl->comp;
                const int first = (c + 1) % 3;
                const float sign = w == first ? 1.0f : -1.0f;
                const int diff = w == first ? (c + 2) % 3 : first;
                const float *src = in[diff];
                float *tgt = out[c];
                int from[3] = {0, 0, 0}, cnt[3] = {nx, ny, nz};
                from[w] = sl->from;
                cnt[w] = sl->count;
                const long stride = w == 0 ? 1 : (w == 1 ? nx : plane);
#pragma omp parallel for schedule(static) collapse(2)
                for (int lk = 0; lk < cnt[2]; lk++)
                    for (int lj = 0; lj < cnt[1]; lj++) {
                        for (int li = 0; li < cnt[0]; li++) {
                            const int m[3] = {from[0] + li, from[1] + lj, from[2] + lk};
                            const long r = (long)m[2] * plane + (long)m[1] * nx + m[0];
                            const int mw = m[w];
                            float other;
                            if (forward) other = mw + 1 < n[w] ? src[r + stride] : 0.f;
                            else other = mw > 0 ? src[r - stride] : 0.f;
                            const float d = (forward ? other - src[r] : src[r] - other) / h;
                            const long qq = ((long)lk * cnt[1] + lj) * cnt[0] + li;
                            const int s = mw - sl->from;
                            sl->psi[qq] = sl->b[s] * sl->psi[qq] + sl->a[s] * d;
                            const float u = sign * sl->psi[qq];
                            tgt[r] = forward ? tgt[r] - dt * u : tgt[r] + cb[c][r] * u;
                        }
                    }
            }
        }
    }
}
