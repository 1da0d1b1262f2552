/* TEST INFRASTRUCTURE ONLY: the CPU speed baseline bench.py times beside the device (its
 * cpu_baseline leg), never the thing measured or shipped.
 *
 * A restatement of the reference's CPU kernel for one Yee + CPML step in f32
 * (src/fdtd/kernel.rs:248-372 curl_update over rows, :461-579 update_slabs): the rows of the
 * grid, each a plain loop over i that the compiler vectorizes, shared among threads in chunks of
 * ceil(4096 / nx) rows (kernel.rs:31, rayon there, OpenMP here), then each CPML slab's rows.
 * A uniform medium (ca = 1, cb = dt/eps) with a CPML on every face: the kernel's work per
 * cell-update, whatever the geometry. Returns the best seconds per step of `repeats` runs of
 * `steps` steps (the reference's protocol takes the best of 3, src/fdtd/gpu/rates.rs:49-100). */
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

double or_cpu_rate_f32(int nx, int ny, int nz, int pml, int steps, int repeats, int threads) {
    if (threads > 0) omp_set_num_threads(threads);
    const long N = (long)nx * ny * nz, plane = (long)nx * ny;
    const int n[3] = {nx, ny, nz};
    const float h = 0.02f, dt = (float)(0.9 * 0.02 / sqrt(3.0));
    float *f[6];
    for (int q = 0; q < 6; q++) {
        f[q] = aligned_alloc(64, ((N * 4 + 63) / 64) * 64);
        for (long r = 0; r < N; r++) f[q][r] = 1e-3f * (float)((r * 2654435761u + q) % 1000) - 0.5f;
    }
    float *cb[3];
    for (int c = 0; c < 3; c++) {
        cb[c] = aligned_alloc(64, ((N * 4 + 63) / 64) * 64);
        for (long r = 0; r < N; r++) cb[c][r] = dt / 2.1f;
    }
    float fac[3][1024];
    for (int a = 0; a < 3; a++)
        for (int m = 0; m < n[a] && m < 1024; m++) fac[a][m] = 1.0f / h;
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
    const int rows_per = (4096 + nx - 1) / nx;
    const long nrows = (long)ny * nz;
    double best = 1e30;
    for (int rep = 0; rep < repeats; rep++) {
        const double t0 = now();
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
                        const long oy = forward ? (j + 1 < ny ? nx : 0) : (j > 0 ? -nx : 0);
                        const long oz = forward ? (k + 1 < nz ? plane : 0) : (k > 0 ? -plane : 0);
                        const float ky = (forward ? (j + 1 < ny) : (j > 0)) ? 1.0f : 0.0f;
                        const float kz = (forward ? (k + 1 < nz) : (k > 0)) ? 1.0f : 0.0f;
                        const float fy = fac[1][j], fz = fac[2][k];
                        for (int c = 0; c < 3; c++) {
                            float *o = out[c] + base;
                            const float *fa = in[(c + 1) % 3] + base, *fb = in[(c + 2) % 3] + base;
                            const float *cbr = cb[c] + base;
                            for (int i = 0; i < nx; i++) {
                                float da, db;
                                /* D_a f_b - D_b f_a, with a = c + 1, b = c + 2 */
                                const int a = (c + 1) % 3, b = (c + 2) % 3;
                                float nb, na;
                                if (a == 0) nb = forward ? (i + 1 < nx ? fb[i + 1] : 0.f) : (i > 0 ? fb[i - 1] : 0.f);
                                else if (a == 1) nb = ky * fb[i + oy];
                                else nb = kz * fb[i + oz];
                                if (b == 0) na = forward ? (i + 1 < nx ? fa[i + 1] : 0.f) : (i > 0 ? fa[i - 1] : 0.f);
                                else if (b == 1) na = ky * fa[i + oy];
                                else na = kz * fa[i + oz];
                                da = forward ? nb - fb[i] : fb[i] - nb;
                                db = forward ? na - fa[i] : fa[i] - na;
                                const float fa_ = a == 0 ? fac[0][i] : (a == 1 ? fy : fz);
                                const float fb_ = b == 0 ? fac[0][i] : (b == 1 ? fy : fz);
                                const float curl = da * fa_ - db * fb_;
                                o[i] = forward ? o[i] - dt * curl : o[i] + cbr[i] * curl;
                            }
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
        const double el = (now() - t0) / steps;
        if (el < best) best = el;
    }
    for (int q = 0; q < 6; q++) free(f[q]);
    for (int c = 0; c < 3; c++) free(cb[c]);
    for (int q = 0; q < ns; q++) { free(slabs[q].b); free(slabs[q].a); free(slabs[q].psi); }
    return best;
}
