/* TEST INFRASTRUCTURE ONLY (the CPU oracle): see oracle.h for what may call it and what pins
 * each part. A sequential restatement of the reference's FDTD in FP64 (and, for checking the
 * device's f32 arithmetic, the same operations rounded to f32).
 *
 * Set-up follows src/fdtd.rs:570-869 (averaged permittivity, dt, walls, CPML) and
 * src/fdfd/three.rs:137-156,563-588 (positions, subpixel averaging); conductivity
 * src/fdtd.rs:949-1047; sources src/fdtd.rs:1346-1395 and src/fdtd/sources.rs:132-136,223-240;
 * probes src/fdtd.rs:1059-1098,1403-1409; transforms src/fdtd/monitors.rs:129-206,415-450;
 * media src/fdtd/media.rs:61-109,592-825.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define MAX_SLABS 24
#define MAX_GROUPS 16

static __thread char last_error[512];
const char *or_last_error(void) { return last_error; }
static void fail(const char *msg) { snprintf(last_error, sizeof last_error, "%s", msg); }

typedef struct {
    int field, comp, axis, from, count;
    double *b, *a, *psi;
    long psi_len;
} slab_t;

typedef struct {
    int field, comp;
    long r;
    or_waveform w;
} source_t;

typedef struct {
    int field, n;
    int32_t *comps;
    int64_t *r;
    double *amps;
    or_waveform w;
} current_t;

typedef struct {
    int field, comp;
    long r;
    double *values;
    long len, cap;
} probe_t;

typedef struct {
    int field, comp;
    int lo[3], hi[3];
    long volume;
    double *values; /* [f][k][j][i] complex */
} series_t;

typedef struct {
    int nf;
    double *freqs;
    or_c64 *phases;
    series_t series[6];
} dft_t;

typedef struct {
    long n;
    int32_t *comp;
    int64_t *r;
    double *scale;
    double chi;
    int np;
    double *poles; /* [np][3]: alpha, xi, gamma */
    double *before, *now, *cur, *prev;
} group_t;

/* a plane wave's auxiliary grid and its couplings onto the box's surface; [0] E's, [1] H's */
typedef struct {
    long g[3];
    double steps[3];
    long base, pad, len;
    double *e[3], *h[3], *ca, *cb, *da, *db;
    int nsheet, sheet_c[6];
    long sheet_i[6];
    double sheet_j[6];
    or_waveform w;
    long non[2];
    int *lc[2], *lf[2];
    long *lr[2], *li[2];
    double *lw[2];
} plane_t;

typedef struct or_sim or_sim;
static void probe_push(probe_t *p, double v);
static void dft_phases(dft_t *d, double t, double dt);
static void dft_accumulate_values(const or_sim *s, series_t *se, const dft_t *d, const void *f,
                                  size_t elem);

#define R double
#define SFX(x) x##_64
#include "state.inc"
#undef R
#undef SFX
#define R float
#define SFX(x) x##_32
#include "state.inc"
#undef R
#undef SFX

struct or_sim {
    or_grid g;
    or_boundaries b;
    int n[3];
    long stride[3];
    long N;
    double dt;
    long steps;
    int f32;
    state_64 st64;
    state_32 st32;
    double *kn[3], *kh[3]; /* 1/kappa at the nodes and halfway after them */
    int nslabs;
    slab_t slabs[MAX_SLABS];
    int nsources, ncurrents, nprobes, ndfts, ngroups;
    source_t *sources;
    current_t *currents;
    probe_t *probes;
    dft_t *dfts;
    group_t groups[MAX_GROUPS];
    int32_t *member[3]; /* each E value's media group, -1 outside */
    double *views[6];   /* f64 copies of an f32 run's fields */
    int nplanes;
    plane_t *planes;
};

static void pw_step_h(plane_t *P);
static void pw_step_e(plane_t *P, double half);
static void pw_free(plane_t *P);

#define R double
#define SFX(x) x##_64
#include "step_impl.inc"
#undef R
#undef SFX
#define R float
#define SFX(x) x##_32
#include "step_impl.inc"
#undef R
#undef SFX

static void probe_push(probe_t *p, double v) {
    if (p->len == p->cap) {
        p->cap = p->cap ? 2 * p->cap : 256;
        p->values = realloc(p->values, (size_t)p->cap * sizeof(double));
    }
    p->values[p->len++] = v;
}

/* e^(i w t) dt per frequency (monitors.rs:178-182) */
static void dft_phases(dft_t *d, double t, double dt) {
    for (int f = 0; f < d->nf; f++) {
        const double w = 6.283185307179586 * d->freqs[f];
        or_c64 p = or_cexp((or_c64){0.0, w * t});
        d->phases[f] = (or_c64){p.re * dt, p.im * dt};
    }
}

/* values += F e^(i w t) dt, F real: (F + 0i)(p) as num-complex multiplies (monitors.rs:185) */
static void dft_accumulate_values(const or_sim *s, series_t *se, const dft_t *d, const void *f,
                                  size_t elem) {
    const int nx = s->n[0], ny = s->n[1];
    const int lx = se->hi[0] - se->lo[0], ly = se->hi[1] - se->lo[1];
    const int lz = se->hi[2] - se->lo[2];
    for (int q = 0; q < d->nf; q++) {
        const or_c64 w = d->phases[q];
        double *vals = se->values + 2 * (size_t)q * se->volume;
        for (int k = 0; k < lz; k++)
            for (int j = 0; j < ly; j++)
                for (int i = 0; i < lx; i++) {
                    const long r = ((long)(k + se->lo[2]) * ny + (j + se->lo[1])) * nx + i + se->lo[0];
                    const double v = elem == sizeof(double) ? ((const double *)f)[r]
                                                            : (double)((const float *)f)[r];
                    const long o = ((long)k * ly + j) * lx + i;
                    vals[2 * o] += v * w.re - 0.0 * w.im;
                    vals[2 * o + 1] += v * w.im + 0.0 * w.re;
                }
    }
}

/* ---------------------------------------------------------------- waveforms (fdtd.rs:259-424) */
static double smoothstep(double t, double ramp) {
    if (ramp > 0.0) {
        double s = t / ramp;
        s = s < 0.0 ? 0.0 : (s > 1.0 ? 1.0 : s);
        return s * s * (3.0 - 2.0 * s);
    }
    return 1.0;
}

double or_waveform_at(const or_waveform *w, double t) {
    const double omega = 6.283185307179586 * w->frequency;
    switch (w->kind) {
    case 0: {
        const double u = (t - w->delay) / w->width;
        double s, c;
        or_sincos(omega * (t - w->delay), &s, &c);
        return or_exp(-u * u) * s;
    }
    case 1: {
        const double u = (t - w->delay) / w->width;
        return -u * or_exp(-u * u) * sqrt(2.0 * 2.718281828459045);
    }
    default: {
        or_c64 e = or_cexp((or_c64){0.0, -omega * t});
        const double re = w->amp_re * e.re - w->amp_im * e.im;
        return smoothstep(t, w->ramp) * re;
    }
    }
}

or_c64 or_waveform_complex_at(const or_waveform *w, double t) {
    const double omega = 6.283185307179586 * w->frequency;
    switch (w->kind) {
    case 0: {
        const double u = (t - w->delay) / w->width;
        const double g = or_exp(-u * u);
        or_c64 e = or_cexp((or_c64){0.0, -omega * (t - w->delay)});
        /* (0 + ig)(e.re + i e.im) */
        return (or_c64){0.0 * e.re - g * e.im, 0.0 * e.im + g * e.re};
    }
    case 1:
        return (or_c64){or_waveform_at(w, t), 0.0};
    default: {
        const double s = smoothstep(t, w->ramp);
        or_c64 e = or_cexp((or_c64){0.0, -omega * t});
        /* (s * a) * e, s real times complex first */
        const double ar = s * w->amp_re, ai = s * w->amp_im;
        return (or_c64){ar * e.re - ai * e.im, ar * e.im + ai * e.re};
    }
    }
}

/* ------------------------------------------------------------------------------- set-up */
static double scene_value(const or_scene *sc, int comp, double x, double y, double z) {
    double v = sc->background[comp];
    for (int q = 0; q < sc->n; q++) {
        const or_shape *sh = &sc->shapes[q];
        int in;
        if (sh->kind == 0)
            in = sh->lo[0] < x && x < sh->hi[0] && sh->lo[1] < y && y < sh->hi[1] &&
                 sh->lo[2] < z && z < sh->hi[2];
        else {
            const double dx = x - sh->lo[0], dy = y - sh->lo[1];
            in = dx * dx + dy * dy < sh->hi[0] * sh->hi[0] && sh->lo[2] < z && z < sh->hi[2];
        }
        if (in) v = sh->value[comp];
    }
    return v;
}

static void e_position(const or_grid *g, int comp, int i, int j, int k, double p[3]) {
    p[0] = g->x0 + (double)i * g->dx;
    p[1] = g->y0 + (double)j * g->dy;
    p[2] = g->z0 + (double)k * g->dz;
    const double h[3] = {g->dx, g->dy, g->dz};
    p[comp] += 0.5 * h[comp];
}

/* the reference's averaging, with num-complex's real/complex division e/(e e + 0 0) */
static double averaged(const or_scene *sc, int eps_comp, const double centre[3],
                       const double h[3], int along) {
    const int samples = 8;
    const double n = (double)samples;
    const int a = along, b = (along + 1) % 3, c = (along + 2) % 3;
    double sum = 0.0;
    double p[3] = {centre[0], centre[1], centre[2]};
    for (int sb = 0; sb < samples; sb++) {
        p[b] = centre[b] + (((double)sb + 0.5) / n - 0.5) * h[b];
        for (int sc_ = 0; sc_ < samples; sc_++) {
            p[c] = centre[c] + (((double)sc_ + 0.5) / n - 0.5) * h[c];
            double inverse = 0.0;
            for (int sa = 0; sa < samples; sa++) {
                p[a] = centre[a] + (((double)sa + 0.5) / n - 0.5) * h[a];
                const double e = scene_value(sc, eps_comp, p[0], p[1], p[2]);
                inverse += 1.0 * e / (e * e + 0.0 * 0.0);
            }
            sum += n * inverse / (inverse * inverse + 0.0 * 0.0);
        }
    }
    return sum / (n * n);
}

static int check_setup(const or_grid *g, const or_boundaries *b, double courant) {
    const int n[3] = {g->nx, g->ny, g->nz};
    if ((long)g->nx * g->ny * g->nz == 0 || !(g->dx > 0) || !(g->dy > 0) || !(g->dz > 0)) {
        fail("the grid needs cells and positive steps");
        return -1;
    }
    if (!(b->reflection > 0 && b->reflection < 1) || !(b->order >= 0) || !(b->kappa >= 1) ||
        !(b->alpha >= 0) || (b->has_sigma && !(b->sigma >= 0))) {
        fail("the CPML needs R in (0, 1), an order >= 0, kappa >= 1, alpha >= 0, sigma >= 0");
        return -1;
    }
    for (int a = 0; a < 3; a++)
        if (!b->periodic[a] && b->low[a] + b->high[a] >= n[a]) {
            fail("CPMLs leave nothing of the axis");
            return -1;
        }
    if (!(courant > 0 && courant <= 1)) {
        fail("the Courant number must be in (0, 1]");
        return -1;
    }
    return 0;
}

static void alloc_state_64(or_sim *s) {
    const long N = s->N;
    state_64 *st = &s->st64;
    for (int c = 0; c < 3; c++) {
        st->e[c] = calloc(N, sizeof(double));
        st->h[c] = calloc(N, sizeof(double));
        st->ca[c] = malloc(N * sizeof(double));
        st->cb[c] = malloc(N * sizeof(double));
        st->fe[c] = malloc(s->n[c] * sizeof(double));
        st->fh[c] = malloc(s->n[c] * sizeof(double));
    }
}

/* src/fdtd.rs:799-869 */
static void build_cpml(or_sim *s) {
    const or_boundaries *cp = &s->b;
    const double hs[3] = {s->g.dx, s->g.dy, s->g.dz};
    for (int axis = 0; axis < 3; axis++) {
        s->kn[axis] = malloc(s->n[axis] * sizeof(double));
        s->kh[axis] = malloc(s->n[axis] * sizeof(double));
        for (int m = 0; m < s->n[axis]; m++) s->kn[axis][m] = s->kh[axis][m] = 1.0;
    }
    for (int axis = 0; axis < 3; axis++) {
        if (cp->periodic[axis]) continue;
        const int low = cp->low[axis], high = cp->high[axis];
        const int n = s->n[axis];
        const double h = hs[axis];
#define PROFILE(depth, cells, sig, kap)                                                       \
    do {                                                                                      \
        const double d_ = (double)(cells) * h;                                                \
        double q_ = (depth) / (double)(cells);                                                \
        q_ = q_ < 0 ? 0 : (q_ > 1 ? 1 : q_);                                                  \
        const double x_ = or_pow(q_, cp->order);                                                 \
        const double smax_ = cp->has_sigma ? cp->sigma                                        \
                                           : (cp->order + 1.0) * (-or_ln(cp->reflection)) /   \
                                                 (2.0 * d_);                                  \
        sig = smax_ * x_;                                                                     \
        kap = 1.0 + (cp->kappa - 1.0) * x_;                                                   \
    } while (0)
        double sig, kap;
#define AT(pos)                                                                               \
    do {                                                                                      \
        if (low > 0 && (pos) < (double)low) PROFILE((double)low - (pos), low, sig, kap);       \
        else if (high > 0 && (pos) > (double)(n - high))                                      \
            PROFILE((pos) - (double)(n - high), high, sig, kap);                              \
        else { sig = 0.0; kap = 1.0; }                                                        \
    } while (0)
        for (int m = 0; m < n; m++) {
            AT((double)m);
            s->kn[axis][m] = 1.0 / kap;
            AT((double)m + 0.5);
            s->kh[axis][m] = 1.0 / kap;
        }
        const int bax = (axis + 1) % 3, cax = (axis + 2) % 3;
        const long others = s->N / n;
        const int froms[2] = {0, n - high}, counts[2] = {low, high};
        for (int side = 0; side < 2; side++) {
            const int from = froms[side], count = counts[side];
            if (count == 0) continue;
            for (int fi = 0; fi < 2; fi++) {
                const int field = fi == 0 ? 0 : 1;
                const double offset = fi == 0 ? 0.0 : 0.5;
                double *bb = malloc(count * sizeof(double)), *aa = malloc(count * sizeof(double));
                for (int m = from; m < from + count; m++) {
                    AT((double)m + offset);
                    const double bv = or_exp(-(sig / kap + cp->alpha) * s->dt);
                    const double den = sig * kap + kap * kap * cp->alpha;
                    const double av = (sig > 0.0 && den > 0.0) ? sig * (bv - 1.0) / den : 0.0;
                    bb[m - from] = bv;
                    aa[m - from] = av;
                }
                const int comps[2] = {bax, cax};
                for (int ci = 0; ci < 2; ci++) {
                    slab_t *sl = &s->slabs[s->nslabs++];
                    sl->field = field;
                    sl->comp = comps[ci];
                    sl->axis = axis;
                    sl->from = from;
                    sl->count = count;
                    sl->b = malloc(count * sizeof(double));
                    sl->a = malloc(count * sizeof(double));
                    memcpy(sl->b, bb, count * sizeof(double));
                    memcpy(sl->a, aa, count * sizeof(double));
                    sl->psi_len = (long)count * others;
                    sl->psi = calloc(sl->psi_len, sizeof(double));
                }
                free(bb);
                free(aa);
            }
        }
#undef AT
#undef PROFILE
    }
    state_64 *st = &s->st64;
    for (int q = 0; q < s->nslabs; q++) {
        st->sb[q] = s->slabs[q].b;
        st->sa[q] = s->slabs[q].a;
        st->psi[q] = s->slabs[q].psi;
    }
    for (int a = 0; a < 3; a++)
        for (int m = 0; m < s->n[a]; m++) {
            st->fe[a][m] = s->kn[a][m] / hs[a];
            st->fh[a][m] = s->kh[a][m] / hs[a];
        }
}

static or_sim *sim_with_eps(const or_grid *g, const or_boundaries *b, double courant,
                            const or_scene *scene, const double *eps3n) {
    if (check_setup(g, b, courant)) return NULL;
    or_sim *s = calloc(1, sizeof(or_sim));
    s->g = *g;
    s->b = *b;
    s->n[0] = g->nx;
    s->n[1] = g->ny;
    s->n[2] = g->nz;
    s->stride[0] = 1;
    s->stride[1] = g->nx;
    s->stride[2] = (long)g->nx * g->ny;
    s->N = (long)g->nx * g->ny * g->nz;
    const double hs[3] = {g->dx, g->dy, g->dz};
    /* bloch.rs:31-43 with k = 0: an axis one cell long and periodic doesn't count */
    double sum = 0.0;
    for (int a = 0; a < 3; a++) {
        if (b->periodic[a] && s->n[a] == 1) continue;
        sum += or_powi(hs[a], -2);
    }
    if (sum == 0.0) {
        fail("the field varies along no axis");
        free(s);
        return NULL;
    }
    s->dt = courant / sqrt(sum);
    alloc_state_64(s);
    state_64 *st = &s->st64;
    st->dt = s->dt;
    for (int a = 0; a < 3; a++) st->step[a] = hs[a];
    const long N = s->N;
    for (int c = 0; c < 3; c++) {
        for (long r = 0; r < N; r++) {
            const int i = (int)(r % g->nx), j = (int)((r / g->nx) % g->ny);
            const int k = (int)(r / ((long)g->nx * g->ny));
            double e;
            if (eps3n)
                e = eps3n[c * N + r];
            else {
                double p[3];
                e_position(g, c, i, j, k, p);
                e = averaged(scene, c, p, hs, c);
            }
            if (!(isfinite(e) && e > 0)) {
                fail("the permittivity must be finite and positive");
                or_sim_free(s);
                return NULL;
            }
            st->ca[c][r] = 1.0;
            st->cb[c][r] = s->dt / e;
        }
    }
    /* walls: E tangential to a wall is zero at node 0 (fdtd.rs:738-756) */
    for (int c = 0; c < 3; c++)
        for (int axis = 0; axis < 3; axis++) {
            if (axis == c || b->periodic[axis]) continue;
            for (long r = 0; r < N; r++) {
                const int at[3] = {(int)(r % g->nx), (int)((r / g->nx) % g->ny),
                                   (int)(r / ((long)g->nx * g->ny))};
                if (at[axis] == 0) st->ca[c][r] = st->cb[c][r] = 0.0;
            }
        }
    build_cpml(s);
    return s;
}

or_sim *or_sim_new(const or_grid *g, const or_boundaries *b, const or_scene *eps,
                   double courant) {
    return sim_with_eps(g, b, courant, eps, NULL);
}

or_sim *or_sim_new_eps(const or_grid *g, const or_boundaries *b, const double *eps3n,
                       double courant) {
    return sim_with_eps(g, b, courant, NULL, eps3n);
}

static void free_state_32(or_sim *s) {
    if (!s->f32) return;
    state_32 *st = &s->st32;
    for (int c = 0; c < 3; c++) {
        free(st->e[c]); free(st->h[c]); free(st->ca[c]); free(st->cb[c]);
        free(st->fe[c]); free(st->fh[c]);
    }
    for (int q = 0; q < s->nslabs; q++) { free(st->sb[q]); free(st->sa[q]); free(st->psi[q]); }
    for (int g = 0; g < s->ngroups; g++) {
        free(st->g_scale[g]); free(st->g_poles[g]); free(st->g_before[g]);
        free(st->g_now[g]); free(st->g_cur[g]); free(st->g_prev[g]);
    }
}

void or_sim_free(or_sim *s) {
    if (!s) return;
    free_state_32(s);
    state_64 *st = &s->st64;
    for (int c = 0; c < 3; c++) {
        free(st->e[c]); free(st->h[c]); free(st->ca[c]); free(st->cb[c]);
        free(st->fe[c]); free(st->fh[c]); free(s->kn[c]); free(s->kh[c]);
        free(s->member[c]);
    }
    for (int v = 0; v < 6; v++) free(s->views[v]);
    for (int q = 0; q < s->nslabs; q++) { free(s->slabs[q].b); free(s->slabs[q].a); free(s->slabs[q].psi); }
    for (int q = 0; q < s->ncurrents; q++) {
        free(s->currents[q].comps); free(s->currents[q].r); free(s->currents[q].amps);
    }
    free(s->currents);
    free(s->sources);
    for (int q = 0; q < s->nprobes; q++) free(s->probes[q].values);
    free(s->probes);
    for (int d = 0; d < s->ndfts; d++) {
        for (int q = 0; q < 6; q++) free(s->dfts[d].series[q].values);
        free(s->dfts[d].freqs); free(s->dfts[d].phases);
    }
    free(s->dfts);
    for (int g = 0; g < s->ngroups; g++) {
        group_t *gr = &s->groups[g];
        free(gr->comp); free(gr->r); free(gr->scale); free(gr->poles);
        free(gr->before); free(gr->now); free(gr->cur); free(gr->prev);
    }
    for (int p = 0; p < s->nplanes; p++) pw_free(&s->planes[p]);
    free(s->planes);
    free(s);
}

#include "planewave.inc"

/* src/fdtd.rs:949-992 */
int or_sim_conductivity(or_sim *s, const or_scene *sigma) {
    if (s->ngroups) { fail("the conductivity comes before the dispersive media"); return -1; }
    state_64 *st = &s->st64;
    const or_grid *g = &s->g;
    for (int c = 0; c < 3; c++)
        for (long r = 0; r < s->N; r++) {
            if (st->cb[c][r] == 0.0) continue;
            double p[3];
            e_position(g, c, (int)(r % g->nx), (int)((r / g->nx) % g->ny),
                       (int)(r / ((long)g->nx * g->ny)), p);
            const double sg = scene_value(sigma, 0, p[0], p[1], p[2]);
            if (!(isfinite(sg) && sg >= 0)) { fail("the conductivity must be finite and >= 0"); return -1; }
            const double half = sg * st->cb[c][r] / 2.0;
            st->ca[c][r] = (1.0 - half) / (1.0 + half);
            st->cb[c][r] /= 1.0 + half;
        }
    return 0;
}

static void media_remove(or_sim *s, int c, long r);

void or_sim_conductor(or_sim *s, int comp, int i, int j, int k) {
    const long r = ((long)k * s->n[1] + j) * s->n[0] + i;
    s->st64.ca[comp][r] = 0.0;
    s->st64.cb[comp][r] = 0.0;
    s->st64.e[comp][r] = 0.0;
    media_remove(s, comp, r);
}

static int inside(const or_sim *s, int i, int j, int k) {
    if (i >= 0 && j >= 0 && k >= 0 && i < s->n[0] && j < s->n[1] && k < s->n[2]) return 1;
    fail("outside the grid");
    return 0;
}

int or_sim_add_source(or_sim *s, int field, int comp, int i, int j, int k, const or_waveform *w) {
    if (!inside(s, i, j, k)) return -1;
    s->sources = realloc(s->sources, (s->nsources + 1) * sizeof(source_t));
    s->sources[s->nsources++] = (source_t){field, comp, ((long)k * s->n[1] + j) * s->n[0] + i, *w};
    return 0;
}

int or_sim_add_current(or_sim *s, int field, int n, const int32_t *comps, const int64_t *index,
                       const double *amps, const or_waveform *w) {
    for (int v = 0; v < n; v++)
        if (index[v] < 0 || index[v] >= s->N || comps[v] < 0 || comps[v] > 2) {
            fail("a current's value is outside the grid");
            return -1;
        }
    s->currents = realloc(s->currents, (s->ncurrents + 1) * sizeof(current_t));
    current_t *cu = &s->currents[s->ncurrents++];
    cu->field = field;
    cu->n = n;
    cu->comps = malloc(n * sizeof(int32_t));
    cu->r = malloc(n * sizeof(int64_t));
    cu->amps = malloc(2 * n * sizeof(double));
    memcpy(cu->comps, comps, n * sizeof(int32_t));
    memcpy(cu->r, index, n * sizeof(int64_t));
    memcpy(cu->amps, amps, 2 * n * sizeof(double));
    cu->w = *w;
    return 0;
}

int or_sim_add_probe(or_sim *s, int field, int comp, int i, int j, int k) {
    if (!inside(s, i, j, k)) return -1;
    s->probes = realloc(s->probes, (s->nprobes + 1) * sizeof(probe_t));
    s->probes[s->nprobes] = (probe_t){field, comp, ((long)k * s->n[1] + j) * s->n[0] + i, NULL, 0, 0};
    return s->nprobes++;
}

int or_sim_add_dft(or_sim *s, const int32_t low[3], const int32_t high[3], int nf,
                   const double *freqs) {
    if (nf <= 0) { fail("a monitor needs at least one frequency"); return -1; }
    for (int f = 0; f < nf; f++)
        if (6.283185307179586 * freqs[f] * s->dt >= 3.141592653589793) {
            fail("a frequency is beyond the time step's Nyquist frequency");
            return -1;
        }
    if (!inside(s, high[0], high[1], high[2])) return -1;
    for (int a = 0; a < 3; a++)
        if (low[a] > high[a] || low[a] < 0) { fail("a transform's box must run low to high"); return -1; }
    s->dfts = realloc(s->dfts, (s->ndfts + 1) * sizeof(dft_t));
    dft_t *d = &s->dfts[s->ndfts];
    d->nf = nf;
    d->freqs = malloc(nf * sizeof(double));
    memcpy(d->freqs, freqs, nf * sizeof(double));
    d->phases = malloc(nf * sizeof(or_c64));
    for (int q = 0; q < 6; q++) {
        series_t *se = &d->series[q];
        se->field = q < 3 ? 0 : 1;
        se->comp = q % 3;
        se->volume = 1;
        for (int a = 0; a < 3; a++) {
            se->lo[a] = low[a];
            se->hi[a] = high[a] + 1;
            se->volume *= se->hi[a] - se->lo[a];
        }
        se->values = calloc(2 * (size_t)nf * se->volume, sizeof(double));
    }
    return s->ndfts++;
}

/* ------------------------------------------------------------------ media (media.rs) */
static void media_remove(or_sim *s, int c, long r) {
    if (!s->member[c]) return;
    const int g = s->member[c][r];
    if (g < 0) return;
    s->member[c][r] = -1;
    group_t *gr = &s->groups[g];
    long w = 0;
    for (long v = 0; v < gr->n; v++) {
        if (gr->comp[v] == c && gr->r[v] == r) continue;
        gr->comp[w] = gr->comp[v];
        gr->r[w] = gr->r[v];
        gr->scale[w] = gr->scale[v];
        w++;
    }
    gr->n = w;
}

/* A planted mistake in the pole coefficients, for tests that show a check fails on it: 0 none,
 * 1 the damping term of xi with its sign flipped, 2 alpha without its denominator, 3 the damping
 * rate doubled in every coefficient, 4 chi (E^(n-1)'s weight in the update) halved. */
static int media_plant = 0;
void or_media_plant(int code) { media_plant = code; }

/* Pole::coefficients (media.rs:101-109) from Pole::parts (media.rs:74-87): a pole is
 * (0, strength, f0, damping) for a Lorentz term a / (w0^2 - w^2 - i G w) with a = strength w0^2,
 * or (1, fp, damping, 0) for a Drude term (a = wp^2, w0 = 0); frequencies in c/um, w = 2 pi f.
 * d = 1 + G dt / 2, alpha = (2 - w0^2 dt^2) / d, xi = (G dt / 2 - 1) / d, gamma = a dt^2 / d. */
void or_pole_coefficients(const double *p, double dt, double out[3]) {
    double a, w0sq, gamma;
    if (p[0] == 0.0) {
        const double w0 = 6.283185307179586 * p[2];
        a = p[1] * w0 * w0;
        w0sq = w0 * w0;
        gamma = 6.283185307179586 * p[3];
    } else {
        const double wp = 6.283185307179586 * p[1];
        a = wp * wp;
        w0sq = 0.0;
        gamma = 6.283185307179586 * p[2];
    }
    if (media_plant == 3) gamma = 2.0 * gamma;
    const double d = 1.0 + gamma * dt / 2.0;
    out[0] = media_plant == 2 ? 2.0 - w0sq * dt * dt : (2.0 - w0sq * dt * dt) / d;
    out[1] = media_plant == 1 ? (-gamma * dt / 2.0 - 1.0) / d : (gamma * dt / 2.0 - 1.0) / d;
    out[2] = a * dt * dt / d;
}

int or_sim_add_medium(or_sim *s, double eps_inf, int np, const double *poles,
                      const or_scene *inside_scene) {
    if (s->steps > 0) { fail("a dispersive medium goes in before the run starts"); return -1; }
    if (s->ngroups >= MAX_GROUPS) { fail("too many media"); return -1; }
    const double dt = s->dt;
    double *co = malloc((np ? np : 1) * 3 * sizeof(double));
    double chi_sum = 0.0;
    for (int q = 0; q < np; q++) {
        or_pole_coefficients(poles + 4 * q, dt, co + 3 * q);
        chi_sum += co[3 * q + 2];
    }
    const double chi = media_plant == 4 ? chi_sum / (8.0 * dt) : chi_sum / (4.0 * dt);
    state_64 *st = &s->st64;
    const or_grid *g = &s->g;
    long cap = 1024, n = 0;
    int32_t *vc = malloc(cap * sizeof(int32_t));
    int64_t *vr = malloc(cap * sizeof(int64_t));
    double *vs = malloc(cap * sizeof(double));
    for (int c = 0; c < 3; c++)
        for (long r = 0; r < s->N; r++) {
            if (st->cb[c][r] == 0.0) continue;
            double p[3];
            e_position(g, c, (int)(r % g->nx), (int)((r / g->nx) % g->ny),
                       (int)(r / ((long)g->nx * g->ny)), p);
            if (scene_value(inside_scene, 0, p[0], p[1], p[2]) == 0.0) continue;
            double old_chi = 0.0;
            if (s->member[c] && s->member[c][r] >= 0) old_chi = s->groups[s->member[c][r]].chi;
            double sigma = (1.0 - st->ca[c][r]) / st->cb[c][r] - old_chi;
            if (n == cap) {
                cap *= 2;
                vc = realloc(vc, cap * sizeof(int32_t));
                vr = realloc(vr, cap * sizeof(int64_t));
                vs = realloc(vs, cap * sizeof(double));
            }
            vc[n] = c;
            vr[n] = r;
            vs[n] = sigma > 0.0 ? sigma : 0.0;
            n++;
        }
    const double e = eps_inf;
    for (long v = 0; v < n; v++) {
        const int c = vc[v];
        const long r = vr[v];
        media_remove(s, c, r);
        const double nu = 1.0 + dt / e * (vs[v] / 2.0 + chi);
        st->ca[c][r] = (1.0 - dt * vs[v] / (2.0 * e)) / nu;
        st->cb[c][r] = dt / e / nu;
        vs[v] = st->cb[c][r];
    }
    if (np > 0 && n > 0) {
        const int gi = s->ngroups++;
        group_t *gr = &s->groups[gi];
        gr->n = n;
        gr->comp = vc;
        gr->r = vr;
        gr->scale = vs;
        gr->chi = chi;
        gr->np = np;
        gr->poles = co;
        gr->before = calloc(n, sizeof(double));
        gr->now = calloc(n, sizeof(double));
        gr->cur = calloc(n * np, sizeof(double));
        gr->prev = calloc(n * np, sizeof(double));
        for (int c = 0; c < 3; c++)
            if (!s->member[c]) {
                s->member[c] = malloc(s->N * sizeof(int32_t));
                for (long r = 0; r < s->N; r++) s->member[c][r] = -1;
            }
        for (long v = 0; v < n; v++) s->member[vc[v]][vr[v]] = gi;
        st->g_scale[gi] = gr->scale;
        st->g_chi[gi] = chi;
        st->g_poles[gi] = co;
        st->g_before[gi] = gr->before;
        st->g_now[gi] = gr->now;
        st->g_cur[gi] = gr->cur;
        st->g_prev[gi] = gr->prev;
    } else {
        free(vc); free(vr); free(vs); free(co);
    }
    return 0;
}

/* ------------------------------------------------------------------------ f32 and run */
static float *to_f32(const double *v, long n) {
    float *o = malloc((n ? n : 1) * sizeof(float));
    for (long q = 0; q < n; q++) o[q] = (float)v[q];
    return o;
}

void or_sim_use_f32(or_sim *s) {
    if (s->f32) return;
    state_64 *a = &s->st64;
    state_32 *b = &s->st32;
    const long N = s->N;
    for (int c = 0; c < 3; c++) {
        b->e[c] = to_f32(a->e[c], N);
        b->h[c] = to_f32(a->h[c], N);
        b->ca[c] = to_f32(a->ca[c], N);
        b->cb[c] = to_f32(a->cb[c], N);
        b->fe[c] = to_f32(a->fe[c], s->n[c]);
        b->fh[c] = to_f32(a->fh[c], s->n[c]);
        b->step[c] = (float)(1.0 / a->step[c]); /* f32 psi multiplies by 1/step (yee.wgsl) */
    }
    b->dt = (float)a->dt;
    for (int q = 0; q < s->nslabs; q++) {
        b->sb[q] = to_f32(s->slabs[q].b, s->slabs[q].count);
        b->sa[q] = to_f32(s->slabs[q].a, s->slabs[q].count);
        b->psi[q] = to_f32(s->slabs[q].psi, s->slabs[q].psi_len);
    }
    for (int g = 0; g < s->ngroups; g++) {
        group_t *gr = &s->groups[g];
        b->g_scale[g] = to_f32(gr->scale, gr->n);
        b->g_chi[g] = (float)gr->chi;
        b->g_poles[g] = to_f32(gr->poles, 3 * gr->np);
        b->g_before[g] = to_f32(gr->before, gr->n);
        b->g_now[g] = to_f32(gr->now, gr->n);
        b->g_cur[g] = to_f32(gr->cur, gr->n * gr->np);
        b->g_prev[g] = to_f32(gr->prev, gr->n * gr->np);
    }
    s->f32 = 1;
}

void or_sim_run(or_sim *s, long count) {
    for (long q = 0; q < count; q++) {
        if (s->f32)
            step_32(s, &s->st32);
        else
            step_64(s, &s->st64);
    }
}

/* ---------------------------------------------------------------------------- access */
const double *or_sim_field(or_sim *s, int which, int comp) {
    if (!s->f32) return which == 0 ? s->st64.e[comp] : s->st64.h[comp];
    const float *src = which == 0 ? s->st32.e[comp] : s->st32.h[comp];
    const int v = which * 3 + comp;
    double **slot = &s->views[v];
    if (!*slot) *slot = malloc(s->N * sizeof(double));
    for (long r = 0; r < s->N; r++) (*slot)[r] = (double)src[r];
    return *slot;
}

void or_sim_set_field(or_sim *s, int which, int comp, const double *values) {
    double *d = which == 0 ? s->st64.e[comp] : s->st64.h[comp];
    memcpy(d, values, s->N * sizeof(double));
    if (s->f32) {
        float *f = which == 0 ? s->st32.e[comp] : s->st32.h[comp];
        for (long r = 0; r < s->N; r++) f[r] = (float)values[r];
    }
}

double or_sim_dt(const or_sim *s) { return s->dt; }
long or_sim_steps(const or_sim *s) { return s->steps; }
const double *or_sim_ca(const or_sim *s, int comp) { return s->st64.ca[comp]; }
const double *or_sim_cb(const or_sim *s, int comp) { return s->st64.cb[comp]; }
const double *or_sim_factors(const or_sim *s, int which, int axis) {
    return which == 0 ? s->st64.fe[axis] : s->st64.fh[axis];
}
int or_sim_slab_count(const or_sim *s) { return s->nslabs; }

void or_sim_slab(or_sim *s, int n, int32_t *info5, const double **b, const double **a,
                 const double **psi, int64_t *psi_len) {
    slab_t *sl = &s->slabs[n];
    info5[0] = sl->field;
    info5[1] = sl->comp;
    info5[2] = sl->axis;
    info5[3] = sl->from;
    info5[4] = sl->count;
    *b = sl->b;
    *a = sl->a;
    if (s->f32) {
        if (!sl->psi) sl->psi = calloc(sl->psi_len, sizeof(double));
        for (long q = 0; q < sl->psi_len; q++) sl->psi[q] = (double)s->st32.psi[n][q];
    }
    *psi = sl->psi;
    *psi_len = sl->psi_len;
}

int or_sim_probe_count(const or_sim *s) { return s->nprobes; }
const double *or_sim_probe(const or_sim *s, int n, int64_t *len) {
    *len = s->probes[n].len;
    return s->probes[n].values;
}

const double *or_sim_dft_series(const or_sim *s, int n, int series, int64_t *len) {
    const series_t *se = &s->dfts[n].series[series];
    *len = 2 * (int64_t)s->dfts[n].nf * se->volume;
    return se->values;
}

/* 1/2 sum eps E.E dV (src/fdtd.rs:1573-1589) */
double or_sim_electric_energy(const or_sim *s) {
    const double volume = s->g.dx * s->g.dy * s->g.dz;
    double sum = 0.0;
    for (int c = 0; c < 3; c++)
        for (long r = 0; r < s->N; r++) {
            const double cb = s->st64.cb[c][r];
            if (cb != 0.0) {
                const double v = s->st64.e[c][r];
                sum += s->dt / cb * v * v;
            }
        }
    return 0.5 * sum * volume;
}

/* ------------------------------------------------------------ dipoles (sources.rs:242-315) */
/* Simulation::add_dipole's restriction to the grid: along each axis that varies (not one cell
 * with a periodic side), the position s = (x - first) / step from the component's first value
 * (half a step on along the component for E, across it for H), the two values floor(s) and
 * floor(s) + 1 (wrapped on a periodic axis, the last one's twin clamped at a wall) with weights
 * 1 - f and f; each of the 8 products wx wy wz that is not zero adds a (w / volume) to its value
 * (volume the product of the varying axes' steps), values in index order as a BTreeMap keeps
 * them. Writes up to 8 values (index, amplitude re, im); returns their count, or -1 for a
 * position outside the component's values along a walled axis or one that isn't finite. */
static int cmp_i64(const void *a, const void *b) {
    const int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
    return x < y ? -1 : (x > y);
}

int or_dipole_values(const or_grid *g, const or_boundaries *b, int field, int comp, const double pos[3],
                     double amp_re, double amp_im, int64_t *index_out, double *amp_out) {
    if (!isfinite(amp_re) || !isfinite(amp_im) || !isfinite(pos[0]) || !isfinite(pos[1]) || !isfinite(pos[2])) {
        fail("a dipole must be finite");
        return -1;
    }
    const int n[3] = {g->nx, g->ny, g->nz};
    const double h[3] = {g->dx, g->dy, g->dz}, o[3] = {g->x0, g->y0, g->z0};
    int64_t at[3][2];
    double w[3][2];
    double volume = 1.0;
    for (int a = 0; a < 3; a++) {
        if (n[a] == 1 && b->periodic[a]) {
            at[a][0] = at[a][1] = 0;
            w[a][0] = 1.0;
            w[a][1] = 0.0;
            continue;
        }
        volume *= h[a];
        const int half = field == 0 ? comp == a : comp != a;
        const double first = o[a] + 0.0 * h[a] + (half ? 0.5 * h[a] : 0.0);
        const double s = (pos[a] - first) / h[a];
        if (!b->periodic[a] && !(s >= 0.0 && s <= (double)(n[a] - 1))) {
            fail("the dipole is outside its component's values along an axis");
            return -1;
        }
        const double m = floor(s), f = s - m;
        const int64_t mi = (int64_t)m;
        if (b->periodic[a]) {
            at[a][0] = ((mi % n[a]) + n[a]) % n[a];
            at[a][1] = (((mi + 1) % n[a]) + n[a]) % n[a];
        } else {
            at[a][0] = mi;
            at[a][1] = mi + 1 < n[a] - 1 ? mi + 1 : n[a] - 1;
        }
        w[a][0] = 1.0 - f;
        w[a][1] = f;
    }
    int64_t keys[8];
    double re[8], im[8];
    int count = 0;
    for (int x = 0; x < 2; x++)
        for (int y = 0; y < 2; y++)
            for (int z = 0; z < 2; z++) {
                const double wv = w[0][x] * w[1][y] * w[2][z];
                if (wv == 0.0) continue;
                const int64_t r = (at[2][z] * n[1] + at[1][y]) * n[0] + at[0][x];
                const double q = wv / volume;
                int e = 0;
                while (e < count && keys[e] != r) e++;
                if (e == count) {
                    keys[count] = r;
                    re[count] = 0.0;
                    im[count] = 0.0;
                    count++;
                }
                re[e] += amp_re * q - amp_im * 0.0;
                im[e] += amp_re * 0.0 + amp_im * q;
            }
    int64_t order[8];
    for (int e = 0; e < count; e++) order[e] = keys[e];
    qsort(order, (size_t)count, sizeof(int64_t), cmp_i64);
    for (int e = 0; e < count; e++) {
        int q = 0;
        while (keys[q] != order[e]) q++;
        index_out[e] = order[e];
        amp_out[2 * e] = re[q];
        amp_out[2 * e + 1] = im[q];
    }
    return count;
}
