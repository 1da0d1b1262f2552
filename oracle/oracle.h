/* TEST INFRASTRUCTURE ONLY. The CPU oracle of the FDTD time-stepper: a sequential C
 * restatement of the reference's FP64 scheme (photonoxide 0.6.4, src/fdtd.rs and
 * src/fdtd/kernel/reference.rs), in the same operation order, compiled with
 * -ffp-contract=off. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only as the checker. The product (photonoxide_amd/) never links or calls it.
 *
 * What pins each part (the reference cannot be compiled here: it is Rust and no Rust
 * toolchain exists in this image):
 *   - grid, averaging, dt, walls, the Yee update, point sources, exp and sincos: the
 *     reference's canary hash 0x7c8db7ae69980888 (tests/canary.rs:1007-1062), reproduced
 *     bit for bit by tests/test_oracle.py;
 *   - the Yee update and dt: the exact discrete dispersion cos(n theta) of a periodic plane
 *     wave (src/fdtd/checks.rs:35-82);
 *   - leapfrog energy conservation in a walled box (src/fdtd/checks.rs:94-136);
 *   - CPML: coefficient closed forms (src/fdtd.rs:799-869) and reflection falling with the
 *     PML's thickness (src/fdtd/checks.rs:144+);
 *   - DFT: the sum by hand from the probes (src/fdtd/monitors_checks.rs:35-71);
 *   - ADE media: a uniform dispersive box against a non-dispersive one where the poles
 *     vanish, and the steady-state leapfrog permittivity (src/fdtd/media.rs:148-159);
 *   - total-field/scattered-field plane waves (src/fdtd/sources.rs:742-1040): the box leaks
 *     only round-off at the reference's five angles (src/fdtd/tests.rs:252-265, bar 1e-12),
 *     a planted coupling error leaks, the reference's refusals (tests.rs:350-400), and a slab's
 *     reflection against Airy's formula at the reference's own measured errors
 *     (src/validation.rs:2848-2856);
 *   - f32 stepping: the same arithmetic rounded to f32 (src/fdtd/kernel.rs `Real for f32`),
 *     checked against the f64 run to the reference's own growth law 2e-8 per step
 *     (src/fdtd/kernel_tests.rs:235-258).
 *   - the CPML profile's pow and ln: fdlibm ports (portable.c), pinned by the reference's grid
 *     hashes 0x0bf1e9d3216efc53 and 0x9133a6eca7e0bb06 (glibc's pow and log give other bits);
 *   - f32 psi terms multiply by the host's 1/step rounded to f32, as the reference's device
 *     kernel does (src/fdtd/gpu/yee.wgsl).
 */
#ifndef PXFDTD_ORACLE_H
#define PXFDTD_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { double re, im; } or_c64;

/* portable maths (portable.c) */
double or_exp(double x);
void or_sincos(double x, double *s, double *c);
double or_sin(double x);
double or_cos(double x);
double or_powi(double x, int n);
double or_ln(double x);
void or_pole_coefficients(const double *pole, double dt, double out[3]);
void or_media_plant(int code);
double or_pow(double x, double y);
or_c64 or_cexp(or_c64 z);

/* src/fdfd/three.rs:70-89 */
typedef struct {
    int32_t nx, ny, nz;
    double dx, dy, dz, x0, y0, z0;
} or_grid;

/* src/fdtd.rs:84-110,177-232; periodic = Bloch with k = 0 */
typedef struct {
    int32_t periodic[3];
    int32_t low[3], high[3];
    double reflection;
    int32_t has_sigma;
    double sigma, order, kappa, alpha;
} or_boundaries;

/* A region: kind 0 is the open box lo < p < hi; kind 1 a cylinder along z,
 * (x-lo0)^2 + (y-lo1)^2 < hi0^2 and lo2 < z < hi2. value[] is the diagonal eps (or sigma in
 * value[0]). Later regions replace earlier ones. */
typedef struct {
    int32_t kind;
    double lo[3], hi[3];
    double value[3];
} or_shape;

typedef struct {
    double background[3];
    int32_t n;
    const or_shape *shapes;
} or_scene;

/* src/fdtd.rs:259-424: kind 0 Gaussian, 1 differentiated Gaussian, 2 continuous */
typedef struct {
    int32_t kind;
    double frequency, width, delay;
    double amp_re, amp_im, ramp;
} or_waveform;

double or_waveform_at(const or_waveform *w, double t);
or_c64 or_waveform_complex_at(const or_waveform *w, double t);

typedef struct or_sim or_sim;

const char *or_last_error(void);
or_sim *or_sim_new(const or_grid *g, const or_boundaries *b, const or_scene *eps, double courant);
or_sim *or_sim_new_eps(const or_grid *g, const or_boundaries *b, const double *eps3n,
                       double courant);
void or_sim_free(or_sim *s);
int or_dipole_values(const or_grid *g, const or_boundaries *b, int field, int comp, const double pos[3],
                     double amp_re, double amp_im, int64_t *index_out, double *amp_out);
int or_sim_conductivity(or_sim *s, const or_scene *sigma);
void or_sim_conductor(or_sim *s, int comp, int i, int j, int k);
int or_sim_add_source(or_sim *s, int field, int comp, int i, int j, int k,
                      const or_waveform *w);
int or_sim_add_current(or_sim *s, int field, int n, const int32_t *comps,
                       const int64_t *index, const double *amps /* [n][2] */,
                       const or_waveform *w);
int or_sim_add_probe(or_sim *s, int field, int comp, int i, int j, int k);
int or_sim_add_dft(or_sim *s, const int32_t low[3], const int32_t high[3], int nf,
                   const double *freqs);
/* poles [np][4]: kind (0 Lorentz: strength, resonance, damping; 1 Drude: plasma, damping) */
int or_sim_add_medium(or_sim *s, double eps_inf, int np, const double *poles,
                      const or_scene *inside);
/* a total-field/scattered-field plane wave (src/fdtd/sources.rs:742-1040) on the box of nodes
 * low to high, travelling along the integer direction, E along the polarization's part across
 * it, in a background of permittivity eps; its auxiliary grid starts at rest now */
int or_sim_add_plane_wave(or_sim *s, const int low[3], const int high[3], const long direction[3],
                          const double polarization[3], double eps, const or_waveform *w);
/* plane p's couplings onto field (0 E, 1 H): returns their count; with arrays, fills
 * component, value index, weight, incident component and auxiliary sample per coupling */
long or_sim_plane_links(const or_sim *s, int p, int field, int32_t *comp, int64_t *r, double *w,
                        int32_t *from, int64_t *sample);
/* plane p's auxiliary grid length; with out, copies its field (0 E, 1 H) component c */
long or_sim_plane_aux(const or_sim *s, int p, int field, int c, double *out);
/* multiplies coupling m of plane p onto field by factor (a planted mistake, for tests) */
void or_sim_plane_scale_link(or_sim *s, int p, int field, long m, double factor);
/* plane p's incident field (0 E, 1 H) component comp at the value (i, j, k) now */
double or_sim_plane_incident(const or_sim *s, int p, int field, int comp, int i, int j, int k);
/* from now on step in f32 (the tables rounded once, the fields rounded now) */
void or_sim_use_f32(or_sim *s);
void or_sim_run(or_sim *s, long count);

/* accessors: which 0 = E, 1 = H; the fields in f64 (an f32 run converts into a buffer) */
const double *or_sim_field(or_sim *s, int which, int comp);
void or_sim_set_field(or_sim *s, int which, int comp, const double *values);
double or_sim_dt(const or_sim *s);
long or_sim_steps(const or_sim *s);
const double *or_sim_ca(const or_sim *s, int comp);
const double *or_sim_cb(const or_sim *s, int comp);
const double *or_sim_factors(const or_sim *s, int which, int axis);
int or_sim_slab_count(const or_sim *s);
/* a slab: field, component, axis, from, count, and its b, a (count each) and psi */
void or_sim_slab(or_sim *s, int n, int32_t *info5, const double **b, const double **a,
                 const double **psi, int64_t *psi_len);
int or_sim_probe_count(const or_sim *s);
const double *or_sim_probe(const or_sim *s, int n, int64_t *len);
/* transform monitor n: its series (6: E x,y,z then H x,y,z), each [f][k][j][i] complex */
const double *or_sim_dft_series(const or_sim *s, int n, int series, int64_t *len);
double or_sim_electric_energy(const or_sim *s);

/* the CPU speed baseline: the reference CPU kernel's row loops (src/fdtd/kernel.rs), rows in
 * chunks of ceil(4096/nx) shared by `threads` OpenMP threads, f32, a uniform medium of
 * nx x ny x nz cells with a CPML of `pml` cells, timed as src/fdtd/gpu/rates.rs:49-100 times
 * (steps doubled from 4 until a run takes > 0.3 s, best of 3); returns cell-updates per second
 * and the steps of a run in *steps_out. */
double or_cpu_rate_f32(int nx, int ny, int nz, int pml, int threads, long *steps_out);
/* the same with the blocking given: s steps at a time in tiles of about `rows` rows (s = 0: the
 * whole grid each step; s < 0: Blocking::auto's choice) */
double or_cpu_rate_blocked_f32(int nx, int ny, int nz, int pml, int threads, int s, int rows, long *steps_out);
/* Blocking::auto's choice for f32 on `threads` threads (s = 0: the whole grid) */
void or_cpu_auto_blocking(int nx, int ny, int nz, int threads, int *s, int *rows);
/* the values of E, H and psi that differ after `steps` steps by diamonds of s steps and about
 * `rows` rows against the whole grid a step at a time (-1 if the fields did not move) */
long or_cpu_blocked_differs(int nx, int ny, int nz, int pml, int steps, int s, int rows, int threads);
/* the STREAM triad a = b + s c over `len` doubles on `threads` threads (src/traffic.rs:59-80):
 * GB/s at 24 bytes an element, best of `repeats` */
double or_cpu_triad(long len, int repeats, int threads);

#ifdef __cplusplus
}
#endif
#endif
