/* pxfdtd: the C ABI of the MI355X-native FDTD time-stepper.
 *
 * Two levels, both plain C (opaque handles, plain pointers and sizes, int status codes):
 *
 *  1. The device seam, pxfdtd_device_* and pxfdtd_resident_*: what photonoxide's GPU resident
 *     state does for a Simulation (src/fdtd/gpu/resident.rs:43-870, reached through
 *     Simulation::set_device / run, src/fdtd.rs:1142-1158,1517-1540, and gpu::Resident::new,
 *     upload, step, download, record). The caller (the reference's Rust Simulation, or the C++
 *     host below) owns the problem and computes its tables in FP64, as the reference's host
 *     does (resident.rs:543-843); the device steps them in f32 or f64.
 *
 *  2. The host mirror, pxfdtd_simulation_*: photonoxide's fdtd::Simulation (src/fdtd.rs:470-1590)
 *     in C++: grid, boundaries, averaged permittivity, CPML, conductivity, conductors, point
 *     sources, currents, dipoles, probes, transform monitors and dispersive media, with the
 *     same meaning and errors, stepping on the device through level 1.
 *
 * Every function returning int returns PXFDTD_OK (0) or an error class; pxfdtd_last_error()
 * gives the calling thread's last message. A refused feature or an allocation failure is a
 * status, never an abort (the reference panics on device failure, gpu.rs:440-450).
 *
 * Threading: a pxfdtd_resident or pxfdtd_simulation is used from one host thread at a time and
 * owns one HIP stream; a pxfdtd_device may be shared.
 *
 * Layout of every host array: the reference's, component by component (x, y, z), then k, j, i
 * with i fastest: index(c, (i, j, k)) = c N + (k ny + j) nx + i (src/fdfd/three.rs:122-124).
 */
#ifndef PXFDTD_ABI_H
#define PXFDTD_ABI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    PXFDTD_OK = 0,
    PXFDTD_ERR_INVALID = 1,     /* a value out of range: Error::InvalidValue */
    PXFDTD_ERR_UNSUPPORTED = 2, /* what the device doesn't step: Error::Gpu */
    PXFDTD_ERR_MEMORY = 3,      /* it doesn't fit */
    PXFDTD_ERR_DEVICE = 4,      /* the HIP runtime failed */
    PXFDTD_ERR_TIMEOUT = 5,     /* a wait ended before its job did */
    PXFDTD_ERR_GROWING = 6      /* the fields grew again after the sources ended: Error::Growing */
};
enum { PXFDTD_F32 = 0, PXFDTD_F64 = 1 };
enum { PXFDTD_E = 0, PXFDTD_H = 1 };

const char *pxfdtd_last_error(void);

/* ------------------------------------------------------------------ level 1: the device */

typedef struct pxfdtd_device pxfdtd_device;
typedef struct pxfdtd_resident pxfdtd_resident;

typedef struct {
    char name[256];
    char arch[64];
    int32_t ordinal;
    int32_t compute_units;
    int64_t hbm_bytes;
    int64_t free_bytes;
} pxfdtd_device_info;

/* Opens HIP device `ordinal` (replaces Gpu::new, src/fdtd/gpu.rs:149-234). */
int pxfdtd_device_open(int ordinal, pxfdtd_device **out);
int pxfdtd_device_query(const pxfdtd_device *dev, pxfdtd_device_info *info);
void pxfdtd_device_close(pxfdtd_device *dev);
/* The achievable HBM rate (no counterpart in the reference: the stepper's roofline is also
 * reported against it): streaming kernels over buffers of `bytes` each (at least 1 MiB, a
 * multiple of 16), 16-byte accesses, the best of `reps` timed runs each. out: read, write,
 * copy (2 x bytes) and triad (3 x bytes) in GB/s of bytes read plus written. Allocates three
 * buffers for the call. */
int pxfdtd_device_hbm_rates(pxfdtd_device *dev, int64_t bytes, int32_t reps, double out[4]);

/* A dispersive medium's values and poles (src/fdtd/media.rs:592-607 Group): E's update there
 * uses p_e, p_m as its ca, cb (given in the problem's ca, cb), then the ADE rest. */
typedef struct {
    int32_t poles;           /* at most 4 */
    double chi;              /* sum gamma_p / (4 dt) */
    double coef[4][3];       /* alpha_p, xi_p, gamma_p */
    int64_t count;
    const int32_t *component;
    const int64_t *index;    /* within a component: (k ny + j) nx + i */
} pxfdtd_medium;

/* Everything Resident::new and plan read (resident.rs:154-293,543-793). */
typedef struct {
    int32_t nx, ny, nz;
    int32_t precision;       /* PXFDTD_F32 or PXFDTD_F64 */
    int32_t periodic[3];     /* Bloch k = 0 sides */
    int32_t pml_low[3], pml_high[3];
    double dt;
    double step[3];          /* dx, dy, dz */
    const double *factors_e[3];   /* kappa/step at the nodes, n_axis each */
    const double *factors_h[3];   /* kappa/step halfway after them */
    /* the CPML recursion's b and a per [axis][side low, high][field E, H], count values each */
    const double *cpml_b[3][2][2];
    const double *cpml_a[3][2][2];
    const double *ca;        /* [3][N] */
    const double *cb;        /* [3][N] */
    int32_t nmedia;
    const pxfdtd_medium *media;
} pxfdtd_problem;

/* A source value (Applied, src/fdtd/sources.rs:132-136): after `field`'s curl update each step,
 * H -= dt Re(a w) or E -= cb Re(a w), w the emitter's value that step; entries apply in
 * order. A point source is a = 1 with w = Waveform::at. */
typedef struct {
    int32_t field, component;
    int64_t index;
    double a_re, a_im;
    int32_t emitter;
} pxfdtd_injection;

typedef struct {
    int32_t field, component;
    int64_t index;
} pxfdtd_probe;

/* A transform series (src/fdtd/monitors.rs:45-55): one component over the box lo..hi
 * (hi exclusive), its sums [f][k][j][i] complex, at monitor `monitor`'s frequencies. */
typedef struct {
    int32_t field, component;
    int32_t lo[3], hi[3];
    int32_t monitor;
} pxfdtd_series;

/* The HBM a problem needs on the device, bytes: for admission by footprint. */
int pxfdtd_resident_footprint(const pxfdtd_problem *p, size_t *bytes);
/* Allocates and tables the problem on the device (Resident::new). */
int pxfdtd_resident_create(pxfdtd_device *dev, const pxfdtd_problem *p, pxfdtd_resident **out);
void pxfdtd_resident_destroy(pxfdtd_resident *r);
/* Fields [3][N] and the CPML's psi, all slabs concatenated in the reference's slab order
 * (axis x, y, z; side low, high; field E, H; component axis+1, axis+2; each [k][j][i] over its
 * box): Resident::upload / download (resident.rs:378-413). e, h or psi may be NULL (zeros). */
int pxfdtd_resident_upload(pxfdtd_resident *r, const double *e, const double *h,
                           const double *psi);
int pxfdtd_resident_download(pxfdtd_resident *r, double *e, double *h, double *psi);
int64_t pxfdtd_resident_psi_len(const pxfdtd_resident *r);
int pxfdtd_resident_set_injections(pxfdtd_resident *r, const pxfdtd_injection *inj, int64_t n,
                                   int32_t emitters);
int pxfdtd_resident_set_probes(pxfdtd_resident *r, const pxfdtd_probe *probes, int32_t n);
/* monitor_nf[m]: monitor m's number of frequencies. Phase slots per step: for each monitor,
 * nf slots at E's time then nf at H's. Resets the sums. */
int pxfdtd_resident_set_series(pxfdtd_resident *r, const pxfdtd_series *series, int32_t n,
                               const int32_t *monitor_nf, int32_t monitors);
int32_t pxfdtd_resident_phase_slots(const pxfdtd_resident *r);
/* `count` steps (Resident::step, resident.rs:430-539): waves [count][emitters][2] (each
 * emitter's w at its field's time that step), phases [count][slots][2] (e^(i w t) dt in FP64,
 * E at (n+1) dt, H at (n+1/2) dt). */
int pxfdtd_resident_step(pxfdtd_resident *r, uint32_t count, const double *waves,
                         const double *phases);
/* the steps recorded since the last read of the probes */
int64_t pxfdtd_resident_probe_steps(const pxfdtd_resident *r);
/* probe values recorded since the last read, [steps][probes]; *steps on return */
int pxfdtd_resident_read_probes(pxfdtd_resident *r, double *out, int64_t capacity,
                                int64_t *steps);
/* the transform sums, FP64, series after series, each [f][k][j][i] complex */
int pxfdtd_resident_read_sums(pxfdtd_resident *r, double *out, int64_t capacity);
int64_t pxfdtd_resident_sums_len(const pxfdtd_resident *r);
/* The transform sums where they live, in the device's memory (the layout above), for
 * post-processing on the device (pxfdtd_flux_device) without a copy to the host. Valid until
 * the series are set again or the resident is destroyed; read it on the resident's stream
 * (pxfdtd_resident_stream) or after pxfdtd_resident_synchronize. */
int pxfdtd_resident_sums_device(const pxfdtd_resident *r, const double **sums);
/* the resident's HIP stream (a hipStream_t) */
void *pxfdtd_resident_stream(const pxfdtd_resident *r);
/* sets the transform sums the device goes on adding to (a run continued from sums held
 * elsewhere: each sum stays one running FP64 sum in step order, as the reference's) */
int pxfdtd_resident_write_sums(pxfdtd_resident *r, const double *in, int64_t count);
/* A total-field/scattered-field plane wave's auxiliary 1D grid (PlaneWaveRun,
 * src/fdtd/sources.rs:742-766): the incident wave on samples 0 .. len - 1 along
 * l = sum_a g_a twice_a (twice_a: twice a value's position along axis a in steps). Each step,
 * before the grid's H: H~ <- da H~ - db curl E; after the grid's E: E <- ca E + cb curl H~, then
 * E_c[i] -= cb[i] j w for each sheet entry (c, i, j), w the emitter's value that step
 * (Waveform::at at t + dt/2); both over samples pad .. len - pad - 1, with differences
 * (f[i + g_a] - f[i - g_a]) / step_a (sources.rs:993-1039). Always FP64. */
typedef struct {
    int32_t g[3];
    int64_t len, pad;
    const double *ca, *cb, *da, *db;  /* len each */
    int32_t nsheet;                   /* at most 6 */
    int32_t sheet_component[6];
    int64_t sheet_sample[6];
    double sheet_j[6];
    int32_t emitter;
} pxfdtd_plane_wave;

/* A coupling across a plane wave's box (PlaneWaveRun::correct, sources.rs:971-990): after the
 * field's sources and currents, H -= dt weight F or E -= cb weight F, F the wave's incident
 * field now (E's for H's update, H~'s for E's) of component from_component at `sample`. */
typedef struct {
    int32_t field, component;
    int64_t index;
    double weight;
    int32_t plane, from_component;
    int64_t sample;
} pxfdtd_coupling;

/* The plane waves and their boxes' couplings (Simulation::add_plane_wave's run,
 * sources.rs:716-721; the auxiliary fields start at zero). Emitters index the waves given to
 * pxfdtd_resident_step, as pxfdtd_resident_set_injections set them. */
int pxfdtd_resident_set_plane_waves(pxfdtd_resident *r, const pxfdtd_plane_wave *waves,
                                    int32_t n, const pxfdtd_coupling *couplings, int64_t count);
/* plane wave n's auxiliary fields [3][len], E and H~ (either may be NULL) */
int pxfdtd_resident_plane_state(pxfdtd_resident *r, int32_t n, double *e, double *h);
int pxfdtd_resident_set_plane_state(pxfdtd_resident *r, int32_t n, const double *e,
                                    const double *h);
/* the step kernel: 0 tiles of 64 x rows exchanging H through LDS; 1 strips of 62 x rows per
 * wavefront marching along z with no barrier, loading a plane ahead; 2 the same without the
 * plane-ahead loads; 3 (the default where a field's three components fit in 2 GiB) aligned
 * strips of 64 x rows, the values beside the strip gathered; 4 (the default for a grid one cell
 * thick with z periodic) the 2D march: strips of 64 marching along y, stepping only the field
 * set (TE: Ex, Ey, Hz; TM: Ez, Hx, Hy) that holds values or sources */
int pxfdtd_resident_set_kernel(pxfdtd_resident *r, int32_t kind);
int pxfdtd_resident_get_kernel(const pxfdtd_resident *r);
/* how the aligned march launches its units: 0 (the default) the units meeting no CPML slab,
 * those meeting only the x slabs and the others as three launches on the resident's stream,
 * each in a kernel variant compiled for just that psi work; 1 the same with the third on a
 * second stream beside the others; 2 every unit in one launch of the full variant */
int pxfdtd_resident_set_concurrent(pxfdtd_resident *r, int32_t mode);
/* the tiling: rows per tile (kernel 0: 2, 4, 8; kernels 1 to 3: 1, 2, 4; kernel 4: 1) and
 * planes per chunk (kernel 4: rows per chunk); 0 keeps the current value */
int pxfdtd_resident_set_tiling(pxfdtd_resident *r, int32_t rows, int32_t planes);
int pxfdtd_resident_get_tiling(const pxfdtd_resident *r, int32_t *rows, int32_t *planes);
/* the coefficient storage chosen: 0 uint8 ids into a table, 1 full ca/cb arrays, 2 cb arrays
 * alone (a lossless problem without media, each ca 0 where its cb is 0 and 1 elsewhere) */
int32_t pxfdtd_resident_coefficient_mode(const pxfdtd_resident *r);
/* the average time per step of the last pxfdtd_resident_step call (all its kernels), ms, from
 * HIP events on the resident's stream */
double pxfdtd_resident_last_kernel_ms(const pxfdtd_resident *r);
int pxfdtd_resident_synchronize(pxfdtd_resident *r);
/* time every fused step kernel with HIP events on the resident's stream (off by default) */
int pxfdtd_resident_set_profiling(pxfdtd_resident *r, int32_t on);
/* the fused step kernel's average duration over the last pxfdtd_resident_step call, ms */
double pxfdtd_resident_kernel_ms(const pxfdtd_resident *r);
/* The autotuner (no counterpart in the reference): times `trials` launches of the aligned
 * march at each tiling of rows {1, 2, 4} (at most max_rows) x planes {4, 8, 16} (the 2D
 * march: chunks of 2 to 32 rows), on the
 * resident's current fields without taking a step (the parity stays, injections are left
 * out), and keeps the fastest. Problems with dispersive media (whose state a trial would
 * advance) and the other kernels keep their tiling. The tiling chosen and its ms per launch. */
int pxfdtd_resident_autotune(pxfdtd_resident *r, int32_t trials, int32_t max_rows, int32_t *rows,
                             int32_t *planes, double *best_ms);

/* ------------------------------------------------------------ level 2: the host mirror */
/* (the serving loop, which batches host-mirror simulations, follows the host mirror) */

typedef struct pxfdtd_simulation pxfdtd_simulation;

typedef struct {
    int32_t nx, ny, nz;
    double dx, dy, dz, x0, y0, z0;
} pxfdtd_grid;

/* Boundaries (src/fdtd.rs:177-232) with the Cpml (src/fdtd.rs:84-110) */
typedef struct {
    int32_t periodic[3];
    int32_t low[3], high[3];
    double reflection;
    int32_t has_sigma;
    double sigma, order, kappa, alpha;
} pxfdtd_boundaries;

/* A region of the scene: kind 0 the open box lo < p < hi, kind 1 a cylinder along z,
 * (x - lo0)^2 + (y - lo1)^2 < hi0^2 and lo2 < z < hi2; value[] its diagonal eps (or sigma in
 * value[0], or membership != 0). Later regions replace earlier ones. */
typedef struct {
    int32_t kind;
    double lo[3], hi[3];
    double value[3];
} pxfdtd_shape;

typedef struct {
    double background[3];
    int32_t n;
    const pxfdtd_shape *shapes;
} pxfdtd_scene;

/* Waveform (src/fdtd.rs:259-424): kind 0 Gaussian, 1 differentiated Gaussian, 2 continuous */
typedef struct {
    int32_t kind;
    double frequency, width, delay;
    double amp_re, amp_im, ramp;
} pxfdtd_waveform;

/* Simulation::diagonal with the scene's eps averaged per component over 8^3 samples
 * (src/fdtd.rs:570-653) */
int pxfdtd_simulation_new(const pxfdtd_grid *g, const pxfdtd_boundaries *b,
                          const pxfdtd_scene *eps, double courant, pxfdtd_simulation **out);
/* pxfdtd_simulation_new with the averaging done on `dev` (coefficient construction on the
 * device: the 512 samples per value and shape are the set-up's dominant cost in 3D), the same
 * bits as the host's: each value's samples are the host's sequence of IEEE 754 operations */
int pxfdtd_simulation_new_on_device(pxfdtd_device *dev, const pxfdtd_grid *g, const pxfdtd_boundaries *b,
                                    const pxfdtd_scene *eps, double courant, pxfdtd_simulation **out);
/* the permittivity each value of E sees given directly, [3][N] */
int pxfdtd_simulation_new_eps(const pxfdtd_grid *g, const pxfdtd_boundaries *b,
                              const double *eps3n, double courant, pxfdtd_simulation **out);
void pxfdtd_simulation_destroy(pxfdtd_simulation *s);
int pxfdtd_simulation_conductivity(pxfdtd_simulation *s, const pxfdtd_scene *sigma);
int pxfdtd_simulation_conductor(pxfdtd_simulation *s, int32_t comp, int32_t i, int32_t j,
                                int32_t k);
int pxfdtd_simulation_add_source(pxfdtd_simulation *s, int32_t field, int32_t comp, int32_t i,
                                 int32_t j, int32_t k, const pxfdtd_waveform *w);
int pxfdtd_simulation_add_current(pxfdtd_simulation *s, int32_t field, int64_t n,
                                  const int32_t *comps, const int64_t *index,
                                  const double *amps, const pxfdtd_waveform *w);
int pxfdtd_simulation_add_dipole(pxfdtd_simulation *s, int32_t field, int32_t comp,
                                 const double position[3], double amp_re, double amp_im,
                                 const pxfdtd_waveform *w);
int pxfdtd_simulation_add_probe(pxfdtd_simulation *s, int32_t field, int32_t comp, int32_t i,
                                int32_t j, int32_t k, int32_t *number);
int pxfdtd_simulation_add_dft(pxfdtd_simulation *s, const int32_t low[3], const int32_t high[3],
                              int32_t nf, const double *freqs, int32_t *number);
/* poles [np][4]: (0, strength, resonance, damping) Lorentz or (1, plasma, damping, 0) Drude */
int pxfdtd_simulation_add_medium(pxfdtd_simulation *s, double eps_inf, int32_t np,
                                 const double *poles, const pxfdtd_scene *inside);
/* Dispersive::amplification (src/fdtd/media.rs:258-299): the largest |z| over the modes of a
 * plane wave stepped in the medium eps_inf with poles [np][4] (as pxfdtd_simulation_add_medium
 * takes them) and conductivity sigma at time step dt, for (K dt)^2 = u: the roots of the
 * update's characteristic polynomial, by Aberth's iteration. pxfdtd_simulation_add_medium
 * admits a medium when this is at most 1 + 1e-6 at 200 values of u up to 4 sum 1/h^2 dt^2,
 * without conductivity and with the largest inside (Dispersive::stable, media.rs:300-309). */
int pxfdtd_medium_amplification(double eps_inf, int32_t np, const double *poles, double sigma, double dt,
                                double u, double *out);
/* Simulation::add_flux (src/fdtd/monitors.rs:470-491): the flux towards +axis through the plane
 * halfway between nodes `plane` and plane + 1, within window[w] = {lo, hi} (the half-planes
 * after nodes lo and hi) along the w-th of the axes after `axis` (lo < 0: the whole axis), at
 * nf frequencies (c/um) from the next step on; its number in *number */
int pxfdtd_simulation_add_flux(pxfdtd_simulation *s, int32_t axis, int32_t plane,
                               const int32_t window[2][2], int32_t nf, const double *freqs,
                               int32_t *number);
/* Simulation::add_flux_box (monitors.rs:493-528): the flux out of the box between the
 * half-planes after nodes planes[a][0] and planes[a][1] along each axis a */
int pxfdtd_simulation_add_flux_box(pxfdtd_simulation *s, const int32_t planes[3][2], int32_t nf,
                                   const double *freqs, int32_t *number);
/* Simulation::flux (monitors.rs:530-547): flux monitor n's flux at each of its frequencies,
 * from the FP64 transform sums (out holds pxfdtd_simulation_flux_frequencies values) */
int pxfdtd_simulation_flux(pxfdtd_simulation *s, int32_t n, double *out);
int32_t pxfdtd_simulation_flux_frequencies(const pxfdtd_simulation *s, int32_t n);
/* A flux plane over transform sums held in device memory (FluxPlane, src/fdtd/monitors.rs:
 * 210-350), halfway between nodes `plane` and plane + 1 along `axis`. Along b = (axis + 1) % 3
 * (w = 0) and c = (axis + 2) % 3 (w = 1) the sum runs over values from[w] <= m < to[w]: the
 * window's wlo[w] .. whi[w] + 1 when windowed[w] (FluxPlane::range), which then also weighs the
 * values (FluxPlane::weight), else the grid's 0 .. n. series[6] are the transform's series in
 * device memory, E x, y, z then H x, y, z, each [f][k][j][i] complex over the nodes
 * lo[a] <= n < hi[a] (the resident's layout); series[axis] and series[3 + axis] are not read.
 * sign: the face's sign in its monitor (1, or -1 for a box's low faces); step_b, step_c: the
 * grid's steps along b and c. */
typedef struct {
    int32_t axis, plane;
    int32_t windowed[2], wlo[2], whi[2];
    int32_t from[2], to[2];
    int32_t lo[3], hi[3];
    int32_t nf;
    double sign, step_b, step_c;
    const double *series[6];
} pxfdtd_flux_face;
enum { PXFDTD_FLUX_MAX_FACES = 6 };
/* Post-processing on the device (Simulation::flux, src/fdtd/monitors.rs:530-547, over
 * FluxPlane::flux, :318-350): the flux of n faces (1 to 6, one frequency count) summed per
 * frequency, sign times each face's flux added to -0.0 in face order, each face's terms in the
 * reference's order, so the result carries the sequential host sum's bits. Written to out[nf]
 * in device memory, queued on `stream` (a hipStream_t; null: the device's default stream).
 * The library allocates a terms buffer of 8 B x nf x 2 x (to[0] - from[0]) x (to[1] - from[1])
 * per face for the call, and waits for the stream before it frees it and returns. */
int pxfdtd_flux_device(pxfdtd_device *dev, const pxfdtd_flux_face *faces, int32_t n, double *out,
                       void *stream);
/* pxfdtd_flux_device without an allocation. It has no counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc). It calls no hipMalloc, hipFree or other allocating routine
 * and does not wait: its terms go to a 16 MiB array the library holds on each GPU until the
 * process ends, so it takes at most 2^21 terms (2 x nf x values per face, summed over the
 * faces; more is PXFDTD_ERR_INVALID), and out[nf] is ready when `stream` reaches it. Every
 * call on one GPU shares that GPU's array: calls that overlap on one GPU (on different streams
 * or host threads) race and give wrong values. Same results as pxfdtd_flux_device. */
int pxfdtd_flux_device_no_alloc(pxfdtd_device *dev, const pxfdtd_flux_face *faces, int32_t n,
                                double *out, void *stream);
/* A port's mode for Simulation::add_mode_source (src/fdtd/sources.rs:550-596), given by its
 * profile as the reference's PortMode3d holds it (src/fdfd/three/ports.rs:46-60): the axis the
 * waveguide runs along (0 x, 1 y, 2 z), the plane of nodes it was solved on, the cells along
 * the plane's two axes b = (axis + 1) % 3 and c = (axis + 2) % 3, the step along the axis and
 * the wavenumber k0 it was solved at, its propagation constant beta (rad/um), E_b then E_c on
 * the plane (2 x size[0] x size[1] complex as re, im; b fastest) and E_a on the half-plane
 * after it (size[0] x size[1] complex), going forward at unit amplitude. The profile comes
 * from a mode solver: the host mirror holds none (the reference's is FDFD's PortMode3d). */
typedef struct {
    int32_t axis, plane;
    int32_t size[2];
    double step, k0;
    double beta[2];
    const double *tangential;
    const double *normal;
} pxfdtd_port_mode;
/* The mode launched going `direction` (0 forward, 1 backward) along its axis by
 * total-field/scattered-field on its plane, at the waveform's carrier: the currents of the
 * surface's couplings (src/fdtd/sources.rs:321-418, 419-453) for the mode's E and the grid's
 * curl of it, H~ = curl E / (i k~). The same refusals as the reference's: a plane not this
 * grid's (size or step), a periodic axis, the plane and the next not clear of the CPMLs and
 * the ends, a waveform with no carrier, beyond the Nyquist frequency, or whose leapfrog
 * wavenumber is not k0 to 1e-9. */
int pxfdtd_simulation_add_mode_source(pxfdtd_simulation *s, const pxfdtd_port_mode *mode,
                                      int32_t direction, const pxfdtd_waveform *w);
/* Simulation::add_mode_monitor (src/fdtd/monitors.rs:569-611): monitors the amplitudes of
 * `count` modes from the next step on, each by the FP64 transform of E over its plane and the
 * next at the frequency whose leapfrog wavenumber is its k0 (Simulation::leapfrog_frequency,
 * :560-567). Refused as the reference refuses: no mode, a plane not this grid's (size or step)
 * or its next plane off the grid, a k0 beyond the Nyquist frequency. Mode monitors are
 * numbered on their own, from 0; the number in *number. */
int pxfdtd_simulation_add_mode_monitor(pxfdtd_simulation *s, const pxfdtd_port_mode *modes,
                                       int32_t count, int32_t *number);
/* The permittivity each value of E sees, [3][N] into out: the per-component averages
 * (Simulation::diagonal, src/fdtd.rs:619-653) or the values given to pxfdtd_simulation_new_eps. */
int pxfdtd_simulation_permittivity(const pxfdtd_simulation *s, double *out);
/* Lattice::port_modes (src/fdfd/three/ports.rs:408-594): up to `count` modes of the guide
 * crossing node plane `plane` normal to `axis`, within window[w] = {lo, hi} (nodes; lo < 0 for
 * the whole axis) along the w-th of the axes after `axis`, on the simulation's grid and
 * permittivity at the leapfrog wavenumber of `frequency` (c/um): the eigenvalues nearest
 * k0^2 max(eps, 1) of the plane's transverse problem (found in *found), each normalized to
 * N(f, g) = 1 with its largest tangential value's real part positive, propagating modes first
 * by |Re beta|. Out: beta [count][2], tangential [count][2][size_c][size_b] complex and normal
 * [count][size_c][size_b] complex as pxfdtd_port_mode holds them, and each mode's power (the
 * scheme's flux at unit amplitude; 1 for a real mode). The plane and the next must be clear of
 * the CPMLs as a mode source's; a periodic axis across must be whole in the window. With
 * `lossy` nonzero the permittivity is the leapfrog's at the frequency where a conductivity was
 * given, eps + i sigma cos(w dt / 2) / k0 (the reference's lossy check, src/fdtd/checks.rs:689-
 * 740); otherwise the real permittivity, as the reference's job solves it. */
int pxfdtd_simulation_port_modes(pxfdtd_simulation *s, int32_t axis, int32_t plane,
                                 const int32_t window[2][2], double frequency, int32_t count,
                                 int32_t lossy, int32_t *found, double *beta, double *tangential,
                                 double *normal, double *power);
/* harmonic_inversion (src/fdtd/harmonic.rs:62-184): the terms sum_k d_k e^(-i w_k t) of a
 * signal of n complex samples (re, im pairs) every dt (um/c) whose frequencies lie in [low,
 * high] (c/um), by Mandelshtam and Taylor's filter diagonalization, sorted by frequency: each
 * as (frequency, decay rate, amplitude re, amplitude im, error) in out, up to `capacity`, the
 * number found in *found. Refused: dt not positive, high not above low or at or past the
 * Nyquist frequency, fewer than 8 samples or one not finite. */
int pxfdtd_harmonic_inversion(const double *signal, int64_t n, double dt, double low, double high,
                              int32_t capacity, int32_t *found, double *out);
/* pxfdtd_harmonic_inversion of a signal in device memory (`signal`: n complex FP64 samples on
 * `dev`, read on `stream`, a hipStream_t or null), for post-processing on the device: the sums
 * over the samples at each of the filter diagonalization's points run there and only they come
 * back; the small eigenproblem runs on the host. The same resonances as the host's, bit for bit,
 * and the same refusals. It allocates 8 B x 12 x points of device memory per call and waits for
 * the stream. */
int pxfdtd_harmonic_inversion_device(pxfdtd_device *dev, const double *signal, int64_t n, double dt, double low,
                                     double high, int32_t capacity, int32_t *found, double *out, void *stream);
/* Lattice::mode_amplitude_weights (src/fdfd/three/ports.rs:293-345): the mode monitor's
 * amplitude of `mode` going `direction` (0 forward, 1 backward) as weights on E's transform:
 * out [3][box] complex, the box the mode's plane and the next along its axis and the whole grid
 * across, [c][k][j][i] box-local; the amplitude is the sum of weight times transform (an
 * adjoint's objective, with component -1). */
int pxfdtd_simulation_mode_weights(pxfdtd_simulation *s, const pxfdtd_port_mode *mode, int32_t direction,
                                   double *out);
/* The modes in mode monitor n, -1 for no monitor n. */
int32_t pxfdtd_simulation_mode_count(const pxfdtd_simulation *s, int32_t n);
/* Simulation::mode_frequencies (monitors.rs:653-663): each mode's frequency (c/um) in out. */
int pxfdtd_simulation_mode_frequencies(const pxfdtd_simulation *s, int32_t n, double *out);
/* Simulation::mode_amplitudes (monitors.rs:613-640): each mode's forward and backward
 * amplitudes in the transform, out[4 m ..] = forward re, im, backward re, im: FDFD's Lorentz
 * form across the mode's plane and the next (src/fdfd/three/ports.rs:253-287, 637-651), with
 * the PMLs graded as the CPMLs are, in the transform's units (divide by the waveform's
 * spectrum for FDFD's amplitudes). */
int pxfdtd_simulation_mode_amplitudes(pxfdtd_simulation *s, int32_t n, double *out);
/* GaussianBeam (src/fdtd/sources.rs:104-128): the plane's normal axis and node, the direction
 * (0 forward, 1 backward), where the beam's axis crosses the plane along the other two axes
 * (um), the waist w0 (the field's 1/e radius at the focus, um), the focus's distance from the
 * plane along the beam (um), the axis among the other two it tilts towards and the tilt
 * (radians, |angle| < pi/2), E's direction (0 perpendicular to the plane of incidence, 1
 * parallel) and the medium's relative permittivity on the plane. */
typedef struct {
    int32_t axis, plane, direction;
    double centre[2];
    double waist, focus;
    int32_t tilt;
    double angle;
    int32_t polarization;
    double eps;
} pxfdtd_gaussian_beam;
/* Simulation::add_beam (src/fdtd/sources.rs:618-699): the paraxial beam's E on the plane
 * (beam_profile, :1043-1108), decomposed into the grid's own plane waves over the plane's
 * window (PlaneSpectrum, :1110-1300: each wave's normal wavenumber from the grid's dispersion
 * at the leapfrog wavenumber, its normal E making it divergence-free on the grid), launched one
 * way by total-field/scattered-field at the waveform's carrier as fixed currents. The same
 * refusals as the reference's (the plane in a CPML or at an end, a periodic axis, a tilt along
 * the axis or not across a varying axis, |angle| >= pi/2, a non-positive waist or
 * permittivity, no carrier, a medium on the plane not uniform with eps). */
int pxfdtd_simulation_add_beam(pxfdtd_simulation *s, const pxfdtd_gaussian_beam *beam,
                               const pxfdtd_waveform *w);
/* Simulation::add_plane_wave (src/fdtd/sources.rs:701-721): a plane wave of `w` on the
 * total-field/scattered-field box of nodes low .. high, travelling along the integer direction,
 * E along the polarization's part across it, in a uniform background of permittivity eps; its
 * number (for pxfdtd_simulation_incident) in *number. The same refusals as the reference's. */
int pxfdtd_simulation_add_plane_wave(pxfdtd_simulation *s, const int32_t low[3],
                                     const int32_t high[3], const int64_t direction[3],
                                     const double polarization[3], double eps,
                                     const pxfdtd_waveform *w, int32_t *number);
/* Simulation::incident (sources.rs:723-737): plane wave n's incident field now at a value of
 * the grid (E at the current time, H~ half a step behind); 0 beyond its auxiliary grid */
int pxfdtd_simulation_incident(pxfdtd_simulation *s, int32_t n, int32_t field, int32_t comp,
                               int32_t i, int32_t j, int32_t k, double *out);
/* Steps on `dev` in `precision` from now on (Simulation::set_device); NULL: back to none */
int pxfdtd_simulation_set_device(pxfdtd_simulation *s, pxfdtd_device *dev, int32_t precision);
int pxfdtd_simulation_run(pxfdtd_simulation *s, int64_t count);
/* Simulation::run_until (src/fdtd.rs:1160-1167): steps until E's time reaches `time` (um/c) */
int pxfdtd_simulation_run_until(pxfdtd_simulation *s, double time);
/* Simulation::run_until_decayed (src/fdtd/monitors.rs:740-841): runs in blocks of `interval`
 * until, over a whole block, |F|^2 at each probe stayed below `fraction` of its largest, or the
 * time reaches `limit`; *decayed says which. After the sources end (none continuous) it watches
 * for the fields growing again and stops with PXFDTD_ERR_GROWING (the factor in *growth_factor,
 * which may be NULL) */
int pxfdtd_simulation_run_until_decayed(pxfdtd_simulation *s, const int32_t *probes, int32_t n,
                                        double fraction, double interval, double limit,
                                        int32_t *decayed, double *growth_factor);
double pxfdtd_simulation_dt(const pxfdtd_simulation *s);
int64_t pxfdtd_simulation_steps(const pxfdtd_simulation *s);
int64_t pxfdtd_simulation_cells(const pxfdtd_simulation *s);
/* fields: copies into out[N] (brought back from the device if it steps there) */
int pxfdtd_simulation_field(pxfdtd_simulation *s, int32_t which, int32_t comp, double *out);
int pxfdtd_simulation_set_field(pxfdtd_simulation *s, int32_t which, int32_t comp,
                                const double *values);
int pxfdtd_simulation_coefficients(const pxfdtd_simulation *s, int32_t comp, double *ca,
                                   double *cb);
int64_t pxfdtd_simulation_probe_len(pxfdtd_simulation *s, int32_t n);
int pxfdtd_simulation_probe(pxfdtd_simulation *s, int32_t n, double *out);
/* transform monitor n's series (0..5: E x, y, z, H x, y, z), [f][k][j][i] complex */
int64_t pxfdtd_simulation_dft_len(pxfdtd_simulation *s, int32_t n, int32_t series);
int pxfdtd_simulation_dft(pxfdtd_simulation *s, int32_t n, int32_t series, double *out);
pxfdtd_resident *pxfdtd_simulation_resident(pxfdtd_simulation *s);
/* the device bytes the simulation needs in `precision` (fields in two copies, psi, coefficient
 * tables, media state, monitor staging and FP64 sums): what admission by footprint counts */
int pxfdtd_simulation_footprint(pxfdtd_simulation *s, int32_t precision, int64_t *bytes);
/* Yee::bytes_per_step (src/fdtd/kernel/yee.rs:141-153), the reference's byte model of one step:
 * (24 N + 6 per E psi value + 5 per H psi value) values of the precision's size */
int pxfdtd_simulation_bytes_per_step(const pxfdtd_simulation *s, int32_t precision, int64_t *bytes);
/* The dispersive media's share of a step's algorithmic bytes (no counterpart in the reference,
 * whose byte model leaves media out): per value in a medium, E at the previous step read and
 * written and each term's two currents read and written, values of the precision's size */
int pxfdtd_simulation_media_bytes_per_step(const pxfdtd_simulation *s, int32_t precision, int64_t *bytes);

/* ------------------------------------------------------------ level 3: the serving loop */
/* Continuous batching of jobs on one device (no counterpart in the reference; its nearest is
 * the process-level farm, src/farm.rs:1-56). A job is a host-mirror simulation with no device,
 * handed over with a step count; a loop thread admits jobs in arrival order while the admitted
 * jobs' device footprint fits the budget and their number max_jobs, steps all admitted jobs
 * together in segments of `segment` steps (one batched launch per kernel class and step,
 * replayed from a captured hipGraph while the job set stays), and admits new arrivals at each
 * segment boundary. A finished job's fields, psi and FP64 transform sums go back to its
 * simulation, whose device is then none again, and its device memory is freed. Each job's
 * results are the same bits as running it alone. Jobs step in f32; probes are not recorded
 * (a job with probes is refused). The caller leaves a submitted simulation alone until its
 * job is done, failed or cancelled. */
typedef struct pxfdtd_server pxfdtd_server;
typedef struct {
    int64_t hbm_budget;  /* device bytes the admitted jobs may hold; 0: 90 % of the free HBM */
    int32_t max_jobs;    /* jobs admitted at once; 0: 256 */
    int32_t segment;     /* steps per batched segment; 0: 32 */
    int32_t direct;      /* 1: launch each segment's kernels directly instead of a hipGraph */
    int32_t autotune;    /* trial launches per tiling for the first 3D job of each class of
                            grid shape (each extent's power of two); 0: two rows per strip and
                            8-plane chunks, which suit a full batch (2D: one row, chunks of up
                            to 16 rows either way) */
    int32_t order;       /* 0: every admitted job steps in every segment, jobs admitted in
                            arrival order; 1: shortest remaining work first: a segment steps the
                            admitted jobs with the fewest cell-updates left until their cells
                            reach fill_cells (at least one job), and the queue admits its
                            smallest job first. The others wait resident, so short jobs finish
                            early; results are the same bits either way */
    int32_t pad;
    int64_t fill_cells;  /* order 1: cells a segment's jobs fill the device with; 0: 2^26 */
} pxfdtd_server_config;
enum {
    PXFDTD_JOB_QUEUED = 0,
    PXFDTD_JOB_RUNNING = 1,
    PXFDTD_JOB_DONE = 2,
    PXFDTD_JOB_FAILED = 3,
    PXFDTD_JOB_CANCELLED = 4
};
typedef struct {
    int64_t submitted, admitted, completed, failed;
    int64_t segments, captures; /* segments stepped; hipGraphs captured (one per job set) */
    double busy_seconds;        /* wall time from each segment's launch to the host seeing its end */
    double cell_updates;        /* cells x steps over all jobs */
    int64_t hbm_in_use;         /* the admitted jobs' footprint now */
    int64_t peak_jobs;          /* the most jobs admitted at once */
    double admit_seconds;       /* host time putting jobs on the device (tables, uploads, tuning) */
    double retire_seconds;      /* host time bringing finished jobs' results back and waiting for the segment */
    double rebuild_seconds;     /* host time building the batched tables and capturing graphs */
    double prepare_seconds;     /* host time choosing a segment's jobs and filling its tables */
    double gap_seconds;         /* host time from one segment's end to the next one's launch */
    double device_seconds;      /* the segments' time on the device, first copy to last kernel */
} pxfdtd_server_stats;
int pxfdtd_server_create(pxfdtd_device *dev, const pxfdtd_server_config *cfg, pxfdtd_server **out);
/* stops the loop: queued jobs are cancelled, running ones cancelled after their segment */
void pxfdtd_server_destroy(pxfdtd_server *s);
int pxfdtd_server_submit(pxfdtd_server *s, pxfdtd_simulation *sim, int64_t steps, int64_t *job);
int pxfdtd_server_state(pxfdtd_server *s, int64_t job, int32_t *state);
/* waits for the job (timeout_seconds <= 0: no limit); PXFDTD_OK once it is done */
int pxfdtd_server_wait(pxfdtd_server *s, int64_t job, double timeout_seconds);
/* Cancels a queued or running job (no counterpart in the reference). A queued job leaves the
 * queue at once; a running one stops at the next segment boundary, its simulation holding the
 * fields and sums of the step it reached, with no device, and its device memory freed. Its
 * state becomes PXFDTD_JOB_CANCELLED (pxfdtd_server_wait reports it); the other jobs go on,
 * their results the same bits. A job already done, failed or cancelled is left as it is. */
int pxfdtd_server_cancel(pxfdtd_server *s, int64_t job);
/* seconds since the server started: submission, admission, completion */
int pxfdtd_server_times(pxfdtd_server *s, int64_t job, double *submitted, double *started,
                        double *finished);
int pxfdtd_server_get_stats(pxfdtd_server *s, pxfdtd_server_stats *out);

/* ------------------------------------------------------- level 4: the reverse-time adjoint */
/* The gradient of J = |S|^2, S = sum over a transform monitor's values of m(r) times its FP64
 * sum of E's component c at one frequency, with respect to cb of chosen E values (for a
 * lossless value cb = dt / eps, so dJ/deps = -dJ/dcb cb^2 / dt). It is the exact transpose of
 * the device's time stepping (fields and CPML), run backwards from the last step with the state
 * recomputed from checkpoints taken every `checkpoint` steps in HBM. photonoxide's adjoint is
 * frequency-domain with no checkpointing (src/fdtd/adjoint.rs:43-46,496-569); this one has no
 * counterpart. The simulation has a device and has not run; the design values lie outside the
 * CPML slabs, carry no E source, and are lossless and not walls; dispersive media are refused. */
typedef struct pxfdtd_adjoint pxfdtd_adjoint;
typedef struct {
    int32_t monitor;          /* the transform monitor (pxfdtd_simulation_add_dft's number) */
    int32_t component;        /* E's component the objective weighs; -1 all three, the weights
                                 then [c][k][j][i] */
    int32_t frequency;        /* the index of the monitor's frequency */
    const double *weights;    /* m over the monitor's box, [k][j][i] complex */
    int64_t ndesign;
    const int32_t *design_component;
    const int64_t *design_index;  /* (k ny + j) nx + i */
    int32_t checkpoint;       /* steps between checkpoints */
} pxfdtd_adjoint_spec;
int pxfdtd_adjoint_create(pxfdtd_simulation *sim, const pxfdtd_adjoint_spec *spec,
                          pxfdtd_adjoint **out);
/* runs `steps` steps from step 0 with checkpoints; the objective J */
int pxfdtd_adjoint_forward(pxfdtd_adjoint *a, int64_t steps, double *objective);
/* dJ/dcb per design value (FP64); the simulation is left as the forward run left it */
int pxfdtd_adjoint_backward(pxfdtd_adjoint *a, double *grad_cb);
int pxfdtd_adjoint_info(const pxfdtd_adjoint *a, int64_t *checkpoints, int64_t *checkpoint_bytes,
                        int64_t *recomputed_steps);
void pxfdtd_adjoint_destroy(pxfdtd_adjoint *a);

/* ------------------------------------- designs and the frequency-domain adjoint (two runs) */
/* Design (src/fdtd/adjoint.rs:63-75): the cells [lo, hi) along each axis, each with a density
 * rho and eps(rho) = eps[0] + rho (eps[1] - eps[0]); densities numbered i fastest, then j, k. */
typedef struct {
    int32_t lo[3], hi[3];
    double eps[2];
} pxfdtd_design;
/* Simulation::with_design (adjoint.rs:366-423): every E value whose cell reaches the design's
 * cells gets the average (8^3 samples, harmonic along the component) of eps(rho) in the design's
 * cells and `background` elsewhere; before the run starts. Refused: a design off the grid or
 * whose values reach a CPML, densities of the wrong count or giving eps <= 0, values with
 * conductivity or a dispersive medium. */
int pxfdtd_simulation_with_design(pxfdtd_simulation *s, const pxfdtd_design *d, const double *density,
                                  int64_t n, const pxfdtd_scene *background);
/* Simulation::add_design_monitor (adjoint.rs:435-470): E's transforms over the values the
 * design's cells reach at the objective's frequencies (c/um), from the next step on. */
int pxfdtd_simulation_add_design_monitor(pxfdtd_simulation *s, const pxfdtd_design *d, int32_t nf,
                                         const double *freqs, int32_t *number);
/* Design::values (adjoint.rs:196-255): how many values design monitor n reads (returned), and
 * with comp_ijk non-null each one's (component, i, j, k), in the order the gradients take them. */
int64_t pxfdtd_simulation_design_values(const pxfdtd_simulation *s, int32_t n, int32_t *comp_ijk);
/* A design monitor's transform of each value it reads at frequency f, out [values][2]. */
int pxfdtd_simulation_design_transform(pxfdtd_simulation *s, int32_t n, int32_t f, double *out);
/* Term (adjoint.rs:285-338), what an objective's derivative dF/da (Wirtinger, conj(a) held) is
 * taken of: kind 0 a mode monitor's amplitude (monitor, mode, direction 0 forward 1 backward),
 * 1 a flux monitor's flux at its frequency number `frequency` (derivative_re real), 2 a
 * transform monitor's value of field's component at `at`, frequency number `frequency`. */
typedef struct {
    int32_t kind;
    int32_t monitor, mode, direction, frequency, field, component, at[3];
    double derivative_re, derivative_im;
} pxfdtd_term;
/* Simulation::adjoint (adjoint.rs:496-569): the adjoint run, at rest, for the objective whose
 * derivatives the terms give: the same grid, media and boundaries, no sources or monitors but
 * design monitor `monitor` (transforms at zero), and currents J = -g/D, M = g/D at the values
 * the terms read, D the CPMLs' symmetrizer (adjoint.rs:704-734), each a sum of Gaussian pulses
 * `bandwidth` wide (c/um), two a quarter period apart per frequency, whose transforms at the
 * design monitor's frequencies are the targets (the real system of adjoint.rs:575-621). Run the
 * forward and the adjoint until their fields die away, then pxfdtd_simulation_design_gradient. */
int pxfdtd_simulation_adjoint(pxfdtd_simulation *s, const pxfdtd_term *terms, int32_t n, int32_t monitor,
                              double bandwidth, pxfdtd_simulation **out);
/* Simulation::permittivity_gradient (adjoint.rs:831-855): dF/deps per value design monitor
 * `monitor` reads (pxfdtd_simulation_design_values' order), sum over the frequencies of
 * 2 Re(i w~ E_adj E), w~ = (2/dt) sin(w dt/2), from this run (the forward) and `adjoint`. */
int pxfdtd_simulation_permittivity_gradient(pxfdtd_simulation *s, pxfdtd_simulation *adjoint, int32_t monitor,
                                            double *out);
/* Simulation::gradient (adjoint.rs:857-879): dF/drho per design cell, the permittivity gradient
 * times (eps[1] - eps[0]) times each cell's share of each value's cell. */
int pxfdtd_simulation_design_gradient(pxfdtd_simulation *s, pxfdtd_simulation *adjoint, int32_t monitor,
                                      double *out);
/* Simulation::run_until_converged (adjoint.rs:904-964): runs in blocks of `interval` (um/c)
 * until, after every source has finished, no series of any transform monitor has moved over a
 * block by more than `tolerance` of its own largest value, or until the time reaches `limit`;
 * *settled whether they had. */
int pxfdtd_simulation_run_until_converged(pxfdtd_simulation *s, double tolerance, double interval, double limit,
                                          int32_t *settled);

/* The library's portable maths, evaluated once (no counterpart in the reference's API: these are
 * the functions of src/portable that the host layer computes its waveforms, coefficients, beams,
 * port modes and resonances with), so that tests can hold their bits to the reference's hashes over
 * its argument grid (src/portable/tests.rs:400-548). `name` is the reference's: "exp", "ln",
 * "ln_1p", "asin", "atan", "acosh" (of a); "atan2" (y = a, x = b), "hypot", "powf", "powi" (n = b,
 * an integer); "c_sqrt", "c_asin", "c_ln", "c_polar" (|z| and arg z) of z = a + ib. A real result
 * in *re, *im = 0. PXFDTD_ERR_INVALID for another name. */
int pxfdtd_portable(const char *name, double a, double b, double *re, double *im);

#ifdef __cplusplus
}
#endif
#endif
