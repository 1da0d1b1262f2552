// Level 2 of the C ABI (include/pxfdtd.h): the host mirror of photonoxide's fdtd::Simulation
// (src/fdtd.rs:470-1590). It owns the problem in FP64 as the reference does: the averaged
// permittivity, dt, the walls, the CPML's coefficients and slabs, conductivity, conductors,
// sources, currents, dipoles, probes, transform monitors and dispersive media. It steps only on
// a device, through level 1 (pxfdtd_resident_*): there is no CPU stepping here, and running
// without a device is an error.
//
// Each set-up computation follows the reference's operation order (cited per function), so the
// tables it hands the device are the reference's bits.
#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <omp.h>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/pxfdtd.h"
#include "portable.hpp"
#include "eigen.hpp"
#include "harmonic.hpp"

using pxfdtd::C64;

int px_fail(int code, const char *fmt, ...);  // resident.hip
int flux_to_host(pxfdtd_device *dev, const pxfdtd_flux_face *faces, int32_t n, double *out, void *stream);  // flux.hip
int averaged_on_device(pxfdtd_device *dev, const pxfdtd_grid *g, const pxfdtd_scene *sc, double *eps3n);  // coefficients.hip

#define TRY(x)                            \
    do {                                  \
        int rc_ = (x);                    \
        if (rc_ != PXFDTD_OK) return rc_; \
    } while (0)

namespace {

constexpr double TAU = 6.283185307179586;

struct Slab {
    int field, comp, axis, side, from, count;
    std::vector<double> b, a, psi;
};

struct SourceRec {
    int field, comp;
    long long r;
    pxfdtd_waveform w;
};

struct CurrentRec {
    int field;
    std::vector<int32_t> comps;
    std::vector<int64_t> r;
    std::vector<C64> amps;
    pxfdtd_waveform w;
};

// A plane wave's run (src/fdtd/sources.rs:742-766): its auxiliary 1D grid, whose state lives
// on the device while it steps there, and its box's couplings.
struct PlaneRec {
    int64_t g[3];
    long long base, pad, len;
    std::vector<double> ca, cb, da, db;
    std::vector<double> e, h;  // [3][len] each
    int nsheet = 0;
    int32_t sheet_c[6];
    int64_t sheet_i[6];
    double sheet_j[6];
    pxfdtd_waveform w;
    std::vector<pxfdtd_coupling> couplings;
};

struct ProbeRec {
    int field, comp;
    long long r;
    std::vector<double> values;
};

struct DftRec {
    std::vector<double> freqs;
    int lo[3], hi[3];
    std::vector<double> acc;  // host-held sums of earlier device layouts, 6 series concatenated
};

// A flux plane (src/fdtd/monitors.rs:210-350): halfway between nodes `plane` and plane + 1
// along `axis`, within a window between half-planes along each of the other two axes (or
// whole); its transforms live in a transform monitor of its own over the plane's nodes either
// side, taken with the flux's sign (-1 for a box's low faces).
struct FluxFace {
    int axis, plane;
    bool windowed[2];
    int wlo[2], whi[2];
    double sign;
    int dft;  // its transform monitor, an index into the simulation's
};

struct FluxRec {
    std::vector<FluxFace> faces;
};

// A mode monitor (Simulation::add_mode_monitor, src/fdtd/monitors.rs:569-611): per mode, a copy
// of the mode and the transform monitor of E over its plane and the next, at the frequency
// whose leapfrog wavenumber is the mode's k0
struct ModeEntry {
    int32_t axis, plane, size[2];
    double step, k0, beta[2];
    std::vector<double> tangential, normal;  // as pxfdtd_port_mode holds them
    double frequency;
    int dft;  // an index into the simulation's transform monitors

    pxfdtd_port_mode view() const {
        return {axis, plane, {size[0], size[1]}, step, k0, {beta[0], beta[1]}, tangential.data(), normal.data()};
    }
};

struct ModeMonRec {
    std::vector<ModeEntry> modes;
};

// A value of E whose cell reaches a design's cells, with each design cell's share of that cell
// (Touched, src/fdtd/adjoint.rs:77-86)
struct Touched {
    int comp;
    int at[3];
    long long r;
    std::vector<std::pair<long long, double>> shares;
};

// A design monitor (DesignMonitor, adjoint.rs:263-283): the design, the values its cells reach
// and the transform monitor of E over each component's box of them
struct DesignRec {
    pxfdtd_design d;
    std::vector<Touched> touched;
    std::vector<double> freqs;
    int dft[3];  // per component, an index into the simulation's transform monitors, or -1
};

struct Group {
    std::vector<int32_t> comp;
    std::vector<int64_t> r;
    double chi;
    int poles;
    double coef[4][3];
};

double smoothstep(double t, double ramp) {
    if (ramp > 0.0) {
        double s = std::clamp(t / ramp, 0.0, 1.0);
        return s * s * (3.0 - 2.0 * s);
    }
    return 1.0;
}

// Waveform::at (src/fdtd.rs:399-423)
double wave_at(const pxfdtd_waveform &w, double t) {
    const double omega = TAU * w.frequency;
    if (w.kind == 0) {
        const double u = (t - w.delay) / w.width;
        double s, c;
        pxfdtd::psincos(omega * (t - w.delay), s, c);
        return pxfdtd::pexp(-u * u) * s;
    }
    if (w.kind == 1) {
        const double u = (t - w.delay) / w.width;
        return -u * pxfdtd::pexp(-u * u) * std::sqrt(2.0 * 2.718281828459045);
    }
    const C64 e = pxfdtd::pcexp({0.0, -omega * t});
    return smoothstep(t, w.ramp) * (w.amp_re * e.re - w.amp_im * e.im);
}

// Waveform::complex_at (src/fdtd.rs:329-351)
C64 wave_complex_at(const pxfdtd_waveform &w, double t) {
    const double omega = TAU * w.frequency;
    if (w.kind == 0) {
        const double u = (t - w.delay) / w.width;
        const double g = pxfdtd::pexp(-u * u);
        const C64 e = pxfdtd::pcexp({0.0, -omega * (t - w.delay)});
        return {0.0 * e.re - g * e.im, 0.0 * e.im + g * e.re};
    }
    if (w.kind == 1) return {wave_at(w, t), 0.0};
    const double s = smoothstep(t, w.ramp);
    const C64 e = pxfdtd::pcexp({0.0, -omega * t});
    const double ar = s * w.amp_re, ai = s * w.amp_im;
    return {ar * e.re - ai * e.im, ar * e.im + ai * e.re};
}

double scene_value(const pxfdtd_scene *sc, int comp, double x, double y, double z) {
    double v = sc->background[comp];
    for (int q = 0; q < sc->n; q++) {
        const pxfdtd_shape &s = sc->shapes[q];
        bool in;
        if (s.kind == 0)
            in = s.lo[0] < x && x < s.hi[0] && s.lo[1] < y && y < s.hi[1] && s.lo[2] < z && z < s.hi[2];
        else {
            const double dx = x - s.lo[0], dy = y - s.lo[1];
            in = dx * dx + dy * dy < s.hi[0] * s.hi[0] && s.lo[2] < z && z < s.hi[2];
        }
        if (in) v = s.value[comp];
    }
    return v;
}

// the permittivity of E's value's cell from samples^3 points of eps(x, y, z), harmonic along
// the component, arithmetic across, with num-complex's division of a real by a complex,
// e / (e e + 0 0) (src/fdfd/three.rs:563-588)
template <typename F>
double averaged_with(const F &eps, int comp, const double centre[3], const double h[3]) {
    const int samples = 8;
    const double n = samples;
    const int a = comp, b = (comp + 1) % 3, c = (comp + 2) % 3;
    double sum = 0.0, p[3] = {centre[0], centre[1], centre[2]};
    for (int sb = 0; sb < samples; sb++) {
        p[b] = centre[b] + ((sb + 0.5) / n - 0.5) * h[b];
        for (int sc_ = 0; sc_ < samples; sc_++) {
            p[c] = centre[c] + ((sc_ + 0.5) / n - 0.5) * h[c];
            double inverse = 0.0;
            for (int sa = 0; sa < samples; sa++) {
                p[a] = centre[a] + ((sa + 0.5) / n - 0.5) * h[a];
                const double e = eps(p[0], p[1], p[2]);
                inverse += 1.0 * e / (e * e + 0.0 * 0.0);
            }
            sum += n * inverse / (inverse * inverse + 0.0 * 0.0);
        }
    }
    return sum / (n * n);
}

double averaged(const pxfdtd_scene *sc, int comp, const double centre[3], const double h[3]) {
    return averaged_with([&](double x, double y, double z) { return scene_value(sc, comp, x, y, z); }, comp, centre,
                         h);
}

}  // namespace

struct pxfdtd_simulation {
    pxfdtd_grid g;
    pxfdtd_boundaries b;
    int n[3];
    long long N;
    double dt;
    long long steps = 0;
    std::vector<double> e, h, ca, cb;
    // [3][N] the permittivity each value of E sees (the port-mode solver's); allocated without
    // a fill, the parallel loop that averages it touching it first
    std::unique_ptr<double[]> eps;
    std::vector<double> sigma;  // [3][N] the conductivity given, when one was (lossy port modes)
    std::vector<double> kn[3], kh[3], fe[3], fh[3];
    std::vector<Slab> slabs;
    std::vector<double> cpml_b[3][2][2], cpml_a[3][2][2];
    std::vector<SourceRec> sources;
    std::vector<CurrentRec> currents;
    std::vector<PlaneRec> planes;
    std::vector<ProbeRec> probes;
    std::vector<DftRec> dfts;
    std::vector<int> shown;  // the caller's transform monitors: their indices in dfts
    std::vector<FluxRec> fluxes;
    std::vector<ModeMonRec> modemons;
    std::vector<DesignRec> designs;
    std::vector<Group> groups;
    std::vector<int32_t> member;  // [3][N] group or -1, when media exist
    // the device
    pxfdtd_device *dev = nullptr;
    int precision = PXFDTD_F32;
    pxfdtd_resident *res = nullptr;
    bool device_newer = false;  // the device's fields are newer than the host's
    bool host_newer = false;    // the host's fields changed since the last upload
    bool fields_zero = true;    // e, h and every slab's psi hold +0.0 alone, as made
    bool inj_dirty = true, taps_dirty = true;
    int planes_on_device = 0;      // the plane waves the resident steps
    bool planes_device_newer = false;  // their auxiliary fields there are newer than the host's
    int emitters = 0;

    void e_position(int comp, long long r, double p[3]) const {
        const int i = (int)(r % g.nx), j = (int)((r / g.nx) % g.ny), k = (int)(r / ((long long)g.nx * g.ny));
        p[0] = g.x0 + i * g.dx;
        p[1] = g.y0 + j * g.dy;
        p[2] = g.z0 + k * g.dz;
        const double hs[3] = {g.dx, g.dy, g.dz};
        p[comp] += 0.5 * hs[comp];
    }
};

namespace {

int check(const pxfdtd_grid *g, const pxfdtd_boundaries *b, double courant) {
    const int n[3] = {g->nx, g->ny, g->nz};
    if ((long long)g->nx * g->ny * g->nz <= 0 || !(g->dx > 0) || !(g->dy > 0) || !(g->dz > 0))
        return px_fail(PXFDTD_ERR_INVALID, "the grid needs cells and positive steps");
    if (!(b->reflection > 0 && b->reflection < 1) || !(b->order >= 0 && std::isfinite(b->order)) ||
        !(b->kappa >= 1 && std::isfinite(b->kappa)) || !(b->alpha >= 0 && std::isfinite(b->alpha)) ||
        (b->has_sigma && !(b->sigma >= 0 && std::isfinite(b->sigma))))
        return px_fail(PXFDTD_ERR_INVALID,
                       "the CPML needs R in (0, 1), an order >= 0, kappa >= 1, alpha >= 0 and sigma >= 0");
    for (int a = 0; a < 3; a++)
        if (!b->periodic[a] && (b->low[a] < 0 || b->high[a] < 0 || b->low[a] + b->high[a] >= n[a]))
            return px_fail(PXFDTD_ERR_INVALID, "CPMLs of %d and %d cells leave nothing of %d cells",
                           b->low[a], b->high[a], n[a]);
    if (!(courant > 0 && courant <= 1))
        return px_fail(PXFDTD_ERR_INVALID, "the Courant number must be in (0, 1], got %g", courant);
    return PXFDTD_OK;
}

// src/fdtd.rs:799-869
void build_cpml(pxfdtd_simulation *s) {
    const pxfdtd_boundaries &cp = s->b;
    const double hs[3] = {s->g.dx, s->g.dy, s->g.dz};
    for (int a = 0; a < 3; a++) {
        s->kn[a].assign(s->n[a], 1.0);
        s->kh[a].assign(s->n[a], 1.0);
    }
    for (int axis = 0; axis < 3; axis++) {
        if (cp.periodic[axis]) continue;
        const int low = cp.low[axis], high = cp.high[axis], n = s->n[axis];
        const double h = hs[axis];
        auto profile = [&](double depth, int cells, double &sig, double &kap) {
            const double d = cells * h;
            const double x = pxfdtd::ppow(std::clamp(depth / cells, 0.0, 1.0), cp.order);
            const double smax = cp.has_sigma ? cp.sigma : (cp.order + 1.0) * (-pxfdtd::pln(cp.reflection)) / (2.0 * d);
            sig = smax * x;
            kap = 1.0 + (cp.kappa - 1.0) * x;
        };
        auto at = [&](double pos, double &sig, double &kap) {
            if (low > 0 && pos < low)
                profile(low - pos, low, sig, kap);
            else if (high > 0 && pos > n - high)
                profile(pos - (n - high), high, sig, kap);
            else {
                sig = 0.0;
                kap = 1.0;
            }
        };
        for (int m = 0; m < n; m++) {
            double sg, kp;
            at(m, sg, kp);
            s->kn[axis][m] = 1.0 / kp;
            at(m + 0.5, sg, kp);
            s->kh[axis][m] = 1.0 / kp;
        }
        const int froms[2] = {0, n - high}, counts[2] = {low, high};
        for (int side = 0; side < 2; side++) {
            if (!counts[side]) continue;
            for (int f = 0; f < 2; f++) {
                std::vector<double> bb, aa;
                for (int m = froms[side]; m < froms[side] + counts[side]; m++) {
                    double sg, kp;
                    at(m + (f == 0 ? 0.0 : 0.5), sg, kp);
                    const double bv = pxfdtd::pexp(-(sg / kp + cp.alpha) * s->dt);
                    const double den = sg * kp + kp * kp * cp.alpha;
                    bb.push_back(bv);
                    aa.push_back(sg > 0.0 && den > 0.0 ? sg * (bv - 1.0) / den : 0.0);
                }
                s->cpml_b[axis][side][f] = bb;
                s->cpml_a[axis][side][f] = aa;
                for (int cc = 1; cc <= 2; cc++) {
                    Slab sl;
                    sl.field = f;
                    sl.comp = (axis + cc) % 3;
                    sl.axis = axis;
                    sl.side = side;
                    sl.from = froms[side];
                    sl.count = counts[side];
                    sl.b = bb;
                    sl.a = aa;
                    sl.psi.assign((size_t)counts[side] * (s->N / n), 0.0);
                    s->slabs.push_back(std::move(sl));
                }
            }
        }
    }
    for (int a = 0; a < 3; a++) {
        s->fe[a].resize(s->n[a]);
        s->fh[a].resize(s->n[a]);
        for (int m = 0; m < s->n[a]; m++) {
            s->fe[a][m] = s->kn[a][m] / hs[a];
            s->fh[a][m] = s->kh[a][m] / hs[a];
        }
    }
}

// src/fdtd.rs:680-796 with the permittivity per value from `eps(c, r)`
template <typename F>
int make(const pxfdtd_grid *g, const pxfdtd_boundaries *b, double courant, F eps,
         pxfdtd_simulation **out) {
    TRY(check(g, b, courant));
    auto *s = new pxfdtd_simulation();
    s->g = *g;
    s->b = *b;
    s->n[0] = g->nx;
    s->n[1] = g->ny;
    s->n[2] = g->nz;
    s->N = (long long)g->nx * g->ny * g->nz;
    const double hs[3] = {g->dx, g->dy, g->dz};
    double sum = 0.0;
    for (int a = 0; a < 3; a++)
        if (!(b->periodic[a] && s->n[a] == 1)) sum += pxfdtd::ppowi(hs[a], -2);
    if (sum == 0.0) {
        delete s;
        return px_fail(PXFDTD_ERR_INVALID, "the field varies along no axis");
    }
    s->dt = courant / std::sqrt(sum);
    const long long N = s->N;
    s->e.assign(3 * N, 0.0);
    s->h.assign(3 * N, 0.0);
    s->ca.assign(3 * N, 1.0);
    s->cb.assign(3 * N, 0.0);
    s->eps.reset(new double[(size_t)(3 * N)]);
    int bad = 0;
    for (int c = 0; c < 3; c++) {
        // each value's average alone: the same bits on any number of threads
#pragma omp parallel for schedule(static) reduction(| : bad)
        for (long long r = 0; r < N; r++) {
            const double v = eps(c, r);
            if (!(std::isfinite(v) && v > 0)) bad |= 1;
            s->eps[c * N + r] = v;
            s->cb[c * N + r] = s->dt / v;
        }
    }
    if (bad) {
        delete s;
        return px_fail(PXFDTD_ERR_INVALID, "the permittivity must be finite and positive");
    }
    for (int c = 0; c < 3; c++)
        for (int axis = 0; axis < 3; axis++) {
            if (axis == c || b->periodic[axis]) continue;
            for (long long r = 0; r < N; r++) {
                const long long at[3] = {r % g->nx, (r / g->nx) % g->ny, r / ((long long)g->nx * g->ny)};
                if (at[axis] == 0) s->ca[c * N + r] = s->cb[c * N + r] = 0.0;
            }
        }
    build_cpml(s);
    *out = s;
    return PXFDTD_OK;
}

}  // namespace

extern "C" int pxfdtd_simulation_new(const pxfdtd_grid *g, const pxfdtd_boundaries *b,
                                     const pxfdtd_scene *eps, double courant,
                                     pxfdtd_simulation **out) {
    if (!eps) return px_fail(PXFDTD_ERR_INVALID, "a scene is required");
    const double hs[3] = {g->dx, g->dy, g->dz};
    const pxfdtd_grid gg = *g;
    if (eps->n == 0) {
        // a scene of its background alone: every value's average is the same arithmetic on the
        // same samples, so it is computed once per component
        const double o[3] = {gg.x0, gg.y0, gg.z0};
        const double u[3] = {averaged(eps, 0, o, hs), averaged(eps, 1, o, hs), averaged(eps, 2, o, hs)};
        return make(g, b, courant, [&](int c, long long) { return u[c]; }, out);
    }
    return make(g, b, courant, [&](int c, long long r) {
        const int i = (int)(r % gg.nx), j = (int)((r / gg.nx) % gg.ny), k = (int)(r / ((long long)gg.nx * gg.ny));
        double p[3] = {gg.x0 + i * gg.dx, gg.y0 + j * gg.dy, gg.z0 + k * gg.dz};
        p[c] += 0.5 * hs[c];
        return averaged(eps, c, p, hs);
    }, out);
}

namespace pxfdtd_host {
double averaged_at(const pxfdtd_scene *sc, int comp, const double centre[3], const double h[3]) {
    return averaged(sc, comp, centre, h);
}
}  // namespace pxfdtd_host

// Simulation::diagonal with the averaging on the device (coefficients.hip): the same bits as
// pxfdtd_simulation_new
extern "C" int pxfdtd_simulation_new_on_device(pxfdtd_device *dev, const pxfdtd_grid *g, const pxfdtd_boundaries *b,
                                               const pxfdtd_scene *eps, double courant, pxfdtd_simulation **out) {
    if (!eps) return px_fail(PXFDTD_ERR_INVALID, "a scene is required");
    if (!dev) return px_fail(PXFDTD_ERR_INVALID, "a device is required");
    TRY(check(g, b, courant));
    const long long N = (long long)g->nx * g->ny * g->nz;
    std::unique_ptr<double[]> e(new (std::nothrow) double[(size_t)(3 * N)]);
    if (!e) return px_fail(PXFDTD_ERR_MEMORY, "no host memory for %lld averaged values", 3 * N);
    TRY(averaged_on_device(dev, g, eps, e.get()));
    return make(g, b, courant, [&](int c, long long r) { return e[c * N + r]; }, out);
}

extern "C" int pxfdtd_simulation_new_eps(const pxfdtd_grid *g, const pxfdtd_boundaries *b,
                                         const double *eps3n, double courant,
                                         pxfdtd_simulation **out) {
    const long long N = (long long)g->nx * g->ny * g->nz;
    return make(g, b, courant, [&](int c, long long r) { return eps3n[c * N + r]; }, out);
}

extern "C" void pxfdtd_simulation_destroy(pxfdtd_simulation *s) {
    if (!s) return;
    if (s->res) pxfdtd_resident_destroy(s->res);
    delete s;
}

static void drop_resident(pxfdtd_simulation *s);
static int sync_host(pxfdtd_simulation *s);
static int pull_planes(pxfdtd_simulation *s);

// the device's transform sums to the host: the device went on from the host's, so they are the
// whole running sums; when its series are set next, it goes on from these again
static int bank_sums(pxfdtd_simulation *s) {
    if (!s->res || s->taps_dirty || s->dfts.empty()) return PXFDTD_OK;
    std::vector<double> sums(pxfdtd_resident_sums_len(s->res));
    TRY(pxfdtd_resident_read_sums(s->res, sums.data(), (int64_t)sums.size()));
    size_t at = 0;
    for (auto &d : s->dfts)
        for (auto &v : d.acc) v = sums[at++];
    s->taps_dirty = true;
    return PXFDTD_OK;
}

// src/fdtd.rs:949-992
extern "C" int pxfdtd_simulation_conductivity(pxfdtd_simulation *s, const pxfdtd_scene *sigma) {
    if (!s->groups.empty())
        return px_fail(PXFDTD_ERR_INVALID,
                       "the conductivity comes before the dispersive media: conductivity, then medium");
    // cb is dt/eps until the conductivity is set, so it is set once (the reference's builder
    // would compound a second conductivity into cb while ca took only the second)
    if (!s->sigma.empty())
        return px_fail(PXFDTD_ERR_INVALID, "the conductivity is set once, for the whole grid");
    TRY(sync_host(s));
    const long long N = s->N;
    // every value's conductivity first, so that a refused one leaves the simulation as it was
    std::vector<double> given((size_t)(3 * N), 0.0);
    int bad = 0;
    double bad_value = 0.0;
#pragma omp parallel for schedule(static) if (N > (1 << 16))
    for (long long q = 0; q < 3 * N; q++) {
        if (s->cb[q] == 0.0) continue;
        double p[3];
        s->e_position((int)(q / N), q % N, p);
        const double sg = scene_value(sigma, 0, p[0], p[1], p[2]);
        given[(size_t)q] = sg;
        if (!(std::isfinite(sg) && sg >= 0)) {
#pragma omp critical
            {
                bad = 1;
                bad_value = sg;
            }
        }
    }
    if (bad) return px_fail(PXFDTD_ERR_INVALID, "the conductivity must be finite and >= 0, got %g", bad_value);
    for (long long q = 0; q < 3 * N; q++) {
        double &cb = s->cb[q];
        if (cb == 0.0) continue;
        const double half = given[(size_t)q] * cb / 2.0;
        s->ca[q] = (1.0 - half) / (1.0 + half);
        cb /= 1.0 + half;
    }
    s->sigma = std::move(given);
    drop_resident(s);
    return PXFDTD_OK;
}

static void media_remove(pxfdtd_simulation *s, int c, long long r) {
    if (s->member.empty()) return;
    const int g = s->member[c * s->N + r];
    if (g < 0) return;
    s->member[c * s->N + r] = -1;
    Group &G = s->groups[g];
    size_t w = 0;
    for (size_t v = 0; v < G.r.size(); v++) {
        if (G.comp[v] == c && G.r[v] == r) continue;
        G.comp[w] = G.comp[v];
        G.r[w] = G.r[v];
        w++;
    }
    G.comp.resize(w);
    G.r.resize(w);
}

// src/fdtd.rs:995-1010
extern "C" int pxfdtd_simulation_conductor(pxfdtd_simulation *s, int32_t comp, int32_t i,
                                           int32_t j, int32_t k) {
    if (comp < 0 || comp > 2 || i < 0 || j < 0 || k < 0 || i >= s->n[0] || j >= s->n[1] || k >= s->n[2])
        return px_fail(PXFDTD_ERR_INVALID, "(%d, %d, %d) is outside the grid", i, j, k);
    TRY(sync_host(s));
    const long long r = ((long long)k * s->n[1] + j) * s->n[0] + i;
    s->ca[comp * s->N + r] = 0.0;
    s->cb[comp * s->N + r] = 0.0;
    s->e[comp * s->N + r] = 0.0;
    media_remove(s, comp, r);
    drop_resident(s);
    return PXFDTD_OK;
}

static int inside(const pxfdtd_simulation *s, int i, int j, int k) {
    if (i >= 0 && j >= 0 && k >= 0 && i < s->n[0] && j < s->n[1] && k < s->n[2]) return PXFDTD_OK;
    return px_fail(PXFDTD_ERR_INVALID, "(%d, %d, %d) is outside the %d x %d x %d grid", i, j, k,
                   s->n[0], s->n[1], s->n[2]);
}

static int check_waveform(const pxfdtd_waveform *w) {
    if (w->kind < 0 || w->kind > 2) return px_fail(PXFDTD_ERR_INVALID, "unknown waveform kind %d", w->kind);
    if (w->kind != 1 && !(w->frequency > 0 && std::isfinite(w->frequency)))
        return px_fail(PXFDTD_ERR_INVALID, "a waveform's frequency must be positive and finite");
    if (w->kind != 2 && !(w->width > 0))
        return px_fail(PXFDTD_ERR_INVALID, "a pulse's width must be positive");
    return PXFDTD_OK;
}

extern "C" int pxfdtd_simulation_add_source(pxfdtd_simulation *s, int32_t field, int32_t comp,
                                            int32_t i, int32_t j, int32_t k,
                                            const pxfdtd_waveform *w) {
    TRY(inside(s, i, j, k));
    TRY(check_waveform(w));
    if (field < 0 || field > 1 || comp < 0 || comp > 2)
        return px_fail(PXFDTD_ERR_INVALID, "a source needs a field and a component");
    s->sources.push_back({field, comp, ((long long)k * s->n[1] + j) * s->n[0] + i, *w});
    s->inj_dirty = true;
    return PXFDTD_OK;
}

// src/fdtd/sources.rs:223-240
extern "C" int pxfdtd_simulation_add_current(pxfdtd_simulation *s, int32_t field, int64_t n,
                                             const int32_t *comps, const int64_t *index,
                                             const double *amps, const pxfdtd_waveform *w) {
    TRY(check_waveform(w));
    if (field < 0 || field > 1) return px_fail(PXFDTD_ERR_INVALID, "a current needs a field");
    CurrentRec c;
    c.field = field;
    c.w = *w;
    for (int64_t v = 0; v < n; v++) {
        if (comps[v] < 0 || comps[v] > 2 || index[v] < 0 || index[v] >= s->N)
            return px_fail(PXFDTD_ERR_INVALID, "a current's value %lld is outside the grid", (long long)v);
        if (!std::isfinite(amps[2 * v]) || !std::isfinite(amps[2 * v + 1]))
            return px_fail(PXFDTD_ERR_INVALID, "a current's amplitude must be finite");
        c.comps.push_back(comps[v]);
        c.r.push_back(index[v]);
        c.amps.push_back({amps[2 * v], amps[2 * v + 1]});
    }
    s->currents.push_back(std::move(c));
    s->inj_dirty = true;
    return PXFDTD_OK;
}

// a dipole restricted to the grid by its trilinear weights (src/fdtd/sources.rs:248-315)
extern "C" int pxfdtd_simulation_add_dipole(pxfdtd_simulation *s, int32_t field, int32_t comp,
                                            const double position[3], double amp_re,
                                            double amp_im, const pxfdtd_waveform *w) {
    TRY(check_waveform(w));
    if (!std::isfinite(amp_re) || !std::isfinite(amp_im) || !std::isfinite(position[0]) ||
        !std::isfinite(position[1]) || !std::isfinite(position[2]))
        return px_fail(PXFDTD_ERR_INVALID, "a dipole must be finite");
    const double hs[3] = {s->g.dx, s->g.dy, s->g.dz};
    const double origin[3] = {s->g.x0, s->g.y0, s->g.z0};
    struct W { long long m; double w; };
    W along[3][2];
    double volume = 1.0;
    for (int axis = 0; axis < 3; axis++) {
        const int n = s->n[axis];
        const bool varies = !(n == 1 && s->b.periodic[axis]);
        if (!varies) {
            along[axis][0] = {0, 1.0};
            along[axis][1] = {0, 0.0};
            continue;
        }
        volume *= hs[axis];
        const bool half = field == PXFDTD_E ? comp == axis : comp != axis;
        const double first = origin[axis] + 0 * hs[axis] + (half ? 0.5 * hs[axis] : 0.0);
        const double sp = (position[axis] - first) / hs[axis];
        const bool periodic = s->b.periodic[axis];
        if (!periodic && !(sp >= 0.0 && sp <= n - 1))
            return px_fail(PXFDTD_ERR_INVALID, "the dipole is outside its component's values along axis %d", axis);
        const double m = std::floor(sp);
        const double f = sp - m;
        long long lo, hi;
        if (periodic) {
            lo = (((long long)m % n) + n) % n;
            hi = (((long long)m + 1) % n + n) % n;
        } else {
            lo = (long long)m;
            hi = std::min<long long>((long long)m + 1, n - 1);
        }
        along[axis][0] = {lo, 1.0 * (1.0 - f)};
        along[axis][1] = {hi, 1.0 * f};
    }
    // weights summed per value, values in index order (a BTreeMap in the reference)
    std::map<long long, C64> values;
    for (auto &x : along[0])
        for (auto &y : along[1])
            for (auto &z : along[2]) {
                // complex weights with zero imaginary parts: wx * wy * wz
                const double wv = x.w * y.w * z.w;
                if (wv == 0.0) continue;
                const long long r = (z.m * s->n[1] + y.m) * s->n[0] + x.m;
                // a * (w / volume): (w/volume) real, a complex
                const double q = wv / volume;
                C64 &acc = values[r];
                acc.re += amp_re * q - amp_im * 0.0;
                acc.im += amp_re * 0.0 + amp_im * q;
            }
    CurrentRec c;
    c.field = field;
    c.w = *w;
    for (auto &kv : values) {
        c.comps.push_back(comp);
        c.r.push_back(kv.first);
        c.amps.push_back(kv.second);
    }
    s->currents.push_back(std::move(c));
    s->inj_dirty = true;
    return PXFDTD_OK;
}

// ------------------------------------------------------------------------ plane waves

// twice a value's position along `axis` in steps from node 0: E_c is halfway along c, H~_c
// halfway along the other two (src/fdtd/sources.rs:155-161)
static long long twice(int field, int comp, int axis, long long m) {
    const bool half = field == PXFDTD_E ? comp == axis : comp != axis;
    return 2 * m + (half ? 1 : 0);
}

static int64_t gcd64(int64_t a, int64_t b) { return b == 0 ? (a < 0 ? -a : a) : gcd64(b, a % b); }

static long long mod2(long long v) { return ((v % 2) + 2) % 2; }

// The couplings across a total-field/scattered-field surface among the values in the region
// [lo, hi) (src/fdtd/sources.rs:321-418): each value whose update reads a value on the other
// side of the surface, `total(field, component, at)` saying which side each value is on, with
// the weight that value's difference carries. Values E is held at zero on (cb = 0: walls,
// conductors) neither take nor give a correction.
struct Link { int target, comp; long long r; double weight; int from_comp; long long from[3]; };

template <typename Total>
static std::vector<Link> collect_links(const pxfdtd_simulation *s, const long long lo[3], const long long hi[3],
                                       const Total &total) {
    auto neighbour = [&](const long long at[3], int axis, bool forward, long long out[3]) {
        for (int a = 0; a < 3; a++) out[a] = at[a];
        const long long n = s->n[axis];
        if (forward) {
            if (at[axis] + 1 < n) out[axis]++;
            else if (s->b.periodic[axis]) out[axis] = 0;
            else return false;
        } else if (at[axis] > 0) {
            out[axis]--;
        } else if (s->b.periodic[axis]) {
            out[axis] = n - 1;
        } else {
            return false;
        }
        return true;
    };
    auto index_of = [&](const long long at[3]) { return (at[2] * s->n[1] + at[1]) * s->n[0] + at[0]; };
    std::vector<Link> links;
    const long long N = s->N;
    for (long long k = lo[2]; k < hi[2]; k++)
        for (long long j = lo[1]; j < hi[1]; j++)
            for (long long i = lo[0]; i < hi[0]; i++) {
                const long long at[3] = {i, j, k};
                const long long r = index_of(at);
                for (int c = 0; c < 3; c++) {
                    const int a = (c + 1) % 3, b = (c + 2) % 3;
                    long long nb[3];
                    // E_c: curl = D_a H_b - D_b H_a with back differences
                    if (s->cb[c * N + r] != 0.0) {
                        const bool in = total(PXFDTD_E, c, at);
                        const double fa = s->fe[a][at[a]], fb = s->fe[b][at[b]];
                        auto take = [&](int fc, const long long *from, double coef) {
                            if (from && total(PXFDTD_H, fc, from) != in)
                                links.push_back({PXFDTD_E, c, r, (in ? -1.0 : 1.0) * coef, fc, {from[0], from[1], from[2]}});
                        };
                        take(b, at, fa);
                        take(b, neighbour(at, a, false, nb) ? nb : nullptr, -fa);
                        take(a, at, -fb);
                        take(a, neighbour(at, b, false, nb) ? nb : nullptr, fb);
                    }
                    // H~_c: curl = D_a E_b - D_b E_a with forward differences; held E gives none
                    const bool in = total(PXFDTD_H, c, at);
                    const double fa = s->fh[a][at[a]], fb = s->fh[b][at[b]];
                    auto take = [&](int fc, const long long *from, double coef) {
                        if (!from || s->cb[fc * N + index_of(from)] == 0.0) return;
                        if (total(PXFDTD_E, fc, from) != in)
                            links.push_back({PXFDTD_H, c, r, (in ? 1.0 : -1.0) * coef, fc, {from[0], from[1], from[2]}});
                    };
                    take(b, neighbour(at, a, true, nb) ? nb : nullptr, fa);
                    take(b, at, -fa);
                    take(a, neighbour(at, b, true, nb) ? nb : nullptr, -fb);
                    take(a, at, fb);
                }
            }
    return links;
}

// Simulation::uniform (src/fdtd/sources.rs:511-522): every E value on the couplings' surface in
// a lossless medium of permittivity eps, not dispersive
static bool surface_uniform(const pxfdtd_simulation *s, const std::vector<Link> &links, double eps) {
    const double expected = s->dt / eps;
    const long long N = s->N;
    for (const Link &l : links) {
        const int c = l.target == PXFDTD_E ? l.comp : l.from_comp;
        const long long r = l.target == PXFDTD_E ? l.r : (l.from[2] * s->n[1] + l.from[1]) * s->n[0] + l.from[0];
        const bool media = !s->member.empty() && s->member[c * N + r] >= 0;
        if (!(std::fabs(s->cb[c * N + r] - expected) <= 1e-12 * expected && s->ca[c * N + r] == 1.0 && !media))
            return false;
    }
    return true;
}

// num-complex's product and quotient, operation for operation
static C64 cmul(C64 a, C64 b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
static C64 cdiv(C64 a, C64 b) {
    const double n = b.re * b.re + b.im * b.im;
    return {(a.re * b.re + a.im * b.im) / n, (a.im * b.re - a.re * b.im) / n};
}

// Simulation::phasor_surface (src/fdtd/sources.rs:419-453): per target field, E then H, the
// couplings' sums per (component, value) in that order, zeros dropped, added as currents with
// the waveform. E targets read the incident H~ (`h_of`), H targets the incident E (`e_of`).
template <typename EOf, typename HOf>
static void add_phasor_surface(pxfdtd_simulation *s, const std::vector<Link> &links, const EOf &e_of,
                               const HOf &h_of, const pxfdtd_waveform *w) {
    for (int field : {PXFDTD_E, PXFDTD_H}) {
        std::map<std::pair<int, long long>, C64> values;
        for (const Link &l : links) {
            if (l.target != field) continue;
            const C64 inc = l.target == PXFDTD_E ? h_of(l.from_comp, l.from) : e_of(l.from_comp, l.from);
            C64 &acc = values[{l.comp, l.r}];
            acc.re += l.weight * inc.re;
            acc.im += l.weight * inc.im;
        }
        CurrentRec c;
        c.field = field;
        c.w = *w;
        for (auto &kv : values) {
            if (kv.second.re == 0.0 && kv.second.im == 0.0) continue;
            c.comps.push_back(kv.first.first);
            c.r.push_back(kv.first.second);
            c.amps.push_back(kv.second);
        }
        if (!c.r.empty()) s->currents.push_back(std::move(c));
    }
    s->inj_dirty = true;
}

// carrier_wavenumber (src/fdtd/sources.rs:485-508): the leapfrog's wavenumber at the waveform's
// carrier, (2 / dt) sin(w dt / 2); with `k0` given, it must be that to 1e-9
static int carrier_wavenumber(const pxfdtd_simulation *s, const pxfdtd_waveform *w, const double *k0,
                              const char *what, double *out) {
    if (w->kind == 1)
        return px_fail(PXFDTD_ERR_INVALID, "%s needs a waveform with a carrier (a Gaussian or a continuous wave)", what);
    const double omega = TAU * w->frequency;
    if (omega * s->dt >= M_PI)
        return px_fail(PXFDTD_ERR_INVALID, "%g c/um is beyond the time step's Nyquist frequency", w->frequency);
    double sn = 0.0, cs = 0.0;
    pxfdtd::psincos(omega * s->dt / 2.0, sn, cs);
    const double tilde = 2.0 / s->dt * sn;
    if (k0 && std::fabs(*k0 - tilde) > 1e-9 * tilde)
        return px_fail(PXFDTD_ERR_INVALID,
                       "%s was solved at k0 = %.17g rad/um, but the waveform's carrier runs at the leapfrog's "
                       "%.17g: solve it at the leapfrog's wavelength", what, *k0, tilde);
    *out = tilde;
    return PXFDTD_OK;
}

// PortMode3d::e and value_at (src/fdfd/three/ports.rs:130-184): the mode's E component `comp`
// at the 3D index `at`, going forward (or backward) at unit amplitude on its plane: the profile
// carried along the axis as e^(+-i beta step) per step, the normal component going backward
// negated and moved half a step (-e_a e^(-i beta step)).
static C64 mode_value(const pxfdtd_port_mode *mode, int comp, const long long at[3], bool forward) {
    const int axis = mode->axis, ab = (axis + 1) % 3, ac = (axis + 2) % 3;
    const long long plane_values = (long long)mode->size[0] * mode->size[1];
    const long long k = at[ac] * mode->size[0] + at[ab];
    C64 ev;
    if (comp == axis) {
        ev = {mode->normal[2 * k], mode->normal[2 * k + 1]};
    } else {
        const long long o = comp == ab ? k : plane_values + k;
        ev = {mode->tangential[2 * o], mode->tangential[2 * o + 1]};
    }
    C64 i_beta = cmul({0.0, 1.0}, {mode->beta[0], mode->beta[1]});
    i_beta = {i_beta.re * mode->step, i_beta.im * mode->step};
    const double st = (double)at[axis] - (double)mode->plane;
    if (forward) return cmul(ev, pxfdtd::pcexp({i_beta.re * st, i_beta.im * st}));
    const C64 neg = {-i_beta.re, -i_beta.im};
    const C64 phase = pxfdtd::pcexp({neg.re * st, neg.im * st});
    if (comp == axis) return cmul(cmul({-ev.re, -ev.im}, pxfdtd::pcexp(neg)), phase);
    return cmul(ev, phase);
}

// Simulation::add_mode_source (src/fdtd/sources.rs:550-596): the mode's E from
// PortMode3d::value_at (src/fdfd/three/ports.rs:165-184), H~ from it by h_of (sources.rs:457-
// 483), the couplings of the planes around the mode's by links (:321-418), their currents by
// phasor_surface (:419-453), after carrier_wavenumber's checks (:485-508)
extern "C" int pxfdtd_simulation_add_mode_source(pxfdtd_simulation *s, const pxfdtd_port_mode *mode,
                                                 int32_t direction, const pxfdtd_waveform *w) {
    TRY(check_waveform(w));
    if (!mode || direction < 0 || direction > 1)
        return px_fail(PXFDTD_ERR_INVALID, "a mode source needs a mode and a direction (0 forward, 1 backward)");
    const int axis = mode->axis;
    if (axis < 0 || axis > 2) return px_fail(PXFDTD_ERR_INVALID, "a mode's axis must be 0, 1 or 2");
    const int ab = (axis + 1) % 3, ac = (axis + 2) % 3;
    const double steps[3] = {s->g.dx, s->g.dy, s->g.dz};
    if (mode->size[0] != s->n[ab] || mode->size[1] != s->n[ac] || mode->step != steps[axis])
        return px_fail(PXFDTD_ERR_INVALID, "the mode's plane must be this grid's: solve it with FDFD on the same grid");
    if (s->b.periodic[axis])
        return px_fail(PXFDTD_ERR_INVALID, "a mode source normal to axis %d needs ends along it, not periodic ones", axis);
    const long long n = s->n[axis], p = mode->plane;
    const long long low = s->b.low[axis], high = s->b.high[axis];
    if (p < low + 1 || p + 2 + high > n)
        return px_fail(PXFDTD_ERR_INVALID,
                       "the mode's plane %lld and the next must be clear of the CPMLs (%lld and %lld cells) and "
                       "the ends of %lld cells along axis %d", p, low, high, n, axis);
    const long long plane_values = (long long)mode->size[0] * mode->size[1];
    if (!mode->tangential || !mode->normal) return px_fail(PXFDTD_ERR_INVALID, "a mode needs its profile");
    for (long long q = 0; q < 4 * plane_values; q++)
        if (!std::isfinite(mode->tangential[q])) return px_fail(PXFDTD_ERR_INVALID, "a mode's profile must be finite");
    for (long long q = 0; q < 2 * plane_values; q++)
        if (!std::isfinite(mode->normal[q])) return px_fail(PXFDTD_ERR_INVALID, "a mode's profile must be finite");
    if (!std::isfinite(mode->beta[0]) || !std::isfinite(mode->beta[1]))
        return px_fail(PXFDTD_ERR_INVALID, "a mode's propagation constant must be finite");
    double k0 = 0.0;
    TRY(carrier_wavenumber(s, w, &mode->k0, "the mode", &k0));
    const bool forward = direction == 0;

    auto e_of = [&](int comp, const long long at[3]) -> C64 { return mode_value(mode, comp, at, forward); };
    // h_of: curl E / (i k0), forward differences, zero beyond a wall
    auto h_of = [&](int comp, const long long at[3]) -> C64 {
        const int a = (comp + 1) % 3, b = (comp + 2) % 3;
        auto next = [&](int ax, int f) -> C64 {
            long long m[3] = {at[0], at[1], at[2]};
            if (at[ax] + 1 < s->n[ax]) m[ax]++;
            else if (s->b.periodic[ax]) m[ax] = 0;
            else return {0.0, 0.0};
            return e_of(f, m);
        };
        const double fa = s->fh[a][at[a]], fb = s->fh[b][at[b]];
        const C64 da = next(a, b), eb = e_of(b, at), db = next(b, a), ea = e_of(a, at);
        const C64 t1 = {(da.re - eb.re) * fa, (da.im - eb.im) * fa};
        const C64 t2 = {(db.re - ea.re) * fb, (db.im - ea.im) * fb};
        return cdiv({t1.re - t2.re, t1.im - t2.im}, {0.0, k0});
    };
    auto total = [&](int field, int comp, const long long at[3]) {
        const long long t = twice(field, comp, axis, at[axis]);
        return forward ? t >= 2 * p : t <= 2 * (p + 1);
    };
    long long lo[3] = {0, 0, 0}, hi[3] = {s->n[0], s->n[1], s->n[2]};
    lo[axis] = p - 1;
    hi[axis] = std::min(p + 3, n);
    const std::vector<Link> links = collect_links(s, lo, hi, total);
    add_phasor_surface(s, links, e_of, h_of, w);
    return PXFDTD_OK;
}

// Simulation::add_beam (src/fdtd/sources.rs:618-699) with beam_profile (:1043-1108) and
// PlaneSpectrum (:1110-1350), with the reference's operations and its portable complex square
// root, arcsine and modulus.
extern "C" int pxfdtd_simulation_add_beam(pxfdtd_simulation *s, const pxfdtd_gaussian_beam *beam,
                                          const pxfdtd_waveform *w) {
    TRY(check_waveform(w));
    if (!beam) return px_fail(PXFDTD_ERR_INVALID, "a beam is required");
    const int axis = beam->axis, tilt = beam->tilt;
    if (axis < 0 || axis > 2 || tilt < 0 || tilt > 2 || beam->direction < 0 || beam->direction > 1 ||
        beam->polarization < 0 || beam->polarization > 1)
        return px_fail(PXFDTD_ERR_INVALID, "a beam's axis, tilt, direction or polarization is out of range");
    auto ok = [](double v) { return std::isfinite(v) && v > 0.0; };
    if (!ok(beam->waist) || !ok(beam->eps) || !std::isfinite(beam->focus))
        return px_fail(PXFDTD_ERR_INVALID, "a beam needs a positive waist and permittivity and a finite focus");
    bool varies[3];
    for (int a = 0; a < 3; a++) varies[a] = !(s->n[a] == 1 && s->b.periodic[a]);
    if (tilt == axis || std::isnan(beam->angle) || std::fabs(beam->angle) >= M_PI / 2 ||
        (beam->angle != 0.0 && !varies[tilt]))
        return px_fail(PXFDTD_ERR_INVALID,
                       "a beam tilts towards an axis across its plane along which the field varies, by less than 90 degrees");
    if (s->b.periodic[axis])
        return px_fail(PXFDTD_ERR_INVALID, "a beam normal to axis %d needs ends along it, not periodic ones", axis);
    const long long n = s->n[axis], p = beam->plane;
    const long long low = s->b.low[axis], high = s->b.high[axis];
    if (p < low + 1 || p + 2 + high > n)
        return px_fail(PXFDTD_ERR_INVALID,
                       "the beam's plane %lld and the next must be clear of the CPMLs (%lld and %lld cells) and "
                       "the ends of %lld cells along axis %d", p, low, high, n, axis);
    double k0 = 0.0;
    TRY(carrier_wavenumber(s, w, nullptr, "a beam", &k0));
    const int ab = (axis + 1) % 3, ac = (axis + 2) % 3;
    const double steps[3] = {s->g.dx, s->g.dy, s->g.dz};
    const double origin[3] = {s->g.x0, s->g.y0, s->g.z0};
    long long wlo[2], whi[2];
    for (int q = 0; q < 2; q++) {
        const int t = q == 0 ? ab : ac;
        wlo[q] = s->b.periodic[t] ? 0 : s->b.low[t];
        whi[q] = s->b.periodic[t] ? s->n[t] : s->n[t] - s->b.high[t];
    }
    const bool forward = beam->direction == 0;
    auto node = [&](int a, long long m) { return origin[a] + (double)m * steps[a]; };

    // beam_profile: E of the paraxial beam at a point
    const double sign = forward ? 1.0 : -1.0;
    double sa = 0.0, ca = 0.0;
    pxfdtd::psincos(beam->angle, sa, ca);
    double u[3] = {0.0, 0.0, 0.0};
    u[axis] = sign * ca;
    u[tilt] = sa;
    double crossing[3] = {0.0, 0.0, 0.0};
    crossing[axis] = node(axis, p);
    crossing[ab] = beam->centre[0];
    crossing[ac] = beam->centre[1];
    double waist[3];
    for (int q = 0; q < 3; q++) waist[q] = crossing[q] + beam->focus * u[q];
    const int s_axis = 3 - axis - tilt;
    double sv[3] = {0.0, 0.0, 0.0};
    sv[s_axis] = 1.0;
    double pv[3] = {sv[1] * u[2] - sv[2] * u[1], sv[2] * u[0] - sv[0] * u[2], sv[0] * u[1] - sv[1] * u[0]};
    if (pv[tilt] < 0.0 || (pv[tilt] == 0.0 && pv[axis] < 0.0))
        for (auto &v : pv) v = -v;
    const double *ev = beam->polarization == 0 ? sv : pv;
    const double kk = std::sqrt(beam->eps) * k0;
    const double rayleigh = kk * beam->waist * beam->waist / 2.0;
    const int transverse = (int)varies[ab] + (int)varies[ac];
    auto profile = [&](const double r[3], int comp) -> C64 {
        double rho[3] = {0.0, 0.0, 0.0};
        for (int q = 0; q < 3; q++)
            if (varies[q]) rho[q] = r[q] - waist[q];
        double z = 0.0;
        for (int q = 0; q < 3; q++) z += rho[q] * u[q];
        double r2 = 0.0;
        for (int q = 0; q < 3; q++) r2 += rho[q] * rho[q];
        r2 -= z * z;
        const C64 qq = {z, -rayleigh};
        const C64 ratio = cdiv({0.0, -rayleigh}, qq);
        C64 envelope = ratio;
        if (transverse < 2) {
            envelope = pxfdtd::pcsqrt(ratio);
        }
        // i (k r2 / (2 q) + k z)
        const C64 t1 = cdiv({kk * r2, 0.0}, {2.0 * qq.re, 2.0 * qq.im});
        const C64 inner = {t1.re + kk * z, t1.im};
        const C64 phase = cmul({0.0, 1.0}, inner);
        const C64 value = cmul(envelope, pxfdtd::pcexp(phase));
        return {value.re * ev[comp], value.im * ev[comp]};
    };

    // PlaneSpectrum::new
    const long long nw[2] = {whi[0] - wlo[0], whi[1] - wlo[1]};
    const double st[2] = {steps[ab], steps[ac]};
    std::vector<double> kt[2];
    for (int q = 0; q < 2; q++) {
        const long long lowm = -((nw[q] - 1) / 2);
        for (long long m = 0; m < nw[q]; m++) kt[q].push_back(TAU * (double)(lowm + m) / ((double)nw[q] * st[q]));
    }
    auto positions = [&](int field, int comp, std::vector<double> x[2]) {
        for (int q = 0; q < 2; q++) {
            const int t = q == 0 ? ab : ac;
            const bool half = field == PXFDTD_E ? comp == t : comp != t;
            x[q].clear();
            for (long long m = wlo[q]; m < whi[q]; m++) x[q].push_back(node(t, m) + (half ? 0.5 * steps[t] : 0.0));
        }
    };
    // the separable sum over the other index, from samples to waves (sgn -1) or back (+1)
    auto transform = [&](const std::vector<C64> &values, const std::vector<double> x[2], double sgn) {
        const bool to_waves = sgn < 0.0;
        std::vector<C64> tab[2];
        for (int q = 0; q < 2; q++) {
            const long long m = nw[q];
            tab[q].resize((size_t)(m * m));
            for (long long a = 0; a < m; a++)
                for (long long b = 0; b < m; b++) {
                    const long long wave = to_waves ? a : b, sample = to_waves ? b : a;
                    tab[q][a * m + b] = pxfdtd::pcexp({0.0, sgn * kt[q][wave] * x[q][sample]});
                }
        }
        const long long n0 = nw[0], n1 = nw[1];
        std::vector<C64> along(values.size()), out(values.size());
        for (long long v = 0; v < n1; v++)
            for (long long pp = 0; pp < n0; pp++) {
                C64 sum = {0.0, 0.0};
                for (long long q = 0; q < n0; q++) {
                    const C64 t = cmul(values[v * n0 + q], tab[0][pp * n0 + q]);
                    sum.re += t.re;
                    sum.im += t.im;
                }
                along[v * n0 + pp] = sum;
            }
        for (long long pp = 0; pp < n1; pp++)
            for (long long uu = 0; uu < n0; uu++) {
                C64 sum = {0.0, 0.0};
                for (long long q = 0; q < n1; q++) {
                    const C64 t = cmul(along[q * n0 + uu], tab[1][pp * n1 + q]);
                    sum.re += t.re;
                    sum.im += t.im;
                }
                out[pp * n0 + uu] = sum;
            }
        return out;
    };
    const size_t nwaves = (size_t)(nw[0] * nw[1]);
    std::vector<C64> coefficients[2];
    for (int q = 0; q < 2; q++) {
        const int comp = q == 0 ? ab : ac;
        std::vector<double> x[2];
        positions(PXFDTD_E, comp, x);
        std::vector<C64> samples(nwaves);
        for (long long v = 0; v < nw[1]; v++)
            for (long long uu = 0; uu < nw[0]; uu++) {
                long long at[3];
                at[axis] = p;
                at[ab] = wlo[0] + uu;
                at[ac] = wlo[1] + v;
                double pos[3] = {node(0, at[0]), node(1, at[1]), node(2, at[2])};
                pos[comp] += 0.5 * steps[comp];
                samples[v * nw[0] + uu] = profile(pos, comp);
            }
        coefficients[q] = transform(samples, x, -1.0);
        const double scale = 1.0 / (double)(nw[0] * nw[1]);
        for (auto &c : coefficients[q]) c = {c.re * scale, c.im * scale};
    }
    const double step_a = steps[axis];
    std::vector<C64> ka(nwaves);
    std::vector<std::array<C64, 3>> amplitude(nwaves), big_k(nwaves);
    for (long long mc = 0; mc < nw[1]; mc++)
        for (long long mb = 0; mb < nw[0]; mb++) {
            double s0, c0, s1, c1;
            pxfdtd::psincos(kt[0][mb] * st[0] / 2.0, s0, c0);
            pxfdtd::psincos(kt[1][mc] * st[1] / 2.0, s1, c1);
            const double kb = 2.0 / st[0] * s0, kc = 2.0 / st[1] * s1;
            const double square = beam->eps * k0 * k0 - kb * kb - kc * kc;
            C64 normal = square >= 0.0 ? C64{std::sqrt(square), 0.0} : C64{0.0, std::sqrt(-square)};
            normal = {normal.re * sign, normal.im * sign};
            const C64 as = pxfdtd::pcasin({normal.re * step_a / 2.0, normal.im * step_a / 2.0});
            const C64 k_normal = {2.0 / step_a * as.re, 2.0 / step_a * as.im};
            const size_t wv = (size_t)(mc * nw[0] + mb);
            const C64 eb = coefficients[0][wv], ec = coefficients[1][wv];
            std::array<C64, 3> vec = {C64{0.0, 0.0}, C64{0.0, 0.0}, C64{0.0, 0.0}};
            std::array<C64, 3> kvec = {C64{0.0, 0.0}, C64{0.0, 0.0}, C64{0.0, 0.0}};
            kvec[axis] = normal;
            kvec[ab] = {kb, 0.0};
            kvec[ac] = {kc, 0.0};
            if (pxfdtd::pcabs(normal) > 1e-12 * k0) {
                vec[ab] = eb;
                vec[ac] = ec;
                const C64 num = {kb * eb.re + kc * ec.re, kb * eb.im + kc * ec.im};
                vec[axis] = cdiv({-num.re, -num.im}, normal);
            }
            ka[wv] = k_normal;
            amplitude[wv] = vec;
            big_k[wv] = kvec;
        }

    // the couplings of the planes around the beam's over the window
    auto total = [&](int field, int comp, const long long at[3]) {
        const long long t = twice(field, comp, axis, at[axis]);
        return forward ? t >= 2 * p : t <= 2 * (p + 1);
    };
    long long lo[3], hi[3];
    lo[axis] = p - 1;
    hi[axis] = std::min(p + 3, n);
    lo[ab] = wlo[0];
    hi[ab] = whi[0];
    lo[ac] = wlo[1];
    hi[ac] = whi[1];
    const std::vector<Link> links = collect_links(s, lo, hi, total);
    if (!surface_uniform(s, links, beam->eps))
        return px_fail(PXFDTD_ERR_INVALID,
                       "the medium where the beam enters must be uniform, with its permittivity %g and no conductivity",
                       beam->eps);
    // PlaneSpectrum::planes and synthesize: each plane of values the couplings read
    std::map<std::tuple<int, int, long long>, std::vector<C64>> planes;
    for (const Link &l : links) {
        const int ff = l.target == PXFDTD_E ? PXFDTD_H : PXFDTD_E;
        const auto key = std::make_tuple(ff, l.from_comp, l.from[axis]);
        if (planes.count(key)) continue;
        const double offset = (double)(twice(ff, l.from_comp, axis, l.from[axis]) - 2 * p) * step_a / 2.0;
        std::vector<C64> co(nwaves);
        for (size_t wv = 0; wv < nwaves; wv++) {
            C64 value;
            if (ff == PXFDTD_E) {
                value = amplitude[wv][l.from_comp];
            } else {
                const int x = (l.from_comp + 1) % 3, y = (l.from_comp + 2) % 3;
                const C64 t1 = cmul(big_k[wv][x], amplitude[wv][y]), t2 = cmul(big_k[wv][y], amplitude[wv][x]);
                value = {(t1.re - t2.re) / k0, (t1.im - t2.im) / k0};
            }
            const C64 ik = cmul({0.0, 1.0}, ka[wv]);
            co[wv] = cmul(value, pxfdtd::pcexp({ik.re * offset, ik.im * offset}));
        }
        std::vector<double> x[2];
        positions(ff, l.from_comp, x);
        planes[key] = transform(co, x, 1.0);
    }
    auto lookup = [&](int field, int comp, const long long at[3]) -> C64 {
        const long long uu = at[ab], v = at[ac];
        if (uu < wlo[0] || uu >= whi[0] || v < wlo[1] || v >= whi[1]) return {0.0, 0.0};
        auto it = planes.find(std::make_tuple(field, comp, at[axis]));
        if (it == planes.end()) return {0.0, 0.0};
        return it->second[(size_t)((v - wlo[1]) * nw[0] + (uu - wlo[0]))];
    };
    auto e_of = [&](int comp, const long long at[3]) { return lookup(PXFDTD_E, comp, at); };
    auto h_of = [&](int comp, const long long at[3]) { return lookup(PXFDTD_H, comp, at); };
    add_phasor_surface(s, links, e_of, h_of, w);
    return PXFDTD_OK;
}

// Simulation::add_plane_wave: PlaneWaveRun::new (src/fdtd/sources.rs:774-950) with the box's
// couplings from Simulation::links (:321-418) and the surface check Simulation::uniform
// (:511-522)
extern "C" int pxfdtd_simulation_add_plane_wave(pxfdtd_simulation *s, const int32_t low[3],
                                                const int32_t high[3], const int64_t direction[3],
                                                const double polarization[3], double eps,
                                                const pxfdtd_waveform *w, int32_t *number) {
    TRY(check_waveform(w));
    int64_t g[3] = {direction[0], direction[1], direction[2]};
    const int64_t divisor = gcd64(gcd64(g[0], g[1]), g[2]);
    if (divisor == 0) return px_fail(PXFDTD_ERR_INVALID, "a plane wave's direction can't be zero");
    for (auto &v : g) v /= divisor;
    if (!(std::isfinite(eps) && eps > 0.0))
        return px_fail(PXFDTD_ERR_INVALID, "the background's permittivity must be positive, got %g", eps);
    bool faces[3];
    for (int a = 0; a < 3; a++) {
        const int n = s->n[a];
        faces[a] = !s->b.periodic[a];
        if (!faces[a]) {
            if (low[a] != 0 || high[a] != n || g[a] != 0)
                return px_fail(PXFDTD_ERR_INVALID,
                               "along the periodic axis %d the box must be whole (0 to %d) and the direction "
                               "zero, got %d to %d and %lld", a, n, low[a], high[a], (long long)g[a]);
        } else if (low[a] < s->b.low[a] + 1 || high[a] + s->b.high[a] + 1 > n || low[a] >= high[a]) {
            return px_fail(PXFDTD_ERR_INVALID,
                           "along axis %d the box's faces, nodes %d and %d, must be in order and a cell clear "
                           "of the CPMLs (%d and %d cells) and the ends of %d cells",
                           a, low[a], high[a], s->b.low[a], s->b.high[a], n);
        }
    }
    const double steps[3] = {s->g.dx, s->g.dy, s->g.dz};
    double gv[3], norm2 = 0.0;
    for (int a = 0; a < 3; a++) {
        gv[a] = (double)g[a] / steps[a];
        norm2 += gv[a] * gv[a];
    }
    const double norm = std::sqrt(norm2);
    double khat[3], pt[3], along = 0.0, size2 = 0.0;
    for (int a = 0; a < 3; a++) khat[a] = gv[a] / norm;
    for (int a = 0; a < 3; a++) along += polarization[a] * khat[a];
    for (int a = 0; a < 3; a++) pt[a] = polarization[a] - along * khat[a];
    for (int a = 0; a < 3; a++) size2 += pt[a] * pt[a];
    const double size = std::sqrt(size2);
    if (!(std::isfinite(size) && size > 1e-12))
        return px_fail(PXFDTD_ERR_INVALID, "the polarization has nothing across the direction");
    for (auto &v : pt) v /= size;
    TRY(sync_host(s));  // the waves already stepping keep their state

    // the couplings: each value near the box whose update reads a value on the other side
    auto total = [&](int field, int comp, const long long at[3]) {
        for (int a = 0; a < 3; a++) {
            if (!faces[a]) continue;
            const long long t = twice(field, comp, a, at[a]);
            if (t < 2LL * low[a] || t > 2LL * high[a]) return false;
        }
        return true;
    };
    long long lo[3], hi[3];
    for (int a = 0; a < 3; a++) {
        lo[a] = faces[a] ? low[a] - 1 : 0;
        hi[a] = faces[a] ? high[a] + 2 : s->n[a];
    }
    const long long N = s->N;
    auto index_of = [&](const long long at[3]) { return (at[2] * s->n[1] + at[1]) * s->n[0] + at[0]; };
    const std::vector<Link> links = collect_links(s, lo, hi, total);
    // every E on the surface in a lossless medium of eps, not dispersive
    if (!surface_uniform(s, links, eps))
        return px_fail(PXFDTD_ERR_INVALID,
                       "the medium on the box's surface must be uniform, with the permittivity %g and no "
                       "conductivity", eps);
    (void)index_of;
    // l over the box and a cell around it: its extremes are at the corners
    long long least = 0, most = 0, stride = 0;
    for (int a = 0; a < 3; a++) {
        const long long t0 = faces[a] ? 2LL * low[a] - 2 : 0, t1 = faces[a] ? 2LL * high[a] + 2 : 0;
        least += std::min(g[a] * t0, g[a] * t1);
        most += std::max(g[a] * t0, g[a] * t1);
        stride = std::max<long long>(stride, g[a] < 0 ? -g[a] : g[a]);
    }
    PlaneRec P;
    P.pad = stride;
    const long long absorber = 200 * stride;
    const long long sheet_at = least - 4 * stride - 4;
    P.base = sheet_at - (absorber + 4 * stride + P.pad);
    P.base -= mod2(P.base);
    const long long top = most + 4 * stride + absorber + P.pad;
    P.len = top - P.base + 1;
    // a matched absorber at each end: cubic sigma for E, sigma/eps for H~
    const double depth = (double)absorber * (1.0 / (2.0 * norm));
    const double sqrt_eps = std::sqrt(eps);
    const double sigma_max = 50.0 * sqrt_eps / depth;
    const double dt = s->dt;
    P.ca.assign(P.len, 1.0);
    P.cb.assign(P.len, dt / eps);
    P.da.assign(P.len, 1.0);
    P.db.assign(P.len, dt);
    const long long edge = P.pad + absorber;
    for (long long i = 0; i < P.len; i++) {
        double into;
        if (i < edge) into = (double)(edge - i);
        else if (i + edge >= P.len) into = (double)(i + edge + 1 - P.len);
        else continue;
        const double sigma = sigma_max * pxfdtd::ppowi(std::min(into / (double)absorber, 1.0), 3);
        const double half_e = sigma * dt / (2.0 * eps);
        P.ca[i] = (1.0 - half_e) / (1.0 + half_e);
        P.cb[i] = dt / eps / (1.0 + half_e);
        const double half_h = sigma / eps * dt / 2.0;
        P.da[i] = (1.0 - half_h) / (1.0 + half_h);
        P.db[i] = dt / (1.0 + half_h);
    }
    // the sheet: J = -2 n |G| p on the samples where each component lives, halved over the two
    // neighbours where the sheet's l has the other parity
    for (int c = 0; c < 3; c++) {
        if (pt[c] == 0.0) continue;
        const double jc = -2.0 * sqrt_eps * norm * pt[c];
        if (mod2(sheet_at - g[c]) == 0) {
            P.sheet_c[P.nsheet] = c; P.sheet_i[P.nsheet] = sheet_at - P.base; P.sheet_j[P.nsheet++] = jc;
        } else {
            P.sheet_c[P.nsheet] = c; P.sheet_i[P.nsheet] = sheet_at - 1 - P.base; P.sheet_j[P.nsheet++] = 0.5 * jc;
            P.sheet_c[P.nsheet] = c; P.sheet_i[P.nsheet] = sheet_at + 1 - P.base; P.sheet_j[P.nsheet++] = 0.5 * jc;
        }
    }
    for (int a = 0; a < 3; a++) P.g[a] = g[a];
    P.w = *w;
    P.e.assign(3 * P.len, 0.0);
    P.h.assign(3 * P.len, 0.0);
    for (const Link &l : links) {
        const int from_field = l.target == PXFDTD_E ? PXFDTD_H : PXFDTD_E;
        long long ell = 0;
        for (int a = 0; a < 3; a++) ell += g[a] * twice(from_field, l.from_comp, a, l.from[a]);
        const long long q = ell - P.base;
        if (q < 0 || q >= P.len) return px_fail(PXFDTD_ERR_INVALID, "the auxiliary grid does not span the box");
        P.couplings.push_back({l.target, l.comp, l.r, l.weight, 0, l.from_comp, q});
    }
    s->planes.push_back(std::move(P));
    s->inj_dirty = true;
    if (number) *number = (int32_t)s->planes.size() - 1;
    return PXFDTD_OK;
}

extern "C" int pxfdtd_simulation_incident(pxfdtd_simulation *s, int32_t n, int32_t field, int32_t comp,
                                          int32_t i, int32_t j, int32_t k, double *out) {
    if (n < 0 || n >= (int32_t)s->planes.size()) return px_fail(PXFDTD_ERR_INVALID, "no plane wave %d", n);
    if (field < 0 || field > 1 || comp < 0 || comp > 2) return px_fail(PXFDTD_ERR_INVALID, "a field and a component");
    TRY(inside(s, i, j, k));
    TRY(pull_planes(s));
    const PlaneRec &P = s->planes[n];
    const long long at[3] = {i, j, k};
    long long ell = 0;
    for (int a = 0; a < 3; a++) ell += P.g[a] * twice(field, comp, a, at[a]);
    const long long q = ell - P.base;
    *out = (q < 0 || q >= P.len) ? 0.0 : (field == PXFDTD_E ? P.e : P.h)[comp * P.len + q];
    return PXFDTD_OK;
}

extern "C" int pxfdtd_simulation_add_probe(pxfdtd_simulation *s, int32_t field, int32_t comp,
                                           int32_t i, int32_t j, int32_t k, int32_t *number) {
    TRY(inside(s, i, j, k));
    if (field < 0 || field > 1 || comp < 0 || comp > 2)
        return px_fail(PXFDTD_ERR_INVALID, "a probe needs a field and a component");
    s->probes.push_back({field, comp, ((long long)k * s->n[1] + j) * s->n[0] + i, {}});
    *number = (int32_t)s->probes.size() - 1;
    s->taps_dirty = true;
    return PXFDTD_OK;
}

// src/fdtd/monitors.rs:395-450
static int check_frequencies(const pxfdtd_simulation *s, int32_t nf, const double *freqs) {
    if (nf <= 0) return px_fail(PXFDTD_ERR_INVALID, "a monitor needs at least one frequency");
    for (int f = 0; f < nf; f++)
        if (!(freqs[f] > 0) || TAU * freqs[f] * s->dt >= 3.141592653589793)
            return px_fail(PXFDTD_ERR_INVALID, "%g c/um is beyond the time step's Nyquist frequency", freqs[f]);
    return PXFDTD_OK;
}

// a transform monitor over nodes low..high (inclusive); its index in dfts
static int add_transform(pxfdtd_simulation *s, const int32_t low[3], const int32_t high[3],
                         int32_t nf, const double *freqs, int *index) {
    TRY(check_frequencies(s, nf, freqs));
    TRY(inside(s, high[0], high[1], high[2]));
    for (int a = 0; a < 3; a++)
        if (low[a] < 0 || low[a] > high[a])
            return px_fail(PXFDTD_ERR_INVALID, "a transform's box must run from low to high indices");
    DftRec d;
    d.freqs.assign(freqs, freqs + nf);
    long long vol = 1;
    for (int a = 0; a < 3; a++) {
        d.lo[a] = low[a];
        d.hi[a] = high[a] + 1;
        vol *= d.hi[a] - d.lo[a];
    }
    d.acc.assign(6 * 2 * (size_t)nf * vol, 0.0);
    // sums so far on the device are kept on the host before the device's layout changes
    TRY(bank_sums(s));
    s->dfts.push_back(std::move(d));
    *index = (int)s->dfts.size() - 1;
    s->taps_dirty = true;
    return PXFDTD_OK;
}

extern "C" int pxfdtd_simulation_add_dft(pxfdtd_simulation *s, const int32_t low[3],
                                         const int32_t high[3], int32_t nf, const double *freqs,
                                         int32_t *number) {
    int index;
    TRY(add_transform(s, low, high, nf, freqs, &index));
    s->shown.push_back(index);
    *number = (int32_t)s->shown.size() - 1;
    return PXFDTD_OK;
}

// FluxPlane::range: the indices along the window's axis slot w whose values it reaches
static void flux_range(const pxfdtd_simulation *s, const FluxFace &F, int w, int &from, int &to) {
    const int along = (F.axis + 1 + w) % 3;
    if (F.windowed[w]) {
        from = F.wlo[w];
        to = F.whi[w] + 1;
    } else {
        from = 0;
        to = s->n[along];
    }
}

// FluxPlane::weight: a value's share of the window, halfway after node m (`half`) or on it
static double flux_weight(const FluxFace &F, int w, int m, bool half) {
    if (!F.windowed[w]) return 1.0;
    const int lo = F.wlo[w], hi = F.whi[w];
    if (half) return (lo < m && m < hi) ? 1.0 : ((m == lo || m == hi) ? 0.5 : 0.0);
    return (lo < m && m <= hi) ? 1.0 : 0.0;
}

// FluxPlane::check and the face's transform monitor (FluxPlane::series, monitors.rs:267-316)
static int add_face(pxfdtd_simulation *s, FluxFace &F, int32_t nf, const double *freqs) {
    const int a = F.axis;
    if (F.plane < 0 || F.plane + 1 >= s->n[a])
        return px_fail(PXFDTD_ERR_INVALID, "a flux plane after node %d along axis %d needs node %d on the grid",
                       F.plane, a, F.plane + 1);
    int32_t low[3], high[3];
    low[a] = F.plane;
    high[a] = F.plane + 1;
    for (int w = 0; w < 2; w++) {
        const int along = (a + 1 + w) % 3;
        if (F.windowed[w] && !(F.wlo[w] >= 0 && F.wlo[w] < F.whi[w] && F.whi[w] < s->n[along]))
            return px_fail(PXFDTD_ERR_INVALID,
                           "a flux plane's window along axis %d must run from one half-plane to a later one on "
                           "the grid, got %d to %d of %d", along, F.wlo[w], F.whi[w], s->n[along]);
        int from, to;
        flux_range(s, F, w, from, to);
        low[along] = from;
        high[along] = to - 1;
    }
    return add_transform(s, low, high, nf, freqs, &F.dft);
}

// Simulation::add_flux (src/fdtd/monitors.rs:470-491): window[w] = {lo, hi} along the w-th of
// the axes after `axis`, or lo < 0 for the whole axis
extern "C" int pxfdtd_simulation_add_flux(pxfdtd_simulation *s, int32_t axis, int32_t plane,
                                          const int32_t window[2][2], int32_t nf, const double *freqs,
                                          int32_t *number) {
    if (axis < 0 || axis > 2) return px_fail(PXFDTD_ERR_INVALID, "no axis %d", axis);
    TRY(check_frequencies(s, nf, freqs));
    FluxFace F;
    F.axis = axis;
    F.plane = plane;
    F.sign = 1.0;
    for (int w = 0; w < 2; w++) {
        F.windowed[w] = window && window[w][0] >= 0;
        F.wlo[w] = F.windowed[w] ? window[w][0] : 0;
        F.whi[w] = F.windowed[w] ? window[w][1] : 0;
    }
    TRY(add_face(s, F, nf, freqs));
    s->fluxes.push_back({{F}});
    *number = (int32_t)s->fluxes.size() - 1;
    return PXFDTD_OK;
}

// Simulation::add_flux_box (monitors.rs:493-528): the six faces of the box between the
// half-planes after nodes planes[a][0] and planes[a][1] along each axis a, outwards
extern "C" int pxfdtd_simulation_add_flux_box(pxfdtd_simulation *s, const int32_t planes[3][2],
                                              int32_t nf, const double *freqs, int32_t *number) {
    TRY(check_frequencies(s, nf, freqs));
    FluxRec R;
    for (int a = 0; a < 3; a++) {
        const int lo = planes[a][0], hi = planes[a][1];
        if (lo >= hi)
            return px_fail(PXFDTD_ERR_INVALID, "a flux box's planes along axis %d must be from low to high, got %d and %d",
                           a, lo, hi);
        for (int side = 0; side < 2; side++) {
            FluxFace F;
            F.axis = a;
            F.plane = side ? hi : lo;
            F.sign = side ? 1.0 : -1.0;
            for (int w = 0; w < 2; w++) {
                const int along = (a + 1 + w) % 3;
                F.windowed[w] = true;
                F.wlo[w] = planes[along][0];
                F.whi[w] = planes[along][1];
            }
            TRY(add_face(s, F, nf, freqs));
            R.faces.push_back(F);
        }
    }
    s->fluxes.push_back(std::move(R));
    *number = (int32_t)s->fluxes.size() - 1;
    return PXFDTD_OK;
}

// ---------------------------------------------------------------- media stability
// |z|, portable::complex::abs
static double cabs_(C64 z) { return pxfdtd::pcabs(z); }

// media.rs:446-455: the product of two polynomials, coefficients from the constant term up
static std::vector<double> poly_multiply(const std::vector<double> &a, const std::vector<double> &b) {
    std::vector<double> out(a.size() + b.size() - 1, 0.0);
    for (size_t i = 0; i < a.size(); i++)
        for (size_t j = 0; j < b.size(); j++) out[i + j] += a[i] * b[j];
    return out;
}

// media.rs:457-510: the roots of a real polynomial by Aberth's simultaneous iteration (O. Aberth,
// Math. Comput. 27, 339 (1973)), from a circle of the roots' geometric mean modulus turned off
// the real axis, at most 500 sweeps, until no root moves by more than 1e-15 of its modulus
static std::vector<C64> poly_roots(std::vector<double> p) {
    while (p.size() > 1 && p.back() == 0.0) p.pop_back();
    const int n = (int)p.size() - 1;
    if (n <= 0) return {};
    const double lead = p[n];
    for (double &c : p) c = c / lead;
    const double radius = std::max(pxfdtd::ppow(std::fabs(p[0]), 1.0 / n), 1e-3);
    std::vector<C64> z(n);
    for (int k = 0; k < n; k++) {
        double sn = 0.0, cs = 1.0;
        pxfdtd::psincos(TAU * (k + 0.25) / n, sn, cs);
        z[k] = {radius * cs, radius * sn};
    }
    auto eval = [&](C64 x, C64 &v, C64 &d) {
        v = {0.0, 0.0};
        d = {0.0, 0.0};
        for (int q = n; q >= 0; q--) {
            const C64 dx = cmul(d, x), vx = cmul(v, x);
            d = {dx.re + v.re, dx.im + v.im};
            v = {vx.re + p[q], vx.im};
        }
    };
    for (int it = 0; it < 500; it++) {
        double largest = 0.0;
        for (int k = 0; k < n; k++) {
            C64 v, d;
            eval(z[k], v, d);
            if (v.re == 0.0 && v.im == 0.0) continue;
            const C64 ratio = cdiv(v, d);
            C64 rep = {0.0, 0.0};
            for (int j = 0; j < n; j++) {
                if (j == k) continue;
                const C64 t = cdiv({1.0, 0.0}, {z[k].re - z[j].re, z[k].im - z[j].im});
                rep = {rep.re + t.re, rep.im + t.im};
            }
            const C64 rr = cmul(ratio, rep);
            const C64 step = cdiv(ratio, {1.0 - rr.re, -rr.im});
            if (std::isfinite(step.re) && std::isfinite(step.im)) {
                z[k] = {z[k].re - step.re, z[k].im - step.im};
                largest = std::max(largest, cabs_(step) / std::max(cabs_(z[k]), 1e-300));
            }
        }
        if (largest < 1e-15) break;
    }
    return z;
}

// Dispersive::amplification (media.rs:258-299): the largest |z| over the modes of a plane wave
// stepped in the medium (eps_inf, terms {alpha, xi, gamma}) with conductivity sigma at time
// step dt, for (K dt)^2 = u: the roots of the update's characteristic polynomial,
// eps_inf (z - 1)^2 + (sigma dt / 2)(z^2 - 1) + u z times every term's z^2 - alpha z - xi, plus
// each term's (gamma / 4)(z^2 - 1)^2 times the others'
static double medium_amplification(double eps_inf, const std::vector<std::array<double, 3>> &terms, double sigma,
                                   double dt, double u) {
    auto product = [&](int skip) {
        std::vector<double> acc = {1.0};
        for (int i = 0; i < (int)terms.size(); i++)
            if (i != skip) acc = poly_multiply(acc, {-terms[i][1], -terms[i][0], 1.0});
        return acc;
    };
    const double e = eps_inf, half = sigma * dt / 2.0;
    std::vector<double> poly = poly_multiply({e - half, u - 2.0 * e, e + half}, product(-1));
    for (int i = 0; i < (int)terms.size(); i++) {
        const double g = terms[i][2];
        const std::vector<double> t = poly_multiply({g / 4.0, 0.0, -g / 2.0, 0.0, g / 4.0}, product(i));
        for (size_t c = 0; c < poly.size() && c < t.size(); c++) poly[c] += t[c];
    }
    double m = 0.0;
    for (const C64 &z : poly_roots(poly)) m = std::max(m, cabs_(z));
    return m;
}

// Dispersive::stable (media.rs:300-309): every (K dt)^2 up to top dt^2 in 200 steps (K = 0,
// a static field's double root at z = 1, left out) amplifies by at most 1 + 1e-6
static bool medium_stable(double eps_inf, const std::vector<std::array<double, 3>> &terms, double sigma, double dt,
                          double top) {
    const int points = 200;
    const double most = top * dt * dt;
    for (int i = 1; i <= points; i++) {
        const double u = most * i / points;
        if (!(medium_amplification(eps_inf, terms, sigma, dt, u) <= 1.0 + 1e-6)) return false;
    }
    return true;
}

// a medium's terms {alpha, xi, gamma} at dt from poles [np][4] (Pole::coefficients, media.rs:
// 61-109): kind 0 Lorentz (strength, resonance, damping), 1 Drude (plasma, damping)
static int medium_terms(int32_t np, const double *poles, double dt, std::vector<std::array<double, 3>> &out) {
    out.clear();
    for (int q = 0; q < np; q++) {
        const double *p = poles + 4 * q;
        double a, w0sq, gamma;
        if (p[0] == 0.0) {
            if (!(p[1] >= 0 && std::isfinite(p[1])) || !(p[2] > 0 && std::isfinite(p[2])) ||
                !(p[3] >= 0 && std::isfinite(p[3])))
                return px_fail(PXFDTD_ERR_INVALID, "a Lorentz term needs strength >= 0, resonance > 0, damping >= 0");
            const double w0 = TAU * p[2];
            a = p[1] * w0 * w0;
            w0sq = w0 * w0;
            gamma = TAU * p[3];
        } else {
            if (!(p[1] > 0 && std::isfinite(p[1])) || !(p[2] >= 0 && std::isfinite(p[2])))
                return px_fail(PXFDTD_ERR_INVALID, "a Drude term needs plasma > 0, damping >= 0");
            const double wp = TAU * p[1];
            a = wp * wp;
            w0sq = 0.0;
            gamma = TAU * p[2];
        }
        const double d = 1.0 + gamma * dt / 2.0;
        out.push_back({(2.0 - w0sq * dt * dt) / d, (gamma * dt / 2.0 - 1.0) / d, a * dt * dt / d});
    }
    return PXFDTD_OK;
}

extern "C" int pxfdtd_medium_amplification(double eps_inf, int32_t np, const double *poles, double sigma, double dt,
                                           double u, double *out) {
    if (!(eps_inf > 0 && std::isfinite(eps_inf)) || np < 0 || !(dt > 0) || !std::isfinite(u))
        return px_fail(PXFDTD_ERR_INVALID, "eps_inf > 0, np >= 0, dt > 0 and a finite u are needed");
    std::vector<std::array<double, 3>> terms;
    TRY(medium_terms(np, poles, dt, terms));
    *out = medium_amplification(eps_inf, terms, sigma, dt, u);
    return PXFDTD_OK;
}

// src/fdtd/media.rs:61-109,733-825
extern "C" int pxfdtd_simulation_add_medium(pxfdtd_simulation *s, double eps_inf, int32_t np,
                                            const double *poles, const pxfdtd_scene *inside_) {
    if (s->steps > 0) return px_fail(PXFDTD_ERR_INVALID, "a dispersive medium goes in before the run starts");
    if (!(eps_inf > 0 && std::isfinite(eps_inf)))
        return px_fail(PXFDTD_ERR_INVALID, "a medium's eps_inf must be positive and finite");
    if (np < 0 || np > 4) return px_fail(PXFDTD_ERR_UNSUPPORTED, "a medium of 0 to 4 terms steps here");
    TRY(sync_host(s));
    const double dt = s->dt;
    Group G;
    G.poles = np;
    std::vector<std::array<double, 3>> terms;
    TRY(medium_terms(np, poles, dt, terms));
    double gsum = 0.0;
    for (int q = 0; q < np; q++) {
        for (int m = 0; m < 3; m++) G.coef[q][m] = terms[q][m];
        gsum += G.coef[q][2];
    }
    // the curl's largest eigenvalue, 4 sum 1/h^2 over the axes a field varies along
    // (bloch::curl_bound, src/fdtd/bloch.rs:31-43, with k = 0 on periodic sides)
    double curl = 0.0;
    const double hs[3] = {s->g.dx, s->g.dy, s->g.dz};
    for (int a = 0; a < 3; a++)
        if (!(s->b.periodic[a] && s->n[a] == 1)) curl += pxfdtd::ppowi(hs[a], -2);
    const double top = 4.0 * curl;
    G.chi = gsum / (4.0 * dt);
    const long long N = s->N;
    std::vector<double> sig;
    // the values inside, in component-then-index order: each thread scans a contiguous range
    // and the ranges are joined in order
    for (int c = 0; c < 3; c++) {
        std::vector<std::vector<long long>> found;
        std::vector<std::vector<double>> fsig;
#pragma omp parallel
        {
#pragma omp single
            {
                found.resize((size_t)omp_get_num_threads());
                fsig.resize(found.size());
            }
            const int t = omp_get_thread_num(), nt = omp_get_num_threads();
            const long long a = N * t / nt, z = N * (t + 1) / nt;
            for (long long r = a; r < z; r++) {
                if (s->cb[c * N + r] == 0.0) continue;
                double p[3];
                s->e_position(c, r, p);
                if (scene_value(inside_, 0, p[0], p[1], p[2]) == 0.0) continue;
                double old = 0.0;
                if (!s->member.empty() && s->member[c * N + r] >= 0) old = s->groups[s->member[c * N + r]].chi;
                const double sigma = (1.0 - s->ca[c * N + r]) / s->cb[c * N + r] - old;
                found[t].push_back(r);
                fsig[t].push_back(sigma > 0.0 ? sigma : 0.0);
            }
        }
        for (size_t t = 0; t < found.size(); t++) {
            G.r.insert(G.r.end(), found[t].begin(), found[t].end());
            G.comp.insert(G.comp.end(), found[t].size(), c);
            sig.insert(sig.end(), fsig[t].begin(), fsig[t].end());
        }
    }
    // stability (media.rs:783-793): without conductivity and with the largest inside
    double largest = 0.0;
    for (double v : sig) largest = std::max(largest, v);
    for (double sigma : {0.0, largest})
        if (!medium_stable(eps_inf, terms, sigma, dt, top))
            return px_fail(PXFDTD_ERR_INVALID,
                           "the medium isn't stable at the time step %g um/c (a term's w0 dt above 2, or a "
                           "Courant number above sqrt(eps_inf)): take a smaller step", dt);
    const double e = eps_inf;
    // the values taken from earlier media leave those, each group compacted once
    std::vector<char> dirty(s->groups.size(), 0);
    if (!s->member.empty())
        for (size_t v = 0; v < G.r.size(); v++) {
            int &m = s->member[G.comp[v] * N + G.r[v]];
            if (m >= 0) {
                dirty[(size_t)m] = 1;
                m = -1;
            }
        }
    for (size_t g = 0; g < dirty.size(); g++) {
        if (!dirty[g]) continue;
        Group &O = s->groups[g];
        size_t w = 0;
        for (size_t v = 0; v < O.r.size(); v++) {
            if (s->member[O.comp[v] * N + O.r[v]] != (int)g) continue;
            O.comp[w] = O.comp[v];
            O.r[w] = O.r[v];
            w++;
        }
        O.comp.resize(w);
        O.r.resize(w);
    }
    for (size_t v = 0; v < G.r.size(); v++) {
        const int c = G.comp[v];
        const long long r = G.r[v];
        const double nu = 1.0 + dt / e * (sig[v] / 2.0 + G.chi);
        s->ca[c * N + r] = (1.0 - dt * sig[v] / (2.0 * e)) / nu;
        s->cb[c * N + r] = dt / e / nu;
    }
    if (np > 0 && !G.r.empty()) {
        if (s->member.empty()) s->member.assign(3 * N, -1);
        const int gi = (int)s->groups.size();
        for (size_t v = 0; v < G.r.size(); v++) s->member[G.comp[v] * N + G.r[v]] = gi;
        s->groups.push_back(std::move(G));
    }
    drop_resident(s);
    return PXFDTD_OK;
}

// ------------------------------------------------------------------------------ device

static void drop_resident(pxfdtd_simulation *s) {
    (void)bank_sums(s);
    if (s->res) {
        pxfdtd_resident_destroy(s->res);
        s->res = nullptr;
    }
    s->planes_on_device = 0;
    s->planes_device_newer = false;
    s->inj_dirty = s->taps_dirty = true;
    s->host_newer = true;
}

// the plane waves' auxiliary fields back on the host
static int pull_planes(pxfdtd_simulation *s) {
    if (!s->res || !s->planes_device_newer) return PXFDTD_OK;
    for (int q = 0; q < s->planes_on_device; q++)
        TRY(pxfdtd_resident_plane_state(s->res, q, s->planes[q].e.data(), s->planes[q].h.data()));
    s->planes_device_newer = false;
    return PXFDTD_OK;
}

// the device's fields, psi and transform sums back on the host
static int sync_host(pxfdtd_simulation *s) {
    TRY(pull_planes(s));
    if (!s->res || !s->device_newer) return PXFDTD_OK;
    std::vector<double> psi(pxfdtd_resident_psi_len(s->res));
    TRY(pxfdtd_resident_download(s->res, s->e.data(), s->h.data(), psi.data()));
    s->fields_zero = false;
    size_t at = 0;
    for (auto &sl : s->slabs)
        for (auto &v : sl.psi) v = psi[at++];
    TRY(bank_sums(s));
    s->device_newer = false;
    // the dispersive media's currents live on the device only: keep it
    return PXFDTD_OK;
}

// the device problem of the simulation (what Resident::new reads); `media` holds the media
// records p points into
static void make_problem(const pxfdtd_simulation *s, int precision, pxfdtd_problem &p,
                         std::vector<pxfdtd_medium> &media) {
    std::memset(&p, 0, sizeof p);
    p.nx = s->n[0];
    p.ny = s->n[1];
    p.nz = s->n[2];
    p.precision = precision;
    p.dt = s->dt;
    p.step[0] = s->g.dx;
    p.step[1] = s->g.dy;
    p.step[2] = s->g.dz;
    for (int a = 0; a < 3; a++) {
        p.periodic[a] = s->b.periodic[a];
        p.pml_low[a] = s->b.periodic[a] ? 0 : s->b.low[a];
        p.pml_high[a] = s->b.periodic[a] ? 0 : s->b.high[a];
        p.factors_e[a] = s->fe[a].data();
        p.factors_h[a] = s->fh[a].data();
        for (int sd = 0; sd < 2; sd++)
            for (int f = 0; f < 2; f++) {
                p.cpml_b[a][sd][f] = s->cpml_b[a][sd][f].empty() ? nullptr : s->cpml_b[a][sd][f].data();
                p.cpml_a[a][sd][f] = s->cpml_a[a][sd][f].empty() ? nullptr : s->cpml_a[a][sd][f].data();
            }
    }
    p.ca = s->ca.data();
    p.cb = s->cb.data();
    media.clear();
    for (size_t g = 0; g < s->groups.size(); g++) {
        const Group &G = s->groups[g];
        pxfdtd_medium m;
        std::memset(&m, 0, sizeof m);
        m.poles = G.poles;
        m.chi = G.chi;
        std::memcpy(m.coef, G.coef, sizeof m.coef);
        m.count = (int64_t)G.r.size();
        m.component = G.comp.data();
        m.index = G.r.data();
        // media with no terms left are plain values
        if (m.count && m.poles) media.push_back(m);
    }
    p.nmedia = (int32_t)media.size();
    p.media = media.data();
}

static int ensure_resident(pxfdtd_simulation *s) {
    if (!s->dev) return px_fail(PXFDTD_ERR_UNSUPPORTED, "no device: set one with pxfdtd_simulation_set_device");
    if (!s->res) {
        pxfdtd_problem p;
        std::vector<pxfdtd_medium> live;
        make_problem(s, s->precision, p, live);
        TRY(pxfdtd_resident_create(s->dev, &p, &s->res));
        s->host_newer = true;
        s->inj_dirty = s->taps_dirty = true;
    }
    if (s->host_newer && s->fields_zero) {  // fields as made: zeros, without their bytes
        TRY(pxfdtd_resident_upload(s->res, nullptr, nullptr, nullptr));
        s->host_newer = false;
    }
    if (s->host_newer) {
        std::vector<double> psi;
        for (auto &sl : s->slabs) psi.insert(psi.end(), sl.psi.begin(), sl.psi.end());
        TRY(pxfdtd_resident_upload(s->res, s->e.data(), s->h.data(), psi.empty() ? nullptr : psi.data()));
        s->host_newer = false;
    }
    if (s->inj_dirty) {
        // H's then E's order within a field: point sources, then currents, each in order
        std::vector<pxfdtd_injection> inj;
        int em = 0;
        for (auto &so : s->sources) inj.push_back({so.field, so.comp, so.r, 1.0, 0.0, em++});
        for (auto &cu : s->currents) {
            for (size_t v = 0; v < cu.r.size(); v++)
                inj.push_back({cu.field, cu.comps[v], cu.r[v], cu.amps[v].re, cu.amps[v].im, em});
            em++;
        }
        const int first_plane = em;
        em += (int)s->planes.size();
        s->emitters = em;
        TRY(pxfdtd_resident_set_injections(s->res, inj.data(), (int64_t)inj.size(), std::max(em, 1)));
        if (!s->planes.empty() || s->planes_on_device) {
            // the waves the device steps keep their state: fetched before the tables change
            TRY(pull_planes(s));
            std::vector<pxfdtd_plane_wave> pw;
            std::vector<pxfdtd_coupling> cp;
            for (size_t q = 0; q < s->planes.size(); q++) {
                const PlaneRec &P = s->planes[q];
                pxfdtd_plane_wave w;
                std::memset(&w, 0, sizeof w);
                for (int a = 0; a < 3; a++) w.g[a] = (int32_t)P.g[a];
                w.len = P.len;
                w.pad = P.pad;
                w.ca = P.ca.data();
                w.cb = P.cb.data();
                w.da = P.da.data();
                w.db = P.db.data();
                w.nsheet = P.nsheet;
                for (int t = 0; t < P.nsheet; t++) {
                    w.sheet_component[t] = P.sheet_c[t];
                    w.sheet_sample[t] = P.sheet_i[t];
                    w.sheet_j[t] = P.sheet_j[t];
                }
                w.emitter = first_plane + (int)q;
                pw.push_back(w);
                for (pxfdtd_coupling c : P.couplings) {
                    c.plane = (int32_t)q;
                    cp.push_back(c);
                }
            }
            TRY(pxfdtd_resident_set_plane_waves(s->res, pw.data(), (int32_t)pw.size(), cp.data(),
                                                (int64_t)cp.size()));
            for (size_t q = 0; q < s->planes.size(); q++)
                TRY(pxfdtd_resident_set_plane_state(s->res, (int32_t)q, s->planes[q].e.data(),
                                                    s->planes[q].h.data()));
            s->planes_on_device = (int)s->planes.size();
        }
        s->inj_dirty = false;
    }
    if (s->taps_dirty) {
        std::vector<pxfdtd_probe> pr;
        for (auto &p : s->probes) pr.push_back({p.field, p.comp, p.r});
        TRY(pxfdtd_resident_set_probes(s->res, pr.data(), (int32_t)pr.size()));
        std::vector<pxfdtd_series> se;
        std::vector<int32_t> nf;
        for (size_t m = 0; m < s->dfts.size(); m++) {
            const DftRec &d = s->dfts[m];
            nf.push_back((int32_t)d.freqs.size());
            for (int q = 0; q < 6; q++) {
                pxfdtd_series x;
                x.field = q < 3 ? PXFDTD_E : PXFDTD_H;
                x.component = q % 3;
                for (int a = 0; a < 3; a++) {
                    x.lo[a] = d.lo[a];
                    x.hi[a] = d.hi[a];
                }
                x.monitor = (int32_t)m;
                se.push_back(x);
            }
        }
        TRY(pxfdtd_resident_set_series(s->res, se.data(), (int32_t)se.size(), nf.data(), (int32_t)nf.size()));
        // the sums go on from the host's (one running sum each, in step order)
        std::vector<double> acc;
        for (auto &d : s->dfts) acc.insert(acc.end(), d.acc.begin(), d.acc.end());
        bool any = false;
        for (double v : acc) any = any || v != 0.0 || std::signbit(v);
        if (any) TRY(pxfdtd_resident_write_sums(s->res, acc.data(), (int64_t)acc.size()));
        s->taps_dirty = false;
    }
    return PXFDTD_OK;
}

extern "C" int pxfdtd_simulation_set_device(pxfdtd_simulation *s, pxfdtd_device *dev, int32_t precision) {
    if (precision != PXFDTD_F32 && precision != PXFDTD_F64)
        return px_fail(PXFDTD_ERR_INVALID, "precision must be PXFDTD_F32 or PXFDTD_F64");
    TRY(sync_host(s));
    drop_resident(s);
    s->dev = dev;
    s->precision = precision;
    if (dev) {
        const int rc = ensure_resident(s);
        if (rc != PXFDTD_OK) {  // the device stays as it was: none
            s->dev = nullptr;
            drop_resident(s);
            return rc;
        }
    }
    return PXFDTD_OK;
}

static int read_probes(pxfdtd_simulation *s) {
    if (!s->res || s->probes.empty()) return PXFDTD_OK;
    const int64_t rows = pxfdtd_resident_probe_steps(s->res);
    if (!rows) return PXFDTD_OK;
    const size_t np = s->probes.size();
    std::vector<double> buf((size_t)rows * np);
    int64_t steps = 0;
    TRY(pxfdtd_resident_read_probes(s->res, buf.data(), (int64_t)buf.size(), &steps));
    for (int64_t n = 0; n < steps; n++)
        for (size_t q = 0; q < np; q++) s->probes[q].values.push_back(buf[n * np + q]);
    return PXFDTD_OK;
}

// Simulation::run on the device (src/fdtd.rs:1142-1158, resident.rs:430-539): waveforms and
// phases in FP64 on the host, in chunks of steps
// each emitter's w and each transform phase for `m` steps from step n0, in FP64 as the
// reference's host computes them (resident.rs:797-843): waves [m][max(1, emitters)][2] at the
// field's time (H at n dt, E at (n + 1/2) dt), phases [m][max(1, slots)][2], e^(i w t) dt with E
// at (n + 1) dt and H at (n + 1/2) dt
static void fill_tables(const pxfdtd_simulation *s, int64_t n0, int64_t m, int slots, double *waves,
                        double *phases) {
    const int ew = std::max(1, s->emitters), sw = std::max(1, slots);
    for (int64_t q = 0; q < m; q++) {
        const long long n = n0 + q;
        const double t = (double)n * s->dt;
        const double half = t + 0.5 * s->dt;
        int em = 0;
        double *w = waves + (size_t)q * ew * 2;
        w[0] = w[1] = 0.0;
        for (auto &so : s->sources) {
            w[2 * em] = wave_at(so.w, so.field == PXFDTD_H ? t : half);
            w[2 * em + 1] = 0.0;
            em++;
        }
        for (auto &cu : s->currents) {
            const C64 v = wave_complex_at(cu.w, cu.field == PXFDTD_H ? t : half);
            w[2 * em] = v.re;
            w[2 * em + 1] = v.im;
            em++;
        }
        for (auto &pl : s->planes) {  // the sheet's w at t + dt/2 (sources.rs:1035)
            w[2 * em] = wave_at(pl.w, half);
            w[2 * em + 1] = 0.0;
            em++;
        }
        const double te = (double)(n + 1) * s->dt;
        const double th = te - 0.5 * s->dt;
        double *ph = phases + (size_t)q * sw * 2;
        ph[0] = ph[1] = 0.0;
        int slot = 0;
        for (auto &d : s->dfts) {
            for (int f2 = 0; f2 < 2; f2++)
                for (double fr : d.freqs) {
                    const C64 p = pxfdtd::pcexp({0.0, TAU * fr * (f2 == 0 ? te : th)});
                    ph[2 * slot] = p.re * s->dt;
                    ph[2 * slot + 1] = p.im * s->dt;
                    slot++;
                }
        }
    }
}

// Simulation::run on the device (src/fdtd.rs:1142-1158, resident.rs:430-539): waveforms and
// phases in FP64 on the host, in chunks of steps
extern "C" int pxfdtd_simulation_run(pxfdtd_simulation *s, int64_t count) {
    if (count < 0) return px_fail(PXFDTD_ERR_INVALID, "a negative step count");
    TRY(ensure_resident(s));
    const int slots = pxfdtd_resident_phase_slots(s->res);
    const int64_t chunk = 512;
    std::vector<double> waves, phases;
    for (int64_t done = 0; done < count;) {
        const int64_t m = std::min(chunk, count - done);
        waves.assign((size_t)m * std::max(1, s->emitters) * 2, 0.0);
        phases.assign((size_t)m * std::max(1, slots) * 2, 0.0);
        fill_tables(s, s->steps + done, m, slots, waves.data(), phases.data());
        TRY(pxfdtd_resident_step(s->res, (uint32_t)m, waves.data(), phases.data()));
        done += m;
        TRY(read_probes(s));
    }
    s->steps += count;
    if (count) s->device_newer = true;
    if (count && s->planes_on_device) s->planes_device_newer = true;
    return PXFDTD_OK;
}

// Simulation::run_until (src/fdtd.rs:1160-1167): steps until E's time reaches `time`
extern "C" int pxfdtd_simulation_run_until(pxfdtd_simulation *s, double time) {
    if (std::isnan(time)) return px_fail(PXFDTD_ERR_INVALID, "a time to run until must be a number");
    int64_t count = 0;
    while ((double)(s->steps + count) * s->dt < time) count++;
    return pxfdtd_simulation_run(s, count);
}

namespace {

// Growth (src/fdtd/monitors.rs:665-717): rising at each of the last `needed` checks, to a
// thousand times the level before the first of them
struct Growth {
    int needed;
    double start = 0.0, last = 0.0;
    int rises = 0;
    explicit Growth(int n) : needed(n) {}
    // for checks every `check` after a pulse that ends at `pulse_end`
    static Growth after(double pulse_end, double check) {
        return Growth(std::max((int)std::ceil(pulse_end / check) + 1, 4));
    }
    // the factor the fields grew by, or 0 while they aren't growing
    double take(double level) {
        if (last > 0.0 && level > last) {
            if (rises == 0) start = last;
            rises++;
        } else {
            rises = 0;
        }
        last = level;
        return (rises >= needed && level >= 1e3 * start) ? level / start : 0.0;
    }
};

}  // namespace

// when the last source ends (monitors.rs:719-738): a pulse 6 widths after its peak; false for
// a continuous wave, which never ends
static bool sources_end(const pxfdtd_simulation *s, double &end) {
    end = 0.0;
    auto take = [&](const pxfdtd_waveform &w) {
        if (w.kind == 2) return false;
        end = std::max(end, w.delay + 6.0 * w.width);
        return true;
    };
    for (auto &so : s->sources)
        if (!take(so.w)) return false;
    for (auto &cu : s->currents)
        if (!take(cu.w)) return false;
    for (auto &pl : s->planes)
        if (!take(pl.w)) return false;
    return true;
}

// Simulation::run_until_decayed (monitors.rs:740-841)
extern "C" int pxfdtd_simulation_run_until_decayed(pxfdtd_simulation *s, const int32_t *probes, int32_t n,
                                                   double fraction, double interval, double limit,
                                                   int32_t *decayed, double *growth_factor) {
    if (n <= 0 || !probes || !(fraction > 0.0 && fraction < 1.0) || !(std::isfinite(interval) && interval > 0.0) ||
        !std::isfinite(limit))
        return px_fail(PXFDTD_ERR_INVALID,
                       "running until the fields decay needs probes, a fraction in (0, 1), a positive interval and a "
                       "finite limit, got %d probes, %g, %g and %g", n, fraction, interval, limit);
    for (int q = 0; q < n; q++)
        if (probes[q] < 0 || probes[q] >= (int32_t)s->probes.size())
            return px_fail(PXFDTD_ERR_INVALID, "no probe %d", probes[q]);
    *decayed = 0;
    if (growth_factor) *growth_factor = 0.0;
    double end;
    const bool ends = sources_end(s, end);
    Growth growth = Growth::after(ends ? end : 0.0, interval);
    while ((double)s->steps * s->dt < limit) {
        const size_t from = s->probes[probes[0]].values.size();
        const double began = (double)s->steps * s->dt;
        TRY(pxfdtd_simulation_run_until(s, std::min(began + interval, limit)));
        bool all = true;
        double level = 0.0;
        for (int q = 0; q < n; q++) {
            const std::vector<double> &v = s->probes[probes[q]].values;
            double peak = 0.0, recent = 0.0;
            for (size_t t = 0; t < v.size(); t++) {
                peak = std::max(peak, v[t] * v[t]);
                if (t >= std::min(from, v.size())) recent = std::max(recent, v[t] * v[t]);
            }
            all = all && peak > 0.0 && recent <= fraction * peak;
            level = std::max(level, recent);
        }
        if (all) {
            *decayed = 1;
            return PXFDTD_OK;
        }
        if (ends && began >= end) {
            const double factor = growth.take(level);
            if (factor > 0.0) {
                if (growth_factor) *growth_factor = factor;
                return px_fail(PXFDTD_ERR_GROWING,
                               "the fields grew %g times over %d checks to t = %g um/c: a wave of the grid amplified "
                               "in a CPML, which a finer grid removes", factor, growth.needed, (double)s->steps * s->dt);
            }
        }
    }
    return PXFDTD_OK;
}

extern "C" double pxfdtd_simulation_dt(const pxfdtd_simulation *s) { return s->dt; }
extern "C" int64_t pxfdtd_simulation_steps(const pxfdtd_simulation *s) { return s->steps; }
extern "C" int64_t pxfdtd_simulation_cells(const pxfdtd_simulation *s) { return s->N; }
extern "C" pxfdtd_resident *pxfdtd_simulation_resident(pxfdtd_simulation *s) { return s->res; }

extern "C" int pxfdtd_simulation_field(pxfdtd_simulation *s, int32_t which, int32_t comp, double *out) {
    if (which < 0 || which > 1 || comp < 0 || comp > 2) return px_fail(PXFDTD_ERR_INVALID, "no such field");
    TRY(sync_host(s));
    const double *src = (which == 0 ? s->e.data() : s->h.data()) + comp * s->N;
    std::memcpy(out, src, s->N * sizeof(double));
    return PXFDTD_OK;
}

extern "C" int pxfdtd_simulation_set_field(pxfdtd_simulation *s, int32_t which, int32_t comp,
                                           const double *values) {
    if (which < 0 || which > 1 || comp < 0 || comp > 2) return px_fail(PXFDTD_ERR_INVALID, "no such field");
    TRY(sync_host(s));
    double *dst = (which == 0 ? s->e.data() : s->h.data()) + comp * s->N;
    std::memcpy(dst, values, s->N * sizeof(double));
    s->host_newer = true;
    s->fields_zero = false;
    return PXFDTD_OK;
}

extern "C" int pxfdtd_simulation_coefficients(const pxfdtd_simulation *s, int32_t comp, double *ca, double *cb) {
    if (comp < 0 || comp > 2) return px_fail(PXFDTD_ERR_INVALID, "no such component");
    std::memcpy(ca, s->ca.data() + comp * s->N, s->N * sizeof(double));
    std::memcpy(cb, s->cb.data() + comp * s->N, s->N * sizeof(double));
    return PXFDTD_OK;
}

extern "C" int64_t pxfdtd_simulation_probe_len(pxfdtd_simulation *s, int32_t n) {
    if (n < 0 || n >= (int32_t)s->probes.size()) return -1;
    return (int64_t)s->probes[n].values.size();
}

extern "C" int pxfdtd_simulation_probe(pxfdtd_simulation *s, int32_t n, double *out) {
    if (n < 0 || n >= (int32_t)s->probes.size()) return px_fail(PXFDTD_ERR_INVALID, "no probe %d", n);
    std::memcpy(out, s->probes[n].values.data(), s->probes[n].values.size() * sizeof(double));
    return PXFDTD_OK;
}

extern "C" int64_t pxfdtd_simulation_dft_len(pxfdtd_simulation *s, int32_t n, int32_t series) {
    if (n < 0 || n >= (int32_t)s->shown.size() || series < 0 || series > 5) return -1;
    return (int64_t)s->dfts[s->shown[n]].acc.size() / 6;
}

// the sums of transform monitor `index` (an index into dfts), series `series`, into out
static int transform_sums(pxfdtd_simulation *s, int index, int series, double *out) {
    const int n = index;
    const DftRec &d = s->dfts[n];
    const size_t len = d.acc.size() / 6;
    if (s->res && !s->taps_dirty) {
        // the device's sums, which went on from the host's
        std::vector<double> sums(pxfdtd_resident_sums_len(s->res));
        TRY(pxfdtd_resident_read_sums(s->res, sums.data(), (int64_t)sums.size()));
        size_t at = 0;
        for (int32_t m = 0; m < n; m++) at += s->dfts[m].acc.size();
        std::memcpy(out, sums.data() + at + series * len, len * sizeof(double));
    } else {
        std::memcpy(out, d.acc.data() + series * len, len * sizeof(double));
    }
    return PXFDTD_OK;
}

extern "C" int pxfdtd_simulation_dft(pxfdtd_simulation *s, int32_t n, int32_t series, double *out) {
    if (n < 0 || n >= (int32_t)s->shown.size() || series < 0 || series > 5)
        return px_fail(PXFDTD_ERR_INVALID, "no transform %d series %d", n, series);
    return transform_sums(s, s->shown[n], series, out);
}

// FluxPlane::flux (monitors.rs:318-350): S.n = 1/2 Re(E_b H~_c* - E_c H~_b*) over the plane, E
// the mean of the nodes either side; summed in the reference's order
static double face_flux(pxfdtd_simulation *s, const FluxFace &F, const std::vector<double> series[6], int f) {
    const DftRec &d = s->dfts[F.dft];
    const int a = F.axis, b = (a + 1) % 3, c = (a + 2) % 3;
    const int ex = d.hi[0] - d.lo[0], ey = d.hi[1] - d.lo[1], ez = d.hi[2] - d.lo[2];
    auto value = [&](int field, int comp, const int at[3]) {
        const std::vector<double> &v = series[field * 3 + comp];
        const long long q = (((long long)f * ez + (at[2] - d.lo[2])) * ey + (at[1] - d.lo[1])) * ex + (at[0] - d.lo[0]);
        return C64{v[2 * q], v[2 * q + 1]};
    };
    int bfrom, bto, cfrom, cto;
    flux_range(s, F, 0, bfrom, bto);
    flux_range(s, F, 1, cfrom, cto);
    double total = 0.0;
    const int comps[2] = {b, c}, hs[2] = {c, b};
    const double signs[2] = {0.5, -0.5};
    for (int q = 0; q < 2; q++) {
        const int comp = comps[q], h = hs[q];
        for (int v = cfrom; v < cto; v++)
            for (int u = bfrom; u < bto; u++) {
                const double weight = flux_weight(F, 0, u, comp == b) * flux_weight(F, 1, v, comp == c);
                if (weight == 0.0) continue;
                int at[3];
                at[a] = F.plane;
                at[b] = u;
                at[c] = v;
                const C64 low = value(PXFDTD_E, comp, at);
                const C64 hv = value(PXFDTD_H, h, at);
                at[a] = F.plane + 1;
                const C64 high = value(PXFDTD_E, comp, at);
                const double mre = 0.5 * (low.re + high.re), mim = 0.5 * (low.im + high.im);
                // Re(m conj(hv)) as num-complex multiplies: m.re hv.re - m.im (-hv.im)
                total += weight * signs[q] * (mre * hv.re - mim * -hv.im);
            }
    }
    const double steps[3] = {s->g.dx, s->g.dy, s->g.dz};
    return total * steps[b] * steps[c];
}

// the flux from the device's sums, on the device (pxfdtd_flux_device): only nf values come back
static int device_flux(pxfdtd_simulation *s, const FluxRec &R, int nf, double *out) {
    const double *sums = nullptr;
    TRY(pxfdtd_resident_sums_device(s->res, &sums));
    std::vector<pxfdtd_flux_face> faces(R.faces.size());
    const double steps[3] = {s->g.dx, s->g.dy, s->g.dz};
    for (size_t q = 0; q < R.faces.size(); q++) {
        const FluxFace &F = R.faces[q];
        pxfdtd_flux_face &D = faces[q];
        D = pxfdtd_flux_face{};
        D.axis = F.axis;
        D.plane = F.plane;
        for (int w = 0; w < 2; w++) {
            D.windowed[w] = F.windowed[w];
            D.wlo[w] = F.wlo[w];
            D.whi[w] = F.whi[w];
            flux_range(s, F, w, D.from[w], D.to[w]);
        }
        const DftRec &d = s->dfts[F.dft];
        for (int a = 0; a < 3; a++) {
            D.lo[a] = d.lo[a];
            D.hi[a] = d.hi[a];
        }
        D.nf = nf;
        D.sign = F.sign;
        D.step_b = steps[(F.axis + 1) % 3];
        D.step_c = steps[(F.axis + 2) % 3];
        size_t at = 0;
        for (int m = 0; m < F.dft; m++) at += s->dfts[m].acc.size();
        const size_t len = d.acc.size() / 6;
        for (int t = 0; t < 6; t++) D.series[t] = sums + at + t * len;
    }
    return flux_to_host(s->dev, faces.data(), (int32_t)faces.size(), out, pxfdtd_resident_stream(s->res));
}

// Simulation::flux (monitors.rs:530-547): per frequency, the faces' signed fluxes summed
extern "C" int pxfdtd_simulation_flux(pxfdtd_simulation *s, int32_t n, double *out) {
    if (n < 0 || n >= (int32_t)s->fluxes.size()) return px_fail(PXFDTD_ERR_INVALID, "no flux monitor %d", n);
    const FluxRec &R = s->fluxes[n];
    const int nf = (int)s->dfts[R.faces[0].dft].freqs.size();
    if (s->res && !s->taps_dirty) return device_flux(s, R, nf, out);
    std::vector<std::array<std::vector<double>, 6>> sums(R.faces.size());
    for (size_t q = 0; q < R.faces.size(); q++) {
        const DftRec &d = s->dfts[R.faces[q].dft];
        for (int t = 0; t < 6; t++) {
            sums[q][t].resize(d.acc.size() / 6);
            TRY(transform_sums(s, R.faces[q].dft, t, sums[q][t].data()));
        }
    }
    for (int f = 0; f < nf; f++) {
        // a sum of f64s from -0.0, as Rust's Sum for f64 folds (Rust 1.83 on)
        double total = -0.0;
        for (size_t q = 0; q < R.faces.size(); q++) total += R.faces[q].sign * face_flux(s, R.faces[q], sums[q].data(), f);
        out[f] = total;
    }
    return PXFDTD_OK;
}

extern "C" int32_t pxfdtd_simulation_flux_frequencies(const pxfdtd_simulation *s, int32_t n) {
    if (n < 0 || n >= (int32_t)s->fluxes.size()) return -1;
    return (int32_t)s->dfts[s->fluxes[n].faces[0].dft].freqs.size();
}

// ---------------------------------------------------------------- mode monitors

// FDFD's lattice for the grid's boundaries at k0 (Lattice::new, src/fdfd/three.rs:276-308, with
// the grid's boundaries as fdfd_boundaries gives them, src/fdtd/monitors.rs:642-651): a CPML of
// low and high cells is a PML of as many (0 cells a wall) graded with the CPML's reflection and
// order and no real stretch, a periodic axis Bloch with k = 0. Its stretches at the nodes and
// halfway after them (graded, src/fdfd/mod.rs:219-242, with the reference's portable ln and
// powf); shift, fixed, neighbour, curl_e, curl_h and
// curl_curl_into (three.rs:310-418) and volume (src/fdfd/three/ports.rs:228-240). Values are
// numbered as Grid3d::index numbers them: component, then k, j, i.
struct ModeLattice {
    int n[3];
    long long cells;
    double h[3];
    bool periodic[3];
    double k0;
    std::vector<C64> nodes[3], halves[3];

    ModeLattice(const pxfdtd_simulation *s, double k0_) : k0(k0_) {
        const double origin[3] = {s->g.x0, s->g.y0, s->g.z0};
        const double steps[3] = {s->g.dx, s->g.dy, s->g.dz};
        cells = s->N;
        for (int a = 0; a < 3; a++) {
            n[a] = s->n[a];
            h[a] = steps[a];
            periodic[a] = s->b.periodic[a] != 0;
            const int low = periodic[a] ? 0 : s->b.low[a], high = periodic[a] ? 0 : s->b.high[a];
            const double span0 = origin[a] + 0.0 * h[a], span1 = origin[a] + (double)n[a] * h[a];
            for (int off = 0; off < 2; off++) {
                std::vector<C64> &out = off == 0 ? nodes[a] : halves[a];
                out.resize(n[a]);
                for (int m = 0; m < n[a]; m++) {
                    const double pos = origin[a] + (double)m * h[a] + (off == 0 ? 0.0 : 0.5) * h[a];
                    const C64 st = graded(pos, span0, span1, low, high, h[a], s->b.reflection, s->b.order);
                    out[m] = {st.re + 0.0 * st.im, st.im};
                }
            }
        }
    }

    C64 graded(double pos, double span0, double span1, int low, int high, double step, double R,
               double order) const {
        auto profile = [&](double depth, double d) -> C64 {
            const double sigma = (order + 1.0) * (-pxfdtd::pln(R)) / (2.0 * k0 * d);
            double q = depth / d;
            q = q < 0.0 ? 0.0 : (q > 1.0 ? 1.0 : q);
            return {1.0, sigma * pxfdtd::ppow(q, order)};
        };
        const double dl = (double)low * step, dh = (double)high * step;
        if (dl > 0.0 && pos < span0 + dl) return profile(span0 + dl - pos, dl);
        if (dh > 0.0 && pos > span1 - dh) return profile(pos - (span1 - dh), dh);
        return {1.0, 0.0};
    }

    long long index(int c, const int at[3]) const {
        return (long long)c * cells + ((long long)at[2] * n[1] + at[1]) * n[0] + at[0];
    }
    void at(long long r, int &c, int a[3]) const {
        c = (int)(r / cells);
        const long long rest = r % cells;
        a[0] = (int)(rest % n[0]);
        a[1] = (int)((rest / n[0]) % n[1]);
        a[2] = (int)(rest / ((long long)n[0] * n[1]));
    }
    // m + step along axis a and the phase there; false beyond a wall
    bool shift(int a, int m, int step, int &to, C64 &phase) const {
        to = m + step;
        if (to >= 0 && to < n[a]) {
            phase = {1.0, 0.0};
            return true;
        }
        if (!periodic[a]) return false;
        const double turns = to < 0 ? -1.0 : 1.0, period = (double)n[a] * h[a];
        const double angle = turns * 0.0 * period;
        phase = {1.0 * std::cos(angle), 1.0 * std::sin(angle)};
        to = ((to % n[a]) + n[a]) % n[a];
        return true;
    }
    bool fixed(long long r) const {
        int c, a[3];
        at(r, c, a);
        for (int ax = 0; ax < 3; ax++)
            if (ax != c && a[ax] == 0 && !periodic[ax]) return true;
        return false;
    }
    bool neighbour(int comp, const int a[3], int ax, int step, long long &r, C64 &phase) const {
        int to;
        if (!shift(ax, a[ax], step, to, phase)) return false;
        int m[3] = {a[0], a[1], a[2]};
        m[ax] = to;
        r = index(comp, m);
        return true;
    }
    // num-complex's real over complex: a (c - i d) / (c^2 + d^2)
    static C64 rdiv(double x, C64 b) {
        const double nn = b.re * b.re + b.im * b.im;
        return {x * b.re / nn, -x * b.im / nn};
    }
    // curl E at H's value r, curl H at E's value r: (value, coefficient) in the reference's order
    int curl_e(long long r, long long *cols, C64 *ws) const {
        int a, m[3], k = 0;
        at(r, a, m);
        const int b = (a + 1) % 3, c = (a + 2) % 3;
        const int alongs[2] = {b, c}, fields[2] = {c, b};
        const double signs[2] = {1.0, -1.0};
        for (int t = 0; t < 2; t++) {
            const int along = alongs[t], field = fields[t];
            const C64 st = halves[along][m[along]];
            const C64 w = rdiv(signs[t], {st.re * h[along], st.im * h[along]});
            long long col;
            C64 phase;
            if (neighbour(field, m, along, 1, col, phase)) {
                cols[k] = col;
                ws[k++] = cmul(w, phase);
            }
            cols[k] = index(field, m);
            ws[k++] = {-w.re, -w.im};
        }
        return k;
    }
    int curl_h(long long r, long long *cols, C64 *ws) const {
        int a, m[3], k = 0;
        at(r, a, m);
        const int b = (a + 1) % 3, c = (a + 2) % 3;
        const int alongs[2] = {b, c}, fields[2] = {c, b};
        const double signs[2] = {1.0, -1.0};
        for (int t = 0; t < 2; t++) {
            const int along = alongs[t], field = fields[t];
            const C64 st = nodes[along][m[along]];
            const C64 w = rdiv(signs[t], {st.re * h[along], st.im * h[along]});
            cols[k] = index(field, m);
            ws[k++] = w;
            long long col;
            C64 phase;
            if (neighbour(field, m, along, -1, col, phase)) {
                cols[k] = col;
                ws[k++] = cmul({-w.re, -w.im}, phase);
            }
        }
        return k;
    }
    // row r of -curl curl: up to 16 entries, none for a value fixed at zero
    int curl_curl(long long r, long long *cols, C64 *ws) const {
        if (fixed(r)) return 0;
        long long fc[4], ec[4];
        C64 fw[4], ew[4];
        const int nf = curl_h(r, fc, fw);
        int k = 0;
        for (int q = 0; q < nf; q++) {
            const int ne = curl_e(fc[q], ec, ew);
            for (int t = 0; t < ne; t++) {
                if (fixed(ec[t])) continue;
                cols[k] = ec[t];
                ws[k++] = cmul({-fw[q].re, -fw[q].im}, ew[t]);
            }
        }
        return k;
    }
    // a node (numbered as E_x's value there) on a wall behind a PML, where a potential is zero
    bool node_fixed(long long q) const {
        int c, a[3];
        at(q, c, a);
        for (int ax = 0; ax < 3; ax++)
            if (a[ax] == 0 && !periodic[ax]) return true;
        return false;
    }
    // the divergence at node q (Lattice::divergence, three.rs:454-466) and the gradient at E's
    // value r (Lattice::gradient, :490-506): (value or node, coefficient); node_fixed :445-451
    int divergence(long long q, long long *cols, C64 *ws) const {
        int c, m[3], k = 0;
        at(q, c, m);
        for (int a = 0; a < 3; a++) {
            const C64 st = nodes[a][m[a]];
            const C64 w = rdiv(1.0, {st.re * h[a], st.im * h[a]});
            cols[k] = index(a, m);
            ws[k++] = w;
            long long col;
            C64 phase;
            if (neighbour(a, m, a, -1, col, phase)) {
                cols[k] = col;
                ws[k++] = cmul({-w.re, -w.im}, phase);
            }
        }
        return k;
    }
    int gradient(long long r, long long *cols, C64 *ws) const {
        int a, m[3], k = 0;
        at(r, a, m);
        const C64 st = halves[a][m[a]];
        const C64 w = rdiv(1.0, {st.re * h[a], st.im * h[a]});
        long long q;
        C64 phase;
        if (neighbour(0, m, a, 1, q, phase) && !node_fixed(q)) {
            cols[k] = q;
            ws[k++] = cmul(w, phase);
        }
        const long long here = index(0, m);
        if (!node_fixed(here)) {
            cols[k] = here;
            ws[k++] = {-w.re, -w.im};
        }
        return k;
    }
    C64 volume(long long r) const {
        int c, m[3];
        at(r, c, m);
        C64 v = {1.0, 0.0};
        for (int a = 0; a < 3; a++) v = cmul(v, a == c ? halves[a][m[a]] : nodes[a][m[a]]);
        return v;
    }
    // Lattice::lorentz (src/fdfd/three/ports.rs:253-287): N(u, w) across the cut between node
    // `plane` and the next along `axis`
    template <typename U, typename W>
    C64 lorentz(int axis, int plane, const U &u, const W &w) const {
        const int b = (axis + 1) % 3, c = (axis + 2) % 3;
        C64 total = {0.0, 0.0};
        long long cols[16];
        C64 ws[16];
        for (int v = 0; v < n[c]; v++)
            for (int uu = 0; uu < n[b]; uu++)
                for (int comp : {b, c, axis}) {
                    int m[3];
                    m[axis] = plane;
                    m[b] = uu;
                    m[c] = v;
                    const long long r = index(comp, m);
                    const C64 ur = u(r), wr = w(r);
                    const int k = curl_curl(r, cols, ws);
                    C64 sum = {0.0, 0.0};
                    for (int q = 0; q < k; q++) {
                        int to, sat[3];
                        at(cols[q], to, sat);
                        if (to != axis && sat[axis] == plane + 1) {
                            const C64 t1 = cmul(wr, u(cols[q])), t2 = cmul(ur, w(cols[q]));
                            const C64 d = cmul(ws[q], {t1.re - t2.re, t1.im - t2.im});
                            sum = {sum.re + d.re, sum.im + d.im};
                        }
                    }
                    const C64 add = cmul(volume(r), sum);
                    total = {total.re + add.re, total.im + add.im};
                }
        const double vol = h[0] * h[1] * h[2];
        return cdiv({total.re * vol, total.im * vol}, {0.0, 4.0 * k0});
    }
};

// Simulation::leapfrog_frequency (src/fdtd/monitors.rs:560-567): the frequency whose leapfrog
// wavenumber is k0, (2 / dt) asin(k0 dt / 2) / 2 pi, or false beyond the Nyquist frequency
static bool leapfrog_frequency(const pxfdtd_simulation *s, double k0, double *f) {
    const double x = 0.5 * k0 * s->dt;
    if (!(x >= 0.0 && x < 1.0)) return false;
    *f = 2.0 / s->dt * pxfdtd::pasin(x) / TAU;
    return true;
}

// Simulation::add_mode_monitor (src/fdtd/monitors.rs:580-611)
extern "C" int pxfdtd_simulation_add_mode_monitor(pxfdtd_simulation *s, const pxfdtd_port_mode *modes,
                                                  int32_t count, int32_t *number) {
    if (!modes || count < 1) return px_fail(PXFDTD_ERR_INVALID, "a mode monitor needs at least one mode");
    const double steps[3] = {s->g.dx, s->g.dy, s->g.dz};
    ModeMonRec rec;
    for (int32_t q = 0; q < count; q++) {
        const pxfdtd_port_mode &m = modes[q];
        if (m.axis < 0 || m.axis > 2) return px_fail(PXFDTD_ERR_INVALID, "a mode's axis must be 0, 1 or 2");
        const int axis = m.axis, b = (axis + 1) % 3, c = (axis + 2) % 3;
        if (m.size[0] != s->n[b] || m.size[1] != s->n[c] || m.step != steps[axis] || m.plane < 0 ||
            m.plane + 1 >= s->n[axis])
            return px_fail(PXFDTD_ERR_INVALID,
                           "a mode's plane and the next must be this grid's: solve it with FDFD on the same grid");
        if (!m.tangential || !m.normal) return px_fail(PXFDTD_ERR_INVALID, "a mode needs its profile");
        ModeEntry e;
        if (!leapfrog_frequency(s, m.k0, &e.frequency))
            return px_fail(PXFDTD_ERR_INVALID, "the mode's k0 = %.17g rad/um is beyond the time step's Nyquist frequency",
                           m.k0);
        e.axis = m.axis;
        e.plane = m.plane;
        e.size[0] = m.size[0];
        e.size[1] = m.size[1];
        e.step = m.step;
        e.k0 = m.k0;
        e.beta[0] = m.beta[0];
        e.beta[1] = m.beta[1];
        const size_t pv = (size_t)m.size[0] * m.size[1];
        e.tangential.assign(m.tangential, m.tangential + 4 * pv);
        e.normal.assign(m.normal, m.normal + 2 * pv);
        rec.modes.push_back(std::move(e));
    }
    // the checks passed: the transforms (E on the plane and the next, all of the other two axes)
    for (ModeEntry &e : rec.modes) {
        int32_t low[3] = {0, 0, 0}, high[3] = {s->n[0] - 1, s->n[1] - 1, s->n[2] - 1};
        low[e.axis] = e.plane;
        high[e.axis] = e.plane + 1;
        TRY(add_transform(s, low, high, 1, &e.frequency, &e.dft));
    }
    s->modemons.push_back(std::move(rec));
    *number = (int32_t)s->modemons.size() - 1;
    return PXFDTD_OK;
}

extern "C" int32_t pxfdtd_simulation_mode_count(const pxfdtd_simulation *s, int32_t n) {
    if (n < 0 || n >= (int32_t)s->modemons.size()) return -1;
    return (int32_t)s->modemons[n].modes.size();
}

extern "C" int pxfdtd_simulation_mode_frequencies(const pxfdtd_simulation *s, int32_t n, double *out) {
    if (n < 0 || n >= (int32_t)s->modemons.size()) return px_fail(PXFDTD_ERR_INVALID, "no mode monitor %d", n);
    for (size_t q = 0; q < s->modemons[n].modes.size(); q++) out[q] = s->modemons[n].modes[q].frequency;
    return PXFDTD_OK;
}

// Simulation::mode_amplitudes (src/fdtd/monitors.rs:623-640) by mode_amplitudes_of and
// Lattice::mode_amplitudes (src/fdfd/three/ports.rs:637-651, 755-767): forward N(u, g) and
// backward -N(u, f), u the transforms of E, f and g the mode going forward and backward
extern "C" int pxfdtd_simulation_mode_amplitudes(pxfdtd_simulation *s, int32_t n, double *out) {
    if (n < 0 || n >= (int32_t)s->modemons.size()) return px_fail(PXFDTD_ERR_INVALID, "no mode monitor %d", n);
    const ModeMonRec &R = s->modemons[n];
    for (size_t q = 0; q < R.modes.size(); q++) {
        const ModeEntry &e = R.modes[q];
        const pxfdtd_port_mode mode = e.view();
        const DftRec &d = s->dfts[e.dft];
        const size_t len = d.acc.size() / 6;
        std::vector<double> series[3];
        for (int c = 0; c < 3; c++) {
            series[c].resize(len);
            TRY(transform_sums(s, e.dft, PXFDTD_E * 3 + c, series[c].data()));
        }
        const int ex = d.hi[0] - d.lo[0], ey = d.hi[1] - d.lo[1];
        const ModeLattice L(s, e.k0);
        auto u = [&](long long r) -> C64 {
            int c, a[3];
            L.at(r, c, a);
            const long long v = ((long long)(a[2] - d.lo[2]) * ey + (a[1] - d.lo[1])) * ex + (a[0] - d.lo[0]);
            return {series[c][2 * v], series[c][2 * v + 1]};
        };
        auto f = [&](long long r) -> C64 {
            int c, a[3];
            L.at(r, c, a);
            const long long at3[3] = {a[0], a[1], a[2]};
            return mode_value(&mode, c, at3, true);
        };
        auto g = [&](long long r) -> C64 {
            int c, a[3];
            L.at(r, c, a);
            const long long at3[3] = {a[0], a[1], a[2]};
            return mode_value(&mode, c, at3, false);
        };
        const C64 forward = L.lorentz(e.axis, e.plane, u, g);
        const C64 back = L.lorentz(e.axis, e.plane, u, f);
        out[4 * q + 0] = forward.re;
        out[4 * q + 1] = forward.im;
        out[4 * q + 2] = -back.re;
        out[4 * q + 3] = -back.im;
    }
    return PXFDTD_OK;
}

// harmonic_inversion (src/fdtd/harmonic.rs:62-184): up to `capacity` resonances into out as
// (frequency, decay, amplitude re, amplitude im, error), all found in *found
extern "C" int pxfdtd_harmonic_inversion(const double *signal, int64_t n, double dt, double low, double high,
                                         int32_t capacity, int32_t *found, double *out) {
    *found = 0;
    if (n < 0 || (n > 0 && !signal)) return px_fail(PXFDTD_ERR_INVALID, "harmonic inversion needs a signal");
    std::vector<std::complex<double>> c((size_t)n);
    for (int64_t q = 0; q < n; q++) c[(size_t)q] = {signal[2 * q], signal[2 * q + 1]};
    std::vector<pxharm::Resonance> r;
    std::string why;
    if (!pxharm::harmonic_inversion(c, dt, low, high, r, &why)) return px_fail(PXFDTD_ERR_INVALID, "%s", why.c_str());
    *found = (int32_t)r.size();
    for (int32_t q = 0; q < std::min<int32_t>(capacity, (int32_t)r.size()); q++) {
        out[5 * q] = r[q].frequency;
        out[5 * q + 1] = r[q].decay;
        out[5 * q + 2] = r[q].amplitude.real();
        out[5 * q + 3] = r[q].amplitude.imag();
        out[5 * q + 4] = r[q].error;
    }
    return PXFDTD_OK;
}

// Lattice::mode_amplitude_weights and lorentz_weights (src/fdfd/three/ports.rs:293-345): the
// amplitude of `mode` going `direction` as weights on E's transform over the mode's plane and
// the next, out [3][box] complex with the box [plane, plane + 1] along the axis and whole
// across ([c][k][j][i], box-local): the amplitude is sum w_r u_r. Each coupling adds
// f a w_r to u_s's weight and -f a w_s to u_r's, f = dx dy dz / (4 i k0) V_r.
extern "C" int pxfdtd_simulation_mode_weights(pxfdtd_simulation *s, const pxfdtd_port_mode *mode, int32_t direction,
                                              double *out) {
    if (!mode || mode->axis < 0 || mode->axis > 2 || direction < 0 || direction > 1)
        return px_fail(PXFDTD_ERR_INVALID, "mode weights need a mode and a direction (0 forward, 1 backward)");
    const int axis = mode->axis, b = (axis + 1) % 3, c = (axis + 2) % 3;
    const double steps[3] = {s->g.dx, s->g.dy, s->g.dz};
    if (mode->size[0] != s->n[b] || mode->size[1] != s->n[c] || mode->step != steps[axis] || mode->plane < 0 ||
        mode->plane + 1 >= s->n[axis])
        return px_fail(PXFDTD_ERR_INVALID, "a mode's plane and the next must be this grid's");
    const ModeLattice L(s, mode->k0);
    int bl[3] = {0, 0, 0}, bn[3] = {s->n[0], s->n[1], s->n[2]};
    bl[axis] = mode->plane;
    bn[axis] = 2;
    const long long vol = (long long)bn[0] * bn[1] * bn[2];
    std::fill(out, out + 6 * vol, 0.0);
    auto slot = [&](long long r) -> double * {
        int comp, m[3];
        L.at(r, comp, m);
        const long long v = ((long long)(m[2] - bl[2]) * bn[1] + (m[1] - bl[1])) * bn[0] + (m[0] - bl[0]);
        return out + 2 * (comp * vol + v);
    };
    const bool forward = direction == 0;
    auto w = [&](long long r) -> C64 {  // the mode going the other way
        int comp, m[3];
        L.at(r, comp, m);
        const long long at3[3] = {m[0], m[1], m[2]};
        return mode_value(mode, comp, at3, !forward);
    };
    const C64 scale = ModeLattice::rdiv(s->g.dx * s->g.dy * s->g.dz, {0.0, 4.0 * L.k0});
    const double sign = forward ? 1.0 : -1.0;
    long long cols[16];
    C64 ws[16];
    for (int v = 0; v < L.n[c]; v++)
        for (int uu = 0; uu < L.n[b]; uu++)
            for (int comp : {b, c, axis}) {
                int m[3];
                m[axis] = mode->plane;
                m[b] = uu;
                m[c] = v;
                const long long r = L.index(comp, m);
                const C64 wr = w(r);
                const int k = L.curl_curl(r, cols, ws);
                const C64 factor = cmul(scale, L.volume(r));
                for (int q = 0; q < k; q++) {
                    int to, sat[3];
                    L.at(cols[q], to, sat);
                    if (to == axis || sat[axis] != mode->plane + 1) continue;
                    const C64 fa = cmul(factor, ws[q]);
                    const C64 add = cmul(fa, wr), sub = cmul(fa, w(cols[q]));
                    double *os = slot(cols[q]), *orr = slot(r);
                    os[0] += sign * add.re;
                    os[1] += sign * add.im;
                    orr[0] -= sign * sub.re;
                    orr[1] -= sign * sub.im;
                }
            }
    return PXFDTD_OK;
}

// the permittivity each value of E sees, [3][N] (Simulation::diagonal's averages, or those given)
extern "C" int pxfdtd_simulation_permittivity(const pxfdtd_simulation *s, double *out) {
    std::memcpy(out, s->eps.get(), (size_t)(3 * s->N) * sizeof(double));
    return PXFDTD_OK;
}

// Lattice::flux_with and h_with (src/fdfd/three.rs:1138-1171): the scheme's flux through the
// half-plane after node `plane` of the field whose E value r is e(r), H~ = curl E / (i k0)
template <typename E>
static double lattice_flux(const ModeLattice &L, int axis, int plane, const E &e) {
    const int b = (axis + 1) % 3, c = (axis + 2) % 3;
    double total = 0.0;
    const int comps[2] = {b, c}, hs[2] = {c, b};
    const double signs[2] = {0.5, -0.5};
    long long cols[4];
    C64 ws[4];
    for (int t = 0; t < 2; t++)
        for (int v = 0; v < L.n[c]; v++)
            for (int u = 0; u < L.n[b]; u++) {
                int m[3];
                m[axis] = plane;
                m[b] = u;
                m[c] = v;
                const C64 low = e(L.index(comps[t], m));
                m[axis] = plane + 1;
                const C64 high = e(L.index(comps[t], m));
                m[axis] = plane;
                const int k = L.curl_e(L.index(hs[t], m), cols, ws);
                C64 curl = {0.0, 0.0};
                for (int q = 0; q < k; q++) {
                    const C64 p = cmul(ws[q], e(cols[q]));
                    curl = {curl.re + p.re, curl.im + p.im};
                }
                const C64 hv = cdiv(curl, {0.0, L.k0});
                const C64 mean = {0.5 * (low.re + high.re), 0.5 * (low.im + high.im)};
                total += signs[t] * (mean.re * hv.re - mean.im * -hv.im);
            }
    return total * L.h[b] * L.h[c];
}

// Lattice::port_modes (src/fdfd/three/ports.rs:408-594) on the simulation's grid and
// permittivity, at the leapfrog wavenumber of `frequency`. The unknowns are E across the plane
// inside the window; the rows are k0^2 eps E_t - curl_t curl_t E_t + grad_t(eps_a^-1 div_t(eps
// E_t)); the eigenvalues nearest k0^2 max(eps, 1) are beta_d^2, beta = (2/h) asin(beta_d h / 2)
// (the portable complex arcsine). E_a = i div_t(eps E_t) / (beta_d eps_a) e^(i beta h / 2)
// on the half-plane after; each mode normalized to N(f, g) = 1 with its largest tangential
// value's real part positive, going the way its power goes (propagating) or decays; sorted
// propagating first by |Re beta|, then the others by Im beta. The plane must be clear of the
// CPMLs as a mode source's (the reference's job solves on planes cut out of the grid there).
extern "C" int pxfdtd_simulation_port_modes(pxfdtd_simulation *s, int32_t axis, int32_t plane,
                                            const int32_t window[2][2], double frequency, int32_t count,
                                            int32_t lossy, int32_t *found, double *beta_out,
                                            double *tangential_out, double *normal_out, double *power_out) {
    *found = 0;
    if (axis < 0 || axis > 2) return px_fail(PXFDTD_ERR_INVALID, "a port's axis must be 0, 1 or 2");
    if (count < 1) return px_fail(PXFDTD_ERR_INVALID, "ask for one mode or more");
    const int b = (axis + 1) % 3, c = (axis + 2) % 3;
    if (s->b.periodic[axis])
        return px_fail(PXFDTD_ERR_INVALID, "a port normal to axis %d needs ends along it, not periodic ones", axis);
    const long long na = s->n[axis], low = s->b.low[axis], high = s->b.high[axis];
    if (plane < low + 1 || plane + 2 + high > na)
        return px_fail(PXFDTD_ERR_INVALID,
                       "the port's plane %d and the next must be clear of the CPMLs (%lld and %lld cells) and the "
                       "ends of %lld cells along axis %d", plane, low, high, na, axis);
    if (!(frequency > 0) || TAU * frequency * s->dt >= M_PI)
        return px_fail(PXFDTD_ERR_INVALID, "%g c/um is beyond the time step's Nyquist frequency", frequency);
    int wlo[2], whi[2];
    const int others[2] = {b, c};
    for (int w = 0; w < 2; w++) {
        const int along = others[w], cells = s->n[along];
        const bool whole = !window || window[w][0] < 0;
        wlo[w] = whole ? 0 : window[w][0];
        whi[w] = whole ? cells : window[w][1];
        if (s->b.periodic[along]) {
            if (wlo[w] != 0 || whi[w] != cells)
                return px_fail(PXFDTD_ERR_INVALID, "a periodic axis %d must be whole in the port's window", along);
        } else if (whi[w] > cells || wlo[w] < 0 || whi[w] - wlo[w] < 3) {
            return px_fail(PXFDTD_ERR_INVALID,
                           "the window along axis %d, cells %d to %d, must have 3 cells or more on the grid's %d",
                           along, wlo[w], whi[w], cells);
        }
    }
    double sn = 0.0, cs = 0.0;
    pxfdtd::psincos(TAU * frequency * s->dt / 2.0, sn, cs);
    const double k0 = 2.0 / s->dt * sn, k2 = k0 * k0;
    const ModeLattice L(s, k0);
    const int size[2] = {s->n[b], s->n[c]};
    const long long per = (long long)size[0] * size[1];
    // the window's values: tangential on the plane, normal on the half-plane after it
    auto inside = [&](long long r) {
        int comp, m[3];
        L.at(r, comp, m);
        if (m[axis] != plane) return false;
        for (int w = 0; w < 2; w++) {
            const int along = others[w];
            if (s->b.periodic[along]) continue;
            const int mm = m[along];
            const bool half = comp == along;
            if (half ? !(mm >= wlo[w] && mm < whi[w]) : !(mm > wlo[w] && mm < whi[w])) return false;
        }
        return true;
    };
    // the unknowns, numbered (v, u, component) so that the matrix's band is about 2 size[0]
    std::vector<long long> local((size_t)(2 * per), -1), unknowns;
    for (int v = 0; v < size[1]; v++)
        for (int u = 0; u < size[0]; u++)
            for (int t = 0; t < 2; t++) {
                int m[3];
                m[axis] = plane;
                m[b] = u;
                m[c] = v;
                const long long r = L.index(others[t], m);
                if (inside(r) && !L.fixed(r)) {
                    local[(size_t)(t * per + (long long)v * size[0] + u)] = (long long)unknowns.size();
                    unknowns.push_back(r);
                }
            }
    if (unknowns.empty()) return px_fail(PXFDTD_ERR_INVALID, "the window holds no field");
    auto number = [&](long long r) -> long long {
        int comp, m[3];
        L.at(r, comp, m);
        if (comp == axis || m[axis] != plane) return -1;
        const int t = comp == c ? 1 : 0;
        return local[(size_t)(t * per + (long long)m[c] * size[0] + m[b])];
    };
    // the permittivity the leapfrog sees at the frequency: eps, or with `lossy` eps + i sigma
    // cos(w dt / 2) / k0 (the conductivity's semi-implicit average; r = component * N + cell)
    const double loss_factor = cs / k0;
    auto eps_of = [&](long long r) -> C64 {
        const double im = lossy && !s->sigma.empty() ? s->sigma[(size_t)r] * loss_factor : 0.0;
        return {s->eps[(size_t)r], im};
    };
    auto normal_index = [&](long long q) {
        int comp, m[3];
        L.at(q, comp, m);
        return L.index(axis, m);
    };
    std::vector<std::tuple<long long, long long, pxeig::cd>> entries;
    double top = 1.0;
    long long cols[16], cols2[16];
    C64 ws[16], ws2[16];
    for (size_t row = 0; row < unknowns.size(); row++) {
        const long long r = unknowns[row];
        const C64 er = eps_of(r);
        top = std::max(top, er.re);
        entries.emplace_back((long long)row, (long long)row, pxeig::cd(k2 * er.re, k2 * er.im));
        const int nf = L.curl_h(r, cols, ws);
        for (int q = 0; q < nf; q++) {
            int fc, fm[3];
            L.at(cols[q], fc, fm);
            if (fc != axis) continue;
            const int ne = L.curl_e(cols[q], cols2, ws2);
            for (int t = 0; t < ne; t++) {
                const long long k = number(cols2[t]);
                if (k < 0) continue;
                const C64 v = cmul({-ws[q].re, -ws[q].im}, ws2[t]);
                entries.emplace_back((long long)row, k, pxeig::cd(v.re, v.im));
            }
        }
        const int ng = L.gradient(r, cols, ws);
        for (int q = 0; q < ng; q++) {
            const long long nq = normal_index(cols[q]);
            if (!inside(nq)) continue;
            const C64 inverse = ModeLattice::rdiv(1.0, eps_of(nq));
            const int nd = L.divergence(cols[q], cols2, ws2);
            for (int t = 0; t < nd; t++) {
                const long long k = number(cols2[t]);
                if (k < 0) continue;
                const C64 v = cmul(cmul(cmul(ws[q], ws2[t]), eps_of(cols2[t])), inverse);
                entries.emplace_back((long long)row, k, pxeig::cd(v.re, v.im));
            }
        }
    }
    std::string why;
    const std::vector<pxeig::Pair> pairs =
        pxeig::nearest((long long)unknowns.size(), entries, pxeig::cd(k2 * top, 0.0), count, 1e-10, &why);
    if (pairs.empty()) return px_fail(PXFDTD_ERR_INVALID, "the port's modes: %s", why.c_str());
    const double hstep = L.h[axis];
    struct Built {
        C64 beta;
        std::vector<double> tangential, normal;
        double power;
    };
    auto propagates = [](C64 bt) { return std::fabs(bt.im) < std::fabs(bt.re); };
    auto build = [&](const pxeig::Pair &pair, C64 beta_d) -> Built {
        Built m;
        const C64 as = pxfdtd::pcasin({beta_d.re * hstep / 2.0, beta_d.im * hstep / 2.0});
        m.beta = {2.0 / hstep * as.re, 2.0 / hstep * as.im};
        m.tangential.assign((size_t)(4 * per), 0.0);
        m.normal.assign((size_t)(2 * per), 0.0);
        auto at_plane = [&](long long r) -> C64 {
            const long long k = number(r);
            if (k < 0) return {0.0, 0.0};
            return {pair.vector[k].real(), pair.vector[k].imag()};
        };
        for (long long t = 0; t < 2 * per; t++) {
            const long long k = local[(size_t)t];
            if (k < 0) continue;
            m.tangential[2 * t] = pair.vector[k].real();
            m.tangential[2 * t + 1] = pair.vector[k].imag();
        }
        const C64 half = pxfdtd::pcexp(cmul({0.0, 0.5}, {m.beta.re * hstep, m.beta.im * hstep}));
        for (int v = 0; v < size[1]; v++)
            for (int u = 0; u < size[0]; u++) {
                int mm[3];
                mm[axis] = plane;
                mm[b] = u;
                mm[c] = v;
                const long long r = L.index(axis, mm);
                if (!inside(r) || L.fixed(r)) continue;
                const int nd = L.divergence(L.index(0, mm), cols, ws);
                C64 div = {0.0, 0.0};
                for (int t = 0; t < nd; t++) {
                    int cc, cm[3];
                    L.at(cols[t], cc, cm);
                    if (cc == axis) continue;
                    const C64 p = cmul(cmul(ws[t], eps_of(cols[t])), at_plane(cols[t]));
                    div = {div.re + p.re, div.im + p.im};
                }
                const C64 val = cmul(cdiv(cmul({0.0, 1.0}, div), cmul(beta_d, eps_of(r))), half);
                const long long o = (long long)v * size[0] + u;
                m.normal[2 * o] = val.re;
                m.normal[2 * o + 1] = val.im;
            }
        pxfdtd_port_mode view = {axis, plane, {size[0], size[1]}, hstep, k0, {m.beta.re, m.beta.im},
                                 m.tangential.data(), m.normal.data()};
        auto fwd = [&](long long r) -> C64 {
            int cc, a[3];
            L.at(r, cc, a);
            const long long at3[3] = {a[0], a[1], a[2]};
            return mode_value(&view, cc, at3, true);
        };
        auto bwd = [&](long long r) -> C64 {
            int cc, a[3];
            L.at(r, cc, a);
            const long long at3[3] = {a[0], a[1], a[2]};
            return mode_value(&view, cc, at3, false);
        };
        const C64 norm = L.lorentz(axis, plane, fwd, bwd);
        C64 scale = cdiv({1.0, 0.0}, pxfdtd::pcsqrt(norm));
        C64 peak = {1.0, 0.0};
        double best = -1.0;
        for (long long t = 0; t < 2 * per; t++) {  // the last of the largest, as max_by gives
            const C64 x = {m.tangential[2 * t], m.tangential[2 * t + 1]};
            const double size = pxfdtd::pcabs(x);
            if (size >= best) {
                best = size;
                peak = x;
            }
        }
        if (cmul(peak, scale).re < 0.0) scale = {-scale.re, -scale.im};
        for (long long t = 0; t < 2 * per; t++) {
            const C64 x = cmul({m.tangential[2 * t], m.tangential[2 * t + 1]}, scale);
            m.tangential[2 * t] = x.re;
            m.tangential[2 * t + 1] = x.im;
        }
        for (long long t = 0; t < per; t++) {
            const C64 x = cmul({m.normal[2 * t], m.normal[2 * t + 1]}, scale);
            m.normal[2 * t] = x.re;
            m.normal[2 * t + 1] = x.im;
        }
        m.power = lattice_flux(L, axis, plane, fwd);
        return m;
    };
    std::vector<Built> modes;
    for (const pxeig::Pair &pair : pairs) {
        C64 bd = pxfdtd::pcsqrt({pair.value.real(), pair.value.imag()});
        const bool prop = propagates(bd);
        if ((!prop && bd.im < 0.0) || (prop && bd.re < 0.0)) bd = {-bd.re, -bd.im};
        Built m = build(pair, bd);
        if (prop && m.power < 0.0) m = build(pair, {-bd.re, -bd.im});
        modes.push_back(std::move(m));
    }
    std::stable_sort(modes.begin(), modes.end(), [&](const Built &p, const Built &q) {
        const bool ep = !propagates(p.beta), eq = !propagates(q.beta);
        if (ep != eq) return !ep;
        if (std::fabs(p.beta.re) != std::fabs(q.beta.re)) return std::fabs(p.beta.re) > std::fabs(q.beta.re);
        return p.beta.im < q.beta.im;
    });
    for (size_t q = 0; q < modes.size(); q++) {
        beta_out[2 * q] = modes[q].beta.re;
        beta_out[2 * q + 1] = modes[q].beta.im;
        std::memcpy(tangential_out + q * 4 * per, modes[q].tangential.data(), (size_t)(4 * per) * sizeof(double));
        std::memcpy(normal_out + q * 2 * per, modes[q].normal.data(), (size_t)(2 * per) * sizeof(double));
        power_out[q] = modes[q].power;
    }
    *found = (int32_t)modes.size();
    return PXFDTD_OK;
}

// ---------------------------------------------------------------- for the serving loop
// (resident.hip's server drives simulations through these; not part of the C ABI)
namespace pxfdtd_host {

int sim_footprint(pxfdtd_simulation *s, int32_t precision, size_t *bytes) {
    TRY(sync_host(s));
    pxfdtd_problem p;
    std::vector<pxfdtd_medium> media;
    make_problem(s, precision, p, media);
    TRY(pxfdtd_resident_footprint(&p, bytes));
    // taps, staging and FP64 sums of the transform monitors
    long long vals = 0, sums = 0;
    for (auto &d : s->dfts) {
        long long v = 1;
        for (int a = 0; a < 3; a++) v *= d.hi[a] - d.lo[a];
        vals += 6 * v;
        sums += 6 * v * (long long)d.freqs.size();
    }
    const size_t esz = precision == PXFDTD_F64 ? 8 : 4;
    *bytes += (size_t)vals * (9 + 32 * esz + 4) + (size_t)sums * 16;
    return PXFDTD_OK;
}

int sim_servable(const pxfdtd_simulation *s) {
    if (!s->probes.empty())
        return px_fail(PXFDTD_ERR_UNSUPPORTED, "the server records transform monitors, not probes: a job with probes runs alone");
    if (!s->planes.empty())
        return px_fail(PXFDTD_ERR_UNSUPPORTED, "a job with a plane wave runs alone: its auxiliary grid steps outside the batched kernels");
    return PXFDTD_OK;
}

int sim_attach(pxfdtd_simulation *s, pxfdtd_device *dev, int32_t precision, pxfdtd_resident **out) {
    TRY(pxfdtd_simulation_set_device(s, dev, precision));
    *out = s->res;
    return PXFDTD_OK;
}

int sim_emitters(const pxfdtd_simulation *s) { return s->emitters; }

void sim_set_steps(pxfdtd_simulation *s, int64_t steps) { s->steps = steps; }

double sim_dft_frequency(const pxfdtd_simulation *s, int monitor, int f) {
    if (monitor < 0 || monitor >= (int)s->shown.size()) return -1.0;
    const DftRec &d = s->dfts[s->shown[monitor]];
    if (f < 0 || f >= (int)d.freqs.size()) return -1.0;
    return d.freqs[f];
}

// the caller's transform monitor's box, lo inclusive and hi exclusive, and its index among
// all the simulation's transform monitors (the device's series order)
int sim_dft_box(const pxfdtd_simulation *s, int monitor, int lo[3], int hi[3]) {
    if (monitor < 0 || monitor >= (int)s->shown.size()) return px_fail(PXFDTD_ERR_INVALID, "no transform monitor %d", monitor);
    const DftRec &d = s->dfts[s->shown[monitor]];
    for (int a = 0; a < 3; a++) {
        lo[a] = d.lo[a];
        hi[a] = d.hi[a];
    }
    return PXFDTD_OK;
}

bool sim_has_media(const pxfdtd_simulation *s) { return !s->groups.empty(); }
bool sim_has_planes(const pxfdtd_simulation *s) { return !s->planes.empty(); }

// the E sources' values (component, index) and whether any E current or source sits there
bool sim_e_source_at(const pxfdtd_simulation *s, int comp, long long r) {
    for (auto &so : s->sources)
        if (so.field == PXFDTD_E && so.comp == comp && so.r == r) return true;
    for (auto &cu : s->currents)
        if (cu.field == PXFDTD_E)
            for (size_t v = 0; v < cu.r.size(); v++)
                if (cu.comps[v] == comp && cu.r[v] == r) return true;
    return false;
}

const double *sim_ca(const pxfdtd_simulation *s) { return s->ca.data(); }
const double *sim_cb(const pxfdtd_simulation *s) { return s->cb.data(); }
int64_t sim_steps(const pxfdtd_simulation *s) { return s->steps; }

void sim_tables(const pxfdtd_simulation *s, int64_t n0, int count, int slots, double *waves, double *phases) {
    fill_tables(s, n0, count, slots, waves, phases);
}

void sim_advanced(pxfdtd_simulation *s, int64_t count) {
    s->steps += count;
    if (count) s->device_newer = true;
}

// the results to the host and the device state freed: the simulation has no device again
int sim_detach(pxfdtd_simulation *s) {
    TRY(sync_host(s));
    drop_resident(s);
    s->dev = nullptr;
    return PXFDTD_OK;
}

// the two halves of sim_detach: the results to the host (device memory untouched, so it may
// run while other work is on the device), then the device state freed
int sim_fetch(pxfdtd_simulation *s) { return sync_host(s); }

void sim_release(pxfdtd_simulation *s) {
    drop_resident(s);
    s->dev = nullptr;
}

}  // namespace pxfdtd_host

extern "C" int pxfdtd_simulation_footprint(pxfdtd_simulation *s, int32_t precision, int64_t *bytes) {
    if (precision != PXFDTD_F32 && precision != PXFDTD_F64)
        return px_fail(PXFDTD_ERR_INVALID, "precision must be PXFDTD_F32 or PXFDTD_F64");
    size_t b = 0;
    TRY(pxfdtd_host::sim_footprint(s, precision, &b));
    *bytes = (int64_t)b;
    return PXFDTD_OK;
}

// the media's share of a step's algorithmic bytes: per value in a medium, E at the previous
// step and each term's two currents, each read and written
extern "C" int pxfdtd_simulation_media_bytes_per_step(const pxfdtd_simulation *s, int32_t precision, int64_t *bytes) {
    if (precision != PXFDTD_F32 && precision != PXFDTD_F64)
        return px_fail(PXFDTD_ERR_INVALID, "precision must be PXFDTD_F32 or PXFDTD_F64");
    long long v = 0;
    for (const Group &G : s->groups) v += (long long)G.r.size() * (2 + 4 * G.poles);
    *bytes = (int64_t)(v * (precision == PXFDTD_F32 ? 4 : 8));
    return PXFDTD_OK;
}

// Yee::bytes_per_step (src/fdtd/kernel/yee.rs:141-153): the reference's byte model of a step,
// (24 N + 6 psi_E + 5 psi_H) values of the precision's size
extern "C" int pxfdtd_simulation_bytes_per_step(const pxfdtd_simulation *s, int32_t precision, int64_t *bytes) {
    if (precision != PXFDTD_F32 && precision != PXFDTD_F64)
        return px_fail(PXFDTD_ERR_INVALID, "precision must be PXFDTD_F32 or PXFDTD_F64");
    long long slabs = 0;
    for (const Slab &sl : s->slabs) slabs += (long long)sl.psi.size() * (sl.field == PXFDTD_E ? 6 : 5);
    *bytes = (int64_t)((s->N * (9 + 15) + slabs) * (precision == PXFDTD_F32 ? 4 : 8));
    return PXFDTD_OK;
}

// ---------------------------------------------------------------- designs and the frequency-domain adjoint
// src/fdtd/adjoint.rs: a design region's densities, the transforms of E over the values its
// cells reach, and the adjoint run whose transforms, with the forward run's, give the gradient.

namespace {

long long design_len(const pxfdtd_design &d) {
    return (long long)(d.hi[0] - d.lo[0]) * (d.hi[1] - d.lo[1]) * (d.hi[2] - d.lo[2]);
}

// Design::index (adjoint.rs:95-103): i fastest, then j, then k; -1 outside the region
long long design_index(const pxfdtd_design &d, long long i, long long j, long long k) {
    if (i < d.lo[0] || i >= d.hi[0] || j < d.lo[1] || j >= d.hi[1] || k < d.lo[2] || k >= d.hi[2]) return -1;
    const long long nx = d.hi[0] - d.lo[0], ny = d.hi[1] - d.lo[1];
    return ((k - d.lo[2]) * ny + (j - d.lo[1])) * nx + (i - d.lo[0]);
}

// Design::permittivity (adjoint.rs:105-108)
double design_eps(const pxfdtd_design &d, double rho) { return d.eps[0] + rho * (d.eps[1] - d.eps[0]); }

// Design::cell_at (adjoint.rs:148-167): the design cell holding p, across a periodic side wrapped
// onto the grid; -1 for none
long long design_cell_at(const pxfdtd_simulation *s, const pxfdtd_design &d, const double p[3]) {
    const double origin[3] = {s->g.x0, s->g.y0, s->g.z0}, hs[3] = {s->g.dx, s->g.dy, s->g.dz};
    long long at[3];
    for (int a = 0; a < 3; a++) {
        const long long n = s->n[a];
        long long m = (long long)std::floor((p[a] - origin[a]) / hs[a]);
        if (s->b.periodic[a])
            m = ((m % n) + n) % n;
        else if (m < 0 || m >= n)
            return -1;
        at[a] = m;
    }
    return design_index(d, at[0], at[1], at[2]);
}

// Design::check (adjoint.rs:110-140): on the grid, clear of the CPMLs with the values of E its
// cells reach, its permittivities finite and positive
int design_check(const pxfdtd_simulation *s, const pxfdtd_design *d) {
    if (!d) return px_fail(PXFDTD_ERR_INVALID, "a design is required");
    for (double e : d->eps)
        if (!(std::isfinite(e) && e > 0.0))
            return px_fail(PXFDTD_ERR_INVALID, "a design's permittivities must be finite and positive, got %g and %g",
                           d->eps[0], d->eps[1]);
    for (int a = 0; a < 3; a++) {
        const int n = s->n[a];
        if (d->lo[a] < 0 || d->lo[a] >= d->hi[a] || d->hi[a] > n)
            return px_fail(PXFDTD_ERR_INVALID,
                           "a design's cells along axis %d must be a nonempty range within the grid's %d, got %d..%d", a,
                           n, d->lo[a], d->hi[a]);
        const int low = s->b.low[a], high = s->b.high[a];
        if (!s->b.periodic[a] && (low > 0 || high > 0) && (d->lo[a] < low || d->hi[a] + high >= n))
            return px_fail(PXFDTD_ERR_INVALID,
                           "a design's cells along axis %d, %d..%d, and the values of E they reach must be clear of the "
                           "CPMLs (%d and %d of %d cells)", a, d->lo[a], d->hi[a], low, high, n);
    }
    return PXFDTD_OK;
}

// Design::check_density (adjoint.rs:142-146)
int design_check_density(const pxfdtd_design &d, const double *density, int64_t n) {
    if (n != design_len(d) || (n > 0 && !density))
        return px_fail(PXFDTD_ERR_INVALID, "a design of %lld cells needs as many densities, got %lld", design_len(d),
                       (long long)n);
    for (int64_t v = 0; v < n; v++)
        if (!(std::isfinite(density[v]) && design_eps(d, density[v]) > 0.0))
            return px_fail(PXFDTD_ERR_INVALID, "a density must be finite and give a positive permittivity, got %g",
                           density[v]);
    return PXFDTD_OK;
}

// Design::values (adjoint.rs:196-255): every value of E whose cell reaches the design's cells,
// each design cell's share of it (the 8 x 8 samples across the component; along it the cell lies
// in one design cell), component by component, then k, j, i
std::vector<Touched> design_values(const pxfdtd_simulation *s, const pxfdtd_design &d) {
    const int samples = 8;
    const double n = samples;
    auto offset = [&](int q) { return (q + 0.5) / n - 0.5; };
    const double h[3] = {s->g.dx, s->g.dy, s->g.dz};
    std::vector<Touched> out;
    for (int c = 0; c < 3; c++) {
        std::vector<int> cand[3];
        for (int a = 0; a < 3; a++) {
            std::vector<int> set;
            const int len = s->n[a];
            const int last = a == c ? d.hi[a] - 1 : d.hi[a];
            for (int m = d.lo[a]; m <= last; m++) {
                if (m < len)
                    set.push_back(m);
                else if (s->b.periodic[a])
                    set.push_back(m % len);
            }
            std::sort(set.begin(), set.end());
            set.erase(std::unique(set.begin(), set.end()), set.end());
            cand[a] = set;
        }
        const int b = (c + 1) % 3, cc = (c + 2) % 3;
        for (int k : cand[2])
            for (int j : cand[1])
                for (int i : cand[0]) {
                    const long long r = ((long long)k * s->n[1] + j) * s->n[0] + i;
                    double centre[3];
                    s->e_position(c, r, centre);
                    std::map<long long, double> shares;
                    double p[3] = {centre[0], centre[1], centre[2]};
                    for (int sb = 0; sb < samples; sb++) {
                        p[b] = centre[b] + offset(sb) * h[b];
                        for (int sc = 0; sc < samples; sc++) {
                            p[cc] = centre[cc] + offset(sc) * h[cc];
                            p[c] = centre[c];
                            const long long v = design_cell_at(s, d, p);
                            if (v >= 0) shares[v] += 1.0 / (n * n);
                        }
                    }
                    if (!shares.empty()) {
                        Touched t;
                        t.comp = c;
                        t.at[0] = i;
                        t.at[1] = j;
                        t.at[2] = k;
                        t.r = r;
                        t.shares.assign(shares.begin(), shares.end());
                        out.push_back(std::move(t));
                    }
                }
    }
    return out;
}

// every transform monitor's sums, concatenated in the order of s->dfts (the device's, which went
// on from the host's, when the device holds them)
int all_sums(pxfdtd_simulation *s, std::vector<double> &out) {
    if (s->res && !s->taps_dirty) {
        out.resize(pxfdtd_resident_sums_len(s->res));
        return pxfdtd_resident_read_sums(s->res, out.data(), (int64_t)out.size());
    }
    out.clear();
    for (const DftRec &d : s->dfts) out.insert(out.end(), d.acc.begin(), d.acc.end());
    return PXFDTD_OK;
}

size_t dft_offset(const pxfdtd_simulation *s, int index) {
    size_t at = 0;
    for (int m = 0; m < index; m++) at += s->dfts[m].acc.size();
    return at;
}

// DesignMonitor::value (adjoint.rs:285-291): the transform at frequency f of the value t reads,
// from the sums `all`
C64 design_value(const pxfdtd_simulation *s, const DesignRec &R, const std::vector<double> &all, const Touched &t,
                 int f) {
    const int index = R.dft[t.comp];
    const DftRec &d = s->dfts[index];
    const size_t len = d.acc.size() / 6;
    const double *series = all.data() + dft_offset(s, index) + (size_t)(PXFDTD_E * 3 + t.comp) * len;
    const long long ex = d.hi[0] - d.lo[0], ey = d.hi[1] - d.lo[1], ez = d.hi[2] - d.lo[2];
    const long long q = (((long long)f * ez + (t.at[2] - d.lo[2])) * ey + (t.at[1] - d.lo[1])) * ex + (t.at[0] - d.lo[0]);
    return {series[2 * q], series[2 * q + 1]};
}

// num-complex's quotient of a real by a complex
C64 rcdiv(double x, C64 z) {
    const double n = z.re * z.re + z.im * z.im;
    return {x * z.re / n, (0.0 - x * z.im) / n};
}

// inverse_stretch (adjoint.rs:682-702): 1/s~ of the CPML along `axis` at index m, on a node
// (E's convolutions) or halfway after it (H~'s), 1/kappa + a/(1 - b e^(i w dt)); 1 outside
C64 inverse_stretch(const pxfdtd_simulation *s, int axis, int m, bool half, double omega) {
    C64 v = {half ? s->kh[axis][m] : s->kn[axis][m], 0.0};
    const int field = half ? PXFDTD_H : PXFDTD_E;
    for (const Slab &sl : s->slabs)
        if (sl.field == field && sl.axis == axis && m >= sl.from && m < sl.from + sl.count) {
            const int q = m - sl.from;
            double sn, cs;
            pxfdtd::psincos(omega * s->dt, sn, cs);
            const C64 bz = {sl.b[q] * cs, sl.b[q] * sn};
            const C64 t = rcdiv(sl.a[q], {1.0 - bz.re, 0.0 - bz.im});
            v = {v.re + t.re, v.im + t.im};
            break;
        }
    return v;
}

// symmetrizer (adjoint.rs:704-718): D at field's component at `at`, the product of the CPMLs'
// stretches along the three axes there; 1 outside the CPMLs
C64 symmetrizer(const pxfdtd_simulation *s, int field, int comp, const int at[3], double omega) {
    C64 d = {1.0, 0.0};
    for (int axis = 0; axis < 3; axis++) {
        const bool half = field == PXFDTD_E ? axis == comp : axis != comp;
        d = cdiv(d, inverse_stretch(s, axis, at[axis], half, omega));
    }
    return d;
}

// Waveform::spectrum (src/fdtd.rs:353-396): sum of w(t_n) e^(i w t_n) dt over the first `steps`
// steps, at the times a current on `field` takes its values (E: (n + 1/2) dt, H~: n dt)
C64 waveform_spectrum(const pxfdtd_waveform &w, int field, double dt, long long steps, double frequency) {
    const double omega = TAU * frequency;
    const double offset = field == PXFDTD_E ? 0.5 * dt : 0.0;
    C64 sum = {0.0, 0.0};
    for (long long n = 0; n < steps; n++) {
        const double t = (double)n * dt + offset;
        const double v = wave_at(w, t);
        const C64 e = pxfdtd::pcexp({0.0, omega * t});
        const C64 p = {v * e.re - 0.0 * e.im, v * e.im + 0.0 * e.re};
        sum = {sum.re + p.re * dt, sum.im + p.im * dt};
    }
    return sum;
}

// basis (adjoint.rs:1011-1031): per frequency, Waveform::pulse on it and the same a quarter
// period later
int adjoint_basis(const std::vector<double> &freqs, double bandwidth, std::vector<pxfdtd_waveform> &out) {
    if (!(std::isfinite(bandwidth) && bandwidth > 0.0))
        return px_fail(PXFDTD_ERR_INVALID, "a pulse's bandwidth must be positive and finite, got %g", bandwidth);
    out.clear();
    for (double f : freqs) {
        const double width = std::sqrt(2.0 * 0.6931471805599453) / (3.141592653589793 * bandwidth);
        for (double shift : {0.0, 0.25 / f}) {
            pxfdtd_waveform w{};
            w.kind = 0;
            w.frequency = f;
            w.width = width;
            w.delay = 6.0 * width + shift;
            w.amp_re = 1.0;
            out.push_back(w);
        }
    }
    return PXFDTD_OK;
}

// Simulation::amplitudes (adjoint.rs:571-621): the real amplitudes b_m of each value's current
// sum b_m w_m(t) whose transform at each frequency is its target there, the real system
// sum b_m W_m(w_j) = target_j (real and imaginary rows) solved by LU with partial pivoting
int adjoint_amplitudes(const pxfdtd_simulation *s, int field, const std::vector<pxfdtd_waveform> &waves,
                       const std::vector<double> &freqs, const std::vector<const std::vector<C64> *> &targets,
                       std::vector<std::vector<double>> &out) {
    const int m = (int)waves.size();
    long long steps = 0;
    for (const pxfdtd_waveform &w : waves)
        steps = std::max(steps, (long long)std::ceil((w.delay + 8.0 * w.width) / s->dt) + 1);
    std::vector<double> a((size_t)m * m), rhs((size_t)m * targets.size());
    for (int q = 0; q < m; q++)
        for (size_t j = 0; j < freqs.size(); j++) {
            const C64 t = waveform_spectrum(waves[q], field, s->dt, steps, freqs[j]);
            a[(2 * j) * m + q] = t.re;
            a[(2 * j + 1) * m + q] = t.im;
        }
    const size_t nt = targets.size();
    for (size_t col = 0; col < nt; col++)
        for (size_t j = 0; j < freqs.size(); j++) {
            rhs[(2 * j) * nt + col] = (*targets[col])[j].re;
            rhs[(2 * j + 1) * nt + col] = (*targets[col])[j].im;
        }
    // Gaussian elimination with partial pivoting on [a | rhs]
    for (int k = 0; k < m; k++) {
        int p = k;
        for (int i = k + 1; i < m; i++)
            if (std::fabs(a[(size_t)i * m + k]) > std::fabs(a[(size_t)p * m + k])) p = i;
        if (p != k) {
            for (int c = 0; c < m; c++) std::swap(a[(size_t)k * m + c], a[(size_t)p * m + c]);
            for (size_t c = 0; c < nt; c++) std::swap(rhs[k * nt + c], rhs[p * nt + c]);
        }
        const double piv = a[(size_t)k * m + k];
        for (int i = k + 1; i < m; i++) {
            const double l = a[(size_t)i * m + k] / piv;
            if (l == 0.0) continue;
            for (int c = k; c < m; c++) a[(size_t)i * m + c] -= l * a[(size_t)k * m + c];
            for (size_t c = 0; c < nt; c++) rhs[i * nt + c] -= l * rhs[k * nt + c];
        }
    }
    out.assign(nt, std::vector<double>(m));
    for (size_t col = 0; col < nt; col++)
        for (int k = m - 1; k >= 0; k--) {
            double v = rhs[k * nt + col];
            for (int c = k + 1; c < m; c++) v -= a[(size_t)k * m + c] * out[col][c];
            out[col][k] = v / a[(size_t)k * m + k];
            if (!std::isfinite(out[col][k]))
                return px_fail(PXFDTD_ERR_INVALID,
                               "the adjoint's pulses can't give its sources' transforms: frequencies too close for the "
                               "bandwidth");
        }
    return PXFDTD_OK;
}

struct Derivative {
    int field, comp;
    long long r;
    C64 g;
};

// flux_derivatives (adjoint.rs:1033-1082): a flux face's derivative with respect to each
// transform its sum pairs, times `scale`: the flux is sum kappa Re(E conj(H~)), so d/dE = kappa
// conj(H~) / 2 and d/dH~ = kappa conj(E) / 2
void flux_derivatives(const pxfdtd_simulation *s, const FluxFace &F, const double *series0, size_t len, int f,
                      double scale, std::vector<Derivative> &out) {
    const DftRec &d = s->dfts[F.dft];
    const int a = F.axis, b = (a + 1) % 3, c = (a + 2) % 3;
    const long long ex = d.hi[0] - d.lo[0], ey = d.hi[1] - d.lo[1], ez = d.hi[2] - d.lo[2];
    auto value = [&](int field, int comp, const int at[3]) {
        const double *v = series0 + (size_t)(field * 3 + comp) * len;
        const long long q = (((long long)f * ez + (at[2] - d.lo[2])) * ey + (at[1] - d.lo[1])) * ex + (at[0] - d.lo[0]);
        return C64{v[2 * q], v[2 * q + 1]};
    };
    auto index = [&](const int at[3]) { return ((long long)at[2] * s->n[1] + at[1]) * s->n[0] + at[0]; };
    const double steps[3] = {s->g.dx, s->g.dy, s->g.dz};
    const double area = steps[b] * steps[c];
    int bfrom, bto, cfrom, cto;
    flux_range(s, F, 0, bfrom, bto);
    flux_range(s, F, 1, cfrom, cto);
    const int comps[2] = {b, c}, hs[2] = {c, b};
    const double signs[2] = {0.5, -0.5};
    for (int q = 0; q < 2; q++) {
        const int comp = comps[q], h = hs[q];
        for (int v = cfrom; v < cto; v++)
            for (int u = bfrom; u < bto; u++) {
                const double weight = flux_weight(F, 0, u, comp == b) * flux_weight(F, 1, v, comp == c);
                if (weight == 0.0) continue;
                const double kappa = scale * weight * signs[q] * 0.5 * area;
                int at[3];
                at[a] = F.plane;
                at[b] = u;
                at[c] = v;
                const C64 hv = value(PXFDTD_H, h, at), low = value(PXFDTD_E, comp, at);
                const long long h_at = index(at);
                at[a] = F.plane + 1;
                const C64 high = value(PXFDTD_E, comp, at);
                const double k2 = 0.5 * kappa;
                out.push_back({PXFDTD_E, comp, h_at, {k2 * hv.re, k2 * -hv.im}});
                out.push_back({PXFDTD_E, comp, index(at), {k2 * hv.re, k2 * -hv.im}});
                const C64 sum = {low.re + high.re, low.im + high.im};
                out.push_back({PXFDTD_H, h, h_at, {k2 * sum.re, k2 * -sum.im}});
            }
    }
}

}  // namespace

// Simulation::term_derivatives (adjoint.rs:623-680): a term's frequency and the objective's
// derivative with respect to each transform it reads
static int term_derivatives(pxfdtd_simulation *s, const pxfdtd_term &t, const std::vector<double> &all, double &freq,
                            std::vector<Derivative> &out) {
    out.clear();
    if (t.kind == 0) {
        if (t.monitor < 0 || t.monitor >= (int32_t)s->modemons.size() || t.mode < 0 ||
            t.mode >= (int32_t)s->modemons[t.monitor].modes.size() || t.direction < 0 || t.direction > 1)
            return px_fail(PXFDTD_ERR_INVALID, "the objective's mode term reads no mode monitor here");
        const ModeEntry &e = s->modemons[t.monitor].modes[t.mode];
        const pxfdtd_port_mode mode = e.view();
        int bl[3] = {0, 0, 0}, bn[3] = {s->n[0], s->n[1], s->n[2]};
        bl[e.axis] = e.plane;
        bn[e.axis] = 2;
        const long long vol = (long long)bn[0] * bn[1] * bn[2];
        std::vector<double> w((size_t)(6 * vol));
        TRY(pxfdtd_simulation_mode_weights(s, &mode, t.direction, w.data()));
        const C64 g = {t.derivative_re, t.derivative_im};
        for (int c = 0; c < 3; c++)
            for (long long v = 0; v < vol; v++) {
                const C64 wv = {w[2 * (c * vol + v)], w[2 * (c * vol + v) + 1]};
                if (wv.re == 0.0 && wv.im == 0.0) continue;
                const long long i = v % bn[0] + bl[0], j = (v / bn[0]) % bn[1] + bl[1], k = v / ((long long)bn[0] * bn[1]) + bl[2];
                out.push_back({PXFDTD_E, c, (k * s->n[1] + j) * s->n[0] + i, cmul(g, wv)});
            }
        freq = e.frequency;
        return PXFDTD_OK;
    }
    if (t.kind == 1) {
        if (t.monitor < 0 || t.monitor >= (int32_t)s->fluxes.size())
            return px_fail(PXFDTD_ERR_INVALID, "the objective's flux term reads no flux monitor here");
        const FluxRec &R = s->fluxes[t.monitor];
        const DftRec &d0 = s->dfts[R.faces[0].dft];
        if (t.frequency < 0 || t.frequency >= (int32_t)d0.freqs.size())
            return px_fail(PXFDTD_ERR_INVALID, "the objective's flux term names no frequency of its monitor");
        freq = d0.freqs[t.frequency];
        for (const FluxFace &F : R.faces) {
            const double *series0 = all.data() + dft_offset(s, F.dft);
            flux_derivatives(s, F, series0, s->dfts[F.dft].acc.size() / 6, t.frequency, F.sign * t.derivative_re, out);
        }
        return PXFDTD_OK;
    }
    if (t.kind == 2) {
        if (t.monitor < 0 || t.monitor >= (int32_t)s->shown.size() || t.field < 0 || t.field > 1 || t.component < 0 ||
            t.component > 2)
            return px_fail(PXFDTD_ERR_INVALID, "the objective's transform term reads no transform monitor here");
        const DftRec &d = s->dfts[s->shown[t.monitor]];
        if (t.frequency < 0 || t.frequency >= (int32_t)d.freqs.size())
            return px_fail(PXFDTD_ERR_INVALID, "the objective's transform term names no frequency of its monitor");
        for (int a = 0; a < 3; a++)
            if (t.at[a] < d.lo[a] || t.at[a] >= d.hi[a])
                return px_fail(PXFDTD_ERR_INVALID, "the objective's transform term reads no value of its monitor");
        freq = d.freqs[t.frequency];
        out.push_back({t.field, t.component, ((long long)t.at[2] * s->n[1] + t.at[1]) * s->n[0] + t.at[0],
                       {t.derivative_re, t.derivative_im}});
        return PXFDTD_OK;
    }
    return px_fail(PXFDTD_ERR_INVALID, "no term kind %d", t.kind);
}

// the transform monitors of design `R`'s values on `t` (its at-rest copy's among them)
static int add_design_transforms(pxfdtd_simulation *t, DesignRec &R) {
    for (int c = 0; c < 3; c++) {
        R.dft[c] = -1;
        int32_t lo[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, hi[3] = {-1, -1, -1};
        for (const Touched &v : R.touched)
            if (v.comp == c)
                for (int a = 0; a < 3; a++) {
                    lo[a] = std::min(lo[a], v.at[a]);
                    hi[a] = std::max(hi[a], v.at[a]);
                }
        if (hi[0] < 0) continue;
        TRY(add_transform(t, lo, hi, (int32_t)R.freqs.size(), R.freqs.data(), &R.dft[c]));
    }
    return PXFDTD_OK;
}

// Simulation::at_rest (adjoint.rs:720-785): the same problem at rest from the start, the grid,
// media and boundaries, no sources, probes or monitors but design monitor `keep`, at zero
static int sim_at_rest(const pxfdtd_simulation *s, int keep, pxfdtd_simulation **out) {
    std::unique_ptr<pxfdtd_simulation> t(new pxfdtd_simulation());
    t->g = s->g;
    t->b = s->b;
    for (int a = 0; a < 3; a++) t->n[a] = s->n[a];
    t->N = s->N;
    t->dt = s->dt;
    t->steps = 0;
    t->e.assign(3 * s->N, 0.0);
    t->h.assign(3 * s->N, 0.0);
    t->ca = s->ca;
    t->cb = s->cb;
    t->eps.reset(new double[(size_t)(3 * s->N)]);
    std::memcpy(t->eps.get(), s->eps.get(), (size_t)(3 * s->N) * sizeof(double));
    t->sigma = s->sigma;
    for (int a = 0; a < 3; a++) {
        t->kn[a] = s->kn[a];
        t->kh[a] = s->kh[a];
        t->fe[a] = s->fe[a];
        t->fh[a] = s->fh[a];
        for (int side = 0; side < 2; side++)
            for (int f = 0; f < 2; f++) {
                t->cpml_b[a][side][f] = s->cpml_b[a][side][f];
                t->cpml_a[a][side][f] = s->cpml_a[a][side][f];
            }
    }
    t->slabs = s->slabs;
    for (Slab &sl : t->slabs) std::fill(sl.psi.begin(), sl.psi.end(), 0.0);
    t->groups = s->groups;
    t->member = s->member;
    t->dev = s->dev;
    t->precision = s->precision;
    for (size_t m = 0; m < s->designs.size(); m++) {
        DesignRec R;
        R.d = s->designs[m].d;
        R.freqs = s->designs[m].freqs;
        R.dft[0] = R.dft[1] = R.dft[2] = -1;
        if ((int)m == keep) {
            R.touched = s->designs[m].touched;
            TRY(add_design_transforms(t.get(), R));
        }
        t->designs.push_back(std::move(R));
    }
    *out = t.release();
    return PXFDTD_OK;
}

// Simulation::with_design (adjoint.rs:366-423); every value is computed and checked before any
// changes
extern "C" int pxfdtd_simulation_with_design(pxfdtd_simulation *s, const pxfdtd_design *d, const double *density,
                                             int64_t n, const pxfdtd_scene *background) {
    TRY(design_check(s, d));
    TRY(design_check_density(*d, density, n));
    if (!background) return px_fail(PXFDTD_ERR_INVALID, "a design needs the background's scene");
    if (s->steps > 0) return px_fail(PXFDTD_ERR_INVALID, "a design is set before the run starts");
    const long long N = s->N;
    const double h[3] = {s->g.dx, s->g.dy, s->g.dz};
    const std::vector<Touched> touched = design_values(s, *d);
    std::vector<double> values(touched.size(), 0.0);
    for (size_t q = 0; q < touched.size(); q++) {
        const Touched &t = touched[q];
        const long long cr = t.comp * N + t.r;
        if (s->cb[cr] == 0.0) continue;  // held at zero: a wall or a conductor
        if (s->ca[cr] != 1.0 || (!s->member.empty() && s->member[cr] >= 0))
            return px_fail(PXFDTD_ERR_INVALID, "a design's values must be free of conductivity and dispersive media");
        double centre[3];
        s->e_position(t.comp, t.r, centre);
        const int comp = t.comp;
        const double value = averaged_with(
            [&](double x, double y, double z) {
                const double p[3] = {x, y, z};
                const long long v = design_cell_at(s, *d, p);
                return v >= 0 ? design_eps(*d, density[v]) : scene_value(background, comp, x, y, z);
            },
            comp, centre, h);
        if (!(std::isfinite(value) && value > 0.0))
            return px_fail(PXFDTD_ERR_INVALID, "the permittivity must be finite and positive, got %g", value);
        values[q] = value;
    }
    for (size_t q = 0; q < touched.size(); q++) {
        const long long cr = touched[q].comp * N + touched[q].r;
        if (s->cb[cr] == 0.0) continue;
        s->cb[cr] = s->dt / values[q];
        s->eps[cr] = values[q];
    }
    if (s->res) drop_resident(s);  // the device's coefficients are the old ones
    return PXFDTD_OK;
}

// Simulation::add_design_monitor (adjoint.rs:435-470)
extern "C" int pxfdtd_simulation_add_design_monitor(pxfdtd_simulation *s, const pxfdtd_design *d, int32_t nf,
                                                    const double *freqs, int32_t *number) {
    TRY(check_frequencies(s, nf, freqs));
    TRY(design_check(s, d));
    DesignRec R;
    R.d = *d;
    R.touched = design_values(s, *d);
    R.freqs.assign(freqs, freqs + nf);
    TRY(add_design_transforms(s, R));
    s->designs.push_back(std::move(R));
    *number = (int32_t)s->designs.size() - 1;
    return PXFDTD_OK;
}

extern "C" int64_t pxfdtd_simulation_design_values(const pxfdtd_simulation *s, int32_t n, int32_t *comp_ijk) {
    if (n < 0 || n >= (int32_t)s->designs.size()) {
        px_fail(PXFDTD_ERR_INVALID, "no design monitor %d", n);
        return -1;
    }
    const std::vector<Touched> &T = s->designs[n].touched;
    if (comp_ijk)
        for (size_t q = 0; q < T.size(); q++) {
            comp_ijk[4 * q] = T[q].comp;
            for (int a = 0; a < 3; a++) comp_ijk[4 * q + 1 + a] = T[q].at[a];
        }
    return (int64_t)T.size();
}

extern "C" int pxfdtd_simulation_design_transform(pxfdtd_simulation *s, int32_t n, int32_t f, double *out) {
    if (n < 0 || n >= (int32_t)s->designs.size()) return px_fail(PXFDTD_ERR_INVALID, "no design monitor %d", n);
    const DesignRec &R = s->designs[n];
    if (f < 0 || f >= (int32_t)R.freqs.size()) return px_fail(PXFDTD_ERR_INVALID, "no frequency %d", f);
    std::vector<double> all;
    TRY(all_sums(s, all));
    for (size_t q = 0; q < R.touched.size(); q++) {
        const C64 v = design_value(s, R, all, R.touched[q], f);
        out[2 * q] = v.re;
        out[2 * q + 1] = v.im;
    }
    return PXFDTD_OK;
}

// Simulation::adjoint (adjoint.rs:496-569)
extern "C" int pxfdtd_simulation_adjoint(pxfdtd_simulation *s, const pxfdtd_term *terms, int32_t n, int32_t monitor,
                                         double bandwidth, pxfdtd_simulation **out) {
    if (monitor < 0 || monitor >= (int32_t)s->designs.size())
        return px_fail(PXFDTD_ERR_INVALID, "there is no design monitor %d", monitor);
    if (n < 0 || (n > 0 && !terms)) return px_fail(PXFDTD_ERR_INVALID, "the adjoint needs its terms");
    const std::vector<double> freqs = s->designs[monitor].freqs;
    const size_t count = freqs.size();
    std::vector<double> all;
    TRY(all_sums(s, all));
    // each value's transform targets, frequency by frequency, keyed (field, component, index)
    std::map<std::tuple<int, int, long long>, std::vector<C64>> targets;
    std::vector<Derivative> ders;
    for (int32_t q = 0; q < n; q++) {
        double frequency;
        TRY(term_derivatives(s, terms[q], all, frequency, ders));
        int f = -1;
        for (size_t m = 0; m < count; m++)
            if (freqs[m] == frequency) {
                f = (int)m;
                break;
            }
        if (f < 0)
            return px_fail(PXFDTD_ERR_INVALID, "the objective's frequency %.17g isn't among design monitor %d's",
                           frequency, monitor);
        const double omega = TAU * frequency;
        for (const Derivative &dv : ders) {
            const int at[3] = {(int)(dv.r % s->n[0]), (int)((dv.r / s->n[0]) % s->n[1]),
                               (int)(dv.r / ((long long)s->n[0] * s->n[1]))};
            const C64 d = symmetrizer(s, dv.field, dv.comp, at, omega);
            // the adjoint's sources: J = -g_E, M = g_H, over D
            const C64 value = dv.field == PXFDTD_E ? cdiv({-dv.g.re, -dv.g.im}, d) : cdiv(dv.g, d);
            auto it = targets.find({dv.field, dv.comp, dv.r});
            if (it == targets.end())
                it = targets.emplace(std::make_tuple(dv.field, dv.comp, dv.r), std::vector<C64>(count, C64{0.0, 0.0}))
                         .first;
            it->second[f] = {it->second[f].re + value.re, it->second[f].im + value.im};
        }
    }
    std::vector<pxfdtd_waveform> waves;
    TRY(adjoint_basis(freqs, bandwidth, waves));
    pxfdtd_simulation *adj = nullptr;
    TRY(sim_at_rest(s, monitor, &adj));
    std::unique_ptr<pxfdtd_simulation, void (*)(pxfdtd_simulation *)> guard(adj, pxfdtd_simulation_destroy);
    for (int field : {PXFDTD_E, PXFDTD_H}) {
        std::vector<std::tuple<int, long long>> keys;
        std::vector<const std::vector<C64> *> mine;
        for (const auto &kv : targets)
            if (std::get<0>(kv.first) == field) {
                keys.push_back({std::get<1>(kv.first), std::get<2>(kv.first)});
                mine.push_back(&kv.second);
            }
        if (mine.empty()) continue;
        std::vector<std::vector<double>> amps;
        TRY(adjoint_amplitudes(s, field, waves, freqs, mine, amps));
        for (size_t k = 0; k < waves.size(); k++) {
            CurrentRec c;
            c.field = field;
            c.w = waves[k];
            for (size_t v = 0; v < keys.size(); v++) {
                if (amps[v][k] == 0.0) continue;
                c.comps.push_back(std::get<0>(keys[v]));
                c.r.push_back(std::get<1>(keys[v]));
                c.amps.push_back({amps[v][k], 0.0});
            }
            if (!c.r.empty()) adj->currents.push_back(std::move(c));
        }
    }
    adj->inj_dirty = true;
    *out = guard.release();
    return PXFDTD_OK;
}

// design_pair (adjoint.rs:881-902): design monitor `monitor` of both runs, the same in both
static int design_pair(const pxfdtd_simulation *s, const pxfdtd_simulation *adj, int32_t monitor) {
    bool same = monitor >= 0 && monitor < (int32_t)s->designs.size() && monitor < (int32_t)adj->designs.size();
    if (same) {
        const DesignRec &F = s->designs[monitor], &B = adj->designs[monitor];
        same = std::memcmp(F.d.lo, B.d.lo, sizeof F.d.lo) == 0 && std::memcmp(F.d.hi, B.d.hi, sizeof F.d.hi) == 0 &&
               F.d.eps[0] == B.d.eps[0] && F.d.eps[1] == B.d.eps[1] && F.freqs == B.freqs &&
               F.touched.size() == B.touched.size();
        for (size_t q = 0; same && q < F.touched.size(); q++) {
            const Touched &a = F.touched[q], &b = B.touched[q];
            same = a.comp == b.comp && a.r == b.r && a.shares == b.shares;
        }
    }
    if (!same)
        return px_fail(PXFDTD_ERR_INVALID, "design monitor %d must be the same in the forward and the adjoint runs",
                       monitor);
    return PXFDTD_OK;
}

// Simulation::permittivity_gradient (adjoint.rs:831-855)
extern "C" int pxfdtd_simulation_permittivity_gradient(pxfdtd_simulation *s, pxfdtd_simulation *adj, int32_t monitor,
                                                       double *out) {
    TRY(design_pair(s, adj, monitor));
    const DesignRec &F = s->designs[monitor], &B = adj->designs[monitor];
    std::vector<double> fa, ba;
    TRY(all_sums(s, fa));
    TRY(all_sums(adj, ba));
    std::vector<double> tilde;
    for (double f : F.freqs) {  // Simulation::leapfrog_wavenumber (src/fdtd/sources.rs:171-173)
        double sn, cs;
        pxfdtd::psincos(TAU * f * s->dt / 2.0, sn, cs);
        tilde.push_back(2.0 / s->dt * sn);
    }
    for (size_t q = 0; q < F.touched.size(); q++) {
        double sum = -0.0;  // a sum of f64s, as Rust's Sum folds them
        for (size_t f = 0; f < tilde.size(); f++) {
            const C64 b = design_value(adj, B, ba, B.touched[q], (int)f);
            const C64 e = design_value(s, F, fa, F.touched[q], (int)f);
            sum += 2.0 * cmul(cmul({0.0, tilde[f]}, b), e).re;
        }
        out[q] = sum;
    }
    return PXFDTD_OK;
}

// Simulation::gradient (adjoint.rs:857-879)
extern "C" int pxfdtd_simulation_design_gradient(pxfdtd_simulation *s, pxfdtd_simulation *adj, int32_t monitor,
                                                 double *out) {
    TRY(design_pair(s, adj, monitor));
    const DesignRec &R = s->designs[monitor];
    std::vector<double> values(R.touched.size());
    TRY(pxfdtd_simulation_permittivity_gradient(s, adj, monitor, values.data()));
    const double scale = R.d.eps[1] - R.d.eps[0];
    std::fill(out, out + design_len(R.d), 0.0);
    for (size_t q = 0; q < R.touched.size(); q++) {
        const double g = values[q];
        for (const auto &vs : R.touched[q].shares) out[vs.first] += g * vs.second * scale;
    }
    return PXFDTD_OK;
}

// Simulation::run_until_converged (adjoint.rs:904-964), over every series of every transform
// monitor the simulation holds
extern "C" int pxfdtd_simulation_run_until_converged(pxfdtd_simulation *s, double tolerance, double interval,
                                                     double limit, int32_t *settled) {
    if (!(std::isfinite(tolerance) && tolerance > 0.0) || !(std::isfinite(interval) && interval > 0.0) ||
        !std::isfinite(limit))
        return px_fail(PXFDTD_ERR_INVALID,
                       "running until the transforms settle needs a positive tolerance and interval and a finite "
                       "limit, got %g, %g and %g", tolerance, interval, limit);
    if (s->dfts.empty()) return px_fail(PXFDTD_ERR_INVALID, "running until the transforms settle needs a monitor");
    *settled = 0;
    double finished;
    if (!sources_end(s, finished)) finished = HUGE_VAL;
    std::vector<double> before, now;
    TRY(all_sums(s, before));
    while ((double)s->steps * s->dt < limit) {
        TRY(pxfdtd_simulation_run_until(s, std::min((double)s->steps * s->dt + interval, limit)));
        TRY(all_sums(s, now));
        bool all = true;
        size_t at = 0;
        for (const DftRec &d : s->dfts) {
            const size_t len = d.acc.size() / 6;
            for (int t = 0; t < 6; t++, at += len) {
                double largest = 0.0, change = 0.0;
                for (size_t q = 0; q < len; q += 2) {
                    largest = std::max(largest, std::hypot(now[at + q], now[at + q + 1]));
                    change = std::max(change, std::hypot(now[at + q] - before[at + q], now[at + q + 1] - before[at + q + 1]));
                }
                all = all && change <= tolerance * largest;
            }
        }
        if ((double)s->steps * s->dt >= finished && all) {
            *settled = 1;
            return PXFDTD_OK;
        }
        before.swap(now);
    }
    return PXFDTD_OK;
}
