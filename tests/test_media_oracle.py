"""The oracle's dispersive media (the ADE of src/fdtd/media.rs:101-109, 592-705, restated in
oracle/fdtd.c and oracle/step_impl.inc) held to the reference's own media checks
(src/fdtd/media_tests.rs:9-57, src/fdtd/media_checks.rs), with damped Lorentz and Drude terms:

- the pole coefficients: the leapfrog's permittivity built from the oracle's alpha, xi, gamma
  (media.rs:148-159) is the medium's own eps(w) = eps_inf + sum a / (w0^2 - w^2 - i G w) to second
  order in the step, error ratios 4 from dt = 0.02 to 0.01 to 0.005 (media_tests.rs:17-32);
- the update: a continuous current in a closed box filled with a Drude + damped Lorentz medium
  and a conductivity settles to FDFD's field at the leapfrog's frequency with that permittivity,
  to round-off (media_checks.rs:29-91, bound 1e-10), and not with the medium's own eps(w~);
- the physics of a Drude half-space: its reflection against Fresnel's (1 - n)/(1 + n) with the
  metal's eps(w), converging at second order (media_checks.rs:170-232, media_tests.rs:34-43).

Each check is shown to fail on a mistake planted in the oracle (or_media_plant): the damping
term of xi with its sign flipped, alpha without its denominator, the damping doubled, chi
halved. CPU only: the oracle is the checker, nothing of the product runs here.
"""
import math
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import oracle_lib as O  # noqa: E402
import modemonitor as MM  # noqa: E402

MM.use_portable(O.lib.or_ln, O.lib.or_pow)
TAU = 2 * math.pi

# media_checks.rs:12-27: eps_inf 2, a Drude term (fp 1.2 c/um, collisions 0.3 c/um) and a damped
# Lorentz term (strength 1.5, f0 0.9 c/um, damping 0.3 c/um)
DRUDE_LORENTZ = (2.0, [("drude", 1.2, 0.3), ("lorentz", 1.5, 0.9, 0.3)])
# media_checks.rs:160-168: the half-space's Drude metal, eps_inf 1, fp 1 c/um, collisions 0.05
DRUDE_METAL = (1.0, [("drude", 1.0, 0.05)])
DRUDE_BAND = [0.4, 0.6, 0.8, 1.0, 1.2, 1.4, 1.6]


@pytest.fixture(autouse=True)
def no_plant():
    O.lib.or_media_plant(0)
    yield
    O.lib.or_media_plant(0)


def pole_array(p):
    return np.array((0.0, p[1], p[2], p[3]) if p[0] == "lorentz" else (1.0, p[1], p[2], 0.0))


def oracle_coefficients(p, dt):
    """alpha, xi, gamma as the oracle computes them for the simulation's own media."""
    out = np.zeros(3)
    a = pole_array(p)
    O.lib.or_pole_coefficients(a.ctypes.data_as(O.pd), dt, out.ctypes.data_as(O.pd))
    return out


def permittivity(medium, f):
    """Dispersive::permittivity (media.rs:89-93, 129-137): the susceptibilities' closed forms,
    a / (w0^2 - w^2 - i G w), Lorentz a = strength w0^2, Drude a = wp^2 and w0 = 0."""
    eps_inf, poles = medium
    w = TAU * f
    eps = complex(eps_inf)
    for p in poles:
        if p[0] == "lorentz":
            w0 = TAU * p[2]
            a, w0sq, g = p[1] * w0 * w0, w0 * w0, TAU * p[3]
        else:
            a, w0sq, g = (TAU * p[1]) ** 2, 0.0, TAU * p[2]
        eps += a / complex(w0sq - w * w, -g * w)
    return eps


def leapfrog_permittivity(medium, f, dt):
    """Dispersive::leapfrog_permittivity (media.rs:148-159) with the oracle's coefficients."""
    eps_inf, poles = medium
    w = TAU * f
    z = complex(math.cos(w * dt), -math.sin(w * dt))
    tilde = 2.0 / dt * math.sin(w * dt / 2.0)
    factor = 1j * math.cos(w * dt / 2.0) / tilde
    eps = complex(eps_inf)
    for p in poles:
        alpha, xi, gamma = oracle_coefficients(p, dt)
        eps += factor * gamma * (z - 1.0 / z) / (2.0 * dt * (z - alpha - xi / z))
    return eps


def second_order_ok(medium):
    for f in (0.3, 0.9, 1.4):
        exact = permittivity(medium, f)
        e = [abs(leapfrog_permittivity(medium, f, dt) - exact) for dt in (0.02, 0.01, 0.005)]
        if not (abs(e[0] / e[1] - 4.0) < 0.1 and abs(e[1] / e[2] - 4.0) < 0.03):
            return False, (f, e)
    return True, None


def test_the_leapfrogs_permittivity_is_the_mediums_to_second_order():
    """media_tests.rs:17-32: the oracle's alpha, xi, gamma make the leapfrog's permittivity the
    medium's own to second order (error ratios 4 +- 0.1 and 4 +- 0.03), at 0.3, 0.9 and 1.4 c/um,
    across the Drude term's and the Lorentz term's damping; a planted sign, denominator or
    damping rate breaks it."""
    ok, info = second_order_ok(DRUDE_LORENTZ)
    assert ok, info
    for plant in (1, 2, 3):
        O.lib.or_media_plant(plant)
        ok, _ = second_order_ok(DRUDE_LORENTZ)
        assert not ok, plant


def dispersive_box(plant=0):
    """media_checks.rs:37-91: 16 x 14 x 12 cells of 50 nm, walls, eps 1, sigma 2/um, the
    Drude + Lorentz medium on every value, a continuous E_y current at the centre (1.55 um, 64
    steps a period, ramp 6 periods), 70 periods, then its steady amplitude over 4 more (the
    transform times 2 / (4 x 64 dt)). Returns the largest difference from FDFD's field (walls,
    A = w~^2 eps - curl curl on the oracle's lattice, b = -i w~ J) with the leapfrog's
    permittivity, and with the medium's own eps(w~), each relative to FDFD's largest value."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    O.lib.or_media_plant(plant)
    h, n, spp, sigma = 0.05, (16, 14, 12), 64, 2.0
    dt0 = 1.55 / spp
    courant = dt0 * math.sqrt(3.0 / (h * h))
    grid = (*n, h, h, h, -n[0] * h / 2, -n[1] * h / 2, -n[2] * h / 2)
    s = O.Sim(grid, O.bounds(), O.scene(1.0), courant)
    s.conductivity(O.scene(sigma))
    s.add_medium(DRUDE_LORENTZ[0], DRUDE_LORENTZ[1], O.scene(1.0))
    O.lib.or_media_plant(0)
    f = 1.0 / 1.55
    at = (n[0] // 2 + 1, n[1] // 2, n[2] // 2)
    s.add_source(0, 1, at, O.wave(2, f, 0, 0, 1.0, ramp=6.0 * 1.55))
    s.run(70 * spp)
    d = s.add_dft((0, 0, 0), (n[0] - 1, n[1] - 1, n[2] - 1), [f])
    s.run(4 * spp)
    dt = s.dt()
    amp = np.concatenate([s.dft(d, 0, c)[0].ravel() for c in range(3)]) * (2.0 / (4 * spp * dt))
    w = TAU * f
    tilde = 2.0 / dt * math.sin(w * dt / 2.0)
    loss = 1j * sigma * math.cos(w * dt / 2.0) / tilde
    lat = MM.Lattice(n, (h,) * 3, grid[3:], (0, 0, 0), (0, 0, 0), (0, 0, 0), 1e-8, 3.0, tilde)
    cells = n[0] * n[1] * n[2]
    rows, cols, base, curl = [], [], [], []
    diag = []
    for r in range(3 * cells):
        if lat.fixed(r):
            diag.append(r)
            continue
        for q, a in lat.curl_curl(r):
            rows.append(r), cols.append(q), curl.append(complex(*a))
    free = np.array([r for r in range(3 * cells) if not lat.fixed(r)])
    b = np.zeros(3 * cells, complex)
    b[lat.index(1, at)] = -1j * tilde

    def solve(eps):
        rr = list(rows) + list(free) + diag
        cc = list(cols) + list(free) + diag
        vv = list(curl) + [tilde * tilde * (eps + loss)] * len(free) + [1.0] * len(diag)
        A = sp.csc_matrix((vv, (rr, cc)), shape=(3 * cells, 3 * cells))
        x = spl.spsolve(A, b)
        return np.abs(x - amp).max() / np.abs(x).max()
    exact = solve(leapfrog_permittivity(DRUDE_LORENTZ, f, dt))
    own = solve(permittivity(DRUDE_LORENTZ, tilde / TAU))
    return exact, own


def test_a_dispersive_boxs_steady_state_is_fdfds_field_with_the_leapfrogs_permittivity():
    """media_tests.rs:8-15: to round-off (bound 1e-10) with the leapfrog's permittivity, and at
    least 1e3 times further with the medium's own eps(w~) (the ADE's O(dt^2)). chi halved in the
    oracle (plant 4, the update's E^(n-1) weight) breaks the first."""
    exact, own = dispersive_box()
    print(f"dispersive box: {exact:.2e} with the leapfrog's permittivity, {own:.2e} with eps(w~)")
    assert exact < 1e-10, exact
    assert own > 1e3 * exact, own
    planted, _ = dispersive_box(plant=4)
    print(f"with chi halved: {planted:.2e}")
    assert planted > 1e-4, planted


def drude_reflection(h, plant=0):
    """media_checks.rs:170-232: a column one cell across (Bloch k = 0 on x and y), CPMLs of
    0.5 um at its ends, Courant 0.9; a pulse (1 c/um, bandwidth 1.2) on E_x 0.5 um in, a probe
    1.5 um in, the metal beyond a face 2 um in (midway between two E_x values); the scattered
    field's transform over the incident one's at DRUDE_BAND, referred to the face with the
    grid's vacuum wavenumber, against Fresnel. The largest |r - (1 - n)/(1 + n)|."""
    cells = lambda x: int(round(x / h))  # noqa: E731
    cpml = cells(0.5)
    nz = 2 * cpml + cells(4.0)
    face_node = cpml + cells(2.0)
    face = face_node * h - h / 2.0
    source, probe = cpml + cells(0.5), cpml + cells(1.5)
    width = math.sqrt(2 * math.log(2)) / (math.pi * 1.2)
    pulse = O.wave(0, 1.0, width, 6 * width)

    def run(with_metal):
        s = O.Sim((1, 1, nz, h, h, h, 0.0, 0.0, 0.0),
                  O.bounds(periodic=(1, 1, 0), low=(0, 0, cpml), high=(0, 0, cpml)), O.scene(1.0), 0.9)
        if with_metal:
            O.lib.or_media_plant(plant)
            inside = O.scene(0.0, [(0, (-1e9, -1e9, face), (1e9, 1e9, 1e9), 1.0)])
            s.add_medium(DRUDE_METAL[0], DRUDE_METAL[1], inside)
            O.lib.or_media_plant(0)
        s.add_source(0, 0, (0, 0, source), pulse)
        p = s.add_probe(0, 0, (0, 0, probe))
        dt = s.dt()
        count = 0
        while count * dt < 150.0:
            count += 1
        s.run(count)
        return s.probe(p), dt
    full, dt = run(True)
    empty, _ = run(False)
    t = dt * np.arange(1, len(full) + 1)
    distance = face - probe * h
    worst = 0.0
    for f in DRUDE_BAND:
        w = TAU * f
        ph = np.exp(1j * w * t)
        sc = np.sum((full - empty) * ph)
        inc = np.sum(empty * ph)
        k = 2.0 / h * math.asin(math.sin(w * dt / 2.0) * h / dt)
        r = sc / inc * np.exp(-2j * k * distance)
        nn = np.sqrt(permittivity(DRUDE_METAL, f))
        worst = max(worst, abs(r - (1 - nn) / (1 + nn)))
    return worst


def test_a_drude_half_space_reflects_as_fresnel_says_to_second_order():
    """media_tests.rs:34-43: the error falls by more than 3.5 from 20 nm to 10 nm cells and is
    below 2e-3 at 10 nm; with the damping term of xi's sign flipped (plant 1) or the damping
    doubled (plant 3) it is not."""
    e = [drude_reflection(h) for h in (0.02, 0.01)]
    print(f"Drude half-space: {e}")
    assert e[1] < e[0] / 3.5, e
    assert e[1] < 2e-3, e
    for plant in (1, 3):
        ep = drude_reflection(0.01, plant)
        print(f"plant {plant}: {ep:.3e}")
        assert ep > 2e-3, (plant, ep)
