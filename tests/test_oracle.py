"""The oracle pinned against the reference's golden vectors and against what the mathematics
fixes (CPU only)."""
import math

import numpy as np
import pytest

import oracle_lib as O
from fixtures import INF, noise

pd = O.pd


def test_canary_fdtd_2d_bit_hash():
    """tests/canary.rs:1007-1062: 200 steps of a 2D strip lit by a Gaussian pulse, hashed over
    Ez, Hx, Hy: the reference's 0x7c8d_b7ae_6998_0888, bit for bit."""
    grid = (48, 40, 1, 0.05, 0.05, 0.05, -1.2, -1.0, -0.025)
    s = O.Sim(grid, O.bounds(periodic=(0, 0, 1)),
              O.scene(1.0, [(0, (-0.3, -0.2, -INF), (INF, 0.2, INF), 6.25)]), 0.9)
    s.add_source(0, 2, (12, 20, 0), O.wave(0, 1.0, 0.6, 1.8))
    s.run(200)
    h = O.fold_hash([s.field(0, 2), s.field(1, 0), s.field(1, 1)])
    peak = np.abs(s.field(0, 2)).max()
    assert 1e-6 < peak < INF
    assert h == 0x7C8DB7AE69980888, hex(h)


def test_canary_detects_a_dropped_term():
    """The hash is sensitive: a source one value off, or one step more, changes it."""
    grid = (48, 40, 1, 0.05, 0.05, 0.05, -1.2, -1.0, -0.025)
    s = O.Sim(grid, O.bounds(periodic=(0, 0, 1)),
              O.scene(1.0, [(0, (-0.3, -0.2, -INF), (INF, 0.2, INF), 6.25)]), 0.9)
    s.add_source(0, 2, (12, 21, 0), O.wave(0, 1.0, 0.6, 1.8))
    s.run(200)
    assert O.fold_hash([s.field(0, 2), s.field(1, 0), s.field(1, 1)]) != 0x7C8DB7AE69980888


def test_portable_exp_and_sincos():
    """fdlibm's exp and musl's sincos: within an ulp of the correctly rounded values, odd and
    even symmetry, sin^2 + cos^2 = 1, and the exact values at 0."""
    s, c = O.C.c_double(), O.C.c_double()
    rng = np.random.default_rng(7)
    xs = np.concatenate([rng.uniform(-700, 700, 2000), rng.uniform(-2, 2, 2000), [0.0, 1e-300, -1e-20]])
    for x in xs:
        e = O.lib.or_exp(x)
        ref = math.exp(x)
        assert abs(e - ref) <= 1.0 * math.ulp(ref) + 1e-300, x
    ys = np.concatenate([rng.uniform(-1e5, 1e5, 2000), rng.uniform(-7, 7, 2000),
                         [math.pi / 2, math.pi, 1.5 * math.pi, 2 * math.pi, 0.0]])
    for x in ys:
        O.lib.or_sincos(x, O.C.byref(s), O.C.byref(c))
        assert abs(s.value - math.sin(x)) <= 1.0 * math.ulp(1.0), x
        assert abs(c.value - math.cos(x)) <= 1.0 * math.ulp(1.0), x
        assert abs(s.value * s.value + c.value * c.value - 1.0) < 4e-16
        s2, c2 = O.C.c_double(), O.C.c_double()
        O.lib.or_sincos(-x, O.C.byref(s2), O.C.byref(c2))
        assert s2.value == -s.value and c2.value == c.value
    O.lib.or_sincos(0.0, O.C.byref(s), O.C.byref(c))
    assert s.value == 0.0 and c.value == 1.0
    # fdlibm's exp(1) is one ulp above e's nearest double, as fdlibm is known to give
    assert O.lib.or_exp(0.0) == 1.0 and O.lib.or_exp(1.0) == 2.7182818284590455


def test_plane_wave_dispersion_is_exact():
    """src/fdtd/checks.rs:35-82: a periodic plane wave E_z = cos(kx x + ky y) evolves as
    cos(n theta) with sin^2(theta/2) = (dt^2/4) sum 4/d^2 sin^2(k d/2), Taflove and Brodwin's
    relation. A wrong difference sign, factor or dt fails it."""
    for (mx, my), courant in [((1, 0), 0.5), ((2, 1), 0.9), ((3, 2), 0.99)]:
        nx, ny, dx, dy = 24, 18, 0.05, 0.04
        s = O.Sim((nx, ny, 1, dx, dy, 0.05, 0, 0, 0), O.bounds(periodic=(1, 1, 1)),
                  O.scene(1.0), courant)
        kx, ky = 2 * math.pi * mx / (nx * dx), 2 * math.pi * my / (ny * dy)
        i, j = np.meshgrid(np.arange(nx), np.arange(ny))
        ez = np.cos(kx * i * dx + ky * j * dy).ravel()
        s.set_field(0, 2, ez)
        at = int(np.argmax(np.abs(ez)))
        p = s.add_probe(0, 2, (at % nx, at // nx, 0))
        s.run(400)
        series = np.concatenate([[ez[at]], s.probe(p)])
        num = np.sum(series[1:-1] * (series[2:] + series[:-2]))
        den = np.sum(2 * series[1:-1] ** 2)
        measured = math.acos(num / den)
        lam = 4 / dx**2 * math.sin(kx * dx / 2) ** 2 + 4 / dy**2 * math.sin(ky * dy / 2) ** 2
        theory = 2 * math.asin(s.dt() * math.sqrt(lam) / 2)
        assert abs(measured - theory) < 1e-10 * theory, (mx, my, measured, theory)


def test_energy_is_conserved_in_a_walled_box():
    """src/fdtd/checks.rs:94-136: 1/2 sum eps E.E + 1/2 sum H(t-dt/2).H(t+dt/2) is the
    leapfrog's invariant; random fields in a walled box with a dielectric block keep it to
    round-off over 300 steps."""
    g = (10, 9, 8, 0.05, 0.05, 0.05, -0.25, -0.225, -0.2)
    s = O.Sim(g, O.bounds(), O.scene(2.0, [(0, (-0.12, -0.1, -INF), (0.12, 0.1, INF), 12.0)]), 0.95)
    n = 10 * 9 * 8
    r = np.arange(n)
    for c in range(3):
        s.set_field(0, c, noise(r + 7 * c))
        s.set_field(1, c, noise(r + 13 * c + 100_000))
    s.run(1)
    vol = 0.05**3

    def energy():
        before = [s.field(1, c) for c in range(3)]
        el = s.energy()
        s.run(1)
        mag = sum(np.dot(b, s.field(1, c)) for c, b in enumerate(before))
        return el + 0.5 * mag * vol

    first = energy()
    drift = max(abs(energy() - first) / first for _ in range(300))
    assert drift < 1e-12, drift


def test_cpml_absorbs_more_when_thicker():
    """src/fdtd/checks.rs:144+: a 2D pulse recorded near a CPML, against the same interior in
    a grid large enough that nothing returns in time: the error falls with the thickness."""
    def run(n, pml, steps=160):
        h = 0.05
        s = O.Sim((n, n, 1, h, h, h, -n * h / 2, -n * h / 2, -h / 2),
                  O.bounds(periodic=(0, 0, 1), low=(pml, pml, 0), high=(pml, pml, 0)),
                  O.scene(1.0), 0.9)
        c = n // 2
        s.add_source(0, 2, (c, c, 0), O.wave(1, 0.0, 0.15, 0.6))
        p = s.add_probe(0, 2, (c + 10, c, 0))
        s.run(steps)
        return s.probe(p)

    big = run(200, 0)
    errs = []
    for pml in (4, 8, 16):
        small = run(40 + 2 * pml, pml)
        errs.append(np.abs(small - big).max() / np.abs(big).max())
    assert errs[0] > errs[1] > errs[2], errs
    assert errs[2] < 2e-3, errs


def test_f32_stays_within_the_reference_growth_law():
    """src/fdtd/kernel_tests.rs:235-258: the same problem in f32 differs from f64 by at most
    2e-8 per step, relative to the largest field."""
    from fixtures import busy_spec, oracle
    spec = busy_spec()
    a = oracle(spec)
    b = oracle(spec)
    b.use_f32()
    steps = 200
    a.run(steps)
    b.run(steps)
    fa, fb = a.fields(), b.fields()
    rel = np.abs(fa - fb).max() / np.abs(fa).max()
    assert 0 < rel <= 2e-8 * steps, rel


def test_dft_equals_the_sum_by_hand():
    """src/fdtd/monitors_checks.rs:35-71: a transform monitor equals sum F(t_n) e^(i w t_n) dt
    over a probe's series at the same value (E at n dt, H at (n - 1/2) dt)."""
    g = (20, 16, 1, 0.05, 0.05, 0.05, -0.5, -0.4, -0.025)
    s = O.Sim(g, O.bounds(periodic=(0, 0, 1), low=(4, 4, 0), high=(4, 4, 0)), O.scene(2.0), 0.7)
    s.add_source(0, 2, (8, 8, 0), O.wave(0, 1.2, 0.5, 1.5))
    at = (11, 9, 0)
    pe = s.add_probe(0, 2, at)
    ph = s.add_probe(1, 1, at)
    d = s.add_dft(at, at, [0.9, 1.2])
    steps = 150
    s.run(steps)
    dt = s.dt()
    te = dt * np.arange(1, steps + 1)
    for f, freq in enumerate([0.9, 1.2]):
        w = 2 * math.pi * freq
        want_e = np.sum(s.probe(pe) * np.exp(1j * w * te) * dt)
        want_h = np.sum(s.probe(ph) * np.exp(1j * w * (te - 0.5 * dt)) * dt)
        assert abs(s.dft(d, 0, 2)[f, 0, 0, 0] - want_e) < 1e-12 * abs(want_e)
        assert abs(s.dft(d, 1, 1)[f, 0, 0, 0] - want_h) < 1e-12 * abs(want_h)


def test_lorentz_medium_reaches_the_leapfrog_permittivity():
    """src/fdtd/media.rs:148-159: a continuous wave in a uniform Lorentz medium settles to the
    plane wave whose wavenumber is the grid's for the leapfrog permittivity
    eps_inf + sum (i cos(w dt/2)/w~) gamma (z - 1/z) / (2 dt (z - alpha - xi/z)), z = e^(-i w dt).
    Checked through the phase velocity along a 1D line: k measured from two probes."""
    n, h = 600, 0.02
    s = O.Sim((n, 1, 1, h, h, h, 0, 0, 0), O.bounds(periodic=(0, 1, 1), low=(40, 0, 0), high=(40, 0, 0)),
              O.scene(1.0), 0.5)
    strength, res, damp = 1.5, 2.0, 0.0
    s.add_medium(2.0, [("lorentz", strength, res, damp)], O.scene(1.0))
    f = 0.8
    s.add_source(0, 1, (60, 0, 0), O.wave(2, f, 0, 0, 1.0, ramp=40.0))
    p1 = s.add_probe(0, 1, (200, 0, 0))
    p2 = s.add_probe(0, 1, (210, 0, 0))
    steps = 5000
    s.run(steps)
    dt = s.dt()
    w = 2 * math.pi * f
    t = dt * np.arange(1, steps + 1)
    tail = slice(steps - 2000, steps)
    a1 = np.sum(s.probe(p1)[tail] * np.exp(1j * w * t[tail]))
    a2 = np.sum(s.probe(p2)[tail] * np.exp(1j * w * t[tail]))
    k_measured = np.angle(a2 / a1) / (10 * h)
    # leapfrog permittivity
    tw = 2 / dt * math.sin(w * dt / 2)
    z = np.exp(-1j * w * dt)
    gam = 2 * math.pi * damp
    w0 = 2 * math.pi * res
    d = 1 + gam * dt / 2
    alpha, xi, gamma = (2 - w0 * w0 * dt * dt) / d, (gam * dt / 2 - 1) / d, strength * w0 * w0 * dt * dt / d
    eps = 2.0 + 1j * math.cos(w * dt / 2) / tw * gamma * (z - 1 / z) / (2 * dt * (z - alpha - xi / z))
    # the grid's wavenumber: sin^2(k h/2)/h^2 = eps sin^2(w dt/2)/dt^2
    k_grid = 2 / h * math.asin(math.sqrt(eps.real) * math.sin(w * dt / 2) * h / dt)
    assert abs(k_measured - k_grid) < 2e-3 * k_grid, (k_measured, k_grid)
