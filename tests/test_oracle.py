"""The oracle pinned against the reference's golden vectors and against what the mathematics
fixes (CPU only)."""
import math
import os
import struct

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


# src/portable/tests.rs:357-548: the reference's grid of arguments (special values, the
# subnormals' edge, 600 from a fixed LCG over five scales) and its hashes of each portable
# function's bits over it (h = rotl(h, 7) ^ bits, NaN as the canonical quiet NaN), taken on
# x86-64 with libm 0.2.16. Pairs are (grid[i], grid[n - 1 - i]).
GRID_HASHES = {"exp": 0x413FBC04627E9685, "sin_cos.0": 0xDA525217FA01FBF0,
               "sin_cos.1": 0xDA5C8775269B6B93, "powi": 0x0193356F37480A41,
               "c_exp": 0x9FC0EEDECCB4CA6E, "ln": 0x0BF1E9D3216EFC53, "powf": 0x9133A6ECA7E0BB06}


def reference_grid():
    xs = [0.0, -0.0, math.inf, -math.inf, math.nan, 2.2250738585072014e-308, struct.unpack("<d", struct.pack("<Q", 1))[0],
          -struct.unpack("<d", struct.pack("<Q", 3))[0], 1.7976931348623157e308, -1.7976931348623157e308,
          1.0, -1.0, 0.5, 2.0, math.pi, math.pi / 2, 1e22, -709.5, 710.0, 26.5]
    s = 0x0822_2026_1113_0001
    for scale in (1.0, 4.0, 30.0, 800.0, 1e6):
        for _ in range(120):
            s = (s * 6_364_136_223_846_793_005 + 1_442_695_040_888_963_407) % (1 << 64)
            xs.append(((s >> 11) / 9_007_199_254_740_992.0 - 0.5) * 2.0 * scale)
    return xs


def grid_fold(h, v):
    b = 0x7FF8000000000000 if math.isnan(v) else struct.unpack("<Q", struct.pack("<d", v))[0]
    return (((h << 7) | (h >> 57)) & ((1 << 64) - 1)) ^ b


def test_portable_bits_are_the_references_over_its_grid():
    """exp, ln, sin_cos (both parts), powi, powf and complex exp give the reference's bits at every
    argument of its grid, large ones (1e22, the largest double: Payne-Hanek) and special ones
    (inf, NaN, subnormals) included: the hashes are the reference's own."""
    xs = reference_grid()
    pairs = list(zip(xs, reversed(xs)))
    s, c = O.C.c_double(), O.C.c_double()

    def sincos(x):
        O.lib.or_sincos(x, O.C.byref(s), O.C.byref(c))
        return s.value, c.value

    def as_i32(v):  # Rust's saturating `as i32`
        return 0 if math.isnan(v) else int(max(min(v, 2**31 - 1), -2**31))

    got = {"exp": 0, "sin_cos.0": 0, "sin_cos.1": 0, "powi": 0, "c_exp": 0, "ln": 0, "powf": 0}
    for x in xs:
        got["exp"] = grid_fold(got["exp"], O.lib.or_exp(x))
        got["ln"] = grid_fold(got["ln"], O.lib.or_ln(x))
        sv, cv = sincos(x)
        got["sin_cos.0"] = grid_fold(got["sin_cos.0"], sv)
        got["sin_cos.1"] = grid_fold(got["sin_cos.1"], cv)
    for a, b in pairs:
        n = as_i32(b * 7.0)
        got["powi"] = grid_fold(got["powi"], O.lib.or_powi(a, int(math.copysign(abs(n) % 64, n))))
        got["powf"] = grid_fold(got["powf"], O.lib.or_pow(a, b))
        r = O.lib.or_exp(a)  # portable::complex::exp: exp(re) (cos im, sin im), im 0 kept
        if b == 0.0:
            w = (r, b)
        else:
            sv, cv = sincos(b)
            w = (r * cv, r * sv)
        got["c_exp"] = grid_fold(grid_fold(got["c_exp"], w[0]), w[1])
    for name, want in GRID_HASHES.items():
        assert got[name] == want, (name, hex(got[name]), hex(want))


def test_portable_within_the_references_bounds_of_its_golden_values():
    """tests/data/canary/portable.txt (copied as tests/golden/portable_exp_sin_cos.txt): the
    correctly rounded exp, sin and cos at the reference's 48 arguments each (the largest known
    errors of OpenLibm and musl among them, and arguments to 1e300), held to the reference's
    bounds (src/portable/tests.rs:200-215: 0.949, 0.831, 0.834 ulps, so one binary64 step)."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "portable_exp_sin_cos.txt")
    s, c = O.C.c_double(), O.C.c_double()
    seen = {"exp": 0, "sin": 0, "cos": 0}
    for line in open(path):
        if line.startswith("#") or not line.strip():
            continue
        f = line.split()
        name, a, r = f[0], int(f[1], 16), int(f[5], 16)
        x = struct.unpack("<d", struct.pack("<Q", a))[0]
        if name == "exp":
            got = O.lib.or_exp(x)
        else:
            O.lib.or_sincos(x, O.C.byref(s), O.C.byref(c))
            got = s.value if name == "sin" else c.value
        gb = struct.unpack("<q", struct.pack("<d", got))[0]
        wb = struct.unpack("<q", struct.pack("<d", struct.unpack("<d", struct.pack("<Q", r))[0]))[0]
        # binary64 numbers between them (same sign: the ordered integer distance)
        assert (gb < 0) == (wb < 0) or got == 0.0, (name, x, got)
        assert abs(gb - wb) <= 1, (name, x, got, hex(r))
        seen[name] += 1
    assert seen == {"exp": 48, "sin": 48, "cos": 48}


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


def _pulse(bandwidth):
    """Waveform::pulse (src/fdtd.rs:300-312): width sqrt(2 ln 2)/(pi B), delay 6 widths."""
    width = math.sqrt(2 * math.log(2)) / (math.pi * bandwidth)
    return O.wave(0, 1.0, width, 6 * width)


def tfsf_leakage(direction, polarization, three_d):
    """src/fdtd/checks.rs:511-563: the largest field in the scattered region over a pulse's run,
    relative to the largest in the box, vacuum, nothing inside."""
    h = 0.05
    if three_d:
        n, low, high, cp = (36, 36, 36), (12, 12, 12), (24, 24, 24), 8
        b = O.bounds(low=(cp,) * 3, high=(cp,) * 3)
    else:
        n, low, high, cp = (80, 80, 1), (25, 25, 0), (55, 55, 1), 10
        b = O.bounds(periodic=(0, 0, 1), low=(cp, cp, 0), high=(cp, cp, 0))
    grid = (*n, h, h, h, -n[0] * h / 2, -n[1] * h / 2, -n[2] * h / 2)
    s = O.Sim(grid, b, O.scene(1.0), 0.9)
    s.add_plane_wave(low, high, direction, polarization, 1.0, _pulse(0.6))
    i = np.arange(n[0] * n[1] * n[2])
    at = [i % n[0], (i // n[0]) % n[1], i // (n[0] * n[1])]
    inside = {}
    for f in (0, 1):
        for c in range(3):
            m = np.ones(i.size, bool)
            for a in range(3):
                if n[a] == 1:
                    continue
                t = 2 * at[a] + ((c == a) if f == 0 else (c != a))
                m &= (t >= 2 * low[a]) & (t <= 2 * high[a])
            inside[f, c] = m
    leak = largest = 0.0
    for _ in range(int(12.0 / s.dt())):
        s.run(1)
        for (f, c), m in inside.items():
            v = np.abs(s.field(f, c))
            largest = max(largest, v[m].max())
            leak = max(leak, v[~m].max())
    return leak / largest


@pytest.mark.parametrize("direction,polarization,three_d", [
    ((1, 0, 0), (0, 0, 1), False), ((2, 1, 0), (0, 0, 1), False), ((1, -3, 0), (1, 0, 0), False),
    ((0, 0, 1), (1, 0, 0), True), ((2, 1, 1), (1, -1, 0), True)])
def test_plane_waves_box_leaks_only_round_off(direction, polarization, three_d):
    """src/fdtd/tests.rs:252-265: at every angle of the reference's test, < 1e-12 (it measured
    1.3e-15). A wrong sign, weight or sample in any coupling leaks O(1)."""
    leak = tfsf_leakage(direction, polarization, three_d)
    assert leak < 1e-12, leak


def tfsf_leak_2d(plant=None):
    n, h, cp = (80, 80, 1), 0.05, 10
    s = O.Sim((*n, h, h, h, 0, 0, 0), O.bounds(periodic=(0, 0, 1), low=(cp, cp, 0), high=(cp, cp, 0)),
              O.scene(1.0), 0.9)
    s.add_plane_wave((25, 25, 0), (55, 55, 1), (2, 1, 0), (0, 0, 1), 1.0, _pulse(0.6))
    if plant:
        field, factor = plant
        c, r, w, f, q = s.plane_links(0, field)
        # a coupling that carries the incident wave: onto Ez (from Hx, Hy), or from Ez
        m = int(np.flatnonzero(c == 2 if field == 0 else f == 2)[len(c) // 7])
        s.plane_scale_link(0, field, m, factor)
    leak = largest = 0.0
    for _ in range(int(8.0 / s.dt()) // 20):
        s.run(20)
        ez = np.abs(s.field(0, 2)).reshape(80, 80)
        largest = max(largest, ez[26:55, 26:55].max())
        leak = max(leak, ez[:, :24].max())
    return leak / largest


@pytest.mark.parametrize("plant", [(0, 1.001), (1, -1.0)])
def test_plane_wave_leak_check_sees_a_planted_mistake(plant):
    """A coupling's weight off by 1e-3 (onto E) or of the wrong sign (onto H), planted in the
    oracle: the scattered region then holds far more than round-off."""
    assert tfsf_leak_2d() < 1e-12
    assert tfsf_leak_2d(plant) > 1e-6


def test_plane_wave_refuses_what_it_cannot_launch():
    """src/fdtd/tests.rs:350-400: a face in or against the CPML, the box not whole along a
    periodic axis, a direction along it, a box across a block, the polarization along the
    direction, a zero direction."""
    n = (40, 40, 1)
    sc = O.scene(1.0, [(0, (0.4 + 1e-9, -INF, -INF), (INF, INF, INF), 4.0)])
    grid = (*n, 0.05, 0.05, 0.05, -1.0, -1.0, -0.025)
    b = O.bounds(periodic=(0, 0, 1), low=(6, 6, 0), high=(6, 6, 0))
    ok = O.Sim(grid, b, sc, 0.9)
    ok.add_plane_wave((10, 10, 0), (25, 25, 1), (2, 1, 0), (0, 0, 1), 1.0, _pulse(0.5))
    for low, high, d, p in [((5, 10, 0), (25, 25, 1), (1, 0, 0), (0, 0, 1)),
                            ((6, 10, 0), (25, 25, 1), (1, 0, 0), (0, 0, 1)),
                            ((10, 10, 0), (25, 25, 0), (1, 0, 0), (0, 0, 1)),
                            ((10, 10, 0), (25, 25, 1), (1, 0, 1), (0, 0, 1)),
                            ((10, 10, 0), (30, 25, 1), (1, 0, 0), (0, 0, 1)),
                            ((10, 10, 0), (25, 25, 1), (1, 0, 0), (1, 0, 0)),
                            ((10, 10, 0), (25, 25, 1), (0, 0, 0), (0, 0, 1))]:
        s = O.Sim(grid, b, sc, 0.9)
        with pytest.raises(ValueError):
            s.add_plane_wave(low, high, d, p, 1.0, _pulse(0.5))


def tfsf_slab(h):
    """src/fdtd/checks.rs:565-650: a slab of eps 4, 0.3 um thick, in a box at normal incidence;
    the largest difference of |DFT(E_s)/DFT(E_i)|^2 from Airy's formula at 0.8, 1, 1.2."""
    thick, eps_slab = 0.3, 4.0
    cells = lambda x: int(round(x / h))
    cp = cells(0.5)
    nz = 2 * cp + cells(2.4)
    low, high = cp + cells(0.4), nz - cp - cells(0.4)
    za = (low + cells(0.5)) * h
    sc = O.scene(1.0, [(0, (-INF, -INF, za - 1e-12), (INF, INF, za + thick - 1e-12), eps_slab)])
    s = O.Sim((1, 1, nz, h, h, h, 0, 0, 0), O.bounds(periodic=(1, 1, 0), low=(0, 0, cp), high=(0, 0, cp)),
              sc, 0.9)
    s.add_plane_wave((0, 0, low), (1, 1, high), (0, 0, 1), (1, 0, 0), 1.0, _pulse(0.8))
    front, inside = low - cells(0.2), (0, 0, low + 2)
    fr = np.array([0.8, 1.0, 1.2])
    sca, inc = np.zeros(3, complex), np.zeros(3, complex)
    dt = s.dt()
    for n in range(int(30.0 / dt)):
        s.run(1)
        w = np.exp(2j * math.pi * fr * (n + 1) * dt)
        sca += s.field(0, 0)[front] * w
        inc += s.incident(0, 0, 0, inside) * w
    n2 = math.sqrt(eps_slab)
    r12 = (1 - n2) / (1 + n2)
    ph = np.exp(2j * n2 * 2 * math.pi * fr * thick)
    r = r12 * (1 - ph) / (1 - r12 * r12 * ph)
    return float(np.max(np.abs(np.abs(sca / inc) ** 2 - np.abs(r) ** 2)))


def test_plane_wave_on_a_slab_reflects_as_airy_says_to_second_order():
    """src/fdtd/tests.rs:267-276 and src/validation.rs:2848-2856: the error falls by more than 3
    per halving of the cells and is < 1e-3 at 5 nm (the reference measured 5.4e-3, 1.3e-3,
    3.3e-4 at 20, 10, 5 nm)."""
    e = [tfsf_slab(h) for h in (0.02, 0.01, 0.005)]
    assert e[1] < e[0] / 3 and e[2] < e[1] / 3, e
    assert e[2] < 1e-3, e
    assert e == pytest.approx([5.4e-3, 1.3e-3, 3.3e-4], rel=0.05)


def box_balance_problem(inside, plant=None):
    """src/fdtd/monitors_checks.rs:188-245 on the oracle: a walled 16 x 14 x 12 box, lossy
    (sigma 3) but where the flux box's values are, a block of eps 4, an E_y pulse in or out of the
    flux box; the flux box and its six faces as planes, run to t = 200. Returns (net, largest)
    over the two frequencies."""
    import sys, os
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
    import flux as F
    n, h = (16, 14, 12), 0.05
    planes = [(5, 10), (4, 9), (3, 8)]
    x0 = [-m * h / 2 for m in n]
    sig = O.scene(3.0, [(0, tuple(x0[a] + planes[a][0] * h - 2e-9 for a in range(3)),
                         tuple(x0[a] + (planes[a][1] + 1) * h + 2e-9 for a in range(3)), 0.0)])
    s = O.Sim((*n, h, h, h, *x0), O.bounds(), O.scene(1.0, [(0, (-0.08, -0.06, -0.05), (0.08, 0.06, 0.05), 4.0)]), 0.9)
    s.conductivity(sig)
    f0 = 1.0 / 1.55
    width = math.sqrt(2 * math.log(2)) / (math.pi * 0.3)
    s.add_source(0, 1, (8, 6, 5) if inside else (2, 6, 5), O.wave(0, f0, width, 6 * width))
    freqs = [f0, f0 + 0.15]
    faces = F.box_faces(planes)
    if plant:
        plant(faces)
    dfts = []
    for face in faces:
        lo, hi = face.box(n)
        dfts.append((s.add_dft(lo, hi, freqs), lo))
    count = 0
    while count * s.dt() < 200.0:
        count += 1
    s.run(count)

    def value_of(q):
        m, lo = dfts[q]
        series = {(fd, c): s.dft(m, fd, c) for fd in (0, 1) for c in range(3)}
        return lambda fd, c, at, f: series[fd, c][f, at[2] - lo[2], at[1] - lo[1], at[0] - lo[0]]

    values = [value_of(q) for q in range(6)]
    net = max(abs(F.total_flux(faces, n, (h, h, h), values, f)) for f in range(2))
    largest = max(abs(face.flux(n, (h, h, h), values[q], f)) for q, face in enumerate(faces) for f in range(2))
    return net, largest


def test_flux_box_balances_in_a_closed_lossless_box():
    """src/fdtd/monitors_tests.rs:18-25: no net flux leaves a lossless box with no source in it;
    with the source inside, what leaves is what it radiates."""
    net, largest = box_balance_problem(False)
    assert largest > 0 and net < 1e-12 * largest, (net, largest)
    net, largest = box_balance_problem(True)
    assert net > 0.1 * largest, (net, largest)


def test_flux_balance_sees_a_planted_mistake():
    """The balance is sensitive: an edge value's half weight taken whole, or a face's sign
    flipped, planted in the oracle's flux, leaves a net flux far above round-off."""
    def whole_edges(faces):
        for f in faces:
            f.weight = (lambda w, m, half, f=f: 1.0 if f.window[w] is None else
                        (1.0 if f.window[w][0] <= m <= f.window[w][1] and half else
                         (1.0 if f.window[w][0] < m <= f.window[w][1] else 0.0)))

    def flipped(faces):
        faces[2].sign = -faces[2].sign

    for plant in (whole_edges, flipped):
        net, largest = box_balance_problem(False, plant)
        assert net > 1e-6 * largest, (plant.__name__, net, largest)


def _control():
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
    import control
    return control


def test_growth_fires_only_after_enough_rises_to_a_thousand_times():
    """Growth (src/fdtd/monitors.rs:679-717) on hand-made levels."""
    K = _control()
    assert K.Growth.after(10.0, 1.0).needed == 11 and K.Growth.after(0.5, 1.0).needed == 4
    g = K.Growth(4)
    levels = [1.0, 10.0, 100.0, 1000.0, 10000.0]
    out = [g.check(v) for v in levels]
    assert out[:4] == [None] * 4 and out[4] == 1e4  # 4 rises from 1.0, reaching 1e4 >= 1e3
    g = K.Growth(4)
    out = [g.check(v) for v in [1.0, 2.0, 3.0, 4.0, 5.0, 6.0]]
    assert out == [None] * 6  # rising, but not a thousand times
    g = K.Growth(2)
    out = [g.check(v) for v in [1.0, 100.0, 50.0, 5000.0, 6e4]]
    assert out == [None, None, None, None, 6e4 / 50.0]  # the fall restarts the count


def _open_box(cw=False):
    n, h, cp = (40, 36, 1), 0.05, 8
    s = O.Sim((*n, h, h, h, 0, 0, 0), O.bounds(periodic=(0, 0, 1), low=(cp, cp, 0), high=(cp, cp, 0)),
              O.scene(1.0, [(0, (0.8, 0.7, -INF), (1.2, 1.1, INF), 3.0)]), 0.9)
    w = O.wave(2, 1.0, 0.0, 0.0, 1.0, ramp=2.0) if cw else O.wave(0, 1.0, 0.3, 1.8)
    s.add_source(0, 2, (14, 15, 0), w)
    p = [s.add_probe(0, 2, (25, 20, 0)), s.add_probe(1, 0, (20, 25, 0))]
    return s, p, [(2, 0.0, 0.0)] if cw else [(0, 0.3, 1.8)]


def test_run_until_stops_at_the_first_step_past_the_time():
    K = _control()
    s, _, _ = _open_box()
    K.run_until(s, 1.0)
    n = s.steps()
    assert n * s.dt() >= 1.0 and (n - 1) * s.dt() < 1.0
    K.run_until(s, 0.5)  # already past: no step
    assert s.steps() == n


def test_run_until_decayed_on_a_pulse_and_a_continuous_wave():
    """A pulse in an open box decays: the last block's values all stay under the fraction of
    their peak and the block before's do not. A continuous wave never does: the run ends at the
    limit."""
    K = _control()
    s, p, w = _open_box()
    assert K.run_until_decayed(s, p, 1e-4, 0.5, 40.0, w)
    t = s.steps() * s.dt()
    assert t < 40.0
    for n in p:
        v = np.asarray(s.probe(n)) ** 2
        last = int(round(0.5 / s.dt()))
        assert v[-last + 1:].max() <= 1e-4 * v.max()
    s, p, w = _open_box(cw=True)
    assert not K.run_until_decayed(s, p, 1e-4, 1.0, 8.0, w)
    assert s.steps() * s.dt() >= 8.0 and (s.steps() - 1) * s.dt() < 8.0
