"""Harmonic inversion (harmonic_inversion, src/fdtd/harmonic.rs:62-184) on the host mirror, held
to the reference's own tests (src/fdtd/harmonic_tests.rs) and its cavity check
(src/fdtd/monitors_checks.rs:316-393, monitors_tests.rs:32-36).

CPU (synthetic signals, closed forms):
- two terms 0.01 c/um apart, closer than the transform's resolution 0.025, and a third outside
  the window: the two recovered with frequency to 1e-11, decay to 1e-10, amplitude to 1e-9 and
  error below 1e-9 (harmonic_tests.rs:68-95);
- a real damped cosine A e^(-gt) cos(wt - phi): one term, frequency and decay to 1e-11,
  amplitude A e^(i phi) / 2 to 1e-9, Q = w / 2g to 1e-7 (:97-124);
- the refusals (:126-132).

GPU: a slab of eps 4, 0.5 um thick, in vacuum along z (one cell across, periodic; 1 um of
vacuum and a 30-cell CPML each side, Courant 0.5), E_x kicked inside it and recorded until it
decays below 1e-28 of its peak: its second resonance against the continuum's, w = (2 pi + i ln
r) / (n L) (1.0 c/um, Q = 2.86), converges at second order (the errors of Re w and of the decay
rate fall 4 times, to 0.1, from 20 to 10 nm cells) to below 1e-3 and 1e-2
(monitors_checks.rs:395-484, monitors_tests.rs:38-45).

GPU: a lossy 2D cavity (24 x 18 cells of 50 nm between walls, eps 2, sigma 0.05, E_z kicked
by a pulse), its probe's 4000 steps after the pulse: at least 4 resonances with error below
1e-8, each e^(-i w dt) within 1e-11 of a root of the leapfrog's own characteristic equation
u^2 - (1 + ca - cb dt lambda) u + ca = 0 for the walled grid's eigenvalues lambda.
"""
import math

import numpy as np
import pytest


def signal(terms, tau, n):
    t = np.arange(n) * tau
    return sum(d * np.exp(-decay * t - 2j * np.pi * f * t) for f, decay, d in terms)


def test_two_terms_closer_than_the_fourier_resolution_are_recovered():
    import photonoxide_amd as P
    tau, n = 0.1, 400
    terms = [(1.0, 0.01, 1.0 + 0.5j), (1.01, 0.03, -0.4 + 0.8j), (1.6, 0.002, 2.0)]
    found = [r for r in P.harmonic_inversion(signal(terms, tau, n), tau, 0.9, 1.1) if r.error < 1e-6]
    print(found)
    assert len(found) == 2
    for r, (f, decay, d) in zip(found, terms):
        assert abs(r.frequency - f) < 1e-11, r
        assert abs(r.decay - decay) < 1e-10, r
        assert abs(r.amplitude - d) < 1e-9 * abs(d), r
        assert r.error < 1e-9, r


def test_a_real_signals_term_is_half_its_amplitude_and_q_is_omega_over_twice_the_decay():
    import photonoxide_amd as P
    tau, n = 0.05, 600
    amplitude, f, decay, phase = 3.0, 0.8, 0.004, 0.7
    t = np.arange(n) * tau
    real = amplitude * np.exp(-decay * t) * np.cos(2 * np.pi * f * t - phase)
    found = [r for r in P.harmonic_inversion(real.astype(complex), tau, 0.5, 1.2) if r.error < 1e-6]
    print(found)
    assert len(found) == 1
    r = found[0]
    assert abs(r.frequency - f) < 1e-11
    assert abs(r.decay - decay) < 1e-11
    assert abs(r.amplitude - 0.5 * amplitude * np.exp(1j * phase)) < 1e-9
    q = 2 * np.pi * f / (2 * decay)
    assert abs(r.q() - q) < 1e-7 * q


def test_harmonic_inversion_refuses_what_the_reference_refuses():
    import photonoxide_amd as P
    s = np.ones(100, complex)
    for args in [(s, 0.1, 1.0, 5.0), (s, 0.1, 2.0, 1.0), (s[:5], 0.1, 1.0, 2.0), (s, 0.0, 1.0, 2.0),
                 (np.array([1.0, np.nan] * 10, complex), 0.1, 1.0, 2.0)]:
        with pytest.raises(Exception):
            P.harmonic_inversion(*args)


@pytest.mark.gpu
def test_a_lossy_cavitys_resonances_are_the_leapfrogs_own():
    import photonoxide_amd as P
    print("case lossy cavity resonances", flush=True)
    nx, ny, h = 24, 18, 0.05
    g = P.Grid3d(nx, ny, 1, h, h, h, -nx * h / 2, -ny * h / 2, -h / 2)
    eps, sigma = 2.0, 0.05
    s = P.Simulation.new(g, P.Scene(eps), P.Boundaries(P.Pml(0, 0), P.Pml(0, 0), P.Periodic()), 0.9)
    s.with_conductivity(P.Scene(sigma))
    pulse = P.Waveform.pulse(0.8, 0.8)
    s.add_source(P.Source(P.Field.E, P.Axis.Z, (7, 5, 0), pulse))
    p = s.add_probe(P.Field.E, P.Axis.Z, (16, 11, 0))
    s.set_device(P.Gpu(0), P.Precision.F64)
    s.run_until(pulse.delay + 6.0 * pulse.width)
    start = len(s.probe(p))
    s.run(4000)
    dt = s.dt()
    found = P.harmonic_inversion(s.probe(p)[start:].astype(complex), dt, 0.4, 1.2)
    cb = dt / eps / (1.0 + sigma * dt / eps / 2.0)
    ca = (1.0 - sigma * dt / eps / 2.0) / (1.0 + sigma * dt / eps / 2.0)
    exact = []
    for m in range(1, nx):
        for n in range(1, ny):
            lam = 4.0 / (h * h) * (math.sin(math.pi * m / (2 * nx)) ** 2 + math.sin(math.pi * n / (2 * ny)) ** 2)
            b = 1.0 + ca - cb * dt * lam
            root = np.sqrt(complex(b * b - 4.0 * ca, 0.0))
            u = 0.5 * (b - root)
            exact.append(u if u.imag < 0 else 0.5 * (b + root))
    exact = np.array(exact)
    kept = [r for r in found if r.error < 1e-8]
    worst = 0.0
    for r in kept:
        u = np.exp(-1j * complex(2 * np.pi * r.frequency, -r.decay) * dt)
        worst = max(worst, np.abs(exact - u).min())
    print(f"{len(found)} found, {len(kept)} kept, worst distance to the leapfrog's roots {worst:.2e}")
    assert len(kept) >= 4
    assert worst < 1e-11


def slab_resonance(h):
    """monitors_checks.rs:403-484: the relative errors of Re w and of the decay rate."""
    import photonoxide_amd as P
    n_index, thickness, gap, cpml = 2.0, 0.5, 1.0, 30
    cells = int(round((thickness + 2.0 * gap) / h)) + 2 * cpml
    g = P.Grid3d(1, 1, cells, h, h, h, 0.0, 0.0, -cells * h / 2)
    centre = 0.5 * h  # the faces halfway between nodes
    s = P.Simulation.new(g, P.Scene(1.0, [P.Box((-np.inf, -np.inf, centre - thickness / 2),
                                                (np.inf, np.inf, centre + thickness / 2), n_index ** 2)]),
                         P.Boundaries(P.Periodic(), P.Periodic(), P.Pml(cpml, cpml)), 0.5)
    pulse = P.Waveform.pulse(1.0, 1.0)
    middle = cells // 2
    s.add_source(P.Source(P.Field.E, P.Axis.X, (0, 0, middle - int(round(0.1 / h))), pulse))
    p = s.add_probe(P.Field.E, P.Axis.X, (0, 0, middle + int(round(0.07 / h))))
    s.set_device(P.Gpu(0), P.Precision.F64)
    s.run_until(pulse.delay + 6.0 * pulse.width)
    start = len(s.probe(p))
    assert s.run_until_decayed([p], 1e-28, 2.0, 200.0), "the slab's field didn't decay"
    found = P.harmonic_inversion(s.probe(p)[start:].astype(complex), s.dt(), 0.75, 1.25)
    r = (n_index - 1.0) / (n_index + 1.0)
    omega = complex(2.0 * np.pi, math.log(r)) / (n_index * thickness)
    best = min((x for x in found if x.error < 1e-6), key=lambda x: abs(x.frequency - 1.0))
    return (abs(2 * np.pi * best.frequency - omega.real) / omega.real,
            abs(best.decay + omega.imag) / -omega.imag)


@pytest.mark.gpu
def test_a_slabs_resonance_and_q_converge_to_the_continuums_at_second_order():
    print("case slab resonance", flush=True)
    coarse, fine = slab_resonance(0.02), slab_resonance(0.01)
    print(f"20 nm: {coarse}, 10 nm: {fine}, ratios {coarse[0] / fine[0]:.3f} {coarse[1] / fine[1]:.3f}")
    for c, f in zip(coarse, fine):
        assert abs(c / f - 4.0) < 0.1
    assert fine[0] < 1e-3 and fine[1] < 1e-2


@pytest.mark.gpu
def test_harmonic_inversion_on_the_device_is_the_hosts_bit_for_bit():
    """A signal held in device memory (pxfdtd_harmonic_inversion_device): the sums over its
    samples run on the device and give every resonance of the host's, bit for bit, on the
    synthetic two-term signal and on a lossy cavity's probe; the refusals are the host's."""
    import photonoxide_amd as P
    from photonoxide_amd.fdtd import harmonic_inversion_device
    from photonoxide_amd._native import NativeError
    from devmem import Buffer
    print("case harmonic inversion on the device", flush=True)
    gpu = P.Gpu(0)
    tau, n = 0.1, 3001
    sig = signal([(0.95, 0.01, 1.0), (0.96, 0.02, 0.5j), (1.3, 0.0, 0.2)], tau, n)
    nx, ny, h = 24, 18, 0.05
    g = P.Grid3d(nx, ny, 1, h, h, h, -nx * h / 2, -ny * h / 2, -h / 2)
    s = P.Simulation.new(g, P.Scene(2.0), P.Boundaries(P.Pml(0, 0), P.Pml(0, 0), P.Periodic()), 0.9)
    s.with_conductivity(P.Scene(0.05))
    pulse = P.Waveform.pulse(0.8, 0.8)
    s.add_source(P.Source(P.Field.E, P.Axis.Z, (7, 5, 0), pulse))
    p = s.add_probe(P.Field.E, P.Axis.Z, (16, 11, 0))
    s.set_device(gpu, P.Precision.F64)
    s.run_until(pulse.delay + 6.0 * pulse.width)
    start = len(s.probe(p))
    s.run(4000)
    cavity = s.probe(p)[start:].astype(complex)
    for x, dt, lo, hi in [(sig, tau, 0.9, 1.1), (cavity, s.dt(), 0.4, 1.2)]:
        host = P.harmonic_inversion(x, dt, lo, hi)
        buf = Buffer.of(np.ascontiguousarray(x).view(np.float64))
        dev = harmonic_inversion_device(gpu, buf.ptr, len(x), dt, lo, hi)
        assert len(dev) == len(host) > 0
        for a, b in zip(dev, host):
            assert (a.frequency, a.decay, a.amplitude, a.error) == (b.frequency, b.decay, b.amplitude, b.error)
    bad = sig.copy()
    bad[17] = np.nan
    buf = Buffer.of(bad.view(np.float64))
    with pytest.raises(NativeError, match="finite samples"):
        harmonic_inversion_device(gpu, buf.ptr, len(bad), tau, 0.9, 1.1)
    with pytest.raises(NativeError, match="below the Nyquist"):
        harmonic_inversion_device(gpu, buf.ptr, len(bad), tau, 1.1, 0.9)
