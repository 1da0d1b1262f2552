"""Media admission (CPU): the reference's stability check for a dispersive medium
(Dispersive::amplification and stable, src/fdtd/media.rs:258-309, roots by Aberth's iteration
:457-510), as pxfdtd_medium_amplification and pxfdtd_simulation_add_medium apply it.

Pinned by:
- the non-dispersive closed form: eps (z - 1)^2 + u z has |z| = 1 up to u = 4 eps, and beyond
  it the larger root's modulus (u - 2 eps + sqrt(u (u - 4 eps))) / (2 eps);
- an independent root finder: numpy's companion-matrix eigenvalues of the same characteristic
  polynomial, for random Lorentz and Drude media with conductivity;
- the admission's edges as the reference states them (media.rs:36-40): a Lorentz term's
  w0 dt = 2, and a Courant number of sqrt(eps_inf);
- a planted mistake: the characteristic polynomial with a term's (gamma / 4)(z^2 - 1)^2 part
  left out gives other roots.
"""
import ctypes as C
import math

import numpy as np
import pytest

TAU = 2 * math.pi


def amplification(eps_inf, poles, sigma, dt, u):
    from photonoxide_amd import _native as N
    arr = np.array(poles, dtype=float).reshape(-1) if poles else np.zeros(4)
    out = C.c_double()
    N.check(N.lib.pxfdtd_medium_amplification(eps_inf, len(poles), N.dptr(arr), sigma, dt, u, C.byref(out)))
    return out.value


def terms(poles, dt):
    """Pole::coefficients (media.rs:61-109): alpha, xi, gamma."""
    out = []
    for kind, p1, p2, p3 in poles:
        if kind == 0:
            w0 = TAU * p2
            a, w0sq, g = p1 * w0 * w0, w0 * w0, TAU * p3
        else:
            a, w0sq, g = (TAU * p1) ** 2, 0.0, TAU * p2
        d = 1 + g * dt / 2
        out.append(((2 - w0sq * dt * dt) / d, (g * dt / 2 - 1) / d, a * dt * dt / d))
    return out


def characteristic(eps_inf, poles, sigma, dt, u, drop_coupling=False):
    P = np.polynomial.polynomial
    T = terms(poles, dt)

    def product(skip):
        acc = np.array([1.0])
        for i, (alpha, xi, _) in enumerate(T):
            if i != skip:
                acc = P.polymul(acc, [-xi, -alpha, 1.0])
        return acc
    half = sigma * dt / 2
    poly = P.polymul([eps_inf - half, u - 2 * eps_inf, eps_inf + half], product(-1))
    if not drop_coupling:
        for i, (_, _, g) in enumerate(T):
            poly = P.polyadd(poly, P.polymul([g / 4, 0, -g / 2, 0, g / 4], product(i)))
    return poly


def test_without_poles_the_closed_form():
    for eps in (0.9, 1.0, 2.5, 12.0):
        for frac in (0.01, 0.5, 0.99, 0.999999):
            assert abs(amplification(eps, [], 0.0, 0.01, 4 * eps * frac) - 1.0) < 1e-12
        for frac in (1.01, 1.3, 3.0):
            u = 4 * eps * frac
            want = (u - 2 * eps + math.sqrt(u * (u - 4 * eps))) / (2 * eps)
            assert abs(amplification(eps, [], 0.0, 0.01, u) - want) < 1e-12 * want
        # conductivity damps: every mode shrinks
        assert amplification(eps, [], 3.0, 0.01, 2 * eps) < 1.0


def test_roots_are_numpys_for_random_media():
    rng = np.random.default_rng(3)
    worst = 0.0
    for _ in range(150):
        eps = rng.uniform(1, 8)
        dt = rng.uniform(0.005, 0.03)
        poles = [(0, rng.uniform(0, 3), rng.uniform(1, 0.3 / dt), rng.uniform(0, 3)) for _ in range(rng.integers(1, 4))]
        if rng.random() < 0.5:
            poles.append((1, rng.uniform(1, 10), rng.uniform(0, 1), 0.0))
        sigma = rng.uniform(0, 5)
        u = rng.uniform(0, 4.5 * eps)
        got = amplification(eps, poles, sigma, dt, u)
        want = max(abs(np.polynomial.polynomial.polyroots(characteristic(eps, poles, sigma, dt, u))))
        worst = max(worst, abs(got - want) / want)
    print(f"largest relative difference from numpy's roots: {worst:.2e}")
    assert worst < 1e-6


def test_a_planted_mistake_changes_the_roots():
    # a strongly damped Lorentz term: its modes decay (|z| < 1) only through the coupling
    poles = [(0, 2.0, 6.0, 3.0)]
    eps, dt, u = 2.0, 0.02, 3.0
    good = max(abs(np.polynomial.polynomial.polyroots(characteristic(eps, poles, 0.0, dt, u))))
    bad = max(abs(np.polynomial.polynomial.polyroots(characteristic(eps, poles, 0.0, dt, u, drop_coupling=True))))
    got = amplification(eps, poles, 0.0, dt, u)
    assert abs(got - good) < 1e-9 and abs(got - bad) > 1e-2


def small_simulation(P):
    h = 0.05
    g = P.Grid3d(10, 10, 10, h, h, h, -0.25, -0.25, -0.25)
    return P.Simulation.new(g, P.Scene(1.0), P.Boundaries.walls(), 0.9)


def admitted(P, eps_inf, poles, sigma=None):
    from photonoxide_amd import _native as N
    s = small_simulation(P)
    if sigma is not None:
        s.with_conductivity(P.Scene(sigma))
    try:
        s.with_medium(eps_inf, poles, P.Scene(0.0, [P.Box((-0.1, -0.1, -0.1), (0.1, 0.1, 0.1), 1.0)]))
        return True
    except N.NativeError as e:
        assert "isn't stable" in str(e)
        return False


def test_admission_at_the_references_edges():
    import photonoxide_amd as P
    s = small_simulation(P)
    dt = s.dt()
    w0 = lambda w0dt: w0dt / dt / TAU  # noqa: E731 (a resonance, c/um, at w0 dt)
    # a Lorentz term's w0 dt up to 2
    assert admitted(P, 2.0, [("lorentz", 1.0, w0(1.98), 0.0)])
    assert not admitted(P, 2.0, [("lorentz", 1.0, w0(2.02), 0.0)])
    # the Courant number (0.9) up to sqrt(eps_inf)
    assert admitted(P, 0.82, [("drude", 1.0, 0.1)])
    assert not admitted(P, 0.80, [("drude", 1.0, 0.1)])
    # damping and conductivity do not take a stable medium out
    assert admitted(P, 2.0, [("lorentz", 3.0, w0(1.5), w0(0.5)), ("drude", 5.0, 1.0)], sigma=4.0)
