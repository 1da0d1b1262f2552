"""Designs on the CPU (no GPU call): the host mirror's Design::values, with_design and the
adjoint's pulse basis against the oracle restatement (oracle/design_adjoint.py) and what the
mathematics fixes.

- with_design's permittivity is the oracle's averaging (oracle/fdtd.c, the canary-pinned
  Simulation::new) of a scene whose design cells are boxes at eps(rho): every ca and cb, bit
  for bit;
- the values the design reaches are the oracle's, in the same order, and each share is the
  derivative of that average with respect to the cell's density over eps1 - eps0 (the average
  is linear in the densities across the component);
- the adjoint's pulses' transforms (Waveform::spectrum) are their Fourier integrals in closed
  form, and the amplitudes solved for reproduce their targets.
"""
import math
import os
import sys

import numpy as np
import pytest

import oracle_lib as O

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import design_adjoint as DA  # noqa: E402

EPS_SI, EPS_OX = 3.48 ** 2, 1.444 ** 2
INF = float("inf")


def problem(seed=7, periodic=False):
    import photonoxide_amd as P
    if periodic:
        # the design reaches the last cell along the periodic z: its values wrap onto k = 0
        g = P.Grid3d(22, 20, 8, 0.05, 0.05, 0.05, -0.55, -0.5, -0.2)
        b = P.Boundaries(P.Pml(4, 4), P.Pml(4, 4), P.Periodic())
        cells = ((9, 13), (8, 12), (6, 8))
    else:
        g = P.Grid3d(22, 20, 18, 0.05, 0.05, 0.05, -0.55, -0.5, -0.45)
        b = P.Boundaries.cpml_all(4)
        cells = ((9, 13), (8, 12), (7, 10))
    # a strip crossing part of the design, so the background varies across its values' cells
    shapes = [P.Box((-INF, -0.11, -0.07), (INF, 0.11, 0.07), EPS_SI)]
    design = P.Design(cells, (EPS_OX, EPS_SI))
    rho = np.random.default_rng(seed).random(design.shape)
    return P, g, b, shapes, design, rho


def oracle_box_scene(g, shapes, design, rho, periodic=False):
    """The scene with each design cell a box at eps(rho), after the background's shapes; across a
    periodic z its images one period either side too (the samples beyond the grid's ends see
    the cells wrapped onto it)."""
    boxes = [(0, s.lo, s.hi, s.value) for s in shapes]
    (x0, x1), (y0, y1), (z0, z1) = design.cells
    images = (-g.nz, 0, g.nz) if periodic else (0,)
    for k in range(z0, z1):
        for j in range(y0, y1):
            for i in range(x0, x1):
                r = rho[k - z0, j - y0, i - x0]
                eps = design.eps[0] + r * (design.eps[1] - design.eps[0])
                for shift in images:
                    lo = (g.x0 + i * g.dx, g.y0 + j * g.dy, g.z0 + (k + shift) * g.dz)
                    hi = (g.x0 + (i + 1) * g.dx, g.y0 + (j + 1) * g.dy, g.z0 + (k + shift + 1) * g.dz)
                    boxes.append((0, lo, hi, eps))
    return O.scene(EPS_OX, boxes)


def oracle_bounds(b, periodic):
    return O.bounds(periodic=(0, 0, 1) if periodic else (0, 0, 0), low=(4, 4, 0 if periodic else 4),
                    high=(4, 4, 0 if periodic else 4))


@pytest.mark.parametrize("periodic", [False, True])
def test_with_design_is_the_oracles_average_of_a_scene_of_boxes(periodic):
    P, g, b, shapes, design, rho = problem(periodic=periodic)
    s = P.Simulation.new(g, P.Scene(EPS_OX, shapes), b, 0.9)
    s.with_design(design, rho, P.Scene(EPS_OX, shapes))
    o = O.Sim(tuple(vars(g).values()), oracle_bounds(b, periodic), oracle_box_scene(g, shapes, design, rho, periodic), 0.9)
    assert s.dt() == o.dt()
    changed = 0
    plain = P.Simulation.new(g, P.Scene(EPS_OX, shapes), b, 0.9)
    for c in range(3):
        ca0, cb0 = o.coefficients(c)
        ca1, cb1 = s.coefficients(c)
        assert np.array_equal(ca0.view(np.uint64), ca1.view(np.uint64)), c
        assert np.array_equal(cb0.view(np.uint64), cb1.view(np.uint64)), c
        changed += int(np.sum(cb1 != plain.coefficients(c)[1]))
    assert changed > 100


@pytest.mark.parametrize("periodic", [False, True])
def test_design_values_are_the_oracles_and_shares_are_the_averages_derivatives(periodic):
    P, g, b, shapes, design, rho = problem(periodic=periodic)
    s = P.Simulation.new(g, P.Scene(EPS_OX, shapes), b, 0.9)
    n = s.add_design_monitor(design, [1 / 1.55])
    got = s.design_values(n)
    per = (False, False, periodic)
    want = DA.design_values(g, per, design.cells)
    assert [(int(v[0]), tuple(int(x) for x in v[1:])) for v in got] == [(c, at) for c, at, _ in want]
    # each share is d eps / d rho over eps1 - eps0, from the oracle's averaging of box scenes
    rng = np.random.default_rng(3)
    picks = rng.choice(len(want), size=4, replace=False)
    base = O.Sim(tuple(vars(g).values()), oracle_bounds(b, periodic), oracle_box_scene(g, shapes, design, rho, periodic), 0.9)
    dt = base.dt()
    delta = 0.5
    for q in picks:
        c, (i, j, k), shares = want[q]
        r = (k * g.ny + j) * g.nx + i
        eps0 = dt / base.coefficients(c)[1][r]
        assert abs(sum(shares.values()) - 1.0) < 1e-15 or sum(shares.values()) < 1.0
        for v, share in shares.items():
            bumped = rho.copy().ravel()
            bumped[v] += delta
            o = O.Sim(tuple(vars(g).values()), oracle_bounds(b, periodic),
                      oracle_box_scene(g, shapes, design, bumped.reshape(design.shape), periodic), 0.9)
            eps1 = dt / o.coefficients(c)[1][r]
            slope = (eps1 - eps0) / (delta * (design.eps[1] - design.eps[0]))
            assert abs(slope - share) < 1e-12, (q, v, slope, share)
    # the shares of a value whose cell lies wholly in the design sum to one
    if not periodic:
        inside = [sh for c, at, sh in want
                  if all(design.cells[a][0] + 1 <= at[a] < design.cells[a][1] - 1 for a in range(3))]
        assert inside and all(abs(sum(sh.values()) - 1.0) == 0.0 for sh in inside)


def test_with_design_refuses_what_the_reference_refuses():
    P, g, b, shapes, design, rho = problem()
    from photonoxide_amd._native import NativeError
    s = P.Simulation.new(g, P.Scene(EPS_OX, shapes), b, 0.9)
    bg = P.Scene(EPS_OX, shapes)
    with pytest.raises(NativeError, match="CPML"):
        s.with_design(P.Design(((2, 8), (8, 12), (7, 10)), design.eps), np.zeros((3, 4, 6)), bg)
    with pytest.raises(NativeError, match="densities"):
        s.with_design(design, np.zeros(7), bg)
    with pytest.raises(NativeError, match="positive permittivity"):
        s.with_design(design, -np.ones(design.shape), bg)
    with pytest.raises(NativeError, match="finite and positive"):
        s.with_design(P.Design(design.cells, (0.0, 1.0)), rho, bg)
    # nothing was changed by the refusals
    plain = P.Simulation.new(g, P.Scene(EPS_OX, shapes), b, 0.9)
    for c in range(3):
        assert np.array_equal(s.coefficients(c)[1], plain.coefficients(c)[1])


def closed_form_transform(w, omega):
    """The Fourier integral of exp(-((t - d)/W)^2) sin(w0 (t - d)) times e^(i omega t)."""
    f0, width, delay = w
    w0 = 2.0 * math.pi * f0
    a = math.exp(-((omega + w0) * width) ** 2 / 4.0) - math.exp(-((omega - w0) * width) ** 2 / 4.0)
    return complex(np.exp(1j * omega * delay) * width * math.sqrt(math.pi) / 2j * a)


def test_pulse_transforms_are_their_fourier_integrals_and_amplitudes_reproduce_targets():
    dt = 0.05 / math.sqrt(3) * 0.9
    freqs = [1 / 1.60, 1 / 1.55, 1 / 1.50]
    waves = DA.basis(freqs, 0.02)
    steps = max(math.ceil((d + 8.0 * wd) / dt) + 1 for (_, wd, d) in waves)
    for field in (0, 1):
        for w in waves:
            for f in freqs:
                got = DA.spectrum(w, field, dt, steps, f)
                want = closed_form_transform(w, 2.0 * math.pi * f)
                assert abs(got - want) < 1e-12 * abs(want) + 1e-13, (field, w, f, got, want)
    rng = np.random.default_rng(1)
    targets = [rng.normal(size=3) + 1j * rng.normal(size=3) for _ in range(4)]
    amps, steps = DA.amplitudes(0, waves, freqs, targets, dt)
    for t, a in zip(targets, amps):
        for j, f in enumerate(freqs):
            got = sum(a[m] * closed_form_transform(w, 2.0 * math.pi * f) for m, w in enumerate(waves))
            assert abs(got - t[j]) < 1e-9 * np.max(np.abs(t)), (got, t[j])
