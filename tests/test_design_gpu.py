"""The frequency-domain adjoint on the device (GPU): the reference's gradient by its own method
(src/fdtd/adjoint.rs: a forward run and an adjoint run, each until its transforms settle, then
2 Re(i w~ E_adj E) over the design's values) against the reverse-time adjoint and against
central differences, in FP64.

A 30 x 26 x 22 grid of 50 nm with a 6-cell CPML, a Si strip in SiO2 crossing a 5 x 5 x 3 design
block at random densities, a pulsed E_z point source at 1.55 um:
- the objective |E_z(p)|^2 at a point (a transform term): the frequency-domain gradient per
  value is the reverse-time adjoint's (the exact gradient of the same run) to the runs'
  settling tolerance, and both runs' J are the same; the product's gradient is the oracle
  restatement's formula on the device's transforms;
- the flux through a whole plane (a flux term, its values reaching into the CPMLs, where the
  adjoint's sources are divided by the symmetrizer D): the density gradient matches central
  differences at design cells;
- with a Drude + Lorentz block beside the design, whose ADE the adjoint run steps as the
  forward does, the density gradient matches central differences.
"""
import math
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import design_adjoint as DA  # noqa: E402

pytestmark = pytest.mark.gpu

EPS_SI, EPS_OX = 3.48 ** 2, 1.444 ** 2
INF = float("inf")
F0 = 1 / 1.55
TOL, INTERVAL, LIMIT = 1e-9, 2.0, 600.0


@pytest.fixture(scope="module")
def gpu():
    import photonoxide_amd as P
    return P.Gpu(0)


def build(rho, objective="point"):
    import photonoxide_amd as P
    g = P.Grid3d(30, 26, 22, 0.05, 0.05, 0.05, -0.75, -0.65, -0.55)
    shapes = [P.Box((-INF, -0.11, -0.07), (INF, 0.11, 0.07), EPS_SI)]
    s = P.Simulation.new(g, P.Scene(EPS_OX, shapes), P.Boundaries.cpml_all(6), 0.9)
    design = P.Design(((12, 17), (10, 15), (9, 12)), (EPS_OX, EPS_SI))
    s.with_design(design, rho, P.Scene(EPS_OX, shapes))
    s.add_source(P.Source(P.Field.E, P.Axis.Z, (8, 13, 11), P.Waveform.pulse(F0, 0.15)))
    if objective == "point":
        mon = s.add_dft((21, 13, 10), (21, 13, 10), [F0])
    else:
        mon = s.add_flux(P.Axis.X, 20, [F0])
    return s, design, mon


def objective_of(s, mon, objective):
    import photonoxide_amd as P
    if objective == "point":
        return abs(s.dft(mon, P.Field.E, P.Axis.Z)[0, 0, 0, 0]) ** 2
    return float(s.flux(mon)[0])


def test_frequency_domain_gradient_is_the_reverse_time_adjoints(gpu):
    import photonoxide_amd as P
    from photonoxide_amd.adjoint import Adjoint
    print("case design 5x5x3 in 30x26x22, point objective, FP64", flush=True)
    rho = np.random.default_rng(11).random((3, 5, 5))
    s, design, mon = build(rho)
    dm = s.add_design_monitor(design, [F0])
    s.set_device(gpu, P.Precision.F64)
    assert s.run_until_converged(TOL, INTERVAL, LIMIT)
    steps = s.steps()
    S = s.dft(mon, P.Field.E, P.Axis.Z)[0, 0, 0, 0]
    adj = s.adjoint([P.Term.transform(mon, P.Field.E, P.Axis.Z, (21, 13, 10), 0, np.conj(S))], dm, 0.15)
    assert adj.run_until_converged(TOL, INTERVAL, LIMIT)
    print(f"forward {steps} steps, adjoint {adj.steps()} steps, |S|^2 {abs(S) ** 2:.6e}", flush=True)
    g_eps = s.permittivity_gradient(adj, dm)
    g_rho = s.gradient(adj, dm, design)
    # the product's formula is the oracle restatement's on the device's transforms
    want = DA.permittivity_gradient(s.design_transform(dm)[None, :], adj.design_transform(dm)[None, :], [F0], s.dt())
    assert np.max(np.abs(g_eps - want)) <= 1e-13 * np.max(np.abs(want))
    values = DA.design_values(s.grid, (False, False, False), design.cells)
    assert [(c, at) for c, at, _ in values] == [(int(v[0]), tuple(int(x) for x in v[1:])) for v in s.design_values(dm)]
    want_rho = DA.density_gradient(values, g_eps, design.eps, len(design)).reshape(design.shape)
    assert np.max(np.abs(g_rho - want_rho)) <= 1e-13 * np.max(np.abs(want_rho))
    # the reverse-time adjoint of the same run, same objective
    s2, _, mon2 = build(rho)
    s2.set_device(gpu, P.Precision.F64)
    nx, ny = s2.grid.nx, s2.grid.ny
    dv = [(int(c), (int(k) * ny + int(j)) * nx + int(i)) for c, i, j, k in s.design_values(dm)]
    rt = Adjoint(s2, mon2, 2, 0, np.ones((1, 1, 1), dtype=complex), dv, checkpoint=64)
    j_rt = rt.forward(steps)
    g_rt = rt.grad_eps(rt.backward())
    assert abs(j_rt - abs(S) ** 2) <= 1e-12 * abs(S) ** 2, (j_rt, abs(S) ** 2)
    scale = np.max(np.abs(g_rt))
    err = np.max(np.abs(g_eps - g_rt)) / scale
    print(f"frequency-domain vs reverse-time dF/deps over {len(dv)} values: max {err:.2e} of the largest",
          flush=True)
    assert err < 1e-8, err


def test_flux_term_gradient_matches_central_differences(gpu):
    import photonoxide_amd as P
    print("case design 5x5x3 in 30x26x22, flux objective through a whole plane, FP64", flush=True)
    rho = np.random.default_rng(12).random((3, 5, 5))
    s, design, mon = build(rho, "flux")
    dm = s.add_design_monitor(design, [F0])
    s.set_device(gpu, P.Precision.F64)
    assert s.run_until_converged(TOL, INTERVAL, LIMIT)
    steps = s.steps()
    flux = objective_of(s, mon, "flux")
    adj = s.adjoint([P.Term.flux(mon, 0, 1.0)], dm, 0.15)
    assert adj.run_until_converged(TOL, INTERVAL, LIMIT)
    g_rho = s.gradient(adj, dm, design)
    print(f"forward {steps} steps, flux {flux:.6e}", flush=True)
    d = 1e-3
    worst = 0.0
    for cell in [(1, 2, 2), (0, 1, 3), (2, 4, 0)]:
        vals = []
        for sign in (1, -1):
            r = rho.copy()
            r[cell] += sign * d
            t, _, m = build(r, "flux")
            t.set_device(gpu, P.Precision.F64)
            t.run(steps)
            vals.append(objective_of(t, m, "flux"))
        fd = (vals[0] - vals[1]) / (2 * d)
        rel = abs(g_rho[cell] - fd) / np.max(np.abs(g_rho))
        worst = max(worst, rel)
        print(f"cell {cell}: adjoint {g_rho[cell]:.9e}, central difference {fd:.9e}, {rel:.2e} of the largest",
              flush=True)
    assert worst < 1e-6, worst


def test_gradient_with_a_dispersive_medium_matches_central_differences(gpu):
    """The frequency-domain adjoint with a Drude + Lorentz block beside the design: its adjoint
    run steps the same medium (the ADE terms sit on K's diagonal blocks, symmetric,
    src/fdtd/adjoint.rs:9-12), and the density gradient matches central differences."""
    import photonoxide_amd as P
    print("case design beside a Drude + Lorentz block, FP64", flush=True)

    def with_media(rho):
        s, design, mon = build(rho)
        box = P.Scene(0.0, [P.Box((0.2, -0.25, -0.15), (0.45, 0.25, 0.15), 1.0)])
        s.with_medium(2.0, [("drude", 1.2, 0.05), ("lorentz", 0.8, 1.1, 0.02)], box)
        return s, design, mon

    rho = np.random.default_rng(13).random((3, 5, 5))
    s, design, mon = with_media(rho)
    dm = s.add_design_monitor(design, [F0])
    s.set_device(gpu, P.Precision.F64)
    assert s.run_until_converged(TOL, INTERVAL, LIMIT)
    steps = s.steps()
    S = s.dft(mon, P.Field.E, P.Axis.Z)[0, 0, 0, 0]
    adj = s.adjoint([P.Term.transform(mon, P.Field.E, P.Axis.Z, (21, 13, 10), 0, np.conj(S))], dm, 0.15)
    assert adj.run_until_converged(TOL, INTERVAL, LIMIT)
    g_rho = s.gradient(adj, dm, design)
    d = 1e-3
    worst = 0.0
    for cell in [(1, 2, 2), (2, 0, 4)]:
        vals = []
        for sign in (1, -1):
            r = rho.copy()
            r[cell] += sign * d
            t, _, m = with_media(r)
            t.set_device(gpu, P.Precision.F64)
            t.run(steps)
            vals.append(abs(t.dft(m, P.Field.E, P.Axis.Z)[0, 0, 0, 0]) ** 2)
        fd = (vals[0] - vals[1]) / (2 * d)
        rel = abs(g_rho[cell] - fd) / np.max(np.abs(g_rho))
        worst = max(worst, rel)
        print(f"cell {cell}: adjoint {g_rho[cell]:.9e}, central difference {fd:.9e}, {rel:.2e} of the largest",
              flush=True)
    assert worst < 1e-5, worst
