"""Gaussian beams (Simulation::add_beam, src/fdtd/sources.rs:618-699), held to the reference's
own test of them (src/fdtd/tests.rs:306-347, measured as src/fdtd/checks.rs:845-1030 does):

A 2D beam (w0 = 1.5 um, focus 2 um past its plane, lambda = 1 um, 32 steps a period, 50 nm
cells, vacuum, a 10-cell CPML) launched as a continuous wave; the steady amplitude's widths at 2,
5, 8, 11, 14 um from the plane, fitted w^2 = w0^2 + theta^2 (d - d0)^2, give the waist to 1e-3,
the focus to 0.1 um, the divergence to 1e-4 of the grid's own (its plane waves' spread under
the grid's dispersion) and to 5 % of the paraxial 1/(pi w0); behind the plane the field is
below 1e-5 of the beam's. Tilted by 10 degrees, its centre moves across at the grid's predicted
slope to 1e-3, still one way. These are the reference's bounds.

The host mirror's refusals (CPU): no carrier, a plane across a non-uniform medium, a tilt
towards an axis the field doesn't vary along or along the normal, 90 degrees, a plane in the
CPML.
"""
import math

import numpy as np
import pytest

H = 0.05
SPP = 32  # steps per period


def run_beam(gpu, angle):
    """The reference's checks::beam: |E_z| of the steady state after 32 periods, its transform
    at the carrier over the next 2 periods (checks::steady_amplitude, :646-667), up to a
    constant factor."""
    import photonoxide_amd as P
    n = (10 + 2 + 320 + 10, 460, 1)
    g = P.Grid3d(n[0], n[1], n[2], H, H, H, -n[0] * H / 2, -n[1] * H / 2, -n[2] * H / 2)
    courant = 1.0 / SPP * math.sqrt(2.0) / H
    s = P.Simulation.new(g, P.Scene(1.0), P.Boundaries.cpml_2d(10), courant)
    plane = 12
    s.add_beam(P.GaussianBeam(axis=0, plane=plane, waist=1.5, focus=2.0, centre=(0.0, 0.0), direction=0,
                              tilt=1, angle=angle, polarization=0, eps=1.0),
               P.Waveform.continuous(1.0, ramp=4.0))
    s.set_device(gpu, P.Precision.F64)
    s.run(32 * SPP)
    m = s.add_dft((0, 0, 0), (n[0] - 1, n[1] - 1, 0), [1.0])
    s.run(2 * SPP)
    amp = np.abs(s.dft(m, P.Field.E, P.Axis.Z)[0, 0])  # [j][i]
    y = g.y0 + np.arange(n[1]) * H
    columns = []
    for d in (2.0, 5.0, 8.0, 11.0, 14.0):
        i = plane + int(round(d / H))
        p = amp[:, i] ** 2
        total, first, second = p.sum(), (p * y).sum(), (p * y * y).sum()
        centre = first / total
        columns.append((d, 2.0 * math.sqrt(second / total - centre * centre), centre))
    inner = amp[10:n[1] - 10, 10:n[0] - 10]
    behind = inner[:, :plane - 10].max()
    ahead = inner[:, plane - 10:].max()
    return columns, behind / ahead


def beam_fit(columns):
    """checks::beam_fit: w^2 = c0 + c1 d + c2 d^2 by least squares."""
    a = np.zeros((3, 3))
    rhs = np.zeros(3)
    for d, w, _ in columns:
        pw = np.array([1.0, d, d * d])
        a += np.outer(pw, pw)
        rhs += pw * w * w
    c0, c1, c2 = np.linalg.solve(a, rhs)
    focus = -c1 / (2.0 * c2)
    return math.sqrt(c0 - c2 * focus * focus), focus, math.sqrt(c2)


def simpson_weights(points):
    f = np.ones(points + 1)
    f[1:-1:2] = 4.0
    f[2:-1:2] = 2.0
    return f


def beam_divergence(waist, wavelength, grid=None):
    """checks::beam_divergence: 2 sqrt(<(dk_x/dk_y)^2>) over e^(-k_y^2 w0^2 / 2), k_x from the
    grid's dispersion (or the continuum's), Simpson's rule over +-10/w0."""
    omega = 2 * math.pi / wavelength
    points = 20000
    span = 10.0 / waist
    ky = -span + np.arange(points + 1) * (2.0 * span / points)
    factor = simpson_weights(points)
    weight = np.exp(-ky * ky * waist * waist / 2.0) * factor
    if grid:
        h, dt = grid
        tilde = 2.0 / dt * math.sin(omega * dt / 2.0)
        s = (tilde * h / 2.0) ** 2 - np.sin(ky * h / 2.0) ** 2
        ok = s > 0
        kx = 2.0 / h * np.arcsin(np.sqrt(np.where(ok, s, 0.0)))
        v = np.where(ok, np.sin(ky * h) / np.where(ok, np.sin(kx * h), 1.0), 0.0)
    else:
        s = omega * omega - ky * ky
        ok = s > 0
        v = np.where(ok, ky / np.sqrt(np.where(ok, s, 1.0)), 0.0)
    return 2.0 * math.sqrt((weight * v * v).sum() / weight.sum())


def beam_centre_slope(waist, angle, wavelength, h, dt):
    """checks::beam_centre_slope: <-dk_x/dk_y> over the beam's plane waves on the grid."""
    omega = 2 * math.pi / wavelength
    tilde = 2.0 / dt * math.sin(omega * dt / 2.0)
    points = 20000
    span = min(10.0 / waist, 0.999 * tilde)
    kappa = -span + np.arange(points + 1) * (2.0 * span / points)
    factor = simpson_weights(points)
    phi = angle + np.arcsin(kappa / tilde)
    ky = tilde * np.sin(phi)
    s = (tilde * h / 2.0) ** 2 - np.sin(ky * h / 2.0) ** 2
    ok = (s > 0) & (np.cos(phi) > 0)
    kx = 2.0 / h * np.arcsin(np.sqrt(np.where(ok, s, 0.0)))
    jac = np.cos(phi - angle) / np.cos(phi)
    weight = np.where(ok, np.exp(-kappa * kappa * waist * waist / 2.0) * jac * factor, 0.0)
    val = np.where(ok, np.sin(ky * h) / np.where(ok, np.sin(kx * h), 1.0), 0.0)
    return (weight * val).sum() / weight.sum()


@pytest.mark.gpu
def test_a_gaussian_beam_keeps_its_waist_and_divergence():
    import photonoxide_amd as P
    gpu = P.Gpu(0)
    print("case beam, straight", flush=True)
    columns, backward = run_beam(gpu, 0.0)
    waist, focus, divergence = beam_fit(columns)
    dt = 1.0 / SPP
    on_grid = beam_divergence(1.5, 1.0, (H, dt))
    paraxial = 1.0 / (math.pi * 1.5)
    print(f"beam: waist {waist:.6f}, focus {focus:.4f}, divergence {divergence:.6f} (grid {on_grid:.6f}, "
          f"relative {abs(divergence - on_grid) / on_grid:.2e}, paraxial {paraxial:.6f}), backward {backward:.2e}")
    assert abs(waist - 1.5) < 1e-3, waist
    assert abs(focus - 2.0) < 0.1, focus
    assert abs(divergence - on_grid) < 1e-4 * on_grid, (divergence, on_grid)
    assert abs(divergence - paraxial) < 0.05 * paraxial
    assert backward < 1e-5, backward
    print("case beam, tilted 10 degrees", flush=True)
    columns, backward = run_beam(gpu, math.radians(10.0))
    (d0, _, y0), (d1, _, y1) = columns[0], columns[4]
    slope = (y1 - y0) / (d1 - d0)
    predicted = beam_centre_slope(1.5, math.radians(10.0), 1.0, H, dt)
    print(f"tilted beam: centre's slope {slope:.6f} (predicted {predicted:.6f}, tan 10 degrees "
          f"{math.tan(math.radians(10.0)):.6f}), backward {backward:.2e}")
    assert abs(slope - predicted) < 1e-3 * predicted, (slope, predicted)
    assert backward < 1e-5, backward


def test_host_mirror_refuses_beams_the_reference_refuses():
    """src/fdtd/tests.rs:350-435 (beams) and add_beam's checks (src/fdtd/sources.rs:618-660)."""
    import photonoxide_amd as P
    n = (40, 40, 1)
    g = P.Grid3d(n[0], n[1], n[2], H, H, H, -n[0] * H / 2, -n[1] * H / 2, -n[2] * H / 2)
    s = P.Simulation.new(g, P.Scene(1.0, [P.Box((-0.3, -0.2, -np.inf), (0.3, 0.3, np.inf), 4.0)]),
                         P.Boundaries.cpml_2d(6), 0.9)
    pulse = P.Waveform.pulse(1.0, 0.5)
    beam = P.GaussianBeam(axis=1, plane=18, waist=0.4, focus=0.0, centre=(0.0, -0.5), tilt=2)
    with pytest.raises(Exception):
        s.add_beam(beam, P.Waveform.differentiated_gaussian(0.3, 1.5))  # no carrier
    with pytest.raises(Exception):
        s.add_beam(beam, pulse)  # its plane crosses the block
    along_x = P.GaussianBeam(axis=0, plane=10, waist=0.4, focus=0.0, centre=(0.0, -0.5), tilt=1)
    s.add_beam(along_x, pulse)
    for bad in [
        P.GaussianBeam(axis=0, plane=10, waist=0.4, centre=(0.0, -0.5), tilt=2, angle=0.2),  # z doesn't vary
        P.GaussianBeam(axis=0, plane=10, waist=0.4, centre=(0.0, -0.5), tilt=0),            # along the normal
        P.GaussianBeam(axis=0, plane=10, waist=0.4, centre=(0.0, -0.5), tilt=1, angle=math.pi / 2),
        P.GaussianBeam(axis=0, plane=6, waist=0.4, centre=(0.0, -0.5), tilt=1),             # in the CPML
        P.GaussianBeam(axis=0, plane=10, waist=-0.4, centre=(0.0, -0.5), tilt=1),
        P.GaussianBeam(axis=2, plane=0, waist=0.4, tilt=0),                                 # periodic normal
    ]:
        with pytest.raises(Exception):
            s.add_beam(bad, pulse)
