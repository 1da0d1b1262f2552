"""Mode sources (Simulation::add_mode_source, src/fdtd/sources.rs:550-596) for a mode given by
its profile.

CPU (the oracle restatement, oracle/modesource.py, run on the oracle's stepper):
- a discrete eigenmode of a uniform guide (E across the axis uniform, beta from the grid's own
  dispersion at the leapfrog wavenumber) launched forward or backward goes one way: behind the
  source plane the field is the CPMLs' reflections, 2.2e-6 of the launched one with 32-cell
  CPMLs (1.2e-5 with 16: it follows their thickness, not the source);
- a planted error breaks that: H's sign flipped sends 4 times as much backward as forward,
  beta 1 % off 5e-3 of it (0.01 % off, 5e-5);
- the host mirror refuses what the reference refuses (the plane's size or step, a periodic
  axis, the plane in a CPML, no carrier, a k0 that is not the leapfrog's).

GPU: the device running the host mirror's mode source is the oracle running the restatement's
currents, bit for bit in f64 and f32, for a 3D mode with every component set (complex, random)
going either way.
"""
import math
import os
import sys

import numpy as np
import pytest

import oracle_lib as O

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import modesource as MS  # noqa: E402


def sincos(x):
    s, c = O.C.c_double(), O.C.c_double()
    O.lib.or_sincos(x, O.C.byref(s), O.C.byref(c))
    return s.value, c.value


def oracle_currents(o, grid, periodic, mode, direction, frequency, **plant):
    """The restatement's currents for the oracle sim `o` (CPML kappa 1: every curl factor is
    1/step)."""
    n = grid[:3]
    steps = grid[3:6]
    cb = [o.coefficients(c)[1] for c in range(3)]
    f = [[1.0 / steps[a]] * n[a] for a in range(3)]
    return MS.mode_currents(n, periodic, cb, f, f, o.dt(), mode, direction, frequency, O.lib.or_exp,
                            sincos, **plant)


def add_to_oracle(o, currents, wave):
    for field, vals in currents:
        o.add_current(field, [v[0] for v in vals], [v[1] for v in vals],
                      [complex(*v[2]) for v in vals], wave)


# ------------------------------------------------------------------- a uniform guide, one way

H_STEP = 0.05
N_MED = 1.5
FREQ = 1.0


def uniform_guide(direction):
    """192 x 2 x 1 cells, periodic across, a 32-cell CPML at each end of x, eps 2.25."""
    grid = (192, 2, 1, H_STEP, H_STEP, H_STEP, 0.0, 0.0, 0.0)
    periodic = (0, 1, 1)
    o = O.Sim(grid, O.bounds(periodic=periodic, low=(32, 0, 0), high=(32, 0, 0)),
              O.scene(N_MED**2), 0.9)
    return o, grid, periodic


def eigenmode(o, p):
    """E_y uniform on the plane; beta from the grid's dispersion, (2/h sin(beta h / 2))^2 =
    eps k~^2, with eps as the scheme has it (dt / cb) and k~ the leapfrog wavenumber."""
    dt = o.dt()
    k0 = 2.0 / dt * math.sin(2.0 * math.pi * FREQ * dt / 2.0)
    eps = dt / o.coefficients(1)[1][5]
    beta = 2.0 / H_STEP * math.asin(H_STEP * math.sqrt(eps) * k0 / 2.0)
    ny, nz = 2, 1
    tangential = [[[(1.0, 0.0)] * ny for _ in range(nz)], [[(0.0, 0.0)] * ny for _ in range(nz)]]
    normal = [[(0.0, 0.0)] * ny for _ in range(nz)]
    return dict(axis=0, plane=p, step=H_STEP, beta=(beta, 0.0), k0=k0,
                tangential=tangential, normal=normal)


def launched(direction, **plant):
    """The largest |E_y| over the last 500 of 5000 steps, 30 cells behind the source plane and
    30 cells ahead."""
    o, grid, periodic = uniform_guide(direction)
    p = 76 if direction == 0 else 116
    mode = eigenmode(o, p)
    wave = O.wave(kind=2, frequency=FREQ, width=0.0, delay=0.0, amp=1.0, ramp=4.0)
    add_to_oracle(o, oracle_currents(o, grid, periodic, mode, direction, FREQ, **plant), wave)
    ahead, behind = (p + 30, p - 30) if direction == 0 else (p - 30, p + 30)
    o.add_probe(0, 1, (ahead, 0, 0))
    o.add_probe(0, 1, (behind, 0, 0))
    o.run(5000)
    return np.abs(o.probe(0)[-500:]).max(), np.abs(o.probe(1)[-500:]).max()


@pytest.mark.parametrize("direction", [0, 1])
def test_eigenmode_goes_one_way(direction):
    ahead, behind = launched(direction)
    print(f"direction {direction}: ahead {ahead:.3e}, behind {behind:.3e}, ratio {behind / ahead:.2e}")
    assert ahead > 0.5
    assert behind < 5e-6 * ahead


@pytest.mark.parametrize("plant,least", [(dict(flip_h=True), 0.5), (dict(beta_scale=1.01), 3e-3)])
def test_planted_errors_send_light_backward(plant, least):
    ahead, behind = launched(0, **plant)
    print(f"{plant}: ahead {ahead:.3e}, behind {behind:.3e}, ratio {behind / ahead:.2e}")
    assert behind > least * ahead


# ------------------------------------------------------------------------- refusals (host)

def product_guide():
    import photonoxide_amd as P
    g = P.Grid3d(160, 2, 1, H_STEP, H_STEP, H_STEP)
    s = P.Simulation.new(g, P.Scene(N_MED**2), P.Boundaries(P.Pml(16, 16), P.Periodic(), P.Periodic()), 0.9)
    return P, s


def product_mode(P, s, p=60, size=(2, 1), step=H_STEP, axis=0, k0=None):
    k0 = s.leapfrog_wavenumber(FREQ) if k0 is None else k0
    return P.PortMode(axis, p, step, k0, 9.0, np.ones((2, size[1], size[0]), complex),
                      np.zeros((size[1], size[0]), complex))


def test_host_mirror_refuses_what_the_reference_refuses():
    """src/fdtd/sources.rs:550-577 and carrier_wavenumber :485-508."""
    P, s = product_guide()
    cw = P.Waveform.continuous(FREQ, ramp=4.0)
    s.add_mode_source(product_mode(P, s), 0, cw)  # a fitting mode is taken
    bad = [
        (product_mode(P, s, size=(3, 1)), cw),            # the plane's size is not the grid's
        (product_mode(P, s, step=0.04), cw),              # nor its step
        (product_mode(P, s, axis=1, size=(1, 160)), cw),  # normal to a periodic axis
        (product_mode(P, s, p=16), cw),                   # the plane in the CPML
        (product_mode(P, s, p=143), cw),                  # the next plane in the CPML
        (product_mode(P, s), P.Waveform.differentiated_gaussian(0.5, 3.0)),  # no carrier
        (product_mode(P, s, k0=s.leapfrog_wavenumber(FREQ) * (1 + 1e-7)), cw),  # k0 not the leapfrog's
    ]
    for mode, w in bad:
        with pytest.raises(Exception):
            s.add_mode_source(mode, 0, w)


# ------------------------------------------------------------------ the device (GPU) parity

def random_mode_problem(direction, precision):
    """28 x 20 x 16 cells, CPMLs on every face (kappa 1), a dielectric block across the
    source planes, a random complex profile in all three components, a Gaussian pulse."""
    import photonoxide_amd as P
    grid = (28, 20, 16, H_STEP, H_STEP, H_STEP, 0.0, 0.0, 0.0)
    n = grid[:3]
    periodic = (0, 0, 0)
    block = (0, (-np.inf, 0.3, 0.25), (np.inf, 0.7, 0.55), 4.0)
    o = O.Sim(grid, O.bounds(low=(5, 4, 4), high=(5, 4, 4)), O.scene(1.0, [block]), 0.9)
    g = P.Grid3d(*grid)
    s = P.Simulation.new(g, P.Scene(1.0, [P.Box(block[1], block[2], block[3])]),
                         P.Boundaries(P.Pml(5, 5), P.Pml(4, 4), P.Pml(4, 4)), 0.9)
    rng = np.random.default_rng(11 + direction)
    tan = rng.standard_normal((2, n[1], n[2])) + 1j * rng.standard_normal((2, n[1], n[2]))
    nor = rng.standard_normal((n[1], n[2])) + 1j * rng.standard_normal((n[1], n[2]))
    # mode along x: b = y, c = z; arrays [c][b]
    tan = np.ascontiguousarray(np.transpose(tan, (0, 2, 1)))
    nor = np.ascontiguousarray(nor.T)
    freq = 1.4
    k0 = s.leapfrog_wavenumber(freq)
    p = 10
    beta = 1.7 * k0 + 0.05j
    mode_o = dict(axis=0, plane=p, step=H_STEP, beta=(beta.real, beta.imag),
                  tangential=[[[(v.real, v.imag) for v in row] for row in comp] for comp in tan],
                  normal=[[(v.real, v.imag) for v in row] for row in nor])
    width = math.sqrt(2.0 * math.log(2.0)) / (math.pi * 0.5)
    wave_o = O.wave(kind=0, frequency=freq, width=width, delay=6.0 * width)
    add_to_oracle(o, oracle_currents(o, grid, periodic, mode_o, direction, freq), wave_o)
    s.add_mode_source(P.PortMode(0, p, H_STEP, k0, beta, tan, nor), direction,
                      P.Waveform(0, freq, width, 6.0 * width))
    for at in [(p + 5, 10, 8), (p - 4, 10, 8)]:
        o.add_probe(0, 1, at)
        s.add_probe(P.Field.E, P.Axis.Y, at)
    if precision == "f32":
        o.use_f32()
    return o, s


@pytest.mark.gpu
@pytest.mark.parametrize("direction", [0, 1])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_device_mode_source_is_the_oracle_bit_for_bit(direction, prec):
    import photonoxide_amd as P
    from fixtures import product_fields
    print(f"case mode source direction {direction} {prec}", flush=True)
    o, s = random_mode_problem(direction, prec)
    s.set_device(P.Gpu(0), P.Precision.F64 if prec == "f64" else P.Precision.F32)
    o.run(150)
    s.run(150)
    fo, fp = o.fields(), product_fields(s)
    assert np.abs(fo).max() > 0
    assert np.array_equal(fo.view(np.uint64), fp.view(np.uint64)), np.abs(fo - fp).max()
    for q in range(2):
        assert np.array_equal(o.probe(q).view(np.uint64), s.probe(q).view(np.uint64))
