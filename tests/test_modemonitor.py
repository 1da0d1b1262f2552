"""Mode monitors (Simulation::add_mode_monitor and mode_amplitudes, src/fdtd/monitors.rs:569-663)
for modes given by their profiles: FDFD's Lorentz form (src/fdfd/three/ports.rs:253-287) of the
FP64 transform of E across a mode's plane and the next.

CPU (the oracle restatement, oracle/modemonitor.py):
- for a discrete plane wave in a lossless periodic guide (beta from the grid's dispersion), the
  form N(f, g) of the mode going forward and backward is the scheme's flux of f, the flux
  monitor's formula (E's mean either side times H~ = curl E / (i k0)), to round-off; N(f, f)
  and N(g, g) are exactly zero, so the mode's own field reads back one way to the bit;
- planted errors break that: the cut's antisymmetry flipped, k0 1 % off;
- the PML stretch the form grades with, times k0, is the CPML's sigma at the same nodes and
  half-nodes (recovered from the oracle's CPML coefficients, which the Roden-Gedney numbers
  pin), to 1e-9 of its largest; a half-cell shift of the half-nodes breaks it;
- the host mirror refuses what the reference refuses.

GPU:
- the host mirror's amplitudes are the oracle restatement's of the same device transforms, bit
  for bit, for random complex modes along two axes in a box with CPMLs on every face (the
  stretches in play), in f64 and f32;
- an eigenmode of a uniform periodic guide launched by the mode source reads back one way: its
  backward amplitude is the CPMLs' reflection (below 1e-5 of the forward), and |forward|^2 /
  N(f, g) is the flux monitor's flux through the same cut to 1e-10 (2.8e-13 measured; the
  backward amplitude 5.3e-7 of the forward).
"""
import ctypes as C
import math
import os
import sys

import numpy as np
import pytest

import oracle_lib as O

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import modemonitor as MM  # noqa: E402

MM.use_portable(O.lib.or_ln, O.lib.or_pow)


def sincos(x):
    s, c = O.C.c_double(), O.C.c_double()
    O.lib.or_sincos(x, O.C.byref(s), O.C.byref(c))
    return s.value, c.value


EXP = O.lib.or_exp
H_STEP = 0.05
N_MED = 1.5


def plane_wave_mode(k0, plane, size=(2, 1)):
    """E_y uniform across a guide along x, beta from the grid's dispersion in a medium of index
    N_MED: (2/h sin(beta h / 2))^2 = eps k0^2."""
    beta = 2.0 / H_STEP * math.asin(H_STEP * N_MED * k0 / 2.0)
    nb, nc = size
    return dict(axis=0, plane=plane, step=H_STEP, beta=(beta, 0.0),
                tangential=[[[(1.0, 0.0)] * nb for _ in range(nc)], [[(0.0, 0.0)] * nb for _ in range(nc)]],
                normal=[[(0.0, 0.0)] * nb for _ in range(nc)])


def own_field(mode, forward):
    return lambda c, at: MM.mode_value(mode, c, at, forward, EXP, sincos)


def scheme_flux(mode, k0, ny, nz):
    """The flux monitor's formula for the mode's field at the half-plane after its plane: 1/2 Re
    (E_y mean conj(H~_z)) dy dz over the plane, H~_z = (E_y(p + 1) - E_y(p)) / (dx i k0)."""
    p = mode["plane"]
    f = own_field(mode, True)
    e0, e1 = f(1, (p, 0, 0)), f(1, (p + 1, 0, 0))
    mean = ((e0[0] + e1[0]) / 2, (e0[1] + e1[1]) / 2)
    hz = MM.cdiv(((e1[0] - e0[0]) / H_STEP, (e1[1] - e0[1]) / H_STEP), (0.0, k0))
    return 0.5 * (mean[0] * hz[0] + mean[1] * hz[1]) * H_STEP * H_STEP * ny * nz


def guide_lattice(k0, n=(16, 2, 1), plant_k0=1.0):
    return MM.Lattice(n, (H_STEP,) * 3, (0.0, 0.0, 0.0), (0, 1, 1), (3, 0, 0), (3, 0, 0), 1e-8, 3.0, k0 * plant_k0)


@pytest.mark.parametrize("n", [(16, 2, 1), (16, 3, 2)])
def test_the_form_is_the_schemes_flux_and_one_way(n):
    k0 = 2.0 * math.pi / 1.55
    mode = plane_wave_mode(k0, 7, size=(n[1], n[2]))
    lat = guide_lattice(k0, n)
    fwd, bwd = MM.mode_amplitudes(lat, mode, own_field(mode, True), EXP, sincos)
    flux = scheme_flux(mode, k0, n[1], n[2])
    print(f"N(f, g) = {fwd}, flux {flux:.17g}, relative {abs(fwd[0] - flux) / flux:.2e}, N(f, f) = {bwd}")
    assert abs(fwd[0] - flux) <= 1e-14 * flux and abs(fwd[1]) <= 1e-14 * flux
    assert bwd[0] == 0.0 and bwd[1] == 0.0
    gf, gb = MM.mode_amplitudes(lat, mode, own_field(mode, False), EXP, sincos)
    assert gf[0] == 0.0 and gf[1] == 0.0
    assert abs(gb[0] - flux) <= 1e-14 * flux and abs(gb[1]) <= 1e-14 * flux


def test_planted_errors_break_the_flux_identity():
    k0 = 2.0 * math.pi / 1.55
    mode = plane_wave_mode(k0, 7)
    flux = scheme_flux(mode, k0, 2, 1)
    _, bwd = MM.mode_amplitudes(guide_lattice(k0), mode, own_field(mode, True), EXP, sincos, plant_sign=-1.0)
    print(f"the cut's antisymmetry flipped: N(f, f) = {bwd}")
    assert abs(complex(*bwd)) > 0.1 * flux
    fwd, _ = MM.mode_amplitudes(guide_lattice(k0, plant_k0=1.01), mode, own_field(mode, True), EXP, sincos)
    print(f"k0 1 % off: N(f, g) / flux = {fwd[0] / flux}")
    assert abs(fwd[0] / flux - 1.0) > 1e-3


@pytest.mark.parametrize("shift", [0.0, 0.5])
def test_the_grading_is_the_cpmls(shift):
    """graded's stretch times k0 against the oracle's CPML sigma = -ln(b) / dt (kappa 1, alpha
    0) along each axis's slabs, E's (field 0) at the nodes and H's at the half-nodes; `shift`
    0.5 plants the half-nodes half a cell off."""
    n, low, high = (20, 14, 12), (5, 3, 4), (4, 6, 3)
    steps = (0.05, 0.04, 0.03)
    o = O.Sim((*n, *steps, 0.0, 0.0, 0.0), O.bounds(low=low, high=high), O.scene(1.0), 0.9)
    dt = o.dt()
    k0 = 3.7
    lat = MM.Lattice(n, steps, (0.0, 0.0, 0.0), (0, 0, 0), low, high, 1e-8, 3.0, k0)
    worst, largest, count = 0.0, 0.0, 0
    for q in range(O.lib.or_sim_slab_count(o.h)):
        info = (C.c_int32 * 5)()
        b, a, ps = C.POINTER(C.c_double)(), C.POINTER(C.c_double)(), C.POINTER(C.c_double)()
        ln = C.c_int64()
        O.lib.or_sim_slab(o.h, q, info, C.byref(b), C.byref(a), C.byref(ps), C.byref(ln))
        field, _, axis, frm, cnt = list(info)
        for m in range(frm, frm + cnt):
            sigma = -math.log(b[m - frm]) / dt
            if field == 0:
                st = lat.nodes[axis][m]
            elif shift == 0.0:
                st = lat.halves[axis][m]
            else:
                pos = m * steps[axis] + (0.5 + shift) * steps[axis]
                st = MM.graded(pos, (0.0, n[axis] * steps[axis]), (low[axis], high[axis]), steps[axis], k0, 1e-8, 3.0)
            worst = max(worst, abs(st[1] * k0 - sigma))
            largest = max(largest, sigma)
            count += 1
    print(f"{count} slab values, largest sigma {largest:.6g}, worst difference {worst:.3e}")
    assert count > 0
    if shift == 0.0:
        assert worst <= 1e-9 * largest
    else:
        assert worst > 1e-2 * largest


def test_host_mirror_refuses_what_the_reference_refuses():
    """src/fdtd/monitors.rs:580-611 and leapfrog_frequency :560-567."""
    import photonoxide_amd as P
    g = P.Grid3d(40, 6, 4, H_STEP, H_STEP, H_STEP)
    s = P.Simulation.new(g, P.Scene(N_MED ** 2), P.Boundaries(P.Pml(8, 8), P.Periodic(), P.Periodic()), 0.9)
    k0 = s.leapfrog_wavenumber(1.0)

    def mode(plane=20, size=(6, 4), step=H_STEP, axis=0, k=k0):
        return P.PortMode(axis, plane, step, k, 9.0, np.ones((2, size[1], size[0]), complex),
                          np.zeros((size[1], size[0]), complex))
    assert s.add_mode_monitor([mode()]) == 0
    assert s.add_mode_monitor([mode(), mode(plane=30)]) == 1
    f = s.mode_frequencies(1)
    assert len(f) == 2 and abs(f[0] - 1.0) < 1e-12
    assert s.mode_amplitudes(1) == [(0j, 0j), (0j, 0j)]
    for bad in [[], [mode(size=(5, 4))], [mode(step=0.04)], [mode(plane=39)], [mode(plane=-1)],
                [mode(k=2.0 / s.dt() * 1.0001)], [mode(), mode(size=(4, 6))]]:
        with pytest.raises(Exception):
            s.add_mode_monitor(bad)
    with pytest.raises(IndexError):
        s.mode_amplitudes(2)


# ------------------------------------------------------------------------------ the device

def random_problem(precision):
    """28 x 20 x 16 cells, CPMLs on every face, a dielectric block, an E_y dipole pulse; a mode
    monitor with two random complex modes (along x on plane 14, along y on plane 9), and the
    caller's transforms over the same planes at the same frequencies."""
    import photonoxide_amd as P
    n = (28, 20, 16)
    g = P.Grid3d(*n, H_STEP, H_STEP, H_STEP)
    s = P.Simulation.new(g, P.Scene(1.0, [P.Box((0.5, 0.3, 0.25), (0.9, 0.7, 0.55), 4.0)]),
                         P.Boundaries(P.Pml(5, 5), P.Pml(4, 4), P.Pml(4, 4)), 0.9)
    width = math.sqrt(2.0 * math.log(2.0)) / (math.pi * 0.5)
    s.add_source(P.Source(P.Field.E, P.Axis.Y, (9, 10, 8), P.Waveform(0, 1.4, width, 6.0 * width)))
    rng = np.random.default_rng(17)
    k0 = s.leapfrog_wavenumber(1.4)
    modes = []
    for axis, plane in ((0, 14), (1, 9)):
        nb, nc = n[(axis + 1) % 3], n[(axis + 2) % 3]
        tan = rng.standard_normal((2, nc, nb)) + 1j * rng.standard_normal((2, nc, nb))
        nor = rng.standard_normal((nc, nb)) + 1j * rng.standard_normal((nc, nb))
        modes.append(P.PortMode(axis, plane, H_STEP, k0, 1.6 * k0 + 0.03j, tan, nor))
    mon = s.add_mode_monitor(modes)
    freqs = s.mode_frequencies(mon)
    boxes = []
    for m, f in zip(modes, freqs):
        lo, hi = [0, 0, 0], [n[0] - 1, n[1] - 1, n[2] - 1]
        lo[m.axis], hi[m.axis] = m.plane, m.plane + 1
        boxes.append((s.add_dft(lo, hi, [f]), lo))
    s.set_device(P.Gpu(0), P.Precision.F64 if precision == "f64" else P.Precision.F32)
    return s, n, modes, mon, boxes


def oracle_mode(m):
    return dict(axis=m.axis, plane=m.plane, step=m.step, beta=(complex(m.beta).real, complex(m.beta).imag),
                tangential=[[[(float(v.real), float(v.imag)) for v in row] for row in comp] for comp in m.tangential],
                normal=[[(float(v.real), float(v.imag)) for v in row] for row in m.normal])


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_device_amplitudes_are_the_oracles_bit_for_bit(prec):
    import photonoxide_amd as P
    print(f"case mode monitor, random modes, {prec}", flush=True)
    s, n, modes, mon, boxes = random_problem(prec)
    s.run(160)
    got = s.mode_amplitudes(mon)
    for m, (d, lo), (fwd, bwd) in zip(modes, boxes, got):
        e = [s.dft(d, P.Field.E, c)[0] for c in range(3)]
        assert np.abs(e[1]).max() > 0

        def field(c, at, e=e, lo=lo):
            v = e[c][at[2] - lo[2], at[1] - lo[1], at[0] - lo[0]]
            return (float(v.real), float(v.imag))
        lat = MM.Lattice(n, (H_STEP,) * 3, (0.0, 0.0, 0.0), (0, 0, 0), (5, 4, 4), (5, 4, 4), 1e-8, 3.0, m.k0)
        of, ob = MM.mode_amplitudes(lat, oracle_mode(m), field, EXP, sincos)
        print(f"axis {m.axis}: forward {fwd} (oracle {complex(*of)}), backward {bwd} (oracle {complex(*ob)})")
        assert abs(fwd) > 0 and abs(bwd) > 0
        assert np.array_equal(np.array([fwd.real, fwd.imag, bwd.real, bwd.imag]).view(np.uint64),
                              np.array([of[0], of[1], ob[0], ob[1]]).view(np.uint64))


@pytest.mark.gpu
def test_an_eigenmode_reads_back_one_way_with_its_flux():
    """A uniform guide along x (192 x 2 x 1 cells, periodic across, 32-cell CPMLs at the ends,
    eps 2.25); its discrete eigenmode launched forward from plane 76 by a Gaussian pulse at
    1 c/um, run until it has left; read on plane 106."""
    import photonoxide_amd as P
    print("case mode monitor, eigenmode one way", flush=True)
    g = P.Grid3d(192, 2, 1, H_STEP, H_STEP, H_STEP)
    s = P.Simulation.new(g, P.Scene(N_MED ** 2), P.Boundaries(P.Pml(32, 32), P.Periodic(), P.Periodic()), 0.9)
    k0 = s.leapfrog_wavenumber(1.0)
    beta = 2.0 / H_STEP * math.asin(H_STEP * N_MED * k0 / 2.0)
    tan = np.zeros((2, 1, 2), complex)
    tan[0] = 1.0
    nor = np.zeros((1, 2), complex)
    width = math.sqrt(2.0 * math.log(2.0)) / (math.pi * 0.3)
    s.add_mode_source(P.PortMode(0, 76, H_STEP, k0, beta, tan, nor), 0, P.Waveform(0, 1.0, width, 6.0 * width))
    read = 106
    mon = s.add_mode_monitor([P.PortMode(0, read, H_STEP, k0, beta, tan, nor)])
    f = s.mode_frequencies(mon)[0]
    flux = s.add_flux(0, read, [f])
    s.set_device(P.Gpu(0), P.Precision.F64)
    steps = int(round((12.0 * width + 60.0) / s.dt()))
    s.run(steps)
    (fwd, bwd), = s.mode_amplitudes(mon)
    mode = plane_wave_mode(k0, read)
    lat = MM.Lattice((192, 2, 1), (H_STEP,) * 3, (0.0, 0.0, 0.0), (0, 1, 1), (32, 0, 0), (32, 0, 0), 1e-8, 3.0, k0)
    nfg = MM.mode_amplitudes(lat, mode, own_field(mode, True), EXP, sincos)[0][0]
    power = s.flux(flux)[0]
    print(f"{steps} steps: forward {fwd}, backward {bwd} ({abs(bwd) / abs(fwd):.2e}), N(f, g) {nfg:.6g}, "
          f"|forward|^2 / N {abs(fwd) ** 2 / nfg:.12g}, flux {power:.12g}, "
          f"relative {abs(abs(fwd) ** 2 / nfg - power) / power:.2e}")
    assert abs(fwd) > 0
    assert abs(bwd) < 1e-5 * abs(fwd)
    assert abs(abs(fwd) ** 2 / nfg - power) < 1e-10 * power


# ----------------------------------------------- a mode's amplitude as weights (an objective)

def test_mode_weights_are_the_forms_coefficients():
    """Lattice::mode_amplitude_weights (ports.rs:293-345): sum(weights * u) is the oracle
    restatement's N(u, g) (forward) and -N(u, f) (backward) for a random field u over the mode's
    two planes and a random complex mode, in a box with CPMLs across (the stretches in play),
    to round-off (the sums run in another order)."""
    import photonoxide_amd as P
    n = (16, 12, 10)
    g = P.Grid3d(*n, H_STEP, H_STEP, H_STEP)
    s = P.Simulation.new(g, P.Scene(2.0), P.Boundaries(P.Pml(3, 3), P.Pml(3, 3), P.Pml(2, 2)), 0.9)
    rng = np.random.default_rng(5)
    k0 = s.leapfrog_wavenumber(1.3)
    tan = rng.standard_normal((2, n[2], n[1])) + 1j * rng.standard_normal((2, n[2], n[1]))
    nor = rng.standard_normal((n[2], n[1])) + 1j * rng.standard_normal((n[2], n[1]))
    mode = P.PortMode(0, 7, H_STEP, k0, 1.4 * k0 + 0.02j, tan, nor)
    lo, hi = s.mode_box(mode)
    u = rng.standard_normal((3, n[2], n[1], 2)) + 1j * rng.standard_normal((3, n[2], n[1], 2))

    def field(c, at):
        v = u[c, at[2] - lo[2], at[1] - lo[1], at[0] - lo[0]]
        return (float(v.real), float(v.imag))
    lat = MM.Lattice(n, (H_STEP,) * 3, (0.0, 0.0, 0.0), (0, 0, 0), (3, 3, 2), (3, 3, 2), 1e-8, 3.0, k0)
    of, ob = MM.mode_amplitudes(lat, oracle_mode(mode), field, EXP, sincos)
    for direction, expected in ((0, complex(*of)), (1, complex(*ob))):
        w = s.mode_weights(mode, direction)
        got = np.sum(w * u)
        scale = np.sum(np.abs(w) * np.abs(u))
        print(f"direction {direction}: weights {got}, form {expected}, difference {abs(got - expected) / scale:.2e} of the terms")
        assert abs(got - expected) < 1e-14 * scale


@pytest.mark.gpu
def test_the_adjoint_differentiates_a_modes_transmission():
    """The adjoint with a mode's weights as its objective (component -1): J = |a|^2, a the strip
    guide's fundamental (this build's port_modes) going forward on the output plane. J is the mode
    monitor's |a|^2 of a plain run, and dJ/drho agrees with central differences of it in FP64
    within 1e-3 at random design cells."""
    import torch
    from adjoint_problem import DesignProblem
    from photonoxide_amd.adjoint import Adjoint, objective
    import photonoxide_amd as P
    print("case adjoint of a mode's transmission", flush=True)

    class ModeProblem(DesignProblem):
        def mode(self, s):
            i_mon = self.n[0] - self.pml - 3
            return s.port_modes(0, i_mon, 1 / 1.55)[0]

        def build(self, rho):
            s = self.simulation(rho)
            m = self.mode(s)
            freq = s.mode_frequencies(s.add_mode_monitor([m]))[0]
            lo, hi = s.mode_box(m)
            d = s.add_dft(lo, hi, [freq])
            adj = Adjoint(s, monitor=d, component=-1, frequency=0, weights=s.mode_weights(m, 0),
                          design=self.design_values(), checkpoint=self.checkpoint)
            return s, adj, self.steps

        def objective(self, rho):
            s = self.simulation(rho)
            mon = s.add_mode_monitor([self.mode(s)])
            s.run(self.steps)
            return float(abs(s.mode_amplitudes(mon)[0][0]) ** 2)

    p = ModeProblem(P.Gpu(0))
    rng = np.random.default_rng(8)
    rho = rng.random(p.shape)
    r = torch.tensor(rho, dtype=torch.float64, requires_grad=True)
    j = objective(r, p)
    j.backward()
    g = r.grad.numpy()
    plain = p.objective(rho)
    jv = float(j.detach())
    print(f"J {jv:.12e}, plain run's |a|^2 {plain:.12e}")
    assert abs(jv - plain) < 1e-10 * plain
    d = 1e-3
    worst = 0.0
    for _ in range(4):
        c = tuple(rng.integers(0, q) for q in p.shape)
        hi, lo = rho.copy(), rho.copy()
        hi[c] += d
        lo[c] -= d
        fd = (p.objective(hi) - p.objective(lo)) / (2 * d)
        worst = max(worst, abs(g[c] - fd) / abs(fd))
        print(f"cell {c}: adjoint {g[c]:.9e}, central differences {fd:.9e}")
    assert worst < 1e-3, worst
