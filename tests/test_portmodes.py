"""The port-mode solver (Lattice::port_modes, src/fdfd/three/ports.rs:408-594) on the host mirror,
held to the reference's own port tests (src/cases/fdfd/three/port_tests.rs) and its mode-source
check (src/fdtd/checks.rs:742-826, tests.rs:285-296).

CPU:
- in a uniform guide periodic across, the fundamental's beta is the grid's own dispersion,
  (2/h sin(beta h / 2))^2 = eps k0^2, to 1e-13 (E_y and E_z are degenerate there: one start
  vector finds one mode of the pair);
- in the reference's walled strip (port_tests.rs:38-70), the mode extended as e^(i beta x)
  solves every row of the 3D system k0^2 eps E - curl curl E clear of the PMLs to 1e-13 of the
  rows' scale (port_tests.rs:71-100; the rows from the oracle's restatement of the lattice);
- a real mode normalized by its Lorentz form carries unit power to 1e-12, its tangential E is
  real and its normal E imaginary half a step on (port_tests.rs:102-133);
- the refusals.

GPU: the reference's open guide (40 x 36 x 40 cells of 50 nm, CPMLs of 8, 64 steps a period of
1.55 um), its fundamental along z launched forward from plane 13 by a continuous wave: the
backward power behind (plane 10's backward amplitude squared times its power) is below 1e-5 of
the forward flux through plane 26, and that flux is the mode's power at the source's amplitude
to 5e-4 (the reference's bounds); the mode monitor's forward amplitude there, over the
transform's own weight, is 1 to 5e-4 as well.
"""
import math
import os
import sys

import numpy as np
import pytest

import oracle_lib as O

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import modemonitor as MM  # noqa: E402

MM.use_portable(O.lib.or_ln, O.lib.or_pow)

H = 0.05
SILICON, OXIDE = 3.48, 1.444


def sincos(x):
    s, c = O.C.c_double(), O.C.c_double()
    O.lib.or_sincos(x, O.C.byref(s), O.C.byref(c))
    return s.value, c.value


def oracle_mode(m):
    return dict(axis=m.axis, plane=m.plane, step=m.step, beta=(complex(m.beta).real, complex(m.beta).imag),
                tangential=[[[(float(v.real), float(v.imag)) for v in row] for row in comp] for comp in m.tangential],
                normal=[[(float(v.real), float(v.imag)) for v in row] for row in m.normal])


def test_a_uniform_guides_mode_has_the_grids_dispersion():
    import photonoxide_amd as P
    g = P.Grid3d(40, 6, 4, H, H, H)
    s = P.Simulation.new(g, P.Scene(1.5 ** 2), P.Boundaries(P.Pml(8, 8), P.Periodic(), P.Periodic()), 0.9)
    k0 = s.leapfrog_wavenumber(1.0)
    exact = 2.0 / H * math.asin(H * 1.5 * k0 / 2.0)
    modes = s.port_modes(0, 20, 1.0, count=1)
    print([m.beta for m in modes], exact)
    assert abs(modes[0].beta - exact) < 1e-13 * exact


def walled_strip():
    """port_tests.rs:38-70: a 0.5 x 0.22 um silicon strip along x in oxide, 14 x 16 x 12 cells of
    50 nm centred across, PMLs of 4 along x, walls across."""
    import photonoxide_amd as P
    n = (14, 16, 12)
    g = P.Grid3d(*n, H, H, H, 0.0, -n[1] * H / 2, -n[2] * H / 2)
    s = P.Simulation.new(g, P.Scene(OXIDE ** 2, [P.Box((-np.inf, -0.25, -0.11), (np.inf, 0.25, 0.11), SILICON ** 2)]),
                         P.Boundaries(P.Pml(4, 4), P.Pml(0, 0), P.Pml(0, 0)), 0.9)
    return s, n


def test_a_port_mode_solves_every_row_of_the_3d_system():
    s, n = walled_strip()
    mode = s.port_modes(0, 6, 1 / 1.55, count=1)[0]
    print(f"beta {mode.beta}, n_eff {mode.beta / mode.k0}")
    eps = s.permittivity().reshape(3, -1)
    lat = MM.Lattice(n, (H,) * 3, (0.0, -n[1] * H / 2, -n[2] * H / 2), (0, 0, 0), (4, 0, 0), (4, 0, 0), 1e-8, 3.0,
                     mode.k0)
    om = oracle_mode(mode)
    cache = {}

    def field(r):
        if r not in cache:
            c, at = lat.at(r)
            cache[r] = complex(*MM.mode_value(om, c, at, True, O.lib.or_exp, sincos))
        return cache[r]
    worst, largest = 0.0, 0.0
    for c in range(3):
        for k in range(n[2]):
            for j in range(n[1]):
                for i in range(5, 9):
                    r = lat.index(c, (i, j, k))
                    if lat.fixed(r):
                        applied, scale = field(r), abs(field(r))
                    else:
                        diag = mode.k0 ** 2 * eps[c][r - c * lat.cells] * field(r)
                        terms = [diag] + [complex(*a) * field(q) for q, a in lat.curl_curl(r)]
                        applied, scale = sum(terms), sum(abs(t) for t in terms)
                    worst = max(worst, abs(applied))
                    largest = max(largest, scale)
    print(f"worst row {worst:.3e} of {largest:.3e}")
    assert worst < 1e-13 * largest


def test_a_real_mode_carries_unit_power_and_is_real():
    s, n = walled_strip()
    modes = s.port_modes(0, 6, 1 / 1.55, count=3)
    real = [m for m in modes if abs(m.beta.imag) < 1e-9 * abs(m.beta)]
    print([(m.beta / m.k0, m.power) for m in modes])
    assert len(real) >= 2
    assert all(abs(modes[q].beta.real) >= abs(modes[q + 1].beta.real) for q in range(len(modes) - 1))
    for m in real:
        assert abs(m.power - 1.0) < 1e-12, m.power
        half = np.exp(-0.5j * H * m.beta)
        largest = max(np.abs(m.tangential).max(), np.abs(m.normal).max())
        wrong = max(np.abs(m.tangential.imag).max(), np.abs((m.normal * half).real).max())
        assert wrong < 1e-12 * largest, (wrong, largest)


def test_the_jobs_te_and_tm_picks():
    """src/job/fdtd.rs:2108-2135: TE has its E mostly across the guide (y for a port along x),
    TM mostly along z; in the walled strip (wider than tall) the TE fundamental has the higher
    index."""
    s, n = walled_strip()
    te = s.port_mode(0, 6, 1 / 1.55, te=True)
    tm = s.port_mode(0, 6, 1 / 1.55, te=False)
    ey, ez = (np.abs(te.tangential) ** 2).sum(axis=(1, 2))
    print(f"TE n_eff {te.beta / te.k0}, TM n_eff {tm.beta / tm.k0}")
    assert ey > ez
    ey, ez = (np.abs(tm.tangential) ** 2).sum(axis=(1, 2))
    assert ez > ey
    assert te.beta.real > tm.beta.real


def test_host_mirror_refuses_ports_that_dont_fit():
    """ports.rs:530-580 (check_port) and the mode source's clearance of the CPMLs."""
    import photonoxide_amd as P
    s, n = walled_strip()
    bad = [dict(axis=0, plane=4), dict(axis=0, plane=9), dict(axis=1, plane=0),
           dict(axis=0, plane=6, window=((0, 2), None)), dict(axis=0, plane=6, window=((0, 17), None)),
           dict(axis=0, plane=6, count=0), dict(axis=0, plane=6, frequency=1e3)]
    for b in bad:
        kw = dict(frequency=1 / 1.55)
        kw.update(b)
        with pytest.raises(Exception):
            s.port_modes(**kw)
    g = P.Grid3d(20, 6, 4, H, H, H)
    p = P.Simulation.new(g, P.Scene(2.0), P.Boundaries(P.Pml(4, 4), P.Periodic(), P.Periodic()), 0.9)
    with pytest.raises(Exception):
        p.port_modes(0, 8, 1.0, window=((1, 5), None))  # periodic across: the window must be whole
    with pytest.raises(Exception):
        p.port_modes(1, 3, 1.0)  # periodic along


# ------------------------------------------------------------------------------ the device

def open_guide():
    """checks.rs:742-770: eps 12 in a 0.4 x 0.3 um core along z in eps 2.1, 40 x 36 x 40 cells
    of 50 nm centred on it, CPMLs of 8, 64 steps a period of 1.55 um."""
    import photonoxide_amd as P
    n = (40, 36, 40)
    g = P.Grid3d(*n, H, H, H, -n[0] * H / 2, -n[1] * H / 2, -n[2] * H / 2)
    courant = 1.55 / 64 * math.sqrt(3.0 / (H * H))
    s = P.Simulation.new(g, P.Scene(2.1, [P.Box((-0.2, -0.15, -np.inf), (0.2, 0.15, np.inf), 12.0)]),
                         P.Boundaries.cpml_all(8), courant)
    return s, n


@pytest.mark.gpu
def test_a_mode_source_launches_its_mode_one_way_with_its_power():
    import photonoxide_amd as P
    print("case port mode launched one way", flush=True)
    s, n = open_guide()
    f = 1 / 1.55
    launch = s.port_modes(2, 13, f)[0]
    behind = s.port_modes(2, 10, f)[0]
    ahead = s.port_modes(2, 26, f)[0]
    print(f"n_eff {launch.beta / launch.k0}, power {launch.power}")
    s.add_mode_source(launch, 0, P.Waveform.continuous(f, ramp=4.0 * 1.55))
    s.set_device(P.Gpu(0), P.Precision.F64)
    spp = 64
    s.run(40 * spp)
    mon = s.add_mode_monitor([behind, ahead])
    freq = s.mode_frequencies(mon)[1]
    flux = s.add_flux(2, 26, [freq])
    periods = 4
    s.run(periods * spp)
    (_, back), (fwd, _) = s.mode_amplitudes(mon)
    forward = s.flux(flux)[0]
    weight = 0.5 * periods * spp * s.dt()  # the transform of a unit steady wave over whole periods
    backward_share = abs(back) ** 2 * behind.power / forward
    forward_share = forward / (weight ** 2 * ahead.power)
    print(f"backward {backward_share:.2e} of forward, forward {forward_share:.6f} of the mode's power, "
          f"forward amplitude {abs(fwd) / weight:.6f}")
    assert backward_share < 1e-5
    assert abs(forward_share - 1.0) < 5e-4
    assert abs(abs(fwd) / weight - 1.0) < 5e-4


@pytest.mark.gpu
def test_a_mode_source_in_a_lossy_box_is_fdfds_mode_source():
    """checks.rs:689-740, tests.rs:279-283: the guide's fundamental (this build's port_modes with
    the leapfrog's loss, eps + i sigma cos(w dt / 2) / k0) launched forward from plane 6 by a
    continuous wave in a closed box of a lossy medium (sigma = 4/um everywhere; 16 x 14 x 24 cells
    of 50 nm, 64 steps a period of 1.55 um), its steady amplitude after 90 periods (the
    transform over 4 more, times 2 / (4 x 64 dt)) against FDFD's field for FDFD's own mode source
    of the same mode (Lattice::mode_source, ports.rs:594-634, restated here; A = k0^2 eps -
    curl curl from the oracle's lattice, walls all round, solved by scipy's sparse LU): within
    1e-10 of the largest field, the reference's bound. The mode monitor's amplitudes over the
    same periods (the same scale), for the mode behind on plane 4 and ahead on plane 16, are
    FDFD's mode_amplitudes of FDFD's field (the oracle's restatement of the form) within 1e-9 of
    the largest (monitors_checks.rs:255-310, there with a pulse over its analytic spectrum)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spl
    import photonoxide_amd as P
    print("case mode source in a lossy box against FDFD", flush=True)
    n, spp, sigma = (16, 14, 24), 64, 4.0
    g = P.Grid3d(*n, H, H, H, -n[0] * H / 2, -n[1] * H / 2, -n[2] * H / 2)
    courant = 1.55 / spp * math.sqrt(3.0 / (H * H))
    walls = P.Pml(0, 0)
    s = P.Simulation.new(g, P.Scene(2.1, [P.Box((-0.2, -0.15, -np.inf), (0.2, 0.15, np.inf), 12.0)]),
                         P.Boundaries(walls, walls, walls), courant)
    s.with_conductivity(P.Scene(sigma))
    f = 1 / 1.55
    mode = s.port_modes(2, 6, f, lossy=True)[0]
    behind = s.port_modes(2, 4, f, lossy=True)[0]
    ahead = s.port_modes(2, 16, f, lossy=True)[0]
    k0 = mode.k0
    dt = s.dt()
    s.add_mode_source(mode, 0, P.Waveform.continuous(f, ramp=4.0 * 1.55))
    s.set_device(P.Gpu(0), P.Precision.F64)
    s.run(90 * spp)
    d = s.add_dft((0, 0, 0), (n[0] - 1, n[1] - 1, n[2] - 1), [f])
    mon = s.add_mode_monitor([behind, ahead])
    s.run(4 * spp)
    amp = np.concatenate([s.dft(d, P.Field.E, c)[0].ravel() for c in range(3)]) * (2.0 / (4 * spp * dt))
    measured = [(a * (2.0 / (4 * spp * dt)), bk * (2.0 / (4 * spp * dt))) for a, bk in s.mode_amplitudes(mon)]
    # FDFD: A x = b on the same lattice
    lat = MM.Lattice(n, (H,) * 3, (g.x0, g.y0, g.z0), (0, 0, 0), (0, 0, 0), (0, 0, 0), 1e-8, 3.0, k0)
    loss = sigma * math.cos(2 * math.pi * f * dt / 2.0) / k0
    eps = s.permittivity().ravel() + 1j * loss
    total = 3 * lat.cells
    rows, cols, vals = [], [], []
    for r in range(total):
        if lat.fixed(r):
            rows.append(r), cols.append(r), vals.append(1.0)
            continue
        rows.append(r), cols.append(r), vals.append(k0 * k0 * eps[r])
        for q, a in lat.curl_curl(r):
            rows.append(r), cols.append(q), vals.append(complex(*a))
    A = sp.csc_matrix((vals, (rows, cols)), shape=(total, total))
    om = oracle_mode(mode)
    p = mode.plane

    def scattered(r):
        c, at = lat.at(r)
        return 2 * at[2] + (1 if c == 2 else 0) < 2 * p
    b = np.zeros(total, complex)
    for m in range(p - 1, p + 3):
        for v in range(n[1]):
            for u in range(n[0]):
                for c in range(3):
                    r = lat.index(c, (u, v, m))
                    qr = scattered(r)
                    acc = 0j
                    for q, a in lat.curl_curl(r):
                        if scattered(q) != qr:
                            cq, at = lat.at(q)
                            val = complex(*MM.mode_value(om, cq, at, True, O.lib.or_exp, sincos))
                            acc += (1.0 if qr else -1.0) * complex(*a) * val
                    b[r] += acc
    x = spl.spsolve(A, b)
    largest = np.abs(x).max()
    worst = np.abs(x - amp).max()
    print(f"n_eff {mode.beta / k0}, largest field {largest:.4e}, worst difference {worst:.3e}, "
          f"relative {worst / largest:.2e}")
    assert worst < 1e-10 * largest

    def fdfd_field(c, at):
        v = x[lat.index(c, at)]
        return (float(v.real), float(v.imag))
    expected = [MM.mode_amplitudes(lat, oracle_mode(m), fdfd_field, O.lib.or_exp, sincos) for m in (behind, ahead)]
    expected = [(complex(*a), complex(*bk)) for a, bk in expected]
    top = max(max(abs(a), abs(bk)) for a, bk in expected)
    err = max(max(abs(e[0] - m[0]), abs(e[1] - m[1])) for e, m in zip(expected, measured))
    print(f"mode amplitudes: FDFD {expected}, monitor {measured}, worst {err / top:.2e} of the largest")
    assert err < 1e-9 * top
