"""The native library on the CPU: it loads and exports what include/pxfdtd.h declares, and the
host mirror's set-up (dt, averaged permittivity, walls, conductivity, CPML) is the oracle's, bit
for bit. No GPU call is made."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle_lib as O
from fixtures import INF, busy_spec, oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "pxfdtd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pxfdtd_\w+)\s*\(", text)))


def test_library_exports_every_declared_function():
    from photonoxide_amd import _native as N
    out = subprocess.run(["nm", "-D", "--defined-only", N.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (pxfdtd_\w+)", out))
    declared = declared_functions()
    assert declared, "no declarations found"
    missing = [f for f in declared if f not in exported]
    assert not missing, missing
    for f in declared:
        assert hasattr(N.lib, f)
        assert f in N.SIGNATURES, f"{f} has no ctypes signature"


def test_setup_is_the_oracles_bit_for_bit():
    import photonoxide_amd as P
    from fixtures import product
    for two_d in (False, True):
        spec = busy_spec(two_d)
        a = oracle(spec, with_fields=False)
        b = product(spec, with_fields=False)
        assert a.dt() == b.dt()
        for c in range(3):
            ca0, cb0 = a.coefficients(c)
            ca1, cb1 = b.coefficients(c)
            assert np.array_equal(ca0.view(np.uint64), ca1.view(np.uint64)), c
            assert np.array_equal(cb0.view(np.uint64), cb1.view(np.uint64)), c


def test_canary_setup_matches():
    import photonoxide_amd as P
    g = P.Grid3d(48, 40, 1, 0.05, 0.05, 0.05, -1.2, -1.0, -0.025)
    b = P.Boundaries(P.Pml(), P.Pml(), P.Periodic())
    s = P.Simulation.new(g, P.Scene(1.0, [P.Box((-0.3, -0.2, -INF), (INF, 0.2, INF), 6.25)]), b, 0.9)
    o = O.Sim(tuple(vars(g).values()), O.bounds(periodic=(0, 0, 1)),
              O.scene(1.0, [(0, (-0.3, -0.2, -INF), (INF, 0.2, INF), 6.25)]), 0.9)
    assert s.dt() == o.dt()
    for c in range(3):
        assert np.array_equal(s.coefficients(c)[1], o.coefficients(c)[1])


def test_errors_are_statuses():
    import photonoxide_amd as P
    from photonoxide_amd._native import NativeError
    g = P.Grid3d(8, 8, 8, 0.05, 0.05, 0.05)
    with pytest.raises(NativeError, match="Courant"):
        P.Simulation.new(g, P.Scene(1.0), P.Boundaries.walls(), 1.5)
    with pytest.raises(NativeError, match="leave nothing"):
        P.Simulation.new(g, P.Scene(1.0), P.Boundaries.cpml_all(4), 0.9)
    s = P.Simulation.new(g, P.Scene(1.0), P.Boundaries.cpml_all(2), 0.9)
    with pytest.raises(NativeError, match="outside"):
        s.add_probe(0, 0, (8, 0, 0))
    with pytest.raises(NativeError, match="no device"):
        s.run(1)
    with pytest.raises(NativeError, match="Nyquist"):
        s.add_dft((0, 0, 0), (1, 1, 1), [1e6])


def test_plane_wave_refusals_are_the_references():
    """src/fdtd/tests.rs:350-400, each refused by the host mirror as by the oracle: a face in or
    against the CPML, the box not whole along a periodic axis, a direction along it, a box
    across a block, the polarization along the direction, a zero direction."""
    import photonoxide_amd as P
    from photonoxide_amd._native import NativeError
    g = P.Grid3d(40, 40, 1, 0.05, 0.05, 0.05, -1.0, -1.0, -0.025)
    b = P.Boundaries(P.Pml(6, 6), P.Pml(6, 6), P.Periodic())
    block = P.Scene(1.0, [P.Box((0.4 + 1e-9, -INF, -INF), (INF, INF, INF), 4.0)])
    pulse = P.Waveform.pulse(1.0, 0.5)
    s = P.Simulation.new(g, block, b, 0.9)
    assert s.add_plane_wave(P.PlaneWave((10, 10, 0), (25, 25, 1), (2, 1, 0), (0, 0, 1), 1.0, pulse)) == 0
    for low, high, d, pol in [((5, 10, 0), (25, 25, 1), (1, 0, 0), (0, 0, 1)),
                              ((6, 10, 0), (25, 25, 1), (1, 0, 0), (0, 0, 1)),
                              ((10, 10, 0), (25, 25, 0), (1, 0, 0), (0, 0, 1)),
                              ((10, 10, 0), (25, 25, 1), (1, 0, 1), (0, 0, 1)),
                              ((10, 10, 0), (30, 25, 1), (1, 0, 0), (0, 0, 1)),
                              ((10, 10, 0), (25, 25, 1), (1, 0, 0), (1, 0, 0)),
                              ((10, 10, 0), (25, 25, 1), (0, 0, 0), (0, 0, 1))]:
        with pytest.raises(NativeError):
            s.add_plane_wave(P.PlaneWave(low, high, d, pol, 1.0, pulse))
    # the incident field before any step: at rest
    assert s.incident(0, 0, 2, (15, 15, 0)) == 0.0


def test_bytes_per_step_is_the_references_model():
    """Yee::bytes_per_step (src/fdtd/kernel/yee.rs:141-153): (24 N + 6 psi_E + 5 psi_H) values,
    the psi counts from the oracle's slabs."""
    import ctypes as C
    import photonoxide_amd as P
    from fixtures import product
    spec = busy_spec(False)
    o = oracle(spec, with_fields=False)
    p = product(spec, with_fields=False)
    psi = {0: 0, 1: 0}
    for q in range(O.lib.or_sim_slab_count(o.h)):
        info = (C.c_int32 * 5)()
        b, a, s = O.pd(), O.pd(), O.pd()
        n = C.c_int64()
        O.lib.or_sim_slab(o.h, q, info, C.byref(b), C.byref(a), C.byref(s), C.byref(n))
        psi[info[0]] += n.value
    cells = spec["grid"][0] * spec["grid"][1] * spec["grid"][2]
    want = 24 * cells + 6 * psi[0] + 5 * psi[1]
    assert p.bytes_per_step(P.Precision.F32) == 4 * want
    assert p.bytes_per_step(P.Precision.F64) == 8 * want


def test_a_refused_conductivity_leaves_the_simulation_as_it_was():
    """src/fdtd.rs:949-992: a conductivity that is negative or not finite anywhere is refused;
    here the simulation keeps its coefficients (no value half-converted), and takes a good one
    afterwards."""
    import photonoxide_amd as P
    from photonoxide_amd import _native as N
    g = P.Grid3d(12, 10, 8, 0.05, 0.05, 0.05, -0.3, -0.25, -0.2)
    s = P.Simulation.new(g, P.Scene(2.0), P.Boundaries.cpml_all(2), 0.9)
    before = [s.coefficients(c) for c in range(3)]
    with pytest.raises(N.NativeError):
        s.with_conductivity(P.Scene(1.0, [P.Box((0.1, -INF, -INF), (INF, INF, INF), -0.5)]))
    for c in range(3):
        ca, cb = s.coefficients(c)
        assert np.array_equal(ca, before[c][0]) and np.array_equal(cb, before[c][1]), c
    s.with_conductivity(P.Scene(1.0))
    assert not np.array_equal(s.coefficients(0)[1], before[0][1])


def test_the_conductivity_is_set_once():
    """A second conductivity would divide cb again while ca took only the second value (the
    reference's builder would compound them): refused, the first kept."""
    import photonoxide_amd as P
    from photonoxide_amd import _native as N
    g = P.Grid3d(12, 10, 8, 0.05, 0.05, 0.05, -0.3, -0.25, -0.2)
    s = P.Simulation.new(g, P.Scene(2.0), P.Boundaries.cpml_all(2), 0.9)
    s.with_conductivity(P.Scene(1.0))
    once = [s.coefficients(c) for c in range(3)]
    with pytest.raises(N.NativeError):
        s.with_conductivity(P.Scene(0.5))
    for c in range(3):
        assert np.array_equal(s.coefficients(c)[0], once[c][0]) and np.array_equal(s.coefficients(c)[1], once[c][1])
