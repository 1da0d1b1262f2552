"""Flux monitors and the run-control rules on the device, against the oracle (GPU).

- A flux box and its six faces as planes (src/fdtd/monitors_checks.rs:188-245): the device's
  fluxes are the oracle's flux restatement over the oracle's FP64 sums, bit for bit in f64;
  no net flux leaves the lossless box with no source in it.
- run_until_decayed stops at the same step as the oracle's restatement of the rule, with the
  same answer, for a pulse (it decays) and a continuous wave (it reaches the limit).
"""
import math
import os
import sys

import numpy as np
import pytest

import oracle_lib as O
from fixtures import INF

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import control as K  # noqa: E402
import flux as F  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import photonoxide_amd as P
    return P.Gpu(0)


N, H = (16, 14, 12), 0.05
PLANES = [(5, 10), (4, 9), (3, 8)]
F0 = 1.0 / 1.55
FREQS = [F0, F0 + 0.15]


def balance_pair(inside):
    """The same closed box on the oracle and on the host mirror."""
    import photonoxide_amd as P
    x0 = [-m * H / 2 for m in N]
    lo = tuple(x0[a] + PLANES[a][0] * H - 2e-9 for a in range(3))
    hi = tuple(x0[a] + (PLANES[a][1] + 1) * H + 2e-9 for a in range(3))
    block = ((-0.08, -0.06, -0.05), (0.08, 0.06, 0.05))
    width = math.sqrt(2 * math.log(2)) / (math.pi * 0.3)
    at = (8, 6, 5) if inside else (2, 6, 5)
    o = O.Sim((*N, H, H, H, *x0), O.bounds(), O.scene(1.0, [(0, *block, 4.0)]), 0.9)
    o.conductivity(O.scene(3.0, [(0, lo, hi, 0.0)]))
    o.add_source(0, 1, at, O.wave(0, F0, width, 6 * width))
    g = P.Grid3d(*N, H, H, H, *x0)
    p = P.Simulation.new(g, P.Scene(1.0, [P.Box(*block, 4.0)]), P.Boundaries.walls(), 0.9)
    p.with_conductivity(P.Scene(3.0, [P.Box(lo, hi, 0.0)]))
    p.add_source(P.Source(P.Field.E, P.Axis.Y, at, P.Waveform(0, F0, width, 6 * width)))
    return o, p


def oracle_fluxes(o, faces):
    boxes = [f.box(N) for f in faces]
    ms = [o.add_dft(lo, hi, FREQS) for lo, hi in boxes]
    return boxes, ms


def value_fn(o, m, lo):
    series = {(fd, c): o.dft(m, fd, c) for fd in (0, 1) for c in range(3)}
    return lambda fd, c, at, f: series[fd, c][f, at[2] - lo[2], at[1] - lo[1], at[0] - lo[0]]


@pytest.mark.parametrize("inside", [False, True])
def test_flux_box_and_planes_are_the_oracles(gpu, inside):
    import photonoxide_amd as P
    print(f"case flux box, source inside {inside}", flush=True)
    o, p = balance_pair(inside)
    faces = F.box_faces(PLANES)
    boxes, ms = oracle_fluxes(o, faces)
    around = p.add_flux_box(PLANES, FREQS)
    planes = [p.add_flux(f.axis, f.plane, FREQS, window=f.window) for f in faces]
    p.set_device(gpu, P.Precision.F64)
    K.run_until(o, 200.0)
    p.run_until(200.0)
    assert p.steps() == o.steps()
    values = [value_fn(o, m, lo) for m, (lo, _) in zip(ms, boxes)]
    want_box = [F.total_flux(faces, N, (H, H, H), values, f) for f in range(2)]
    got_box = p.flux(around)
    assert list(got_box) == want_box, (got_box, want_box)
    for q, face in enumerate(faces):
        want = [face.flux(N, (H, H, H), values[q], f) for f in range(2)]
        assert list(p.flux(planes[q])) == want, q
    largest = max(abs(v) for q in range(6) for v in p.flux(planes[q]))
    net = max(abs(v) for v in got_box)
    if inside:
        assert net > 0.1 * largest
    else:
        assert net < 1e-12 * largest, (net, largest)


def test_flux_in_f32_balances_to_f32_round_off(gpu):
    import photonoxide_amd as P
    _, p = balance_pair(False)
    around = p.add_flux_box(PLANES, FREQS)
    faces = [p.add_flux(f.axis, f.plane, FREQS, window=f.window) for f in F.box_faces(PLANES)]
    p.set_device(gpu, P.Precision.F32)
    p.run_until(200.0)
    largest = max(abs(v) for q in faces for v in p.flux(q))
    net = max(abs(v) for v in p.flux(around))
    print(f"f32 flux balance: {net / largest:.2e}")
    assert net < 1e-5 * largest, (net, largest)


def open_box_pair(cw):
    import photonoxide_amd as P
    n, h, cp = (40, 36, 1), 0.05, 8
    block = ((0.8, 0.7, -INF), (1.2, 1.1, INF))
    o = O.Sim((*n, h, h, h, 0, 0, 0), O.bounds(periodic=(0, 0, 1), low=(cp, cp, 0), high=(cp, cp, 0)),
              O.scene(1.0, [(0, *block, 3.0)]), 0.9)
    g = P.Grid3d(*n, h, h, h, 0, 0, 0)
    p = P.Simulation.new(g, P.Scene(1.0, [P.Box(*block, 3.0)]),
                         P.Boundaries(P.Pml(cp, cp), P.Pml(cp, cp), P.Periodic()), 0.9)
    if cw:
        o.add_source(0, 2, (14, 15, 0), O.wave(2, 1.0, 0.0, 0.0, 1.0, ramp=2.0))
        p.add_source(P.Source(P.Field.E, P.Axis.Z, (14, 15, 0), P.Waveform.continuous(1.0, 1.0, 2.0)))
        w = [(2, 0.0, 0.0)]
    else:
        o.add_source(0, 2, (14, 15, 0), O.wave(0, 1.0, 0.3, 1.8))
        p.add_source(P.Source(P.Field.E, P.Axis.Z, (14, 15, 0), P.Waveform.gaussian(1.0, 0.3, 1.8)))
        w = [(0, 0.3, 1.8)]
    probes = [(0, 2, (25, 20, 0)), (1, 0, (20, 25, 0))]
    po = [o.add_probe(f, c, at) for f, c, at in probes]
    pp = [p.add_probe(f, c, at) for f, c, at in probes]
    return o, p, po, pp, w


@pytest.mark.parametrize("cw", [False, True])
def test_run_until_decayed_stops_where_the_oracle_does(gpu, cw):
    import photonoxide_amd as P
    print(f"case run_until_decayed cw {cw}", flush=True)
    o, p, po, pp, w = open_box_pair(cw)
    p.set_device(gpu, P.Precision.F64)
    limit = 8.0 if cw else 40.0
    want = K.run_until_decayed(o, po, 1e-4, 0.5, limit, w)
    got = p.run_until_decayed(pp, 1e-4, 0.5, limit)
    assert got == want == (not cw)
    assert p.steps() == o.steps()
    for a, b in zip(po, pp):
        assert np.array_equal(np.asarray(o.probe(a)).view(np.uint64), p.probe(b).view(np.uint64))
