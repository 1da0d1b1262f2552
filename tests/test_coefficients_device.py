"""Coefficient construction on the device (GPU): pxfdtd_simulation_new_on_device averages each
E value's permittivity over 8^3 samples on the GPU (Simulation::diagonal, src/fdtd.rs:619-653;
fdfd::averaged, src/fdfd/three.rs:563-588) and must give the host's bits, for every value.

The host path (pxfdtd_simulation_new) is the reference's arithmetic (pinned by the canary hash
in tests/test_oracle.py). Cases exercise each of the device's exact shortcuts: cells no shape
reaches (the uniform average), cells wholly inside a box or a cylinder (the covering shape),
cells cut by faces and curved walls, overlapping shapes where the later one wins, more shapes
reaching one cell than the device keeps in registers, anisotropic values, and infinite box
faces; then the dispersive workload's pillar-and-block scene at a reduced size.
"""
import numpy as np
import pytest

from fixtures import INF

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import photonoxide_amd as P
    return P.Gpu(0)


def build(P, grid, scene, gpu=None):
    g = P.Grid3d(*grid)
    b = P.Boundaries.cpml_all(2)
    return P.Simulation.new(g, scene, b, 0.9, device=gpu)


def same_permittivity(P, gpu, grid, scene):
    import time
    t0 = time.perf_counter()
    host = build(P, grid, scene).permittivity()
    t1 = time.perf_counter()
    dev = build(P, grid, scene, gpu).permittivity()
    t2 = time.perf_counter()
    print(f"set-up: host {t1 - t0:.3f} s, averaging on the device {t2 - t1:.3f} s", flush=True)
    assert host.shape == dev.shape
    diff = np.flatnonzero(host.view(np.uint64) != dev.view(np.uint64))
    assert diff.size == 0, (diff.size, diff[:5], host.ravel()[diff[:5]], dev.ravel()[diff[:5]])
    return host


def test_boxes_cylinders_and_overlaps_are_the_hosts_bits(gpu):
    import photonoxide_amd as P
    print("case device averaging: boxes, cylinders, overlaps", flush=True)
    h = 0.05
    grid = (40, 34, 22, h, h, h, -1.0, -0.85, -0.55)
    shapes = [
        P.Box((-0.52, -0.31, -INF), (0.47, 0.29, INF), (12.0, 11.0, 10.0)),   # faces through cells
        P.Box((-0.3, -0.2, -0.2), (0.3, 0.2, 0.2), 2.25),                      # nested, later wins
        P.Cylinder((0.33, 0.11), 0.17, (-0.4, 0.31), (6.0, 6.5, 7.0)),         # curved wall
        P.Cylinder((-0.6, -0.4), 0.33, (-INF, INF), 3.0),                      # cells wholly inside
        P.Box((-0.2, -0.1, -0.0), (-0.05, 0.05, 0.125), 9.0),                  # within one cell row
        P.Box((-INF, 0.6, -INF), (INF, INF, INF), 4.0),                        # a half space
    ]
    host = same_permittivity(P, gpu, grid, P.Scene((1.0, 1.5, 2.0), shapes))
    # the scene is not trivial: averaged values strictly between the shapes' values exist
    assert len(np.unique(host)) > 50


def test_many_shapes_reaching_one_cell(gpu):
    """Thin slabs of 1/40 of a cell, 20 of them in one cell row along x, and small cylinders
    crowded into a few cells: more shapes reach a cell than the device keeps in registers."""
    import photonoxide_amd as P
    print("case device averaging: many shapes in one cell", flush=True)
    h = 0.04
    grid = (24, 20, 16, h, h, h, -0.48, -0.4, -0.32)
    shapes = [P.Box((-0.1 + q * 0.002, -INF, -INF), (-0.1 + q * 0.002 + 0.001, INF, INF), 2.0 + 0.1 * q)
              for q in range(20)]
    rng = np.random.default_rng(5)
    for q in range(30):
        c = rng.uniform(-0.05, 0.05, 2)
        shapes.append(P.Cylinder((c[0], c[1]), rng.uniform(0.003, 0.03), (-INF, INF), 5.0 + 0.2 * q))
    same_permittivity(P, gpu, grid, P.Scene(1.0, shapes))


def test_the_dispersive_workloads_scene(gpu):
    """The dispersive workload's geometry (pillars and blocks) as a permittivity scene, at
    128 x 128 x 64."""
    import photonoxide_amd as P
    from configs import dispersive_3d
    print("case device averaging: the dispersive workload's scene at 128x128x64", flush=True)
    spec = dispersive_3d(n=(128, 128, 64), pillars=20, blocks=6, seed=3)
    shapes = []
    for q, (_, _, (_, sh)) in enumerate(spec["media"]):
        for kind, lo, hi, _ in sh:
            val = 12.0 if q == 0 else 4.0
            if kind == 0:
                shapes.append(P.Box(lo, hi, val))
            else:
                shapes.append(P.Cylinder((lo[0], lo[1]), hi[0], (lo[2], hi[2]), val))
    same_permittivity(P, gpu, spec["grid"], P.Scene(2.085, shapes))
