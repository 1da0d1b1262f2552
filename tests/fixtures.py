"""Shared problems for the parity tests, built the same way on the oracle and on the device.

`busy` follows the reference's GPU check problem (src/fdtd/gpu/checks.rs:36-147): a dielectric
block and a lossy one, a conductor, CPMLs with kappa and alpha, a wall and (in 2D) periodic
sides, random fields to start from, point sources on E and H and a current sharing a value with
one of them, probes and a transform box, at sizes that leave tiles part-filled.
"""
import math

import numpy as np

import oracle_lib as O

INF = float("inf")


def noise(r):
    """A pseudo-random value in [-1, 1) from an integer (src/fdtd/checks.rs:84-90)."""
    r = np.asarray(r, dtype=np.uint64)
    with np.errstate(over="ignore"):
        x = (r * np.uint64(0x9E3779B97F4A7C15)) ^ np.uint64(0x5DEECE66)
        x ^= x >> np.uint64(31)
        x = x * np.uint64(0xBF58476D1CE4E5B9)
        x ^= x >> np.uint64(29)
    return (x >> np.uint64(11)).astype(np.float64) / float(1 << 52) - 1.0


def pulse():
    width = math.sqrt(2.0 * math.log(2.0)) / (math.pi * 0.6)
    return dict(kind=0, frequency=1.0, width=width, delay=6.0 * width)


def busy_spec(two_d=False, periodic=False):
    h = 0.05
    if periodic:
        # periodic along y and z at sizes no tile divides, a CPML and a wall along x
        n = (37, 21, 19)
        b = dict(periodic=(0, 1, 1), low=(4, 0, 0), high=(0, 0, 0), kappa=1.5, alpha=0.1)
    elif two_d:
        n = (70, 45, 1)
        b = dict(periodic=(1, 0, 1), low=(0, 6, 0), high=(0, 6, 0), kappa=2.0, alpha=0.3)
    else:
        n = (37, 21, 19)
        b = dict(periodic=(0, 0, 0), low=(4, 4, 3), high=(5, 0, 3), kappa=3.0, alpha=0.2)
    grid = (n[0], n[1], n[2], h, h, h, -n[0] * h / 2, -n[1] * h / 2, -n[2] * h / 2)
    nx, ny, nz = n
    mid = (nx // 2, ny // 2, nz // 2)
    spec = dict(
        grid=grid, bounds=b, courant=0.9,
        eps=(1.0, [(0, (-0.3, -0.2, -INF), (0.3, 0.2, INF), 12.0)]),
        sigma=(0.0, [(0, (0.6, 0.3, -INF), (INF, INF, INF), 2.0)]),
        conductor=(0, (nx // 3, ny // 3, nz // 2)),
        sources=[(0, 2, mid, pulse()),
                 (1, 0, (nx // 4, ny // 2, nz // 2), dict(kind=1, frequency=0.0, width=0.3, delay=1.2))],
        currents=[(0, [(2, mid, 0.5 - 0.25j), (1, (nx // 2 + 3, ny // 2, nz // 2), -0.3 + 0.7j)], pulse())],
        probes=[(f, c, at) for at in [(3, 3, 0), (nx - 8, ny // 3, nz // 3), mid] for (f, c) in [(0, 2), (1, 1)]],
        dft=((nx // 2 - 6, ny // 2 - 4, 0), (nx // 2 + 6, ny // 2 + 4, nz - 1), [0.8, 1.0, 1.3]),
    )
    return spec


def smooth_eps_spec(two_d=False, lossy=False, seed=7):
    """The busy problem with a smoothed random permittivity per E value (far more than 256
    distinct coefficients), lossless or with the busy problem's conductivity."""
    spec = dict(busy_spec(two_d))
    nx, ny, nz = spec["grid"][:3]
    rng = np.random.default_rng(seed)
    eps = 1.0 + 4.0 * rng.random((3, nz, ny, nx))
    spec["eps3n"] = eps
    if not lossy:
        spec["sigma"] = None
    return spec


def initial_fields(n):
    """0.01 noise in E and H, as the reference's check problem starts."""
    r = np.arange(n)
    e = [0.01 * noise(r + 3 * c) for c in range(3)]
    h = [0.01 * noise(r + 7 * c + 50_000) for c in range(3)]
    return e, h


def index(grid, at):
    return (at[2] * grid[1] + at[1]) * grid[0] + at[0]


def oracle(spec, with_fields=True):
    if spec.get("eps3n") is not None:  # a permittivity per E value instead of a scene
        s = O.Sim(spec["grid"], O.bounds(**spec["bounds"]), courant=spec["courant"], eps3n=spec["eps3n"])
    else:
        s = O.Sim(spec["grid"], O.bounds(**spec["bounds"]), O.scene(*spec["eps"]), spec["courant"])
    if spec.get("sigma"):
        s.conductivity(O.scene(*spec["sigma"]))
    n = s.n
    if with_fields:
        e, h = initial_fields(n)
        for c in range(3):
            s.set_field(0, c, e[c])
            s.set_field(1, c, h[c])
    if spec.get("conductor"):
        c, at = spec["conductor"]
        s.conductor(c, *at)
    for f, c, at, w in spec.get("sources", []):
        s.add_source(f, c, at, O.wave(**{k: v for k, v in w.items() if k != "kind"}, kind=w["kind"]))
    for f, vals, w in spec.get("currents", []):
        s.add_current(f, [v[0] for v in vals], [index(spec["grid"], v[1]) for v in vals],
                      [v[2] for v in vals], O.wave(**{k: v for k, v in w.items() if k != "kind"}, kind=w["kind"]))
    for lo, hi, d, pol, eps, w in spec.get("planes", []):
        s.add_plane_wave(lo, hi, d, pol, eps, O.wave(**{k: v for k, v in w.items() if k != "kind"}, kind=w["kind"]))
    for f, c, at in spec.get("probes", []):
        s.add_probe(f, c, at)
    if spec.get("dft"):
        lo, hi, fr = spec["dft"]
        s.add_dft(lo, hi, fr)
    for lo, hi, fr in spec.get("dfts", []):
        s.add_dft(lo, hi, fr)
    return s


def product(spec, with_fields=True):
    import photonoxide_amd as P
    g = P.Grid3d(*spec["grid"])
    b = spec["bounds"]
    edges = []
    for a in range(3):
        edges.append(P.Periodic() if b["periodic"][a] else P.Pml(b["low"][a], b["high"][a]))
    bd = P.Boundaries(edges[0], edges[1], edges[2],
                      P.Cpml(kappa=b.get("kappa", 1.0), alpha=b.get("alpha", 0.0)))

    def to_scene(sc):
        bg, shapes = sc
        return P.Scene(bg, [P.Box(lo, hi, v) if k == 0 else P.Cylinder((lo[0], lo[1]), hi[0], (lo[2], hi[2]), v)
                            for k, lo, hi, v in shapes])

    if spec.get("eps3n") is not None:
        s = P.Simulation.with_eps(g, spec["eps3n"], bd, spec["courant"])
    else:
        s = P.Simulation.new(g, to_scene(spec["eps"]), bd, spec["courant"])
    if spec.get("sigma"):
        s.with_conductivity(to_scene(spec["sigma"]))
    if with_fields:
        e, h = initial_fields(g.cells())
        for c in range(3):
            s.set_e(c, e[c])
            s.set_h(c, h[c])
    if spec.get("conductor"):
        c, at = spec["conductor"]
        s.conductor(c, at)
    for f, c, at, w in spec.get("sources", []):
        s.add_source(P.Source(P.Field(f), P.Axis(c), at, P.Waveform(w["kind"], w["frequency"], w["width"], w["delay"])))
    for f, vals, w in spec.get("currents", []):
        s.add_current(P.Current(P.Field(f), [(P.Axis(v[0]), v[1], v[2]) for v in vals],
                                P.Waveform(w["kind"], w["frequency"], w["width"], w["delay"])))
    for lo, hi, d, pol, eps, w in spec.get("planes", []):
        s.add_plane_wave(P.PlaneWave(lo, hi, d, pol, eps, P.Waveform(w["kind"], w["frequency"], w["width"], w["delay"])))
    for f, c, at in spec.get("probes", []):
        s.add_probe(f, c, at)
    if spec.get("dft"):
        lo, hi, fr = spec["dft"]
        s.add_dft(lo, hi, fr)
    for lo, hi, fr in spec.get("dfts", []):
        s.add_dft(lo, hi, fr)
    return s


def tfsf_spec(two_d=False, direction=None, polarization=None, scatterer=True, second=False):
    """A plane wave on a box in vacuum around a dielectric scatterer (eps 4) and a lossy one,
    CPMLs with kappa and alpha, a point source outside the box, probes in and out of it and a
    transform box across its face; optionally a second plane wave on the same box."""
    h = 0.05
    if two_d:
        n = (83, 70, 1)
        b = dict(periodic=(0, 0, 1), low=(8, 7, 0), high=(9, 7, 0), kappa=2.0, alpha=0.2)
        low, high = (20, 16, 0), (62, 52, 1)
        direction = direction or (2, 1, 0)
        polarization = polarization or (0, 0, 1)
    else:
        n = (45, 38, 33)
        b = dict(periodic=(0, 0, 0), low=(6, 5, 5), high=(6, 6, 5), kappa=2.0, alpha=0.2)
        low, high = (12, 11, 11), (33, 27, 22)
        direction = direction or (2, 1, 1)
        polarization = polarization or (1, -1, 0)
    grid = (n[0], n[1], n[2], h, h, h, -n[0] * h / 2, -n[1] * h / 2, -n[2] * h / 2)
    nx, ny, nz = n
    mid = (nx // 2, ny // 2, nz // 2)
    width = math.sqrt(2.0 * math.log(2.0)) / (math.pi * 1.5)  # Waveform::pulse(1.2, 1.5)
    pw = dict(kind=0, frequency=1.2, width=width, delay=6.0 * width)
    spec = dict(
        grid=grid, bounds=b, courant=0.9,
        eps=(1.0, [(0, (-0.25, -0.2, -INF if two_d else -0.15), (0.2, 0.25, INF if two_d else 0.15), 4.0)]
             if scatterer else []),
        sigma=(0.0, [(0, (0.25, -0.3, -INF if two_d else -0.1), (0.4, -0.1, INF if two_d else 0.1), 1.5)])
        if scatterer else None,
        sources=[(0, 2, (4 + b["low"][0], 3 + b["low"][1], 0 if two_d else nz // 2), pulse())],
        planes=[(low, high, direction, polarization, 1.0, pw)],
        probes=[(f, c, at) for at in [mid, (low[0] - 2, mid[1], mid[2]), (high[0] + 1, high[1], mid[2])]
                for (f, c) in [(0, 2), (1, 1), (0, 0)]],
        dft=((low[0] - 3, mid[1] - 4, 0 if two_d else mid[2] - 2), (low[0] + 3, mid[1] + 4, 0 if two_d else mid[2] + 2),
             [0.9, 1.2]),
    )
    if second:  # another wave on the same box, along another direction
        spec["planes"].append((low, high, (1, -3, 0), (0, 0, 1) if two_d else (1, 0, 1), 1.0, pulse()))
    return spec


def product_fields(s):
    return np.stack([s.e(c) for c in range(3)] + [s.h(c) for c in range(3)])
