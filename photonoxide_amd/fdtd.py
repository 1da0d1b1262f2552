"""photonoxide's FDTD interface (src/fdtd.rs), stepping on an MI355X through the native library.

The names, arguments and errors follow the reference's ``fdtd`` module: a :class:`Simulation` on
a :class:`Grid3d` inside :class:`Boundaries`, with :class:`Source`, :class:`Current` and
:class:`Dipole` sources, probes, :meth:`Simulation.add_dft` transform monitors and
:meth:`Simulation.with_medium` dispersive media, stepped by :meth:`Simulation.run` on the
device set by :meth:`Simulation.set_device`. Fields are numpy arrays in the reference's layout
(k, j, i with i fastest). There is no CPU stepping: ``run`` without a device raises.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field as dc_field
from enum import IntEnum
from typing import Optional, Sequence

import numpy as np

from . import _native as N


class Axis(IntEnum):
    X = 0
    Y = 1
    Z = 2

    def others(self):
        """The next two axes in cyclic order (src/fdfd/three.rs:52-58)."""
        return Axis((self + 1) % 3), Axis((self + 2) % 3)


class Field(IntEnum):
    E = 0
    H = 1


class Precision(IntEnum):
    F32 = N.F32
    F64 = N.F64


@dataclass(frozen=True)
class Grid3d:
    """A uniform grid of nx x ny x nz cells of dx x dy x dz um, low corner (x0, y0, z0)
    (src/fdfd/three.rs:70-89)."""
    nx: int
    ny: int
    nz: int
    dx: float
    dy: float
    dz: float
    x0: float = 0.0
    y0: float = 0.0
    z0: float = 0.0

    def cells(self) -> int:
        return self.nx * self.ny * self.nz

    def n(self, axis) -> int:
        return (self.nx, self.ny, self.nz)[axis]

    def step(self, axis) -> float:
        return (self.dx, self.dy, self.dz)[axis]

    def index(self, i, j, k) -> int:
        return (k * self.ny + j) * self.nx + i

    def e_position(self, component, i, j, k):
        p = [self.x0 + i * self.dx, self.y0 + j * self.dy, self.z0 + k * self.dz]
        p[component] += 0.5 * self.step(component)
        return p

    @staticmethod
    def centred(n, h):
        """The grid of n cells of h um centred on the origin (src/fdtd/checks.rs:8-20)."""
        return Grid3d(n[0], n[1], n[2], h, h, h, -n[0] * h / 2.0, -n[1] * h / 2.0, -n[2] * h / 2.0)


@dataclass(frozen=True)
class Pml:
    """A CPML of low and high cells inside the grid (0 cells: a wall)."""
    low: int = 0
    high: int = 0


@dataclass(frozen=True)
class Periodic:
    """A periodic side (a Bloch boundary with k = 0)."""


@dataclass(frozen=True)
class Cpml:
    """The convolutional PML's parameters (src/fdtd.rs:84-110): FDFD's R = 1e-8, order 3."""

    @staticmethod
    def sigma_optimal(order: float, eps: float, step: float) -> float:
        """Taflove's sigma_opt = (m + 1) / (150 pi sqrt(eps) step) in this build's units (sigma
        times eta0), src/fdtd.rs:115-117."""
        return (order + 1.0) * 376.730313668 / (150.0 * math.pi) / (math.sqrt(eps) * step)

    reflection: float = 1e-8
    sigma: Optional[float] = None
    order: float = 3.0
    kappa: float = 1.0
    alpha: float = 0.0


@dataclass(frozen=True)
class Boundaries:
    x: object = Pml()
    y: object = Pml()
    z: object = Pml()
    cpml: Cpml = Cpml()

    @staticmethod
    def cpml_all(cells: int) -> "Boundaries":
        e = Pml(cells, cells)
        return Boundaries(e, e, e)

    @staticmethod
    def walls() -> "Boundaries":
        return Boundaries.cpml_all(0)

    @staticmethod
    def cpml_2d(cells: int) -> "Boundaries":
        e = Pml(cells, cells)
        return Boundaries(e, e, Periodic())

    def native(self) -> N.Bounds:
        b = N.Bounds()
        for a, e in enumerate((self.x, self.y, self.z)):
            if isinstance(e, Periodic):
                b.periodic[a] = 1
            else:
                b.low[a], b.high[a] = e.low, e.high
        c = self.cpml
        b.reflection, b.order, b.kappa, b.alpha = c.reflection, c.order, c.kappa, c.alpha
        b.has_sigma = 0 if c.sigma is None else 1
        b.sigma = 0.0 if c.sigma is None else c.sigma
        return b


class Waveform:
    """How a source's value varies in time (src/fdtd.rs:259-424)."""

    def __init__(self, kind, frequency=0.0, width=0.0, delay=0.0, amplitude=1 + 0j, ramp=0.0):
        self.kind, self.frequency, self.width, self.delay = kind, frequency, width, delay
        self.amplitude, self.ramp = complex(amplitude), ramp

    @staticmethod
    def gaussian(frequency, width, delay):
        return Waveform(0, frequency, width, delay)

    @staticmethod
    def differentiated_gaussian(width, delay):
        return Waveform(1, 0.0, width, delay)

    @staticmethod
    def continuous(frequency, amplitude=1 + 0j, ramp=0.0):
        return Waveform(2, frequency, 0.0, 0.0, amplitude, ramp)

    @staticmethod
    def pulse(frequency, bandwidth):
        """A Gaussian on a carrier whose power spectrum is `bandwidth` wide at half its peak:
        width sqrt(2 ln 2)/(pi bandwidth), its peak 6 widths in (src/fdtd.rs:300-312)."""
        if not (math.isfinite(bandwidth) and bandwidth > 0):
            raise ValueError(f"a pulse's bandwidth must be positive and finite, got {bandwidth}")
        width = math.sqrt(2.0 * math.log(2.0)) / (math.pi * bandwidth)
        return Waveform(0, frequency, width, 6.0 * width)

    def native(self) -> N.Wave:
        return N.Wave(self.kind, self.frequency, self.width, self.delay,
                      self.amplitude.real, self.amplitude.imag, self.ramp)


@dataclass
class Source:
    field: Field
    component: Axis
    at: tuple
    waveform: Waveform


@dataclass
class Current:
    """A current over many values: (component, (i, j, k), complex amplitude) each."""
    field: Field
    values: list
    waveform: Waveform


@dataclass
class Dipole:
    field: Field
    component: Axis
    position: tuple
    amplitude: complex
    waveform: Waveform


@dataclass
class PlaneWave:
    """A plane wave on a total-field/scattered-field box (src/fdtd/sources.rs PlaneWave): the
    box of nodes low to high, the integer direction, E along the polarization's part across it,
    in a uniform background of permittivity eps."""
    low: tuple
    high: tuple
    direction: tuple
    polarization: tuple
    eps: float
    waveform: Waveform


@dataclass
class PortMode:
    """A port's mode as the reference's PortMode3d holds it (src/fdfd/three/ports.rs:46-60), from
    a mode solver: the waveguide's axis, the plane of nodes it was solved on, the step along the
    axis and the wavenumber k0 it was solved at, its propagation constant beta (rad/um), E on
    the plane across the axis as `tangential[2][size_c][size_b]` (E_b then E_c, b = (axis + 1)
    % 3, c = (axis + 2) % 3) and E_a on the half-plane after it as `normal[size_c][size_b]`,
    going forward at unit amplitude; `power`, the scheme's flux at unit amplitude, when a solver
    gave it."""
    axis: int
    plane: int
    step: float
    k0: float
    beta: complex
    tangential: np.ndarray
    normal: np.ndarray
    power: Optional[float] = None


@dataclass
class Resonance:
    """One term of a signal sum_k d_k e^(-i w_k t) (src/fdtd/harmonic.rs:24-46): Re w / 2 pi
    (c/um), the decay rate -Im w (per um/c), the amplitude d at the first sample, and the
    filter diagonalization's error (near round-off for a term the signal holds)."""
    frequency: float
    decay: float
    amplitude: complex
    error: float

    def q(self) -> float:
        """Q = w / (2 gamma)."""
        return math.pi * self.frequency / self.decay


def harmonic_inversion(signal, dt: float, low: float, high: float):
    """harmonic_inversion (src/fdtd/harmonic.rs:62-184): the terms of `signal` (sampled every
    `dt`) with frequencies in [low, high], by filter diagonalization, sorted by frequency."""
    c = np.ascontiguousarray(np.asarray(signal, dtype=np.complex128))
    flat = c.view(np.float64).ravel()
    capacity = 64
    while True:
        out = np.zeros(5 * capacity)
        found = C.c_int32()
        N.check(N.lib.pxfdtd_harmonic_inversion(N.dptr(flat), len(c), float(dt), float(low), float(high),
                                                capacity, C.byref(found), N.dptr(out)))
        if found.value <= capacity:
            break
        capacity = found.value
    return [Resonance(float(out[5 * q]), float(out[5 * q + 1]), complex(out[5 * q + 2], out[5 * q + 3]),
                      float(out[5 * q + 4])) for q in range(found.value)]


def harmonic_inversion_device(gpu: "Gpu", signal_ptr, n: int, dt: float, low: float, high: float, stream=None):
    """harmonic_inversion of n complex FP64 samples held in `gpu`'s memory at `signal_ptr` (a
    device address): the sums over the samples run on the device, only they come back
    (pxfdtd_harmonic_inversion_device); the same resonances as harmonic_inversion."""
    capacity = 64
    while True:
        out = np.zeros(5 * capacity)
        found = C.c_int32()
        N.check(N.lib.pxfdtd_harmonic_inversion_device(gpu.handle, signal_ptr, int(n), float(dt), float(low),
                                                       float(high), capacity, C.byref(found), N.dptr(out), stream))
        if found.value <= capacity:
            break
        capacity = found.value
    return [Resonance(float(out[5 * q]), float(out[5 * q + 1]), complex(out[5 * q + 2], out[5 * q + 3]),
                      float(out[5 * q + 4])) for q in range(found.value)]


def _port_mode(mode: PortMode):
    """The C ABI's pxfdtd_port_mode for `mode`, and the arrays it points into (kept alive by
    the caller while the struct is in use)."""
    tan = np.ascontiguousarray(mode.tangential, dtype=np.complex128)
    nor = np.ascontiguousarray(mode.normal, dtype=np.complex128)
    if tan.ndim != 3 or tan.shape[0] != 2 or nor.shape != tan.shape[1:]:
        raise ValueError("a mode's tangential E is [2][size_c][size_b] and its normal E [size_c][size_b]")
    tv = tan.view(np.float64).ravel()
    nv = nor.view(np.float64).ravel()
    beta = complex(mode.beta)
    m = N.PortMode(int(mode.axis), int(mode.plane), (C.c_int32 * 2)(tan.shape[2], tan.shape[1]),
                   float(mode.step), float(mode.k0), (C.c_double * 2)(beta.real, beta.imag),
                   N.dptr(tv), N.dptr(nv))
    return m, (tv, nv)


@dataclass
class GaussianBeam:
    """A Gaussian beam (src/fdtd/sources.rs:104-128): launched from the plane of nodes `plane`
    normal to `axis`, going `direction` (0 forward, 1 backward), its axis crossing the plane at
    `centre` (um, along the other two axes), waist w0 (um) at `focus` um along the beam, tilted
    by `angle` radians towards `tilt`, E perpendicular (0) or parallel (1) to the plane of
    incidence, in a medium of permittivity `eps` on the plane."""
    axis: int
    plane: int
    waist: float
    focus: float = 0.0
    centre: tuple = (0.0, 0.0)
    direction: int = 0
    tilt: int = -1
    angle: float = 0.0
    polarization: int = 0
    eps: float = 1.0

    def native(self) -> N.GaussianBeam:
        tilt = (int(self.axis) + 1) % 3 if self.tilt < 0 else int(self.tilt)
        return N.GaussianBeam(int(self.axis), int(self.plane), int(self.direction),
                              (C.c_double * 2)(*self.centre), float(self.waist), float(self.focus), tilt,
                              float(self.angle), int(self.polarization), float(self.eps))


@dataclass
class Box:
    """An open box lo < p < hi (inf allowed) with a diagonal permittivity (or one value)."""
    lo: tuple
    hi: tuple
    value: object

    def native(self) -> N.Shape:
        v = self.value if isinstance(self.value, (tuple, list)) else (self.value,) * 3
        return N.Shape(0, N.d3(*self.lo), N.d3(*self.hi), N.d3(*v))


@dataclass
class Cylinder:
    """A cylinder along z: (x - cx)^2 + (y - cy)^2 < r^2 and zlo < z < zhi."""
    centre: tuple
    radius: float
    z: tuple
    value: object

    def native(self) -> N.Shape:
        v = self.value if isinstance(self.value, (tuple, list)) else (self.value,) * 3
        return N.Shape(1, N.d3(self.centre[0], self.centre[1], self.z[0]),
                       N.d3(self.radius, 0.0, self.z[1]), N.d3(*v))


class Scene:
    """A background value and regions, later ones replacing earlier ones."""

    def __init__(self, background, shapes=()):
        bg = background if isinstance(background, (tuple, list)) else (background,) * 3
        self.background = tuple(float(b) for b in bg)
        self.shapes = list(shapes)

    def native(self):
        arr = (N.Shape * max(1, len(self.shapes)))(*[s.native() for s in self.shapes])
        sc = N.Scene(N.d3(*self.background), len(self.shapes), arr)
        return sc, arr  # keep arr alive with the scene


class Gpu:
    """A HIP device (replaces gpu::Gpu, src/fdtd/gpu.rs:84-234)."""

    def __init__(self, ordinal: int = 0):
        h = C.c_void_p()
        N.check(N.lib.pxfdtd_device_open(ordinal, C.byref(h)))
        self.handle = h

    def info(self) -> dict:
        i = N.DeviceInfo()
        N.check(N.lib.pxfdtd_device_query(self.handle, C.byref(i)))
        return {"name": i.name.decode(), "arch": i.arch.decode(), "ordinal": i.ordinal,
                "compute_units": i.compute_units, "hbm_bytes": i.hbm_bytes,
                "free_bytes": i.free_bytes}

    def hbm_rates(self, nbytes: int = 1 << 30, reps: int = 10) -> dict:
        """The device's achievable HBM rates in GB/s (bytes read plus written), from streaming
        kernels over buffers of nbytes each: read, write, copy and triad."""
        out = (C.c_double * 4)()
        N.check(N.lib.pxfdtd_device_hbm_rates(self.handle, nbytes, reps, C.byref(out)))
        return dict(zip(("read", "write", "copy", "triad"), (float(v) for v in out)))

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            N.lib.pxfdtd_device_close(h)
            self.handle = None


@dataclass
class Design:
    """A design region (src/fdtd/adjoint.rs:63-75): the cells [lo, hi) along x, y and z, each with
    a density rho and the permittivity eps[0] + rho (eps[1] - eps[0]); densities are numbered i
    fastest, then j, then k."""
    cells: tuple  # ((lo, hi), (lo, hi), (lo, hi))
    eps: tuple    # (eps at density 0, eps at density 1)

    def __len__(self):
        return int(np.prod([hi - lo for lo, hi in self.cells]))

    @property
    def shape(self):
        """The densities' array shape, (nz, ny, nx) of the region."""
        return tuple(hi - lo for lo, hi in reversed(self.cells))

    def native(self) -> N.Design:
        return N.Design(N.i3(*[c[0] for c in self.cells]), N.i3(*[c[1] for c in self.cells]),
                        (C.c_double * 2)(*self.eps))


@dataclass
class Term:
    """What an objective's derivative dF/da is taken of (Term, src/fdtd/adjoint.rs:285-338), with
    dF/da the Wirtinger derivative (conj(a) held; conj(a) for F = |a|^2). Build with
    Term.mode, Term.flux or Term.transform."""
    kind: int
    monitor: int
    derivative: complex
    mode: int = 0
    direction: int = 0
    frequency: int = 0
    field: int = 0
    component: int = 0
    at: tuple = (0, 0, 0)

    @staticmethod
    def mode_amplitude(monitor: int, mode: int, direction: int, derivative: complex) -> "Term":
        """A mode monitor's amplitude of `mode` going `direction` (0 forward, 1 backward)."""
        return Term(0, monitor, complex(derivative), mode=mode, direction=direction)

    @staticmethod
    def flux(monitor: int, frequency: int, derivative: float) -> "Term":
        """A flux monitor's flux at its frequency number `frequency`; dF/dP real."""
        return Term(1, monitor, complex(derivative), frequency=frequency)

    @staticmethod
    def transform(monitor: int, field, component, at, frequency: int, derivative: complex) -> "Term":
        """A transform monitor's value of field's component at `at`, frequency number `frequency`."""
        return Term(2, monitor, complex(derivative), frequency=frequency, field=int(field), component=int(component),
                    at=tuple(at))

    def native(self) -> N.Term:
        return N.Term(self.kind, self.monitor, self.mode, self.direction, self.frequency, self.field, self.component,
                      N.i3(*self.at), self.derivative.real, self.derivative.imag)


class Simulation:
    """An FDTD problem (src/fdtd.rs:470-512) on the host, stepping on a device."""

    def __init__(self, handle, grid: Grid3d, boundaries: Boundaries):
        self._h = handle
        self.grid = grid
        self.boundaries = boundaries
        self._device = None

    # -- construction ---------------------------------------------------------------------
    @staticmethod
    def new(grid: Grid3d, eps, boundaries: Boundaries, courant: float, device: "Gpu" = None) -> "Simulation":
        """The permittivity averaged per component over 8^3 samples of each cell
        (Simulation::new / diagonal, src/fdtd.rs:570-653). `eps` is a :class:`Scene`. With a
        `device`, the averaging runs there (the same bits); the simulation still has no device
        to step on until set_device."""
        h = C.c_void_p()
        scene, keep = eps.native()
        if device is not None:
            N.check(N.lib.pxfdtd_simulation_new_on_device(device.handle, C.byref(_grid(grid)),
                                                          C.byref(boundaries.native()), C.byref(scene),
                                                          courant, C.byref(h)))
        else:
            N.check(N.lib.pxfdtd_simulation_new(C.byref(_grid(grid)), C.byref(boundaries.native()),
                                                C.byref(scene), courant, C.byref(h)))
        return Simulation(h, grid, boundaries)

    @staticmethod
    def with_eps(grid: Grid3d, eps3n, boundaries: Boundaries, courant: float) -> "Simulation":
        """The permittivity each value of E sees, given per component as [3][nz][ny][nx]."""
        e = np.ascontiguousarray(np.broadcast_to(np.asarray(eps3n, dtype=np.float64),
                                                 (3, grid.nz, grid.ny, grid.nx)))
        h = C.c_void_p()
        N.check(N.lib.pxfdtd_simulation_new_eps(C.byref(_grid(grid)), C.byref(boundaries.native()),
                                                N.dptr(e), courant, C.byref(h)))
        return Simulation(h, grid, boundaries)

    def with_conductivity(self, sigma: Scene) -> "Simulation":
        sc, keep = sigma.native()
        N.check(N.lib.pxfdtd_simulation_conductivity(self._h, C.byref(sc)))
        return self

    def with_medium(self, eps_inf: float, poles: Sequence[tuple], inside: Scene) -> "Simulation":
        """A dispersive medium (src/fdtd/media.rs:733-825). Each pole is
        ('lorentz', strength, resonance, damping) or ('drude', plasma, damping)."""
        arr = np.zeros((max(1, len(poles)), 4))
        for q, p in enumerate(poles):
            if p[0] == "lorentz":
                arr[q] = (0.0, p[1], p[2], p[3])
            elif p[0] == "drude":
                arr[q] = (1.0, p[1], p[2], 0.0)
            else:
                raise ValueError(f"unknown pole {p!r}")
        sc, keep = inside.native()
        N.check(N.lib.pxfdtd_simulation_add_medium(self._h, eps_inf, len(poles), N.dptr(arr),
                                                   C.byref(sc)))
        return self

    def conductor(self, component, at):
        N.check(N.lib.pxfdtd_simulation_conductor(self._h, int(component), *at))

    def plate(self, normal, node: int, first: range, second: range):
        """A perfectly conducting plate on the plane of nodes `node` normal to `normal`, over
        cells `first` and `second` along the other two axes (Simulation::plate,
        src/fdtd.rs:1013-1047): E along each in-plane axis zeroed on the cells along it and on
        the nodes (edges included) along the other, within the grid."""
        a, b = (int(normal) + 1) % 3, (int(normal) + 2) % 3
        n = (self.grid.nx, self.grid.ny, self.grid.nz)

        def put(component, ra, rb):
            for ia in ra:
                for ib in rb:
                    at = [0, 0, 0]
                    at[int(normal)], at[a], at[b] = node, ia, ib
                    if at[0] < n[0] and at[1] < n[1] and at[2] < n[2]:
                        self.conductor(component, tuple(at))
        put(a, range(first.start, max(first.stop - 1, 0) + 1), range(second.start, second.stop + 1))
        put(b, range(first.start, first.stop + 1), range(second.start, max(second.stop - 1, 0) + 1))

    def add_source(self, source: Source):
        N.check(N.lib.pxfdtd_simulation_add_source(self._h, int(source.field), int(source.component),
                                                   *source.at, C.byref(source.waveform.native())))

    def add_current(self, current: Current):
        n = len(current.values)
        comps = np.array([int(v[0]) for v in current.values], dtype=np.int32)
        idx = np.array([self.grid.index(*v[1]) for v in current.values], dtype=np.int64)
        for v in current.values:
            i, j, k = v[1]
            if not (0 <= i < self.grid.nx and 0 <= j < self.grid.ny and 0 <= k < self.grid.nz):
                raise N.NativeError(N.ERR_INVALID, f"{v[1]} is outside the grid")
        amps = np.array([[complex(v[2]).real, complex(v[2]).imag] for v in current.values],
                        dtype=np.float64).reshape(n, 2)
        N.check(N.lib.pxfdtd_simulation_add_current(
            self._h, int(current.field), n, comps.ctypes.data_as(N.pi32),
            idx.ctypes.data_as(N.pi64), N.dptr(amps), C.byref(current.waveform.native())))

    def add_current_indices(self, field, comps, index, amps, waveform: Waveform):
        """A current given as arrays: components, flat indices within a component, complex
        amplitudes (the form a mode source's restriction takes)."""
        comps = np.ascontiguousarray(comps, dtype=np.int32)
        index = np.ascontiguousarray(index, dtype=np.int64)
        a = np.ascontiguousarray(np.stack([np.real(amps), np.imag(amps)], axis=-1), dtype=np.float64)
        N.check(N.lib.pxfdtd_simulation_add_current(
            self._h, int(field), len(index), comps.ctypes.data_as(N.pi32),
            index.ctypes.data_as(N.pi64), N.dptr(a), C.byref(waveform.native())))

    def add_dipole(self, dipole: Dipole):
        pos = np.array(dipole.position, dtype=np.float64)
        a = complex(dipole.amplitude)
        N.check(N.lib.pxfdtd_simulation_add_dipole(self._h, int(dipole.field), int(dipole.component),
                                                   N.dptr(pos), a.real, a.imag,
                                                   C.byref(dipole.waveform.native())))

    def add_flux(self, axis, plane: int, frequencies, window=(None, None)) -> int:
        """Simulation::add_flux (src/fdtd/monitors.rs:470-491): the flux towards +axis through
        the plane halfway after node `plane`; window[w] = (lo, hi) half-planes along the w-th
        of the axes after `axis`, or None for the whole axis."""
        win = (C.c_int32 * 2 * 2)(*[(C.c_int32 * 2)(*(w if w is not None else (-1, -1))) for w in window])
        f = np.ascontiguousarray(frequencies, dtype=np.float64)
        n = C.c_int32()
        N.check(N.lib.pxfdtd_simulation_add_flux(self._h, int(axis), int(plane), win, len(f), N.dptr(f),
                                                 C.byref(n)))
        return n.value

    def add_flux_box(self, planes, frequencies) -> int:
        """Simulation::add_flux_box (monitors.rs:493-528): the flux out of the box between the
        half-planes after nodes planes[a][0] and planes[a][1] along each axis a."""
        pl = (C.c_int32 * 2 * 3)(*[(C.c_int32 * 2)(*p) for p in planes])
        f = np.ascontiguousarray(frequencies, dtype=np.float64)
        n = C.c_int32()
        N.check(N.lib.pxfdtd_simulation_add_flux_box(self._h, pl, len(f), N.dptr(f), C.byref(n)))
        return n.value

    def flux(self, n: int) -> np.ndarray:
        """Simulation::flux (monitors.rs:530-547): per frequency."""
        m = N.lib.pxfdtd_simulation_flux_frequencies(self._h, n)
        if m < 0:
            raise IndexError(f"no flux monitor {n}")
        out = np.zeros(m)
        N.check(N.lib.pxfdtd_simulation_flux(self._h, n, N.dptr(out)))
        return out

    def add_plane_wave(self, wave: PlaneWave) -> int:
        """Simulation::add_plane_wave (src/fdtd/sources.rs:701-721): its number."""
        n = C.c_int32()
        N.check(N.lib.pxfdtd_simulation_add_plane_wave(
            self._h, (C.c_int32 * 3)(*wave.low), (C.c_int32 * 3)(*wave.high),
            (C.c_int64 * 3)(*wave.direction), (C.c_double * 3)(*wave.polarization), float(wave.eps),
            C.byref(wave.waveform.native()), C.byref(n)))
        return n.value

    def add_mode_source(self, mode: PortMode, direction: int, waveform: Waveform):
        """Simulation::add_mode_source (src/fdtd/sources.rs:550-596): the mode launched going
        `direction` (0 forward, 1 backward) along its axis by total-field/scattered-field on its
        plane, at the waveform's carrier, as fixed currents."""
        m, keep = _port_mode(mode)
        N.check(N.lib.pxfdtd_simulation_add_mode_source(self._h, C.byref(m), int(direction),
                                                        C.byref(waveform.native())))
        del keep

    def port_mode(self, axis: int, plane: int, frequency: float, te: bool = True, window=(None, None)) -> PortMode:
        """The fundamental mode of one polarization, as the reference's FDTD job picks it
        (src/job/fdtd.rs:2097-2135): of the 4 modes port_modes gives, the first whose E lies
        mostly (te) or not (tm) along the in-plane axis across the guide (y for a port normal
        to x, else x), by the share of |E_t|^2."""
        modes = self.port_modes(axis, plane, frequency, count=4, window=window)
        b, c = (axis + 1) % 3, (axis + 2) % 3
        across = 1 if axis == 0 else 0
        for m in modes:
            e = np.abs(m.tangential) ** 2  # [t][v][u], t = 0 along b, 1 along c
            along = e[0].sum() if b == across else e[1].sum()
            share = along / max(e.sum(), np.finfo(float).tiny)
            if (share > 0.5) == te:
                return m
        raise ValueError(f"no {'TE' if te else 'TM'} mode guides on the plane at node {plane} along axis {axis}")

    def permittivity(self) -> np.ndarray:
        """The permittivity each value of E sees: [3][nz][ny][nx]."""
        g = self.grid
        out = np.zeros(3 * g.nx * g.ny * g.nz)
        N.check(N.lib.pxfdtd_simulation_permittivity(self._h, N.dptr(out)))
        return out.reshape(3, g.nz, g.ny, g.nx)

    def port_modes(self, axis: int, plane: int, frequency: float, count: int = 1, window=(None, None),
                   lossy: bool = False):
        """Lattice::port_modes (src/fdfd/three/ports.rs:408-594) on this grid and permittivity:
        up to `count` modes of the guide crossing node plane `plane` normal to `axis`, at the
        leapfrog wavenumber of `frequency` (the k0 a mode source at that frequency needs),
        within window[w] = (lo, hi) nodes along the w-th of the axes after `axis` (None: whole);
        propagating modes first by |Re beta|, each normalized to N(f, g) = 1; `lossy`: with the
        conductivity as the leapfrog sees it, eps + i sigma cos(w dt / 2) / k0."""
        g = self.grid
        n = (g.nx, g.ny, g.nz)
        nb, nc = n[(axis + 1) % 3], n[(axis + 2) % 3]
        per = nb * nc
        win = (C.c_int32 * 2 * 2)(*[(C.c_int32 * 2)(*(w if w is not None else (-1, -1))) for w in window])
        beta = np.zeros(2 * count)
        tan = np.zeros(count * 4 * per)
        nor = np.zeros(count * 2 * per)
        power = np.zeros(count)
        found = C.c_int32()
        N.check(N.lib.pxfdtd_simulation_port_modes(self._h, int(axis), int(plane), C.cast(win, C.c_void_p),
                                                   float(frequency), int(count), int(bool(lossy)), C.byref(found),
                                                   N.dptr(beta),
                                                   N.dptr(tan), N.dptr(nor), N.dptr(power)))
        k0 = self.leapfrog_wavenumber(frequency)
        step = (g.dx, g.dy, g.dz)[axis]
        out = []
        for q in range(found.value):
            t = tan[q * 4 * per:(q + 1) * 4 * per].view(np.complex128).reshape(2, nc, nb).copy()
            m = nor[q * 2 * per:(q + 1) * 2 * per].view(np.complex128).reshape(nc, nb).copy()
            out.append(PortMode(int(axis), int(plane), step, k0, complex(beta[2 * q], beta[2 * q + 1]), t, m,
                                float(power[q])))
        return out

    def add_mode_monitor(self, modes) -> int:
        """Simulation::add_mode_monitor (src/fdtd/monitors.rs:569-611): the amplitudes of
        `modes` (each solved at the leapfrog wavenumber of the frequency it is wanted at) from
        the next step on, by the transform of E over each mode's plane and the next. Its number
        for mode_amplitudes; mode monitors are numbered on their own."""
        modes = list(modes)
        if not modes:
            raise ValueError("a mode monitor needs at least one mode")
        packed = [_port_mode(m) for m in modes]
        arr = (N.PortMode * len(packed))(*[p[0] for p in packed])
        n = C.c_int32()
        N.check(N.lib.pxfdtd_simulation_add_mode_monitor(self._h, arr, len(packed), C.byref(n)))
        del packed
        return n.value

    def mode_box(self, mode: PortMode):
        """The nodes (low, high) of the box a mode's amplitude reads: its plane and the next along
        its axis, the whole grid across (for add_dft and mode_weights)."""
        g = self.grid
        lo, hi = [0, 0, 0], [g.nx - 1, g.ny - 1, g.nz - 1]
        lo[mode.axis], hi[mode.axis] = mode.plane, mode.plane + 1
        return tuple(lo), tuple(hi)

    def mode_weights(self, mode: PortMode, direction: int = 0) -> np.ndarray:
        """Lattice::mode_amplitude_weights (src/fdfd/three/ports.rs:293-345): the amplitude of
        `mode` going `direction` as weights on E's transform over mode_box(mode), [3][k][j][i]
        complex: the amplitude is sum(weights * transform)."""
        lo, hi = self.mode_box(mode)
        shape = (3,) + tuple(h - l + 1 for l, h in zip(lo[::-1], hi[::-1]))
        out = np.zeros(2 * int(np.prod(shape)))
        m, keep = _port_mode(mode)
        N.check(N.lib.pxfdtd_simulation_mode_weights(self._h, C.byref(m), int(direction), N.dptr(out)))
        del keep
        return out.view(np.complex128).reshape(shape)

    def mode_amplitudes(self, n: int):
        """Simulation::mode_amplitudes (src/fdtd/monitors.rs:613-640): (forward, backward) of
        each of monitor n's modes, in the transform's units."""
        count = N.lib.pxfdtd_simulation_mode_count(self._h, n)
        if count < 0:
            raise IndexError(f"no mode monitor {n}")
        out = np.zeros(4 * count)
        N.check(N.lib.pxfdtd_simulation_mode_amplitudes(self._h, n, N.dptr(out)))
        return [(complex(out[4 * q], out[4 * q + 1]), complex(out[4 * q + 2], out[4 * q + 3]))
                for q in range(count)]

    def mode_frequencies(self, n: int):
        """Simulation::mode_frequencies (src/fdtd/monitors.rs:653-663): each mode's frequency."""
        count = N.lib.pxfdtd_simulation_mode_count(self._h, n)
        if count < 0:
            raise IndexError(f"no mode monitor {n}")
        out = np.zeros(count)
        N.check(N.lib.pxfdtd_simulation_mode_frequencies(self._h, n, N.dptr(out)))
        return [float(v) for v in out]

    def add_beam(self, beam: GaussianBeam, waveform: Waveform):
        """Simulation::add_beam (src/fdtd/sources.rs:618-699): the beam launched one way from its
        plane by total-field/scattered-field at the waveform's carrier, as fixed currents."""
        N.check(N.lib.pxfdtd_simulation_add_beam(self._h, C.byref(beam.native()), C.byref(waveform.native())))

    def leapfrog_wavenumber(self, frequency: float) -> float:
        """The wavenumber a source at `frequency` settles to on the grid, (2/dt) sin(w dt/2)
        (src/fdtd/sources.rs:171-173), the k0 a mode for it is solved at."""
        return 2.0 / self.dt() * math.sin(2.0 * math.pi * frequency * self.dt() / 2.0)

    def incident(self, n: int, field, component, at) -> float:
        """Plane wave n's incident field now at a value (Simulation::incident, sources.rs:723-737)."""
        v = C.c_double()
        N.check(N.lib.pxfdtd_simulation_incident(self._h, n, int(field), int(component), *at, C.byref(v)))
        return v.value

    def add_probe(self, field, component, at) -> int:
        n = C.c_int32()
        N.check(N.lib.pxfdtd_simulation_add_probe(self._h, int(field), int(component), *at, C.byref(n)))
        return n.value

    def probe(self, n: int) -> np.ndarray:
        m = N.lib.pxfdtd_simulation_probe_len(self._h, n)
        if m < 0:
            raise IndexError(f"no probe {n}")
        out = np.zeros(m)
        N.check(N.lib.pxfdtd_simulation_probe(self._h, n, N.dptr(out)))
        return out

    def add_dft(self, low, high, frequencies) -> int:
        f = np.ascontiguousarray(frequencies, dtype=np.float64)
        lo = (C.c_int32 * 3)(*low)
        hi = (C.c_int32 * 3)(*high)
        n = C.c_int32()
        N.check(N.lib.pxfdtd_simulation_add_dft(self._h, lo, hi, len(f), N.dptr(f), C.byref(n)))
        self._dft_shapes = getattr(self, "_dft_shapes", [])
        self._dft_shapes.append((len(f), tuple(h - l + 1 for l, h in zip(low, high))))
        return n.value

    def dft(self, n: int, field, component) -> np.ndarray:
        """Monitor n's transform of field's component: [f][k][j][i] complex."""
        series = int(field) * 3 + int(component)
        m = N.lib.pxfdtd_simulation_dft_len(self._h, n, series)
        if m < 0:
            raise IndexError(f"no transform {n}")
        out = np.zeros(m)
        N.check(N.lib.pxfdtd_simulation_dft(self._h, n, series, N.dptr(out)))
        nf, (lx, ly, lz) = self._dft_shapes[n]
        return (out[0::2] + 1j * out[1::2]).reshape(nf, lz, ly, lx)

    # -- designs and the frequency-domain adjoint -------------------------------------------
    def with_design(self, design: Design, density, background: Scene) -> "Simulation":
        """Simulation::with_design (src/fdtd/adjoint.rs:366-423): every value of E whose cell
        reaches the design's cells takes the average of eps(rho) there and `background`
        elsewhere. `density` is one value per cell, [nz][ny][nx] of the region or flat."""
        d = np.ascontiguousarray(np.asarray(density, dtype=np.float64).ravel())
        sc, keep = background.native()
        N.check(N.lib.pxfdtd_simulation_with_design(self._h, C.byref(design.native()), N.dptr(d), len(d),
                                                    C.byref(sc)))
        return self

    def add_design_monitor(self, design: Design, frequencies) -> int:
        """Simulation::add_design_monitor (adjoint.rs:435-470): E's transforms over the values the
        design's cells reach, at the objective's frequencies."""
        f = np.ascontiguousarray(np.atleast_1d(np.asarray(frequencies, dtype=np.float64)))
        n = C.c_int32()
        N.check(N.lib.pxfdtd_simulation_add_design_monitor(self._h, C.byref(design.native()), len(f), N.dptr(f),
                                                           C.byref(n)))
        return n.value

    def design_values(self, n: int) -> np.ndarray:
        """The values design monitor n reads, [(component, i, j, k)], in the gradients' order."""
        m = N.lib.pxfdtd_simulation_design_values(self._h, n, None)
        if m < 0:
            raise IndexError(f"no design monitor {n}")
        out = np.zeros((m, 4), dtype=np.int32)
        N.lib.pxfdtd_simulation_design_values(self._h, n, out.ctypes.data_as(N.pi32))
        return out

    def design_transform(self, n: int, frequency: int = 0) -> np.ndarray:
        """Design monitor n's transform of each value it reads at its frequency number `frequency`."""
        m = N.lib.pxfdtd_simulation_design_values(self._h, n, None)
        if m < 0:
            raise IndexError(f"no design monitor {n}")
        out = np.zeros(2 * m)
        N.check(N.lib.pxfdtd_simulation_design_transform(self._h, n, frequency, N.dptr(out)))
        return out[0::2] + 1j * out[1::2]

    def adjoint(self, terms: Sequence[Term], monitor: int, bandwidth: float) -> "Simulation":
        """Simulation::adjoint (adjoint.rs:496-569): the adjoint run, at rest, whose currents'
        transforms at design monitor `monitor`'s frequencies are the objective's derivatives.
        It steps on this simulation's device; run both until their fields die away
        (run_until_converged), then gradient()."""
        arr = (N.Term * max(1, len(terms)))(*[t.native() for t in terms])
        h = C.c_void_p()
        N.check(N.lib.pxfdtd_simulation_adjoint(self._h, arr, len(terms), monitor, float(bandwidth), C.byref(h)))
        adj = Simulation(h, self.grid, self.boundaries)
        adj._device = self._device
        return adj

    def permittivity_gradient(self, adjoint: "Simulation", monitor: int) -> np.ndarray:
        """Simulation::permittivity_gradient (adjoint.rs:831-855): dF/deps per value of
        design_values(monitor)."""
        m = N.lib.pxfdtd_simulation_design_values(self._h, monitor, None)
        if m < 0:
            raise IndexError(f"no design monitor {monitor}")
        out = np.zeros(m)
        N.check(N.lib.pxfdtd_simulation_permittivity_gradient(self._h, adjoint._h, monitor, N.dptr(out)))
        return out

    def gradient(self, adjoint: "Simulation", monitor: int, design: Design) -> np.ndarray:
        """Simulation::gradient (adjoint.rs:857-879): dF/drho per design cell, [nz][ny][nx] of
        the region."""
        out = np.zeros(len(design))
        N.check(N.lib.pxfdtd_simulation_design_gradient(self._h, adjoint._h, monitor, N.dptr(out)))
        return out.reshape(design.shape)

    def run_until_converged(self, tolerance: float, interval: float, limit: float) -> bool:
        """Simulation::run_until_converged (adjoint.rs:904-964): whether every transform had
        settled to `tolerance` of its largest value over a block of `interval` before `limit`."""
        s = C.c_int32()
        N.check(N.lib.pxfdtd_simulation_run_until_converged(self._h, float(tolerance), float(interval), float(limit),
                                                            C.byref(s)))
        return bool(s.value)

    # -- stepping -------------------------------------------------------------------------
    def set_device(self, device: Optional[Gpu], precision: Precision = Precision.F32):
        """Steps on `device` from now on (src/fdtd.rs:1517-1540); refused problems leave none."""
        N.check(N.lib.pxfdtd_simulation_set_device(self._h, device.handle if device else None,
                                                   int(precision)))
        self._device = device

    def run(self, count: int):
        N.check(N.lib.pxfdtd_simulation_run(self._h, int(count)))

    def bytes_per_step(self, precision=None) -> int:
        """The reference's byte model of a step (Yee::bytes_per_step, src/fdtd/kernel/yee.rs:141-153)."""
        b = C.c_int64()
        N.check(N.lib.pxfdtd_simulation_bytes_per_step(self._h, int(Precision.F32 if precision is None else precision),
                                                       C.byref(b)))
        return b.value

    def run_until(self, time: float):
        """Steps until E's time reaches `time` (src/fdtd.rs:1160-1167)."""
        N.check(N.lib.pxfdtd_simulation_run_until(self._h, float(time)))

    def run_until_decayed(self, probes, fraction: float, interval: float, limit: float) -> bool:
        """Simulation::run_until_decayed (src/fdtd/monitors.rs:740-841): whether the fields at
        the probes decayed below `fraction` of their peak over a block of `interval` before
        `limit`; raises NativeError (code ERR_GROWING) if they grow again after the sources end."""
        pr = np.ascontiguousarray(probes, dtype=np.int32)
        d = C.c_int32()
        N.check(N.lib.pxfdtd_simulation_run_until_decayed(self._h, pr.ctypes.data_as(N.pi32), len(pr),
                                                          float(fraction), float(interval), float(limit),
                                                          C.byref(d), None))
        return bool(d.value)

    def dt(self) -> float:
        return N.lib.pxfdtd_simulation_dt(self._h)

    def steps(self) -> int:
        return N.lib.pxfdtd_simulation_steps(self._h)

    def _field(self, which, component) -> np.ndarray:
        out = np.zeros(self.grid.cells())
        N.check(N.lib.pxfdtd_simulation_field(self._h, which, int(component), N.dptr(out)))
        return out

    def e(self, component) -> np.ndarray:
        return self._field(0, component)

    def h(self, component) -> np.ndarray:
        return self._field(1, component)

    def set_e(self, component, values):
        v = np.ascontiguousarray(values, dtype=np.float64).ravel()
        N.check(N.lib.pxfdtd_simulation_set_field(self._h, 0, int(component), N.dptr(v)))

    def set_h(self, component, values):
        v = np.ascontiguousarray(values, dtype=np.float64).ravel()
        N.check(N.lib.pxfdtd_simulation_set_field(self._h, 1, int(component), N.dptr(v)))

    def coefficients(self, component):
        ca = np.zeros(self.grid.cells())
        cb = np.zeros(self.grid.cells())
        N.check(N.lib.pxfdtd_simulation_coefficients(self._h, int(component), N.dptr(ca), N.dptr(cb)))
        return ca, cb

    def resident(self):
        return N.lib.pxfdtd_simulation_resident(self._h)

    def footprint(self, precision: "Precision" = None) -> int:
        """Device bytes this simulation needs (what the server's admission counts)."""
        b = C.c_int64()
        p = Precision.F32 if precision is None else precision
        N.check(N.lib.pxfdtd_simulation_footprint(self._h, int(p), C.byref(b)))
        return b.value

    def autotune(self, trials: int = 3, max_rows: int = 4):
        """Times the aligned march's tilings on the current fields without stepping and keeps
        the fastest: (rows, planes, ms per step kernel)."""
        r = self.resident()
        if not r:
            raise N.NativeError(N.ERR_UNSUPPORTED, "no device")
        rows, planes, ms = C.c_int32(), C.c_int32(), C.c_double()
        N.check(N.lib.pxfdtd_resident_autotune(r, int(trials), int(max_rows), C.byref(rows),
                                               C.byref(planes), C.byref(ms)))
        return rows.value, planes.value, ms.value

    def set_tiling(self, rows=0, planes=0):
        r = self.resident()
        if not r:
            raise N.NativeError(N.ERR_UNSUPPORTED, "no device")
        N.check(N.lib.pxfdtd_resident_set_tiling(r, rows, planes))

    def set_kernel(self, kind: int):
        """0: tiles exchanging H through LDS; 1: barrier-free strips marching along z."""
        r = self.resident()
        if not r:
            raise N.NativeError(N.ERR_UNSUPPORTED, "no device")
        N.check(N.lib.pxfdtd_resident_set_kernel(r, kind))

    def kernel(self) -> int:
        return N.lib.pxfdtd_resident_get_kernel(self.resident())

    def tiling(self):
        r = self.resident()
        rows, planes = C.c_int32(), C.c_int32()
        N.check(N.lib.pxfdtd_resident_get_tiling(r, C.byref(rows), C.byref(planes)))
        return rows.value, planes.value

    def last_step_ms(self) -> float:
        return N.lib.pxfdtd_resident_last_kernel_ms(self.resident())

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            N.lib.pxfdtd_simulation_destroy(h)
            self._h = None


def _grid(g: Grid3d) -> N.Grid:
    return N.Grid(g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, g.x0, g.y0, g.z0)
