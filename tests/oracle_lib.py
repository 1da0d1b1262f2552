"""ctypes bindings of the CPU oracle (oracle/liboracle.so): test infrastructure only, the
checker the device path is compared against. See oracle/oracle.h."""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB_PATH = os.path.join(ORACLE_DIR, "liboracle.so")


def _load():
    if not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)
    return C.CDLL(LIB_PATH)


lib = _load()
d3 = C.c_double * 3
i3 = C.c_int32 * 3
pd = C.POINTER(C.c_double)


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("dx", C.c_double),
                ("dy", C.c_double), ("dz", C.c_double), ("x0", C.c_double), ("y0", C.c_double),
                ("z0", C.c_double)]


class Bounds(C.Structure):
    _fields_ = [("periodic", i3), ("low", i3), ("high", i3), ("reflection", C.c_double),
                ("has_sigma", C.c_int32), ("sigma", C.c_double), ("order", C.c_double),
                ("kappa", C.c_double), ("alpha", C.c_double)]


class Shape(C.Structure):
    _fields_ = [("kind", C.c_int32), ("lo", d3), ("hi", d3), ("value", d3)]


class Scene(C.Structure):
    _fields_ = [("background", d3), ("n", C.c_int32), ("shapes", C.POINTER(Shape))]


class Wave(C.Structure):
    _fields_ = [("kind", C.c_int32), ("frequency", C.c_double), ("width", C.c_double),
                ("delay", C.c_double), ("amp_re", C.c_double), ("amp_im", C.c_double),
                ("ramp", C.c_double)]


class C64(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


vp = C.c_void_p
for name, res, args in [
    ("or_last_error", C.c_char_p, []),
    ("or_sim_new", vp, [C.POINTER(Grid), C.POINTER(Bounds), C.POINTER(Scene), C.c_double]),
    ("or_sim_new_eps", vp, [C.POINTER(Grid), C.POINTER(Bounds), pd, C.c_double]),
    ("or_sim_free", None, [vp]),
    ("or_sim_conductivity", C.c_int, [vp, C.POINTER(Scene)]),
    ("or_sim_conductor", None, [vp, C.c_int, C.c_int, C.c_int, C.c_int]),
    ("or_sim_add_source", C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(Wave)]),
    ("or_sim_add_current", C.c_int, [vp, C.c_int, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_int64), pd, C.POINTER(Wave)]),
    ("or_sim_add_probe", C.c_int, [vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    ("or_sim_add_dft", C.c_int, [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.c_int, pd]),
    ("or_sim_add_medium", C.c_int, [vp, C.c_double, C.c_int, pd, C.POINTER(Scene)]),
    ("or_sim_add_plane_wave", C.c_int, [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_long), pd, C.c_double, C.POINTER(Wave)]),
    ("or_sim_plane_links", C.c_long, [vp, C.c_int, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_int64), pd, C.POINTER(C.c_int32), C.POINTER(C.c_int64)]),
    ("or_sim_plane_aux", C.c_long, [vp, C.c_int, C.c_int, C.c_int, pd]),
    ("or_sim_plane_scale_link", None, [vp, C.c_int, C.c_int, C.c_long, C.c_double]),
    ("or_sim_plane_incident", C.c_double, [vp] + [C.c_int] * 6),
    ("or_sim_use_f32", None, [vp]),
    ("or_sim_run", None, [vp, C.c_long]),
    ("or_sim_field", pd, [vp, C.c_int, C.c_int]),
    ("or_sim_set_field", None, [vp, C.c_int, C.c_int, pd]),
    ("or_sim_dt", C.c_double, [vp]),
    ("or_sim_steps", C.c_long, [vp]),
    ("or_sim_ca", pd, [vp, C.c_int]),
    ("or_sim_cb", pd, [vp, C.c_int]),
    ("or_sim_factors", pd, [vp, C.c_int, C.c_int]),
    ("or_sim_slab_count", C.c_int, [vp]),
    ("or_sim_slab", None, [vp, C.c_int, C.POINTER(C.c_int32), C.POINTER(pd), C.POINTER(pd), C.POINTER(pd), C.POINTER(C.c_int64)]),
    ("or_sim_probe_count", C.c_int, [vp]),
    ("or_sim_probe", pd, [vp, C.c_int, C.POINTER(C.c_int64)]),
    ("or_sim_dft_series", pd, [vp, C.c_int, C.c_int, C.POINTER(C.c_int64)]),
    ("or_sim_electric_energy", C.c_double, [vp]),
    ("or_exp", C.c_double, [C.c_double]),
    ("or_sincos", None, [C.c_double, pd, pd]),
    ("or_powi", C.c_double, [C.c_double, C.c_int]),
    ("or_ln", C.c_double, [C.c_double]),
    ("or_pole_coefficients", None, [pd, C.c_double, pd]),
    ("or_media_plant", None, [C.c_int]),
    ("or_dipole_values", C.c_int, [C.POINTER(Grid), C.POINTER(Bounds), C.c_int, C.c_int, pd, C.c_double, C.c_double,
                                   C.POINTER(C.c_int64), pd]),
    ("or_pow", C.c_double, [C.c_double, C.c_double]),
    ("or_waveform_at", C.c_double, [C.POINTER(Wave), C.c_double]),
    ("or_waveform_complex_at", C64, [C.POINTER(Wave), C.c_double]),
    ("or_cpu_rate_f32", C.c_double, [C.c_int] * 5 + [C.POINTER(C.c_long)]),
    ("or_cpu_triad", C.c_double, [C.c_long, C.c_int, C.c_int]),
    ("or_cpu_rate_blocked_f32", C.c_double, [C.c_int] * 7 + [C.POINTER(C.c_long)]),
    ("or_cpu_auto_blocking", None, [C.c_int] * 4 + [C.POINTER(C.c_int)] * 2),
    ("or_cpu_blocked_differs", C.c_long, [C.c_int] * 8),
]:
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args


def scene(background, shapes=()):
    """A scene from (kind, lo, hi, value) tuples, value a number or a 3-tuple."""
    bg = background if isinstance(background, (tuple, list)) else (background,) * 3
    arr = (Shape * max(1, len(shapes)))()
    for q, (kind, lo, hi, val) in enumerate(shapes):
        v = val if isinstance(val, (tuple, list)) else (val,) * 3
        arr[q] = Shape(kind, d3(*lo), d3(*hi), d3(*v))
    sc = Scene(d3(*bg), len(shapes), arr)
    sc._keep = arr
    return sc


def bounds(periodic=(0, 0, 0), low=(0, 0, 0), high=(0, 0, 0), reflection=1e-8, sigma=None,
           order=3.0, kappa=1.0, alpha=0.0):
    return Bounds(i3(*periodic), i3(*low), i3(*high), reflection, 0 if sigma is None else 1,
                  0.0 if sigma is None else sigma, order, kappa, alpha)


class Sim:
    """The oracle's Simulation."""

    def __init__(self, grid, b, eps=None, courant=0.9, eps3n=None):
        self.grid = grid
        g = Grid(*grid)
        if eps3n is not None:
            e = np.ascontiguousarray(eps3n, dtype=np.float64)
            self.h = lib.or_sim_new_eps(C.byref(g), C.byref(b), e.ctypes.data_as(pd), courant)
        else:
            self.h = lib.or_sim_new(C.byref(g), C.byref(b), C.byref(eps), courant)
        if not self.h:
            raise ValueError(lib.or_last_error().decode())
        self.n = grid[0] * grid[1] * grid[2]
        self._dft_shapes = []

    def __del__(self):
        if getattr(self, "h", None):
            lib.or_sim_free(self.h)
            self.h = None

    def field(self, which, comp):
        p = lib.or_sim_field(self.h, which, comp)
        return np.ctypeslib.as_array(p, shape=(self.n,)).copy()

    def set_field(self, which, comp, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        lib.or_sim_set_field(self.h, which, comp, v.ctypes.data_as(pd))

    def fields(self):
        return np.stack([self.field(w, c) for w in (0, 1) for c in range(3)])

    def coefficients(self, comp):
        ca = np.ctypeslib.as_array(lib.or_sim_ca(self.h, comp), shape=(self.n,)).copy()
        cb = np.ctypeslib.as_array(lib.or_sim_cb(self.h, comp), shape=(self.n,)).copy()
        return ca, cb

    def dt(self):
        return lib.or_sim_dt(self.h)

    def steps(self):
        return lib.or_sim_steps(self.h)

    def run(self, n):
        lib.or_sim_run(self.h, n)

    def use_f32(self):
        lib.or_sim_use_f32(self.h)

    def conductivity(self, sc):
        if lib.or_sim_conductivity(self.h, C.byref(sc)):
            raise ValueError(lib.or_last_error().decode())

    def conductor(self, comp, i, j, k):
        lib.or_sim_conductor(self.h, comp, i, j, k)

    def add_source(self, field, comp, at, wave):
        if lib.or_sim_add_source(self.h, field, comp, *at, C.byref(wave)):
            raise ValueError(lib.or_last_error().decode())

    def add_current(self, field, comps, index, amps, wave):
        comps = np.ascontiguousarray(comps, dtype=np.int32)
        index = np.ascontiguousarray(index, dtype=np.int64)
        a = np.ascontiguousarray(np.stack([np.real(amps), np.imag(amps)], -1), dtype=np.float64)
        if lib.or_sim_add_current(self.h, field, len(index), comps.ctypes.data_as(C.POINTER(C.c_int32)),
                                  index.ctypes.data_as(C.POINTER(C.c_int64)), a.ctypes.data_as(pd),
                                  C.byref(wave)):
            raise ValueError(lib.or_last_error().decode())

    def add_probe(self, field, comp, at):
        return lib.or_sim_add_probe(self.h, field, comp, *at)

    def probe(self, n):
        ln = C.c_int64()
        p = lib.or_sim_probe(self.h, n, C.byref(ln))
        return np.ctypeslib.as_array(p, shape=(ln.value,)).copy() if ln.value else np.zeros(0)

    def add_dft(self, low, high, freqs):
        f = np.ascontiguousarray(freqs, dtype=np.float64)
        n = lib.or_sim_add_dft(self.h, (C.c_int32 * 3)(*low), (C.c_int32 * 3)(*high), len(f),
                               f.ctypes.data_as(pd))
        if n < 0:
            raise ValueError(lib.or_last_error().decode())
        self._dft_shapes.append((len(f), tuple(h - l + 1 for l, h in zip(low, high))))
        return n

    def dft(self, n, field, comp):
        ln = C.c_int64()
        p = lib.or_sim_dft_series(self.h, n, field * 3 + comp, C.byref(ln))
        a = np.ctypeslib.as_array(p, shape=(ln.value,)).copy()
        nf, (lx, ly, lz) = self._dft_shapes[n]
        return (a[0::2] + 1j * a[1::2]).reshape(nf, lz, ly, lx)

    def add_medium(self, eps_inf, poles, inside):
        arr = np.zeros((max(1, len(poles)), 4))
        for q, p in enumerate(poles):
            arr[q] = (0.0, p[1], p[2], p[3]) if p[0] == "lorentz" else (1.0, p[1], p[2], 0.0)
        if lib.or_sim_add_medium(self.h, eps_inf, len(poles), arr.ctypes.data_as(pd), C.byref(inside)):
            raise ValueError(lib.or_last_error().decode())

    def energy(self):
        return lib.or_sim_electric_energy(self.h)

    def add_plane_wave(self, low, high, direction, polarization, eps, wave):
        if lib.or_sim_add_plane_wave(self.h, (C.c_int * 3)(*low), (C.c_int * 3)(*high),
                                     (C.c_long * 3)(*direction), d3(*polarization), eps,
                                     C.byref(wave)):
            raise ValueError(lib.or_last_error().decode())

    def plane_links(self, p, field):
        """(component, value index, weight, incident component, sample) of plane p's couplings
        onto field (0 E, 1 H)."""
        n = lib.or_sim_plane_links(self.h, p, field, None, None, None, None, None)
        c, r, f = np.zeros(n, np.int32), np.zeros(n, np.int64), np.zeros(n, np.int32)
        w, i = np.zeros(n), np.zeros(n, np.int64)
        lib.or_sim_plane_links(self.h, p, field, c.ctypes.data_as(C.POINTER(C.c_int32)),
                               r.ctypes.data_as(C.POINTER(C.c_int64)), w.ctypes.data_as(pd),
                               f.ctypes.data_as(C.POINTER(C.c_int32)),
                               i.ctypes.data_as(C.POINTER(C.c_int64)))
        return c, r, w, f, i

    def plane_scale_link(self, p, field, m, factor):
        lib.or_sim_plane_scale_link(self.h, p, field, m, factor)

    def incident(self, p, field, comp, at):
        return lib.or_sim_plane_incident(self.h, p, field, comp, *at)

    def plane_aux(self, p, field, comp):
        n = lib.or_sim_plane_aux(self.h, p, field, comp, None)
        out = np.zeros(n)
        lib.or_sim_plane_aux(self.h, p, field, comp, out.ctypes.data_as(pd))
        return out


def dipole_values(grid, b, field, comp, position, amp):
    """The oracle's restriction of a dipole to the grid: (indices, complex amplitudes)."""
    g = Grid(*grid)
    idx = np.zeros(8, np.int64)
    a = np.zeros(16)
    pos = np.ascontiguousarray(position, dtype=np.float64)
    n = lib.or_dipole_values(C.byref(g), C.byref(b), field, comp, pos.ctypes.data_as(pd), complex(amp).real,
                             complex(amp).imag, idx.ctypes.data_as(C.POINTER(C.c_int64)), a.ctypes.data_as(pd))
    if n < 0:
        raise ValueError(lib.or_last_error().decode())
    return idx[:n].copy(), a[0:2 * n:2] + 1j * a[1:2 * n:2]


def wave(kind=0, frequency=1.0, width=0.6, delay=1.8, amp=1 + 0j, ramp=0.0):
    return Wave(kind, frequency, width, delay, complex(amp).real, complex(amp).imag, ramp)


def fold_hash(arrays):
    """h = rotl(h, 7) ^ bits, over the values in order (tests/canary.rs:138-140)."""
    h = 0
    mask = (1 << 64) - 1
    for a in arrays:
        for u in np.ascontiguousarray(a, dtype=np.float64).view(np.uint64).tolist():
            h = (((h << 7) | (h >> 57)) & mask) ^ u
    return h
