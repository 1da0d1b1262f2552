"""FDTD jobs: photonoxide's `photonoxide run job.yaml` for a task of kind fdtd
(src/job/fdtd.rs), stepping on the MI355X.

A job file names a window, a grid step, a stack of layers and shapes drawn on them, sources
(a guide's mode, a dipole), monitors (a guide's mode, a flux plane or box, a field's transform,
a resonance) and when to stop; the run records its events, one JSON object a line, in
`events.jsonl`: `started`, `solver`, `fdtd_progress` (with the cell-updates per second),
`fdtd_spectrum` (as the transforms accumulate, and once more at the end), `fdtd_resonances`
and `finished`, each with the reference's fields (src/job/event.rs:18-383).

The task's keys are the reference's FdtdTask's (src/job/fdtd.rs:186-232); `wavelength`,
`x`, `y`, `z`, `step` with units ("1.55 um", "20 nm") are read as `wavelength_um` and so on,
as the job templates write them. What differs, and why:
- the stacks: photonoxide's catalogue of dispersive materials is not part of this build. A
  stack is `soi_220` (a 220 nm Si layer, n = 3.48, in SiO2, n = 1.444, the indices the
  workloads state) or given in the job as `stack: {substrate, cladding, layers: [{name,
  thickness_um, material, background}]}` with `materials: {name: index}`;
- in 2D, the effective index method solves each column's slab mode in closed form (a transfer
  matrix across the layers) where the reference solves it by 1D FDFD (src/job/fdfd.rs
  slab_index): the two agree to the 1D grid's discretization error;
- in 2D, shapes are drawn on the job's layer only (each column is then the layer's background
  or its material); etches, paths, components, bodies, Bloch phases, plane waves, beams,
  frames and npz exports are not read.

The run loop is the reference's (src/job/fdtd.rs:2705-2858): checks every 10 periods of the
carrier once the pulse has ended, a decay measured as the largest |field|^2 at the monitors'
points over a check against its peak, growth detection, the spectra referred to the incident
power. The points' values are recorded on the device every step (probes) and read once per
check: the run does not synchronize with the host per step.
"""
from __future__ import annotations

import json
import math
import os
import re
import time
from dataclasses import dataclass
from typing import Callable, Optional

import numpy as np

from . import fdtd as F

PROGRESS_SECONDS = 1.0
SPECTRA_SECONDS = 2.0
RESONANCE_SAMPLES = 4000
MOST_CELLS = 4e9
GROWTH = 1e3

EPS_SI, EPS_OX = 3.48 ** 2, 1.444 ** 2
STACKS = {
    "soi_220": {"substrate": "SiO2", "cladding": "SiO2",
                "layers": [{"name": "Si", "thickness_um": 0.22, "material": "Si", "background": "SiO2"}]},
}
MATERIALS = {"Si": 3.48, "SiO2": 1.444}


class JobError(ValueError):
    """A job the run refuses, with the reference's reason."""


def load(path: str) -> dict:
    """A job file (YAML) as a dict, loaded without executing anything from it."""
    import yaml
    with open(path) as f:
        return yaml.safe_load(f)


# -------------------------------------------------------------------------------- parsing
_UNIT = re.compile(r"^\s*([-+0-9.eE]+)\s*(um|nm)\s*$")


def _length_um(v, what):
    if isinstance(v, (int, float)):
        return float(v)
    m = _UNIT.match(str(v))
    if not m:
        raise JobError(f"{what}: expected a length such as \"1.55 um\" or \"20 nm\", got {v!r}")
    x = float(m.group(1))
    return x / 1000.0 if m.group(2) == "nm" else x


def _normalized(task: dict) -> dict:
    """The task with `wavelength`, `x`, `y`, `z` and `step` (lengths with units) as the
    reference's `wavelength_um`, `x_um`, `y_um`, `z_um` and `step_nm`, recursively in items."""
    out = {}
    for k, v in task.items():
        if k == "step" and "step_nm" not in task:
            out["step_nm"] = 1000.0 * _length_um(v, k)
        elif k in ("wavelength", "at", "waist", "focus", "radius", "width") and f"{k}_um" not in task:
            out[f"{k}_um"] = _length_um(v, k)
        elif k in ("x", "y", "z", "center", "size", "position") and f"{k}_um" not in task:
            out[f"{k}_um"] = [_length_um(e, k) for e in v]
        elif k == "wavelengths" and "wavelengths_um" not in task:
            out["wavelengths_um"] = [_length_um(e, k) for e in v]
        elif k in ("from", "to") and f"{k}_um" not in task:
            out[f"{k}_um"] = _length_um(v, k)
        elif isinstance(v, dict):
            out[k] = _normalized(v)
        elif isinstance(v, list) and v and all(isinstance(e, dict) for e in v):
            out[k] = [_normalized(e) for e in v]
        else:
            out[k] = v
    return out


def _axis(name, what, allowed):
    a = {"x": 0, "y": 1, "z": 2}.get(str(name))
    if a is None or a not in allowed:
        raise JobError(f"{what}: unknown normal \"{name}\": {' or '.join('xyz'[q] for q in allowed)}")
    return a


def _component(name, what):
    m = re.fullmatch(r"([EH])([xyz])", str(name))
    if not m:
        raise JobError(f"{what}: unknown component \"{name}\": Ex, Ey, Ez, Hx, Hy or Hz")
    return (0 if m.group(1) == "E" else 1), "xyz".index(m.group(2))


def _only(spec, what, allowed):
    extra = [k for k in spec if k != "type" and k not in allowed]
    if extra:
        raise JobError(f"{what}: a {spec.get('type')} doesn't take {extra[0]} (it takes {', '.join(allowed)})")


# ------------------------------------------------------------------ the effective index method
def slab_index(column, wavelength_um, te=True):
    """The fundamental slab mode's effective index of a column of layers, bottom to top:
    column = (n_substrate, [(n, thickness_um), ...], n_cladding), the field decaying into the
    substrate and the cladding; TE with E along the layers, TM with H along them. A transfer
    matrix across the layers of (field, its normal derivative over n^2 for TM), its root in
    n_eff the largest between the outer indices and the largest layer's, by bisection."""
    ns, layers, nc = column
    k0 = 2.0 * math.pi / wavelength_um
    lo = max(ns, nc)
    hi = max([n for n, _ in layers] + [lo])
    if hi <= lo:
        return lo

    def mismatch(neff):
        gs = k0 * math.sqrt(max(neff * neff - ns * ns, 0.0))
        gc = k0 * math.sqrt(max(neff * neff - nc * nc, 0.0))
        f, d = 1.0, gs / (1.0 if te else ns * ns)
        for n, t in layers:
            w = 1.0 if te else n * n
            k2 = k0 * k0 * (n * n - neff * neff)
            if k2 > 0:
                k = math.sqrt(k2)
                c, s = math.cos(k * t), math.sin(k * t)
                f, d = c * f + w * s / k * d, -k * s / w * f + c * d
            elif k2 < 0:
                g = math.sqrt(-k2)
                c, s = math.cosh(g * t), math.sinh(g * t)
                f, d = c * f + w * s / g * d, g * s / w * f + c * d
            else:
                f, d = f + w * t * d, d
        return d + gc / (1.0 if te else nc * nc) * f

    grid = np.linspace(hi, lo, 4001)[1:-1]
    prev_x, prev_v = grid[0], mismatch(grid[0])
    for x in grid[1:]:
        v = mismatch(x)
        if prev_v == 0.0:
            return float(prev_x)
        if (v < 0) != (prev_v < 0):
            a, b, fa = x, prev_x, v
            for _ in range(200):
                m = 0.5 * (a + b)
                fm = mismatch(m)
                if (fm < 0) == (fa < 0):
                    a, fa = m, fm
                else:
                    b = m
                if b - a < 1e-15 * b:
                    break
            return 0.5 * (a + b)
        prev_x, prev_v = x, v
    raise JobError("the layer's slab guides no mode at this wavelength: the 2D effective index needs one")


# --------------------------------------------------------------------------------- the setup
@dataclass
class Plan:
    kind: str
    spec: dict
    what: str
    name: str = ""
    axis: int = 0
    plane: int = 0
    direction: int = 0
    window: tuple = (None, None)
    te: bool = True
    at: tuple = (0, 0, 0)
    field: int = 0
    component: int = 0
    position: tuple = ()
    wavelengths_um: tuple = ()


class Setup:
    """A job's FDTD task, checked as the reference checks it (Setup::new, src/job/fdtd.rs:414-713)."""

    def __init__(self, job: dict):
        if not isinstance(job, dict) or not isinstance(job.get("task"), dict):
            raise JobError("a job needs a [task]")
        self.name = str(job.get("name", "fdtd"))
        self.timeout_minutes = job.get("timeout_minutes")
        t = _normalized(job["task"])
        if t.get("kind") != "fdtd":
            raise JobError(f"this runner takes tasks of kind fdtd, got {t.get('kind')!r}")
        self.task = t
        for key in ("wavelength_um", "x_um", "y_um", "step_nm", "layer"):
            if key not in t:
                raise JobError(f"an fdtd task needs {key}")
        for key in ("x_um", "y_um"):
            w = t[key]
            if not (len(w) == 2 and all(math.isfinite(v) for v in w) and w[1] > w[0]):
                raise JobError(f"{key} must be [low, high] with high above low, got {w}")
        if not (t["step_nm"] > 0 and math.isfinite(t["step_nm"])):
            raise JobError(f"step_nm must be positive, got {t['step_nm']}")
        dims = t.get("dimensions", 2)
        if dims not in (2, 3):
            raise JobError(f"dimensions must be 2 or 3, got {dims}")
        self.three = dims == 3
        pol = t.get("polarization")
        if self.three and pol is not None:
            raise JobError("polarization is for a 2D job: a 3D job's sources and monitors each say theirs")
        if not self.three and pol not in (None, "te", "tm"):
            raise JobError(f"unknown polarization \"{pol}\": te or tm")
        self.te = pol != "tm"
        if not self.three and "z_um" in t:
            raise JobError("z_um is for a 3D job (dimensions = 3): a 2D job is the layer's plane")
        self.stack, self.materials = self._stack(t)
        self.layer = str(t["layer"])
        names = [l["name"] for l in self.stack["layers"]]
        if self.layer not in names:
            raise JobError(f"the stack has no layer {self.layer}")
        tops = np.cumsum([l["thickness_um"] for l in self.stack["layers"]])
        li = names.index(self.layer)
        self.tops = tops
        self.layer_z = (float(tops[li - 1]) if li else 0.0, float(tops[li]))
        self.courant = float(t.get("courant", 0.9))
        if not (0.0 < self.courant <= 1.0):
            raise JobError(f"courant must be in (0, 1], got {self.courant}")
        z_window = t.get("z_um", [self.layer_z[0] - 1.0, self.layer_z[1] + 1.0])
        self.grid = self._grid(t, z_window)
        g = self.grid
        if g.nx * g.ny * g.nz > MOST_CELLS:
            raise JobError(f"the grid is {g.nx} x {g.ny} x {g.nz} cells, more than {MOST_CELLS:e}")
        self.pml = int(t.get("pml_cells", 20))
        self.boundaries = self._boundaries()
        hs = [g.dx, g.dy] + ([g.dz] if self.three else [])
        self.dt = self.courant / math.sqrt(sum(1.0 / (h * h) for h in hs))
        sp = t.get("spectrum")
        if sp is None:
            self.wavelengths = [float(t["wavelength_um"])]
        else:
            pts = int(sp.get("points", 0))
            a, b = float(sp.get("from_um", 0)), float(sp.get("to_um", 0))
            if pts == 0 or not (a > 0 and b > 0):
                raise JobError("the spectrum needs positive from_um and to_um and at least 1 point")
            if pts > 1000:
                raise JobError(f"the spectrum's points must be at most 1000, got {pts}")
            self.wavelengths = [a] if pts == 1 else [a + (b - a) * k / (pts - 1) for k in range(pts)]
        self.carrier = 1.0 / float(t["wavelength_um"])
        reach = max(abs(1.0 / w - self.carrier) for w in self.wavelengths)
        self.bandwidth = max(2.0 * reach, 0.05 * self.carrier)
        nyquist = 0.5 / self.dt
        for f in [1.0 / w for w in self.wavelengths] + [self.carrier + self.bandwidth]:
            if f >= 0.8 * nyquist:
                raise JobError(f"the pulse and the spectrum reach {f:.4f} c/um, too near the time step's Nyquist "
                               f"frequency {nyquist:.4f} c/um: take a smaller step_nm or a narrower spectrum")
        self.until = self._until(t.get("stop"))
        self.margins = [self.pml + 2 if self.pml > 0 else 0] * 2
        self.view_k = g.nz // 2 if self.three else 0
        sources = t.get("source", [])
        if not sources:
            raise JobError("an fdtd job needs at least one [[task.source]]")
        self.sources = [self._source(k, s) for k, s in enumerate(sources)]
        self.monitors = [self._monitor(k, m) for k, m in enumerate(t.get("monitor", []))]
        seen = set()
        for m in self.monitors:
            if m.name and m.kind != "field":
                if m.name in seen:
                    raise JobError(f"two monitors are named \"{m.name}\": each needs its own name")
                seen.add(m.name)

    # -- the structure
    def _stack(self, t):
        st = t.get("stack")
        if st is None:
            raise JobError("needs a stack, e.g. stack = \"soi_220\"")
        materials = dict(MATERIALS)
        materials.update({k: float(v) for k, v in (t.get("materials") or {}).items()})
        if isinstance(st, str):
            if st not in STACKS:
                raise JobError(f"unknown stack \"{st}\": {', '.join(STACKS)} or a stack given in the job")
            st = STACKS[st]
        stack = {"substrate": st["substrate"], "cladding": st["cladding"], "layers": [dict(l) for l in st["layers"]]}
        if t.get("core_nm") is not None:
            stack["layers"][0]["thickness_um"] = float(t["core_nm"]) / 1000.0
        for name in [stack["substrate"], stack["cladding"]] + [l[k] for l in stack["layers"]
                                                               for k in ("material", "background")]:
            if name not in materials:
                raise JobError(f"the stack's material {name} has no index: give it under materials")
        if len(stack["layers"]) > 16:
            raise JobError("an fdtd job's stack may have at most 16 layers")
        return stack, materials

    def _index(self, name):
        return self.materials[name]

    def _grid(self, t, z):
        """grid_of (src/job/fdtd.rs:1726-1747)."""
        h = t["step_nm"] / 1000.0
        cells = lambda w: max(int(round((w[1] - w[0]) / h)), 1)
        nx, ny = cells(t["x_um"]), cells(t["y_um"])
        if self.three:
            if not (len(z) == 2 and z[1] > z[0]):
                raise JobError(f"z_um must be [low, high] with high above low, got {z}")
            nz = cells(z)
            nzd, z0 = (z[1] - z[0]) / nz, z[0]
        else:
            nz, nzd, z0 = 1, h, 0.5 * (self.layer_z[0] + self.layer_z[1]) - 0.5 * h
        return F.Grid3d(nx, ny, nz, (t["x_um"][1] - t["x_um"][0]) / nx, (t["y_um"][1] - t["y_um"][0]) / ny, nzd,
                        t["x_um"][0], t["y_um"][0], z0)

    def _boundaries(self):
        """boundaries_of (src/job/fdtd.rs:1749-1816), CPMLs or walls."""
        spec = self.task.get("boundaries") or {}
        if not self.three and "z" in spec:
            raise JobError("boundaries.z is for a 3D job: a 2D job is periodic along z")
        edges = []
        for a in range(3 if self.three else 2):
            kind = spec.get("xyz"[a], "cpml")
            if kind == "cpml":
                if 2 * self.pml + 4 > self.grid.n(a):
                    raise JobError(f"CPMLs of {self.pml} cells at each end leave nothing of the {self.grid.n(a)} "
                                   f"cells along {'xyz'[a]}")
                edges.append(F.Pml(self.pml, self.pml))
            elif kind == "wall":
                edges.append(F.Pml(0, 0))
            elif kind == "periodic":
                edges.append(F.Periodic())
            else:
                raise JobError(f"unknown boundary \"{kind}\" along {'xyz'[a]}: cpml, wall or periodic")
        if not self.three:
            edges.append(F.Periodic())
        return F.Boundaries(*edges)

    def cpml(self, a):
        e = (self.boundaries.x, self.boundaries.y, self.boundaries.z)[a]
        return (0, 0) if isinstance(e, F.Periodic) else (e.low, e.high)

    def _until(self, stop):
        """stop_of (src/job/fdtd.rs:1679-1724): ('decay', fraction, limit) or ('time', t)."""
        stop = stop or {}
        until = stop.get("until", "decay")
        if until == "decay":
            if "time_um" in stop:
                raise JobError("time_um is for until = \"time\"; a run until the fields decay takes fraction and limit_um")
            fraction = float(stop.get("fraction", 1e-8))
            if not 0.0 < fraction < 1.0:
                raise JobError(f"fraction must be between 0 and 1, got {fraction}")
            limit = float(stop.get("limit_um", 2000.0))
            if not (math.isfinite(limit) and limit > 0):
                raise JobError(f"limit_um must be positive, got {limit}")
            return ("decay", fraction, limit)
        if until == "time":
            if "fraction" in stop or "limit_um" in stop:
                raise JobError("fraction and limit_um are for until = \"decay\"; a run for a time takes time_um")
            if "time_um" not in stop:
                raise JobError("until = \"time\" needs time_um")
            t = float(stop["time_um"])
            if not (math.isfinite(t) and t > 0):
                raise JobError(f"time_um must be positive, got {t}")
            return ("time", None, t)
        raise JobError(f"unknown until \"{until}\": decay or time")

    # -- planes and windows (src/job/fdtd.rs:721-880)
    def origin(self, a):
        return (self.grid.x0, self.grid.y0, self.grid.z0)[a]

    def node(self, a, v, what):
        m = round((v - self.origin(a)) / self.grid.step(a))
        if not (math.isfinite(v) and 0 <= m <= self.grid.n(a)):
            raise JobError(f"{what}: {'xyz'[a]} = {v} um is outside the window")
        return int(m)

    def clear_plane(self, a, v, what):
        p = self.node(a, v, what)
        low, high = self.cpml(a)
        if p < low + 6 or p + 9 + high > self.grid.n(a):
            raise JobError(f"{what}: its plane at {'xyz'[a]} = {v} um must be at least 6 cells clear of the CPML or "
                           f"the window's end before it ({low} cells of CPML) and 9 of those after it ({high})")
        return p

    def half_plane(self, a, v, what):
        m = round((v - self.origin(a)) / self.grid.step(a) - 0.5)
        low, high = self.cpml(a)
        n = self.grid.n(a)
        if not (math.isfinite(v) and m >= low and m + 1 + high <= n and m + 1 < n):
            raise JobError(f"{what}: {'xyz'[a]} = {v} um is outside the window or inside its CPMLs "
                           f"({low} and {high} cells)")
        return int(m)

    def cells(self, a, w, what):
        n = self.grid.n(a)
        if w is None:
            return (0, n)
        lo, hi = (round((v - self.origin(a)) / self.grid.step(a)) for v in w)
        if lo < 0 or hi > n:
            raise JobError(f"{what}: its {'xyz'[a]}_um, {w[0]} to {w[1]}, must lie inside the window")
        if hi < lo + 3:
            raise JobError(f"{what}: its {'xyz'[a]}_um, {w[0]} to {w[1]}, covers fewer than 3 cells")
        return (int(lo), int(hi))

    def window(self, spec, a, what):
        given = lambda q: spec.get(f"{'xyz'[q]}_um")
        if given(a) is not None:
            raise JobError(f"{what}: its plane is normal to {'xyz'[a]}: give at_um, and windows along the other axes")
        if not self.three and spec.get("z_um") is not None:
            raise JobError(f"{what}: z_um is for a 3D job")
        b, c = (a + 1) % 3, (a + 2) % 3
        return (self.cells(b, given(b), what), self.cells(c, given(c), what))

    def polarization(self, spec, what):
        p = spec.get("polarization")
        if not self.three:
            if p is not None:
                raise JobError(f"{what}: a 2D job's polarization is the task's; leave it out here")
            return self.te
        if p in (None, "te"):
            return True
        if p == "tm":
            return False
        raise JobError(f"{what}: unknown polarization \"{p}\": te or tm")

    def position(self, p, what):
        if p is None:
            raise JobError(f"{what}: needs position_um")
        if not self.three and len(p) == 2:
            x, y, z = p[0], p[1], self.grid.z0 + 0.5 * self.grid.dz
        elif self.three and len(p) == 3:
            x, y, z = p
        else:
            raise JobError(f"{what}: position_um must be [x, y]{', z' if self.three else ''}")
        for a, v in enumerate((x, y, z)):
            lo = self.origin(a)
            hi = lo + self.grid.n(a) * self.grid.step(a)
            if not (math.isfinite(v) and lo <= v <= hi):
                raise JobError(f"{what}: its {'xyz'[a]} = {v} um is outside the window")
        return (float(x), float(y), float(z))

    def in_polarization(self, field, comp, what):
        if self.three:
            return
        te_field = (field == 0 and comp in (0, 1)) or (field == 1 and comp == 2)
        if te_field != self.te:
            raise JobError(f"{what}: {'EH'[field]}{'xyz'[comp]} isn't a field of a 2D {'TE' if self.te else 'TM'} job")

    def cell_of(self, p):
        return tuple(min(max(int(math.floor((p[a] - self.origin(a)) / self.grid.step(a))), 0), self.grid.n(a) - 1)
                     for a in range(3))

    def port_window(self, plan):
        """A plane's window for the port solver: (lo, hi) along each of the axes after its normal,
        None where it is the whole axis."""
        out = []
        for slot, w in enumerate(plan.window):
            t = (plan.axis + 1 + slot) % 3
            out.append(None if tuple(w) == (0, self.grid.n(t)) else tuple(w))
        return tuple(out)

    def _source(self, k, spec):
        what = f"source {k + 1} ({spec.get('type')})"
        kind = spec.get("type")
        if kind == "mode":
            _only(spec, what, ["normal", "at_um", "direction", "x_um", "y_um", "z_um", "polarization"])
            a = _axis(spec.get("normal"), what, (0, 1))
            if "at_um" not in spec:
                raise JobError(f"{what}: needs at_um, its plane")
            d = spec.get("direction", "+")
            if d not in ("+", "-"):
                raise JobError(f"{what}: unknown direction \"{d}\": \"+\" or \"-\"")
            return Plan("mode", spec, what, axis=a, plane=self.clear_plane(a, float(spec["at_um"]), what),
                        direction=0 if d == "+" else 1, window=self.window(spec, a, what),
                        te=self.polarization(spec, what))
        if kind == "dipole":
            _only(spec, what, ["position_um", "component"])
            if "component" not in spec:
                raise JobError(f"{what}: needs a component, e.g. \"Ez\"")
            field, comp = _component(spec["component"], what)
            self.in_polarization(field, comp, what)
            return Plan("dipole", spec, what, field=field, component=comp,
                        position=self.position(spec.get("position_um"), what))
        raise JobError(f"{what}: this runner's sources are mode and dipole")

    def _monitor(self, k, spec):
        kind = spec.get("type")
        what = f"monitor {k + 1} ({kind})"
        name = str(spec.get("name", f"{str(kind).replace('_', ' ')} {k + 1}"))
        if kind == "mode":
            _only(spec, what, ["name", "normal", "at_um", "x_um", "y_um", "z_um", "polarization"])
            a = _axis(spec.get("normal"), what, (0, 1))
            if "at_um" not in spec:
                raise JobError(f"{what}: needs at_um, its plane")
            return Plan("mode", spec, what, name=name, axis=a, plane=self.clear_plane(a, float(spec["at_um"]), what),
                        window=self.window(spec, a, what), te=self.polarization(spec, what))
        if kind == "flux":
            _only(spec, what, ["name", "normal", "at_um", "x_um", "y_um", "z_um"])
            a = _axis(spec.get("normal"), what, (0, 1, 2) if self.three else (0, 1))
            if "at_um" not in spec:
                raise JobError(f"{what}: needs at_um, its plane")
            plane = self.half_plane(a, float(spec["at_um"]), what)
            win = []
            for q in ((a + 1) % 3, (a + 2) % 3):
                w = spec.get(f"{'xyz'[q]}_um")
                if w is None or (q == 2 and not self.three):
                    win.append(None)
                    continue
                lo, hi = self.half_plane(q, w[0], what), self.half_plane(q, w[1], what)
                if hi <= lo:
                    raise JobError(f"{what}: its {'xyz'[q]}_um covers less than a cell")
                win.append((lo, hi))
            return Plan("flux", spec, what, name=name, axis=a, plane=plane, window=tuple(win))
        if kind == "field":
            _only(spec, what, ["name", "wavelengths_um"])
            wl = spec.get("wavelengths_um", [float(self.task["wavelength_um"])])
            if not 1 <= len(wl) <= 20:
                raise JobError(f"{what}: wavelengths_um must hold 1 to 20 wavelengths")
            lo = min(1.0 / max(self.wavelengths), self.carrier - self.bandwidth)
            hi = max(1.0 / min(self.wavelengths), self.carrier + self.bandwidth)
            for w in wl:
                if not lo <= 1.0 / w <= hi:
                    raise JobError(f"{what}: {w} um is outside the pulse's band, {1 / hi:.4f} to {1 / lo:.4f} um")
            return Plan("field", spec, what, name=name, wavelengths_um=tuple(float(w) for w in wl))
        if kind == "resonance":
            _only(spec, what, ["name", "position_um", "component"])
            c = spec.get("component", "Ey" if self.three else ("Hz" if self.te else "Ez"))
            field, comp = _component(c, what)
            self.in_polarization(field, comp, what)
            p = self.position(spec.get("position_um"), what)
            return Plan("resonance", spec, what, name=name, field=field, component=comp, at=self.cell_of(p),
                        position=p if self.three else p[:2])
        raise JobError(f"monitor {k + 1}: unknown type \"{kind}\": mode, flux, field or resonance")

    # -- the permittivity (Setup::permittivity, src/job/fdtd.rs:1344-1487)
    def scene(self) -> F.Scene:
        """The permittivity as a scene the simulation averages: in 3D the substrate, each layer's
        background and its shapes, in 2D each column's effective index squared."""
        shapes = self.task.get("rect", []), self.task.get("circle", []), self.task.get("ring", [])
        drawn = []
        for q in self.task.get("shapes", []) or []:
            (kind, spec), = q.items()
            drawn.append((kind, _normalized(spec)))
        drawn = [("rect", s) for s in shapes[0]] + [("circle", s) for s in shapes[1]] + \
                [("ring", s) for s in shapes[2]] + drawn
        layers = self.stack["layers"]
        names = [l["name"] for l in layers]
        inf = float("inf")
        if not self.three:
            bad = [s for _, s in drawn if s.get("layer", self.layer) != self.layer]
            if bad:
                raise JobError("this runner's 2D effective index takes shapes on the job's layer only")
            li = names.index(self.layer)
            lam = float(self.task["wavelength_um"])

            def column(core):
                return (self._index(self.stack["substrate"]),
                        [(self._index(core if q == li else l["background"]), l["thickness_um"])
                         for q, l in enumerate(layers)],
                        self._index(self.stack["cladding"]))
            outside = slab_index(column(layers[li]["background"]), lam, self.te) ** 2
            inside = slab_index(column(layers[li]["material"]), lam, self.te) ** 2
            self.effective = (outside, inside)
            return F.Scene(outside, self._shapes(drawn, inside, outside, (-inf, inf)))
        out = [F.Box((-inf, -inf, -inf), (inf, inf, 0.0), self._index(self.stack["substrate"]) ** 2)]
        bottom = 0.0
        for l, top in zip(layers, self.tops):
            out.append(F.Box((-inf, -inf, bottom), (inf, inf, float(top)), self._index(l["background"]) ** 2))
            bottom = float(top)
        bottom = 0.0
        for l, top in zip(layers, self.tops):
            mine = [(k, s) for k, s in drawn if s.get("layer") == l["name"]]
            out += self._shapes(mine, self._index(l["material"]) ** 2, self._index(l["background"]) ** 2,
                                (bottom, float(top)))
            bottom = float(top)
        return F.Scene(self._index(self.stack["cladding"]) ** 2, out)

    def _shapes(self, drawn, eps, background, z):
        out = []
        for kind, s in drawn:
            if kind == "rect":
                (cx, cy), (sx, sy) = s["center_um"], s["size_um"]
                out.append(F.Box((cx - sx / 2, cy - sy / 2, z[0]), (cx + sx / 2, cy + sy / 2, z[1]), eps))
            elif kind == "circle":
                out.append(F.Cylinder(tuple(s["center_um"]), float(s["radius_um"]), z, eps))
            elif kind == "ring":
                r, w = float(s["radius_um"]), float(s["width_um"])
                out.append(F.Cylinder(tuple(s["center_um"]), r + w / 2, z, eps))
                out.append(F.Cylinder(tuple(s["center_um"]), r - w / 2, z, background))
            else:
                raise JobError(f"unknown shape \"{kind}\": rect, circle or ring")
        return out

    def waveform(self) -> F.Waveform:
        return F.Waveform.pulse(self.carrier, self.bandwidth)

    def probe_components(self):
        """Setup::probe_components (src/job/fdtd.rs:1519-1527)."""
        if self.three:
            return [(0, 0), (0, 1), (0, 2)]
        return [(1, 2)] if self.te else [(0, 2)]

    def probe_points(self):
        """Setup::probe_points (src/job/fdtd.rs:1529-1593)."""
        g = self.grid
        k_mid = g.nz // 2
        pts = []

        def at(a, plane, b, c):
            p = [0, 0, 0]
            p[a] = plane
            p[(a + 1) % 3] = b
            p[(a + 2) % 3] = c
            return (p[0], p[1], min(p[2], g.nz - 1))
        for m in self.monitors:
            if m.kind == "mode":
                mid = [(w[0] + w[1]) // 2 for w in m.window]
                pts.append(at(m.axis, m.plane, mid[0], mid[1]))
            elif m.kind == "flux":
                cen = []
                for slot, t in enumerate(((m.axis + 1) % 3, (m.axis + 2) % 3)):
                    w = m.window[slot]
                    cen.append((w[0] + w[1]) // 2 if w is not None else (k_mid if t == 2 else g.n(t) // 2))
                pts.append(at(m.axis, m.plane, cen[0], cen[1]))
            elif m.kind == "resonance":
                pts.append(tuple(m.at))
        if not pts:
            pts.append((g.nx // 2, g.ny // 2, k_mid))
        return pts


class Growth:
    """Growth (src/fdtd/monitors.rs:678-716): rising at each of the last `needed` checks to 1e3
    times the level before the first of them."""

    def __init__(self, pulse_end, check):
        self.needed = max(math.ceil(pulse_end / check) + 1, 4)
        self.start = self.last = 0.0
        self.rises = 0

    def check(self, level):
        if self.last > 0.0 and level > self.last:
            if self.rises == 0:
                self.start = self.last
            self.rises += 1
        else:
            self.rises = 0
        self.last = level
        return level / self.start if (self.rises >= self.needed and level >= GROWTH * self.start) else None


def _spectrum(w: F.Waveform, field, dt, steps, f):
    """Waveform::spectrum (src/fdtd.rs:353-396) of a pulse, at a current's times on `field`."""
    t = np.arange(steps) * dt + (0.5 * dt if field == 0 else 0.0)
    u = (t - w.delay) / w.width
    v = np.exp(-u * u) * np.sin(2.0 * np.pi * w.frequency * (t - w.delay))
    return complex(np.sum(v * np.exp(2j * np.pi * f * t)) * dt)


def _harmonic(signal, dt, lo, hi):
    from .fdtd import harmonic_inversion
    return harmonic_inversion(signal, dt, lo, hi)


# ------------------------------------------------------------------------------------- the run
def run_job(job, out_dir: Optional[str] = None, gpu: Optional[F.Gpu] = None,
            precision: F.Precision = F.Precision.F32, stop: Optional[Callable[[], Optional[str]]] = None):
    """Runs an FDTD job (a dict, or the path of its YAML file) on `gpu` (device 0 by default);
    its events, also written to out_dir/events.jsonl when given. `stop` returns a reason
    ("requested") to end the run early."""
    started = time.perf_counter()
    if isinstance(job, str):
        job = load(job)
    setup = Setup(job)
    gpu = gpu or F.Gpu(0)
    events = []
    sink = None
    if out_dir:
        os.makedirs(out_dir, exist_ok=True)
        sink = open(os.path.join(out_dir, "events.jsonl"), "w")

    def record(e):
        events.append(e)
        if sink:
            sink.write(json.dumps(e) + "\n")
            sink.flush()

    deadline = None if setup.timeout_minutes is None else started + 60.0 * float(setup.timeout_minutes)

    def reason():
        if stop is not None and stop():
            return stop()
        if deadline is not None and time.perf_counter() > deadline:
            return "timeout"
        return None
    try:
        record({"type": "started", "job": setup.name, "kind": "fdtd"})
        _run(setup, gpu, precision, record, reason)
        record({"type": "finished", "stopped": reason(), "seconds": time.perf_counter() - started})
    finally:
        if sink:
            sink.close()
    return events


def _run(setup: Setup, gpu, precision, record, reason):
    g = setup.grid
    sim = F.Simulation.new(g, setup.scene(), setup.boundaries, setup.courant, device=gpu)
    dt = sim.dt()
    wave = setup.waveform()
    freqs = [1.0 / w for w in setup.wavelengths]
    incident = []  # (mode monitor, direction) of each mode source
    for plan in setup.sources:
        if plan.kind == "mode":
            mode = sim.port_mode(plan.axis, plan.plane, setup.carrier, plan.te, setup.port_window(plan))
            sim.add_mode_source(mode, plan.direction, wave)
            after = plan.plane + 3 if plan.direction == 0 else plan.plane - 3
            modes = [sim.port_mode(plan.axis, after, f, plan.te, setup.port_window(plan)) for f in freqs]
            incident.append((sim.add_mode_monitor(modes), plan.direction))
        else:
            sim.add_dipole(F.Dipole(F.Field(plan.field), F.Axis(plan.component), plan.position, 1.0, wave))
    added = []
    for m in setup.monitors:
        if m.kind == "mode":
            modes = [sim.port_mode(m.axis, m.plane, f, m.te, setup.port_window(m)) for f in freqs]
            added.append(("mode", m, sim.add_mode_monitor(modes)))
        elif m.kind == "flux":
            added.append(("flux", m, sim.add_flux(m.axis, m.plane, freqs, m.window)))
        elif m.kind == "field":
            fs = [1.0 / w for w in m.wavelengths_um]
            added.append(("field", m, sim.add_dft((0, 0, setup.view_k), (g.nx - 1, g.ny - 1, setup.view_k), fs)))
        else:
            added.append(("resonance", m, sim.add_probe(F.Field(m.field), F.Axis(m.component), m.at)))
    points = setup.probe_points()
    comps = setup.probe_components()
    probes = [[sim.add_probe(F.Field(f), F.Axis(c), p) for (f, c) in comps] for p in points]
    sim.set_device(gpu, precision)
    pulse_end = wave.delay + 6.0 * wave.width
    check = 10.0 / setup.carrier
    limit = setup.until[2]
    record({"type": "solver", "module": "fdtd (MI355X, HIP)", "cells": [g.nx, g.ny], "step_um": g.dx,
            "unknowns": 6 * g.nx * g.ny * g.nz,
            "details": [["dimensions", f"3D: {g.nx} x {g.ny} x {g.nz} cells" if setup.three else
                         f"2D by the effective index method ({'TE: H along z, E in the plane' if setup.te else 'TM: E along z'}): "
                         f"{g.nx} x {g.ny} cells"],
                        ["time step", f"{dt:.6f} um/c, Courant number {setup.courant}"],
                        ["x", f"CPMLs of {setup.pml} cells" if setup.cpml(0)[0] else "walls"],
                        ["y", f"CPMLs of {setup.pml} cells" if setup.cpml(1)[0] else "walls"],
                        ["pulse", f"{setup.carrier:.6f} c/um, {setup.bandwidth:.6f} c/um wide, over by {pulse_end:.3f} um/c"],
                        ["precision", "f32 fields, FP64 transforms" if precision == F.Precision.F32 else "f64"]]})
    npts = len(points)
    peaks = np.zeros(npts)
    recent = np.zeros(npts)
    block = np.zeros(npts)
    block_end = pulse_end + check
    growth = Growth(pulse_end, check)
    decayed, grew, read = False, None, 0
    step_seconds = frames_seconds = spectra_seconds = 0.0
    last_spectra = time.perf_counter()
    while True:
        if reason():
            break
        t_now = sim.steps() * dt
        target = min(block_end if t_now >= pulse_end else min(t_now + check, block_end), limit)
        began = time.perf_counter()
        sim.run_until(target)
        for q, ps in enumerate(probes):
            vals = [sim.probe(p)[read:] for p in ps]
            lv = sum(v * v for v in vals)
            times = (read + 1 + np.arange(len(lv))) * dt
            if len(lv):
                peaks[q] = max(peaks[q], float(lv.max()))
                recent[q] = max(recent[q], float(lv.max()))
                late = lv[times > pulse_end]
                if len(late):
                    block[q] = max(block[q], float(late.max()))
        read = sim.steps()
        step_seconds += time.perf_counter() - began
        t = sim.steps() * dt
        if t >= block_end:
            worst = max((b / p if p > 0 else 0.0) for p, b in zip(peaks, block))
            if setup.until[0] == "decay" and worst <= setup.until[1] and np.any(peaks > 0):
                decayed = True
            grew = grew or growth.check(float(block.max()))
            block[:] = 0.0
            block_end += check
        done = decayed or grew is not None or t >= limit
        b0 = time.perf_counter()
        decay = max((r / p if p > 0 else 0.0) for p, r in zip(peaks, recent))
        recent[:] = 0.0
        if not done and time.perf_counter() - last_spectra >= max(SPECTRA_SECONDS, 10.0 * spectra_seconds):
            a = time.perf_counter()
            for e in _spectra(sim, setup, incident, added, wave, freqs, False):
                record(e)
            spectra_seconds = time.perf_counter() - a
            last_spectra = time.perf_counter()
        frames_seconds += time.perf_counter() - b0
        record({"type": "fdtd_progress", "step": sim.steps(), "time_um": t, "until_um": limit, "decay": decay,
                "fraction": setup.until[1], "steps_seconds": step_seconds, "frames_seconds": frames_seconds,
                "cell_updates_per_second": g.nx * g.ny * g.nz * sim.steps() / max(step_seconds, 1e-9)})
        if done:
            break
    for e in _spectra(sim, setup, incident, added, wave, freqs, True):
        record(e)
    if grew is not None:
        record({"type": "warning", "message": f"the fields grew {grew:.3g} times after the pulse: a wave of the grid "
                                              "amplified in a CPML, which a finer grid removes"})
    for kind, m, n in added:
        if kind != "resonance":
            continue
        start = min(math.ceil(pulse_end / dt), len(sim.probe(n)))
        signal = sim.probe(n)[start:].astype(complex)
        lo, hi = min(freqs), max(freqs)
        if not hi > lo:
            lo, hi = setup.carrier - 0.5 * setup.bandwidth, setup.carrier + 0.5 * setup.bandwidth
        q = max(int(math.floor(0.25 / (hi * dt))), 1)
        signal = signal[::q][:RESONANCE_SAMPLES]
        found = _harmonic(signal, dt * q, lo, hi) if len(signal) >= 8 else []
        held = [r for r in found if r.error < 1e-2 and r.decay > 0.0]
        largest = max((abs(r.amplitude) for r in held), default=0.0)
        record({"type": "fdtd_resonances", "monitor": m.name, "position_um": list(m.position),
                "component": f"{'EH'[m.field]}{'xyz'[m.component]}",
                "resonances": [{"wavelength_um": 1.0 / r.frequency, "frequency": r.frequency, "q": r.q(),
                                "decay": r.decay, "amplitude": abs(r.amplitude), "error": r.error}
                               for r in held if abs(r.amplitude) > 1e-6 * largest]})


def _spectra(sim, setup, incident, added, wave, freqs, last):
    """spectra (src/job/fdtd.rs:2248-2435): each mode source's reflection, then each monitor's
    power referred to the incident power (a mode source's launched |a|^2, or the pulse's
    |spectrum|^2 for a dipole's run)."""
    steps, dt = sim.steps(), sim.dt()
    pulse = min(steps, math.ceil((wave.delay + 6.0 * wave.width) / dt) + 1)
    launched = [sim.mode_amplitudes(n) for n, _ in incident]
    referred = all(p.kind != "dipole" for p in setup.sources)
    inc = []
    for k, f in enumerate(freqs):
        w = abs(_spectrum(wave, 1, dt, pulse, f)) ** 2
        total, j = 0.0, 0
        for p in setup.sources:
            if p.kind == "mode":
                a = launched[j][k]
                total += abs(a[0] if incident[j][1] == 0 else a[1]) ** 2
                j += 1
            else:
                total = float("nan")
        inc.append(w if math.isnan(total) else total)
    share = lambda a, b: a / b if b > 0 else 0.0
    per = lambda v: [share(a, b) for a, b in zip(v, inc)]
    quantity = (lambda what: f"{what} / incident power") if referred else \
        (lambda what: f"{what} per unit source spectrum, |field|^2 um^2")
    wl = [1.0 / f for f in freqs]
    t = steps * dt
    out = []
    j = 0
    for s, p in enumerate(setup.sources):
        if p.kind != "mode":
            continue
        n, d = incident[j]
        j += 1
        amps = sim.mode_amplitudes(n)
        fw = [abs(a[0] if d == 0 else a[1]) ** 2 for a in amps]
        bk = [abs(a[1] if d == 0 else a[0]) ** 2 for a in amps]
        out.append({"type": "fdtd_spectrum", "monitor": f"source {s + 1}", "kind": "reflection",
                    "quantity": "power back into the source's mode / power the source launches into it",
                    "wavelengths_um": wl, "series": [{"label": f"back along {'xyz'[p.axis]}",
                                                      "values": [share(b, f) for b, f in zip(bk, fw)]}],
                    "time_um": t, "last": last})
    for kind, m, n in added:
        if kind == "mode":
            amps = sim.mode_amplitudes(n)
            out.append({"type": "fdtd_spectrum", "monitor": m.name, "kind": "mode",
                        "quantity": quantity("power in the guide's mode"), "wavelengths_um": wl,
                        "series": [{"label": f"towards +{'xyz'[m.axis]}", "values": per([abs(a[0]) ** 2 for a in amps])},
                                   {"label": f"towards -{'xyz'[m.axis]}", "values": per([abs(a[1]) ** 2 for a in amps])}],
                        "time_um": t, "last": last})
        elif kind == "flux":
            out.append({"type": "fdtd_spectrum", "monitor": m.name, "kind": "flux", "quantity": quantity("flux"),
                        "wavelengths_um": wl, "series": [{"label": f"towards +{'xyz'[m.axis]}",
                                                          "values": per(list(sim.flux(n)))}],
                        "time_um": t, "last": last})
    return out


def main(argv=None):
    """`python -m photonoxide_amd.job JOB.yaml [RUN_DIR]`: runs the job on device 0, its events
    into RUN_DIR/events.jsonl (the job's name by default), each printed as it is recorded."""
    import sys
    args = list(sys.argv[1:] if argv is None else argv)
    if not args or args[0] in ("-h", "--help"):
        print(main.__doc__)
        return 2
    job = load(args[0])
    out = args[1] if len(args) > 1 else str(job.get("name", "fdtd"))
    for e in run_job(job, out):
        if e["type"] != "fdtd_spectrum" or e["last"]:
            print(json.dumps(e), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
