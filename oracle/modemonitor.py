"""TEST INFRASTRUCTURE ONLY (the CPU oracle): the mode monitor's amplitudes restated from the
reference, as plain Python, with the reference's operations in the reference's order.

- Simulation::mode_amplitudes and fdfd_boundaries (src/fdtd/monitors.rs:623-651): FDFD's
  boundaries for the grid's (a CPML of low and high cells is a PML of as many, 0 a wall; a
  periodic axis is Bloch with k = 0), graded with the CPML's reflection and order, no real
  stretch;
- mode_amplitudes_of (src/fdfd/three/ports.rs:755-767) and Lattice::mode_amplitudes
  (:637-651): forward = N(u, g), backward = -N(u, f), f the mode going forward and g backward;
- Lattice::lorentz (:253-287): N(u, w) = dx dy dz / (4 i k0) sum over the plane's (v, u) and
  components (b, c, axis) of V_r sum_s A_rs (w_r u_s - u_r w_s), s the row's tangential values
  on the next plane; Lattice::volume (:228-240), V_r the product of the stretches;
- Lattice::new (src/fdfd/three.rs:276-308) with graded (src/fdfd/mod.rs:219-242): the stretch
  1 + i sigma (depth/d)^m at the nodes and halfway after them, sigma = (m + 1) (-ln R) /
  (2 k0 d); shift, fixed, neighbour, curl_e, curl_h and curl_curl_into (three.rs:310-418);
- PortMode3d::e and value_at (src/fdfd/three/ports.rs:130-184), as the mode source's.

Complex numbers are (re, im) pairs with num-complex's product and quotient, a real over a
complex as num-complex divides it (a (c - i d) / (c^2 + d^2)), a product of stretches folded
from 1; ln and powf are the reference's portable ones (src/portable/mod.rs:141-220, :299-301),
the oracle's C restatements or_ln and or_pow, which the caller installs with use_portable; complex
exp is the reference's portable one from the oracle's C exp and sincos, which the caller passes.

Only tests/ use this, as the checker of the host mirror's amplitudes. What pins it, beyond its
own formula (tests/test_modemonitor.py):
- N(f, f) is zero in exact arithmetic and here bit for bit (the mode's own field has no
  backward amplitude);
- for a discrete plane wave in a lossless periodic guide, N(f, g) is the scheme's flux of f
  (the flux monitor's formula, the E mean either side times H~ = curl E / (i k0)), to round-off:
  this fixes the form's scale (the 4 i k0, the cell volume, the curl-curl weights);
- a planted error (the cut's sign, k0 off, a stretch dropped) breaks that or the parity with
  the device's monitor;
- on the device, an eigenmode launched by the mode source reads back with its backward
  amplitude at the CPMLs' reflection, and |forward|^2 / N(f, g) is the flux monitor's flux.
"""
import math

from modesource import cdiv, cmul

ONE = (1.0, 0.0)
ZERO = (0.0, 0.0)


def rdiv(a, b):
    """num-complex's real over complex: (a c / n, -a d / n), n = c^2 + d^2."""
    n = b[0] * b[0] + b[1] * b[1]
    return (a * b[0] / n, -a * b[1] / n)


def cadd(a, b):
    return (a[0] + b[0], a[1] + b[1])


def csub(a, b):
    return (a[0] - b[0], a[1] - b[1])


def cneg(a):
    return (-a[0], -a[1])


def cscale(a, x):
    return (a[0] * x, a[1] * x)


_PORTABLE = {}


def use_portable(ln, powf):
    """The portable ln and powf the stretches are graded with (the oracle's or_ln, or_pow)."""
    _PORTABLE["ln"], _PORTABLE["powf"] = ln, powf


def graded(pos, span, layers, h, k0, reflection, order):
    """graded: the stretch at pos on an axis spanning span with layers cells of PML."""
    ln, powf = _PORTABLE["ln"], _PORTABLE["powf"]

    def profile(depth, d):
        sigma = (order + 1.0) * (-ln(reflection)) / (2.0 * k0 * d)
        return (1.0, sigma * powf(min(max(depth / d, 0.0), 1.0), order))
    dl, dh = layers[0] * h, layers[1] * h
    if dl > 0.0 and pos < span[0] + dl:
        return profile(span[0] + dl - pos, dl)
    if dh > 0.0 and pos > span[1] - dh:
        return profile(pos - (span[1] - dh), dh)
    return (1.0, 0.0)


class Lattice:
    """The grid with its boundaries at k0. n, steps, origin per axis; periodic[a] (Bloch k = 0)
    or low[a], high[a] PML cells (0: a wall)."""

    def __init__(self, n, steps, origin, periodic, low, high, reflection, order, k0):
        self.n, self.h, self.periodic = n, steps, periodic
        self.k0 = k0
        self.cells = n[0] * n[1] * n[2]

        def along(a, offset):
            span = (origin[a] + 0.0 * steps[a], origin[a] + n[a] * steps[a])
            layers = (0, 0) if periodic[a] else (low[a], high[a])
            out = []
            for m in range(n[a]):
                pos = origin[a] + m * steps[a] + offset * steps[a]
                s = graded(pos, span, layers, steps[a], k0, reflection, order)
                out.append((s[0] + 0.0 * s[1], s[1]))
            return out
        self.nodes = [along(a, 0.0) for a in range(3)]
        self.halves = [along(a, 0.5) for a in range(3)]

    def index(self, comp, at):
        return comp * self.cells + (at[2] * self.n[1] + at[1]) * self.n[0] + at[0]

    def at(self, r):
        c, rest = divmod(r, self.cells)
        i = rest % self.n[0]
        j = (rest // self.n[0]) % self.n[1]
        k = rest // (self.n[0] * self.n[1])
        return c, (i, j, k)

    def shift(self, a, m, step):
        to = m + step
        if 0 <= to < self.n[a]:
            return to, ONE
        if self.periodic[a]:
            turns = -1.0 if to < 0 else 1.0
            period = self.n[a] * self.h[a]
            ang = turns * 0.0 * period
            return to % self.n[a], (math.cos(ang), math.sin(ang))
        return None

    def fixed(self, r):
        c, at = self.at(r)
        return any(a != c and at[a] == 0 and not self.periodic[a] for a in range(3))

    def neighbour(self, comp, at, a, step):
        s = self.shift(a, at[a], step)
        if s is None:
            return None
        to = list(at)
        to[a] = s[0]
        return self.index(comp, to), s[1]

    def curl_e(self, r):
        a, at = self.at(r)
        b, c = (a + 1) % 3, (a + 2) % 3
        out = []
        for along, field, sign in ((b, c, 1.0), (c, b, -1.0)):
            w = rdiv(sign, cscale(self.halves[along][at[along]], self.h[along]))
            nb = self.neighbour(field, at, along, 1)
            if nb is not None:
                out.append((nb[0], cmul(w, nb[1])))
            out.append((self.index(field, at), cneg(w)))
        return out

    def curl_h(self, r):
        a, at = self.at(r)
        b, c = (a + 1) % 3, (a + 2) % 3
        out = []
        for along, field, sign in ((b, c, 1.0), (c, b, -1.0)):
            w = rdiv(sign, cscale(self.nodes[along][at[along]], self.h[along]))
            out.append((self.index(field, at), w))
            nb = self.neighbour(field, at, along, -1)
            if nb is not None:
                out.append((nb[0], cmul(cneg(w), nb[1])))
        return out

    def curl_curl(self, r):
        if self.fixed(r):
            return []
        row = []
        for face, wh in self.curl_h(r):
            for col, we in self.curl_e(face):
                if not self.fixed(col):
                    row.append((col, cmul(cneg(wh), we)))
        return row

    def volume(self, r):
        comp, at = self.at(r)
        v = ONE
        for a in range(3):
            v = cmul(v, self.halves[a][at[a]] if a == comp else self.nodes[a][at[a]])
        return v

    def lorentz(self, axis, plane, u, w, plant_sign=1.0):
        b, c = (axis + 1) % 3, (axis + 2) % 3
        total = ZERO
        for v in range(self.n[c]):
            for uu in range(self.n[b]):
                for comp in (b, c, axis):
                    at = [0, 0, 0]
                    at[axis], at[b], at[c] = plane, uu, v
                    r = self.index(comp, at)
                    ur, wr = u(r), w(r)
                    acc = ZERO
                    for s, a in self.curl_curl(r):
                        to, sat = self.at(s)
                        if to != axis and sat[axis] == plane + 1:
                            d = csub(cmul(wr, u(s)), cscale(cmul(ur, w(s)), plant_sign))
                            acc = cadd(acc, cmul(a, d))
                    total = cadd(total, cmul(self.volume(r), acc))
        volume = self.h[0] * self.h[1] * self.h[2]
        return cdiv(cscale(total, volume), (0.0, 4.0 * self.k0))


def mode_value(mode, comp, at, forward, exp, sincos):
    """PortMode3d::value_at: the mode's E going forward (or backward) at unit amplitude."""
    axis, p = mode["axis"], mode["plane"]
    ab, ac = (axis + 1) % 3, (axis + 2) % 3

    def cexp(z):
        r = exp(z[0])
        if z[1] == 0.0:
            return (r, z[1])
        sn, cs = sincos(z[1])
        return (r * cs, r * sn)
    u, v = at[ab], at[ac]
    if comp == axis:
        ev = mode["normal"][v][u]
    else:
        ev = mode["tangential"][0 if comp == ab else 1][v][u]
    ib = cmul((0.0, 1.0), mode["beta"])
    ib = (ib[0] * mode["step"], ib[1] * mode["step"])
    st = float(at[axis]) - float(p)
    if forward:
        return cmul(ev, cexp((ib[0] * st, ib[1] * st)))
    neg = (-ib[0], -ib[1])
    phase = cexp((neg[0] * st, neg[1] * st))
    if comp == axis:
        return cmul(cmul((-ev[0], -ev[1]), cexp(neg)), phase)
    return cmul(ev, phase)


def mode_amplitudes(lattice, mode, e, exp, sincos, plant_sign=1.0):
    """(forward, backward) amplitudes of `mode` in the field whose E component c at (i, j, k)
    is e(c, (i, j, k)), an (re, im) pair."""
    axis, plane = mode["axis"], mode["plane"]

    def u(r):
        c, at = lattice.at(r)
        return e(c, at)

    def f(r):
        c, at = lattice.at(r)
        return mode_value(mode, c, at, True, exp, sincos)

    def g(r):
        c, at = lattice.at(r)
        return mode_value(mode, c, at, False, exp, sincos)
    forward = lattice.lorentz(axis, plane, u, g, plant_sign)
    backward = cneg(lattice.lorentz(axis, plane, u, f, plant_sign))
    return forward, backward
