"""TEST INFRASTRUCTURE ONLY (the CPU oracle): a mode source restated from the reference, as
plain Python: the currents Simulation::add_mode_source (src/fdtd/sources.rs:550-596) adds for a
mode given by its profile, with the reference's operations in the reference's order.

- PortMode3d::e and value_at (src/fdfd/three/ports.rs:125-184): the mode's E on the planes
  around its own, e^(i beta step (m - p)) going forward; going backward e^(-i beta ...), with
  the normal component's sign and half-step phase changed;
- Simulation::h_of (sources.rs:457-483): H~ = curl E / (i k~), forward differences, zero
  beyond a wall, the curl's factors kappa/step at the halves;
- Simulation::links (sources.rs:321-418) over the planes p - 1 to p + 2 (all of the other two
  axes), the total field from the mode's plane on (forward) or up to the plane after it
  (backward), twice (:155-161);
- Simulation::phasor_surface (sources.rs:419-453): per target field, E then H, the links' sums
  per (component, value) in that order, zeros dropped;
- carrier_wavenumber (:485-508): k~ = (2/dt) sin(w dt / 2).

Complex numbers are (re, im) pairs with num-complex's product and quotient written out, and
complex exp is the reference's portable one (exp(re) (cos im, sin im)), from the oracle's C
exp and sincos, which the caller passes. Only tests/ use this, as the checker: its currents run
on the oracle's stepper. Pinned by physics, not by its own formula: a discrete eigenmode
launched with it goes one way (the backward field is round-off and the CPMLs' reflections),
and a planted sign or phase error sends a large part backward (tests/test_modesource.py).
"""
import math

E, H = 0, 1


def cmul(a, b):
    return (a[0] * b[0] - a[1] * b[1], a[0] * b[1] + a[1] * b[0])


def cdiv(a, b):
    n = b[0] * b[0] + b[1] * b[1]
    return ((a[0] * b[0] + a[1] * b[1]) / n, (a[1] * b[0] - a[0] * b[1]) / n)


def twice(field, comp, axis, m):
    half = comp == axis if field == E else comp != axis
    return 2 * m + (1 if half else 0)


def links(n, periodic, cb, fe, fh, region, total):
    """Simulation::links: (target, component, value, weight, from component, from) tuples.
    cb[c][r] the E coefficient (0 where E is held), fe/fh[a][i] kappa/step at nodes/halves."""
    def index(at):
        return (at[2] * n[1] + at[1]) * n[0] + at[0]

    def neighbour(at, axis, forward):
        m = list(at)
        if forward:
            if at[axis] + 1 < n[axis]:
                m[axis] += 1
            elif periodic[axis]:
                m[axis] = 0
            else:
                return None
        elif at[axis] > 0:
            m[axis] -= 1
        elif periodic[axis]:
            m[axis] = n[axis] - 1
        else:
            return None
        return tuple(m)

    out = []
    for k in region[2]:
        for j in region[1]:
            for i in region[0]:
                at = (i, j, k)
                r = index(at)
                for c in range(3):
                    a, b = (c + 1) % 3, (c + 2) % 3
                    if cb[c][r] != 0.0:
                        inside = total(E, c, at)
                        fa, fb = fe[a][at[a]], fe[b][at[b]]
                        for fc, frm, coef in ((b, at, fa), (b, neighbour(at, a, False), -fa),
                                              (a, at, -fb), (a, neighbour(at, b, False), fb)):
                            if frm is None:
                                continue
                            if total(H, fc, frm) != inside:
                                out.append((E, c, r, (-1.0 if inside else 1.0) * coef, fc, frm))
                    inside = total(H, c, at)
                    fa, fb = fh[a][at[a]], fh[b][at[b]]
                    for fc, frm, coef in ((b, neighbour(at, a, True), fa), (b, at, -fa),
                                          (a, neighbour(at, b, True), -fb), (a, at, fb)):
                        if frm is None or cb[fc][index(frm)] == 0.0:
                            continue
                        if total(E, fc, frm) != inside:
                            out.append((H, c, r, (1.0 if inside else -1.0) * coef, fc, frm))
    return out


def mode_currents(n, periodic, cb, fe, fh, dt, mode, direction, frequency, exp, sincos,
                  flip_h=False, beta_scale=1.0):
    """The currents [(field, [(component, value, (re, im))])] of a mode source. `mode`: dict
    with axis, plane, step, beta (re, im), tangential[2][size_c][size_b] and normal[size_c][size_b]
    as (re, im) pairs. `exp(x)` and `sincos(x) -> (s, c)` are the reference's portable ones.
    flip_h and beta_scale plant errors (for the tests that show the checks catch them)."""
    axis, p = mode["axis"], mode["plane"]
    ab, ac = (axis + 1) % 3, (axis + 2) % 3
    s, _ = sincos(2.0 * math.pi * frequency * dt / 2.0)
    k0 = 2.0 / dt * s
    forward = direction == 0
    beta = (mode["beta"][0] * beta_scale, mode["beta"][1] * beta_scale)

    def cexp(z):
        r = exp(z[0])
        if z[1] == 0.0:
            return (r, z[1])
        sn, cs = sincos(z[1])
        return (r * cs, r * sn)

    def profile(comp, u, v):
        if comp == axis:
            return mode["normal"][v][u]
        return mode["tangential"][0 if comp == ab else 1][v][u]

    ib = cmul((0.0, 1.0), beta)
    ib = (ib[0] * mode["step"], ib[1] * mode["step"])

    def e_of(comp, at):
        st = float(at[axis]) - float(p)
        ev = profile(comp, at[ab], at[ac])
        if forward:
            return cmul(ev, cexp((ib[0] * st, ib[1] * st)))
        neg = (-ib[0], -ib[1])
        phase = cexp((neg[0] * st, neg[1] * st))
        if comp == axis:
            return cmul(cmul((-ev[0], -ev[1]), cexp(neg)), phase)
        return cmul(ev, phase)

    def h_of(comp, at):
        a, b = (comp + 1) % 3, (comp + 2) % 3

        def nxt(ax, f):
            m = list(at)
            if at[ax] + 1 < n[ax]:
                m[ax] += 1
            elif periodic[ax]:
                m[ax] = 0
            else:
                return (0.0, 0.0)
            return e_of(f, tuple(m))
        fa, fb = fh[a][at[a]], fh[b][at[b]]
        da, eb, db, ea = nxt(a, b), e_of(b, at), nxt(b, a), e_of(a, at)
        t1 = ((da[0] - eb[0]) * fa, (da[1] - eb[1]) * fa)
        t2 = ((db[0] - ea[0]) * fb, (db[1] - ea[1]) * fb)
        h = cdiv((t1[0] - t2[0], t1[1] - t2[1]), (0.0, k0))
        return (-h[0], -h[1]) if flip_h else h

    def total(field, comp, at):
        t = twice(field, comp, axis, at[axis])
        return t >= 2 * p if forward else t <= 2 * (p + 1)

    region = [range(n[0]), range(n[1]), range(n[2])]
    region[axis] = range(p - 1, min(p + 3, n[axis]))
    ls = links(n, periodic, cb, fe, fh, region, total)
    out = []
    for field in (E, H):
        values = {}
        for (target, c, r, w, fc, frm) in ls:
            if target != field:
                continue
            inc = h_of(fc, frm) if target == E else e_of(fc, frm)
            acc = values.get((c, r), (0.0, 0.0))
            values[(c, r)] = (acc[0] + w * inc[0], acc[1] + w * inc[1])
        vals = [(c, r, a) for (c, r), a in sorted(values.items()) if a != (0.0, 0.0)]
        if vals:
            out.append((field, vals))
    return out
