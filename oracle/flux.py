"""TEST INFRASTRUCTURE ONLY (the CPU oracle): flux monitors restated from the reference, as
plain Python over the oracle's FP64 transform sums, in the reference's order of summation.

FluxPlane (src/fdtd/monitors.rs:210-263 range and weight, :267-316 its transforms, :318-350
flux) and Simulation::flux (:530-547). Only tests/ use it, as the checker.

Pinned by the reference's closed-box balance (src/fdtd/monitors_checks.rs:188-250,
monitors_tests.rs:18-25): no net flux leaves a lossless box with no source in it (< 1e-12 of
the largest face flux), and with the source inside, what leaves is what it radiates; a planted
error in a weight or a sign breaks the balance (tests/test_oracle.py).
"""


class Face:
    """A flux plane halfway between nodes `plane` and plane + 1 along `axis`; window[w] =
    (lo, hi) half-planes along the w-th of the axes after it, or None for the whole axis."""

    def __init__(self, axis, plane, window=(None, None), sign=1.0):
        self.axis, self.plane, self.window, self.sign = axis, plane, tuple(window), sign

    def others(self):
        return (self.axis + 1) % 3, (self.axis + 2) % 3

    def range(self, n, w):
        """FluxPlane::range: the indices the window reaches along the w-th other axis."""
        if self.window[w] is None:
            return range(0, n[self.others()[w]])
        lo, hi = self.window[w]
        return range(lo, hi + 1)

    def weight(self, w, m, half):
        """FluxPlane::weight: a value's share, halfway after node m (`half`) or on it."""
        if self.window[w] is None:
            return 1.0
        lo, hi = self.window[w]
        if half:
            return 1.0 if lo < m < hi else (0.5 if m in (lo, hi) else 0.0)
        return 1.0 if lo < m <= hi else 0.0

    def box(self, n):
        """The transform box (inclusive node ranges) the face needs: the nodes either side."""
        low, high = [0] * 3, [0] * 3
        low[self.axis], high[self.axis] = self.plane, self.plane + 1
        for w, along in enumerate(self.others()):
            r = self.range(n, w)
            low[along], high[along] = r.start, r.stop - 1
        return tuple(low), tuple(high)

    def flux(self, n, steps, value, f):
        """FluxPlane::flux: value(field, component, at, f) -> complex from the face's sums."""
        a = self.axis
        b, c = self.others()
        total = 0.0
        for comp, h, sign in ((b, c, 0.5), (c, b, -0.5)):
            for v in self.range(n, 1):
                for u in self.range(n, 0):
                    weight = self.weight(0, u, comp == b) * self.weight(1, v, comp == c)
                    if weight == 0.0:
                        continue
                    at = [0, 0, 0]
                    at[a], at[b], at[c] = self.plane, u, v
                    low = value(0, comp, tuple(at), f)
                    hv = value(1, h, tuple(at), f)
                    at[a] = self.plane + 1
                    high = value(0, comp, tuple(at), f)
                    m_re, m_im = 0.5 * (low.real + high.real), 0.5 * (low.imag + high.imag)
                    # Re(m conj(hv)) as num-complex multiplies: m.re hv.re - m.im (-hv.im)
                    total += weight * sign * (m_re * hv.real - m_im * -hv.imag)
        return total * steps[b] * steps[c]


def box_faces(planes):
    """Simulation::add_flux_box's six faces (monitors.rs:503-528), outwards."""
    faces = []
    for a in range(3):
        lo, hi = planes[a]
        b, c = (a + 1) % 3, (a + 2) % 3
        for plane, sign in ((lo, -1.0), (hi, 1.0)):
            faces.append(Face(a, plane, (planes[b], planes[c]), sign))
    return faces


def total_flux(faces, n, steps, values, f):
    """Simulation::flux at frequency f: the faces' signed fluxes summed from -0.0 (Rust's Sum
    for f64); values[q] is face q's value function."""
    total = -0.0
    for face, value in zip(faces, values):
        total += face.sign * face.flux(n, steps, value, f)
    return total
