"""TEST INFRASTRUCTURE ONLY (the CPU oracle): designs and the frequency-domain adjoint restated
from the reference (src/fdtd/adjoint.rs), as plain Python and numpy. Only tests/ use it, as the
checker.

- design_values: Design::values (adjoint.rs:196-255) with Design::cell_at (:148-167) and
  Design::index (:95-103): every value of E whose cell reaches the design's cells, and each
  design cell's share of it.
- basis (adjoint.rs:1011-1031), spectrum (Waveform::spectrum, src/fdtd.rs:353-396) and
  amplitudes (Simulation::amplitudes, adjoint.rs:571-621): the real amplitudes of the adjoint's
  currents whose transforms are the targets.
- permittivity_gradient (adjoint.rs:831-855) and density_gradient (Simulation::gradient,
  :857-879): sum over frequencies of 2 Re(i w~ E_adj E), chained through the shares.

Pinned by (tests/test_design.py): with_design's permittivity is the oracle's averaging of a scene
with one box per design cell, bit for bit; every share is the derivative of that average with
respect to the cell's density over (eps1 - eps0) (the average is linear in the densities across
the component); the solved amplitudes reproduce their targets through an independent sum of
the pulses' transforms; the gradient, run by the reference's method on the device's transforms,
agrees with the reverse-time adjoint and with central differences (tests/test_design_gpu.py).
A planted error (a share off by one sample, a dropped image term) fails them.
"""
import math

import numpy as np

SAMPLES = 8


def design_index(cells, i, j, k):
    """Design::index: i fastest, then j, then k; None outside."""
    (x0, x1), (y0, y1), (z0, z1) = cells
    if x0 <= i < x1 and y0 <= j < y1 and z0 <= k < z1:
        return ((k - z0) * (y1 - y0) + (j - y0)) * (x1 - x0) + (i - x0)
    return None


def cell_at(grid, periodic, cells, p):
    """Design::cell_at: the design cell holding point p (a periodic side wrapped)."""
    origin = (grid.x0, grid.y0, grid.z0)
    at = []
    for a in range(3):
        n = grid.n(a)
        m = math.floor((p[a] - origin[a]) / grid.step(a))
        if periodic[a]:
            m %= n
        elif not 0 <= m < n:
            return None
        at.append(m)
    return design_index(cells, *at)


def design_values(grid, periodic, cells):
    """Design::values: [(component, (i, j, k), {design cell: share})], component by component,
    then k, j, i."""
    n = float(SAMPLES)
    offset = lambda s: (s + 0.5) / n - 0.5
    h = (grid.dx, grid.dy, grid.dz)
    out = []
    for c in range(3):
        cand = []
        for a in range(3):
            lo, hi = cells[a]
            last = hi - 1 if a == c else hi
            ln = grid.n(a)
            vals = set()
            for m in range(lo, last + 1):
                if m < ln:
                    vals.add(m)
                elif periodic[a]:
                    vals.add(m % ln)
            cand.append(sorted(vals))
        b, cc = (c + 1) % 3, (c + 2) % 3
        for k in cand[2]:
            for j in cand[1]:
                for i in cand[0]:
                    centre = grid.e_position(c, i, j, k)
                    shares = {}
                    p = list(centre)
                    for sb in range(SAMPLES):
                        p[b] = centre[b] + offset(sb) * h[b]
                        for sc in range(SAMPLES):
                            p[cc] = centre[cc] + offset(sc) * h[cc]
                            p[c] = centre[c]
                            v = cell_at(grid, periodic, cells, p)
                            if v is not None:
                                shares[v] = shares.get(v, 0.0) + 1.0 / (n * n)
                    if shares:
                        out.append((c, (i, j, k), dict(sorted(shares.items()))))
    return out


def pulse(frequency, bandwidth):
    """Waveform::pulse (src/fdtd.rs:292-312): (frequency, width, delay)."""
    width = math.sqrt(2.0 * math.log(2.0)) / (math.pi * bandwidth)
    return frequency, width, 6.0 * width


def basis(frequencies, bandwidth):
    """basis: per frequency, the pulse on it and the same a quarter period later."""
    out = []
    for f in frequencies:
        f0, width, delay = pulse(f, bandwidth)
        for shift in (0.0, 0.25 / f):
            out.append((f0, width, delay + shift))
    return out


def gaussian_at(w, t):
    """Waveform::at of a Gaussian pulse (src/fdtd.rs:399-412)."""
    f, width, delay = w
    u = (t - delay) / width
    return math.exp(-u * u) * math.sin(2.0 * math.pi * f * (t - delay))


def spectrum(w, field, dt, steps, frequency):
    """Waveform::spectrum: sum of w(t_n) e^(i w t_n) dt at the times a current on `field` takes its
    values (E: (n + 1/2) dt, H~: n dt)."""
    t = np.arange(steps) * dt + (0.5 * dt if field == 0 else 0.0)
    f, width, delay = w
    u = (t - delay) / width
    v = np.exp(-u * u) * np.sin(2.0 * np.pi * f * (t - delay))
    return complex(np.sum(v * np.exp(1j * 2.0 * np.pi * frequency * t)) * dt)


def amplitudes(field, waves, frequencies, targets, dt):
    """Simulation::amplitudes: per target (one complex value per frequency) the real amplitude of
    each pulse, from the real system sum_m b_m W_m(w_j) = target_j."""
    m = len(waves)
    steps = max(math.ceil((d + 8.0 * wd) / dt) + 1 for (_, wd, d) in waves)
    a = np.zeros((m, m))
    for q, w in enumerate(waves):
        for j, f in enumerate(frequencies):
            s = spectrum(w, field, dt, steps, f)
            a[2 * j, q], a[2 * j + 1, q] = s.real, s.imag
    rhs = np.zeros((m, len(targets)))
    for col, t in enumerate(targets):
        for j, v in enumerate(t):
            rhs[2 * j, col], rhs[2 * j + 1, col] = v.real, v.imag
    return np.linalg.solve(a, rhs).T, steps


def leapfrog_wavenumber(frequency, dt):
    """Simulation::leapfrog_wavenumber (src/fdtd/sources.rs:171-173)."""
    return 2.0 / dt * math.sin(2.0 * math.pi * frequency * dt / 2.0)


def permittivity_gradient(forward, adjoint, frequencies, dt):
    """Simulation::permittivity_gradient: forward, adjoint [frequency][value] complex transforms
    of the values a design monitor reads; dF/deps per value."""
    out = np.zeros(forward.shape[1])
    for q in range(forward.shape[1]):
        s = -0.0
        for f, freq in enumerate(frequencies):
            w = leapfrog_wavenumber(freq, dt)
            s += 2.0 * ((1j * w) * adjoint[f, q] * forward[f, q]).real
        out[q] = s
    return out


def density_gradient(values, per_value, eps, count):
    """Simulation::gradient: dF/deps per value chained through each design cell's share."""
    out = np.zeros(count)
    scale = eps[1] - eps[0]
    for (c, at, shares), g in zip(values, per_value):
        for v, share in shares.items():
            out[v] += g * share * scale
    return out
