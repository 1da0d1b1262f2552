"""The workloads the time-stepper is quoted on, as parity-test specs (the format of fixtures.py):

- `te_strip_2d`: the plain 2D TE problem, 400 x 400 cells of 20 nm, a 500 nm Si strip
  (n = 3.48) along x in SiO2 (n = 1.444), a 12-cell CPML on x and y (z periodic), Courant 0.5, a
  Gaussian-pulse E_y line current across the strip centred at 1.55 um with a 200 nm bandwidth,
  and a transform line across the strip downstream at 5 wavelengths from 1.50 to 1.60 um;
- `mmi_1x2`: the 1x2 MMI on SOI (the geometry of bench.py's workload: a 220 nm Si core,
  n = 3.48, in SiO2, 500 nm arms, a 3 x 5 um body), permittivity sampled per E value, a
  Gaussian-profiled E_y pulse current across the input arm at 1.55 um, and 21-frequency transform
  windows over both output ports; at 96 x 64 x 48 cells of 80 nm the oracle steps it in seconds;
- `dispersive_3d`: the dispersive 3D problem at a size the oracle steps in seconds: Si pillars
  (a 2-pole Lorentz medium) and Au blocks (Drude + 2 Lorentz) in SiO2, seeded positions and
  radii, a 16-cell CPML on every face (fewer on small grids), a broadband pulse 1.2-2.0 um.
"""
import math

import numpy as np

INF = float("inf")
N_SI, N_OX = 3.48, 1.444


def pulse(frequency, bandwidth):
    """Waveform::pulse (src/fdtd.rs:300-312): width sqrt(2 ln 2)/(pi bw), peak 6 widths in."""
    width = math.sqrt(2.0 * math.log(2.0)) / (math.pi * bandwidth)
    return dict(kind=0, frequency=frequency, width=width, delay=6.0 * width)


def te_strip_2d(n=400, h=0.02, pml=12):
    x0 = -n * h / 2
    grid = (n, n, 1, h, h, h, x0, x0, 0.0)
    i_src = int(round((-3.0 - x0) / h))
    i_mon = int(round((3.0 - x0) / h))
    rows = range(pml, n - pml)
    wave = pulse(1 / 1.55, 1 / 1.45 - 1 / 1.65)
    return dict(
        grid=grid, bounds=dict(periodic=(0, 0, 1), low=(pml, pml, 0), high=(pml, pml, 0)),
        courant=0.5,
        eps=(N_OX**2, [(0, (-INF, -0.25, -INF), (INF, 0.25, INF), N_SI**2)]),
        currents=[(0, [(1, (i_src, j, 0), 1.0 + 0j) for j in rows], wave)],
        probes=[(0, 1, (i_mon, n // 2, 0)), (1, 2, (i_mon, n // 2, 0))],
        dft=((i_mon, pml, 0), (i_mon, n - pml - 1, 0), list(1 / np.linspace(1.50, 1.60, 5))),
    )


# Si: two Lorentz poles; Au: a Drude pole and two Lorentz poles (eps_inf, poles) in the oracle's
# pole format: ("lorentz", strength, resonance, damping) / ("drude", plasma, damping), c/um units
SI_MEDIUM = (7.0, [("lorentz", 3.2, 3.9, 0.0), ("lorentz", 1.8, 2.2, 0.01)])
AU_MEDIUM = (1.5, [("drude", 8.9, 0.07), ("lorentz", 1.2, 2.6, 0.5), ("lorentz", 2.4, 3.4, 1.1)])


def dispersive_3d(n=(64, 64, 32), h=0.015, pml=8, pillars=6, blocks=3, seed=1):
    """Seeded pillars (full height, radius U[80, 200] nm scaled to the grid) and blocks."""
    nx, ny, nz = n
    x0, y0, z0 = -nx * h / 2, -ny * h / 2, -nz * h / 2
    grid = (nx, ny, nz, h, h, h, x0, y0, z0)
    rng = np.random.default_rng(seed)
    inner = [(x0 + (pml + 2) * h, -x0 - (pml + 2) * h), (y0 + (pml + 2) * h, -y0 - (pml + 2) * h)]
    scale = min(nx, ny) / 512
    si, au = [], []
    for _ in range(pillars):
        r = rng.uniform(0.08, 0.2) * max(scale, 0.25)
        cx, cy = (rng.uniform(a + r, b - r) for a, b in inner)
        si.append((1, (cx, cy, -INF), (r, 0, INF), 1.0))
    for _ in range(blocks):
        w = rng.uniform(0.05, 0.12)
        cx, cy = (rng.uniform(a + w, b - w) for a, b in inner)
        cz = rng.uniform(z0 + (pml + 2) * h + w, -z0 - (pml + 2) * h - w)
        au.append((0, (cx - w, cy - w, cz - w), (cx + w, cy + w, cz + w), 1.0))
    mid = (nx // 2, ny // 2, nz // 2)
    return dict(
        grid=grid, bounds=dict(periodic=(0, 0, 0), low=(pml,) * 3, high=(pml,) * 3),
        courant=0.9, eps=(N_OX**2, []),
        sources=[(0, 1, (pml + 3, ny // 2, nz // 2), pulse(1 / 1.6, 1 / 1.2 - 1 / 2.0))],
        probes=[(0, 1, mid), (1, 2, (nx - pml - 3, ny // 2, nz // 2))],
        dft=((nx - pml - 4, pml, pml), (nx - pml - 4, ny - pml - 1, nz - pml - 1), [1 / 1.3, 1 / 1.55, 1 / 1.9]),
        media=[(SI_MEDIUM[0], SI_MEDIUM[1], (0.0, si)), (AU_MEDIUM[0], AU_MEDIUM[1], (0.0, au))],
    )


def mmi_1x2(nx=96, ny=64, nz=48, h=0.08, pml=12, port=(6, 5)):
    """bench.py's MMI workload on any grid: x from 0, y and z centred; the ports' windows are
    2 port[0] values across y by 2 port[1] across z around each output arm, 10 cells before the
    high x CPML."""
    x0, y0, z0 = 0.0, -ny * h / 2, -nz * h / 2
    grid = (nx, ny, nz, h, h, h, x0, y0, z0)
    e_si, e_ox = N_SI**2, N_OX**2
    x_body0, x_body1, w_arm, w_body, t = 1.2, 6.2, 0.5, 3.0, 0.22
    y_out = w_body / 4

    def eps_at(c):
        i = np.arange(nx)[None, None, :]
        j = np.arange(ny)[None, :, None]
        k = np.arange(nz)[:, None, None]
        x = x0 + i * h + (0.5 * h if c == 0 else 0.0)
        y = y0 + j * h + (0.5 * h if c == 1 else 0.0)
        z = z0 + k * h + (0.5 * h if c == 2 else 0.0)
        core = np.abs(z) < t / 2
        inp = (x < x_body0) & (np.abs(y) < w_arm / 2)
        body = (x >= x_body0) & (x <= x_body1) & (np.abs(y) < w_body / 2)
        out = (x > x_body1) & ((np.abs(y - y_out) < w_arm / 2) | (np.abs(y + y_out) < w_arm / 2))
        return np.where(core & (inp | body | out), e_si, e_ox).astype(np.float64)
    eps3n = np.stack([np.broadcast_to(eps_at(c), (nz, ny, nx)) for c in range(3)])
    i_src = pml + min(10, max(1, (nx - 2 * pml) // 8))
    jj, kk = np.meshgrid(np.arange(ny), np.arange(nz))
    prof = np.exp(-((y0 + jj * h + 0.5 * h) / 0.3) ** 2 - ((z0 + kk * h) / 0.15) ** 2)
    keep = prof > 1e-3
    cur = [(1, (i_src, int(j), int(k)), complex(v)) for j, k, v in zip(jj[keep], kk[keep], prof[keep])]
    freqs = list(1 / np.linspace(1.50, 1.60, 21))
    i_mon = nx - pml - min(10, max(1, (nx - 2 * pml) // 8))
    dfts = []
    for yc in (y_out, -y_out):
        jc = int(round((yc - y0) / h))
        dfts.append(((i_mon, jc - port[0], nz // 2 - port[1]), (i_mon, jc + port[0] - 1, nz // 2 + port[1] - 1), freqs))
    return dict(
        grid=grid, bounds=dict(periodic=(0, 0, 0), low=(pml,) * 3, high=(pml,) * 3), courant=0.9,
        eps3n=eps3n, currents=[(0, cur, pulse(1 / 1.55, 1 / 1.45 - 1 / 1.65))],
        probes=[(0, 1, (i_src + 2, ny // 2, nz // 2)), (0, 1, (i_mon, int(round((y_out - y0) / h)), nz // 2))],
        dfts=dfts,
    )
