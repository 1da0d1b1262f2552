"""Synthetic serving jobs: 2D and 3D design problems of the batched-serving workload.

Each job is a grid of 20 nm cells with a 12-cell CPML, a Si/SiO2 design region (a seeded binary
pattern at 50 % fill, Gaussian-smoothed with sigma = 2 cells, eps = eps_SiO2 + rho (eps_Si -
eps_SiO2)), an E_y line (2D) or plane (3D) current with a Gaussian pulse at 1.55 um, and a
5-frequency transform monitor (1.50-1.60 um) on a line or plane past the design region. The
same builder makes the job for the device (photonoxide_amd.Simulation, no device set) and for
the FP64 oracle, so spectra can be compared.
"""
import numpy as np
from scipy.ndimage import gaussian_filter

EPS_SI, EPS_OX = 3.48 ** 2, 1.444 ** 2
H = 0.02
PML = 12
FREQS = 1 / np.linspace(1.50, 1.60, 5)


def job_spec(nx, ny, nz, steps, seed):
    """The job's parameters: per-component permittivity [3][nz][ny][nx], source and monitor."""
    rng = np.random.default_rng(seed)
    two_d = nz == 1
    eps = np.full((nz, ny, nx), EPS_OX)
    # the design region: the middle third along x, the middle half across
    x0, x1 = nx // 3, 2 * nx // 3
    y0, y1 = ny // 4, 3 * ny // 4
    if two_d:
        rho = gaussian_filter((rng.random((y1 - y0, x1 - x0)) < 0.5).astype(np.float64), 2.0)
        eps[0, y0:y1, x0:x1] = EPS_OX + rho * (EPS_SI - EPS_OX)
    else:
        z0, z1 = nz // 3, 2 * nz // 3
        rho = gaussian_filter((rng.random((z1 - z0, y1 - y0, x1 - x0)) < 0.5).astype(np.float64), 2.0)
        eps[z0:z1, y0:y1, x0:x1] = EPS_OX + rho * (EPS_SI - EPS_OX)
    eps3 = np.stack([eps, eps, eps])
    i_src = PML + 4
    i_mon = nx - PML - 4
    jj = np.arange(PML, ny - PML)
    if two_d:
        kk = np.zeros(1, dtype=np.int64)
    else:
        kk = np.arange(PML, nz - PML)
    J, Kk = np.meshgrid(jj, kk)
    index = ((Kk * ny + J) * nx + i_src).ravel()
    lo = (i_mon, PML, 0 if two_d else PML)
    hi = (i_mon, ny - PML - 1, 0 if two_d else nz - PML - 1)
    return dict(n=(nx, ny, nz), steps=steps, eps=eps3, index=index, lo=lo, hi=hi, two_d=two_d)


def pulse_params():
    f0 = 1 / 1.55
    bw = 1 / 1.45 - 1 / 1.65
    width = np.sqrt(2.0 * np.log(2.0)) / (np.pi * bw)
    return f0, width, 6.0 * width


def product_job(spec):
    import photonoxide_amd as P
    nx, ny, nz = spec["n"]
    g = P.Grid3d(nx, ny, nz, H, H, H)
    if spec["two_d"]:
        b = P.Boundaries(P.Pml(PML, PML), P.Pml(PML, PML), P.Periodic())
    else:
        b = P.Boundaries.cpml_all(PML)
    s = P.Simulation.with_eps(g, spec["eps"], b, 0.9)
    f0, width, delay = pulse_params()
    wave = P.Waveform(0, f0, width, delay)
    idx = spec["index"]
    s.add_current_indices(P.Field.E, np.full(idx.size, 1), idx, np.ones(idx.size, dtype=complex), wave)
    s.add_dft(spec["lo"], spec["hi"], FREQS)
    return s


def oracle_job(spec):
    import oracle_lib as O
    nx, ny, nz = spec["n"]
    grid = (nx, ny, nz, H, H, H, 0.0, 0.0, 0.0)
    if spec["two_d"]:
        b = O.bounds(periodic=(0, 0, 1), low=(PML, PML, 0), high=(PML, PML, 0))
    else:
        b = O.bounds(low=(PML,) * 3, high=(PML,) * 3)
    o = O.Sim(grid, b, courant=0.9, eps3n=spec["eps"])
    f0, width, delay = pulse_params()
    idx = spec["index"]
    o.add_current(0, np.full(idx.size, 1), idx, np.ones(idx.size, dtype=complex),
                  O.wave(kind=0, frequency=f0, width=width, delay=delay))
    o.add_dft(spec["lo"], spec["hi"], FREQS)
    return o


def spectra(sim, n=0):
    """Ey and Hz of monitor n, [f][...] complex (the fields a line/plane monitor weighs)."""
    return np.stack([sim.dft(n, 0, 1), sim.dft(n, 1, 2)])


def stream(n2d, n3d, seed, size2d=(128, 1024), size3d=(64, 256), steps=(1000, 4000), rate=None):
    """The serving workload's jobs in arrival order: (arrival s, (nx, ny, nz), steps, seed)."""
    rng = np.random.default_rng(seed)
    jobs = []
    for q in range(n2d):
        nx, ny = rng.integers(size2d[0], size2d[1] + 1, 2)
        jobs.append(((int(nx), int(ny), 1), int(rng.integers(steps[0], steps[1] + 1))))
    for q in range(n3d):
        n = rng.integers(size3d[0], size3d[1] + 1, 3)
        jobs.append(((int(n[0]), int(n[1]), int(n[2])), int(rng.integers(steps[0], steps[1] + 1))))
    order = rng.permutation(len(jobs))
    jobs = [jobs[q] for q in order]
    t = 0.0
    out = []
    for q, (n, st) in enumerate(jobs):
        if rate:
            t += rng.exponential(1.0 / rate)
        out.append((t, n, st, int(seed) * 100003 + q))
    return out
