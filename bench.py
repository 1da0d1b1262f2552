"""Throughput of the FDTD time-stepper on one MI355X: 3D Yee + CPML in f32, Gcell-updates/s and
the fused step kernel's fraction of the HBM roofline, beside the CPU baseline.

Workload (the single-GPU configuration the metric is quoted on): a 1x2 MMI on SOI,
384 x 256 x 192 cells of 20 nm (18.9 M cells), a 220 nm Si core (n = 3.48) in SiO2 (n = 1.444),
500 nm arms, a 3 um wide MMI body, a 12-cell CPML on every face, Courant 0.9, a Gaussian-pulse
current at 1.55 um (200 nm bandwidth) with a fundamental-mode-like profile across the input arm,
and 21-frequency transform monitors on both output ports. The MMI body is 5 um long: 10 um does
not fit 384 x 20 nm = 7.68 um. The permittivity is sampled at each E value (no subpixel average:
the geometry is grid-aligned, and the kernel's work does not depend on it).

A step is one pass of the whole path over the grid: the fused Yee + CPML kernel, the monitors'
taps and their FP64 accumulation. `value` is cells x steps / wall time of K steps with the fields
resident in HBM. The roofline entry is for the fused step (the march kernel's launches of one step): algorithmic
bytes per step (E and H read and written, 48 B per cell, plus the coefficients read: 1 B per cell
for a uint8 material id, 24 B for full ca, cb arrays, plus 8 B per CPML psi value, read and
written) over the step's kernel time, from HIP events on the resident's stream around the
launches. `traffic` is the HBM bytes per step that rocprofv3's FETCH_SIZE and WRITE_SIZE counted
(profiles/traffic_march64_step_mmi_f32.json).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0


def mmi_problem(nx=384, ny=256, nz=192, h=0.02, pml=12, plain=False, pml_axes="xyz"):
    import photonoxide_amd as P
    g = P.Grid3d(nx, ny, nz, h, h, h, 0.0, -ny * h / 2, -nz * h / 2)
    n_si, n_ox = 3.48, 1.444
    e_si, e_ox = n_si**2, n_ox**2
    x_body0, x_body1 = 1.2, 6.2
    w_arm, w_body = 0.5, 3.0
    y_out = w_body / 4
    t = 0.22

    def eps_at(c):
        i = np.arange(nx)[None, None, :]
        j = np.arange(ny)[None, :, None]
        k = np.arange(nz)[:, None, None]
        x = g.x0 + i * h + (0.5 * h if c == 0 else 0.0)
        y = g.y0 + j * h + (0.5 * h if c == 1 else 0.0)
        z = g.z0 + k * h + (0.5 * h if c == 2 else 0.0)
        core = np.abs(z) < t / 2
        inp = (x < x_body0) & (np.abs(y) < w_arm / 2)
        body = (x >= x_body0) & (x <= x_body1) & (np.abs(y) < w_body / 2)
        out = (x > x_body1) & ((np.abs(y - y_out) < w_arm / 2) | (np.abs(y + y_out) < w_arm / 2))
        si = core & (inp | body | out)
        return np.where(si, e_si, e_ox).astype(np.float64)

    eps = np.stack([np.broadcast_to(eps_at(c), (nz, ny, nx)) for c in range(3)])
    if pml_axes == "xyz":
        edges = P.Boundaries.cpml_all(pml)
    else:  # the other axes periodic: one CPML slab variant's units at a time, for characterization
        edges = P.Boundaries(*[P.Pml(pml, pml) if a in pml_axes else P.Periodic() for a in "xyz"])
    s = P.Simulation.with_eps(g, eps, edges, 0.9)
    if plain:
        return s, g
    # the source: E_y across the input arm with a Gaussian profile, a pulse at 1.55 um
    f0 = 1 / 1.55
    wave = P.Waveform.pulse(f0, 1 / 1.45 - 1 / 1.65)
    i_src = pml + 10
    jj, kk = np.meshgrid(np.arange(ny), np.arange(nz))
    y = g.y0 + jj * h + 0.5 * h
    z = g.z0 + kk * h
    prof = np.exp(-(y / 0.3) ** 2 - (z / 0.15) ** 2)
    keep = prof > 1e-3
    idx = (kk[keep] * ny + jj[keep]) * nx + i_src
    s.add_current_indices(P.Field.E, np.full(idx.size, 1), idx, prof[keep].astype(complex), wave)
    # 21-frequency monitors on the two output ports, over the arm and its surroundings
    freqs = 1 / np.linspace(1.50, 1.60, 21)
    i_mon = nx - pml - 10
    for yc in (y_out, -y_out):
        jc = int(round((yc - g.y0) / h))
        s.add_dft((i_mon, jc - 25, nz // 2 - 15), (i_mon, jc + 24, nz // 2 + 14), freqs)
    return s, g


# Si: eps_inf and two Lorentz terms; Au: a Drude term and two Lorentz terms, each as
# ("lorentz", strength, resonance, damping) or ("drude", plasma, damping), frequencies in c/um
SI_MEDIUM = (7.0, [("lorentz", 3.2, 3.9, 0.0), ("lorentz", 1.8, 2.2, 0.01)])
AU_MEDIUM = (1.5, [("drude", 7.25, 0.053), ("lorentz", 1.2, 2.6, 0.5), ("lorentz", 2.4, 3.4, 1.1)])


def dispersive_problem(nx=512, ny=512, nz=256, h=0.015, pml=16, seed=1):
    """The dispersive 3D workload: 512 x 512 x 256 cells of 15 nm in SiO2 (n = 1.444), 40 Si
    pillars (2-pole Lorentz medium, full height, radius U[80, 200] nm) and 10 Au blocks (Drude +
    2 Lorentz, sides U[100, 300] nm) at seeded positions, a 16-cell CPML on every face, Courant
    0.9, a broadband pulse (1.2-2.0 um) as an E_y current sheet across the grid near its low x
    face, and a 5-frequency transform plane near the high x face."""
    import photonoxide_amd as P
    x0, y0, z0 = -nx * h / 2, -ny * h / 2, -nz * h / 2
    g = P.Grid3d(nx, ny, nz, h, h, h, x0, y0, z0)
    s = P.Simulation.new(g, P.Scene(1.444**2), P.Boundaries.cpml_all(pml), 0.9)
    rng = np.random.default_rng(seed)
    lo_x, hi_x = x0 + (pml + 20) * h, -x0 - (pml + 20) * h
    lo_y, hi_y = y0 + (pml + 4) * h, -y0 - (pml + 4) * h
    pillars = []
    for _ in range(40):
        r = rng.uniform(0.08, 0.2)
        c = (rng.uniform(lo_x + r, hi_x - r), rng.uniform(lo_y + r, hi_y - r))
        pillars.append(P.Cylinder(c, r, (-np.inf, np.inf), 1.0))
    blocks = []
    for _ in range(10):
        w = rng.uniform(0.1, 0.3) / 2
        c = (rng.uniform(lo_x + w, hi_x - w), rng.uniform(lo_y + w, hi_y - w),
             rng.uniform(z0 + (pml + 4) * h + w, -z0 - (pml + 4) * h - w))
        blocks.append(P.Box(tuple(q - w for q in c), tuple(q + w for q in c), 1.0))
    s.with_medium(SI_MEDIUM[0], SI_MEDIUM[1], P.Scene(0.0, pillars))
    s.with_medium(AU_MEDIUM[0], AU_MEDIUM[1], P.Scene(0.0, blocks))
    wave = P.Waveform.pulse((1 / 1.2 + 1 / 2.0) / 2, 1 / 1.2 - 1 / 2.0)
    i_src = pml + 8
    jj, kk = np.meshgrid(np.arange(pml, ny - pml), np.arange(pml, nz - pml))
    idx = ((kk * ny + jj) * nx + i_src).ravel()
    s.add_current_indices(P.Field.E, np.full(idx.size, 1), idx, np.ones(idx.size, complex), wave)
    i_mon = nx - pml - 8
    s.add_dft((i_mon, pml, pml), (i_mon, ny - pml - 1, nz - pml - 1), 1 / np.linspace(1.2, 2.0, 5))
    return s, g


def ade_bytes_per_step(sim):
    """The media state each step reads and writes, f32: E at the previous step (4 B read, 4 B
    written) and each term's two currents (8 B read, 8 B written) per value in a medium."""
    import ctypes
    from photonoxide_amd import _native as N
    b = ctypes.c_int64()
    N.check(N.lib.pxfdtd_simulation_media_bytes_per_step(sim._h, 0, ctypes.byref(b)))
    return int(b.value)


def algorithmic_bytes_per_step(nx, ny, nz, pml, coef_bytes=1):
    n = nx * ny * nz
    psi = 0
    for axis_n in (nx, ny, nz):
        psi += 4 * (n // axis_n) * (2 * pml)  # 2 fields x 2 components, both sides
    return n * (48 + coef_bytes) + 8 * psi, psi


def host_threads():
    """The threads the CPU baseline uses: the cores this process may run on, within the
    host's thread budget (OMP_NUM_THREADS where it is set)."""
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    budget = os.environ.get("OMP_NUM_THREADS", "")
    return min(cores, int(budget)) if budget.isdigit() and int(budget) > 0 else cores


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(nx, ny, nz, pml):
    """The reference's CPU kernel restated (oracle/cpu_baseline.c), timed with the reference's
    own protocol, on the host's threads, with the host's STREAM triad beside it."""
    import ctypes
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    threads = host_threads()
    steps = ctypes.c_long()
    s, rows = ctypes.c_int(), ctypes.c_int()
    O.lib.or_cpu_auto_blocking(nx, ny, nz, threads, ctypes.byref(s), ctypes.byref(rows))
    rate = O.lib.or_cpu_rate_f32(nx, ny, nz, pml, threads, ctypes.byref(steps)) / 1e9
    triad = O.lib.or_cpu_triad(1 << 26, 5, threads)
    how = (f"Malas diamonds of {s.value} steps in tiles of {rows.value} rows (Blocking::auto, "
           f"src/fdtd/kernel/blocked.rs)" if s.value else "the whole grid a step at a time")
    return {"value": round(rate, 4), "unit": "Gcell-updates/s", "cores": threads, "kind": "port",
            "sample": f"the reference's CPU row kernel restated in C/OpenMP (oracle/cpu_baseline.c), f32, "
                      f"{nx}x{ny}x{nz} cells with a {pml}-cell CPML, stepped by {how}, timed as "
                      f"src/fdtd/gpu/rates.rs times: steps doubled from 4 until a run takes > 0.3 s, best of 3 "
                      f"runs of {steps.value} steps, fields first touched by the threads that step them",
            "blocking": {"steps": s.value, "rows": rows.value},
            "nproc": os.cpu_count(), "affinity_cores": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None,
            "cpu_model": cpu_model(), "triad_gb_per_s": round(triad, 1),
            "triad": "a = b + s c over 2^26 doubles, 24 B an element, best of 5 (src/traffic.rs:59-80)"}


# the files the step kernels and their launches are built from
STEP_SOURCES = ["common.hpp", "kernels.hpp", "march64.hpp", "march2d.hpp", "march2s.hpp", "resident.hip",
                "step_launch.hpp", "step_launch_impl.hpp"] + [f"step_f{b}_m{m}.hip" for b in (32, 64) for m in range(3)]


def source_hash():
    """A hash of the step kernels' sources: byte counts from rocprofv3 hold for the sources they
    were counted on."""
    import hashlib
    h = hashlib.sha256()
    csrc = os.path.join(ROOT, "photonoxide_amd", "csrc")
    for name in STEP_SOURCES:
        h.update(name.encode())
        h.update(open(os.path.join(csrc, name), "rb").read())
    return h.hexdigest()[:16]


def counted_kib(d, name):
    """Per kernel variant (template arguments) of the kernels named `name`: the mean over its
    dispatches of each counter rocprofv3 --pmc wrote under directory d, in KiB."""
    import csv
    import glob
    import re
    sums = {}
    for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        for row in csv.DictReader(open(f)):
            k = row["Kernel_Name"]
            if name not in k:
                continue
            m = re.search(r"<([^>]*)>", k)
            variant = m.group(1) if m else k
            key = (variant, row["Counter_Name"])
            sums.setdefault(key, []).append(float(row["Counter_Value"]))
    out = {}
    for (variant, counter), vals in sums.items():
        out.setdefault(variant, {})[counter] = sum(vals) / len(vals)
    return out


def traffic_summary(out, fetch_dir, write_dir, copy_fetch_dir, copy_write_dir, copy_mib, kind, tiling,
                    kernel, commands):
    """The step's HBM bytes from rocprofv3's FETCH_SIZE and WRITE_SIZE (one pass each), each
    scaled by what the counter reads for the library's copy kernel of known size."""
    fetch, write = counted_kib(fetch_dir, kernel), counted_kib(write_dir, kernel)
    cf = counted_kib(copy_fetch_dir, "hbm_copy")
    cw = counted_kib(copy_write_dir, "hbm_copy")
    copy_kib = copy_mib * 1024.0
    f_copy = sum(v["FETCH_SIZE"] for v in cf.values()) / len(cf)
    w_copy = sum(v["WRITE_SIZE"] for v in cw.values()) / len(cw)
    fetch_scale, write_scale = copy_kib / f_copy, copy_kib / w_copy
    f_step = sum(v["FETCH_SIZE"] for v in fetch.values())
    w_step = sum(v["WRITE_SIZE"] for v in write.values())
    read_b, write_b = f_step * fetch_scale * 1024, w_step * write_scale * 1024
    alg, _ = algorithmic_bytes_per_step(384, 256, 192, 12)
    rec = {"kernel": f"{kernel} (the launches of one step)", "workload": "mmi_1x2_soi_384x256x192_dx20nm_cpml12_f32",
           "kernel_kind": kind, "tiling": list(tiling), "source_hash": source_hash(),
           "fetch_kib_per_step": round(f_step, 1), "write_kib_per_step": round(w_step, 1),
           "fetch_scale": round(fetch_scale, 4), "write_scale": round(write_scale, 4),
           "bytes_per_launch": int(read_b + write_b), "read_bytes": int(read_b), "write_bytes": int(write_b),
           "algorithmic_bytes": alg, "traffic_over_algorithmic": round((read_b + write_b) / alg, 4),
           "per_variant_kib": {k: {c: round(x, 1) for c, x in v.items()} for k, v in fetch.items()},
           "per_variant_write_kib": {k: {c: round(x, 1) for c, x in v.items()} for k, v in write.items()},
           "scale_from": f"the library's copy kernel over {copy_mib} MiB (bench.py --copy-mib {copy_mib}): "
                         f"FETCH_SIZE {f_copy:.1f} KiB per dispatch for {copy_kib:.0f} KiB read, WRITE_SIZE "
                         f"{w_copy:.1f} KiB for {copy_kib:.0f} KiB written",
           "commands": commands}
    json.dump(rec, open(out, "w"), indent=1)
    return rec


def hbm_rates(gpu, mib=1024, reps=10):
    """The device's achievable HBM rates from the library's own streaming kernels."""
    r = gpu.hbm_rates(mib << 20, reps)
    return {k: round(v, 1) for k, v in r.items()} | {"buffer_mib": mib, "unit": "GB/s"}


def serving(gpu, n2d, n3d, rate, seed=2, autotune=0, direct=False, shortest_first=True, fill_cells=0, segment=0):
    """The batched-serving workload: a seeded Poisson stream of 2D (U[128,1024]^2) and 3D
    (U[64,256]^3) design jobs, 1000-4000 steps, 5-frequency monitors, f32. Each job first runs
    alone on the plain path (one at a time: the sequential throughput), then the stream goes
    through the server (the batched throughput and each job's latency, arrival to results on
    the host), and every served spectrum is compared with its job's alone run."""
    import threading
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import serving_jobs as SJ
    import photonoxide_amd as P
    from photonoxide_amd.serving import Server
    stream = SJ.stream(n2d, n3d, seed, rate=rate)
    specs = [SJ.job_spec(*n, st, sd) for (_, n, st, sd) in stream]
    cells = [sp["n"][0] * sp["n"][1] * sp["n"][2] * sp["steps"] for sp in specs]
    # one job at a time on the plain path
    alone, t_seq = [], 0.0
    parts = np.zeros(3)  # set-up on the device, stepping, results back
    for sp in specs:
        sim = SJ.product_job(sp)
        t = time.perf_counter()
        sim.set_device(gpu, P.Precision.F32)
        t1 = time.perf_counter()
        sim.run(sp["steps"])
        t2 = time.perf_counter()
        out = SJ.spectra(sim)
        t3 = time.perf_counter()
        parts += (t1 - t, t2 - t1, t3 - t2)
        t_seq += t3 - t
        alone.append(out)
        del sim
    # the stream through the server
    sims = [SJ.product_job(sp) for sp in specs]
    srv = Server(gpu, autotune=autotune, direct=direct, shortest_first=shortest_first, fill_cells=fill_cells,
                 segment=segment)
    ids = [None] * len(specs)
    t0 = time.perf_counter()

    def arrive():
        for q, (at, _, _, _) in enumerate(stream):
            d = at - (time.perf_counter() - t0)
            if d > 0:
                time.sleep(d)
            ids[q] = srv.submit(sims[q], specs[q]["steps"])
    th = threading.Thread(target=arrive)
    th.start()
    th.join()
    lat = []
    for q in range(len(specs)):
        srv.wait(ids[q], timeout=600)
        sub, _, fin = srv.times(ids[q])
        lat.append(fin - sub)
    span = max(srv.times(j)[2] for j in ids) - min(srv.times(j)[0] for j in ids)
    st = srv.stats()
    srv.close()
    same = all(np.array_equal(SJ.spectra(s).view(np.uint8), a.view(np.uint8)) for s, a in zip(sims, alone))
    worst = max(float(np.abs(SJ.spectra(s) - a).max() / max(np.abs(a).max(), 1e-300)) for s, a in zip(sims, alone))
    total = float(sum(cells))
    lat = np.array(lat)
    return {
        "metric": "batched-job serving throughput (Gcell-updates/s) and p50/p99 latency",
        "value": round(total / span / 1e9, 3), "unit": "Gcell-updates/s", "higher_is_better": True,
        "sequential": round(total / t_seq / 1e9, 3),
        "speedup_vs_sequential": round(t_seq / span, 2),
        "p50_s": round(float(np.percentile(lat, 50)), 4), "p99_s": round(float(np.percentile(lat, 99)), 4),
        "jobs": {"2d": n2d, "3d": n3d, "poisson_rate_per_s": rate, "seed": seed},
        "cell_updates": total, "served_span_s": round(span, 3), "sequential_s": round(t_seq, 3),
        "sequential_parts_s": {"to_device": round(parts[0], 3), "stepping": round(parts[1], 3),
                               "results_back": round(parts[2], 3)},
        "identical_to_alone": bool(same), "max_rel_spectral_diff_vs_alone": worst,
        "dtype": "f32", "data": "synthetic",
        "order": "shortest remaining work first" if shortest_first else "arrival order, every job every segment",
        "fill_cells": fill_cells,
        "server": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in st.items()},
    }


def adjoint_workload(gpu, fd_cells=16, steps=6000, checkpoint=64):
    """The adjoint-gradient workload: 256 x 192 x 96 cells of 20 nm, a 500 x 220 nm Si strip
    in SiO2 along x crossing a 2 x 2 um x 220 nm design block (100 x 100 x 11 cells) whose density
    is seeded U(0, 1) then Gaussian-filtered (sigma 2 cells), a pulse at 1.55 um, the objective
    |a|^2 of the guide's fundamental mode at the output port at 1.55 um (the mode monitor's
    amplitude), 6000 steps with a checkpoint every 64. The f32 gradient (forward with checkpoints, then the reverse-time
    adjoint with recomputation) is timed; central finite differences of f64 runs at fd_cells
    random design cells check it, beside the f64 adjoint's gradient (none with fd_cells 0: the
    timed part alone, for a profile)."""
    from scipy.ndimage import gaussian_filter
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from adjoint_problem import ModeDesignProblem
    import photonoxide_amd as P
    width = math.sqrt(2.0 * math.log(2.0)) / (math.pi * (1 / 1.45 - 1 / 1.65))
    kw = dict(n=(256, 192, 96), pml=12, design=((78, 178), (46, 146), (42, 53)), steps=steps,
              checkpoint=checkpoint, core=(0.5, 0.22), pulse=(width, 6.0 * width), window=(30, 20))
    p32 = ModeDesignProblem(gpu, precision=P.Precision.F32, **kw)
    rng = np.random.default_rng(4)
    rho = gaussian_filter(rng.random(p32.shape), 2.0)
    t0 = time.perf_counter()
    sim, adj, st = p32.build(rho)
    t1 = time.perf_counter()
    j32 = adj.forward(st)
    t2 = time.perf_counter()
    g32 = p32.chain(adj.grad_eps(adj.backward()))
    t3 = time.perf_counter()
    info = adj.info()
    cells = 256 * 192 * 96
    j64 = None
    d = 1e-3
    rel32, rel64 = [], []
    picks = []
    if fd_cells:
        p64 = ModeDesignProblem(gpu, precision=P.Precision.F64, **kw)
        sim64, adj64, _ = p64.build(rho)
        j64 = adj64.forward(st)
        g64 = p64.chain(adj64.grad_eps(adj64.backward()))
        del sim64, adj64
        rng2 = np.random.default_rng(5)
        picks = [tuple(int(rng2.integers(0, s)) for s in p64.shape) for _ in range(fd_cells)]
    for c in picks:
        hi, lo = rho.copy(), rho.copy()
        hi[c] += d
        lo[c] -= d
        fd = (p64.objective(hi) - p64.objective(lo)) / (2 * d)
        rel32.append(abs(g32[c] - fd) / abs(fd))
        rel64.append(abs(g64[c] - fd) / abs(fd))
    return {
        "metric": "reverse-time adjoint gradient: time and agreement with finite differences",
        "config": {"workload": "adjoint_256x192x96_dx20nm_design100x100x11_6000steps_K64", "cells": cells,
                   "objective": "|a|^2 of the output port's fundamental mode at 1.55 um"},
        "objective_f32": j32, "objective_f64": j64,
        "forward_s": round(t2 - t1, 3), "backward_s": round(t3 - t2, 3), "setup_s": round(t1 - t0, 3),
        "forward_gcell_per_s": round(cells * st / (t2 - t1) / 1e9, 2),
        "checkpoints": info["checkpoints"], "checkpoint_gb": round(info["checkpoint_bytes"] / 1e9, 2),
        "recomputed_steps": info["recomputed_steps"],
        "fd_cells": fd_cells, "fd_step": d,
        "max_rel_err_f32_adjoint_vs_fd_f64": float(max(rel32)) if rel32 else None,
        "max_rel_err_f64_adjoint_vs_fd_f64": float(max(rel64)) if rel64 else None,
        "dtype": "f32", "data": "synthetic",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serving", action="store_true",
                    help="leave the batched-serving stream (the metric's latency half) out of the default line")
    ap.add_argument("--rows", type=int, default=0)
    ap.add_argument("--planes", type=int, default=0)
    ap.add_argument("--plain", action="store_true",
                    help="the kernel alone: no source, no monitors (for characterization)")
    ap.add_argument("--pml", type=int, default=12)
    ap.add_argument("--pml-axes", default="xyz",
                    help="with --plain: the axes with a CPML, the others periodic (e.g. y: only the "
                         "y-slab units take a slab variant)")
    ap.add_argument("--kernel", type=int, default=-1)
    ap.add_argument("--launch", type=int, default=-1,
                    help="the aligned march's launch mode: 0 three launches, 1 the slab one on a "
                         "second stream, 2 one launch")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="only time the library's streaming kernels (read, write, copy, triad) over "
                         "buffers of this many MiB, --steps times each (the kernels of known size for "
                         "scaling the HBM byte counters, and the achievable rates)")
    ap.add_argument("--serving", choices=["small", "full"], default=None,
                    help="the batched-serving workload instead: small is 64 2D + 4 3D jobs, full "
                         "256 2D + 16 3D")
    ap.add_argument("--rate", type=float, default=200.0, help="serving arrivals per second")
    ap.add_argument("--workload", choices=["mmi", "dispersive"], default="mmi",
                    help="mmi: the 1x2 MMI, 384 x 256 x 192 (the default); dispersive: 512 x 512 x "
                         "256 with Lorentz/Drude media")
    ap.add_argument("--adjoint", action="store_true",
                    help="the adjoint-gradient workload instead (with its finite-difference check)")
    ap.add_argument("--fd-cells", type=int, default=16,
                    help="adjoint: design cells checked by finite differences (0: none)")
    ap.add_argument("--direct", action="store_true",
                    help="serving: launch each segment's kernels directly instead of a hipGraph")
    ap.add_argument("--arrival-order", action="store_true",
                    help="serving: every admitted job steps every segment (no shortest-first order)")
    ap.add_argument("--segment", type=int, default=0, help="serving: steps per segment (0: the server's default, 32)")
    ap.add_argument("--fill", type=int, default=0,
                    help="serving, shortest first: the cells a segment's jobs fill the device with (0: the server's default)")
    ap.add_argument("--traffic", nargs=5, metavar=("OUT", "FETCH_DIR", "WRITE_DIR", "COPY_FETCH_DIR", "COPY_WRITE_DIR"),
                    help="write the step's counted HBM traffic to OUT from four rocprofv3 --pmc runs (FETCH_SIZE "
                         "then WRITE_SIZE of the bench, then of --copy-mib 1024); no GPU needed")
    ap.add_argument("--traffic-commands", nargs="*", default=[], help="the commands those runs were")
    args = ap.parse_args()
    if args.traffic:
        out, fd, wd, cfd, cwd = args.traffic
        kind = args.kernel if args.kernel >= 0 else 3
        tiling = (args.rows or 2, args.planes or 8)
        name = {3: "march64_step", 5: "march2s_step"}.get(kind, "march")
        print(json.dumps(traffic_summary(out, fd, wd, cfd, cwd, 1024, kind, tiling, name, args.traffic_commands)))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: each box has one MI355X, and "
                          "--gpus above 1 is not supported (a grid would shard by z slabs with a "
                          "one-plane halo exchange per half step)"}))
        sys.exit(2)
    if args.copy_mib:
        import photonoxide_amd as P
        print(json.dumps(hbm_rates(P.Gpu(0), args.copy_mib, max(1, args.steps))), flush=True)
        return
    if args.adjoint:
        import photonoxide_amd as P
        print(json.dumps(adjoint_workload(P.Gpu(0), fd_cells=args.fd_cells)), flush=True)
        return
    if args.serving:
        import photonoxide_amd as P
        n2d, n3d = (256, 16) if args.serving == "full" else (64, 4)
        print(json.dumps(serving(P.Gpu(0), n2d, n3d, args.rate, direct=args.direct,
                                 shortest_first=not args.arrival_order, fill_cells=args.fill,
                                 segment=args.segment)), flush=True)
        return
    import photonoxide_amd as P
    from photonoxide_amd import _native as N

    if args.workload == "dispersive":
        nx, ny, nz, pml = 512, 512, 256, 16
        t_setup = time.perf_counter()
        s, g = dispersive_problem(nx, ny, nz, 0.015, pml)
        t_setup = time.perf_counter() - t_setup
    else:
        nx, ny, nz, pml = 384, 256, 192, args.pml
        if args.pml_axes != "xyz" and not (args.plain and set(args.pml_axes) <= set("xyz")):
            ap.error("--pml-axes takes a subset of xyz and needs --plain")
        s, g = mmi_problem(nx, ny, nz, 0.02, pml, plain=args.plain, pml_axes=args.pml_axes)
    gpu = P.Gpu(0)
    s.set_device(gpu, P.Precision.F32)
    if args.kernel >= 0:
        s.set_kernel(args.kernel)
    if args.rows or args.planes:
        s.set_tiling(args.rows, args.planes)
    res = s.resident()
    if args.launch >= 0:
        N.check(N.lib.pxfdtd_resident_set_concurrent(res, args.launch))
    s.run(args.warmup)
    N.check(N.lib.pxfdtd_resident_synchronize(res))
    N.check(N.lib.pxfdtd_resident_set_profiling(res, 1))
    t0 = time.perf_counter()
    s.run(args.steps)
    N.check(N.lib.pxfdtd_resident_synchronize(res))
    elapsed = time.perf_counter() - t0
    kernel_ms = N.lib.pxfdtd_resident_kernel_ms(res)
    cells = nx * ny * nz
    value = cells * args.steps / elapsed / 1e9
    mode = N.lib.pxfdtd_resident_coefficient_mode(res)
    bytes_step, psi = algorithmic_bytes_per_step(nx, ny, nz, pml, {0: 1, 1: 24, 2: 12}[mode])
    media_bytes = ade_bytes_per_step(s) if args.workload == "dispersive" else 0
    bytes_step += media_bytes
    # the two-step march (kernel 5) takes two steps per pass over the same bytes as one step:
    # a launch is a pair, and kernel_ms is its duration per step
    steps_per_launch = 2 if s.kernel() == 5 else 1
    launch_ms = kernel_ms * steps_per_launch
    achieved = bytes_step / (launch_ms * 1e-3) / 1e9
    traffic, traffic_note = None, "no counted profile for this workload"
    tf = os.path.join(ROOT, "profiles", "traffic_march64_step_mmi_f32.json")
    if os.path.exists(tf) and args.workload == "mmi":
        # the counted bytes hold for the sources, workload and tiling they were counted on
        t = json.load(open(tf))
        if t.get("source_hash") != source_hash():
            traffic_note = "counted on other kernel sources: dropped"
        elif not (not args.plain and pml == 12 and s.kernel() == t.get("kernel_kind")
                  and list(s.tiling()) == t.get("tiling")):
            traffic_note = "counted on another kernel or tiling: dropped"
        else:
            traffic = t.get("bytes_per_launch")
            traffic_note = (f"rocprofv3 FETCH_SIZE and WRITE_SIZE of the step's launches, scaled by what they "
                            f"read for the library's copy kernel of known size; sources {t['source_hash']}")
    measured = hbm_rates(gpu)
    out = {
        "metric": "3D FDTD Gcell-updates/s (% HBM roofline); batched-job p50/p99 latency",
        "value": round(value, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": (f"mmi_1x2_soi_384x256x192_dx20nm_cpml{pml}_f32" + ("_plain" if args.plain else "")
                                + ("" if args.pml_axes == "xyz" else f"_cpml_{args.pml_axes}_only")
                                if args.workload == "mmi" else
                                "dispersive_512x512x256_dx15nm_40si_lorentz2_10au_drude_lorentz2_cpml16_f32"),
                   "cells": cells, "kernel": ["lds_tiles", "register_march", "register_march_noprefetch", "aligned_march", "march_2d",
                              "two_step_march"][s.kernel()], "tiling": list(s.tiling()), "coefficients": {0: "uint8 ids", 1: "ca, cb arrays", 2: "cb arrays"}[mode],
                   "monitors": ("2 ports x 6 components x 50x30 values x 21 frequencies, FP64 sums"
                                if args.workload == "mmi" else
                                "a plane of 6 components x 480x224 values x 5 frequencies, FP64 sums"),
                   "psi_values": psi},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic, "traffic_note": traffic_note,
                     "measured_hbm": measured,
                     "frac_of_measured_copy": round(achieved / measured["copy"], 4),
                     "kernel": ["fused_step", "march_step", "march_step", "march64_step", "march2d_step",
                                "march2s_step"][s.kernel()],
                     "kernel_ms": round(launch_ms, 5), "steps_per_launch": steps_per_launch,
                     "bytes_per_launch": bytes_step,
                     "gcell_per_s_kernel": round(cells / (kernel_ms * 1e-3) / 1e9, 3),
                     # the reference's own byte model (two-pass, f32 ca/cb arrays), for comparison
                     "reference_bytes_per_step": s.bytes_per_step(P.Precision.F32)},
    }
    if media_bytes:
        out["roofline"]["media_bytes_per_step"] = media_bytes
        out["config"]["setup_s"] = round(t_setup, 2)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(nx, ny, nz, pml)
        if media_bytes:
            out["cpu_baseline"]["sample"] += " (Yee + CPML only: the restated CPU kernel has no media)"
    if not args.no_serving and args.workload == "mmi" and not args.plain:
        # the metric's second half: the batched-serving stream's throughput against one job at a
        # time, and its p50/p99 latency (the MMI's device memory freed first)
        del s, res
        srv = serving(gpu, 256, 16, args.rate)
        out["serving"] = {k: v for k, v in srv.items() if k not in ("metric", "dtype", "data", "higher_is_better")}
        out["serving"]["workload"] = "poisson_stream_256x2d_128to1024_16x3d_64to256_1000to4000steps_dft5_f32"
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
