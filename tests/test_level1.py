"""Level 1 of the C ABI driven as the reference's Rust seam would drive it (GPU): the caller
builds a pxfdtd_problem from its own Simulation's tables (here the oracle's: dt, kappa/step
factors, the CPML's b and a per axis side and field, ca and cb), then create, upload, set the
injections, probes and transform series, step with the waves and DFT phases computed on the
host in FP64 (as src/fdtd/gpu/resident.rs:797-843 computes them), read the probes and the sums
back and download, with no host-mirror call in between (resident.rs:154-869's sequence). The
device's fields, probes and sums are the oracle's bits, in f64 and f32.

Also the sizes at which a step kernel's 32-bit buffer offsets stop reaching every value: the
default kernel falls back to one that addresses them, and asking for the others is refused.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
from fixtures import busy_spec, index, oracle

pytestmark = pytest.mark.gpu
TAU = 2 * np.pi


def problem_from_oracle(o, grid, bounds, precision, keep):
    """A pxfdtd_problem holding the oracle's tables (resident.rs:543-793 reads the same)."""
    from photonoxide_amd import _native as N
    nx, ny, nz = grid[:3]
    p = N.Problem()
    p.nx, p.ny, p.nz, p.precision = nx, ny, nz, precision
    for a in range(3):
        p.periodic[a] = bounds["periodic"][a]
        p.pml_low[a], p.pml_high[a] = bounds["low"][a], bounds["high"][a]
        p.step[a] = grid[3 + a]
        n = grid[a]
        for which, dst in ((0, p.factors_e), (1, p.factors_h)):
            f = np.ctypeslib.as_array(O.lib.or_sim_factors(o.h, which, a), shape=(n,)).copy()
            keep.append(f)
            dst[a] = f.ctypes.data_as(N.pd)
    p.dt = o.dt()
    info = (C.c_int32 * 5)()
    b, a_, psi = O.pd(), O.pd(), O.pd()
    ln = C.c_int64()
    for q in range(O.lib.or_sim_slab_count(o.h)):
        O.lib.or_sim_slab(o.h, q, info, C.byref(b), C.byref(a_), C.byref(psi), C.byref(ln))
        field, comp, axis, frm, cnt = list(info)
        side = 0 if frm == 0 else 1
        bb = np.ctypeslib.as_array(b, shape=(cnt,)).copy()
        aa = np.ctypeslib.as_array(a_, shape=(cnt,)).copy()
        keep += [bb, aa]
        p.cpml_b[axis][side][field] = bb.ctypes.data_as(N.pd)
        p.cpml_a[axis][side][field] = aa.ctypes.data_as(N.pd)
    ca = np.concatenate([o.coefficients(c)[0] for c in range(3)])
    cb = np.concatenate([o.coefficients(c)[1] for c in range(3)])
    keep += [ca, cb]
    p.ca, p.cb = ca.ctypes.data_as(N.pd), cb.ctypes.data_as(N.pd)
    p.nmedia = 0
    return p


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_level1_sequence_is_the_oracle_bit_for_bit(prec):
    from photonoxide_amd import _native as N
    print(f"case level 1 busy problem {prec}", flush=True)
    spec = busy_spec(False)
    spec = dict(spec, sources=[], currents=[], probes=[], dft=None)
    grid, bounds = spec["grid"], spec["bounds"]
    nx, ny, nz = grid[:3]
    n = nx * ny * nz
    o = oracle(spec)
    # the caller's sources: a point source on E_z, a current on H_y at two values
    wave_e = dict(kind=0, frequency=1.0, width=0.3, delay=1.0)
    wave_h = dict(kind=0, frequency=1.3, width=0.25, delay=0.9)
    at_e = (nx // 2, ny // 2, nz // 2)
    h_vals = [(index(grid, (nx // 3, ny // 2, nz // 2)), 0.5 - 0.25j), (index(grid, (nx // 3 + 1, ny // 2, nz // 2)), -0.3 + 0.7j)]
    we, wh = O.wave(**wave_e), O.wave(**wave_h)
    o.add_source(0, 2, at_e, we)
    o.add_current(1, [1, 1], [r for r, _ in h_vals], [a for _, a in h_vals], wh)
    probe_at = [(0, 2, (3, 3, 2)), (1, 1, (nx - 8, ny // 3, nz // 3))]
    for f, c, at in probe_at:
        o.add_probe(f, c, at)
    lo, hi = (nx // 2 - 6, ny // 2 - 4, 0), (nx // 2 + 6, ny // 2 + 4, nz - 1)
    freqs = [0.8, 1.0, 1.3]
    o.add_dft(lo, hi, freqs)
    precision = N.F64 if prec == "f64" else N.F32
    keep = []
    p = problem_from_oracle(o, grid, bounds, precision, keep)
    # the caller's fields as they stand (its set-up zeroed E at the conductor)
    f0 = o.fields()
    e0, h0 = np.ascontiguousarray(f0[:3].ravel()), np.ascontiguousarray(f0[3:].ravel())
    if prec == "f32":
        o.use_f32()

    dev, res = C.c_void_p(), C.c_void_p()
    N.check(N.lib.pxfdtd_device_open(0, C.byref(dev)))
    try:
        N.check(N.lib.pxfdtd_resident_create(dev, C.byref(p), C.byref(res)))
        N.check(N.lib.pxfdtd_resident_upload(res, N.dptr(e0), N.dptr(h0), None))
        inj = (N.Injection * 3)()
        inj[0] = N.Injection(0, 2, index(grid, at_e), 1.0, 0.0, 0)
        for q, (r, a) in enumerate(h_vals):
            inj[1 + q] = N.Injection(1, 1, r, a.real, a.imag, 1)
        N.check(N.lib.pxfdtd_resident_set_injections(res, inj, 3, 2))
        pr = (N.Probe * 2)(*[N.Probe(f, c, index(grid, at)) for f, c, at in probe_at])
        N.check(N.lib.pxfdtd_resident_set_probes(res, pr, 2))
        ser = (N.Series * 6)()
        for q in range(6):
            ser[q] = N.Series(q // 3, q % 3, (C.c_int32 * 3)(*lo), (C.c_int32 * 3)(*[h + 1 for h in hi]), 0)
        nf = (C.c_int32 * 1)(len(freqs))
        N.check(N.lib.pxfdtd_resident_set_series(res, ser, 6, nf, 1))
        slots = N.lib.pxfdtd_resident_phase_slots(res)
        assert slots == 2 * len(freqs)
        steps = 90
        dt = o.dt()
        waves = np.zeros((steps, 2, 2))
        phases = np.zeros((steps, slots, 2))
        s_, c_ = C.c_double(), C.c_double()
        for q in range(steps):
            t = q * dt
            waves[q, 0, 0] = O.lib.or_waveform_at(C.byref(we), t + 0.5 * dt)  # J at t + dt/2
            w = O.lib.or_waveform_complex_at(C.byref(wh), t)                   # M at t
            waves[q, 1] = (w.re, w.im)
            te, th = (q + 1) * dt, (q + 1) * dt - 0.5 * dt
            for f2, tt in enumerate((te, th)):
                for m, fr in enumerate(freqs):
                    O.lib.or_sincos(TAU * fr * tt, C.byref(s_), C.byref(c_))
                    phases[q, f2 * len(freqs) + m] = (c_.value * dt, s_.value * dt)
        # two calls, as a run split in two
        for a, b in ((0, 41), (41, steps)):
            wv = np.ascontiguousarray(waves[a:b])
            ph = np.ascontiguousarray(phases[a:b])
            N.check(N.lib.pxfdtd_resident_step(res, b - a, N.dptr(wv), N.dptr(ph)))
        o.run(steps)
        rec = np.zeros(steps * 2)
        got = C.c_int64()
        N.check(N.lib.pxfdtd_resident_read_probes(res, N.dptr(rec), rec.size, C.byref(got)))
        assert got.value == steps
        rec = rec.reshape(steps, 2)
        for q in range(2):
            assert np.array_equal(rec[:, q].view(np.uint64), o.probe(q).view(np.uint64)), q
        sums = np.zeros(N.lib.pxfdtd_resident_sums_len(res))
        N.check(N.lib.pxfdtd_resident_read_sums(res, N.dptr(sums), sums.size))
        want = np.concatenate([np.stack([o.dft(0, q // 3, q % 3).real, o.dft(0, q // 3, q % 3).imag], -1).ravel()
                               for q in range(6)])
        assert np.array_equal(sums.view(np.uint64), want.view(np.uint64))
        e, h = np.zeros(3 * n), np.zeros(3 * n)
        N.check(N.lib.pxfdtd_resident_download(res, N.dptr(e), N.dptr(h), None))
        fo = o.fields()
        assert np.array_equal(np.concatenate([e, h]).view(np.uint64), fo.ravel().view(np.uint64)), \
            np.abs(np.concatenate([e, h]) - fo.ravel()).max()
    finally:
        if res:
            N.lib.pxfdtd_resident_destroy(res)
        N.lib.pxfdtd_device_close(dev)


def big_problem(nx, ny, nz, precision, pml, keep):
    """A vacuum grid with a CPML of `pml` cells on every face, as a caller would table it."""
    from photonoxide_amd import _native as N
    p = N.Problem()
    p.nx, p.ny, p.nz, p.precision = nx, ny, nz, precision
    h, dt = 0.05, 0.05 * 0.9 / np.sqrt(3)
    for a, n in enumerate((nx, ny, nz)):
        p.step[a] = h
        p.pml_low[a] = p.pml_high[a] = pml
        f = np.full(n, 1.0 / h)
        keep.append(f)
        p.factors_e[a] = p.factors_h[a] = f.ctypes.data_as(N.pd)
        for side in range(2):
            for field in range(2):
                b, a_ = np.full(pml, 0.9), np.full(pml, -0.05)
                keep += [b, a_]
                p.cpml_b[a][side][field] = b.ctypes.data_as(N.pd)
                p.cpml_a[a][side][field] = a_.ctypes.data_as(N.pd)
    N_ = nx * ny * nz
    ca, cb = np.ones(3 * N_), np.full(3 * N_, dt)
    keep += [ca, cb]
    p.ca, p.cb = ca.ctypes.data_as(N.pd), cb.ctypes.data_as(N.pd)
    p.dt = dt
    return p


def test_kernels_beyond_their_32_bit_offsets_are_refused():
    """f32 with 3 x values x 4 B beyond 2 GiB (640 x 640 x 448 cells: 183.5 M values a
    component): the aligned march cannot address a field's three components, so the default is
    the 62-wide march, and asking for the aligned one is refused; the 62-wide march still
    addresses each component (733 MB)."""
    from photonoxide_amd import _native as N
    print("case 640x640x448 f32 kernel choice", flush=True)
    keep = []
    p = big_problem(640, 640, 448, N.F32, 4, keep)
    dev, res = C.c_void_p(), C.c_void_p()
    N.check(N.lib.pxfdtd_device_open(0, C.byref(dev)))
    try:
        N.check(N.lib.pxfdtd_resident_create(dev, C.byref(p), C.byref(res)))
        assert N.lib.pxfdtd_resident_get_kernel(res) == 1
        assert N.lib.pxfdtd_resident_set_kernel(res, 3) == 2
        assert N.lib.pxfdtd_resident_get_kernel(res) == 1
        assert N.lib.pxfdtd_resident_set_kernel(res, 0) == 0
        assert N.lib.pxfdtd_resident_set_kernel(res, 2) == 0
    finally:
        if res:
            N.lib.pxfdtd_resident_destroy(res)
        N.lib.pxfdtd_device_close(dev)
