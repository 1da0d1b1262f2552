"""The serving loop (GPU): continuous batching gives each job the same bits as running it alone.

- Jobs of mixed 2D and 3D sizes and step counts, submitted while others run, finish with the
  fields and FP64 transform sums of the same job stepped alone on the plain path.
- A replayed hipGraph and direct launches give the same bits.
- Admission keeps the admitted jobs' footprint within the HBM budget, and every job still runs.
- The served spectra are within 1e-4 relative of the FP64 oracle's.
- The autotuner's choice changes no bit.
"""
import time

import numpy as np
import pytest

from serving_jobs import job_spec, oracle_job, product_job, spectra

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import photonoxide_amd as P
    return P.Gpu(0)


MIX = [((70, 45, 1), 160, 11), ((200, 130, 1), 230, 12), ((130, 260, 1), 97, 13),
       ((40, 36, 30), 140, 14), ((64, 64, 64), 75, 15), ((333, 77, 1), 301, 16),
       ((96, 40, 28), 64, 17)]


def alone(gpu, spec):
    import photonoxide_amd as P
    s = product_job(spec)
    s.set_device(gpu, P.Precision.F32)
    s.run(spec["steps"])
    return s


def fields_of(s):
    return np.stack([s.e(c) for c in range(3)] + [s.h(c) for c in range(3)])


def same_bits(a, b):
    return np.array_equal(np.asarray(a).view(np.uint8), np.asarray(b).view(np.uint8))


def test_each_job_is_the_same_alone_and_in_the_mix(gpu):
    from photonoxide_amd.serving import Server
    specs = [job_spec(*n, st, seed) for n, st, seed in MIX]
    ref = []
    for sp in specs:
        print(f"alone {sp['n']} {sp['steps']} steps", flush=True)
        s = alone(gpu, sp)
        ref.append((spectra(s), fields_of(s)))
    with Server(gpu, segment=32) as srv:
        sims = [product_job(sp) for sp in specs]
        jobs = [srv.submit(sims[q], specs[q]["steps"]) for q in range(3)]
        # the rest arrive while the first ones are on the device
        time.sleep(0.05)
        jobs += [srv.submit(sims[q], specs[q]["steps"]) for q in range(3, len(specs))]
        for q, j in enumerate(jobs):
            srv.wait(j, timeout=120)
            print(f"served {specs[q]['n']}: times {srv.times(j)}", flush=True)
        st = srv.stats()
    print(st)
    assert st["completed"] == len(specs) and st["failed"] == 0
    assert st["captures"] >= 1
    for q, s in enumerate(sims):
        assert s.steps() == specs[q]["steps"]
        sp, fl = spectra(s), fields_of(s)
        assert same_bits(sp, ref[q][0]), (q, np.abs(sp - ref[q][0]).max())
        assert same_bits(fl, ref[q][1]), (q, np.abs(fl - ref[q][1]).max())


def test_graph_and_direct_launches_agree(gpu):
    from photonoxide_amd.serving import Server
    specs = [job_spec(*n, st, seed) for n, st, seed in MIX[:4]]
    out = []
    for direct in (False, True):
        with Server(gpu, segment=16, direct=direct) as srv:
            sims = [product_job(sp) for sp in specs]
            jobs = [srv.submit(s, sp["steps"]) for s, sp in zip(sims, specs)]
            for j in jobs:
                srv.wait(j, timeout=120)
            out.append([spectra(s) for s in sims])
            st = srv.stats()
            assert (st["captures"] > 0) != direct
    for a, b in zip(*out):
        assert same_bits(a, b)


def test_admission_keeps_to_the_budget(gpu):
    from photonoxide_amd import _native as N
    from photonoxide_amd.serving import Server
    specs = [job_spec(160, 120, 1, 120, 30 + q) for q in range(5)]
    sims = [product_job(sp) for sp in specs]
    foot = sims[0].footprint()
    assert foot > 0
    # a budget that holds two of these jobs
    with Server(gpu, hbm_budget=int(2.5 * foot), segment=8) as srv:
        jobs = [srv.submit(s, sp["steps"]) for s, sp in zip(sims, specs)]
        for j in jobs:
            srv.wait(j, timeout=120)
        st = srv.stats()
    assert st["completed"] == 5 and st["hbm_in_use"] == 0
    assert st["peak_jobs"] == 2, st
    with Server(gpu, hbm_budget=int(0.5 * foot)) as srv:
        with pytest.raises(N.NativeError, match="budget"):
            srv.submit(product_job(specs[0]), 10)


def test_served_spectra_within_1e4_of_the_fp64_oracle(gpu):
    from photonoxide_amd.serving import Server
    sp = job_spec(120, 90, 1, 700, 21)
    o = oracle_job(sp)
    o.run(sp["steps"])
    ref = np.stack([o.dft(0, 0, 1), o.dft(0, 1, 2)])
    with Server(gpu) as srv:
        s = product_job(sp)
        srv.wait(srv.submit(s, sp["steps"]), timeout=120)
    got = spectra(s).reshape(ref.shape)
    rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print(f"spectral error vs FP64 oracle: {rel:.3e}")
    assert rel <= 1e-4, rel


def test_autotuned_tiling_changes_no_bit(gpu):
    import photonoxide_amd as P
    from photonoxide_amd.serving import Server
    sp = job_spec(150, 110, 1, 90, 41)
    base = alone(gpu, sp)
    s = product_job(sp)
    s.set_device(gpu, P.Precision.F32)
    r = s.resident()
    import ctypes as C
    from photonoxide_amd import _native as N
    rows, planes, ms = C.c_int32(), C.c_int32(), C.c_double()
    N.check(N.lib.pxfdtd_resident_autotune(r, 3, 4, C.byref(rows), C.byref(planes), C.byref(ms)))
    assert rows.value in (1, 2, 4) and planes.value >= 1 and ms.value > 0
    assert s.tiling() == (rows.value, planes.value)
    s.run(sp["steps"])
    assert same_bits(fields_of(s), fields_of(base))
    with Server(gpu, autotune=2) as srv:
        t = product_job(sp)
        srv.wait(srv.submit(t, sp["steps"]), timeout=60)
    assert same_bits(spectra(t), spectra(base))
