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


def test_cancel_a_queued_and_a_running_job(gpu):
    """One job at a time (max_jobs 1), admitted in arrival order: a long job is cancelled while
    it runs, and stops at a segment boundary with the fields of that many steps alone; the job
    queued behind it is cancelled before admission and never steps; a third job still finishes
    with its own bits."""
    from photonoxide_amd import _native as N
    from photonoxide_amd.serving import CANCELLED, DONE, RUNNING, Server
    print("case cancel: a running 512x512 job of 20000 steps, a queued one, a finishing one", flush=True)
    long_spec = job_spec(512, 512, 1, 20000, 51)
    queued_spec = job_spec(100, 80, 1, 50, 52)
    last_spec = job_spec(90, 70, 1, 120, 53)
    K = 16
    with Server(gpu, segment=K, max_jobs=1, shortest_first=False) as srv:
        s_long, s_q, s_last = product_job(long_spec), product_job(queued_spec), product_job(last_spec)
        j_long = srv.submit(s_long, long_spec["steps"])
        j_q = srv.submit(s_q, queued_spec["steps"])
        j_last = srv.submit(s_last, last_spec["steps"])
        t0 = time.perf_counter()
        while srv.state(j_long) != RUNNING and time.perf_counter() - t0 < 60:
            time.sleep(0.002)
        time.sleep(0.05)
        srv.cancel(j_q)
        assert srv.state(j_q) == CANCELLED
        srv.cancel(j_long)
        with pytest.raises(N.NativeError, match="cancelled"):
            srv.wait(j_long, timeout=120)
        srv.wait(j_last, timeout=120)
        assert srv.state(j_long) == CANCELLED and srv.state(j_last) == DONE
        srv.cancel(j_last)  # done: left as it is
        assert srv.state(j_last) == DONE
        st = srv.stats()
    reached = s_long.steps()
    print(f"the long job stopped after {reached} of {long_spec['steps']} steps; {st}", flush=True)
    assert 0 < reached < long_spec["steps"] and reached % K == 0
    assert s_q.steps() == 0
    assert same_bits(fields_of(s_long), fields_of(alone(gpu, dict(long_spec, steps=reached))))
    assert same_bits(spectra(s_last), spectra(alone(gpu, last_spec)))


def test_shortest_first_finishes_short_jobs_first_with_the_same_bits(gpu):
    """Shortest remaining work first with a fill of one cell: each segment steps the one job
    with the least work left, so the jobs finish in order of their work (the long one, first
    submitted, last), each with the bits of its alone run."""
    from photonoxide_amd.serving import Server
    print("case shortest first, fill 1", flush=True)
    specs = [job_spec(300, 200, 1, 400, 61), job_spec(80, 60, 1, 100, 62), job_spec(120, 90, 1, 60, 63),
             job_spec(48, 40, 32, 40, 64)]
    ref = [spectra(alone(gpu, sp)) for sp in specs]
    with Server(gpu, segment=16, shortest_first=True, fill_cells=1) as srv:
        sims = [product_job(sp) for sp in specs]
        jobs = [srv.submit(s, sp["steps"]) for s, sp in zip(sims, specs)]
        for j in jobs:
            srv.wait(j, timeout=120)
        ends = [srv.times(j)[2] for j in jobs]
    work = [sp["n"][0] * sp["n"][1] * sp["n"][2] * sp["steps"] for sp in specs]
    print("work", work, "ends", ends, flush=True)
    assert np.argmax(ends) == 0
    order = np.argsort(ends)
    assert all(work[order[q]] <= work[order[q + 1]] for q in range(len(specs) - 1)), (work, ends)
    for q, s in enumerate(sims):
        assert same_bits(spectra(s), ref[q]), q
