"""FDTD jobs run end to end on the device (GPU): the job runner's events (started, solver,
fdtd_progress with the cell-updates per second, fdtd_spectrum as the transforms accumulate and
at the end, fdtd_resonances, finished) written to events.jsonl, and what the mathematics fixes
of the results:
- a straight guide carries its mode's power through: transmission 1 within the CPMLs' and the
  port modes' error, the flux through a plane the same, no reflection to speak of;
- a 1x2 MMI symmetric about its axis splits the power evenly, and no more comes out than went in;
- a disk's resonances, found by harmonic inversion of a point's field, lie in the band, decay
  and are held by the signal (error below 1e-2).
"""
import copy
import json
import os

import numpy as np
import pytest

from jobs import DISK, MMI, STRAIGHT

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import photonoxide_amd as P
    return P.Gpu(0)


def last_spectrum(events, name):
    found = [e for e in events if e["type"] == "fdtd_spectrum" and e["monitor"] == name and e["last"]]
    assert len(found) == 1, name
    return found[0]


def test_a_straight_guide_job_records_its_events_and_carries_its_power(gpu, tmp_path):
    from photonoxide_amd.job import run_job
    print("case job straight 160x112 2D TE", flush=True)
    events = run_job(copy.deepcopy(STRAIGHT), str(tmp_path), gpu)
    written = [json.loads(line) for line in open(os.path.join(tmp_path, "events.jsonl"))]
    assert written == json.loads(json.dumps(events))
    types = [e["type"] for e in events]
    assert types[0] == "started" and types[1] == "solver" and types[-1] == "finished"
    assert events[-1]["stopped"] is None
    progress = [e for e in events if e["type"] == "fdtd_progress"]
    assert len(progress) >= 3
    last = progress[-1]
    assert last["decay"] <= 1e-6 and last["fraction"] == 1e-6 and last["time_um"] < last["until_um"]
    assert last["cell_updates_per_second"] > 1e8
    assert all(a["step"] < b["step"] for a, b in zip(progress, progress[1:]))
    out = last_spectrum(events, "output")
    t = np.array(out["series"][0]["values"])
    back = np.array(out["series"][1]["values"])
    through = np.array(last_spectrum(events, "through")["series"][0]["values"])
    refl = np.array(last_spectrum(events, "source 1")["series"][0]["values"])
    print("transmission", t, "flux", through, "reflection", refl, flush=True)
    assert np.all(np.abs(t - 1.0) < 1e-5), t
    assert np.all(np.abs(through - t) < 2e-4), (through, t)
    assert np.all(back < 1e-5) and np.all(refl < 1e-8)
    assert np.allclose(out["wavelengths_um"], [1.5 + 0.1 * k / 5 for k in range(6)], rtol=1e-15, atol=0)


def test_a_symmetric_mmi_splits_evenly(gpu):
    from photonoxide_amd.job import run_job
    print("case job mmi 384x200 2D TE", flush=True)
    events = run_job(copy.deepcopy(MMI), None, gpu)
    up = np.array(last_spectrum(events, "upper")["series"][0]["values"])
    lo = np.array(last_spectrum(events, "lower")["series"][0]["values"])
    refl = np.array(last_spectrum(events, "source 1")["series"][0]["values"])
    print("upper", up, "lower", lo, "reflection", refl, flush=True)
    assert np.all(np.abs(up - lo) <= 1e-6 * np.abs(up)), (up, lo)
    assert np.all(up + lo + refl <= 1.0 + 1e-2)
    assert np.all(up > 0.1)


def test_a_disks_resonances_are_found(gpu):
    from photonoxide_amd.job import run_job
    print("case job disk 128x128 2D TM, dipole, resonance monitor", flush=True)
    events = run_job(copy.deepcopy(DISK), None, gpu)
    res = [e for e in events if e["type"] == "fdtd_resonances"]
    assert len(res) == 1 and res[0]["component"] == "Ez" and res[0]["position_um"] == [0.6, -0.2]
    found = res[0]["resonances"]
    print(found, flush=True)
    assert found
    for r in found:
        assert 1.45 <= r["wavelength_um"] <= 1.65 and r["decay"] > 0 and r["error"] < 1e-2 and r["q"] > 1
