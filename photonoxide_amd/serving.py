"""The serving loop: continuous batching of FDTD jobs on one MI355X (pxfdtd_server_*).

A job is a :class:`~photonoxide_amd.fdtd.Simulation` built with no device, handed over with a
step count. The native loop thread admits jobs while the admitted jobs' device footprint fits
the HBM budget, steps the admitted jobs together in segments (batched launches replayed from a
captured hipGraph; by default those with the least work left, enough of them to fill the
device), admits new arrivals at each segment boundary, and hands each finished job's fields
and FP64 transform sums back to its simulation. A job's
results are the same bits as running it alone. photonoxide has no counterpart; its nearest is
the process-level farm (src/farm.rs:1-56).
"""
import ctypes as C
from typing import Dict, Optional

from . import _native as N
from .fdtd import Gpu, Simulation

QUEUED, RUNNING, DONE, FAILED, CANCELLED = (N.JOB_QUEUED, N.JOB_RUNNING, N.JOB_DONE,
                                             N.JOB_FAILED, N.JOB_CANCELLED)


class Server:
    """A serving loop on `gpu`.

    hbm_budget: device bytes the admitted jobs may hold (0: 90 % of the free HBM);
    max_jobs: jobs admitted at once (0: 256); segment: steps per batched segment (0: 32);
    direct: launch each segment's kernels directly instead of replaying a hipGraph;
    autotune: trial launches per tiling for the first 3D job of each class of grid shape (each
    extent's power of two; 0: two rows per strip and 8-plane chunks, which suit a full batch);
    2D jobs take one row and chunks of up to 16 rows either way;
    shortest_first: each segment steps the admitted jobs with the least work left until their
    cells reach fill_cells (0: 2^26), and the queue admits its smallest job first (False: every
    admitted job steps every segment, admitted in arrival order).
    """

    def __init__(self, gpu: Gpu, hbm_budget: int = 0, max_jobs: int = 0, segment: int = 0,
                 direct: bool = False, autotune: int = 0, shortest_first: bool = True, fill_cells: int = 0):
        self._gpu = gpu
        cfg = N.ServerConfig(int(hbm_budget), int(max_jobs), int(segment), 1 if direct else 0,
                             int(autotune), 1 if shortest_first else 0, 0, int(fill_cells))
        h = C.c_void_p()
        N.check(N.lib.pxfdtd_server_create(gpu.handle, C.byref(cfg), C.byref(h)))
        self._h = h
        self._sims: Dict[int, Simulation] = {}

    def submit(self, sim: Simulation, steps: int) -> int:
        """Queues `sim` (which must have no device) for `steps` steps; the job's number."""
        job = C.c_int64()
        N.check(N.lib.pxfdtd_server_submit(self._h, sim._h, int(steps), C.byref(job)))
        self._sims[job.value] = sim  # kept alive while the server may step it
        return job.value

    def state(self, job: int) -> int:
        st = C.c_int32()
        N.check(N.lib.pxfdtd_server_state(self._h, int(job), C.byref(st)))
        return st.value

    def wait(self, job: int, timeout: float = 0.0) -> Simulation:
        """Blocks until the job is done (raises if it failed, was cancelled or timed out)."""
        N.check(N.lib.pxfdtd_server_wait(self._h, int(job), float(timeout)))
        return self._sims[job]

    def cancel(self, job: int):
        """Cancels a queued or running job; a running one stops at the next segment boundary."""
        N.check(N.lib.pxfdtd_server_cancel(self._h, int(job)))

    def times(self, job: int):
        """(submitted, admitted, finished), seconds since the server started."""
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        N.check(N.lib.pxfdtd_server_times(self._h, int(job), C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def stats(self) -> dict:
        s = N.ServerStats()
        N.check(N.lib.pxfdtd_server_get_stats(self._h, C.byref(s)))
        return {name: getattr(s, name) for name, _ in N.ServerStats._fields_}

    def close(self):
        h = getattr(self, "_h", None)
        if h:
            N.lib.pxfdtd_server_destroy(h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        self.close()
