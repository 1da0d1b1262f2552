"""The reverse-time adjoint (pxfdtd_adjoint_*) and its torch.autograd.Function.

The objective is J = |S|^2, S = sum over a transform monitor's values of m(r) times the FP64
transform of E's component c at one frequency (a mode overlap when m is a mode's profile). Its
gradient with respect to the permittivity of chosen E values comes from the exact transpose of
the device's time stepping run backwards, with the state recomputed from checkpoints taken
every K steps in HBM. photonoxide's adjoint is frequency-domain from two forward runs
(src/fdtd/adjoint.rs:496-569, gradient :831-883); this one differentiates the discrete time
stepping itself, so it agrees with finite differences of the same discrete run.
"""
import ctypes as C
from typing import Callable, Sequence

import numpy as np

from . import _native as N
from .fdtd import Precision, Simulation


class Adjoint:
    """The adjoint of `sim` (which has a device and has not run).

    monitor, component, frequency: the transform monitor, E's component (-1: all three) and the
    frequency index the objective weighs; weights: m over the monitor's box, complex [k][j][i]
    ([c][k][j][i] for all three, as Simulation.mode_weights gives a mode's); design: (component,
    flat index (k ny + j) nx + i) of each design value; checkpoint: steps between checkpoints.
    """

    def __init__(self, sim: Simulation, monitor: int, component: int, frequency: int, weights,
                 design: Sequence, checkpoint: int = 64):
        self.sim = sim
        w = np.ascontiguousarray(np.stack([np.real(weights), np.imag(weights)], axis=-1).ravel(),
                                 dtype=np.float64)
        comps = np.ascontiguousarray([d[0] for d in design], dtype=np.int32)
        index = np.ascontiguousarray([d[1] for d in design], dtype=np.int64)
        self._keep = (w, comps, index)
        spec = N.AdjointSpec(int(monitor), int(component), int(frequency), w.ctypes.data_as(N.pd),
                             len(index), comps.ctypes.data_as(N.pi32), index.ctypes.data_as(N.pi64),
                             int(checkpoint))
        h = C.c_void_p()
        N.check(N.lib.pxfdtd_adjoint_create(sim._h, C.byref(spec), C.byref(h)))
        self._h = h
        self.design = list(design)

    def forward(self, steps: int) -> float:
        j = C.c_double()
        N.check(N.lib.pxfdtd_adjoint_forward(self._h, int(steps), C.byref(j)))
        return j.value

    def backward(self) -> np.ndarray:
        """dJ/dcb per design value."""
        g = np.zeros(len(self.design))
        N.check(N.lib.pxfdtd_adjoint_backward(self._h, N.dptr(g)))
        return g

    def grad_eps(self, g_cb: np.ndarray) -> np.ndarray:
        """dJ/deps per design value from dJ/dcb: a lossless value has cb = dt / eps."""
        dt = self.sim.dt()
        cbs = [self.sim.coefficients(c)[1] for c in range(3)]
        cb = np.array([cbs[c][i] for c, i in self.design])
        return -g_cb * cb * cb / dt

    def info(self) -> dict:
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        N.check(N.lib.pxfdtd_adjoint_info(self._h, C.byref(a), C.byref(b), C.byref(c)))
        return {"checkpoints": a.value, "checkpoint_bytes": b.value, "recomputed_steps": c.value}

    def close(self):
        h = getattr(self, "_h", None)
        if h:
            N.lib.pxfdtd_adjoint_destroy(h)
            self._h = None

    def __del__(self):
        self.close()


def _objective_function():
    import torch

    class DesignObjective(torch.autograd.Function):
        """J(rho) for a design problem: forward runs the simulation built from rho with
        checkpoints; backward runs the adjoint and returns dJ/drho."""

        @staticmethod
        def forward(ctx, rho, problem):
            r = rho.detach().cpu().numpy().astype(np.float64)
            sim, adj, steps = problem.build(r)
            j = adj.forward(steps)
            ctx.adj = adj
            ctx.problem = problem
            ctx.sim = sim
            ctx.rho_shape = tuple(rho.shape)
            ctx.rho_device = rho.device
            return torch.tensor(j, dtype=torch.float64, device=rho.device)

        @staticmethod
        def backward(ctx, grad_out):
            g_eps = ctx.adj.grad_eps(ctx.adj.backward())
            g_rho = ctx.problem.chain(g_eps).reshape(ctx.rho_shape)
            g = torch.from_numpy(g_rho).to(ctx.rho_device)
            return grad_out * g, None

    return DesignObjective


def objective(rho, problem):
    """J(rho) as a differentiable torch scalar; `problem` has build(rho) -> (sim, adjoint, steps)
    and chain(dJ/deps per design value) -> dJ/drho."""
    return _objective_function().apply(rho, problem)
