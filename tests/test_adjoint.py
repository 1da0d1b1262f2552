"""The reverse-time adjoint (GPU): the discrete transpose of the device's time stepping.

- Its gradient (through torch.autograd) agrees with central finite differences of the same
  discrete run, in FP64, within 1e-3 relative (the north star's bar) at random design cells.
- The checkpoint interval changes no bit of the gradient (recomputation is the forward's bits).
- The objective it reports is the plain run's.
- The f32 gradient stays close to the f64 one.
"""
import numpy as np
import pytest

from adjoint_problem import DesignProblem

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import photonoxide_amd as P
    return P.Gpu(0)


def grad_of(problem, rho):
    import torch
    from photonoxide_amd.adjoint import objective
    r = torch.tensor(rho, dtype=torch.float64, requires_grad=True)
    j = objective(r, problem)
    j.backward()
    return float(j.item()), r.grad.numpy().copy()


def test_gradient_matches_central_differences(gpu):
    p = DesignProblem(gpu)
    rng = np.random.default_rng(4)
    rho = rng.random(p.shape)
    j, g = grad_of(p, rho)
    print(f"J = {j:.6e}, |grad| max {np.abs(g).max():.3e}", flush=True)
    assert j > 0 and np.all(np.isfinite(g))
    d = 1e-3
    cells = [tuple(rng.integers(0, s) for s in p.shape) for _ in range(4)]
    worst = 0.0
    for c in cells:
        hi, lo = rho.copy(), rho.copy()
        hi[c] += d
        lo[c] -= d
        fd = (p.objective(hi) - p.objective(lo)) / (2 * d)
        rel = abs(fd - g[c]) / max(abs(fd), 1e-30)
        worst = max(worst, rel)
        print(f"cell {c}: adjoint {g[c]:.9e} fd {fd:.9e} rel {rel:.2e}", flush=True)
    assert worst < 1e-3, worst


def test_checkpoint_interval_changes_no_bit(gpu):
    rng = np.random.default_rng(5)
    out = []
    for k in (1, 37, 400):
        p = DesignProblem(gpu, checkpoint=k, steps=400)
        rho = rng.random(p.shape) if not out else rho0
        rho0 = rho
        sim, adj, steps = p.build(rho)
        j = adj.forward(steps)
        g = adj.backward()
        info = adj.info()
        assert info["checkpoints"] == -(-steps // k)
        assert info["recomputed_steps"] == steps
        out.append((j, g))
    for j, g in out[1:]:
        assert j == out[0][0]
        assert np.array_equal(g.view(np.uint64), out[0][1].view(np.uint64))


def test_objective_is_the_plain_runs(gpu):
    p = DesignProblem(gpu)
    rho = np.random.default_rng(6).random(p.shape)
    sim, adj, steps = p.build(rho)
    j = adj.forward(steps)
    # the same transform sums; the overlap is summed in another order here
    assert j == pytest.approx(p.objective(rho), rel=1e-12)
    # the simulation is left as the forward run left it, after the backward sweep too
    e_before = sim.e(1)
    adj.backward()
    assert np.array_equal(sim.e(1), e_before)
    assert sim.steps() == steps


def test_f32_gradient_close_to_f64(gpu):
    import photonoxide_amd as P
    rho = np.random.default_rng(7).random(DesignProblem(gpu).shape)
    _, g64 = grad_of(DesignProblem(gpu), rho)
    _, g32 = grad_of(DesignProblem(gpu, precision=P.Precision.F32), rho)
    rel = np.linalg.norm(g32 - g64) / np.linalg.norm(g64)
    print(f"f32 vs f64 gradient: relative L2 {rel:.3e}")
    assert rel < 1e-2, rel
