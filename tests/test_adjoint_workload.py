"""The adjoint workload at full size (GPU): 256 x 192 x 96 cells of 20 nm, a 500 x 220 nm Si strip
in SiO2 crossing a 2 x 2 um x 220 nm design block (100 x 100 x 11 cells) whose density is seeded
U(0, 1) then Gaussian-filtered (sigma 2 cells), eps = eps_SiO2 + rho (eps_Si - eps_SiO2), a pulse at
1.55 um, 6000 steps, a checkpoint every 64 steps, and the objective |a|^2 of the guide's
fundamental mode at the output port at 1.55 um (the mode monitor's amplitude).

The f32 reverse-time adjoint's gradient (f32 fields, FP64 transform sums) agrees with central
finite differences of FP64 runs within 1e-3 relative at 4 random design cells (the north star's
bar), and the adjoint's J is the plain f32 run's mode monitor |a|^2.
"""
import numpy as np
import pytest

from adjoint_problem import ModeDesignProblem

pytestmark = pytest.mark.gpu


def test_full_size_adjoint_matches_central_differences_of_the_mode_objective():
    import math
    from scipy.ndimage import gaussian_filter
    import photonoxide_amd as P
    print("case adjoint workload 256x192x96, design 100x100x11, 6000 steps, K 64", flush=True)
    gpu = P.Gpu(0)
    width = math.sqrt(2.0 * math.log(2.0)) / (math.pi * (1 / 1.45 - 1 / 1.65))
    kw = dict(n=(256, 192, 96), pml=12, design=((78, 178), (46, 146), (42, 53)), steps=6000,
              checkpoint=64, core=(0.5, 0.22), pulse=(width, 6.0 * width), window=(30, 20))
    p32 = ModeDesignProblem(gpu, precision=P.Precision.F32, **kw)
    rho = gaussian_filter(np.random.default_rng(4).random(p32.shape), 2.0)
    sim, adj, steps = p32.build(rho)
    j32 = adj.forward(steps)
    g32 = p32.chain(adj.grad_eps(adj.backward()))
    info = adj.info()
    print(f"J {j32:.9e}, {info['checkpoints']} checkpoints, {info['checkpoint_bytes'] / 1e9:.1f} GB", flush=True)
    assert info["checkpoints"] == -(-steps // 64)
    plain = p32.objective(rho)
    assert abs(j32 - plain) < 1e-6 * plain, (j32, plain)
    del sim, adj
    p64 = ModeDesignProblem(gpu, precision=P.Precision.F64, **kw)
    rng = np.random.default_rng(5)
    d = 1e-3
    worst = 0.0
    for _ in range(4):
        c = tuple(int(rng.integers(0, s)) for s in p64.shape)
        hi, lo = rho.copy(), rho.copy()
        hi[c] += d
        lo[c] -= d
        fd = (p64.objective(hi) - p64.objective(lo)) / (2 * d)
        rel = abs(g32[c] - fd) / abs(fd)
        worst = max(worst, rel)
        print(f"cell {c}: adjoint (f32) {g32[c]:.9e}, central difference (f64) {fd:.9e}, relative {rel:.2e}", flush=True)
    assert worst < 1e-3, worst
