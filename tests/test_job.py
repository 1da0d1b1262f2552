"""The FDTD job front-end on the CPU (no GPU call): a job's task read and checked as the
reference's Setup::new (src/job/fdtd.rs:414-713), its grid (grid_of, :1726-1747), boundaries,
pulse and stop, its refusals, and the 2D effective index (the slab's fundamental mode) pinned by
the closed-form dispersion of a symmetric slab."""
import copy
import math

import pytest
from scipy.optimize import brentq

from jobs import DISK, MMI, STRAIGHT


def test_a_job_reads_as_the_reference_reads_it():
    from photonoxide_amd import job as J
    s = J.Setup(copy.deepcopy(STRAIGHT))
    g = s.grid
    assert (g.nx, g.ny, g.nz) == (160, 112, 1)
    assert g.dx == 0.025 and g.x0 == -2.0 and g.y0 == -1.4
    assert abs(g.z0 - (0.11 - 0.0125)) < 1e-15  # the layer's middle, half a cell below
    assert abs(s.dt - 0.9 / math.sqrt(2.0 / 0.025 ** 2)) < 1e-17
    assert s.wavelengths == [1.5 + 0.1 * k / 5 for k in range(6)]
    assert s.carrier == 1 / 1.55
    reach = max(abs(1 / w - 1 / 1.55) for w in s.wavelengths)
    assert s.bandwidth == max(2 * reach, 0.05 / 1.55)
    assert s.until == ("decay", 1e-6, 2000.0)
    assert s.pml == 20 and s.cpml(0) == (20, 20) and s.cpml(2) == (0, 0)
    src = s.sources[0]
    assert (src.kind, src.axis, src.plane, src.direction) == ("mode", 0, 40, 0)
    mon = s.monitors[0]
    assert (mon.kind, mon.plane, mon.window) == ("mode", 120, ((0, 112), (0, 1)))
    assert s.monitors[1].plane == round((1.2 + 2.0) / 0.025 - 0.5)
    assert s.probe_points() == [(120, 56, 0), (s.monitors[1].plane, 56, 0)]
    assert s.probe_components() == [(1, 2)]
    d = J.Setup(copy.deepcopy(DISK))
    assert not d.te and d.probe_components() == [(0, 2)] and d.until == ("time", None, 120.0)
    assert d.monitors[0].at == (88, 56, 0)


@pytest.mark.parametrize("change, reason", [
    (lambda t: t.update(polarization="xx"), "unknown polarization"),
    (lambda t: t.update(z_um=[0, 1]), "z_um is for a 3D job"),
    (lambda t: t.update(dimensions=3, polarization="te"), "polarization is for a 2D job"),
    (lambda t: t.update(courant=1.5), "courant must be in"),
    (lambda t: t.update(pml_cells=60), "leave nothing"),
    (lambda t: t.update(stop={"until": "time"}), "needs time_um"),
    (lambda t: t.update(stop={"until": "decay", "fraction": 2.0}), "fraction must be between"),
    (lambda t: t.update(stop={"until": "decay", "time_um": 3.0}), "time_um is for until"),
    (lambda t: t.update(source=[]), "at least one"),
    (lambda t: t["source"][0].update(at="-1.9 um"), "clear of the CPML"),
    (lambda t: t["source"][0].update(normal="q"), "unknown normal"),
    (lambda t: t["source"][0].update(waist_um=1.0), "doesn't take waist_um"),
    (lambda t: t["monitor"].append({"type": "mode", "name": "output", "normal": "x", "at": "0.5 um"}), "two monitors"),
    (lambda t: t["monitor"].append({"type": "field", "wavelengths_um": [3.0]}), "outside the pulse's band"),
    (lambda t: t.update(step="70 nm", pml_cells=5, spectrum={"from": "0.1 um", "to": "1.6 um", "points": 3}),
     "Nyquist"),
    (lambda t: t.update(layer="Ge"), "no layer Ge"),
    (lambda t: t.update(stack="sin_400"), "unknown stack"),
])
def test_the_check_refuses_what_the_reference_refuses(change, reason):
    from photonoxide_amd import job as J
    job = copy.deepcopy(STRAIGHT)
    change(job["task"])
    with pytest.raises(J.JobError, match=reason):
        J.Setup(job)


def symmetric_slab(n1, n2, d, lam, te):
    """The fundamental (even) mode of a symmetric slab in closed form: kappa tan(kappa d / 2) =
    gamma (TE), over n1^2 and n2^2 for TM."""
    k0 = 2 * math.pi / lam
    kappa = lambda n: k0 * math.sqrt(n1 * n1 - n * n)
    gamma = lambda n: k0 * math.sqrt(n * n - n2 * n2)
    w1, w2 = (1.0, 1.0) if te else (n1 * n1, n2 * n2)
    f = lambda n: kappa(n) / w1 * math.tan(kappa(n) * d / 2) - gamma(n) / w2
    top = n1 - 1e-13
    bottom = max(n2, math.sqrt(max(n1 * n1 - (math.pi / (k0 * d)) ** 2, 0.0))) + 1e-13
    return brentq(f, bottom, top, xtol=1e-15, rtol=1e-15)


@pytest.mark.parametrize("te", [True, False])
@pytest.mark.parametrize("d", [0.22, 0.15, 0.4])
def test_effective_index_is_the_slabs_closed_form(te, d):
    from photonoxide_amd import job as J
    got = J.slab_index((1.444, [(3.48, d)], 1.444), 1.55, te)
    want = symmetric_slab(3.48, 1.444, d, 1.55, te)
    assert abs(got - want) < 1e-12, (got, want)
    # the same slab cut into layers of the same index is the same slab
    assert abs(J.slab_index((1.444, [(3.48, d / 3)] * 3, 1.444), 1.55, te) - want) < 1e-12
    # TE's is above TM's, and both between the cladding's index and the core's
    assert 1.444 < J.slab_index((1.444, [(3.48, d)], 1.444), 1.55, False) < \
        J.slab_index((1.444, [(3.48, d)], 1.444), 1.55, True) < 3.48


def test_the_2d_scene_is_the_columns_effective_indices():
    from photonoxide_amd import job as J
    s = J.Setup(copy.deepcopy(STRAIGHT))
    sc = s.scene()
    n_in = symmetric_slab(3.48, 1.444, 0.22, 1.55, True)
    assert abs(sc.shapes[0].value - n_in ** 2) < 1e-10
    assert sc.background == (1.444 ** 2,) * 3
    box = sc.shapes[0]
    assert box.lo[:2] == (-2.5, -0.225) and box.hi[:2] == (2.5, 0.225)
    m = J.Setup(copy.deepcopy(MMI))
    assert len(m.scene().shapes) == 4
