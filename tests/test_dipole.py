"""Dipoles (Simulation::add_dipole, src/fdtd/sources.rs:242-315): a point current restricted to
the grid by trilinear weights over the component's 8 nearest values, divided by the cell volume.

CPU: the oracle's restatement (or_dipole_values) is pinned by what the restriction must do,
with the values' positions computed here from the Yee convention (E_c half a step on along c,
H_c half a step on across c; src/fdfd/three.rs:137-156), not from the restatement:
- partition of unity: the weights sum to a / volume wherever the dipole sits;
- first moments: sum w x_value = the dipole's position along each axis (linear interpolation
  reproduces linear functions), which fixes which values and which half-step offsets are used;
- a dipole on a value is that one value; at a wall's last value the clamped twin takes nothing;
  across a periodic end the twin wraps to the first value; an axis one cell thick and periodic
  (2D) takes no step in the volume; outside a walled axis's values or not finite it is refused.
A planted mistake (the E half-step taken across instead of along the component) breaks the
first moment.

GPU: the device with dipoles (the host mirror's add_dipole) is the oracle with the restated
values as a current, bit for bit, in f64 and f32.
"""
import numpy as np
import pytest

import oracle_lib as O
from fixtures import busy_spec, oracle, product, product_fields

H = 0.05
N3 = (11, 9, 7)
GRID = (*N3, H, H, H, -0.3, -0.2, -0.15)


def value_position(field, comp, r, grid=GRID):
    nx, ny, nz = grid[:3]
    i, j, k = r % nx, (r // nx) % ny, r // (nx * ny)
    p = np.array([grid[6] + i * grid[3], grid[7] + j * grid[4], grid[8] + k * grid[5]])
    for a in range(3):
        if (field == 0 and a == comp) or (field == 1 and a != comp):
            p[a] += 0.5 * grid[3 + a]
    return p


@pytest.mark.parametrize("field", [0, 1])
@pytest.mark.parametrize("comp", [0, 1, 2])
def test_weights_are_a_partition_of_unity_with_the_positions_first_moments(field, comp):
    rng = np.random.default_rng(10 + 3 * field + comp)
    b = O.bounds(low=(2, 2, 2), high=(2, 2, 2))
    volume = H ** 3
    for _ in range(50):
        # anywhere between the component's first and last values
        lo = value_position(field, comp, 0)
        hi = value_position(field, comp, N3[0] * N3[1] * N3[2] - 1)
        pos = lo + rng.random(3) * (hi - lo)
        amp = complex(rng.standard_normal(), rng.standard_normal())
        idx, a = O.dipole_values(GRID, b, field, comp, pos, amp)
        assert len(idx) <= 8 and np.all(np.diff(idx) > 0)  # index order, each value once
        assert abs(a.sum() - amp / volume) < 1e-13 * abs(amp / volume)
        w = (a / amp).real * volume
        for ax in range(3):
            moment = sum(wi * value_position(field, comp, int(r))[ax] for wi, r in zip(w, idx))
            assert abs(moment - pos[ax]) < 1e-13, (ax, moment, pos[ax])


def test_special_positions():
    """On a grid whose positions are exact in binary (steps of 1/16 from 0), so that a position
    on a value is exactly there."""
    he = 0.0625
    grid = (*N3, he, he, he, 0.0, 0.0, 0.0)
    b = O.bounds(low=(2, 2, 2), high=(2, 2, 2))
    vol = he ** 3
    # on a value: that value alone
    r = (3 * N3[1] + 4) * N3[0] + 5
    idx, a = O.dipole_values(grid, b, 0, 1, value_position(0, 1, r, grid), 2.0 - 1.0j)
    assert list(idx) == [r] and a[0] == (2.0 - 1.0j) / vol
    # at the last value along x (a wall): the clamped twin gets a zero weight and is dropped
    last = value_position(0, 1, N3[0] - 1, grid)
    idx, a = O.dipole_values(grid, b, 0, 1, last, 1.0)
    assert list(idx) == [N3[0] - 1]
    # beyond the last value along a walled axis: refused; not finite: refused
    for bad in (last + np.array([0.5 * he, 0, 0]), np.array([np.nan, 0.0, 0.0])):
        with pytest.raises(ValueError):
            O.dipole_values(grid, b, 0, 1, bad, 1.0)
    # periodic x: halfway past the last value, half goes to the first (its wrapped twin)
    bp = O.bounds(periodic=(1, 0, 0), low=(0, 2, 2), high=(0, 2, 2))
    idx, a = O.dipole_values(grid, bp, 0, 1, last + np.array([0.5 * he, 0, 0]), 1.0)
    assert sorted(idx) == [0, N3[0] - 1] and np.allclose(a, 0.5 / vol, rtol=1e-15)
    # a grid one cell thick, periodic along z: z takes weight 1 and no step in the volume
    g2 = (11, 9, 1, H, H, H, -0.3, -0.2, 0.0)
    b2 = O.bounds(periodic=(0, 0, 1), low=(2, 2, 0), high=(2, 2, 0))
    idx, a = O.dipole_values(g2, b2, 0, 2, (0.0, 0.0, 123.0), 1.0)
    assert abs(a.sum() - 1.0 / H ** 2) < 1e-12 / H ** 2


def test_a_planted_half_step_breaks_the_first_moment():
    """The restriction with E's half step taken across the component (H's convention) puts the
    weights' centre half a step off the dipole along two axes."""
    b = O.bounds(low=(2, 2, 2), high=(2, 2, 2))
    pos = np.array([0.012, -0.031, 0.017])
    idx, a = O.dipole_values(GRID, b, 1, 0, pos, 1.0)  # an H_x restriction used for E_x
    w = a.real * H ** 3
    off = [abs(sum(wi * value_position(0, 0, int(r))[ax] for wi, r in zip(w, idx)) - pos[ax]) for ax in range(3)]
    assert max(off) > 0.4 * H, off


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_device_dipoles_are_the_oracles_restriction_bit_for_bit(prec):
    import photonoxide_amd as P
    print(f"case dipoles {prec}", flush=True)
    spec = busy_spec(False)
    o = oracle(spec)
    p = product(spec)
    w = dict(kind=0, frequency=1.0, width=0.3, delay=1.0)
    dipoles = [(0, 2, (0.013, -0.021, 0.04), 0.7 - 0.2j), (1, 0, (-0.29, 0.11, -0.06), -1.3 + 0.4j),
               (0, 1, (0.2, 0.0, 0.0), 1.0)]
    b = O.bounds(**spec["bounds"])
    for field, comp, pos, amp in dipoles:
        idx, a = O.dipole_values(spec["grid"], b, field, comp, pos, amp)
        o.add_current(field, np.full(len(idx), comp), idx, a, O.wave(**{k: v for k, v in w.items() if k != "kind"}, kind=0))
        p.add_dipole(P.Dipole(P.Field(field), P.Axis(comp), pos, amp, P.Waveform(0, w["frequency"], w["width"], w["delay"])))
    gpu = P.Gpu(0)
    if prec == "f32":
        o.use_f32()
    p.set_device(gpu, P.Precision.F64 if prec == "f64" else P.Precision.F32)
    o.run(70)
    p.run(70)
    fo, fp = o.fields(), product_fields(p)
    assert np.array_equal(fo.view(np.uint64), fp.view(np.uint64)), np.abs(fo - fp).max()
