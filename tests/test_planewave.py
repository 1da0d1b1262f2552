"""Total-field/scattered-field plane waves on the device against the oracle, through the C ABI
(GPU).

- f64 on the device is the oracle's f64 bit for bit, f32 the oracle's f32 arithmetic (the
  auxiliary grid in FP64 in both, as the reference keeps it), with every step kernel: fields,
  probes, transforms and the incident field Simulation::incident reports.
- Two waves on one box, a wave added after the run started, and a run split in pieces.
- The box leaks only round-off at the reference's angles (src/fdtd/tests.rs:252-265) in f64,
  and f32 round-off in f32.
"""
import numpy as np
import pytest

from fixtures import oracle, product, product_fields, pulse, tfsf_spec

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import photonoxide_amd as P
    return P.Gpu(0)


def same_bits(a, b):
    return np.array_equal(np.asarray(a, dtype=np.float64).view(np.uint64),
                          np.asarray(b, dtype=np.float64).view(np.uint64))


def pair(spec, gpu, prec, kernel=None):
    import photonoxide_amd as P
    precision = P.Precision.F64 if prec == "f64" else P.Precision.F32
    o, p = oracle(spec), product(spec)
    if prec == "f32":
        o.use_f32()
    p.set_device(gpu, precision)
    if kernel is not None:
        p.set_kernel(kernel)
    return o, p


def check(spec, o, p):
    fo, fp = o.fields(), product_fields(p)
    assert same_bits(fo, fp), f"largest difference {np.abs(fo - fp).max()}"
    for n in range(len(spec["probes"])):
        assert same_bits(o.probe(n), p.probe(n)), n
    for f in (0, 1):
        for c in range(3):
            assert same_bits(o.dft(0, f, c).view(np.float64), p.dft(0, f, c).view(np.float64)), (f, c)


@pytest.mark.parametrize("kernel", [3, 4, 1, 2, 0])
@pytest.mark.parametrize("shape", ["3d", "2d"])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_plane_wave_is_the_oracle_bit_for_bit(gpu, shape, prec, kernel):
    if kernel == 4 and shape != "2d":
        pytest.skip("the 2D march steps grids one cell thick with z periodic")
    spec = tfsf_spec(shape == "2d")
    print(f"case tfsf {shape} {prec} kernel {kernel}", flush=True)
    o, p = pair(spec, gpu, prec, kernel)
    o.run(120)
    p.run(120)
    check(spec, o, p)
    # the incident field the box sees, at values in, on and beyond it
    lo, hi = spec["planes"][0][0], spec["planes"][0][1]
    for at in [lo, hi[:2] + (0,), ((lo[0] + hi[0]) // 2, (lo[1] + hi[1]) // 2, 0), (1, 1, 0)]:
        for f in (0, 1):
            for c in range(3):
                assert same_bits(o.incident(0, f, c, at), p.incident(0, f, c, at)), (at, f, c)


@pytest.mark.parametrize("shape", ["3d", "2d"])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_two_plane_waves_and_a_split_run(gpu, shape, prec):
    spec = tfsf_spec(shape == "2d", second=True)
    print(f"case tfsf two waves {shape} {prec}", flush=True)
    o, p = pair(spec, gpu, prec)
    o.run(120)
    for m in (37, 1, 82):
        p.run(m)
    check(spec, o, p)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_plane_wave_added_after_the_run_started(gpu, prec):
    """Its auxiliary grid starts at rest when it is added (the reference allows it at any time);
    the values stepped before stay the device's."""
    spec = tfsf_spec(True)
    late = spec["planes"]
    spec = dict(spec, planes=[])
    print(f"case tfsf late {prec}", flush=True)
    o, p = pair(spec, gpu, prec)
    o.run(30)
    p.run(30)
    import photonoxide_amd as P
    for lo, hi, d, pol, eps, w in late:
        import oracle_lib as O
        o.add_plane_wave(lo, hi, d, pol, eps, O.wave(**{k: v for k, v in w.items() if k != "kind"}, kind=w["kind"]))
        p.add_plane_wave(P.PlaneWave(lo, hi, d, pol, eps, P.Waveform(w["kind"], w["frequency"], w["width"], w["delay"])))
    o.run(90)
    p.run(90)
    check(spec, o, p)


def leakage(sim, n, low, high, steps):
    """The largest field in the scattered region over the run relative to the largest in the
    box (src/fdtd/checks.rs:511-563)."""
    i = np.arange(n[0] * n[1] * n[2])
    at = [i % n[0], (i // n[0]) % n[1], i // (n[0] * n[1])]
    inside = {}
    for f in (0, 1):
        for c in range(3):
            m = np.ones(i.size, bool)
            for a in range(3):
                if n[a] == 1:
                    continue
                t = 2 * at[a] + ((c == a) if f == 0 else (c != a))
                m &= (t >= 2 * low[a]) & (t <= 2 * high[a])
            inside[f, c] = m
    leak = largest = 0.0
    for _ in range(steps // 10):
        sim.run(10)
        for (f, c), m in inside.items():
            v = np.abs(sim.e(c) if f == 0 else sim.h(c))
            largest = max(largest, v[m].max())
            leak = max(leak, v[~m].max())
    return leak / largest


@pytest.mark.parametrize("direction,polarization,three_d", [
    ((1, 0, 0), (0, 0, 1), False), ((2, 1, 0), (0, 0, 1), False), ((1, -3, 0), (1, 0, 0), False),
    ((0, 0, 1), (1, 0, 0), True), ((2, 1, 1), (1, -1, 0), True)])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_box_leaks_only_round_off(gpu, direction, polarization, three_d, prec):
    """The reference's leakage check on the device: < 1e-12 in f64 (the reference measured
    1.3e-15); in f32, < 1e-5 (the couplings cancel the incident field to f32 round-off)."""
    import photonoxide_amd as P
    h = 0.05
    if three_d:
        n, low, high, cp = (36, 36, 36), (12, 12, 12), (24, 24, 24), 8
        b = P.Boundaries(P.Pml(cp, cp), P.Pml(cp, cp), P.Pml(cp, cp))
    else:
        n, low, high, cp = (80, 80, 1), (25, 25, 0), (55, 55, 1), 10
        b = P.Boundaries(P.Pml(cp, cp), P.Pml(cp, cp), P.Periodic())
    g = P.Grid3d(*n, h, h, h, -n[0] * h / 2, -n[1] * h / 2, -n[2] * h / 2)
    s = P.Simulation.new(g, P.Scene(1.0, []), b, 0.9)
    s.add_plane_wave(P.PlaneWave(low, high, direction, polarization, 1.0, P.Waveform.pulse(1.0, 0.6)))
    s.set_device(gpu, P.Precision.F64 if prec == "f64" else P.Precision.F32)
    print(f"case leakage {direction} {prec}", flush=True)
    leak = leakage(s, n, low, high, int(12.0 / s.dt()))
    print(f"leakage {direction} {polarization} {prec}: {leak:.2e}")
    assert leak < (1e-12 if prec == "f64" else 1e-5), leak
