"""The device path against the oracle, through the C ABI (GPU).

- f64 on the device is the oracle's f64, bit for bit: fields, CPML, sources, probes, transforms
  and dispersive media (every update the same operations in the same order, no contraction).
- f32 on the device is the oracle's f32 arithmetic, bit for bit, and within the north star's
  1e-5 relative L2 of the FP64 oracle.
- The tiling, splitting a run in pieces, and repeating a run change no bit.
"""
import numpy as np
import pytest

import oracle_lib as O
from fixtures import INF, busy_spec, oracle, product, product_fields, smooth_eps_spec

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def gpu():
    import photonoxide_amd as P
    return P.Gpu(0)


def dispersive_spec():
    spec = busy_spec(False)
    spec = dict(spec)
    spec["media"] = [
        (2.0, [("lorentz", 1.2, 1.6, 0.05), ("lorentz", 0.4, 2.5, 0.0)],
         (0.0, [(0, (-0.5, -0.3, -INF), (-0.1, 0.3, INF), 1.0)])),
        (1.0, [("drude", 2.0, 0.1), ("lorentz", 0.5, 1.9, 0.1), ("lorentz", 0.3, 2.4, 0.2)],
         (0.0, [(1, (0.35, -0.1, -INF), (0.2, 0, INF), 1.0)])),
    ]
    return spec


def with_media(spec, o, p):
    import photonoxide_amd as P
    for eps_inf, poles, inside in spec.get("media", []):
        o.add_medium(eps_inf, poles, O.scene(*inside))
        bg, shapes = inside
        p.with_medium(eps_inf, poles, P.Scene(bg, [P.Box(lo, hi, v) if k == 0 else
                                                    P.Cylinder((lo[0], lo[1]), hi[0], (lo[2], hi[2]), v)
                                                    for k, lo, hi, v in shapes]))


def build_pair(spec, gpu, precision):
    import photonoxide_amd as P
    o = oracle(spec)
    p = product(spec)
    with_media(spec, o, p)
    if precision == P.Precision.F32:
        o.use_f32()
    p.set_device(gpu, precision)
    return o, p


def same_bits(a, b):
    return np.array_equal(np.asarray(a, dtype=np.float64).view(np.uint64),
                          np.asarray(b, dtype=np.float64).view(np.uint64))


@pytest.mark.parametrize("kernel", [3, 5, 4, 1, 2, 0])
@pytest.mark.parametrize("shape", ["3d", "2d", "periodic"])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_device_is_the_oracle_bit_for_bit(gpu, shape, prec, kernel):
    import photonoxide_amd as P
    if kernel == 4 and shape != "2d":
        pytest.skip("the 2D march steps grids one cell thick with z periodic")
    if kernel == 5 and shape == "2d":
        pytest.skip("the two-step march steps 3D grids")
    precision = P.Precision.F64 if prec == "f64" else P.Precision.F32
    spec = busy_spec(shape == "2d", shape == "periodic")
    print(f"case busy {shape} {prec} kernel {kernel}", flush=True)
    o, p = build_pair(spec, gpu, precision)
    p.set_kernel(kernel)
    o.run(80)
    for piece in (33, 8, 1, 38):  # pairs, an odd piece, a lone step, across staging blocks
        p.run(piece)
    fo, fp = o.fields(), product_fields(p)
    diff = np.abs(fo - fp).max()
    assert same_bits(fo, fp), f"largest difference {diff}"
    for n in range(len(spec["probes"])):
        assert same_bits(o.probe(n), p.probe(n)), n
    for f in (0, 1):
        for c in range(3):
            assert same_bits(o.dft(0, f, c).view(np.float64), p.dft(0, f, c).view(np.float64)), (f, c)


@pytest.mark.parametrize("kernel", [3, 1, 0])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_dispersive_media_bit_for_bit(gpu, prec, kernel):
    """The media's update runs as a pass of its own after every step kernel."""
    import photonoxide_amd as P
    precision = P.Precision.F64 if prec == "f64" else P.Precision.F32
    print(f"case dispersive {prec} kernel {kernel}", flush=True)
    o, p = build_pair(dispersive_spec(), gpu, precision)
    p.set_kernel(kernel)
    o.run(120)
    p.run(57)
    p.run(63)
    fo, fp = o.fields(), product_fields(p)
    assert same_bits(fo, fp), np.abs(fo - fp).max()
    for n in range(len(dispersive_spec()["probes"])):
        assert same_bits(o.probe(n), p.probe(n)), n


@pytest.mark.parametrize("kernel", [3, 4, 1, 0])
@pytest.mark.parametrize("lossy", [False, True])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_coefficient_arrays_bit_for_bit(gpu, prec, lossy, kernel):
    """Beyond 256 distinct coefficients: full ca, cb arrays (mode 1) for a lossy problem, cb
    alone (mode 2) for a lossless one."""
    import photonoxide_amd as P
    from photonoxide_amd import _native as N
    if kernel == 4:
        spec = smooth_eps_spec(True, lossy)
    else:
        spec = smooth_eps_spec(False, lossy)
    precision = P.Precision.F64 if prec == "f64" else P.Precision.F32
    print(f"case coefficient arrays {prec} lossy {lossy} kernel {kernel}", flush=True)
    o, p = build_pair(spec, gpu, precision)
    p.set_kernel(kernel)
    assert N.lib.pxfdtd_resident_coefficient_mode(p.resident()) == (1 if lossy else 2)
    o.run(80)
    p.run(80)
    fo, fp = o.fields(), product_fields(p)
    assert same_bits(fo, fp), np.abs(fo - fp).max()
    for n in range(len(spec["probes"])):
        assert same_bits(o.probe(n), p.probe(n)), n


@pytest.mark.parametrize("kernel", [4, 3])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_dispersive_media_2d_bit_for_bit(gpu, prec, kernel):
    import photonoxide_amd as P
    precision = P.Precision.F64 if prec == "f64" else P.Precision.F32
    spec = dict(busy_spec(True))
    spec["media"] = [
        (2.0, [("lorentz", 1.2, 1.6, 0.05), ("lorentz", 0.4, 2.5, 0.0)],
         (0.0, [(0, (-0.8, -0.3, -INF), (-0.1, 0.3, INF), 1.0)])),
        (1.0, [("drude", 2.0, 0.1), ("lorentz", 0.5, 1.9, 0.1)],
         (0.0, [(1, (0.5, 0.1, -INF), (0.35, 0, INF), 1.0)])),
    ]
    print(f"case dispersive 2d {prec} kernel {kernel}", flush=True)
    o, p = build_pair(spec, gpu, precision)
    p.set_kernel(kernel)
    o.run(150)
    p.run(150)
    fo, fp = o.fields(), product_fields(p)
    assert same_bits(fo, fp), np.abs(fo - fp).max()


def test_f32_within_north_star_tolerance_of_f64_oracle(gpu):
    """Fields in f32 within 1e-5 relative L2 of the FP64 reference scheme after N steps."""
    import photonoxide_amd as P
    spec = busy_spec(False)
    o = oracle(spec)
    p = product(spec)
    p.set_device(gpu, P.Precision.F32)
    steps = 400
    o.run(steps)
    p.run(steps)
    fo, fp = o.fields(), product_fields(p)
    rel = np.linalg.norm(fo - fp) / np.linalg.norm(fo)
    print(f"f32 vs f64 oracle, {steps} steps: relative L2 {rel:.3e}")
    assert rel < 1e-5, rel


@pytest.mark.parametrize("kernel,rows,planes", [(0, 2, 1), (0, 2, 3), (0, 8, 19), (0, 4, 5),
                                                (1, 1, 1), (1, 2, 3), (1, 4, 19), (2, 1, 7), (2, 2, 2),
                                                (3, 1, 1), (3, 2, 3), (3, 4, 19), (3, 1, 7), (3, 2, 2),
                                                (5, 2, 1), (5, 2, 5), (5, 3, 16), (5, 3, 3), (5, 2, 19)])
def test_tiling_changes_no_bit(gpu, kernel, rows, planes):
    import photonoxide_amd as P
    spec = busy_spec(False)
    ref = product(spec)
    ref.set_device(gpu, P.Precision.F32)
    ref.run(50)
    p = product(spec)
    p.set_device(gpu, P.Precision.F32)
    p.set_kernel(kernel)
    p.set_tiling(rows, planes)
    assert p.tiling() == (rows, planes) and p.kernel() == kernel
    p.run(50)
    assert same_bits(product_fields(ref), product_fields(p))


def test_pieces_equal_one_run_and_repeat(gpu):
    import photonoxide_amd as P
    spec = busy_spec(False)
    a = product(spec)
    a.set_device(gpu, P.Precision.F32)
    a.run(60)
    b = product(spec)
    b.set_device(gpu, P.Precision.F32)
    b.run(17)
    b.run(43)
    c = product(spec)
    c.set_device(gpu, P.Precision.F32)
    c.run(60)
    fa = product_fields(a)
    assert same_bits(fa, product_fields(b))
    assert same_bits(fa, product_fields(c))
    for n in range(len(spec["probes"])):
        assert same_bits(a.probe(n), b.probe(n))


def te_tm_spec(which):
    """The 2D busy problem with sources on one field set only and no starting fields: TE is Ex,
    Ey, Hz, TM is Ez, Hx, Hy."""
    spec = dict(busy_spec(True))
    nx, ny, _ = spec["grid"][:3]
    mid = (nx // 2, ny // 2, 0)
    w = dict(kind=0, frequency=1.0, width=0.3, delay=1.2)
    if which == "te":
        spec["sources"] = [(0, 1, mid, w), (1, 2, (nx // 4, ny // 2, 0), w)]
        spec["currents"] = [(0, [(0, mid, 0.5 - 0.25j), (1, (nx // 2 + 3, ny // 3, 0), -0.3 + 0.7j)], w)]
    else:
        spec["sources"] = [(0, 2, mid, w), (1, 0, (nx // 4, ny // 2, 0), w)]
        spec["currents"] = [(1, [(1, mid, 0.5 - 0.25j)], w)]
    return spec


@pytest.mark.parametrize("which", ["te", "tm"])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_2d_march_steps_one_field_set_bit_for_bit(gpu, which, prec):
    import photonoxide_amd as P
    precision = P.Precision.F64 if prec == "f64" else P.Precision.F32
    spec = te_tm_spec(which)
    print(f"case 2d {which} {prec}", flush=True)
    o = oracle(spec, with_fields=False)
    p = product(spec, with_fields=False)
    if precision == P.Precision.F32:
        o.use_f32()
    p.set_device(gpu, precision)
    assert p.kernel() == 4
    o.run(90)
    p.run(90)
    fo, fp = o.fields(), product_fields(p)
    assert same_bits(fo, fp), np.abs(fo - fp).max()
    other = [2, 3, 4] if which == "te" else [0, 1, 5]  # Ez, Hx, Hy or Ex, Ey, Hz
    assert not np.any(fp[other]), "the field set without values or sources stays zero"
    assert np.any(fp[[q for q in range(6) if q not in other]])
    for n in range(len(spec["probes"])):
        assert same_bits(o.probe(n), p.probe(n)), n
    for f in (0, 1):
        for c in range(3):
            assert same_bits(o.dft(0, f, c).view(np.float64), p.dft(0, f, c).view(np.float64)), (f, c)


def test_device_again_midrun_changes_no_bit(gpu):
    """Fields, psi and monitor sums brought back to the host and put on a new resident carry on
    as if the run had not stopped (the host's fields are no longer the zeros a new simulation
    holds, so they go up in full)."""
    import photonoxide_amd as P
    spec = busy_spec(False)
    a = product(spec)
    a.set_device(gpu, P.Precision.F32)
    a.run(60)
    b = product(spec)
    b.set_device(gpu, P.Precision.F32)
    b.run(23)
    b.set_device(gpu, P.Precision.F32)
    b.run(37)
    assert same_bits(product_fields(a), product_fields(b))
    for n in range(len(spec["probes"])):
        assert same_bits(a.probe(n), b.probe(n))


@pytest.mark.parametrize("rows_per_chunk", [1, 5, 45])
def test_2d_march_chunks_change_no_bit(gpu, rows_per_chunk):
    import photonoxide_amd as P
    spec = busy_spec(True)
    ref = product(spec)
    ref.set_device(gpu, P.Precision.F32)
    ref.set_kernel(3)
    ref.run(50)
    p = product(spec)
    p.set_device(gpu, P.Precision.F32)
    assert p.kernel() == 4
    p.set_tiling(1, rows_per_chunk)
    p.run(50)
    assert same_bits(product_fields(ref), product_fields(p))
