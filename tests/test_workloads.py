"""The 3D workloads the time-stepper is quoted on, on the device (GPU).

The 1x2 MMI on SOI (384 x 256 x 192 cells of 20 nm, 12-cell CPML, uint8 material ids, 21
frequencies on both output ports) and the dispersive problem (512 x 512 x 256 cells of 15 nm, Si
pillars with a 2-pole Lorentz medium and Au blocks with Drude + 2 Lorentz, 16-cell CPML):

- at a size the oracle steps in seconds, the same problem (tests/configs.py) is the oracle's
  bits on the device, in f64 and in f32 (fields, probes, every transform sum), and the f32
  device is within the north star's 1e-5 relative L2 of the FP64 oracle (fields) and 1e-4
  (each port's spectrum at each frequency); every error is printed;
- at full size, where the oracle would take hours, size-independent properties: the f32 run
  is within 1e-5 relative L2 of the f64 run of the same kernels (fields) and 1e-4 (spectra)
  after the pulse has entered the grid, another tiling and another step kernel give the f32
  run's bits, and the f64 run is the f64 oracle's bits where the oracle can follow it (the
  reduced problem above, the same code).

The adjoint workload (256 x 192 x 96 cells, a 100 x 100 x 11 design block, 6000 steps with a
checkpoint every 64) is in tests/test_adjoint_workload.py.
"""
import os
import sys

import numpy as np
import pytest

from configs import dispersive_3d, mmi_1x2
from fixtures import oracle, product, product_fields

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def same_bits(a, b):
    return np.array_equal(np.asarray(a, dtype=np.float64).view(np.uint64),
                          np.asarray(b, dtype=np.float64).view(np.uint64))


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def spectral_error(sp, ref, n, nf):
    """The largest over frequencies of the transform's difference (E_y over the port's window)
    relative to the largest magnitude of the reference's at that frequency."""
    worst = 0.0
    a, b = sp(n), ref(n)
    for f in range(nf):
        worst = max(worst, float(np.abs(a[f] - b[f]).max() / np.abs(b[f]).max()))
    return worst


@pytest.fixture(scope="module")
def gpu():
    import photonoxide_amd as P
    return P.Gpu(0)


def with_media(spec, o, p):
    import oracle_lib as O
    import photonoxide_amd as P
    for eps_inf, poles, inside in spec.get("media", []):
        bg, shapes = inside
        if o is not None:
            o.add_medium(eps_inf, poles, O.scene(*inside))
        if p is not None:
            p.with_medium(eps_inf, poles, P.Scene(bg, [P.Box(lo, hi, v) if k == 0 else
                                                        P.Cylinder((lo[0], lo[1]), hi[0], (lo[2], hi[2]), v)
                                                        for k, lo, hi, v in shapes]))


# ------------------------------------------------------------------ reduced sizes vs the oracle

REDUCED = {"mmi": (mmi_1x2, 1200), "dispersive": (dispersive_3d, 400)}


@pytest.fixture(scope="module")
def oracle_runs():
    """The oracle's f64 and f32 runs of each reduced workload, once for the module."""
    out = {}
    for name, (make, steps) in REDUCED.items():
        spec = make()
        runs = {}
        for prec in ("f64", "f32"):
            o = oracle(spec, with_fields=False)
            with_media(spec, o, None)
            if prec == "f32":
                o.use_f32()
            o.run(steps)
            runs[prec] = o
        out[name] = (spec, steps, runs)
    return out


@pytest.mark.parametrize("name", ["mmi", "dispersive"])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_reduced_workload_is_the_oracle_bit_for_bit(gpu, oracle_runs, name, prec):
    import photonoxide_amd as P
    spec, steps, runs = oracle_runs[name]
    print(f"case {name} reduced {spec['grid'][:3]} {prec}, {steps} steps", flush=True)
    o = runs[prec]
    p = product(spec, with_fields=False)
    with_media(spec, None, p)
    p.set_device(gpu, P.Precision.F64 if prec == "f64" else P.Precision.F32)
    p.run(steps // 2)
    p.run(steps - steps // 2)
    fo, fp = o.fields(), product_fields(p)
    assert same_bits(fo, fp), np.abs(fo - fp).max()
    for n in range(len(spec["probes"])):
        assert same_bits(o.probe(n), p.probe(n)), n
    ndft = len(spec.get("dfts", [])) + (1 if spec.get("dft") else 0)
    for n in range(ndft):
        for f in (0, 1):
            for c in range(3):
                assert same_bits(o.dft(n, f, c).view(np.float64), p.dft(n, f, c).view(np.float64)), (n, f, c)


@pytest.mark.parametrize("name", ["mmi", "dispersive"])
def test_reduced_workload_f32_within_the_bar_of_fp64(gpu, oracle_runs, name):
    import photonoxide_amd as P
    spec, steps, runs = oracle_runs[name]
    p = product(spec, with_fields=False)
    with_media(spec, None, p)
    p.set_device(gpu, P.Precision.F32)
    p.run(steps)
    o64 = runs["f64"]
    err = rel_l2(product_fields(p), o64.fields())
    ndft = len(spec.get("dfts", [])) + (1 if spec.get("dft") else 0)
    nf = len((spec.get("dfts") or [spec.get("dft")])[0][2])
    spec_err = max(spectral_error(lambda n: p.dft(n, 0, 1), lambda n: o64.dft(n, 0, 1), n, nf) for n in range(ndft))
    print(f"{name} reduced, {steps} steps: f32 device vs FP64 oracle: fields {err:.2e} relative L2, "
          f"spectra {spec_err:.2e}")
    assert err < 1e-5
    assert spec_err < 1e-4


# ------------------------------------------------------------------ full sizes: properties

def full_problem(name):
    sys.path.insert(0, ROOT)
    import bench
    return bench.mmi_problem()[0] if name == "mmi" else bench.dispersive_problem()[0]


FULL = {"mmi": 3000, "dispersive": 1500}


@pytest.fixture(scope="module")
def full_runs(gpu):
    """Each full-size workload run in f64 and in f32 on the device, fields and spectra kept."""
    import photonoxide_amd as P
    out = {}
    for name, steps in FULL.items():
        res = {}
        for prec in (P.Precision.F64, P.Precision.F32):
            print(f"case {name} full size {prec.name}, {steps} steps", flush=True)
            s = full_problem(name)
            s.set_device(gpu, prec)
            s.run(steps)
            ndft = 2 if name == "mmi" else 1
            res[prec.name] = (product_fields(s), [s.dft(n, 0, 1) for n in range(ndft)])
            del s
        out[name] = res
    return out


@pytest.mark.parametrize("name", ["mmi", "dispersive"])
def test_full_size_f32_within_the_bar_of_f64(full_runs, name):
    r = full_runs[name]
    f64, d64 = r["F64"]
    f32, d32 = r["F32"]
    err = rel_l2(f32, f64)
    spec_err = 0.0
    for a, b in zip(d32, d64):
        for f in range(a.shape[0]):
            spec_err = max(spec_err, float(np.abs(a[f] - b[f]).max() / np.abs(b[f]).max()))
    print(f"{name} full size, {FULL[name]} steps: f32 vs f64 fields {err:.2e} relative L2, spectra {spec_err:.2e}; "
          f"largest field {np.abs(f64).max():.3e}")
    assert np.abs(f64).max() > 1e-3  # the pulse is in
    assert err < 1e-5
    assert spec_err < 1e-4


@pytest.mark.parametrize("name", ["mmi", "dispersive"])
def test_full_size_tiling_and_kernel_change_no_bit(gpu, name):
    """Another tiling of the aligned march and the 62-wide march (kernel 1) give the default
    run's f32 bits over the whole grid, at a quarter of the steps."""
    import photonoxide_amd as P
    steps = FULL[name] // 4
    out = []
    for setting in ("default", "tiling", "kernel1"):
        s = full_problem(name)
        s.set_device(gpu, P.Precision.F32)
        if setting == "tiling":
            s.set_tiling(1, 16)
        elif setting == "kernel1":
            s.set_kernel(1)
        s.run(steps)
        out.append(product_fields(s))
        del s
    assert same_bits(out[0], out[1])
    assert same_bits(out[0], out[2])
