"""The plain 2D TE workload (400 x 400, 2000 steps) on the device against the oracle (GPU).

- f32 on the device is the oracle's f32, bit for bit: fields, probes and the transform sums,
  with both step kernels that take a grid one cell thick;
- against the FP64 oracle the f32 fields are within the 1e-5 relative L2 the time-stepper is
  held to, and each of the 5 spectra (the transform line's E_y at one wavelength) within 1e-4
  relative; the error of each is printed.

The 2000 steps end before the pulse's peak (its delay is 6 widths = 26.9 um/c, the run
14.1 um/c), so the spectra are those of the pulse's leading edge, which is what the workload
defines.
"""
import numpy as np
import pytest

from configs import te_strip_2d
from fixtures import oracle, product, product_fields

pytestmark = pytest.mark.gpu

STEPS = 2000


def same_bits(a, b):
    return np.array_equal(np.asarray(a).view(np.uint8), np.asarray(b).view(np.uint8))


@pytest.fixture(scope="module")
def gpu():
    import photonoxide_amd as P
    return P.Gpu(0)


@pytest.fixture(scope="module")
def oracles():
    spec = te_strip_2d()
    o64 = oracle(spec, with_fields=False)
    o64.run(STEPS)
    o32 = oracle(spec, with_fields=False)
    o32.use_f32()
    o32.run(STEPS)
    return spec, o64, o32


@pytest.mark.parametrize("kernel", [4, 3])
def test_te_strip_400x400_is_the_oracle_f32_bit_for_bit(gpu, oracles, kernel):
    import photonoxide_amd as P
    spec, o64, o32 = oracles
    print(f"case te_strip_2d 400x400 f32 kernel {kernel}", flush=True)
    p = product(spec, with_fields=False)
    p.set_device(gpu, P.Precision.F32)
    p.set_kernel(kernel)
    p.run(STEPS)
    fo, fp = o32.fields(), product_fields(p)
    assert np.abs(fo).max() > 0
    assert same_bits(fo, fp), np.abs(fo - fp).max()
    for n in range(len(spec["probes"])):
        assert same_bits(o32.probe(n), p.probe(n)), n
    for f in (0, 1):
        for c in range(3):
            assert same_bits(o32.dft(0, f, c), p.dft(0, f, c)), (f, c)


def test_te_strip_400x400_f32_against_the_fp64_oracle(gpu, oracles):
    import photonoxide_amd as P
    spec, o64, _ = oracles
    p = product(spec, with_fields=False)
    p.set_device(gpu, P.Precision.F32)
    p.run(STEPS)
    fo, fp = o64.fields(), product_fields(p)
    rel = np.linalg.norm(fo - fp) / np.linalg.norm(fo)
    print(f"fields, f32 device vs FP64 oracle after {STEPS} steps: relative L2 {rel:.3e}")
    assert rel < 1e-5, rel
    ref, got = o64.dft(0, 0, 1), p.dft(0, 0, 1)  # E_y on the line, [f][k][j][i]
    lam = 1 / np.asarray(spec["dft"][2])
    for q in range(ref.shape[0]):
        err = np.linalg.norm(got[q] - ref[q]) / np.linalg.norm(ref[q])
        print(f"spectrum at {lam[q]:.3f} um: relative L2 error vs FP64 oracle {err:.3e}")
        assert err < 1e-4, (q, err)
