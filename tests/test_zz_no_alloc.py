"""The allocation-free flux (GPU): pxfdtd_flux_device_no_alloc against pxfdtd_flux_device,
through the C ABI, bit for bit, on two calls in a row, the second smaller than the first (the
scratch keeps the first call's terms, which the second must not see), and the refusal of a
flux beyond its 2^21 terms."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import flux as F  # noqa: E402
from devmem import Buffer  # noqa: E402
from test_flux_device import N3, device_face, random_series  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    from photonoxide_amd import _native as N
    d = C.c_void_p()
    N.check(N.lib.pxfdtd_device_open(0, C.byref(d)))
    yield d
    N.lib.pxfdtd_device_close(d)


def both(dev, faces, nf):
    from photonoxide_amd import _native as N
    arr = (N.FluxFace * len(faces))(*faces)
    a, b = Buffer(8 * nf), Buffer(8 * nf)
    N.check(N.lib.pxfdtd_flux_device(dev, arr, len(faces), a.address, None))
    N.check(N.lib.pxfdtd_flux_device_no_alloc(dev, arr, len(faces), b.address, None))
    # hipMemcpy on the null stream waits for the kernels queued there
    return a.read(np.float64, nf), b.read(np.float64, nf)


def test_no_alloc_flux_is_the_allocating_flux_on_two_calls_the_second_smaller(dev):
    print("case allocation-free flux: a box of 6 faces, nf 5, then one windowed plane, nf 2", flush=True)
    rng = np.random.default_rng(21)
    keep = []
    faces = F.box_faces(((1, 8), (1, 7), (1, 6)))
    dfaces = []
    for face in faces:
        lo, hi = face.box(N3)
        dfaces.append(device_face(face, random_series(rng, lo, hi, 5), lo, hi, 5, keep))
    want, got = both(dev, dfaces, 5)
    assert np.all(np.isfinite(want))
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (got, want)
    face = F.Face(1, 3, ((2, 5), (2, 4)))
    lo, hi = face.box(N3)
    small = device_face(face, random_series(rng, lo, hi, 2), lo, hi, 2, keep)
    want, got = both(dev, [small], 2)
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (got, want)


def test_no_alloc_flux_refuses_more_terms_than_its_scratch(dev):
    """2 x 1024 frequencies x 40 x 40 values = 3.3 M terms > 2^21: refused on the host before
    any launch (the allocating call takes it)."""
    from photonoxide_amd import _native as N
    print("case allocation-free flux beyond its scratch", flush=True)
    D = N.FluxFace()
    D.axis, D.plane, D.nf = 2, 0, 1024
    D.from_[0], D.to[0], D.from_[1], D.to[1] = 0, 40, 0, 40
    D.lo[:] = (0, 0, 0)
    D.hi[:] = (40, 40, 2)
    D.sign, D.step_b, D.step_c = 1.0, 0.1, 0.1
    tiny = Buffer(16)
    for t in range(6):
        D.series[t] = tiny.address  # never read: the call is refused first
    out = Buffer(8 * 1024)
    arr = (N.FluxFace * 1)(D)
    assert N.lib.pxfdtd_flux_device_no_alloc(dev, arr, 1, out.address, None) == 1
    assert b"at most" in N.lib.pxfdtd_last_error()
