"""Post-processing on the device (GPU): pxfdtd_flux_device, the flux of transform sums held in
device memory (FluxPlane::flux, src/fdtd/monitors.rs:318-350; Simulation::flux, :530-547).

Checked against the oracle's flux restatement (oracle/flux.py, pinned by the closed-box
balance) over the same random FP64 series, bit for bit: planes across each axis, windowed and
whole, a box's six faces summed with their signs, values that are not finite where a window
gives them no weight (the reference skips them), and the refusals. The host mirror's flux goes
through the same call (tests/test_monitors.py checks it against the oracle in a run).
"""
import ctypes as C
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import flux as F  # noqa: E402
from devmem import Buffer  # noqa: E402

pytestmark = pytest.mark.gpu

N3 = (10, 9, 8)
STEPS = (0.05, 0.04, 0.03)


@pytest.fixture(scope="module")
def dev():
    from photonoxide_amd import _native as N
    d = C.c_void_p()
    N.check(N.lib.pxfdtd_device_open(0, C.byref(d)))
    yield d
    N.lib.pxfdtd_device_close(d)


def random_series(rng, lo, hi, nf):
    """Six series (E x, y, z, H x, y, z) over the nodes lo..hi inclusive, [f][k][j][i]."""
    shape = (nf, hi[2] - lo[2] + 1, hi[1] - lo[1] + 1, hi[0] - lo[0] + 1)
    return [rng.standard_normal(shape) + 1j * rng.standard_normal(shape) for _ in range(6)]


def value_fn(series, lo):
    return lambda fd, c, at, f: series[fd * 3 + c][f, at[2] - lo[2], at[1] - lo[1], at[0] - lo[0]]


def device_face(face, series, lo, hi, nf, keep):
    """The face as the C ABI takes it, its series copied to device memory."""
    from photonoxide_amd import _native as N
    D = N.FluxFace()
    D.axis, D.plane = face.axis, face.plane
    for w in range(2):
        r = face.range(N3, w)
        D.from_[w], D.to[w] = r.start, r.stop
        if face.window[w] is not None:
            D.windowed[w] = 1
            D.wlo[w], D.whi[w] = face.window[w]
    for a in range(3):
        D.lo[a], D.hi[a] = lo[a], hi[a] + 1
    D.nf = nf
    D.sign = face.sign
    b, c = face.others()
    D.step_b, D.step_c = STEPS[b], STEPS[c]
    for t in range(6):
        d = Buffer.of(np.ascontiguousarray(series[t]).view(np.float64).ravel())
        keep.append(d)
        D.series[t] = d.address
    return D


def device_flux(dev, faces, nf):
    from photonoxide_amd import _native as N
    arr = (N.FluxFace * len(faces))(*faces)
    out = Buffer(8 * nf)
    N.check(N.lib.pxfdtd_flux_device(dev, arr, len(faces), out.address, None))
    return out.read(np.float64, nf)


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("windowed", [False, True])
def test_a_plane_is_the_oracles_flux_bit_for_bit(dev, axis, windowed):
    print(f"case flux plane axis {axis} windowed {windowed}", flush=True)
    rng = np.random.default_rng(40 + 2 * axis + windowed)
    b, c = (axis + 1) % 3, (axis + 2) % 3
    window = ((1, N3[b] - 3), (2, N3[c] - 2)) if windowed else (None, None)
    face = F.Face(axis, 3, window)
    lo, hi = face.box(N3)
    nf = 4
    series = random_series(rng, lo, hi, nf)
    keep = []
    got = device_flux(dev, [device_face(face, series, lo, hi, nf, keep)], nf)
    want = np.array([F.total_flux([face], N3, STEPS, [value_fn(series, lo)], f) for f in range(nf)])
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (got, want)


def test_a_boxs_six_faces_sum_with_their_signs_bit_for_bit(dev):
    print("case flux box six faces", flush=True)
    rng = np.random.default_rng(7)
    faces = F.box_faces(((2, 6), (1, 6), (2, 5)))
    nf = 3
    keep, dfaces, values = [], [], []
    for face in faces:
        lo, hi = face.box(N3)
        series = random_series(rng, lo, hi, nf)
        dfaces.append(device_face(face, series, lo, hi, nf, keep))
        values.append(value_fn(series, lo))
    got = device_flux(dev, dfaces, nf)
    want = np.array([F.total_flux(faces, N3, STEPS, values, f) for f in range(nf)])
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (got, want)


def test_values_without_weight_are_skipped_even_when_not_finite(dev):
    """At the window's first node along b, E_c's weight there is 0 (FluxPlane::weight on a
    node: lo < m <= hi): the reference never reads it, so a NaN there changes nothing."""
    print("case flux NaN where the weight is 0", flush=True)
    rng = np.random.default_rng(11)
    face = F.Face(2, 4, ((2, 7), (1, 6)))
    lo, hi = face.box(N3)
    nf = 2
    series = random_series(rng, lo, hi, nf)
    want = np.array([F.total_flux([face], N3, STEPS, [value_fn(series, lo)], f) for f in range(nf)])
    series[1][:, :, :, 2 - lo[0]] = np.nan  # E_y (c = y for a z plane) at x = 2 = wlo along b
    keep = []
    got = device_flux(dev, [device_face(face, series, lo, hi, nf, keep)], nf)
    assert np.all(np.isfinite(got))
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (got, want)


def test_refusals(dev):
    from photonoxide_amd import _native as N
    print("case flux refusals", flush=True)
    rng = np.random.default_rng(3)
    face = F.Face(0, 4, ((1, 6), None))
    lo, hi = face.box(N3)
    series = random_series(rng, lo, hi, 2)
    keep = []
    good = device_face(face, series, lo, hi, 2, keep)
    out = Buffer(16)

    def call(faces):
        arr = (N.FluxFace * max(1, len(faces)))(*faces)
        return N.lib.pxfdtd_flux_device(dev, arr, len(faces), out.address, None)

    assert call([]) == 1
    assert call([good] * 7) == 1
    for change in (lambda d: setattr(d, "nf", 3), lambda d: setattr(d, "plane", hi[0]),
                   lambda d: d.to.__setitem__(0, 6), lambda d: d.to.__setitem__(1, N3[2] + 1),
                   lambda d: setattr(d, "axis", 3)):
        bad = N.FluxFace.from_buffer_copy(good)
        change(bad)
        assert call([good, bad]) == 1
    bad = N.FluxFace.from_buffer_copy(good)
    bad.series[2] = None  # E_z: E along c for an x plane
    assert call([bad]) == 1
    assert call([good]) == 0
