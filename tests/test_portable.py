"""The host layer's portable maths (photonoxide_amd/csrc/portable.cpp, through pxfdtd_portable)
held to the reference's own bit hashes over its argument grid (src/portable/tests.rs:400-548):
every argument of the grid (special values, subnormals, 600 seeded ones over five scales) and
every pair (grid[i], grid[n - 1 - i]) gives the reference's bits, so each hash below equals the
reference's GRID_HASHES entry. These are the functions the host mirror computes beams
(src/fdtd/sources.rs:1101,1182,1190), port modes (src/fdfd/ports.rs:272-287), the mode
monitor's frequency (src/fdtd/monitors.rs:563) and resonances (src/fdtd/harmonic.rs:102-163)
with, as well as its waveforms and CPML coefficients."""
import ctypes as C
import math

import pytest

from test_oracle import GRID_HASHES, grid_fold, reference_grid

# src/portable/tests.rs:521-547, beside the ones test_oracle already names
MORE_HASHES = {"ln_1p": 0x8DF40320AC11090D, "asin": 0x691EE862FFA6959C, "atan": 0x399728F358C0C9C1,
               "acosh": 0xDAB92EE03691324E, "atan2": 0x5401923556E0EE65, "hypot": 0xAC2E5DE82B7D5C15,
               "c_ln": 0x591636A61903229A, "c_sqrt": 0xC84405D54E382130, "c_asin": 0x76F21B923D0BFC4F,
               "c_polar": 0xB2A10E13B32679F9}
REALS = ["exp", "ln", "ln_1p", "asin", "atan", "acosh"]
TWOS = ["atan2", "hypot", "powf", "powi"]
COMPLEX = ["c_sqrt", "c_asin", "c_ln", "c_polar"]


def evaluate(name, a, b=0.0):
    from photonoxide_amd import _native as N
    re, im = C.c_double(), C.c_double()
    N.check(N.lib.pxfdtd_portable(name.encode(), a, b, C.byref(re), C.byref(im)))
    return re.value, im.value


@pytest.mark.parametrize("name", REALS + TWOS + COMPLEX)
def test_the_host_layers_portable_function_has_the_references_bits_over_its_grid(name):
    want = {**GRID_HASHES, **MORE_HASHES}[name]
    xs = reference_grid()
    pairs = list(zip(xs, reversed(xs)))
    h = 0
    if name in REALS:
        for x in xs:
            h = grid_fold(h, evaluate(name, x)[0])
    elif name in TWOS:
        for a, b in pairs:
            if name == "powi":  # (n * 7.0) as i32 % 64: Rust's saturating cast, a remainder of its sign
                n = 0 if math.isnan(b) else int(max(min(b * 7.0, 2**31 - 1), -2**31))
                b = float(int(math.copysign(abs(n) % 64, n)))
            h = grid_fold(h, evaluate(name, a, b)[0])
    else:
        for a, b in pairs:
            re, im = evaluate(name, a, b)
            h = grid_fold(grid_fold(h, re), im)
    assert h == want, (name, hex(h), hex(want))


def test_an_unknown_name_is_refused():
    from photonoxide_amd._native import NativeError
    with pytest.raises(NativeError, match="no portable function"):
        evaluate("c_tanh", 0.5, 0.5)
