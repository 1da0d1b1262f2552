"""The CPU baseline's kernel (oracle/cpu_baseline.c, test infrastructure): Malas et al.'s
diamonds as the reference's CPU path steps by them (src/fdtd/kernel/blocked.rs) give the same
bits as stepping the whole grid a step at a time, for any tiling and thread count, with every
face's CPML; and the tiles are Blocking::auto's (blocked.rs:76-114)."""
import ctypes as C

import pytest

import oracle_lib as O


@pytest.mark.parametrize("n, pml, steps, s, rows, threads", [
    ((48, 44, 20), 6, 7, 2, 6, 4),
    ((40, 30, 16), 4, 11, 3, 8, 8),
    ((64, 64, 24), 8, 5, 5, 12, 8),
    ((33, 47, 9), 5, 13, 1, 4, 3),
    ((50, 40, 20), 6, 10, 2, 6, 1),
    ((24, 70, 12), 3, 9, 4, 10, 6),
])
def test_diamonds_are_the_whole_grids_bits(n, pml, steps, s, rows, threads):
    assert O.lib.or_cpu_blocked_differs(*n, pml, steps, s, rows, threads) == 0


def test_tiles_are_blocking_autos():
    s, r = C.c_int(), C.c_int()
    O.lib.or_cpu_auto_blocking(384, 256, 192, 16, C.byref(s), C.byref(r))
    assert (s.value, r.value) == (5, 12)  # 12 x 12 planes of 384 x 48 B in flight: 2.65 MB
    O.lib.or_cpu_auto_blocking(64, 64, 64, 16, C.byref(s), C.byref(r))
    assert (s.value, r.value) == (0, 0)  # 12 MiB of fields, under 40 MiB: stepped whole
    O.lib.or_cpu_auto_blocking(2048, 256, 512, 16, C.byref(s), C.byref(r))
    assert (s.value, r.value) == (2, 6)  # wide rows: fewer steps keep a tile's planes in 4.5 MiB
    O.lib.or_cpu_auto_blocking(1024, 48, 256, 16, C.byref(s), C.byref(r))
    assert (s.value, r.value) == (0, 0)  # 48 rows: no tiling leaves each of 16 threads a tile
