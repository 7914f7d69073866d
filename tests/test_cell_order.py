"""The GPU library's internal cell numbering and the ray-list tile order
(csrc/frt_order.h), compiled for the host: a bijection onto 0..n^3-1 that
keeps every 4x4x4 tile contiguous, agrees with the row-major tile key, and
falls back to row-major when n is not a multiple of 4."""
import ctypes as C

import numpy as np
import pytest

from oracle import pyoracle


def idx(n, i, j, k):
    return pyoracle.load().orc_cell_index(n, i, j, k)


def ijk(n, c):
    out = (C.c_int * 3)()
    pyoracle.load().orc_cell_ijk(n, c, out)
    return tuple(out)


@pytest.mark.parametrize("n", [4, 8, 12, 16])
def test_tiled_numbering_is_a_bijection_with_contiguous_tiles(n):
    seen = np.zeros(n ** 3, dtype=bool)
    lib = pyoracle.load()
    for i in range(n):
        for j in range(n):
            for k in range(n):
                c = idx(n, i, j, k)
                assert 0 <= c < n ** 3 and not seen[c]
                seen[c] = True
                assert ijk(n, c) == (i, j, k)
                # the 64 cells of a tile share c >> 6, which is the row-major tile key
                assert c >> 6 == lib.orc_tile_key(n, (i * n + j) * n + k)
                assert c >> 6 == ((i // 4) * (n // 4) + j // 4) * (n // 4) + k // 4
    assert seen.all()


@pytest.mark.parametrize("n", [5, 6, 10])
def test_other_sizes_are_row_major(n):
    for c in range(0, n ** 3, 7):
        i, j, k = ijk(n, c)
        assert (i * n + j) * n + k == c
        assert idx(n, i, j, k) == c


def test_tile_key_of_row_major_cells():
    lib = pyoracle.load()
    n = 10  # tiles per side = 3 (the last tile is partial)
    for (i, j, k) in [(0, 0, 0), (3, 3, 3), (4, 0, 0), (9, 9, 9), (5, 8, 2)]:
        want = ((i // 4) * 3 + j // 4) * 3 + k // 4
        assert lib.orc_tile_key(n, (i * n + j) * n + k) == want
