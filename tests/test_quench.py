"""Host logic of the learned-truncation experiment (paulistrings_amd/quench.py): the 6 x 6 training terms are moved to
the same sites of the 8 x 8 lattice so that the network's 8 x 8 row and column groups count the 6 x 6 rows and
columns."""

from __future__ import annotations

import numpy as np

from paulistrings_amd.learned import lattice_groups
from paulistrings_amd.quench import embed_in_8x8


def popcount(v: int) -> int:
    return bin(int(v)).count("1")


def test_embedding_keeps_rows_columns_and_weight():
    rng = np.random.default_rng(3)
    words = rng.integers(0, 1 << 36, size=200, dtype=np.uint64)
    out = embed_in_8x8(words, 6, 6)
    g8 = lattice_groups(8, 8)
    g6 = lattice_groups(6, 6)
    for w, e in zip(words, out):
        assert popcount(w) == popcount(e)
        for r in range(6):
            assert popcount(int(w) & int(g6[r])) == popcount(int(e) & int(g8[r]))  # row r
            assert popcount(int(w) & int(g6[8 + r])) == popcount(int(e) & int(g8[8 + r]))  # column r
        for r in (6, 7):  # the rows and columns beyond the 6 x 6 lattice stay empty
            assert int(e) & int(g8[r]) == 0 and int(e) & int(g8[8 + r]) == 0


def test_embedding_of_single_sites():
    for y in range(6):
        for x in range(6):
            e = embed_in_8x8(np.array([1 << (y * 6 + x)], np.uint64), 6, 6)[0]
            assert int(e) == 1 << (y * 8 + x)
