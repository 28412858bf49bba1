"""The prepared tables of general channels (paulistrings_amd.tables) against known transfer matrices and an
independent construction (oracle/ptm.py): Hadamard, CNOT, amplitude damping (non-unital), random unitaries; the
entry grouping by local delta; rounding to float32.  CPU only."""

from __future__ import annotations

import math

import numpy as np
import pytest

import ptm
from paulistrings_amd import _native, tables

H = np.array([[1, 1], [1, -1]]) / math.sqrt(2)
CNOT = np.array([[1, 0, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0], [0, 1, 0, 0]])  # control = support qubit 0 (low bit)
I, X, Z, Y = 0, 1, 2, 3


def entries(p):
    return [(int(p.deltas[e].local_delta), [float(v) for v in p.deltas[e].amplitude]) for e in range(p.n_deltas)]


def test_hadamard_swaps_x_and_z_and_negates_y():
    t = tables.transfer_matrix([H], 1)
    want = np.zeros((4, 4))
    want[I, I], want[X, Z], want[Z, X], want[Y, Y] = 1, 1, 1, -1
    assert np.allclose(t, want, atol=1e-15)
    p = tables.unitary_table(H, [5])
    es = dict(entries(p))
    assert sorted(es) == [0, 3]  # I and Y keep their key (delta 0); X <-> Z is delta X ^ Z = 3
    assert np.allclose([es[0][I], es[0][Y], es[3][X], es[3][Z]], [1.0, -1.0, 1.0, 1.0], atol=1e-15)


def test_cnot_is_a_signed_permutation_of_the_sixteen_patterns():
    t = tables.transfer_matrix([CNOT], 2)
    assert np.allclose(np.abs(t).sum(axis=1), 1.0) and np.allclose(np.abs(t).sum(axis=0), 1.0)
    # X on the control spreads to X X; Z on the target spreads to Z Z
    assert math.isclose(t[X, X | X << 2], 1.0, abs_tol=1e-15)
    assert math.isclose(t[Z << 2, Z | Z << 2], 1.0, abs_tol=1e-15)


def test_amplitude_damping_is_non_unital():
    g = 0.3
    k0 = np.array([[1, 0], [0, math.sqrt(1 - g)]])
    k1 = np.array([[0, math.sqrt(g)], [0, 0]])
    t = tables.transfer_matrix([k0, k1], 1)
    want = np.zeros((4, 4))
    want[I, I], want[I, Z] = 1.0, g  # E(I) = I + g Z
    want[X, X] = want[Y, Y] = math.sqrt(1 - g)
    want[Z, Z] = 1 - g
    assert np.allclose(t, want, atol=1e-15)


@pytest.mark.parametrize("k", [1, 2])
def test_random_unitaries_match_the_independent_construction(k):
    rng = np.random.default_rng(11 + k)
    for _ in range(5):
        m = rng.standard_normal((2 ** k, 2 ** k)) + 1j * rng.standard_normal((2 ** k, 2 ** k))
        u, _ = np.linalg.qr(m)
        a = tables.transfer_matrix([u], k)
        b = ptm.conjugation_ptm(u, k)[: 4 ** k, : 4 ** k]
        assert np.allclose(a, b, atol=1e-13)
        p = tables.unitary_table(u, list(range(3, 3 + k)))
        got = entries(p)
        want = ptm.entries_of(ptm.conjugation_ptm(u, k), k, clean=1e-14)
        assert [d for d, _ in got] == [d for d, _ in want]
        for (_, ga), (_, wa) in zip(got, want):
            assert np.allclose(ga, wa, atol=1e-13)


def test_float32_tables_hold_float32_values():
    rng = np.random.default_rng(3)
    u, _ = np.linalg.qr(rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4)))
    p = tables.unitary_table(u, [0, 9], dtype=np.float32)
    for _, amp in entries(p):
        assert all(float(np.float32(a)) == a for a in amp)


def test_rotation_table_comes_from_the_library():
    p = tables.rotation_table(40, "ZZ", [3, 7], 0.4)
    assert p.kind == _native.PREPARED_LOCAL and p.k == 2 and p.n_deltas == 2
    assert int(p.deltas[0].local_delta) == 0
