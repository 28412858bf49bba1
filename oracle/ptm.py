"""Test infrastructure, not product code: a naive local-table layer and Pauli-transfer matrices from matrices.

Only the tests load this module.  It restates, for one- and two-qubit tables, what the reference's differential
oracle does for any channel (`naive_apply_layer`, crates/paulistrings/src/test_support/oracle.rs:21-54): every term
through the channel, outputs accumulated per key, each sum kept iff nonzero and `keep_term`, in plain Python.  The
tables themselves come from 2^k x 2^k matrices with numpy, independently of the product's preparation:
table[s][t] = Tr(P_t U P_s U^dag) / 2^k for a unitary (Channel::apply maps P_s to U P_s U^dag, channel/rotation.rs:43-67),
with the reference's local Pauli index s = x | z << 1 per support qubit (I = 0, X = 1, Z = 2, Y = 3, the first
support qubit in the low bits; channel/ptm.rs:19-24).

Tables here are lists of (local_delta, amplitudes[16]) entries ascending by local_delta, which is the
DeltaEntry grouping of a table (channel/prepared.rs:35-46, push_deltas :257-281).
"""

from __future__ import annotations

import numpy as np

_I = np.eye(2, dtype=complex)
_X = np.array([[0, 1], [1, 0]], dtype=complex)
_Z = np.array([[1, 0], [0, -1]], dtype=complex)
_Y = np.array([[0, -1j], [1j, 0]], dtype=complex)
_ONE_QUBIT = {0: _I, 1: _X, 2: _Z, 3: _Y}  # x | z << 1


def pauli_matrix(pattern: int, k: int) -> np.ndarray:
    """The matrix of local pattern `pattern` on k qubits, support qubit 0 the rightmost tensor factor."""
    m = np.eye(1, dtype=complex)
    for j in range(k):
        m = np.kron(_ONE_QUBIT[(pattern >> (2 * j)) & 3], m)
    return m


def conjugation_ptm(u: np.ndarray, k: int) -> np.ndarray:
    """table[s][t] = Tr(P_t U P_s U^dag) / 2^k, real for a unitary."""
    dim = 1 << (2 * k)
    t = np.zeros((16, 16))
    for s in range(dim):
        img = u @ pauli_matrix(s, k) @ u.conj().T
        for q in range(dim):
            v = np.trace(pauli_matrix(q, k) @ img) / (1 << k)
            assert abs(v.imag) < 1e-12
            t[s, q] = v.real
    return t


def entries_of(table: np.ndarray, k: int, clean: float = 0.0):
    """Group a PTM by key delta (push_deltas): entry d holds amplitude[s] = table[s][s ^ d]; |values| <= clean -> 0."""
    dim = 1 << (2 * k)
    out = []
    for d in range(dim):
        amp = [0.0] * 16
        for s in range(dim):
            v = float(table[s, s ^ d])
            amp[s] = 0.0 if abs(v) <= clean else v
        if any(a != 0.0 for a in amp):
            out.append((d, amp))
    return out


def pattern_of(x_words, z_words, qubits) -> int:
    s = 0
    for j, q in enumerate(qubits):
        s |= ((int(x_words[q // 64]) >> (q % 64)) & 1) << (2 * j)
        s |= ((int(z_words[q // 64]) >> (q % 64)) & 1) << (2 * j + 1)
    return s


def lift(delta: int, qubits, width: int):
    mx = [0] * width
    mz = [0] * width
    for j, q in enumerate(qubits):
        if (delta >> (2 * j)) & 1:
            mx[q // 64] |= 1 << (q % 64)
        if (delta >> (2 * j + 1)) & 1:
            mz[q // 64] |= 1 << (q % 64)
    return mx, mz


def naive_table_layer(terms: dict, qubits, entries, keep, real=float):
    """One layer of a local table over `terms` ({(x tuple, z tuple): c}): every term emits, per entry with a nonzero
    amplitude on its pattern, key ^ lift(d) with c * amplitude (DeltaEntry::emit, prepared.rs:55-71); outputs are
    summed in input order, then entry order; a sum is kept iff nonzero and keep(x, z, c).  `real` rounds every
    product and sum to the coefficient type (float32 sums: numpy.float32)."""
    width = len(next(iter(terms))[0]) if terms else 1
    acc: dict = {}
    for (x, z), c in terms.items():
        s = pattern_of(x, z, qubits)
        for d, amp in entries:
            a = amp[s]
            if a == 0.0:
                continue
            mx, mz = lift(d, qubits, width)
            key = (tuple(int(v) ^ m for v, m in zip(x, mx)), tuple(int(v) ^ m for v, m in zip(z, mz)))
            v = real(real(c) * real(a))
            acc[key] = real(acc[key] + v) if key in acc else v
    return {k: v for k, v in acc.items() if v != 0 and keep(k[0], k[1], v)}
