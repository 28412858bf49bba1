"""Prepared local tables of channels on one or two qubits (the reference's ``Prepared::Local`` / ``LocalPtm``,
crates/paulistrings/src/channel/prepared.rs:35-46, :79-86), for ``GpuPauliSum.apply_table`` and ``apply_run``.

A table maps a term whose pattern on the support is s (per support qubit x | z << 1: I = 0, X = 1, Z = 2, Y = 3, the
first support qubit in the low bits; channel/ptm.rs:19-24) to the terms with pattern s ^ d, one entry per local delta
d, with amplitude ``amp[d][s] = T[s][s ^ d]`` where ``T[s][t] = Tr(P_t E(P_s)) / 2^k`` is the channel's Pauli transfer
matrix in the reference's orientation (a row per input pattern; channel/ptm.rs, ``Channel::ptm``).  For a unitary
``E(P) = U P U^dag`` (channel/ptm/unitary.rs); for Kraus operators ``E(P) = sum_i K_i P K_i^dag``.

Amplitudes are rounded to the coefficient type, and those at most ``clean`` in magnitude are dropped (the
reference's tables hold exact zeros where the transfer matrix does).
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _native

_SINGLE = {
    0: np.eye(2, dtype=complex),
    1: np.array([[0, 1], [1, 0]], dtype=complex),
    2: np.array([[1, 0], [0, -1]], dtype=complex),
    3: np.array([[0, -1j], [1j, 0]], dtype=complex),
}


def _pauli(pattern: int, k: int) -> np.ndarray:
    m = np.eye(1, dtype=complex)
    for j in range(k):
        m = np.kron(_SINGLE[(pattern >> (2 * j)) & 3], m)  # support qubit 0 is the rightmost factor
    return m


def transfer_matrix(kraus, k: int) -> np.ndarray:
    """T[s][t] = Tr(P_t E(P_s)) / 2^k for E(P) = sum_i K_i P K_i^dag on k <= 2 qubits; real for a channel that
    preserves Hermiticity."""
    if k not in (1, 2):
        raise ValueError("a local table acts on one or two qubits")
    dim = 1 << (2 * k)
    paulis = np.stack([_pauli(p, k) for p in range(dim)])
    images = sum(np.einsum("ab,sbc,dc->sad", kk, paulis, kk.conj()) for kk in kraus)
    t = np.einsum("tba,sab->st", paulis, images) / (1 << k)  # Tr(P_t X) = sum P_t[b, a] X[a, b]
    if np.max(np.abs(t.imag)) > 1e-12:
        raise ValueError("the channel's transfer matrix is not real")
    return np.ascontiguousarray(t.real)


def table(t: np.ndarray, qubits, dtype=np.float64, clean: float = 1e-14) -> _native.Prepared:
    """The prepared table of transfer matrix ``t`` (``t[s][t']``, 4 x 4 or 16 x 16) on ``qubits`` (ascending)."""
    qubits = [int(q) for q in qubits]
    k = len(qubits)
    dim = 1 << (2 * k)
    if t.shape != (dim, dim) or sorted(qubits) != qubits or len(set(qubits)) != k:
        raise ValueError("a 4^k x 4^k transfer matrix on ascending support qubits")
    real = np.float32 if np.dtype(dtype) == np.float32 else np.float64
    p = _native.Prepared()
    p.kind, p.k = _native.PREPARED_LOCAL, k
    for j, q in enumerate(qubits):
        p.qubits[j] = q
    n = 0
    for d in range(dim):
        amp = [0.0] * 16
        for s in range(dim):
            v = float(t[s, s ^ d])
            amp[s] = 0.0 if abs(v) <= clean else float(real(v))
        if any(a != 0.0 for a in amp):
            p.deltas[n].local_delta = d
            for s in range(16):
                p.deltas[n].amplitude[s] = amp[s]
            n += 1
    p.n_deltas = n
    return p


def unitary_table(u, qubits, dtype=np.float64) -> _native.Prepared:
    """A unitary gate on one or two qubits (the first support qubit is the low bit of ``u``'s index)."""
    u = np.asarray(u, dtype=complex)
    return table(transfer_matrix([u], len(qubits)), qubits, dtype)


def kraus_table(kraus, qubits, dtype=np.float64) -> _native.Prepared:
    """A channel given by Kraus operators on one or two qubits."""
    return table(transfer_matrix([np.asarray(k, dtype=complex) for k in kraus], len(qubits)), qubits, dtype)


def rotation_table(num_qubits: int, paulis: str, qubits, theta: float, dtype=np.float64,
                   adjoint: bool = False) -> _native.Prepared:
    """exp(-i theta P / 2) as the library prepares it (pp_prepare_channel: the sin/cos snap, a wide generator's
    rotation form)."""
    ch = _native.Channel()
    for q, p in zip(qubits, paulis):
        q = int(q)
        if p in "XY":
            ch.gx[q // 64] |= 1 << (q % 64)
        if p in "ZY":
            ch.gz[q // 64] |= 1 << (q % 64)
    ch.theta = float(theta)
    width = 1 if num_qubits <= 64 else 2
    kind = 1 if np.dtype(dtype) == np.float32 else 0
    out = _native.Prepared()
    _native.check(_native.lib().pp_prepare_channel(num_qubits, width, kind, ctypes.byref(ch), 1 if adjoint else 0,
                                                   ctypes.byref(out)))
    return out
