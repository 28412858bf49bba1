"""Gate factories for rotation channels (crates/paulistrings-py/src/channel_spec.rs:300-327).

Each returns a one-channel :class:`~paulistrings_amd.circuits.Circuit` fragment that ``Circuit.append`` adds:
``rx(theta, q)``, ``ry``, ``rz``, ``rzz(theta, a, b)`` and ``pauli_rotation("ZZ", [a, b], theta)``, all
``exp(-i theta P / 2)``.
"""

from __future__ import annotations

from dataclasses import dataclass


@dataclass(frozen=True)
class Rotation:
    paulis: str
    qubits: tuple
    theta: float

    def apply_to(self, circuit) -> None:
        circuit.pauli_rotation(self.paulis, list(self.qubits), self.theta)


def pauli_rotation(paulis: str, qubits, theta: float) -> Rotation:
    return Rotation(paulis, tuple(int(q) for q in qubits), float(theta))


def rx(theta: float, qubit: int) -> Rotation:
    return Rotation("X", (int(qubit),), float(theta))


def ry(theta: float, qubit: int) -> Rotation:
    return Rotation("Y", (int(qubit),), float(theta))


def rz(theta: float, qubit: int) -> Rotation:
    return Rotation("Z", (int(qubit),), float(theta))


def rzz(theta: float, a: int, b: int) -> Rotation:
    return Rotation("ZZ", (int(a), int(b)), float(theta))
