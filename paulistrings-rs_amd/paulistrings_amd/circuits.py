"""Circuits of Pauli rotations and the lattices the benchmark workloads run on.

A :class:`Circuit` is an ordered list of Pauli rotations ``exp(-i theta P / 2)``,
one channel per gate, mirroring the reference's ``Circuit`` with its ``rx``/``rz``/
``pauli_rotation`` spellings (crates/paulistrings-py/src/channel_spec.rs:300-327,
python/paulistrings/tests/test_pauli_rotation.py:1-20 for the convention).  The
channel decomposition is the truncation schedule: the engine truncates after every
channel (examples/common/circuits.py:6-13).

Workload builders:

* :func:`heavy_hex_127_edges` -- the 127-qubit Eagle heavy-hex coupling graph,
  generated from its row/bridge structure (the reference reads the same 144 edges
  from examples/data/heavy_hex_127.edges).
* :func:`heavy_hex_kicked_ising` -- examples/common/circuits.py:199-283 (greedy
  edge colouring in sorted order, :158-189).
* :func:`tfim_square_step` -- the periodic square-lattice transverse-field Ising
  Trotter step of docs/figures/ising-quench-convergence/quench_convergence.py:67-82.
"""

from __future__ import annotations

import math
from collections import defaultdict
from typing import Iterable, Sequence

import numpy as np

#: Layout of one channel as the C ABI takes it (``pp_channel`` in include/pp_hip.h).
CHANNEL_DTYPE = np.dtype([("gx", "<u8", (4,)), ("gz", "<u8", (4,)), ("theta", "<f8")])

_PAULI_BITS = {"I": (0, 0), "X": (1, 0), "Y": (1, 1), "Z": (0, 1)}


def width_for(num_qubits: int) -> int:
    """Words per x/z half: the smallest supported W with 64 W >= num_qubits."""
    if num_qubits < 0:
        raise ValueError(f"num_qubits must be non-negative, got {num_qubits}")
    if num_qubits <= 64:
        return 1
    if num_qubits <= 128:
        return 2
    raise ValueError(f"num_qubits={num_qubits}: at most 128 qubits are supported")


class Circuit:
    """An ordered list of Pauli rotations on ``num_qubits`` qubits."""

    def __init__(self, num_qubits: int):
        self.num_qubits = int(num_qubits)
        self.width = width_for(self.num_qubits)
        self._rows: list[tuple[tuple[int, ...], tuple[int, ...], float]] = []

    def __len__(self) -> int:
        return len(self._rows)

    def _check(self, q: int) -> None:
        if not 0 <= q < self.num_qubits:
            raise ValueError(f"qubit {q} is outside the {self.num_qubits}-qubit register")

    def pauli_rotation(self, paulis: str, qubits: Sequence[int], theta: float) -> "Circuit":
        """Push ``exp(-i theta P / 2)`` with ``P`` = ``paulis[j]`` on ``qubits[j]`` (positional)."""
        if len(paulis) != len(qubits):
            raise ValueError(f"{len(paulis)} Paulis for {len(qubits)} qubits")
        if len(set(qubits)) != len(qubits):
            raise ValueError(f"repeated qubit in {list(qubits)}")
        gx = [0, 0, 0, 0]
        gz = [0, 0, 0, 0]
        for ch, q in zip(paulis.upper(), qubits):
            if ch not in _PAULI_BITS:
                raise ValueError(f"unknown Pauli {ch!r}")
            self._check(int(q))
            bx, bz = _PAULI_BITS[ch]
            gx[q // 64] |= bx << (q % 64)
            gz[q // 64] |= bz << (q % 64)
        self._rows.append((tuple(gx), tuple(gz), float(theta)))
        return self

    def rx(self, theta: float, qubit: int) -> "Circuit":
        return self.pauli_rotation("X", [qubit], theta)

    def ry(self, theta: float, qubit: int) -> "Circuit":
        return self.pauli_rotation("Y", [qubit], theta)

    def rz(self, theta: float, qubit: int) -> "Circuit":
        return self.pauli_rotation("Z", [qubit], theta)

    def rzz(self, theta: float, a: int, b: int) -> "Circuit":
        return self.pauli_rotation("ZZ", [a, b], theta)

    def append(self, gate) -> "Circuit":
        """Append a gate from :mod:`paulistrings_amd.gates`."""
        gate.apply_to(self)
        return self

    def extend(self, other: "Circuit") -> "Circuit":
        if other.num_qubits != self.num_qubits:
            raise ValueError("circuits act on different registers")
        self._rows.extend(other._rows)
        return self

    def prefix(self, k: int) -> "Circuit":
        """The first ``k`` channels as a circuit of their own."""
        out = Circuit(self.num_qubits)
        out._rows = self._rows[: max(0, int(k))]
        return out

    def channels(self) -> np.ndarray:
        """The channels as a ``CHANNEL_DTYPE`` array, in circuit order."""
        out = np.zeros(len(self._rows), dtype=CHANNEL_DTYPE)
        for i, (gx, gz, theta) in enumerate(self._rows):
            out[i]["gx"] = gx
            out[i]["gz"] = gz
            out[i]["theta"] = theta
        return out

    def __repr__(self) -> str:
        return f"Circuit(num_qubits={self.num_qubits}, channels={len(self)})"


# ---------------------------------------------------------------------------------------------
# Lattices
# ---------------------------------------------------------------------------------------------


def heavy_hex_127_edges() -> list[tuple[int, int]]:
    """The 144 edges ``(lo, hi)`` of the 127-qubit Eagle heavy-hex graph, sorted.

    Seven rows of up to 15 columns: row 0 holds columns 0..13, row 6 columns 1..14,
    every other row all 15.  Four bridge qubits sit between consecutive rows, on
    columns 0, 4, 8, 12 below an even row and 2, 6, 10, 14 below an odd row.
    Qubits are numbered row by row, each row followed by its bridges.
    """
    rows, cols = 7, 15
    number: dict[tuple[str, int, int], int] = {}
    q = 0
    for r in range(rows):
        lo = 1 if r == rows - 1 else 0
        hi = cols - 1 if r == 0 else cols
        for c in range(lo, hi):
            number[("row", r, c)] = q
            q += 1
        if r + 1 < rows:
            for c in (range(0, cols, 4) if r % 2 == 0 else range(2, cols, 4)):
                number[("bridge", r, c)] = q
                q += 1
    edges = []
    for (kind, r, c), a in number.items():
        if kind == "row":
            b = number.get(("row", r, c + 1))
            if b is not None:
                edges.append((a, b))
        else:
            edges.append((number[("row", r, c)], a))
            edges.append((a, number[("row", r + 1, c)]))
    return sorted((min(a, b), max(a, b)) for a, b in edges)


def heavy_hex_edge_coloring(edges: Iterable[tuple[int, int]]) -> list[list[tuple[int, int]]]:
    """Greedy first-fit proper edge colouring in sorted edge order (circuits.py:158-189)."""
    used: dict[int, set[int]] = defaultdict(set)
    classes: list[list[tuple[int, int]]] = []
    for a, b in sorted(edges):
        color = 0
        while color in used[a] or color in used[b]:
            color += 1
        while len(classes) <= color:
            classes.append([])
        classes[color].append((a, b))
        used[a].add(color)
        used[b].add(color)
    return classes


def heavy_hex_kicked_ising(
    trotter_steps: int,
    theta_h: float,
    theta_zz: float = -math.pi / 2,
    *,
    order: str = "x-then-zz",
    num_qubits: int = 127,
) -> Circuit:
    """Kicked Ising on the heavy-hex graph: per step RX(theta_h) on every qubit and
    RZZ(theta_zz) on every edge in colour order (examples/common/circuits.py:199-283)."""
    if order not in ("x-then-zz", "zz-then-x"):
        raise ValueError(f"order must be 'x-then-zz' or 'zz-then-x', got {order!r}")
    lattice = [(a, b) for a, b in heavy_hex_127_edges() if a < num_qubits and b < num_qubits]
    zz_order = [e for group in heavy_hex_edge_coloring(lattice) for e in group]
    c = Circuit(num_qubits)
    for _ in range(trotter_steps):
        if order == "x-then-zz":
            for q in range(num_qubits):
                c.rx(theta_h, q)
            for a, b in zz_order:
                c.pauli_rotation("ZZ", [a, b], theta_zz)
        else:
            for a, b in zz_order:
                c.pauli_rotation("ZZ", [a, b], theta_zz)
            for q in range(num_qubits):
                c.rx(theta_h, q)
    return c


def tfim_square_step(lx: int, ly: int, dt: float, j: float, h: float) -> Circuit:
    """One Trotter step of the periodic ``lx`` x ``ly`` transverse-field Ising model:
    for every site (x, y) the bonds to (x+1, y) and (x, y+1) as RZZ(2 J dt), then RX(2 h dt)
    on every site (quench_convergence.py:67-82)."""
    n = lx * ly

    def idx(x: int, y: int) -> int:
        return (y % ly) * lx + (x % lx)

    c = Circuit(n)
    for y in range(ly):
        for x in range(lx):
            i = idx(x, y)
            for nx, ny in ((x + 1, y), (x, y + 1)):
                c.pauli_rotation("ZZ", [i, idx(nx, ny)], 2 * j * dt)
    for site in range(n):
        c.rx(2 * h * dt, site)
    return c


def repeat(step: Circuit, times: int) -> Circuit:
    out = Circuit(step.num_qubits)
    for _ in range(times):
        out.extend(step)
    return out


def batched_instances(seed: int = 0xC0FFEE, count: int = 64, depths: tuple[int, int] = (5, 30),
                      thresholds: tuple[float, float] = (1e-6, 1e-3)):
    """The batched workload: ``count`` independent propagations with qubit counts uniform in {16..64}, depths uniform
    in ``depths`` (inclusive) brick layers (RZZ on the even then the odd bonds of a chain, alternating by layer, then
    RX or RZ by a coin flip on every qubit, angles uniform in [0, 2 pi)), observables of 1..8 random Paulis of weight
    at most 3 with N(0, 1) coefficients, and coefficient thresholds log-uniform in ``thresholds``.  Drawn with numpy's
    default generator from ``seed``.  Returns a list of (num_qubits, circuit, (x, z, coeff), threshold)."""
    rng = np.random.default_rng(seed)
    out = []
    lo_eps, hi_eps = math.log10(thresholds[0]), math.log10(thresholds[1])
    for _ in range(count):
        n = int(rng.integers(16, 65))
        depth = int(rng.integers(depths[0], depths[1] + 1))
        c = Circuit(n)
        for layer in range(depth):
            for i in range(layer % 2, n - 1, 2):
                c.rzz(float(rng.uniform(0.0, 2 * math.pi)), i, i + 1)
            for q in range(n):
                theta = float(rng.uniform(0.0, 2 * math.pi))
                if rng.random() < 0.5:
                    c.rx(theta, q)
                else:
                    c.rz(theta, q)
        terms = int(rng.integers(1, 9))
        W = width_for(n)
        x = np.zeros((terms, W), np.uint64)
        z = np.zeros((terms, W), np.uint64)
        for t in range(terms):
            k = int(rng.integers(1, 4))
            for q in rng.choice(n, size=k, replace=False):
                p = int(rng.integers(1, 4))  # X, Y, Z
                if p in (1, 2):
                    x[t, q // 64] |= np.uint64(1) << np.uint64(q % 64)
                if p in (2, 3):
                    z[t, q // 64] |= np.uint64(1) << np.uint64(q % 64)
        coeff = rng.standard_normal(terms)
        eps = float(10.0 ** rng.uniform(lo_eps, hi_eps))
        out.append((n, c, (x, z, coeff), eps))
    return out
