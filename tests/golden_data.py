"""Loaders for the committed golden fixtures (tests/golden/*.json) and small test helpers."""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden"


def records(engine: str = "paulistrings"):
    recs = json.loads((GOLDEN / "kicked_ising_records.json").read_text())
    return [r for r in recs if r["engine"] == engine]


def observable(name: str, num_qubits: int = 127):
    """(x, z, coeff) of a Kim et al. observable; coefficient 1 (examples/common/observables.py:145-177)."""
    obs = json.loads((GOLDEN / "kim2023_observables.json").read_text())
    key = {"z62": "weight_1_z62"}.get(name, name)
    support = {int(q): p for q, p in obs[key]["support"].items()}
    W = 1 if num_qubits <= 64 else 2
    x = np.zeros((1, W), np.uint64)
    z = np.zeros((1, W), np.uint64)
    for q, p in support.items():
        bit = np.uint64(1) << np.uint64(q % 64)
        if p in "XY":
            x[0, q // 64] |= bit
        if p in "ZY":
            z[0, q // 64] |= bit
    return x, z, np.array([1.0])


def heavy_hex_edges():
    return [tuple(e) for e in json.loads((GOLDEN / "heavy_hex_127_edges.json").read_text())]
