"""The resident layout's invariants (pp_sum_check, the reference's debug-build k_check_invariants) after every kind of
structural change: upload, out-of-place and in-place rotation layers (with queued and moved pairs), refinements,
TopN and ApproxTopN passes, permutation runs, class passes (rotation runs and general tables), a device copy.  Every
count must be zero.  A propagation with PP_CHECK_INVARIANTS=1 checks after each of its runs.  Each case prints its
name before it touches the GPU."""

from __future__ import annotations

import ctypes
import math
import sys

import numpy as np
import pytest

from paulistrings_amd import Circuit, PauliSum, _native, truncation
from test_gpu_parity import product_policy, random_circuit, random_sum

pytestmark = pytest.mark.gpu


def announce(msg: str) -> None:
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def sound(g):
    r = g.check_layout()
    bad = {k: v for k, v in r.items() if v and k != "unscanned"}
    assert not bad, bad


def test_layout_invariants_through_every_structural_change(monkeypatch):
    announce("layout invariants: upload, layers in and out of place, TopN, runs, class passes, copy")
    rng = np.random.default_rng(2024)
    n = 127
    x, z, c = random_sum(rng, 1 << 15, n, 0.1, np.float64)
    g = PauliSum(n, x, z, c).to_device(0)
    try:
        sound(g)
        circ = random_circuit(rng, n, 40, max_support=3)
        g.propagate(circ, truncation.coeff(1e-5), "forward")
        sound(g)
        monkeypatch.setenv("PP_IN_PLACE_MIN_ROWS", "0")  # in place, with refinements and moved pairs
        grow = Circuit(n)
        for q in range(0, 24, 3):  # about 1.2x a layer: ~16x over the circuit
            grow.rx(0.7, q)
            grow.pauli_rotation("ZZ", [q, q + 1], 0.5)
        g.propagate(grow, truncation.coeff(1e-4), "forward")
        assert len(g) > 4 * len(x)
        sound(g)
        monkeypatch.delenv("PP_IN_PLACE_MIN_ROWS")
        g.propagate(grow, truncation.coeff(1e-6) & truncation.topn(20000), "heisenberg")
        sound(g)
        g.propagate(grow, truncation.coeff(1e-6) & truncation.approx_topn(15000), "forward")
        sound(g)
        cliff = Circuit(n)
        for q in range(0, n - 1, 4):
            cliff.pauli_rotation("ZZ", [q, q + 1], -math.pi / 2)
        g.propagate(cliff, truncation.coeff(1e-6), "forward")  # a permutation run
        sound(g)
        runs = Circuit(n)
        for q in range(8):
            runs.rx(0.3 + 0.1 * q, 100 + q)  # rotation runs through the class pass
        g.propagate(runs, truncation.coeff(1e-6), "forward")
        sound(g)
        h = g.clone()
        sound(h)
        h.free()
        monkeypatch.setenv("PP_CHECK_INVARIANTS", "1")
        g.propagate(random_circuit(rng, n, 30, max_support=2), product_policy(eps=1e-5, topn=30000), "forward")
        sound(g)
    finally:
        g.free()
