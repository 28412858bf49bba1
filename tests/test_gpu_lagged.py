"""Lagged lengths: pp_propagate queues small rotation layers ahead of their counts (each layer's length reaches the
host one layer late, the next layer sized from bounds).  The results must not depend on it: keys, coefficients,
canonical order, every layer's length and the reference's bucket count equal those of the synchronous loop
(PP_LAGGED_LENGTHS=0) and the oracle's.  The circuits cross what drains the queue: key-preserving layers (θ = π),
permutation runs, device refinements as the sum grows, output columns that must grow, and the in-place threshold.
Each case prints its name before it touches the GPU."""

from __future__ import annotations

import math
import sys

import numpy as np
import pytest

import pp_oracle as po
from paulistrings_amd import Circuit, PauliSum, circuits, truncation

pytestmark = pytest.mark.gpu


def announce(msg: str) -> None:
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def random_sum(rng, n, num_qubits, density):
    W = circuits.width_for(num_qubits)
    x = np.zeros((n, W), np.uint64)
    z = np.zeros((n, W), np.uint64)
    for w in range(W):
        live = min(64, num_qubits - 64 * w)
        weights = np.uint64(1) << np.arange(live, dtype=np.uint64)
        x[:, w] = ((rng.random((n, live)) < density) * weights).sum(axis=1, dtype=np.uint64)
        z[:, w] = ((rng.random((n, live)) < density) * weights).sum(axis=1, dtype=np.uint64)
    return x, z, rng.standard_normal(n)


def mixed_circuit(rng, num_qubits, layers):
    """Rotations on one or two qubits, with quarter turns (permutation steps, runs of them) and half turns
    (key-preserving) among generic angles."""
    c = Circuit(num_qubits)
    for _ in range(layers):
        k = int(rng.integers(1, 3))
        qs = [int(q) for q in rng.choice(num_qubits, size=k, replace=False)]
        ps = "".join(rng.choice(list("XYZ"), size=k))
        theta = float(rng.choice([rng.uniform(-math.pi, math.pi)] * 3 + [-math.pi / 2, math.pi / 2, math.pi]))
        c.pauli_rotation(ps, qs, theta)
    return c


def run(monkeypatch, lagged, num_qubits, x, z, c, circ, direction, policy):
    monkeypatch.setenv("PP_LAGGED_LENGTHS", "1" if lagged else "0")
    s = PauliSum(num_qubits, x, z, c).to_device(0)
    stats = s.propagate_with_stats(circ, policy=policy, direction=direction)
    bad = {k: v for k, v in s.check_layout().items() if v and k != "unscanned"}
    assert not bad, bad  # segments, rooms and lengths the queued layers laid out on the device
    bits = s.bucket_bits()
    h = s.to_host()
    s.free()
    return h.x_array(), h.z_array(), h.coefficients_array(), stats.terms_out, bits


def oracle(num_qubits, x, z, c, circ, direction, eps=None, max_weight=None):
    s = po.OracleSum(num_qubits, x, z, c)
    lens = s.propagate(circ.channels(), direction, po.coeff_policy(eps=eps, max_weight=max_weight))
    ox, oz, oc, _ = s.export()
    return ox, oz, oc, [int(v) for v in lens]


def assert_equal_runs(a, b):
    assert len(a[2]) == len(b[2])
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), "keys or canonical order differ"
    assert np.array_equal(a[2], b[2]), "coefficients differ"
    assert a[3] == b[3], "per-layer lengths differ"


def test_tfim_6x6_lagged_equals_synchronous_and_oracle(monkeypatch):
    announce("lagged lengths: 6x6 TFIM, 10 steps")
    circ = circuits.repeat(circuits.tfim_square_step(6, 6, 0.1, 1.0, 3.0), 10)
    n = 36
    x = np.array([[1 << q] for q in range(n)], np.uint64)
    z = np.zeros((n, 1), np.uint64)
    c = np.full(n, 1.0 / n)
    pol = truncation.coeff(1e-4) & truncation.weight(8)
    lag = run(monkeypatch, True, n, x, z, c, circ, "heisenberg", pol)
    sync = run(monkeypatch, False, n, x, z, c, circ, "heisenberg", pol)
    assert_equal_runs(lag, sync)
    assert lag[4][0] == sync[4][0]  # the reference's bucket count
    assert_equal_runs(lag, oracle(n, x, z, c, circ, "heisenberg", 1e-4, 8))


@pytest.mark.parametrize("num_qubits,seed,layers", [(24, 1, 60), (64, 2, 60), (127, 3, 80)])
def test_growth_across_refinements_and_column_growth(monkeypatch, num_qubits, seed, layers):
    """From 40 terms to 3·10^4 .. 2·10^5 with no truncation above 1e-9: the queue drains for every refinement and
    every growth of the output columns, and resumes after."""
    announce(f"lagged lengths: growth, {num_qubits} qubits, seed {seed}")
    rng = np.random.default_rng(seed)
    x, z, c = random_sum(rng, 40, num_qubits, 0.15)
    circ = mixed_circuit(rng, num_qubits, layers)
    pol = truncation.coeff(1e-9)
    lag = run(monkeypatch, True, num_qubits, x, z, c, circ, "forward", pol)
    sync = run(monkeypatch, False, num_qubits, x, z, c, circ, "forward", pol)
    assert_equal_runs(lag, sync)
    assert lag[4][0] == sync[4][0]
    assert max(lag[3]) > 20_000
    assert_equal_runs(lag, oracle(num_qubits, x, z, c, circ, "forward", eps=1e-9))


def test_crossing_the_in_place_threshold(monkeypatch):
    """A sum that grows past the in-place layer's threshold: the lagged queue stops below it and the in-place layer
    takes over above it, with the same result as the synchronous loop."""
    announce("lagged lengths: across the in-place threshold")
    monkeypatch.setenv("PP_IN_PLACE_MIN_ROWS", "20000")
    rng = np.random.default_rng(11)
    x, z, c = random_sum(rng, 3000, 60, 0.1)
    circ = Circuit(60)
    for q in range(0, 24):
        circ.rx(0.3 + 0.1 * q, q)
    pol = truncation.coeff(1e-7)
    lag = run(monkeypatch, True, 60, x, z, c, circ, "forward", pol)
    sync = run(monkeypatch, False, 60, x, z, c, circ, "forward", pol)
    assert_equal_runs(lag, sync)
    assert min(lag[3]) < 20_000 < max(lag[3])
    assert_equal_runs(lag, oracle(60, x, z, c, circ, "forward", eps=1e-7))


@pytest.mark.parametrize("direction", ["forward", "heisenberg"])
def test_kicked_ising_with_permutation_runs(monkeypatch, direction):
    """Heavy-hex kicked Ising steps (144 quarter-turn RZZ layers walked as one run, then RX layers queued)."""
    announce(f"lagged lengths: kicked Ising, {direction}")
    rng = np.random.default_rng(5)
    x, z, c = random_sum(rng, 300, 127, 0.02)
    circ = circuits.heavy_hex_kicked_ising(3, 0.8)
    pol = truncation.coeff(3e-3) & truncation.weight(12)
    lag = run(monkeypatch, True, 127, x, z, c, circ, direction, pol)
    sync = run(monkeypatch, False, 127, x, z, c, circ, direction, pol)
    assert_equal_runs(lag, sync)
    assert_equal_runs(lag, oracle(127, x, z, c, circ, direction, eps=3e-3, max_weight=12))


def test_empty_and_vanishing_sums(monkeypatch):
    """A sum that the threshold empties mid-circuit, and a propagation of no layers."""
    announce("lagged lengths: empty sums")
    x = np.array([[3]], np.uint64)
    z = np.array([[0]], np.uint64)
    c = np.array([1e-3])
    circ = Circuit(10).rz(0.5, 0).rz(0.5, 1).rzz(0.2, 0, 1).rx(0.1, 2)
    pol = truncation.coeff(1e-2)
    lag = run(monkeypatch, True, 10, x, z, c, circ, "forward", pol)
    sync = run(monkeypatch, False, 10, x, z, c, circ, "forward", pol)
    assert_equal_runs(lag, sync)
    assert lag[3][-1] == 0
    none = run(monkeypatch, True, 10, x, z, c, Circuit(10), "forward", pol)
    assert none[3] == [] and len(none[2]) == 1


def test_in_place_layer_after_a_lagged_propagation(monkeypatch):
    """The layout the one-launch small joins leave (segment starts and rooms, the end of the last segment) carries
    the in-place layer that follows, which writes new rows into the rooms and after that end."""
    announce("lagged lengths: an in-place layer after the queue")
    rng = np.random.default_rng(21)
    x, z, c = random_sum(rng, 500, 64, 0.1)
    circ = Circuit(64)
    for q in range(10):
        circ.rx(0.4 + 0.05 * q, q)
    more = Circuit(64).rzz(0.9, 3, 40).rx(0.3, 11)
    pol = truncation.coeff(1e-8)
    monkeypatch.setenv("PP_LAGGED_LENGTHS", "1")
    s = PauliSum(64, x, z, c).to_device(0)
    s.propagate(circ, policy=pol, direction="forward")
    monkeypatch.setenv("PP_IN_PLACE_MIN_ROWS", "0")
    for ch in more.channels():
        s.apply_rotation(ch, policy=pol)
    bad = {k: v for k, v in s.check_layout().items() if v and k != "unscanned"}
    assert not bad, bad
    h = s.to_host()
    s.free()
    whole = Circuit(64)
    for q in range(10):
        whole.rx(0.4 + 0.05 * q, q)
    whole.rzz(0.9, 3, 40).rx(0.3, 11)
    o = oracle(64, x, z, c, whole, "forward", eps=1e-8)

    def by_key(kx, kz, kc):  # the same terms whatever the bucket count behind the canonical order
        order = np.lexsort((kz[:, 0], kx[:, 0]))
        return kx[order], kz[order], kc[order]

    got, want = by_key(h.x_array(), h.z_array(), h.coefficients_array()), by_key(o[0], o[1], o[2])
    assert all(np.array_equal(a, b) for a, b in zip(got, want))
