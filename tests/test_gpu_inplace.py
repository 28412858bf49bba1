"""The rotation layer in place (k_rotation_inplace, the default of pp_propagate, pp_apply_rotation and the layer seam
when the sum's segments have room) against the same layers into the spare columns (tuning join_in_place = 0) and
the oracle: keys, canonical order, coefficients bit-exact, every layer's length.  The cases grow sums through many
refinements (segments with no room after a split: queued pairs, moved to fresh segments), drop rows at a threshold
(holes filled by new rows), alternate with TopN passes (out of place while a selection is pending) and reset a
working sum from a resident input before each layer as the benchmark does.  Each case prints its name first."""

from __future__ import annotations

import math
import sys

import numpy as np
import pytest

import pp_oracle as po
from paulistrings_amd import Circuit, PauliSum, _native, circuits, truncation
from test_gpu_parity import oracle_policy, product_policy, random_circuit, random_sum

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def in_place_at_every_size(monkeypatch):
    """The library runs layers of fewer than 2^18 rows out of place (launch-bound there); these cases are smaller."""
    monkeypatch.setenv("PP_IN_PLACE_MIN_ROWS", "0")


def announce(msg: str) -> None:
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def with_in_place(ctx, on: bool):
    t = ctx.get_tuning()
    ctx.set_tuning(t.bucket_target, t.stage_rows, t.groups_per_cu, t.join_block, 1 if on else 0)
    return t


def propagate_both_ways(num_qubits, x, z, c, circ, direction, policy):
    host = PauliSum(num_qubits, x, z, c, dtype=c.dtype)
    out = []
    for on in (True, False):
        g = host.to_device(0)
        saved = with_in_place(g.ctx, on)
        try:
            stats = g.propagate_with_stats(circ, policy, direction)
            h = g.to_host()
            out.append((h._x, h._z, h._c, stats.terms_out, g.bucket_bits()[0]))
        finally:
            g.ctx.set_tuning(saved.bucket_target, saved.stage_rows, saved.groups_per_cu, saved.join_block,
                             saved.join_in_place)
            g.free()
    a, b = out
    assert a[3] == b[3], "per-layer lengths differ"
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), "keys or canonical order differ"
    assert np.array_equal(a[2], b[2]), "coefficients differ"
    assert a[4] == b[4]
    return a


@pytest.mark.parametrize("num_qubits", [20, 64, 127])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("direction", ["forward", "heisenberg"])
def test_in_place_equals_out_of_place_and_the_oracle(num_qubits, dtype, direction):
    announce(f"in place vs spare columns: random circuit n={num_qubits} {np.dtype(dtype).name} {direction}")
    rng = np.random.default_rng(9090 + num_qubits)
    x, z, c = random_sum(rng, 2000, num_qubits, 0.1, dtype)
    circ = random_circuit(rng, num_qubits, 80, max_support=3)
    for eps, topn in ((1e-4, None), (1e-3, 3000), (-1.0, None)):
        pol = product_policy(eps=eps if eps >= 0 else None, max_weight=num_qubits // 2, topn=topn)
        a = propagate_both_ways(num_qubits, x, z, c, circ, direction, pol)
        o = po.OracleSum(num_qubits, x, z, c)
        olens = o.propagate(circ.channels(), direction,
                            oracle_policy(eps=eps if eps >= 0 else None, max_weight=num_qubits // 2, topn=topn))
        ox, oz, oc, _ = o.export()
        assert [int(v) for v in olens] == a[3]
        assert np.array_equal(a[0], ox) and np.array_equal(a[1], oz) and np.array_equal(a[2], oc)


def test_in_place_growth_through_refinements():
    """A single term grown through many untruncated-ish layers: the device refines again and again (each split
    leaves one half without room), so pairs queue and move; the result equals the spare-column path and the
    oracle."""
    announce("in place: growth through refinements (queued and moved pairs)")
    n = 40
    circ = Circuit(n)
    rng = np.random.default_rng(5)
    for _ in range(125):  # random generators across the register: ~10^5 terms at the end
        a, b = sorted(int(v) for v in rng.choice(n, 2, replace=False))
        p = str(rng.choice(["ZZ", "XX", "YZ", "X", "Y"]))
        circ.pauli_rotation(p, [a] if len(p) == 1 else [a, b], float(rng.uniform(0.3, 1.2)))
    x = np.array([[1 << 3], [1 << 11]], np.uint64)
    z = np.array([[1 << 7], [0]], np.uint64)
    c = np.array([1.0, -0.5])
    a = propagate_both_ways(n, x, z, c, circ, "forward", truncation.coeff(1e-5) & truncation.weight(10))
    assert max(a[3]) > 50000
    o = po.OracleSum(n, x, z, c)
    olens = o.propagate(circ.channels(), "forward", oracle_policy(eps=1e-5, max_weight=10))
    ox, oz, oc, _ = o.export()
    assert [int(v) for v in olens] == a[3]
    assert np.array_equal(a[2], oc) and np.array_equal(a[0], ox)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_reset_and_layer_in_place_equals_the_out_of_place_layer(dtype):
    """The benchmark's protocol: a working sum reset from the resident input (pp_sum_copy) before every layer, the
    layer in place; against pp_apply_rotation_to from the same input into another sum."""
    announce(f"reset + in-place layer vs out-of-place layer, 2^17 keys on 127 qubits {np.dtype(dtype).name}")
    rng = np.random.default_rng(31)
    x, z, c = random_sum(rng, 1 << 17, 127, 0.1, dtype)
    src = PauliSum(127, x, z, c, dtype=dtype).to_device(0)
    work = src.clone()
    dst = src.clone()
    policy = truncation.coeff(1e-6)
    try:
        circ = Circuit(127)
        for k in range(0, 30, 2):
            circ.rzz(0.3 + 0.1 * k, k, k + 1)
        for q in range(0, 127, 9):
            circ.rx(0.2 + 0.01 * q, q)
        for ch in circ.channels():
            work.copy_from(src)
            work.apply_rotation(ch, policy=policy)
            src.apply_rotation_to(dst, ch, policy=policy)
            a, b = work.to_host(), dst.to_host()
            assert len(a) == len(b) > len(c)
            assert np.array_equal(a._x, b._x) and np.array_equal(a._z, b._z) and np.array_equal(a._c, b._c)
    finally:
        src.free()
        work.free()
        dst.free()
