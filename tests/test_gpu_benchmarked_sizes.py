"""Parity at the sizes the benchmark lines are quoted on, through the C ABI, against the oracle.

- The saturated dense layer: 2^24 distinct keys on 127 qubits generated exactly as bench.py generates them (same
  seed, same device generator), 16 device bucket bits as in the benchmark.  Two RZZ and two RX channels of the dense
  layer, each applied out of place from the resident input (the benchmark's timed call), and the same four channels
  as one propagation; f64 and f32.  Keys, canonical order and coefficients bit-exact, every layer's length equal.
- The heavy-hex kicked Ising as the benchmark runs it (zz-then-x, RX(0.8), threshold 1e-5 with a TopN cap): with the
  cap cut to 2^18 for all 20 steps, so that it binds in most layers, and at the full 2^24 cap for the first steps
  the oracle finishes in about a minute.  Per-layer lengths, keys and coefficients bit-exact, <Z62> within 1e-12.

Each case prints its name before it touches the GPU."""

from __future__ import annotations

import math
import sys

import numpy as np
import pytest

import bench
import pp_oracle as po
from paulistrings_amd import DEFAULT_HASH_SEED, PauliSum, circuits, truncation

pytestmark = pytest.mark.gpu


def announce(msg: str) -> None:
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def distinct_host_sum(num_qubits, x, z, c) -> PauliSum:
    """A PauliSum over keys already known to be distinct (the benchmark's input), without the host re-sort."""
    s = PauliSum.__new__(PauliSum)
    s.num_qubits, s.width, s.hash_seed = num_qubits, circuits.width_for(num_qubits), DEFAULT_HASH_SEED
    s._x, s._z, s._c = x, z, c
    return s


def assert_canonical_equal(got: PauliSum, ox, oz, oc, what: str):
    assert len(got) == len(oc), (what, len(got), len(oc))
    assert np.array_equal(got._x, ox) and np.array_equal(got._z, oz), f"{what}: keys or canonical order differ"
    assert np.array_equal(got._c, oc), f"{what}: coefficients differ"


@pytest.mark.timeout(900)
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_saturated_layer_at_2e24_keys(dtype):
    announce(f"saturated layer 2^24 keys {dtype}: generating the benchmark input")
    x, z, c = bench.saturated_keys(1 << 24, bench.SEED, dtype)
    assert len(c) > (1 << 24) - 1000  # distinct draws: a handful of collisions at most
    chans = bench.dense_layer_channels(bench.SEED)
    picked = [0, 1, 63, 64]  # RZZ(Z0 Z1), RZZ(Z2 Z3), RX(X0), RX(X1)
    policy = truncation.coeff(bench.THRESHOLD)
    opol = po.coeff_policy(eps=bench.THRESHOLD)
    base = po.OracleSum(127, x, z, c)
    src = distinct_host_sum(127, x, z, c).to_device(0)
    dst = src.clone()
    try:
        assert src.bucket_bits()[1] == 16, src.bucket_bits()
        for i in picked:
            announce(f"saturated layer 2^24 {dtype}: channel {i} out of place")
            src.apply_rotation_to(dst, chans[i], policy=policy)
            got = dst.to_host()
            o = base.clone()
            olen = o.propagate(chans[i:i + 1], "forward", opol)
            ox, oz, oc, _ = o.export()
            del o
            assert int(olen[0]) == len(got)
            assert_canonical_equal(got, ox, oz, oc, f"channel {i}")
            del got, ox, oz, oc
        announce(f"saturated layer 2^24 {dtype}: the four channels as one propagation")
        circ = chans[picked]
        olens = base.propagate(circ, "forward", opol)
        ox, oz, oc, _ = base.export()
        host = distinct_host_sum(127, x, z, c)
        stats = host.propagate_with_stats(_as_circuit(circ), policy, "forward")
        assert [int(v) for v in olens] == stats.terms_out
        assert_canonical_equal(host, ox, oz, oc, "propagation")
    finally:
        src.free()
        dst.free()


def _as_circuit(chans) -> circuits.Circuit:
    circ = circuits.Circuit(127)
    for row in np.asarray(chans):
        circ._rows.append((tuple(int(v) for v in row["gx"]), tuple(int(v) for v in row["gz"]), float(row["theta"])))
    return circ


def kicked_ising_both(cap: int, steps: int):
    circ = circuits.heavy_hex_kicked_ising(steps, 0.8, -math.pi / 2, order="zz-then-x")
    x = np.zeros((1, 2), np.uint64)
    z = np.zeros((1, 2), np.uint64)
    z[0, 0] = np.uint64(1) << np.uint64(62)
    c = np.array([1.0])
    o = po.OracleSum(127, x, z, c)
    olens = o.propagate(circ.channels(), "heisenberg", po.coeff_policy(eps=1e-5, topn=cap))
    ov = o.expectation("z")
    ox, oz, oc, _ = o.export()
    del o
    g = PauliSum(127, x, z, c).to_device(0)
    try:
        stats = g.propagate_with_stats(circ, truncation.coeff(1e-5) & truncation.topn(cap), "heisenberg")
        gv = g.expectation("z+")
        got = g.to_host()
    finally:
        g.free()
    assert [int(v) for v in olens] == stats.terms_out, "per-layer lengths differ"
    assert_canonical_equal(got, ox, oz, oc, f"kicked Ising cap {cap}")
    assert abs(gv - ov) <= 1e-12, (gv, ov)
    return olens


@pytest.mark.timeout(900)
def test_kicked_ising_cap_binds_every_step():
    announce("kicked Ising zz-then-x, RX(0.8), 20 steps, threshold 1e-5, TopN 2^18")
    cap = 1 << 18
    olens = kicked_ising_both(cap, 20)
    assert (olens == cap).sum() > 2000  # the cap binds in most layers of the later steps


@pytest.mark.timeout(900)
def test_kicked_ising_full_cap_first_steps():
    announce("kicked Ising zz-then-x, RX(0.8), threshold 1e-5, TopN 2^24, first 6 steps")
    cap = 1 << 24
    olens = kicked_ising_both(cap, 6)
    assert (olens == cap).sum() > 0  # the full cap binds before the sixth step ends
