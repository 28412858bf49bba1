"""Results do not depend on the device path's sizes: the same propagation under several tunings (bucket target,
staging rows, grid) downloads the same terms, equal to the oracle's; the autotuner returns one of its candidates
and leaves its sample as it was.  Each case prints its name before it touches the GPU."""

from __future__ import annotations

import math
import sys

import numpy as np
import pytest

import pp_oracle as po
from paulistrings_amd import Circuit, PauliSum, _native, truncation

pytestmark = pytest.mark.gpu

TUNINGS = [
    (256, 0, _native.ONE_PAIR_PER_GROUP, 256),
    (64, 64, _native.ONE_PAIR_PER_GROUP, 256),  # small staging: many pairs take the redo launch
    (384, 128, 0, 512),
    (128, 256, 3, 256),
    (256, 0, _native.ONE_PAIR_PER_GROUP, 512),
    (512, 64, 2, 512),
]


def announce(msg: str) -> None:
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def random_keys(rng, n, num_qubits, density):
    W = 2 if num_qubits > 64 else 1
    x = np.zeros((n, W), np.uint64)
    z = np.zeros((n, W), np.uint64)
    for w in range(W):
        live = min(64, num_qubits - 64 * w)
        weights = np.uint64(1) << np.arange(live, dtype=np.uint64)
        x[:, w] = ((rng.random((n, live)) < density) * weights).sum(axis=1, dtype=np.uint64)
        z[:, w] = ((rng.random((n, live)) < density) * weights).sum(axis=1, dtype=np.uint64)
    return x, z


def circuit(rng, nq, layers):
    c = Circuit(nq)
    for _ in range(layers):
        k = int(rng.integers(1, 3))
        qs = [int(q) for q in rng.choice(nq, size=k, replace=False)]
        c.pauli_rotation("".join(rng.choice(list("XYZ"), size=k)), qs, float(rng.uniform(-math.pi, math.pi)))
    return c


def test_tunings_give_the_oracle_terms():
    rng = np.random.default_rng(55)
    nq = 90
    x, z = random_keys(rng, 40000, nq, 0.15)
    c = rng.standard_normal(40000)
    circ = circuit(rng, nq, 12)
    s = po.OracleSum(nq, x, z, c)
    s.propagate(circ.channels(), "heisenberg", po.coeff_policy(eps=1e-5, approx_topn=120000))
    ox, oz, oc, _ = s.export()
    ctx = _native.Context.get(0)
    saved = ctx.get_tuning()
    try:
        for t in TUNINGS:
            announce(f"tuning {t}")
            ctx.set_tuning(*t)
            g = PauliSum(nq, x, z, c).to_device(0)
            g.propagate(circ, truncation.coeff(1e-5) & truncation.approx_topn(120000), "heisenberg")
            h = g.to_host()
            g.free()
            assert np.array_equal(h.x_array(), ox) and np.array_equal(h.z_array(), oz), t
            assert np.array_equal(h.coefficients_array(), oc), t
    finally:
        ctx.set_tuning(saved.bucket_target, saved.stage_rows, saved.groups_per_cu, saved.join_block)


def test_autotune_picks_a_candidate_and_keeps_the_sample():
    announce("autotune")
    rng = np.random.default_rng(8)
    nq = 127
    x, z = random_keys(rng, 1 << 15, nq, 0.1)
    c = rng.standard_normal(1 << 15)
    g = PauliSum(nq, x, z, c).to_device(0)
    before = g.to_host()
    ctx = g.ctx
    saved = ctx.get_tuning()
    try:
        t = g.autotune(Circuit(nq).rzz(0.9, 0, 1).rx(0.4, 77), truncation.coeff(1e-6))
        assert t.bucket_target in (128, 192, 256, 384) and t.stage_rows in (0, 128, 256)
        assert t.groups_per_cu in (0, _native.ONE_PAIR_PER_GROUP) and t.join_block in (256, 512)
        got = ctx.get_tuning()
        assert (got.bucket_target, got.stage_rows, got.groups_per_cu) == (t.bucket_target, t.stage_rows,
                                                                          t.groups_per_cu)
        after = g.to_host()
        assert np.array_equal(after.x_array(), before.x_array())
        assert np.array_equal(after.coefficients_array(), before.coefficients_array())
    finally:
        ctx.set_tuning(saved.bucket_target, saved.stage_rows, saved.groups_per_cu, saved.join_block)
        g.free()
