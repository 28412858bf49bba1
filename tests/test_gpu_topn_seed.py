"""TopN's seeded first pass: a selection by |c| starts from the previous selection's boundary, and when the first
digit of the new boundary is the previous one's, one pass over the terms decides two digits.  Selections whose
boundary stays (small steps in n), moves within the first digit and jumps (large steps) all keep exactly the terms
the oracle keeps, as do the unseeded passes (PP_TOPN_SEEDING=0).  Each case prints its name first."""

from __future__ import annotations

import math
import os
import sys

import numpy as np
import pytest

import pp_oracle as po
from paulistrings_amd import PauliSum, truncation
from paulistrings_amd.circuits import heavy_hex_kicked_ising
from test_gpu_parity import oracle_policy, random_sum

pytestmark = pytest.mark.gpu


def announce(msg: str) -> None:
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("seeding", ["1", "0"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_repeated_selections_equal_the_oracle(seeding, dtype, monkeypatch):
    announce(f"repeated TopN selections, seeding={seeding} {np.dtype(dtype).name}")
    monkeypatch.setenv("PP_TOPN_SEEDING", seeding)
    rng = np.random.default_rng(404)
    x, z, c = random_sum(rng, 1 << 17, 127, 0.1, dtype)
    c[::7] = c[3]  # exact ties across the boundary
    host = PauliSum(127, x, z, c, dtype=dtype)
    g = host.to_device(0)
    o = po.OracleSum(127, host._x, host._z, host._c)
    try:
        n = len(host)
        for step in [5, 1, 1, 400, 2, 30000, 3, 1, 9000, 17]:
            n -= step
            g.truncate_topn(n)
            o.topn(n)
            h = g.to_host()
            ox, oz, oc, _ = o.export()
            assert len(h) == len(oc) == n
            assert np.array_equal(h._x, ox) and np.array_equal(h._z, oz) and np.array_equal(h._c, oc)
    finally:
        g.free()


def test_kicked_ising_at_a_binding_cap_seeded_and_not():
    announce("kicked Ising zz-then-x, 127 qubits, 10 steps, TopN 2^15: seeded vs unseeded vs oracle")
    circ = heavy_hex_kicked_ising(10, 0.8, -math.pi / 2, order="zz-then-x")
    x = np.zeros((1, 2), np.uint64)
    z = np.array([[1 << 62, 0]], np.uint64)
    pol = truncation.coeff(1e-5) & truncation.topn(1 << 15)
    res = []
    for seeding in ("1", "0"):
        os.environ["PP_TOPN_SEEDING"] = seeding
        try:
            g = PauliSum(127, x, z, np.ones(1)).to_device(0)
            st = g.propagate_with_stats(circ, pol, "heisenberg")
            res.append((list(st.terms_out), g.to_host()))
            g.free()
        finally:
            os.environ.pop("PP_TOPN_SEEDING", None)
    (la, a), (lb, b) = res
    assert la == lb and la.count(1 << 15) > 100
    assert np.array_equal(a._x, b._x) and np.array_equal(a._z, b._z) and np.array_equal(a._c, b._c)
    o = po.OracleSum(127, x, z, np.ones(1))
    olens = o.propagate(circ.channels(), "heisenberg", oracle_policy(eps=1e-5, topn=1 << 15))
    assert [int(v) for v in olens] == la
    ox, oz, oc, _ = o.export()
    assert np.array_equal(a._x, ox) and np.array_equal(a._c, oc)
