"""The class pass (pp_class.hip): a program of local tables applied in one pass, each class bucket's strings held in
LDS through every layer.

* Rotation runs (three to eight rotation tables): with PP_ROTATION_RUNS=class pp_apply_run and pp_propagate give them
  to the class pass; by default they go layer by layer through the pair join.  Both give the same keys, coefficients
  and per-layer lengths, and pp_propagate's equal the oracle's.  A sum whose strings crowd one class past the largest LDS
  set falls back to the layers one by one, with the same result.
* General local tables (dense one- and two-qubit transfer matrices of random unitaries, a non-unital channel) through
  pp_apply_prepared: bit for bit against a naive layer that sums each output's contributions in entry order (the
  class pass's order), and within 1e-12 (float64) of the reference's differential oracle naive_apply_layer
  (test_support/oracle.rs:21-54, restated in oracle/ptm.py), which sums in input order: only the summation order
  differs.

Each case prints its name before it touches the GPU."""

from __future__ import annotations

import ctypes
import math
import os
import sys

import numpy as np
import pytest

import pp_oracle as po
import ptm
from paulistrings_amd import Circuit, PauliSum, _native, truncation
from test_gpu_parity import oracle_policy, product_policy, random_sum
from test_gpu_seam import H, S, CNOT, keep_fn, table_of

pytestmark = pytest.mark.gpu


def announce(msg: str) -> None:
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def check(g, st):
    _native.check(st, g.ctx.handle)


class rotation_runs:
    """PP_ROTATION_RUNS=mode for the duration (the library reads it at every run)."""

    def __init__(self, mode: str):
        self.mode = mode

    def __enter__(self):
        self.saved = os.environ.get("PP_ROTATION_RUNS")
        os.environ["PP_ROTATION_RUNS"] = self.mode

    def __exit__(self, *exc):
        if self.saved is None:
            os.environ.pop("PP_ROTATION_RUNS", None)
        else:
            os.environ["PP_ROTATION_RUNS"] = self.saved


def terms_of(h, real):
    return {(tuple(int(v) for v in h._x[i]), tuple(int(v) for v in h._z[i])): real(h._c[i]) for i in range(len(h._c))}


def rotation_table(num_qubits, W, kind, pauli, qubits, theta):
    ch = _native.Channel()
    for q, p in zip(qubits, pauli):
        if p in "XY":
            ch.gx[q // 64] |= 1 << (q % 64)
        if p in "ZY":
            ch.gz[q // 64] |= 1 << (q % 64)
    ch.theta = theta
    p = _native.Prepared()
    assert _native.lib().pp_prepare_channel(num_qubits, W, kind, ctypes.byref(ch), 0, ctypes.byref(p)) == 0
    return p


def run_vs_layers(host, preps, policy):
    """pp_apply_run over `preps` in one class pass and, on a second copy, layer by layer: (lengths, terms) each."""
    L = _native.lib()
    pol = policy.lower()
    out = []
    for one_by_one in (False, True):
        g = host.to_device(0)
        try:
            arr = (_native.Prepared * len(preps))(*preps)
            lens = np.zeros(len(preps), np.uint64)
            with rotation_runs("layers" if one_by_one else "class"):
                check(g, L.pp_apply_run(g._h, _native.RUN_ROTATIONS, arr, len(preps), ctypes.byref(pol),
                                        lens.ctypes.data))
            h = g.to_host()
            out.append(([int(v) for v in lens], h))
        finally:
            g.free()
    (la, a), (lb, b) = out
    assert la == lb, "per-layer lengths differ"
    assert np.array_equal(a._x, b._x) and np.array_equal(a._z, b._z), "keys or canonical order differ"
    assert np.array_equal(a._c, b._c), "coefficients differ"
    return la, a


@pytest.mark.parametrize("num_qubits", [40, 127])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_rotation_run_in_one_pass_equals_its_layers(num_qubits, dtype):
    """Eight rotation tables on eight qubits (X, Z, Y and two-qubit generators, a Clifford quarter turn among them)
    over 2^16 terms: one class pass against the same tables one by one through the pair join."""
    announce(f"rotation run of 8 tables in one class pass, 2^16 terms n={num_qubits} {np.dtype(dtype).name}")
    rng = np.random.default_rng(4242 + num_qubits)
    W = 1 if num_qubits <= 64 else 2
    kind = 1 if dtype == np.float32 else 0
    x, z, c = random_sum(rng, 1 << 16, num_qubits, 0.1, dtype)
    host = PauliSum(num_qubits, x, z, c, dtype=dtype)
    qs = sorted(int(v) for v in rng.choice(num_qubits, 8, replace=False))
    if num_qubits > 64:
        qs[3], qs[4] = 63, 64  # a generator across the word boundary
        qs = sorted(set(qs))
    preps = []
    for j in range(8):
        a, b = qs[j % len(qs)], qs[(j + 1) % len(qs)]
        pauli, sup = [("X", [a]), ("Z", [a]), ("Y", [a]), ("ZZ", sorted([a, b])), ("XY", sorted([a, b]))][j % 5]
        theta = -math.pi / 2 if j == 5 else float(rng.uniform(0.2, 1.3))
        preps.append(rotation_table(num_qubits, W, kind, pauli, sup, theta))
    lens, _ = run_vs_layers(host, preps, truncation.coeff(1e-4))
    assert lens[-1] > len(host)


def test_rotation_run_with_a_crowded_class_falls_back():
    """Strings wholly inside the run's eight qubits share one class (their key off the run's qubits is zero): 3000 of
    them outgrow every LDS set the class pass tries, so the run goes layer by layer, with the same result."""
    announce("rotation run over one crowded class: the layer-by-layer fallback")
    rng = np.random.default_rng(77)
    n, qs = 64, list(range(8, 16))
    keys = set()
    while len(keys) < 3000:
        xv = sum(int(rng.integers(0, 2)) << q for q in qs)
        zv = sum(int(rng.integers(0, 2)) << q for q in qs)
        keys.add((xv, zv))
    keys = sorted(keys)
    x = np.array([[k[0]] for k in keys], np.uint64)
    z = np.array([[k[1]] for k in keys], np.uint64)
    c = rng.standard_normal(len(keys))
    host = PauliSum(n, x, z, c)
    preps = [rotation_table(n, 1, 0, "X", [q], 0.4 + 0.05 * q) for q in qs]
    lens, h = run_vs_layers(host, preps, truncation.coeff(1e-6))
    o = po.OracleSum(n, x, z, c)
    circ = Circuit(n)
    for q in qs:
        circ.rx(0.4 + 0.05 * q, q)
    olens = o.propagate(circ.channels(), "forward", oracle_policy(eps=1e-6))
    assert [int(v) for v in olens] == lens
    ox, oz, oc, _ = o.export()
    assert np.array_equal(h._x, ox) and np.array_equal(h._c, oc)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("direction", ["forward", "heisenberg"])
def test_propagate_with_rotation_runs_equals_layers_and_the_oracle(dtype, direction):
    """The 6x6 TFIM's steps (runs of three to eight RZZ or RX tables) and a kicked Ising prefix: pp_propagate with the
    class pass and layer by layer agree bit for bit, and both with the oracle."""
    from paulistrings_amd.circuits import heavy_hex_kicked_ising, repeat, tfim_square_step

    cases = [
        ("tfim6x6 x4", 36, repeat(tfim_square_step(6, 6, dt=0.1, j=1.0, h=3.0), 4),
         product_policy(eps=1e-4, max_weight=8), oracle_policy(eps=1e-4, max_weight=8),
         (np.uint64(1) << np.arange(36, dtype=np.uint64)).reshape(-1, 1), np.zeros((36, 1), np.uint64),
         np.full(36, 1.0 / 36)),
        ("kicked ising x5", 127, heavy_hex_kicked_ising(5, 0.8, -math.pi / 2, order="zz-then-x"),
         product_policy(eps=1e-5, topn=1 << 16), oracle_policy(eps=1e-5, topn=1 << 16),
         np.zeros((1, 2), np.uint64), np.array([[1 << 62, 0]], np.uint64), np.ones(1)),
    ]
    for name, n, circ, pol, opol, x, z, c in cases:
        announce(f"pp_propagate with rotation runs: {name} {np.dtype(dtype).name} {direction}")
        c = c.astype(dtype)
        host = PauliSum(n, x, z, c, dtype=dtype)
        res = []
        for one_by_one in (False, True):
            g = host.to_device(0)
            try:
                with rotation_runs("layers" if one_by_one else "class"):
                    st = g.propagate_with_stats(circ, pol, direction)
                res.append((list(st.terms_out), g.to_host()))
            finally:
                g.free()
        (la, a), (lb, b) = res
        assert la == lb
        assert np.array_equal(a._x, b._x) and np.array_equal(a._z, b._z) and np.array_equal(a._c, b._c)
        o = po.OracleSum(n, x, z, c)
        olens = o.propagate(circ.channels(), direction, opol)
        assert [int(v) for v in olens] == la
        ox, oz, oc, _ = o.export()
        assert np.array_equal(a._x, ox) and np.array_equal(a._z, oz) and np.array_equal(a._c, oc)


# ---- general local tables ------------------------------------------------------------------------------------------

def haar_unitary(rng, dim):
    m = rng.standard_normal((dim, dim)) + 1j * rng.standard_normal((dim, dim))
    q, r = np.linalg.qr(m)
    return q * (np.diag(r) / np.abs(np.diag(r)))


def channel_ptm(kraus, k):
    """table[s][t] = Tr(P_t E(P_s)) / 2^k for E(P) = sum K P K^dag (real for a Hermiticity-preserving map)."""
    dim = 1 << (2 * k)
    t = np.zeros((16, 16))
    for s in range(dim):
        img = sum(kk @ ptm.pauli_matrix(s, k) @ kk.conj().T for kk in kraus)
        for q in range(dim):
            v = np.trace(ptm.pauli_matrix(q, k) @ img) / (1 << k)
            assert abs(v.imag) < 1e-12
            t[s, q] = v.real
    return t


def table_from(t, qubits, kind):
    k = len(qubits)
    p = _native.Prepared()
    p.kind, p.k = _native.PREPARED_LOCAL, k
    for j, q in enumerate(qubits):
        p.qubits[j] = q
    es = ptm.entries_of(t, k)
    if kind:
        es = [(d, [float(np.float32(a)) for a in amp]) for d, amp in es]
    p.n_deltas = len(es)
    for e, (d, amp) in enumerate(es):
        p.deltas[e].local_delta = d
        for s in range(16):
            p.deltas[e].amplitude[s] = amp[s]
    return p, es


def entry_order_layer(terms, qubits, entries, keep, real):
    """The class pass's order: every output K gathers, entry by entry, the present source K ^ lift(d) with amplitude
    amp[t ^ d] (t = K's pattern), one rounded multiply and one rounded add each; kept iff nonzero and keep."""
    width = len(next(iter(terms))[0]) if terms else 1
    lifted = {d: ptm.lift(d, qubits, width) for d, _ in entries}

    def xor(key, d):
        mx, mz = lifted[d]
        return (tuple(a ^ m for a, m in zip(key[0], mx)), tuple(a ^ m for a, m in zip(key[1], mz)))

    cands = set()
    for key in terms:
        s = ptm.pattern_of(key[0], key[1], qubits)
        for d, amp in entries:
            if amp[s] != 0.0:
                cands.add(xor(key, d))
    out = {}
    for K in cands:
        t = ptm.pattern_of(K[0], K[1], qubits)
        w = None
        for d, amp in entries:
            a = amp[t ^ d]
            src = xor(K, d)
            if a == 0.0 or src not in terms:
                continue
            v = real(real(terms[src]) * real(a))
            w = v if w is None else real(w + v)
        if w != 0 and keep(K[0], K[1], w):
            out[K] = w
    return out


@pytest.mark.parametrize("num_qubits", [40, 127])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_general_tables_equal_the_naive_layer(num_qubits, dtype):
    """Dense transfer matrices (random two- and one-qubit unitaries, two-qubit amplitude damping, a Clifford among
    them) through pp_apply_prepared, one class pass each: lengths after every layer, keys and coefficients against
    the entry-order naive layer bit for bit, and against the input-order naive layer within 1e-12 (float64)."""
    announce(f"general local tables n={num_qubits} {np.dtype(dtype).name}")
    rng = np.random.default_rng(6060 + num_qubits)
    kind = 1 if dtype == np.float32 else 0
    real = np.float32 if kind else float
    x, z, c = random_sum(rng, 2500, num_qubits, 0.08, dtype)
    host = PauliSum(num_qubits, x, z, c, dtype=dtype)
    g = host.to_device(0)
    terms = terms_of(host, real)
    ref = dict(terms)
    eps = 1e-3
    pol = truncation.coeff(eps).lower()
    keep = keep_fn(eps, None, real)
    L = _native.lib()
    gamma = 0.3
    damp1 = [np.array([[1, 0], [0, math.sqrt(1 - gamma)]]), np.array([[0, math.sqrt(gamma)], [0, 0]])]
    layers = []
    for j in range(5):
        if j in (0, 3):
            a, b = sorted(int(v) for v in rng.choice(num_qubits, 2, replace=False))
            if j == 3 and num_qubits > 64:
                a, b = 63, 64
            layers.append((channel_ptm([haar_unitary(rng, 4)], 2), [a, b]))
        elif j == 1:
            layers.append((channel_ptm([haar_unitary(rng, 2)], 1), [int(rng.integers(0, num_qubits))]))
        elif j == 2:
            a, b = sorted(int(v) for v in rng.choice(num_qubits, 2, replace=False))
            kr = [np.kron(k1, k2) for k1 in damp1 for k2 in damp1]
            layers.append((channel_ptm(kr, 2), [a, b]))
        else:
            a, b = sorted(int(v) for v in rng.choice(num_qubits, 2, replace=False))
            p, es = table_of(CNOT, [a, b], kind)
            layers.append((None, (p, es, [a, b])))
    try:
        for t, qs in layers:
            if t is None:
                p, es, qs = qs
            else:
                p, es = table_from(t, qs, kind)
            check(g, L.pp_apply_prepared(g._h, ctypes.byref(p), ctypes.byref(pol)))
            terms = entry_order_layer(terms, qs, es, keep, real)
            ref = ptm.naive_table_layer(ref, qs, es, keep, real)
            assert len(g) == len(terms)
        got = terms_of(g.to_host(), real)
        assert set(got) == set(terms)
        assert all(got[k] == terms[k] for k in terms), "coefficients differ from the entry-order layer"
        if not kind:
            assert set(ref) == set(terms)
            assert all(abs(got[k] - ref[k]) <= 1e-12 * max(1.0, abs(ref[k])) for k in ref)
    finally:
        g.free()


def test_general_table_at_2e16_terms():
    """One dense two-qubit unitary on 2^16 terms of 127 qubits (hundreds of class buckets, several apply workgroups
    each): lengths, keys and coefficients against the entry-order naive layer."""
    announce("general table on 2^16 terms, 127 qubits")
    rng = np.random.default_rng(99)
    x, z, c = random_sum(rng, 1 << 16, 127, 0.1, np.float64)
    host = PauliSum(127, x, z, c)
    g = host.to_device(0)
    try:
        t = channel_ptm([haar_unitary(rng, 4)], 2)
        p, es = table_from(t, [17, 90], 0)
        pol = truncation.coeff(1e-5).lower()
        check(g, _native.lib().pp_apply_prepared(g._h, ctypes.byref(p), ctypes.byref(pol)))
        want = entry_order_layer(terms_of(host, float), [17, 90], es, keep_fn(1e-5, None, float), float)
        got = terms_of(g.to_host(), float)
        assert len(got) == len(want) > len(host)
        assert got == want
    finally:
        g.free()
