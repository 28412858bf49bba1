"""The layer seam of the C ABI (include/pp_hip.h: pp_prepare_channel, pp_apply_prepared, pp_apply_run,
pp_finalize_layer, pp_settle, pp_sum_refine, pp_sum_coarsen_to), driven by a host loop that plays the reference's
run_layers (engine/driver.rs:319-592) over PartitionBackend (engine/backend.rs:26-108): the reference's bucket count
moved at the start of every run, Runs::take's permutation and rotation runs (engine/driver.rs:202-258), the layer
pass after every run, settle at the end.  The loop's result equals pp_propagate's bit for bit, with every layer's
length, and pp_propagate's equals the oracle's.  One-to-one and key-preserving local tables of other channels
(Clifford gates, Pauli noise) built from their matrices are checked against the naive table layer.  Each case prints
its name before it touches the GPU."""

from __future__ import annotations

import ctypes
import math
import sys

import numpy as np
import pytest

import pp_oracle as po
import ptm
from paulistrings_amd import Circuit, PauliSum, _native, circuits, truncation
from test_gpu_parity import oracle_policy, product_policy, random_circuit, random_sum

pytestmark = pytest.mark.gpu

EPS = 2.220446049250313e-16


def announce(msg: str) -> None:
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def check(g, st):
    _native.check(st, g.ctx.handle)


# ---- the host's questions about a prepared channel (what a Rust host answers from its own Prepared) ----------------

def entries(p):
    return [(int(p.deltas[e].local_delta), [float(v) for v in p.deltas[e].amplitude]) for e in range(p.n_deltas)]


def rotation_layer_of(p):
    """RotationLayer::of (engine/bucketed/rotation_layer.rs:101-140): (shaped, clifford)."""
    if p.kind != _native.PREPARED_LOCAL:
        return False, False
    es = entries(p)
    if len(es) != 2 or es[0][0] != 0:
        return False, False
    g, a0, a1 = es[1][0], es[0][1], es[1][1]
    clifford = True
    for s in range(1 << (2 * p.k)):
        if a1[s] == 0.0:
            if a0[s] != 1.0:
                return False, False
        else:
            if a1[s ^ g] == 0.0:
                return False, False
            clifford = clifford and a0[s] * a0[s] <= EPS * EPS * a1[s] * a1[s]
    return True, clifford


def max_fanout(p):
    """max_fanout (engine/bucketed.rs:628-641)."""
    if p.kind != _native.PREPARED_LOCAL:
        return 2
    shaped, clifford = rotation_layer_of(p)
    if shaped and clifford:
        return 1
    es = entries(p)
    return max(1, max(sum(1 for _, a in es if a[s] != 0.0) for s in range(16)))


def is_permutation_step(p):
    """PermutationStep::of (engine/bucketed/permutation_run.rs:53-121)."""
    if p.kind != _native.PREPARED_LOCAL or p.k == 0:
        return False
    es = entries(p)
    shaped, clifford = rotation_layer_of(p)
    if shaped and not clifford:
        return False
    reached = set()
    for s in range(1 << (2 * p.k)):
        if shaped:
            e = 1 if es[1][1][s] != 0.0 else 0
        else:
            hits = [e for e, (_, a) in enumerate(es) if a[s] != 0.0]
            if len(hits) > 1:
                return False
            if not hits:
                continue
            e = hits[0]
        t = s ^ es[e][0]
        if t in reached:
            return False
        reached.add(t)
    return True


def span_dimension(v):
    v = list(v)
    rank = 0
    for bit in range(31, -1, -1):
        piv = next((d for d in v if (d >> bit) & 1), None)
        if piv is None:
            continue
        v.remove(piv)
        v = [d ^ piv if (d >> bit) & 1 else d for d in v]
        rank += 1
    return rank


class SeamHost:
    """The host side of a device partition: the sum's handle, its hash rows and the reference's bucket count."""

    def __init__(self, g, target=1024, min_buckets=128):
        self.g, self.L = g, _native.lib()
        self.W, self.nq = g.width, g.num_qubits
        rx = np.zeros(31 * self.W, np.uint64)
        rz = np.zeros(31 * self.W, np.uint64)
        self.max_bits = self.L.pp_hash_rows(self.nq, 0x9E3779B97F4A7C15, self.W, rx.ctypes.data, rz.ctypes.data)
        self.rx, self.rz = rx.reshape(-1, self.W), rz.reshape(-1, self.W)
        self.target, self.min_buckets = target, min_buckets
        self.kind = _native.COEFF_F64 if g.dtype == np.float64 else _native.COEFF_F32

    def bits(self):
        return self.g.bucket_bits()[0]

    def bucket_of(self, mx, mz, bits):
        b = 0
        for r in range(bits):
            f = 0
            for w in range(self.W):
                f ^= (mx[w] & int(self.rx[r, w])) ^ (mz[w] & int(self.rz[r, w]))
            b |= (bin(f).count("1") & 1) << r
        return b

    def generator_delta(self, p, bits):
        mx, mz = ptm.lift(entries(p)[1][0], list(p.qubits)[: p.k], self.W)
        return self.bucket_of(mx, mz, bits)

    def prepare(self, row, adjoint):
        ch = _native.Channel()
        for w in range(4):
            ch.gx[w] = int(row["gx"][w])
            ch.gz[w] = int(row["gz"][w])
        ch.theta = float(row["theta"])
        p = _native.Prepared()
        assert self.L.pp_prepare_channel(self.nq, self.W, self.kind, ctypes.byref(ch), 1 if adjoint else 0,
                                         ctypes.byref(p)) == 0
        return p

    def rebucket(self):
        """refine_unprepared's count at the start of a run (engine/backend.rs:155-165)."""
        n, bits = len(self.g), self.bits()
        want = min(po.lib().ppo_rebucketed_bits(n, bits, self.target, self.min_buckets), self.max_bits)
        while self.bits() < want:
            check(self.g, self.L.pp_sum_refine(self.g._h))
        if want < self.bits():
            check(self.g, self.L.pp_sum_coarsen_to(self.g._h, want))

    def run_layers(self, circ: Circuit, policy, direction):
        """run_layers over the seam; returns every layer's length."""
        rows = circ.channels()
        adjoint = direction == "heisenberg"
        if adjoint:
            rows = rows[::-1]
        pol = truncation.lower(policy)
        limit = int(pol.topn) if pol.has_topn else 1 << 62
        preps = [self.prepare(r, adjoint) for r in rows]
        lens = []
        i, n = 0, len(preps)
        while i < n:
            self.rebucket()
            length = len(self.g)
            p = preps[i]
            end, kind = i + 1, None
            if limit > 0:
                if length <= limit and is_permutation_step(p):
                    j = i + 1
                    while j < n and is_permutation_step(preps[j]):
                        j += 1
                    if j > i + 1:
                        end, kind = j, _native.RUN_PERMUTATIONS
                worst = length * max_fanout(p)
                if kind is None and worst <= limit and rotation_layer_of(p)[0]:
                    bits = self.bits()
                    deltas = [0, self.generator_delta(p, bits)]
                    j = i + 1
                    while j < n and j < i + 8:
                        q = preps[j]
                        grown = worst * max_fanout(q)
                        if grown > limit or not rotation_layer_of(q)[0]:
                            break
                        nxt = deltas + [0, self.generator_delta(q, bits)]
                        if span_dimension(nxt) > 3:
                            break
                        deltas, worst = nxt, grown
                        j += 1
                    if j > i + 1:
                        end, kind = j, _native.RUN_ROTATIONS
            if kind is None:
                check(self.g, self.L.pp_apply_prepared(self.g._h, ctypes.byref(p), ctypes.byref(pol)))
                lens.append(len(self.g))
            else:
                arr = (_native.Prepared * (end - i))(*preps[i:end])
                out = np.zeros(end - i, np.uint64)
                check(self.g, self.L.pp_apply_run(self.g._h, kind, arr, end - i, ctypes.byref(pol),
                                                  out.ctypes.data))
                lens.extend(int(v) for v in out)
            check(self.g, self.L.pp_finalize_layer(self.g._h, ctypes.byref(pol)))
            if pol.has_topn:
                lens[-1] = len(self.g)
            i = end
        check(self.g, self.L.pp_settle(self.g._h))
        return lens


def seam_vs_propagate(num_qubits, x, z, c, circ, direction, policy, opolicy):
    host = PauliSum(num_qubits, x, z, c, dtype=c.dtype)
    a = host.to_device(0)
    b = host.to_device(0)
    try:
        seam_lens = SeamHost(a).run_layers(circ, policy, direction)
        stats = b.propagate_with_stats(circ, policy, direction)
        assert seam_lens == stats.terms_out, "per-layer lengths differ"
        ga, gb = a.to_host(), b.to_host()
        assert np.array_equal(ga._x, gb._x) and np.array_equal(ga._z, gb._z), "keys or canonical order differ"
        assert np.array_equal(ga._c, gb._c), "coefficients differ"
        assert a.bucket_bits()[0] == b.bucket_bits()[0]
    finally:
        a.free()
        b.free()
    o = po.OracleSum(num_qubits, x, z, c)
    olens = o.propagate(circ.channels(), direction, opolicy)
    ox, oz, oc, _ = o.export()
    assert [int(v) for v in olens] == seam_lens
    assert np.array_equal(ga._x, ox) and np.array_equal(ga._z, oz) and np.array_equal(ga._c, oc)
    return seam_lens


@pytest.mark.parametrize("num_qubits", [12, 64, 127])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("direction", ["forward", "heisenberg"])
def test_host_loop_over_the_seam_equals_propagate(num_qubits, dtype, direction):
    announce(f"seam host loop: random circuit n={num_qubits} {np.dtype(dtype).name} {direction}, TopN")
    rng = np.random.default_rng(5150 + num_qubits)
    x, z, c = random_sum(rng, 2500, num_qubits, 0.12, dtype)
    circ = random_circuit(rng, num_qubits, 60, max_support=3)
    for q in range(0, num_qubits - 1, 5):  # runs of quarter turns (permutation runs) between the others
        circ.pauli_rotation("ZZ", [q, q + 1], -math.pi / 2)
    for q in range(0, num_qubits, 7):  # and rotation runs
        circ.rx(0.3 + 0.01 * q, q)
    lens = seam_vs_propagate(num_qubits, x, z, c, circ, direction,
                             product_policy(eps=1e-4, max_weight=num_qubits // 2, topn=1500),
                             oracle_policy(eps=1e-4, max_weight=num_qubits // 2, topn=1500))
    assert max(lens) == 1500  # the cap binds: the deferred selections are read by the next layers
    seam_vs_propagate(num_qubits, x, z, c, circ, direction, product_policy(eps=1e-3, approx_topn=3000),
                      oracle_policy(eps=1e-3, approx_topn=3000))


def test_host_loop_over_the_seam_kicked_ising():
    announce("seam host loop: kicked Ising zz-then-x, 127 qubits, 8 steps, TopN 2^13")
    circ = circuits.heavy_hex_kicked_ising(8, 0.8, -math.pi / 2, order="zz-then-x")
    x = np.zeros((1, 2), np.uint64)
    z = np.zeros((1, 2), np.uint64)
    z[0, 0] = np.uint64(1) << np.uint64(62)
    lens = seam_vs_propagate(127, x, z, np.array([1.0]), circ, "heisenberg",
                             truncation.coeff(1e-5) & truncation.topn(1 << 13),
                             oracle_policy(eps=1e-5, topn=1 << 13))
    assert lens.count(1 << 13) > 100


def test_host_loop_over_the_seam_tfim():
    announce("seam host loop: 6x6 TFIM, 10 steps, threshold 1e-4, weight 8")
    circ = circuits.tfim_square_step(6, 6, 0.1, 1.0, 3.0)
    full = Circuit(36)
    for _ in range(10):
        full._rows.extend(circ._rows)
    x = np.array([[1 << q] for q in range(36)], np.uint64)
    z = np.zeros((36, 1), np.uint64)
    seam_vs_propagate(36, x, z, np.full(36, 1 / 36), full, "heisenberg",
                      truncation.coeff(1e-4) & truncation.weight(8), oracle_policy(eps=1e-4, max_weight=8))


# ---- local tables of other channels --------------------------------------------------------------------------------

H = np.array([[1, 1], [1, -1]]) / math.sqrt(2)
S = np.diag([1, 1j])
CNOT = np.array([[1, 0, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0], [0, 1, 0, 0]])  # control = support qubit 0 (low bit)
CZ = np.diag([1, 1, 1, -1])
SWAP = np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]])


def table_of(u, qubits, kind=0):
    """A pp_prepared of the unitary's transfer matrix (exact for a Clifford: entries rounded to +-1, 0)."""
    k = len(qubits)
    t = np.round(ptm.conjugation_ptm(np.asarray(u, complex), k), 12)
    p = _native.Prepared()
    p.kind, p.k = _native.PREPARED_LOCAL, k
    for j, q in enumerate(qubits):
        p.qubits[j] = q
    es = ptm.entries_of(t, k)
    p.n_deltas = len(es)
    for e, (d, amp) in enumerate(es):
        p.deltas[e].local_delta = d
        for s in range(16):
            p.deltas[e].amplitude[s] = float(np.float32(amp[s])) if kind else amp[s]
    return p, es


def pauli_noise_table(qubits, scale):
    """A Pauli-diagonal channel (depolarizing / dephasing): pattern s keeps its key, scaled by scale[s]."""
    p = _native.Prepared()
    p.kind, p.k = _native.PREPARED_LOCAL, len(qubits)
    for j, q in enumerate(qubits):
        p.qubits[j] = q
    p.n_deltas = 1
    p.deltas[0].local_delta = 0
    for s in range(16):
        p.deltas[0].amplitude[s] = scale[s] if s < (1 << (2 * len(qubits))) else 0.0
    return p, [(0, list(p.deltas[0].amplitude))]


def keep_fn(eps, max_weight, real):
    def keep(xw, zw, v):
        if eps is not None and not (float(v) * float(v) > eps * eps):
            return False
        if max_weight is not None and sum(bin(a | b).count("1") for a, b in zip(xw, zw)) > max_weight:
            return False
        return True
    return keep


@pytest.mark.parametrize("num_qubits", [40, 127])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_clifford_and_noise_tables_equal_the_naive_layer(num_qubits, dtype):
    """Runs of Clifford gates given as signed-permutation tables (H, S, CNOT, CZ, SWAP on random qubits, some across
    qubits 63/64) through pp_apply_run's permutation walk, and Pauli-noise tables (key-preserving, a scale per
    pattern) through pp_apply_prepared, against the naive table layer applied gate by gate: keys and coefficients
    bitwise (every output has one contribution), every layer's length."""
    announce(f"Clifford and noise tables n={num_qubits} {np.dtype(dtype).name}")
    rng = np.random.default_rng(880 + num_qubits)
    kind = 1 if dtype == np.float32 else 0
    real = np.float32 if kind else float
    x, z, c = random_sum(rng, 4000, num_qubits, 0.1, dtype)
    host = PauliSum(num_qubits, x, z, c, dtype=dtype)  # repeated draws of one key are summed here
    g = host.to_device(0)
    hx, hz, hc = host.x_array(), host.z_array(), host.coefficients_array()
    terms = {(tuple(int(v) for v in hx[i]), tuple(int(v) for v in hz[i])): real(hc[i]) for i in range(len(hc))}
    eps, mw = 1e-3, num_qubits // 3
    pol = (truncation.coeff(eps) & truncation.weight(mw)).lower()
    keep = keep_fn(eps, mw, real)
    L = _native.lib()
    try:
        for block in range(4):
            gates = []
            for _ in range(12):
                kind_g = int(rng.integers(0, 5))
                if kind_g < 2:
                    q = int(rng.integers(0, num_qubits))
                    gates.append(table_of([H, S][kind_g], [q], kind))
                else:
                    a, b = sorted(int(v) for v in rng.choice(num_qubits, size=2, replace=False))
                    if block == 1 and num_qubits > 64:
                        a, b = 63, 64
                    gates.append(table_of([CNOT, CZ, SWAP][kind_g - 2], [a, b], kind))
            arr = (_native.Prepared * len(gates))(*[p for p, _ in gates])
            out = np.zeros(len(gates), np.uint64)
            check(g, L.pp_apply_run(g._h, _native.RUN_PERMUTATIONS, arr, len(gates), ctypes.byref(pol),
                                    out.ctypes.data))
            want_lens = []
            for p, es in gates:
                qs = list(p.qubits)[: p.k]
                terms = ptm.naive_table_layer(terms, qs, es, keep, real)
                want_lens.append(len(terms))
            assert [int(v) for v in out] == want_lens
            q = sorted(int(v) for v in rng.choice(num_qubits, size=2, replace=False))
            pr = float(rng.uniform(0.05, 0.3))
            scale = [1.0] + [1.0 - pr] * 15  # two-qubit depolarizing: every non-identity pattern shrinks
            if kind:
                scale = [float(np.float32(v)) for v in scale]
            p, es = pauli_noise_table(q, scale)
            check(g, L.pp_apply_prepared(g._h, ctypes.byref(p), ctypes.byref(pol)))
            terms = ptm.naive_table_layer(terms, q, es, keep, real)
            assert len(g) == len(terms)
        got = g.to_host()
        gd = {(tuple(int(v) for v in got._x[i]), tuple(int(v) for v in got._z[i])): got._c[i]
              for i in range(len(got._c))}
        assert set(gd) == set(terms)
        assert all(gd[k] == terms[k] for k in terms), "coefficients differ"
    finally:
        g.free()


def test_malformed_tables_are_refused():
    announce("malformed prepared tables")
    g = PauliSum.from_strings({"XZ": 1.0}, num_qubits=2).to_device(0)
    L = _native.lib()
    try:
        p, _ = table_of(CNOT, [1, 0])  # support not ascending
        assert L.pp_apply_prepared(g._h, ctypes.byref(p), None) == 5
        p, _ = table_of(CNOT, [0, 5])  # outside the register
        assert L.pp_apply_prepared(g._h, ctypes.byref(p), None) == 5
        p, _ = table_of(H, [0])
        arr = (_native.Prepared * 1)(p)
        assert L.pp_apply_run(g._h, _native.RUN_ROTATIONS, arr, 1, None, None) == 5  # not rotation-shaped
        assert len(g) == 1
    finally:
        g.free()
