"""Parity of the HIP path (through the C ABI) with the oracle: keys bit-exact, coefficients bit-exact (every
output has at most two contributions, so no summation order is involved), canonical order identical, and the
reference's golden kicked-Ising records reproduced.  Each case prints its name before it touches the GPU."""

from __future__ import annotations

import math
import sys

import numpy as np
import pytest

import golden_data
import pp_oracle as po
from paulistrings_amd import Circuit, PauliSum, circuits, truncation

pytestmark = pytest.mark.gpu

F64 = np.float64
F32 = np.float32


def announce(msg: str) -> None:
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def oracle_policy(eps=None, max_weight=None, topn=None, approx_topn=None):
    return po.coeff_policy(eps=eps, max_weight=max_weight, topn=topn, approx_topn=approx_topn)


def product_policy(eps=None, max_weight=None, topn=None, approx_topn=None):
    parts = []
    if eps is not None:
        parts.append(truncation.coeff(eps))
    if max_weight is not None:
        parts.append(truncation.weight(max_weight))
    if topn is not None:
        parts.append(truncation.topn(topn))
    if approx_topn is not None:
        parts.append(truncation.approx_topn(approx_topn))
    if not parts:
        return None
    p = parts[0]
    for q in parts[1:]:
        p = p & q
    return p


def run_both(num_qubits, x, z, c, circ: Circuit, direction="forward", eps=None, max_weight=None, topn=None,
             approx_topn=None):
    s = po.OracleSum(num_qubits, x, z, c)
    olens = s.propagate(circ.channels(), direction, oracle_policy(eps, max_weight, topn, approx_topn))
    ox, oz, oc, _ = s.export()
    host = PauliSum(num_qubits, x, z, c, dtype=c.dtype)
    stats = host.propagate_with_stats(circ, product_policy(eps, max_weight, topn, approx_topn), direction)
    return (ox, oz, oc, olens), (host.x_array(), host.z_array(), host.coefficients_array(), stats)


def assert_same(oracle, product):
    ox, oz, oc, olens = oracle
    px, pz, pc, stats = product
    assert len(pc) == len(oc), (len(pc), len(oc))
    assert np.array_equal(px, ox) and np.array_equal(pz, oz), "keys or canonical order differ"
    assert np.array_equal(pc, oc), f"coefficients differ: max |d| = {np.max(np.abs(pc.astype(float) - oc)) if len(oc) else 0}"
    assert [int(v) for v in olens] == stats.terms_out


def random_sum(rng, n, num_qubits, density, dtype):
    W = circuits.width_for(num_qubits)
    x = np.zeros((n, W), np.uint64)
    z = np.zeros((n, W), np.uint64)
    for w in range(W):
        live = min(64, num_qubits - 64 * w)
        if live <= 0:
            continue
        weights = np.uint64(1) << np.arange(live, dtype=np.uint64)
        x[:, w] = ((rng.random((n, live)) < density) * weights).sum(axis=1, dtype=np.uint64)
        z[:, w] = ((rng.random((n, live)) < density) * weights).sum(axis=1, dtype=np.uint64)
    return x, z, rng.standard_normal(n).astype(dtype)


def random_circuit(rng, num_qubits, layers, max_support=2):
    c = Circuit(num_qubits)
    for _ in range(layers):
        k = int(rng.integers(1, max_support + 1))
        qs = [int(q) for q in rng.choice(num_qubits, size=k, replace=False)]
        ps = "".join(rng.choice(list("XYZ"), size=k))
        theta = float(rng.choice([rng.uniform(-math.pi, math.pi), -math.pi / 2, math.pi / 2, math.pi]))
        c.pauli_rotation(ps, qs, theta)
    return c


def test_hand_computed_single_gates():
    announce("hand-computed single gates")
    t = 0.6
    cases = [
        ({"X": 1.0}, "Z", [0], t, "forward", {"X": math.cos(t), "Y": math.sin(t)}),
        ({"Z": 1.0}, "X", [0], t, "forward", {"Z": math.cos(t), "Y": -math.sin(t)}),
        ({"XI": 1.0}, "ZZ", [0, 1], t, "forward", {"XI": math.cos(t), "YZ": math.sin(t)}),
        ({"XI": 1.0}, "ZZ", [0, 1], -math.pi / 2, "forward", {"YZ": -1.0}),
        ({"X": 1.0}, "Z", [0], t, "heisenberg", {"X": math.cos(t), "Y": -math.sin(t)}),
    ]
    for terms, paulis, qubits, theta, direction, want in cases:
        n = len(next(iter(terms)))
        s = PauliSum.from_strings(terms, num_qubits=n)
        s.propagate(Circuit(n).pauli_rotation(paulis, qubits, theta), direction=direction)
        got = {}
        for (xw, zw), v in s.to_dict().items():
            got["".join("IXZY"[((xw[0] >> q) & 1) | (((zw[0] >> q) & 1) << 1)] for q in range(n))] = v
        assert got == want


def test_rotation_amplitudes_bitwise():
    """cos and sin of 1024 random angles as the layer applies them: X_q under RZ(theta_q) becomes
    cos X_q + sin Y_q, so the coefficients are the host libm's values, compared bitwise with the oracle's."""
    announce("rotation amplitudes of 1024 angles")
    rng = np.random.default_rng(17)
    x = np.array([[1 << q] for q in range(64)], np.uint64)
    z = np.zeros((64, 1), np.uint64)
    c = np.ones(64)
    for _ in range(16):
        circ = Circuit(64)
        for q in range(64):
            circ.rz(float(rng.uniform(-4.0, 4.0)), q)
        o, p = run_both(64, x, z, c, circ, "forward")
        assert_same(o, p)


@pytest.mark.parametrize("num_qubits", [12, 64, 65, 127])
@pytest.mark.parametrize("dtype", [F64, F32])
@pytest.mark.parametrize("direction", ["forward", "heisenberg"])
def test_random_circuits_bitwise(num_qubits, dtype, direction):
    announce(f"random circuit n={num_qubits} {np.dtype(dtype).name} {direction}")
    rng = np.random.default_rng(1000 + num_qubits)
    x, z, c = random_sum(rng, 3000, num_qubits, 0.15, dtype)
    circ = random_circuit(rng, num_qubits, 40)
    o, p = run_both(num_qubits, x, z, c, circ, direction, eps=1e-3, max_weight=num_qubits // 2)
    assert_same(o, p)


@pytest.mark.parametrize("num_qubits", [40, 127])
@pytest.mark.parametrize("dtype", [F64, F32])
@pytest.mark.parametrize("direction", ["forward", "heisenberg"])
def test_wide_generators_bitwise(num_qubits, dtype, direction):
    """Generators of weight 1..5 (a weight of 3 or more prepares as a non-local rotation of fanout 2 that never joins
    a rotation or permutation run, channel/rotation.rs:111-122), quarter and half turns among the angles, some
    straddling qubits 63/64, with a threshold, a weight cap and a TopN cap that bounds the runs: keys, coefficients
    and every layer's length against the oracle."""
    announce(f"wide generators n={num_qubits} {np.dtype(dtype).name} {direction}")
    rng = np.random.default_rng(211 + num_qubits)
    x, z, c = random_sum(rng, 3000, num_qubits, 0.12, dtype)
    circ = random_circuit(rng, num_qubits, 50, max_support=5)
    if num_qubits > 64:
        circ.pauli_rotation("XYZ", [62, 63, 64], 0.9).pauli_rotation("ZZXY", [63, 64, 65, 0], math.pi / 2)
        circ.pauli_rotation("YXZZX", [126, 64, 63, 1, 2], 2.2).pauli_rotation("ZZZ", [63, 64, 126], math.pi)
    o, p = run_both(num_qubits, x, z, c, circ, direction, eps=1e-4, max_weight=num_qubits // 2, topn=6000)
    assert_same(o, p)
    o, p = run_both(num_qubits, x, z, c, circ, direction, eps=1e-3)
    assert_same(o, p)


def clifford_run_circuit(rng, num_qubits, steps):
    """Long runs of one-to-one steps -- quarter turns (the Clifford arm) and half turns (key-preserving sign flips)
    on one or two qubits -- between single generic rotations."""
    c = Circuit(num_qubits)
    for i in range(steps):
        k = int(rng.integers(1, 3))
        qs = [int(q) for q in rng.choice(num_qubits, size=k, replace=False)]
        ps = "".join(rng.choice(list("XYZ"), size=k))
        if i % 97 == 96:
            theta = float(rng.uniform(-math.pi, math.pi))
        else:
            theta = float(rng.choice([math.pi / 2, -math.pi / 2, 3 * math.pi / 2, math.pi, -math.pi]))
        c.pauli_rotation(ps, qs, theta)
    return c


@pytest.mark.parametrize("num_qubits", [40, 127])
@pytest.mark.parametrize("dtype", [F64, F32])
def test_permutation_runs_walked_in_one_pass(num_qubits, dtype):
    """Runs of up to ~100 one-to-one steps, walked in one pass per run (and a run longer than one walk launch's
    256 steps), against the oracle stepping layer by layer: keys, coefficients and every layer's length."""
    announce(f"permutation runs n={num_qubits} {np.dtype(dtype).name}")
    rng = np.random.default_rng(71 + num_qubits)
    x, z, c = random_sum(rng, 4000, num_qubits, 0.1, dtype)
    circ = clifford_run_circuit(rng, num_qubits, 300)
    o, p = run_both(num_qubits, x, z, c, circ, "heisenberg", eps=1e-3, max_weight=num_qubits // 4)
    assert_same(o, p)
    long_run = Circuit(num_qubits)
    for i in range(600):
        q = int(rng.integers(0, num_qubits - 1))
        long_run.pauli_rotation("ZZ" if i % 2 else "XY", [q, q + 1], math.pi / 2 if i % 3 else -math.pi / 2)
    o, p = run_both(num_qubits, x, z, c, long_run, "forward", eps=1e-3, max_weight=num_qubits // 3)
    assert_same(o, p)


def test_permutation_runs_match_the_layer_by_layer_device_path(monkeypatch):
    """The one-pass walk and the per-layer join (PP_PERMUTATION_RUNS=layers) give the same sum bit for bit, and
    both the oracle's, on kicked-Ising steps whose 144 quarter-turn RZZ layers form one run each (peak 1.6M terms)."""
    announce("permutation runs: walk vs layers")
    rng = np.random.default_rng(5)
    x, z, c = random_sum(rng, 500, 127, 0.02, F64)
    circ = circuits.heavy_hex_kicked_ising(4, 0.8)
    o, walk = run_both(127, x, z, c, circ, "heisenberg", eps=3e-3, max_weight=12)
    assert_same(o, walk)
    monkeypatch.setenv("PP_PERMUTATION_RUNS", "layers")
    host = PauliSum(127, x, z, c)
    stats = host.propagate_with_stats(circ, product_policy(3e-3, 12), "heisenberg")
    assert np.array_equal(host.x_array(), walk[0]) and np.array_equal(host.z_array(), walk[1])
    assert np.array_equal(host.coefficients_array(), walk[2])
    assert stats.terms_out == walk[3].terms_out


@pytest.mark.parametrize("dtype", [F64, F32])
def test_untruncated_growth_bitwise(dtype):
    announce(f"untruncated growth {np.dtype(dtype).name}")
    rng = np.random.default_rng(7)
    x, z, c = random_sum(rng, 200, 20, 0.2, dtype)
    circ = random_circuit(rng, 20, 14)
    o, p = run_both(20, x, z, c, circ, "forward")
    assert_same(o, p)


def test_tfim_6x6_ten_steps():
    """The CPU-runnable configuration: 6x6 periodic TFIM, J=1, h=3, dt=0.1, 10 steps, threshold 1e-4 and
    weight 8, average X, read out in |+>^36."""
    announce("6x6 TFIM 10 steps")
    step = circuits.tfim_square_step(6, 6, 0.1, 1.0, 3.0)
    circ = circuits.repeat(step, 10)
    n = 36
    x = np.array([[1 << q] for q in range(n)], np.uint64)
    z = np.zeros((n, 1), np.uint64)
    c = np.full(n, 1.0 / n)
    o, p = run_both(n, x, z, c, circ, "heisenberg", eps=1e-4, max_weight=8)
    assert_same(o, p)
    s = po.OracleSum(n, x, z, c)
    s.propagate(circ.channels(), "heisenberg", oracle_policy(1e-4, 8))
    host = PauliSum(n, p[0], p[1], p[2])
    assert abs(host.expectation("x+") - s.expectation("x")) < 1e-12


GOLDEN = [r for r in golden_data.records() if r["peak_terms"] <= 700_000]


@pytest.mark.parametrize("rec", GOLDEN[::3], ids=lambda r: f"{r['observable']}-{r['theta_h_label']}-{r['trotter_steps']}"
                         f"-{r['min_abs_coeff']}-{r['max_weight']}")
def test_golden_kicked_ising(rec):
    announce(f"golden {rec['observable']} {rec['theta_h_label']} steps={rec['trotter_steps']} "
             f"eps={rec['min_abs_coeff']} w={rec['max_weight']}")
    x, z, c = golden_data.observable(rec["observable"])
    circ = circuits.heavy_hex_kicked_ising(rec["trotter_steps"], rec["theta_h"])
    g = PauliSum(127, x, z, c).to_device(0)
    stats = g.propagate_with_stats(circ, product_policy(rec["min_abs_coeff"], rec["max_weight"]), "heisenberg")
    assert len(g) == rec["final_terms"]
    assert stats.peak_terms == rec["peak_terms"]
    assert abs(g.expectation("z+") - rec["expectation_value"]) <= 1e-12 * max(1.0, abs(rec["expectation_value"]))
    g.free()


@pytest.mark.parametrize("n", [1, 700, 2047, 5000])
def test_topn_ties_in_canonical_order(n):
    announce(f"topn ties n={n}")
    rng = np.random.default_rng(3)
    x, z, _ = random_sum(rng, 12000, 40, 0.3, F64)
    c = rng.choice([0.5, -0.5, 0.25, 0.125, -0.125, 1e-3], size=12000)
    s = po.OracleSum(40, x, z, c)
    s.topn(n)
    ox, oz, oc, _ = s.export()
    g = PauliSum(40, x, z, c).to_device(0)
    g.truncate_topn(n)
    h = g.to_host()
    g.free()
    assert np.array_equal(h.x_array(), ox) and np.array_equal(h.z_array(), oz)
    assert np.array_equal(h.coefficients_array(), oc)


@pytest.mark.parametrize("case", ["near_w1", "near_w2", "crowded_w1"])
def test_topn_boundary_bin_ranked_by_magnitude_then_canonical_order(case):
    """TopN decides the top 33 bits of |c|^2 on the device and ranks the boundary bin's terms on the host: here the
    bin holds |c|^2 values that differ only in the low mantissa bits, with exact repeats among them.  In the crowded
    case the bin holds over 2^16 terms, which the device narrows to equal |c|^2 before ranking."""
    announce(f"topn boundary bin {case}")
    rng = np.random.default_rng(41)
    nq, m = (40, 150000) if case == "crowded_w1" else (100 if case == "near_w2" else 40, 20000)
    x, z, _ = random_sum(rng, m, nq, 0.2, F64)
    if case == "crowded_w1":
        steps = rng.integers(0, 3, size=m)  # three |c| values 2^-40 apart, 50000 terms each
    else:
        steps = rng.integers(0, 400, size=m)  # 400 values 2^-40 apart, 50 terms each, in one 33-bit bin
    c = 0.5 * (1.0 + steps * 2.0 ** -40) * rng.choice([-1.0, 1.0], size=m)
    far = rng.random(m) < 0.3
    c[far] = rng.uniform(-1.0, 1.0, size=int(far.sum()))
    for n in (1, 777, m // 2, m - 7):
        s = po.OracleSum(nq, x, z, c)
        s.topn(n)
        ox, oz, oc, _ = s.export()
        g = PauliSum(nq, x, z, c).to_device(0)
        g.truncate_topn(n)
        h = g.to_host()
        g.free()
        assert np.array_equal(h.x_array(), ox) and np.array_equal(h.z_array(), oz), (case, n)
        assert np.array_equal(h.coefficients_array(), oc), (case, n)


def test_topn_inside_propagation():
    announce("topn inside a propagation")
    rng = np.random.default_rng(9)
    x, z, c = random_sum(rng, 64, 30, 0.1, F64)
    circ = random_circuit(rng, 30, 30)
    o, p = run_both(30, x, z, c, circ, "heisenberg", eps=1e-6, topn=900)
    assert_same(o, p)


@pytest.mark.parametrize("num_qubits,dtype", [(40, F64), (127, F64), (127, F32)])
def test_topn_every_run_with_every_layer_kind(num_qubits, dtype):
    """TopN after every run, where the next run is a join layer (which reads its input through the selection), a
    key-preserving layer or a permutation run (both apply the selection first), or a refinement of the device
    buckets: keys, coefficients and every layer's length against the oracle."""
    announce(f"topn every run n={num_qubits} {np.dtype(dtype).name}")
    rng = np.random.default_rng(83 + num_qubits)
    x, z, c = random_sum(rng, 3000, num_qubits, 0.1, dtype)
    circ = random_circuit(rng, num_qubits, 60)
    o, p = run_both(num_qubits, x, z, c, circ, "heisenberg", eps=1e-6, max_weight=num_qubits // 3, topn=2500)
    assert_same(o, p)
    o, p = run_both(num_qubits, x, z, c, circ, "forward", eps=1e-6, topn=20000)
    assert_same(o, p)


@pytest.mark.parametrize("topn", [0, 1, 2999, 3000])
def test_topn_at_its_edges_inside_a_propagation(topn):
    """TopN of 0 (clears the sum after the first run), 1, and just below and at the input size, inside a propagation
    (each selection's retain left to the next join): keys, coefficients and every layer's length."""
    announce(f"topn edge {topn}")
    rng = np.random.default_rng(97)
    x, z, c = random_sum(rng, 3000, 50, 0.1, F64)
    circ = random_circuit(rng, 50, 12)
    o, p = run_both(50, x, z, c, circ, "heisenberg", eps=1e-5, topn=topn)
    assert_same(o, p)


@pytest.mark.parametrize("topn", [1500, 100000])
def test_nan_coefficients_through_joins_and_topn(topn):
    """NaN coefficients (5% of the input, no threshold): the joins carry them as IEEE arithmetic does, TopN never keeps
    one (with its retain left to the next join), and with TopN inactive they reach the output: keys, values (NaN
    equal to NaN) and every layer's length against the oracle."""
    announce(f"nan coefficients topn={topn}")
    rng = np.random.default_rng(91)
    x, z, c = random_sum(rng, 2000, 30, 0.15, F64)
    c[rng.random(2000) < 0.05] = np.nan
    circ = random_circuit(rng, 30, 20)
    s = po.OracleSum(30, x, z, c)
    olens = s.propagate(circ.channels(), "heisenberg", oracle_policy(topn=topn))
    ox, oz, oc, _ = s.export()
    host = PauliSum(30, x, z, c)
    stats = host.propagate_with_stats(circ, product_policy(topn=topn), "heisenberg")
    assert np.array_equal(host.x_array(), ox) and np.array_equal(host.z_array(), oz)
    np.testing.assert_array_equal(host.coefficients_array(), oc)
    assert [int(v) for v in olens] == stats.terms_out
    if topn == 100000:
        assert np.isnan(oc).any()


@pytest.mark.parametrize("case", ["normal_w2_f64", "ties_w1_f64", "normal_w2_f32", "big_w2_f64"])
def test_approx_topn_standalone(case):
    announce(f"approx topn {case}")
    rng = np.random.default_rng(31)
    if case == "ties_w1_f64":
        nq, m = 40, 12000
        x, z, _ = random_sum(rng, m, nq, 0.3, F64)
        c = rng.choice([0.5, -0.5, 0.25, 0.125, -0.125, 1e-3, 0.75], size=m)
    elif case == "big_w2_f64":  # many reference buckets: the quota runs out inside one of them
        nq, m = 100, 300000
        x, z, c = random_sum(rng, m, nq, 0.1, F64)
    else:
        nq, m = 127, 20000
        x, z, c = random_sum(rng, m, nq, 0.1, F32 if case.endswith("f32") else F64)
    for n in (0, 1, 777, m // 3, m - 1, m, m + 5):
        s = po.OracleSum(nq, x, z, c)
        s.approx_topn(n)
        ox, oz, oc, _ = s.export()
        g = PauliSum(nq, x, z, c, dtype=c.dtype).to_device(0)
        g.truncate_approx_topn(n)
        h = g.to_host()
        g.free()
        assert len(h.coefficients_array()) == min(n, m), (case, n)
        assert np.array_equal(h.x_array(), ox) and np.array_equal(h.z_array(), oz), (case, n)
        assert np.array_equal(h.coefficients_array(), oc), (case, n)


def test_topn_after_approx_topn_on_one_sum():
    """ApproxTopN and TopN share the sum's histogram buffer; each selection starts from a clean one: alternating
    passes on one resident sum keep exactly what the oracle keeps."""
    announce("topn after approx topn on one sum")
    rng = np.random.default_rng(37)
    nq, m = 127, 20000
    x, z, c = random_sum(rng, m, nq, 0.1, F64)
    s = po.OracleSum(nq, x, z, c)
    g = PauliSum(nq, x, z, c).to_device(0)
    for kind, n in (("topn", 15000), ("approx", 9000), ("topn", 5000), ("approx", 3000), ("topn", 1234)):
        getattr(s, "topn" if kind == "topn" else "approx_topn")(n)
        if kind == "topn":
            g.truncate_topn(n)
        else:
            g.truncate_approx_topn(n)
        ox, oz, oc, _ = s.export()
        h = g.to_host()
        assert np.array_equal(h.x_array(), ox) and np.array_equal(h.z_array(), oz), (kind, n)
        assert np.array_equal(h.coefficients_array(), oc), (kind, n)
    g.free()


@pytest.mark.parametrize("dtype", [F64, F32])
def test_approx_topn_inside_propagation(dtype):
    announce(f"approx topn inside a propagation {dtype.__name__}")
    rng = np.random.default_rng(19)
    x, z, c = random_sum(rng, 64, 30, 0.1, dtype)
    circ = random_circuit(rng, 30, 30)
    o, p = run_both(30, x, z, c, circ, "heisenberg", eps=1e-6, approx_topn=900)
    assert_same(o, p)


@pytest.mark.parametrize("stage_rows", [0, 64])
@pytest.mark.parametrize("paulis", ["Z", "ZX"])
def test_every_row_anticommuting(paulis, stage_rows):
    """Every term anticommutes with the generator, so every bucket pair stages all of its rows; with 64 staging rows
    that is more than the first join launch holds per workgroup, which leaves every pair to the redo launch."""
    announce(f"all rows anticommuting {paulis} stage_rows={stage_rows}")
    from paulistrings_amd import _native

    ctx = _native.Context.get(0)
    saved = ctx.get_tuning()
    ctx.set_tuning(saved.bucket_target, stage_rows, saved.groups_per_cu, saved.join_block)
    try:
        _every_row_anticommuting(paulis)
    finally:
        ctx.set_tuning(saved.bucket_target, saved.stage_rows, saved.groups_per_cu, saved.join_block)


def _every_row_anticommuting(paulis):
    rng = np.random.default_rng(23)
    nq, m = 40, 1 << 17
    x, z, c = random_sum(rng, m, nq, 0.3, F64)
    x[:, 0] |= np.uint64(1)  # X or Y on qubit 0: anticommutes with Z0 and with Z0 X1 when qubit 1 commutes
    z[:, 0] &= ~np.uint64(2)
    x[:, 0] &= ~np.uint64(2)
    circ = Circuit(nq).pauli_rotation(paulis, [0, 1][: len(paulis)], 0.45)
    o, p = run_both(nq, x, z, c, circ, "forward", eps=1e-9)
    assert_same(o, p)


def test_edge_cases():
    announce("edge cases")
    # empty sum
    e = PauliSum(10)
    e.propagate(Circuit(10).rx(0.3, 2))
    assert len(e) == 0
    # the identity generator and quarter turns
    rng = np.random.default_rng(4)
    x, z, c = random_sum(rng, 500, 10, 0.3, F64)
    circ = Circuit(10).pauli_rotation("", [], 0.7).rz(math.pi, 3).rx(-math.pi / 2, 4).rzz(math.pi / 2, 0, 9)
    o, p = run_both(10, x, z, c, circ, "forward", eps=1e-9)
    assert_same(o, p)
    # generators straddling the word boundary, the last live qubit of W = 2
    x, z, c = random_sum(rng, 2000, 128, 0.05, F64)
    circ = Circuit(128).pauli_rotation("XZ", [63, 64], 0.4).pauli_rotation("YY", [127, 0], 1.1).rx(0.2, 127)
    o, p = run_both(128, x, z, c, circ, "forward")
    assert_same(o, p)


def test_saturated_layer_reduced():
    """The saturated-layer workload at 2^16 keys: 127 qubits, bits Bernoulli(0.1), N(0,1) coefficients, the first
    RZZ and RX of the dense layer, threshold 1e-6, f64 and f32."""
    announce("saturated layer 2^16")
    rng = np.random.default_rng(0xC0FFEE)
    for dtype in (F64, F32):
        x, z, c = random_sum(rng, 1 << 16, 127, 0.1, dtype)
        circ = Circuit(127).rzz(float(rng.uniform(0, math.pi)), 0, 1).rx(float(rng.uniform(0, math.pi)), 0)
        o, p = run_both(127, x, z, c, circ, "forward", eps=1e-6)
        assert_same(o, p)
