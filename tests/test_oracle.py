"""Pin the CPU oracle before it is trusted as the checker of the HIP path.

Every check here is independent of the oracle's own formulas: the reference's golden
vectors and known answers, closed forms, brute force on 2^n x 2^n matrices, and the
reference's naive differential oracle restated separately.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

import golden_data
import pp_oracle as po
from paulistrings_amd import circuits as C

DEFAULT_SEED = 0x9E3779B97F4A7C15
MASK64 = (1 << 64) - 1


def py_mix64(z: int) -> int:
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
    return z ^ (z >> 31)


# ---- rng and hash ---------------------------------------------------------------------------


def test_splitmix64_known_answer():
    # crates/paulistrings/src/rng/tests.rs:3-8
    import ctypes

    state = ctypes.c_uint64(0)
    assert po.lib().ppo_splitmix64(ctypes.byref(state)) == 0xE220A8397B1DCDAF


def test_hash_rows_at_127_qubits_are_the_plain_draw():
    # pauli_sum/hash/tests.rs:209-221: over 127 qubits no row is redrawn.
    rx, rz = po.hash_rows(127, DEFAULT_SEED, 2)
    assert rx.shape == (31, 2)
    gamma = 0x9E3779B97F4A7C15
    masks = [MASK64, (1 << 63) - 1]
    for i in range(31):
        for w in range(2):
            for half, rows in ((0, rx), (1, rz)):
                pos = (i << 16) | (w << 1) | half
                want = py_mix64((DEFAULT_SEED + gamma * (pos + 1)) & MASK64) & masks[w]
                assert int(rows[i, w]) == want, (i, w, half)


def gf2_rank(rows):
    rows = [int(r) for r in rows if r]
    rank = 0
    while rows:
        pivot = max(rows)
        top = pivot.bit_length() - 1
        rows = [r ^ pivot if (r >> top) & 1 else r for r in rows if r != pivot]
        rows = [r for r in rows if r]
        rank += 1
    return rank


@pytest.mark.parametrize("n", [1, 3, 5, 11, 16, 17, 64])
def test_hash_rows_are_full_rank(n):
    # pauli_sum/hash/tests.rs:196-206
    rx, rz = po.hash_rows(n, 0x1, 1)
    assert len(rx) == min(2 * n, 31)
    packed = [(int(rx[i, 0]) << 64) | int(rz[i, 0]) for i in range(len(rx))]
    assert gf2_rank(packed) == len(rx)


def independent_hash_rows(num_qubits: int, seed: int):
    """Gf2Hash::new's rows from their published rule (pauli_sum/hash.rs:22-26 row_word, :102-130 accept_rows), with
    a GF(2) elimination of its own: row r takes the first draw (attempt 0, 1, ... up to 4096) that is independent of
    the rows already taken.  Rows are Python ints, x bits in the low 64 * W and z bits above, masked to the live
    qubits.  Returns the rows and how many of them needed a redraw."""
    W = 1 if num_qubits <= 64 else 2
    gamma = 0x9E3779B97F4A7C15
    masks = [(1 << min(64, max(0, num_qubits - 64 * w))) - 1 for w in range(W)]
    rows = min(2 * num_qubits, 31)
    pivots: dict[int, int] = {}  # leading bit -> reduced row

    def reduce(v: int) -> int:
        while v:
            top = v.bit_length() - 1
            if top not in pivots:
                return v
            v ^= pivots[top]
        return 0

    out, redrawn = [], 0
    for r in range(rows):
        for attempt in range(4097):
            key = 0
            for w in range(W):
                for half in (0, 1):
                    pos = (attempt << 32) | (r << 16) | (w << 1) | half
                    word = py_mix64((seed + gamma * (pos + 1)) & MASK64) & masks[w]
                    key |= word << (64 * (half * W + w))
            red = reduce(key)
            if red or attempt == 4096:
                break
        if red:
            pivots[red.bit_length() - 1] = red
        redrawn += attempt > 0
        out.append(key)
    return out, W, redrawn


@pytest.mark.parametrize("seed", [DEFAULT_SEED, 0x1, 0xC0FFEE, 0xDEADBEEF12345678])
def test_hash_rows_with_rank_completion_redraws(seed):
    """The oracle's rows equal the independent restatement for every qubit count 1..32 (where draws are redrawn
    until independent) and at 36 and 64 qubits (the one-word configurations), and the redraw branch is taken."""
    total_redrawn = 0
    for n in list(range(1, 33)) + [36, 64]:
        want, W, redrawn = independent_hash_rows(n, seed)
        total_redrawn += redrawn
        rx, rz = po.hash_rows(n, seed, W)
        got = []
        for i in range(len(rx)):
            key = 0
            for w in range(W):
                key |= int(rx[i, w]) << (64 * w)
                key |= int(rz[i, w]) << (64 * (W + w))
            got.append(key)
        assert got == want, (n, seed)
    assert total_redrawn > 0


def test_rows_do_not_depend_on_the_width():
    # pauli_sum/hash/tests.rs:252-278
    for n in (1, 5, 64):
        r1 = po.hash_rows(n, DEFAULT_SEED, 1)
        r2 = po.hash_rows(n, DEFAULT_SEED, 2)
        assert np.array_equal(r1[0][:, 0], r2[0][:, 0]) and not r2[0][:, 1].any()
        assert np.array_equal(r1[1][:, 0], r2[1][:, 0]) and not r2[1][:, 1].any()


def test_bucket_ids_are_linear_and_canonical_order_holds():
    rng = np.random.default_rng(5)
    n = 20000
    x = rng.integers(0, 1 << 62, size=(n, 1), dtype=np.uint64)
    z = rng.integers(0, 1 << 62, size=(n, 1), dtype=np.uint64)
    s = po.OracleSum(62, x, z, rng.standard_normal(n))
    ex, ez, ec, eb = s.export()
    assert s.bits == po.lib().ppo_desired_bits(n, 1024, 128)
    rx, rz = po.hash_rows(62, DEFAULT_SEED, 1)

    def bucket(xx, zz):
        b = 0
        for r in range(s.bits):
            b |= (bin((int(xx) & int(rx[r, 0])) ^ (int(zz) & int(rz[r, 0]))).count("1") & 1) << r
        return b

    for i in range(0, n, 997):
        assert eb[i] == bucket(ex[i, 0], ez[i, 0])
        j = (i * 7 + 3) % n
        assert bucket(ex[i, 0] ^ ex[j, 0], ez[i, 0] ^ ez[j, 0]) == eb[i] ^ eb[j]
    keys = list(zip(eb.tolist(), ex[:, 0].tolist(), ez[:, 0].tolist()))
    assert keys == sorted(keys) and len(set(keys)) == n


def test_bucket_count_policy():
    # pauli_sum/storage.rs:15-41
    d = po.lib().ppo_desired_bits
    assert d(0, 1024, 128) == 0 and d(1024, 1024, 128) == 0 and d(1025, 1024, 128) == 1
    assert d(8191, 1024, 128) == 3 and d(8192, 1024, 128) == 7
    assert d(1 << 24, 1024, 128) == 14
    rb = po.lib().ppo_rebucketed_bits
    assert rb(5000, 3, 1024, 128) == 3  # inside the hysteresis gap
    assert rb(100, 3, 1024, 128) == 0  # sparse: coarsen
    assert rb(9000, 3, 1024, 128) == 7  # grow


def test_sin_cos_snaps_quarter_turns():
    # channel/rotation.rs:70-78
    assert po.sin_cos(-math.pi / 2) == (-1.0, 0.0)
    assert po.sin_cos(math.pi) == (0.0, -1.0)
    assert po.sin_cos(0.0) == (0.0, 1.0)
    s, c = po.sin_cos(0.6)
    assert s == math.sin(0.6) and c == math.cos(0.6)


# ---- hand-computed single gates -------------------------------------------------------------


def one_layer(num_qubits, terms, paulis, qubits, theta, direction="forward", dtype=np.float64):
    W = C.width_for(num_qubits)
    x = np.zeros((len(terms), W), np.uint64)
    z = np.zeros((len(terms), W), np.uint64)
    cs = np.zeros(len(terms), dtype)
    for i, (label, c) in enumerate(terms.items()):
        for q, ch in enumerate(label):
            if ch in "XY":
                x[i, q // 64] |= np.uint64(1 << (q % 64))
            if ch in "ZY":
                z[i, q // 64] |= np.uint64(1 << (q % 64))
        cs[i] = c
    circ = C.Circuit(num_qubits).pauli_rotation(paulis, qubits, theta)
    s = po.OracleSum(num_qubits, x, z, cs)
    s.propagate(circ.channels(), direction)
    ex, ez, ec, _ = s.export()
    out = {}
    for i in range(len(ec)):
        label = ""
        for q in range(num_qubits):
            bx = (int(ex[i, q // 64]) >> (q % 64)) & 1
            bz = (int(ez[i, q // 64]) >> (q % 64)) & 1
            label += "IXZY"[bx | (bz << 1)]
        out[label] = float(ec[i])
    return out


def test_hand_computed_phases():
    # engine/bucketed/rotation_layer/tests.rs:343-373, python tests test_pauli_rotation.py:60-140
    t = 0.6
    assert one_layer(1, {"X": 1.0}, "Z", [0], t) == {"X": math.cos(t), "Y": math.sin(t)}
    assert one_layer(1, {"Z": 1.0}, "X", [0], t) == {"Z": math.cos(t), "Y": -math.sin(t)}
    assert one_layer(2, {"XI": 1.0}, "ZZ", [0, 1], t) == {"XI": math.cos(t), "YZ": math.sin(t)}
    # the kicked-Ising bond at its Clifford point maps XI to exactly -YZ
    assert one_layer(2, {"XI": 1.0}, "ZZ", [0, 1], -math.pi / 2) == {"YZ": -1.0}
    # a Heisenberg layer is the adjoint: theta -> -theta
    assert one_layer(1, {"X": 1.0}, "Z", [0], t, "heisenberg") == {"X": math.cos(t), "Y": -math.sin(t)}


# ---- brute force on 2^n x 2^n matrices --------------------------------------------------------

PAULI = {
    (0, 0): np.eye(2, dtype=complex),
    (1, 0): np.array([[0, 1], [1, 0]], dtype=complex),
    (0, 1): np.array([[1, 0], [0, -1]], dtype=complex),
    (1, 1): np.array([[0, -1j], [1j, 0]], dtype=complex),
}


def pauli_matrix(n, xb, zb):
    m = np.ones((1, 1), dtype=complex)
    for q in range(n):
        m = np.kron(m, PAULI[((xb >> q) & 1, (zb >> q) & 1)])
    return m


def brute_force(n, terms, gates, direction):
    dim = 1 << n
    O = np.zeros((dim, dim), dtype=complex)
    for (xb, zb), c in terms.items():
        O += c * pauli_matrix(n, xb, zb)
    seq = gates if direction == "forward" else list(reversed(gates))
    for gx, gz, theta in seq:
        P = pauli_matrix(n, gx, gz)
        U = math.cos(theta / 2) * np.eye(dim) - 1j * math.sin(theta / 2) * P
        O = U @ O @ U.conj().T if direction == "forward" else U.conj().T @ O @ U
    out = {}
    for xb in range(dim):
        for zb in range(dim):
            v = np.trace(pauli_matrix(n, xb, zb) @ O) / dim
            if abs(v) > 1e-13:
                assert abs(v.imag) < 1e-12
                out[(xb, zb)] = v.real
    return out


@pytest.mark.parametrize("direction", ["forward", "heisenberg"])
@pytest.mark.parametrize("seed", range(6))
def test_untruncated_propagation_matches_brute_force(seed, direction):
    rng = np.random.default_rng(seed)
    n = 4
    terms = {}
    for _ in range(5):
        terms[(int(rng.integers(0, 16)), int(rng.integers(0, 16)))] = float(rng.standard_normal())
    gates = []
    circ = C.Circuit(n)
    for _ in range(12):
        k = int(rng.integers(1, 5))  # weight 3 and 4 generators prepare as non-local rotations
        qs = [int(q) for q in rng.choice(n, size=k, replace=False)]
        ps = "".join(rng.choice(list("XYZ"), size=k))
        theta = float(rng.uniform(-math.pi, math.pi))
        circ.pauli_rotation(ps, qs, theta)
        gx = gz = 0
        for p, q in zip(ps, qs):
            gx |= (p in "XY") << q
            gz |= (p in "ZY") << q
        gates.append((gx, gz, theta))
    want = brute_force(n, terms, gates, direction)
    x = np.array([[k[0]] for k in terms], np.uint64)
    z = np.array([[k[1]] for k in terms], np.uint64)
    s = po.OracleSum(n, x, z, np.array(list(terms.values())))
    s.propagate(circ.channels(), direction)
    ex, ez, ec, _ = s.export()
    got = {(int(ex[i, 0]), int(ez[i, 0])): float(ec[i]) for i in range(len(ec)) if abs(ec[i]) > 1e-13}
    assert set(got) == set(want)
    for k in want:
        assert abs(got[k] - want[k]) < 1e-12


# ---- golden vectors of the reference ----------------------------------------------------------


def run_record(rec, threads=0):
    circ = C.heavy_hex_kicked_ising(rec["trotter_steps"], rec["theta_h"]).channels()
    x, z, c = golden_data.observable(rec["observable"])
    s = po.OracleSum(127, x, z, c)
    pol = po.coeff_policy(eps=rec["min_abs_coeff"], max_weight=rec["max_weight"])
    lens = s.propagate(circ, "heisenberg", pol, threads=threads)
    peak = max(1, int(lens.max()) if len(lens) else 1)
    return s, peak


SMALL = [r for r in golden_data.records() if r["peak_terms"] <= 250_000]


def test_heavy_hex_graph_matches_the_reference_edge_list():
    assert C.heavy_hex_127_edges() == golden_data.heavy_hex_edges()
    assert [len(g) for g in C.heavy_hex_edge_coloring(golden_data.heavy_hex_edges())] == [54, 51, 39]


@pytest.mark.parametrize("rec", SMALL, ids=lambda r: f"{r['observable']}-{r['theta_h_label']}-{r['trotter_steps']}"
                         f"-{r['min_abs_coeff']}-{r['max_weight']}")
def test_golden_kicked_ising(rec):
    s, peak = run_record(rec)
    assert len(s) == rec["final_terms"]
    assert peak == rec["peak_terms"]
    assert abs(s.expectation("z") - rec["expectation_value"]) <= 1e-12 * max(1.0, abs(rec["expectation_value"]))


# ---- the naive differential oracle ---------------------------------------------------------------


def random_terms(rng, n_terms, num_qubits, density=0.3):
    W = C.width_for(num_qubits)
    x = np.zeros((n_terms, W), np.uint64)
    z = np.zeros((n_terms, W), np.uint64)
    for w in range(W):
        live = min(64, num_qubits - 64 * w)
        if live <= 0:
            continue
        bx = rng.random((n_terms, live)) < density
        bz = rng.random((n_terms, live)) < density
        weights = (np.uint64(1) << np.arange(live, dtype=np.uint64))
        x[:, w] = (bx * weights).sum(axis=1, dtype=np.uint64)
        z[:, w] = (bz * weights).sum(axis=1, dtype=np.uint64)
    return x, z


@pytest.mark.parametrize("num_qubits", [20, 100])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("theta", [0.37, -math.pi / 2, 2.9])
def test_bucketed_layer_equals_the_naive_oracle(num_qubits, dtype, theta):
    rng = np.random.default_rng(11)
    x, z = random_terms(rng, 30000, num_qubits, 0.2)
    c = rng.standard_normal(30000).astype(dtype)
    W = x.shape[1]
    s = po.OracleSum(num_qubits, x, z, c)
    x0, z0, c0, _ = s.export()
    wide = [0, 4, 11, 17, 19] if num_qubits < 64 else [62, 63, 64, 65, 99]
    for paulis, qubits in (("ZZ", [3, 7]), ("X", [5]), ("Y", [0]), ("XY", [num_qubits - 1, 1]),
                           ("XZY", wide[:3]), ("ZXXZY", wide)):
        ch = C.Circuit(num_qubits).pauli_rotation(paulis, qubits, theta).channels()
        pol = po.coeff_policy(eps=1e-3, max_weight=12)
        nx, nz, nc = po.naive_layer(W, x0, z0, c0, ch, False, pol)
        s.propagate(ch, "forward", pol)
        bx, bz, bc, _ = s.export()
        order = np.lexsort(tuple(bz[:, w] for w in reversed(range(W))) + tuple(bx[:, w] for w in reversed(range(W))))
        assert np.array_equal(bx[order], nx) and np.array_equal(bz[order], nz)
        assert np.array_equal(bc[order], nc)  # at most two contributions per key: bitwise
        x0, z0, c0 = bx, bz, bc


# ---- TopN --------------------------------------------------------------------------------------


def first_n_by_rank(c, n):
    """test_support/compare.rs:33-55: stable sort descending by |c|^2 bits, take n, input order."""
    rank = (c.astype(np.float64) ** 2).view(np.uint64)
    order = sorted(range(len(c)), key=lambda i: -int(rank[i]))
    return sorted(order[:n])


@pytest.mark.parametrize("n", [0, 1, 500, 4095, 5000, 10000])
def test_topn_keeps_the_first_ties_in_bucket_order(n):
    rng = np.random.default_rng(3)
    x, z = random_terms(rng, 12000, 40, 0.3)
    # few distinct magnitudes, both signs: many exact ties at any boundary
    c = rng.choice([0.5, -0.5, 0.25, 0.125, -0.125, 1e-3], size=12000)
    s = po.OracleSum(40, x, z, c)
    ex, ez, ec, _ = s.export()
    keep = first_n_by_rank(ec, n)
    s.topn(n)
    gx, gz, gc, _ = s.export()
    want = min(n, len(ec))
    assert len(gc) == want
    assert np.array_equal(gx, ex[keep]) and np.array_equal(gz, ez[keep]) and np.array_equal(gc, ec[keep])


def first_n_by_octave(c, n):
    """test_support/compare.rs:33-55 with octave_rank (:63-66): stable sort descending by the exponent bits of
    |c|^2, take n, input order.  ApproxTopN's kept set over a canonical (bucket-order) export."""
    rank = ((c.astype(np.float64) ** 2).view(np.uint64) >> np.uint64(52)).astype(np.int64)
    order = sorted(range(len(c)), key=lambda i: -int(rank[i]))
    return sorted(order[:n])


@pytest.mark.parametrize("case", ["random_w1", "tie_heavy_w2", "single_octave"])
def test_approx_topn_keeps_the_first_n_by_octave(case):
    # truncation/tests.rs:116-158: random sums, one magnitude group per octave, a sum inside one octave
    rng = np.random.default_rng(0xA9C7)
    if case == "random_w1":
        x, z = random_terms(rng, 2000, 32, 0.3)
        c = rng.standard_normal(2000)
        ns = [0, 1, 37, 300, 999, 1500, 2000, 2010]
        nq = 32
    elif case == "tie_heavy_w2":
        x, z = random_terms(rng, 2000, 100, 0.2)
        c = rng.choice([1.0, -0.5, 0.25, -0.125], size=2000)
        ns = [0, 3, 250, 700, 1200, 1900, 2500]
        nq = 100
    else:
        x = np.arange(1, 5, dtype=np.uint64).reshape(-1, 1)
        z = np.zeros((4, 1), np.uint64)
        c = np.array([1.0, 1.125, 1.25, 1.375])
        ns = [0, 1, 2, 3, 4]
        nq = 32
    for n in ns:
        s = po.OracleSum(nq, x, z, c)
        ex, ez, ec, _ = s.export()
        keep = first_n_by_octave(ec, n)
        s.approx_topn(n)
        gx, gz, gc, _ = s.export()
        assert len(gc) == min(n, len(ec)), (case, n)
        assert np.array_equal(gx, ex[keep]) and np.array_equal(gz, ez[keep]) and np.array_equal(gc, ec[keep])


def test_approx_topn_as_a_layer_pass():
    # The pass runs after every layer (or run): one layer then ApproxTopN(n) equals the unpassed layer cut alone.
    rng = np.random.default_rng(77)
    x, z = random_terms(rng, 5000, 40, 0.25)
    c = rng.standard_normal(5000)
    ch = C.Circuit(40).rx(0.9, 3).channels()
    a = po.OracleSum(40, x, z, c)
    a.propagate(ch, "forward", po.make_policy(approx_topn=3000))
    b = po.OracleSum(40, x, z, c)
    b.propagate(ch, "forward", po.make_policy())
    b.approx_topn(3000)
    ax, az, ac, _ = a.export()
    bx, bz, bc, _ = b.export()
    assert len(ac) == 3000
    assert np.array_equal(ax, bx) and np.array_equal(az, bz) and np.array_equal(ac, bc)


# ---- keep programs and read-out against independent predicates ----------------------------------


def _weight(x, z):
    return sum(bin(int(a) | int(b)).count("1") for a, b in zip(x, z))


@pytest.mark.parametrize("program", ["or", "and", "nested"])
def test_keep_programs_equal_a_filter_of_the_unfiltered_layer(program):
    # merge.rs:232-239: a merged term is emitted iff its sum is nonzero and keep_term holds, so one filtered layer
    # equals the unfiltered layer with the predicate applied afterwards (builtin.rs:19-59, :470-526).
    rng = np.random.default_rng(21)
    x, z = random_terms(rng, 4000, 30, 0.25)
    c = rng.standard_normal(4000) * np.exp(rng.uniform(-9, 0, 4000))
    eps, k = 3e-3, 7
    if program == "or":
        nodes = [(po.NODE_COEFF, 0, eps), (po.NODE_WEIGHT, k, 0.0), (po.NODE_OR, 0, 0.0)]
        pred = lambda cc, w: cc * cc > eps * eps or w <= k  # noqa: E731
    elif program == "and":
        nodes = [(po.NODE_WEIGHT, k, 0.0), (po.NODE_COEFF, 0, eps), (po.NODE_AND, 0, 0.0)]
        pred = lambda cc, w: cc * cc > eps * eps and w <= k  # noqa: E731
    else:  # (coeff(eps) | weight(k)) & weight(k + 3), and a disabled threshold
        nodes = [(po.NODE_COEFF, 0, eps), (po.NODE_WEIGHT, k, 0.0), (po.NODE_OR, 0, 0.0),
                 (po.NODE_WEIGHT, k + 3, 0.0), (po.NODE_AND, 0, 0.0), (po.NODE_COEFF, 0, -1.0), (po.NODE_AND, 0, 0.0)]
        pred = lambda cc, w: (cc * cc > eps * eps or w <= k) and w <= k + 3  # noqa: E731
    ch = C.Circuit(30).pauli_rotation("ZX", [2, 9], 0.7).channels()
    ux, uz, uc = po.naive_layer(1, x, z, c, ch, False, po.make_policy())
    keep = [i for i in range(len(uc)) if uc[i] != 0 and pred(float(uc[i]), _weight(ux[i], uz[i]))]
    s = po.OracleSum(30, x, z, c)
    s.propagate(ch, "forward", po.make_policy(nodes))
    bx, bz, bc, _ = s.export()
    order = np.lexsort((bz[:, 0], bx[:, 0]))
    assert np.array_equal(bx[order], ux[keep]) and np.array_equal(bz[order], uz[keep])
    assert np.array_equal(bc[order], uc[keep])


def _product_state(n, axis, negative):
    """|psi> = tensor of the +1 (or -1 where `negative`) eigenstates of `axis` on every qubit."""
    vecs = {"z": [np.array([1, 0], complex), np.array([0, 1], complex)],
            "x": [np.array([1, 1], complex) / math.sqrt(2), np.array([1, -1], complex) / math.sqrt(2)],
            "y": [np.array([1, 1j], complex) / math.sqrt(2), np.array([1, -1j], complex) / math.sqrt(2)]}[axis]
    psi = np.ones(1, complex)
    for q in range(n):
        psi = np.kron(psi, vecs[(negative >> q) & 1])
    return psi


@pytest.mark.parametrize("axis", ["x", "y", "z"])
@pytest.mark.parametrize("negative", [0, 0b1011])
def test_expectation_product_against_the_statevector(axis, negative):
    # readout/product_state.rs:84-101: <psi| sum c P |psi> for a product state of axis eigenstates
    n = 4
    rng = np.random.default_rng(5)
    terms = {}
    for _ in range(40):
        terms[(int(rng.integers(0, 16)), int(rng.integers(0, 16)))] = float(rng.standard_normal())
    psi = _product_state(n, axis, negative)
    want = 0.0
    for (xb, zb), c in terms.items():
        # pauli_matrix follows the qubit-0-first kron order of _product_state
        want += c * (psi.conj() @ pauli_matrix(n, xb, zb) @ psi).real
    x = np.array([[k[0]] for k in terms], np.uint64)
    z = np.array([[k[1]] for k in terms], np.uint64)
    s = po.OracleSum(n, x, z, np.array(list(terms.values())))
    got = s.expectation(axis, negative=[negative])
    assert abs(got - want) < 1e-12
