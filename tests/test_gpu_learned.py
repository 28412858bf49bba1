"""The learned truncation policy's device kernels against the CPU restatement (oracle/learned.py): the bf16 MFMA
forward (which also checks the MFMA operand layout the kernels assume), the backward, a training run that learns
a teacher network, and the learned TopN pass's selection.  Each case prints its name before it touches the GPU."""

from __future__ import annotations

import sys

import numpy as np
import pytest

import learned as L
from paulistrings_amd import PauliSum
from paulistrings_amd import learned as PL

pytestmark = pytest.mark.gpu


def announce(msg: str) -> None:
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def samples(rng, n):
    x = rng.integers(0, 2**63, size=n, dtype=np.uint64) & rng.integers(0, 2**63, size=n, dtype=np.uint64)
    z = rng.integers(0, 2**63, size=n, dtype=np.uint64) & rng.integers(0, 2**63, size=n, dtype=np.uint64)
    c = rng.standard_normal(n) * np.exp(rng.uniform(-12, 0, n))
    rem = rng.integers(0, 60, size=n).astype(np.float32)
    return x, z, c, rem


def oracle_params(m: PL.Model) -> L.Params:
    return L.Params.from_flat(m.get_params())


def test_forward_matches_the_oracle():
    announce("mlp forward")
    rng = np.random.default_rng(1)
    groups = L.lattice_groups(8, 8)
    m = PL.Model(groups, depth_scale=40.0, seed=3)
    p = oracle_params(m)
    p.b1 = (rng.standard_normal(L.HID) * 0.1).astype(np.float32)
    p.b2 = (rng.standard_normal(L.HID) * 0.1).astype(np.float32)
    p.b3 = np.float32(0.3)
    m.set_params(p.flat())
    x, z, c, rem = samples(rng, 1000)
    got = m.predict(x, z, c, rem)
    want, _ = L.forward(p, L.features(x, z, c, groups, rem, 40.0))
    err = np.abs(got - want)
    assert np.median(err) <= 1e-4 * (1 + np.median(np.abs(want)))
    assert (err <= 2e-2 * (1 + np.abs(want))).all(), float(err.max())
    m.free()


def test_gradient_matches_the_oracle():
    announce("mlp backward")
    rng = np.random.default_rng(2)
    groups = L.lattice_groups(6, 6)
    m = PL.Model(groups, depth_scale=20.0, seed=5)
    p = oracle_params(m)
    x, z, c, rem = samples(rng, 700)
    x &= np.uint64((1 << 36) - 1)
    z &= np.uint64((1 << 36) - 1)
    labels = rng.standard_normal(700).astype(np.float32) * 3
    g, loss = m.gradient(x, z, c, rem, labels)
    f = L.features(x, z, c, groups, rem, 20.0)
    want = L.backward(p, f, labels)
    assert abs(loss - L.loss(p, f, labels)) <= 1e-3 * (1 + loss)
    gp = L.Params.from_flat(g)
    for name in ("W1", "b1", "W2", "b2", "w3"):
        a, b = getattr(gp, name).astype(np.float64), getattr(want, name).astype(np.float64)
        assert np.linalg.norm(a - b) <= 2e-2 * np.linalg.norm(b) + 1e-7, name
    assert abs(float(gp.b3) - float(want.b3)) <= 1e-3 * (1 + abs(float(want.b3)))
    m.free()


def test_training_learns_a_teacher():
    announce("mlp training")
    rng = np.random.default_rng(3)
    groups = L.lattice_groups(8, 8)
    teacher = L.Params(11)
    x, z, c, rem = samples(rng, 8192)
    labels = L.forward(teacher, L.features(x, z, c, groups, rem, 40.0))[0].astype(np.float32)
    m = PL.Model(groups, depth_scale=40.0, seed=4)
    _, before = m.gradient(x, z, c, rem, labels)
    after = m.train(x, z, c, rem, labels, epochs=30, batch=512, lr=2e-3, seed=9)
    assert after < 0.1 * before, (before, after)
    m.free()


def test_training_is_deterministic():
    """The same samples, initialization and shuffle seed give the same weights bit for bit: gradients are summed per
    64-sample tile and then over tiles in a fixed order (a batch of 512 spans 8 tiles), never by float atomics."""
    announce("mlp training determinism")
    rng = np.random.default_rng(6)
    groups = L.lattice_groups(8, 8)
    x, z, c, rem = samples(rng, 3000)
    labels = rng.standard_normal(3000).astype(np.float32)
    out = []
    for _ in range(2):
        m = PL.Model(groups, depth_scale=40.0, seed=7)
        loss = m.train(x, z, c, rem, labels, epochs=5, batch=512, lr=1e-3, seed=8)
        g, gl = m.gradient(x, z, c, rem, labels)
        out.append((m.get_params(), loss, g, gl))
        m.free()
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][1] == out[1][1] and out[0][3] == out[1][3]
    assert np.array_equal(out[0][2], out[1][2])


@pytest.mark.parametrize("n", [1, 500, 3000])
def test_learned_pass_keeps_the_top_scores(n):
    announce(f"learned pass n={n}")
    rng = np.random.default_rng(4)
    groups = L.lattice_groups(8, 8)
    m = PL.Model(groups, depth_scale=40.0, seed=6)
    x, z, c, _ = samples(rng, 6000)
    s = PauliSum(64, x.reshape(-1, 1), z.reshape(-1, 1), c)
    g = s.to_device(0)
    h = g.to_host()  # the canonical order
    hx, hz, hc = h.x_array()[:, 0], h.z_array()[:, 0], h.coefficients_array()
    # the pass ranks by ln|c| + the network's output, summed in fp32
    score = m.predict(hx, hz, hc, np.float32(12)) + np.log(np.abs(hc)).astype(np.float32)
    key = score.astype(np.float64)
    order = sorted(range(len(key)), key=lambda i: -key[i])  # stable: ties keep the canonical order
    keep = sorted(order[:n])
    PL.truncate_learned(g, m, n, 12)
    got = g.to_host()
    g.free()
    m.free()
    assert np.array_equal(got.x_array()[:, 0], hx[keep]) and np.array_equal(got.z_array()[:, 0], hz[keep])
    assert np.array_equal(got.coefficients_array(), hc[keep])


@pytest.mark.parametrize("clifford_block", [False, True])
def test_propagate_learned_matches_a_layer_by_layer_emulation(clifford_block):
    """pp_propagate_learned against the oracle stepped one channel at a time, with the learned pass emulated on the
    host: after every layer that leaves more than `keep` terms, keep the `keep` of largest score, the score being the
    device network's prediction for the oracle's terms (pp_model_predict, the same per-term arithmetic as the pass's
    score kernel) plus ln|c| in fp32.  While a run's worst-case growth stays within `keep` the device joins layers
    into runs without a pass, and then no layer of the run exceeds `keep` either, so the two agree layer by layer.
    Kept keys and coefficients must be equal bit for bit.  With clifford_block, a block of 24 quarter-turn ZZ layers
    between the two steps forms a permutation run, walked in one pass on the device."""
    import pp_oracle as po
    from paulistrings_amd import truncation
    from paulistrings_amd.circuits import repeat, tfim_square_step

    announce(f"propagate_learned against a layer-by-layer emulation clifford_block={clifford_block}")
    n_q, keep, eps = 16, 300, 1e-7
    step = tfim_square_step(4, 4, dt=0.1, j=1.0, h=2.0)
    circ = repeat(step, 2)
    if clifford_block:
        from paulistrings_amd import Circuit
        block = Circuit(n_q)
        for q in range(n_q):
            if q % 4 != 3:
                block.rzz(np.pi / 2 if q % 2 else -np.pi / 2, q, q + 1)
            if q < 12:
                block.rzz(np.pi / 2, q, q + 4)
        circ = Circuit(n_q).extend(step).extend(block).extend(step)
    chans = circ.channels()
    m = PL.Model(PL.lattice_groups(4, 4), depth_scale=48.0, seed=21)
    x0 = (np.uint64(1) << np.arange(n_q, dtype=np.uint64)).reshape(-1, 1)
    z0 = np.zeros((n_q, 1), np.uint64)
    c0 = np.full(n_q, 1.0 / n_q)

    g = PauliSum(n_q, x0, z0, c0).to_device(0)
    PL.propagate_learned(g, circ, m, keep, truncation.coeff(eps), "heisenberg")
    got = g.to_host()
    g.free()

    x, z, c = x0.copy(), z0.copy(), c0.copy()
    passes = 0
    for k, j in enumerate(range(len(chans) - 1, -1, -1)):  # Heisenberg order: the last channel first
        o = po.OracleSum(n_q, x, z, c)
        o.propagate(chans[j:j + 1], "heisenberg", po.coeff_policy(eps=eps))
        x, z, c, _ = o.export()
        if len(c) > keep:
            remaining = np.float32(len(chans) - (k + 1))
            score = m.predict(x[:, 0], z[:, 0], c, remaining) + np.log(np.abs(c)).astype(np.float32)
            order = np.argsort(-score.astype(np.float64), kind="stable")  # ties: the canonical order
            kept = np.sort(order[:keep])
            x, z, c = x[kept], z[kept], c[kept]
            passes += 1
    m.free()
    assert passes > 10  # the pass did decide most layers
    got_terms = {(int(a), int(b)): float(v) for a, b, v in zip(got.x_array()[:, 0], got.z_array()[:, 0],
                                                               got.coefficients_array())}
    want_terms = {(int(a), int(b)): float(v) for a, b, v in zip(x[:, 0], z[:, 0], c)}
    assert got_terms == want_terms  # same keys, coefficients equal as doubles
