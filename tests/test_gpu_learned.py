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
