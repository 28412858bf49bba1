"""Batched propagation (pp_batch_propagate): many sums stepped together, one launch of each batched kernel per
layer step, each sum ending bitwise as its solo oracle run ends.  The workload is the batched recipe (qubits
16..64, brick-wall RZZ + RX/RZ circuits, weight <= 3 observables) at depths and thresholds that the oracle finishes
in seconds; a few instances carry a TopN or ApproxTopN pass or a key-preserving channel.  Each case prints its
name before it touches the GPU."""

from __future__ import annotations

import math
import sys

import numpy as np
import pytest

import pp_oracle as po
from paulistrings_amd import Circuit, PauliSum, circuits, propagate_batch, truncation

pytestmark = pytest.mark.gpu


def announce(msg: str) -> None:
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def check_batch(instances, policies_o, policies_p, direction="heisenberg"):
    gpus = []
    for (n, circ, (x, z, c), _eps) in instances:
        gpus.append(PauliSum(n, x, z, c).to_device(0))
    propagate_batch(gpus, [inst[1] for inst in instances], policies_p, [direction] * len(instances))
    for k, ((n, circ, (x, z, c), _eps), g) in enumerate(zip(instances, gpus)):
        s = po.OracleSum(n, x, z, c)
        s.propagate(circ.channels(), direction, policies_o[k])
        ox, oz, oc, _ = s.export()
        h = g.to_host()
        g.free()
        assert len(h.coefficients_array()) == len(oc), (k, len(h.coefficients_array()), len(oc))
        assert np.array_equal(h.x_array(), ox) and np.array_equal(h.z_array(), oz), k
        assert np.array_equal(h.coefficients_array(), oc), k


def test_batched_instances_equal_their_solo_runs():
    announce("batch of 24 brick-wall instances")
    inst = circuits.batched_instances(seed=7, count=24, depths=(3, 8), thresholds=(3e-4, 1e-2))
    # a key-preserving layer (RX(pi): its sine snaps to 0) in two of them
    for k in (3, 11):
        n, circ, obs, eps = inst[k]
        circ.rx(math.pi, 0)
    pol_o, pol_p = [], []
    for k, (_n, _c, _o, eps) in enumerate(inst):
        if k in (5, 11, 17):  # 11 also has a key-preserving layer after its last TopN's join layers
            pol_o.append(po.coeff_policy(eps=eps, topn=300))
            pol_p.append(truncation.coeff(eps) & truncation.topn(300))
        elif k == 6:
            pol_o.append(po.coeff_policy(eps=eps, approx_topn=250))
            pol_p.append(truncation.coeff(eps) & truncation.approx_topn(250))
        else:
            pol_o.append(po.coeff_policy(eps=eps))
            pol_p.append(truncation.coeff(eps))
    check_batch(inst, pol_o, pol_p)


def test_batched_permutation_runs_walked_per_instance():
    """Instances whose circuits carry runs of quarter-turn RZZ (permutation runs, walked in one pass per instance
    inside the batch) between their brick-wall layers, one of them under TopN: each equals its solo oracle run."""
    announce("batch with permutation runs")
    inst = circuits.batched_instances(seed=19, count=10, depths=(3, 6), thresholds=(3e-4, 1e-2))
    for k, (n, circ, obs, eps) in enumerate(inst):
        if k % 2 == 0:
            for q in range(0, n - 1, 3):
                circ.pauli_rotation("ZZ", [q, q + 1], -math.pi / 2)
            circ.rx(0.3, 1)
    pol_o = [po.coeff_policy(eps=eps, topn=200) if k == 4 else po.coeff_policy(eps=eps)
             for k, (_n, _c, _o, eps) in enumerate(inst)]
    pol_p = [truncation.coeff(eps) & truncation.topn(200) if k == 4 else truncation.coeff(eps)
             for k, (_n, _c, _o, eps) in enumerate(inst)]
    check_batch(inst, pol_o, pol_p)


def test_batch_of_wide_sums_forward():
    announce("batch of 3 sums on 100-127 qubits, forward")
    rng = np.random.default_rng(12)
    inst = []
    for n in (100, 127, 113):
        m = 20000
        x = np.zeros((m, 2), np.uint64)
        z = np.zeros((m, 2), np.uint64)
        for w in range(2):
            live = min(64, n - 64 * w)
            weights = np.uint64(1) << np.arange(live, dtype=np.uint64)
            x[:, w] = ((rng.random((m, live)) < 0.1) * weights).sum(axis=1, dtype=np.uint64)
            z[:, w] = ((rng.random((m, live)) < 0.1) * weights).sum(axis=1, dtype=np.uint64)
        c = Circuit(n)
        for _ in range(6):
            q = int(rng.integers(0, n - 1))
            c.rzz(float(rng.uniform(0, 2 * math.pi)), q, q + 1).rx(float(rng.uniform(0, 2 * math.pi)), int(rng.integers(0, n)))
        inst.append((n, c, (x, z, rng.standard_normal(m)), 1e-4))
    check_batch(inst, [po.coeff_policy(eps=1e-4)] * 3, [truncation.coeff(1e-4)] * 3, direction="forward")


@pytest.mark.timeout(900)
def test_batched_workload_at_its_own_parameters():
    """The batched workload as the benchmark draws it: 64 instances, 16..64 qubits, depths 5..30, thresholds
    log-uniform in [1e-6, 1e-3], Heisenberg, f64.  Uncapped its deepest instances grow past 10^9 terms, and the
    benchmark caps each at TopN 2^20; here the cap is 2^12 so that the oracle (whose TopN sorts every term after
    every run) finishes all 64 solo runs in well under a minute.  Every instance equals its solo oracle run."""
    announce("batch of the 64 benchmark instances (depths 5..30, thresholds 1e-6..1e-3), TopN 2^12")
    inst = circuits.batched_instances(seed=0xC0FFEE, count=64)
    cap = 1 << 12
    pol_o = [po.coeff_policy(eps=eps, topn=cap) for (_n, _c, _o, eps) in inst]
    pol_p = [truncation.coeff(eps) & truncation.topn(cap) for (_n, _c, _o, eps) in inst]
    check_batch(inst, pol_o, pol_p)


def test_batched_topn_selections_equal_one_by_one(monkeypatch):
    """The TopN selections of the instances that finish a run in the same step run together (PP_BATCH_TOPN, the
    default): one launch per radix pass for all of them, one gather of their boundary bins.  They keep the terms the
    one-by-one selections keep and the oracle keeps, including an instance whose boundary bin holds more than 2^16
    exact ties (which takes the one-by-one path) and one whose boundary bin is kept whole."""
    announce("batched TopN selections vs one by one")
    rng = np.random.default_rng(12)
    inst, pol_o, pol_p = [], [], []
    # exact ties: 2^18 terms of coefficient 1, half of them untouched by each RZ (so still 1 after it)
    n = 40
    m = 1 << 18
    x = np.unique(rng.integers(0, 1 << 40, size=m + 4096, dtype=np.uint64))[:m].reshape(-1, 1)
    z = np.zeros_like(x)
    circ = Circuit(n).rz(0.3, 0).rz(0.4, 1).rz(0.5, 2).rz(0.6, 3)
    inst.append((n, circ, (x, z, np.ones(len(x))), 1e-9))
    pol_o.append(po.coeff_policy(eps=1e-9, topn=100_000))
    pol_p.append(truncation.coeff(1e-9) & truncation.topn(100_000))
    # the rest: brick-wall instances at their caps
    for k, (nq, c, o, eps) in enumerate(circuits.batched_instances(seed=31, count=12, depths=(4, 9),
                                                                      thresholds=(1e-5, 1e-3))):
        cap = 150 + 40 * k
        inst.append((nq, c, o, eps))
        pol_o.append(po.coeff_policy(eps=eps, topn=cap))
        pol_p.append(truncation.coeff(eps) & truncation.topn(cap))
    check_batch(inst, pol_o, pol_p)
    monkeypatch.setenv("PP_BATCH_TOPN", "0")
    check_batch(inst, pol_o, pol_p)
