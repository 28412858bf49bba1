"""The C ABI library builds, loads and exports every entry point include/pp_hip.h declares (no GPU needed),
and the host halves of the product (hash rows, policy lowering, circuits) agree with the oracle."""

from __future__ import annotations

import ctypes
import math
import re
from pathlib import Path

import numpy as np
import pytest

import pp_oracle as po
from paulistrings_amd import Circuit, PauliSum, _native, circuits, truncation

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "pp_hip.h").read_text()
    return sorted(set(re.findall(r"\b(pp_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    syms = declared_symbols()
    assert "pp_propagate" in syms and "pp_sum_upload" in syms
    L = _native.lib()
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


def test_context_without_a_device_reports_no_device():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a device is visible")
    h = ctypes.c_void_p()
    assert _native.lib().pp_context_create(0, ctypes.byref(h)) == 1  # PP_ERR_NO_DEVICE


@pytest.mark.parametrize("n,width", [(1, 1), (5, 1), (36, 1), (64, 1), (100, 2), (127, 2)])
def test_product_hash_rows_equal_the_oracle(n, width):
    rx = np.zeros(31 * width, np.uint64)
    rz = np.zeros(31 * width, np.uint64)
    cnt = _native.lib().pp_hash_rows(n, 0x9E3779B97F4A7C15, width, rx.ctypes.data, rz.ctypes.data)
    ox, oz = po.hash_rows(n, 0x9E3779B97F4A7C15, width)
    assert cnt == len(ox)
    assert np.array_equal(rx[: cnt * width].reshape(cnt, width), ox)
    assert np.array_equal(rz[: cnt * width].reshape(cnt, width), oz)


def test_policy_lowering_is_postfix():
    p = (truncation.coeff(1e-4) & truncation.weight(8)).lower()
    assert [(p.nodes[i].op, p.nodes[i].k) for i in range(p.n_nodes)] == [(1, 0), (2, 8), (3, 0)]
    assert p.has_topn == 0
    q = (truncation.coeff(1e-5) & truncation.topn(1 << 24)).lower()
    assert q.n_nodes == 1 and q.has_topn == 1 and q.topn == 1 << 24
    with pytest.raises(ValueError):
        truncation.coeff(1e-3) | truncation.topn(5)
    r = (truncation.coeff(1e-2) | truncation.weight(3)).lower()
    assert [r.nodes[i].op for i in range(r.n_nodes)] == [1, 2, 4]
    a = (truncation.coeff(1e-5) & truncation.approx_topn(1000) & truncation.approx_topn(700)).lower()
    assert a.n_nodes == 1 and a.has_topn == 2 and a.topn == 700
    with pytest.raises(ValueError):
        truncation.weight(3) | truncation.approx_topn(5)
    with pytest.raises(ValueError):
        (truncation.topn(5) & truncation.approx_topn(5)).lower()


def test_pauli_sum_construction_sums_duplicates_and_drops_zeros():
    s = PauliSum.from_strings({"XI": 1.0, "ZZ": 0.5}, num_qubits=2)
    x = np.array([[1], [1], [0]], np.uint64)
    z = np.array([[0], [0], [3]], np.uint64)
    t = PauliSum(2, x, z, np.array([0.25, -0.25, 2.0]))
    assert len(s) == 2 and t.to_dict() == {((0,), (3,)): 2.0}


def test_repeated_keys_sum_in_input_order():
    """Long runs of one key sum one addition at a time, as the reference's hashmap accumulator and the oracle do."""
    rng = np.random.default_rng(1012)
    n = 3000
    x = (rng.integers(0, 8, size=(n, 1))).astype(np.uint64)
    z = (rng.integers(0, 4, size=(n, 1))).astype(np.uint64)
    c = rng.standard_normal(n)
    host = PauliSum(12, x, z, c)
    ox, oz, oc, _ = po.OracleSum(12, x, z, c).export()
    order = np.lexsort((oz[:, 0], ox[:, 0]))
    assert np.array_equal(host.x_array(), ox[order]) and np.array_equal(host.coefficients_array(), oc[order])


def test_circuit_encoding():
    c = Circuit(70).pauli_rotation("XZ", [69, 3], 0.5).rx(0.1, 64)
    ch = c.channels()
    assert int(ch[0]["gx"][1]) == 1 << 5 and int(ch[0]["gz"][0]) == 1 << 3 and ch[0]["theta"] == 0.5
    assert int(ch[1]["gx"][1]) == 1
    assert len(circuits.heavy_hex_kicked_ising(1, 0.1)) == 127 + 144
    assert len(circuits.tfim_square_step(6, 6, 0.1, 1.0, 3.0)) == 108
    with pytest.raises(ValueError):
        Circuit(4).rx(0.1, 4)
    assert math.isclose(circuits.tfim_square_step(2, 2, 0.05, 1.0, 1.0).channels()[-1]["theta"], 0.1)


def prepared_entries(p, k):
    """(local_delta, amplitudes) of a pp_prepared local table."""
    return [(int(p.deltas[e].local_delta), [float(v) for v in p.deltas[e].amplitude]) for e in range(p.n_deltas)]


@pytest.mark.parametrize("kind", [0, 1])
@pytest.mark.parametrize("adjoint", [False, True])
def test_prepared_rotation_tables_equal_the_conjugation_ptm(kind, adjoint):
    """pp_prepare_channel's table of exp(-i theta P / 2) on every one- and two-qubit generator: the entries and their
    zero pattern equal the transfer matrix U P_s U^dag computed from 2^k x 2^k matrices (adjoint: theta -> -theta),
    the values agree to 1e-14, and the values themselves are exactly the snapped glibc sin and cos (the oracle's,
    rounded to float32 for kind 1) with the matrices' signs.  Host-only: no device."""
    import ptm

    L = _native.lib()
    rng = np.random.default_rng(77 + kind + 2 * adjoint)
    nq = 70
    cases = []
    for k in (1, 2):
        for g in range(1, 1 << (2 * k)):
            if k == 2 and (g & 3 == 0 or g >> 2 == 0):
                continue  # weight 1: covered by k = 1
            qs = sorted(int(q) for q in rng.choice(nq, size=k, replace=False))
            for theta in (float(rng.uniform(-math.pi, math.pi)), math.pi / 2, -math.pi / 2, math.pi, 0.0):
                cases.append((k, qs, g, theta))
    for k, qs, g, theta in cases:
        ch = _native.Channel()
        for j, q in enumerate(qs):
            ch.gx[q // 64] |= ((g >> (2 * j)) & 1) << (q % 64)
            ch.gz[q // 64] |= ((g >> (2 * j + 1)) & 1) << (q % 64)
        ch.theta = theta
        p = _native.Prepared()
        assert L.pp_prepare_channel(nq, 2, kind, ctypes.byref(ch), 1 if adjoint else 0, ctypes.byref(p)) == 0
        assert p.kind == _native.PREPARED_LOCAL and p.k == k and list(p.qubits)[:k] == qs
        th = -theta if adjoint else theta
        gm = ptm.pauli_matrix(g, k)
        u = math.cos(th / 2) * np.eye(1 << k) - 1j * math.sin(th / 2) * gm
        want = ptm.entries_of(ptm.conjugation_ptm(u, k), k, clean=1e-14)
        got = prepared_entries(p, k)
        assert [d for d, _ in got] == [d for d, _ in want], (k, qs, g, theta)
        s, c = po.sin_cos(th)
        rnd = (lambda v: float(np.float32(v))) if kind == 1 else float
        for (d, a), (_, b) in zip(got, want):
            for pat in range(16):
                assert (a[pat] == 0.0) == (b[pat] == 0.0), (k, g, theta, d, pat)
                assert abs(a[pat] - b[pat]) < 1e-6 if kind == 1 else abs(a[pat] - b[pat]) < 1e-14
                if a[pat] != 0.0 and d != 0:
                    assert abs(a[pat]) == rnd(abs(s))  # the generator entry carries +-sin
                elif a[pat] != 0.0 and a[pat] != 1.0:
                    assert a[pat] == rnd(c)  # the identity entry carries cos where it acts


def test_prepared_wide_rotation_keeps_cos_and_sin():
    L = _native.lib()
    ch = _native.Channel()
    ch.gx[0] = 0b1011
    ch.gz[1] = 1 << 3
    ch.theta = 0.7
    p = _native.Prepared()
    assert L.pp_prepare_channel(100, 2, 0, ctypes.byref(ch), 1, ctypes.byref(p)) == 0
    s, c = po.sin_cos(-0.7)
    assert p.kind == _native.PREPARED_ROTATION and p.sin == s and p.cos == c
    assert p.gx[0] == 0b1011 and p.gz[1] == 8
    ch.gx[1] = 1 << 40  # outside a 100-qubit register's second word
    assert L.pp_prepare_channel(100, 2, 0, ctypes.byref(ch), 0, ctypes.byref(p)) == 5
