"""pp_apply_rotation_to_noalloc: the out-of-place rotation layer with its scratch (offset-scan tiles, counters, redo
queue) in the library's fixed per-GPU array instead of allocations.  Through the C ABI it must give exactly what
pp_apply_rotation_to gives, on two calls in a row, the second with a smaller input; an output without room and an
input whose device bucketing is not settled are refused on the host, before anything is queued."""

from __future__ import annotations

import ctypes
import math
import sys

import numpy as np
import pytest

from paulistrings_amd import Circuit, PauliSum, _native, truncation
from test_gpu_parity import random_sum

pytestmark = pytest.mark.gpu


def announce(msg: str) -> None:
    print(f"[gpu case] {msg}", flush=True)
    sys.stdout.flush()


def noalloc(src, dst, ch, policy):
    L = _native.lib()
    c = src._channel(ch)
    pol = truncation.lower(policy)
    return L.pp_apply_rotation_to_noalloc(src._h, dst._h, ctypes.byref(c), 0, ctypes.byref(pol))


def terms(g):
    h = g.to_host()
    return h._x, h._z, h._c


@pytest.mark.parametrize("num_qubits,dtype", [(127, np.float64), (40, np.float32)])
def test_noalloc_layer_equals_the_allocating_one(num_qubits, dtype):
    rng = np.random.default_rng(4040 + num_qubits)
    policy = truncation.coeff(1e-4)
    sizes = [1 << 17, 1 << 14]  # the second call has the smaller input
    # outputs sized once by the allocating call from the larger input
    x, z, c = random_sum(rng, sizes[0], num_qubits, 0.1, dtype)
    big = PauliSum(num_qubits, x, z, c, dtype=dtype).to_device(0)
    out_a = big.clone()
    out_b = big.clone()
    chans = [Circuit(num_qubits).rzz(0.7, 3, 9).channels(), Circuit(num_qubits).rx(1.1, num_qubits - 1).channels(),
             Circuit(num_qubits).pauli_rotation("XYZ", [0, 5, num_qubits - 2], -math.pi / 2).channels()]
    try:
        for ch in chans:
            big.apply_rotation_to(out_a, ch, policy=policy)
            big.apply_rotation_to(out_b, ch, policy=policy)  # sizes out_b's columns and buckets
        for call, n in enumerate(sizes):
            announce(f"noalloc layer n={num_qubits} {np.dtype(dtype).name} call {call} ({n} terms)")
            xs, zs, cs = (x, z, c) if call == 0 else random_sum(rng, n, num_qubits, 0.1, dtype)
            src = PauliSum(num_qubits, xs, zs, cs, dtype=dtype).to_device(0)
            try:
                for ch in chans:
                    src.apply_rotation_to(out_a, ch, policy=policy)
                    want = terms(out_a)
                    assert noalloc(src, out_b, ch, policy) == 0, _native.lib().pp_last_error(src.ctx.handle)
                    got = terms(out_b)
                    assert len(got[2]) == len(want[2]) > 0
                    assert all(np.array_equal(a, b) for a, b in zip(got, want))
            finally:
                src.free()
    finally:
        big.free()
        out_a.free()
        out_b.free()


def test_noalloc_refuses_an_output_without_room():
    announce("noalloc layer refuses an output without room")
    rng = np.random.default_rng(11)
    x, z, c = random_sum(rng, 1 << 15, 60, 0.1, np.float64)
    src = PauliSum(60, x, z, c).to_device(0)
    small = PauliSum.from_strings({"X" + "I" * 59: 1.0}, num_qubits=60).to_device(0)
    try:
        ch = Circuit(60).rzz(0.4, 1, 2).channels()
        assert noalloc(src, small, ch, None) == 5  # PP_ERR_INVALID_ARGUMENT, nothing queued
        assert len(small) == 1 and len(src) == len(c)
    finally:
        src.free()
        small.free()
