"""Upload and download through the pinned, chunked staging (two 32 MiB host chunks that alternate): sums whose word
planes span several chunks, and a chunk boundary that falls inside a plane, come back with every key and coefficient
intact, in the canonical order the device path defines (the same order as the host constructor's)."""

from __future__ import annotations

import sys

import numpy as np
import pytest

from paulistrings_amd import PauliSum
from test_gpu_parity import random_sum

pytestmark = pytest.mark.gpu


def canonical(x, z, c):
    keys = [z[:, w] for w in reversed(range(z.shape[1]))] + [x[:, w] for w in reversed(range(x.shape[1]))]
    order = np.lexsort(keys)
    return x[order], z[order], c[order]


@pytest.mark.parametrize("n,num_qubits,dtype", [(5_000_003, 127, np.float64), (9_000_001, 64, np.float32),
                                                (1, 100, np.float64)])
def test_round_trip_through_the_staging_chunks(n, num_qubits, dtype):
    print(f"[gpu case] staging round trip n={n} qubits={num_qubits} {np.dtype(dtype).name}", flush=True)
    sys.stdout.flush()
    rng = np.random.default_rng(n)
    x, z, c = random_sum(rng, n, num_qubits, 0.1, dtype)
    host = PauliSum(num_qubits, x, z, c, dtype=dtype)  # duplicate draws summed into one term
    g = host.to_device(0)
    try:
        h = g.to_host()
    finally:
        g.free()
    assert len(h) == len(host) > 0.99 * n
    ex, ez, ec = canonical(host._x, host._z, host._c)
    gx, gz, gc = canonical(h._x, h._z, h._c)
    assert np.array_equal(gx, ex) and np.array_equal(gz, ez) and np.array_equal(gc, ec)
