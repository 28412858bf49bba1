"""The learned policy's CPU restatement (oracle/learned.py) pinned on its own: bf16 rounding against known values,
features against hand counts, and backward against central finite differences of the loss (rounding off, float64
forward), so that the device kernels can be checked against it."""

from __future__ import annotations

import numpy as np
import pytest

import learned as L


def test_bf16_rounding_known_values():
    v = np.array([1.0, 1.00390625, 1.0029296875, 1.005859375, -2.5, 3.0e38, 1e-40], np.float32)
    r = L.bf16_round(v)
    # bf16 spacing at 1 is 2^-7.  1 + 2^-8 is halfway: ties to even -> 1; 1 + 3*2^-10 is below halfway -> 1;
    # 1 + 3*2^-9 is above halfway -> 1 + 2^-7
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 and r[3] == np.float32(1.0078125)
    assert r[4] == -2.5
    assert (r.view(np.uint32) & 0xFFFF == 0).all()


def test_features_count_rows_and_columns():
    groups = L.lattice_groups(8, 8)
    # X on (0,0) and (3,0), Y on (3,2), Z on (7,7)
    x = np.uint64((1 << 0) | (1 << 3) | (1 << (2 * 8 + 3)))
    z = np.uint64((1 << (2 * 8 + 3)) | (1 << 63))
    f = L.features(np.array([x]), np.array([z]), np.array([0.5]), groups, remaining=10, depth_scale=40)[0]
    assert f[0] * 8 == 2 and f[2] * 8 == 1 and f[16 + 2] * 8 == 1 and f[16 + 7] * 8 == 1  # rows
    assert f[8 + 0] * 8 == 1 and f[8 + 3] * 8 == 2 and f[16 + 8 + 3] * 8 == 1 and f[16 + 8 + 7] * 8 == 1  # cols
    assert f[32] * 16 == 4 and abs(f[33] * 8 - np.log(0.5)) < 1e-6 and f[34] == np.float32(0.25)
    assert (f[35:] == 0).all()


def test_remaining_feature_saturates_at_the_horizon():
    groups = L.lattice_groups(4, 4)
    rem = np.array([0.0, 20.0, 40.0, 41.0, 1000.0], np.float32)
    f = L.features(np.zeros(5, np.uint64), np.zeros(5, np.uint64), np.ones(5), groups, rem, 40.0)
    assert list(f[:, 34]) == [0.0, 0.5, 1.0, 1.0, 1.0]


@pytest.mark.parametrize("seed", [0, 1])
def test_backward_matches_finite_differences(seed):
    rng = np.random.default_rng(seed)
    p = L.Params(seed)
    p.b1 = rng.standard_normal(L.HID).astype(np.float32) * 0.1
    p.b2 = rng.standard_normal(L.HID).astype(np.float32) * 0.1
    f = rng.standard_normal((40, L.IN_PAD)).astype(np.float32)
    f[:, L.IN:] = 0
    t = rng.standard_normal(40)
    g = L.backward(p, f, t, exact=True)
    flat = p.flat()
    gflat = np.concatenate([g.W1.ravel(), g.b1, g.W2.ravel(), g.b2, g.w3, [g.b3]])
    idx = np.concatenate([rng.choice(L.IN * L.HID, 12, replace=False),            # W1 (live rows)
                          L.IN_PAD * L.HID + rng.choice(L.HID, 4, replace=False),  # b1
                          L.IN_PAD * L.HID + L.HID + rng.choice(L.HID * L.HID, 12, replace=False),  # W2
                          [len(flat) - 1 - L.HID - 3, len(flat) - 5, len(flat) - 1]])  # b2, w3, b3
    for i in idx:
        h = 1e-3
        up, dn = flat.copy(), flat.copy()
        up[i] += h
        dn[i] -= h
        num = (L.loss(L.Params.from_flat(up), f, t, exact=True) - L.loss(L.Params.from_flat(dn), f, t, exact=True)) / (2 * h)
        assert abs(num - gflat[i]) <= 2e-3 * max(1.0, abs(num)), (i, num, gflat[i])
