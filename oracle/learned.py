"""CPU restatement of the learned truncation policy's network (numpy): features, forward, loss and backward.

TEST INFRASTRUCTURE ONLY: tests/ import it as the checker of the device kernels (paulistrings-rs_amd/csrc/
pp_learned.hip); the product never does.  The policy has no counterpart in the reference: it plugs into the
TruncationPolicy seam (crates/paulistrings/src/truncation/mod.rs:30-113), and plain |c| TopN
(truncation/builtin.rs:61-150) is both its fallback and its judge.

Network: 35 inputs -> 128 -> 128 -> 1, ReLU between layers.  The two hidden GEMMs run on the device in bf16 with
fp32 accumulation: inputs and hidden activations are rounded to bf16 (round to nearest even) before each GEMM, and
so are the weight copies they multiply; biases, the last layer, the loss and every gradient are fp32.  This module
rounds at the same places, so the device's results differ from it only by fp32 summation order.

Features of a term (x, z words, coefficient c) on a lattice of up to 16 row/column groups of qubits:
  [0, 16)   popcount(x & group_g) / 8      (X or Y on the group's qubits)
  [16, 32)  popcount(z & group_g) / 8      (Z or Y)
  32        Pauli weight / 16
  33        ln|c| / 8
  34        min(remaining layers / depth scale, 1): the depth scale is the horizon past which depths read alike
  [35, 64)  zero padding (the first GEMM's K is 64)
"""

from __future__ import annotations

import numpy as np

IN, IN_PAD, HID = 35, 64, 128


def bf16_round(a: np.ndarray) -> np.ndarray:
    """float32 -> nearest bf16 (ties to even), returned as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    bias = ((u >> 16) & 1) + 0x7FFF
    r = ((u + bias) >> 16) << 16
    nan = np.isnan(a)
    out = r.astype(np.uint32).view(np.float32).copy()
    out[nan] = np.nan
    return out


def lattice_groups(lx: int, ly: int) -> np.ndarray:
    """Row masks then column masks of an lx-by-ly lattice, qubit index y * lx + x, padded to 16 groups."""
    g = np.zeros(16, np.uint64)
    for y in range(ly):
        for x in range(lx):
            q = y * lx + x
            g[y] |= np.uint64(1) << np.uint64(q)
            g[8 + x] |= np.uint64(1) << np.uint64(q)
    return g


def popcount64(a: np.ndarray) -> np.ndarray:
    a = np.asarray(a, dtype=np.uint64)
    c = np.zeros(a.shape, np.int64)
    for sh in range(0, 64, 8):
        c += np.array([bin(v).count("1") for v in range(256)], np.int64)[((a >> np.uint64(sh)) & np.uint64(255)).astype(np.int64)]
    return c


def features(x: np.ndarray, z: np.ndarray, c: np.ndarray, groups: np.ndarray, remaining: float,
             depth_scale: float) -> np.ndarray:
    """The (n, 64) float32 feature rows of n one-word terms (x, z of shape (n,) or (n, 1))."""
    x = np.asarray(x, np.uint64).reshape(-1)
    z = np.asarray(z, np.uint64).reshape(-1)
    n = len(x)
    f = np.zeros((n, IN_PAD), np.float32)
    for g in range(16):
        f[:, g] = popcount64(x & groups[g]) / 8.0
        f[:, 16 + g] = popcount64(z & groups[g]) / 8.0
    f[:, 32] = popcount64(x | z) / 16.0
    f[:, 33] = (np.log(np.abs(np.asarray(c, np.float64))) / 8.0).astype(np.float32)
    f[:, 34] = np.minimum(np.float32(remaining / depth_scale), np.float32(1.0))
    return f


class Params:
    """fp32 master weights: W1 (64, 128) with rows >= 35 zero, b1 (128,), W2 (128, 128), b2 (128,), w3 (128,), b3."""

    def __init__(self, seed: int = 0):
        rng = np.random.default_rng(seed)
        self.W1 = np.zeros((IN_PAD, HID), np.float32)
        self.W1[:IN] = (rng.standard_normal((IN, HID)) * np.sqrt(2.0 / IN)).astype(np.float32)
        self.b1 = np.zeros(HID, np.float32)
        self.W2 = (rng.standard_normal((HID, HID)) * np.sqrt(2.0 / HID)).astype(np.float32)
        self.b2 = np.zeros(HID, np.float32)
        self.w3 = (rng.standard_normal(HID) * np.sqrt(1.0 / HID)).astype(np.float32)
        self.b3 = np.float32(0.0)

    def flat(self) -> np.ndarray:
        return np.concatenate([self.W1.ravel(), self.b1, self.W2.ravel(), self.b2, self.w3, [self.b3]]).astype(np.float32)

    @staticmethod
    def from_flat(v: np.ndarray) -> "Params":
        p = Params.__new__(Params)
        v = np.asarray(v, np.float32)
        o = 0
        p.W1 = v[o:o + IN_PAD * HID].reshape(IN_PAD, HID).copy(); o += IN_PAD * HID
        p.b1 = v[o:o + HID].copy(); o += HID
        p.W2 = v[o:o + HID * HID].reshape(HID, HID).copy(); o += HID * HID
        p.b2 = v[o:o + HID].copy(); o += HID
        p.w3 = v[o:o + HID].copy(); o += HID
        p.b3 = np.float32(v[o])
        return p


def forward(p: Params, f: np.ndarray, exact: bool = False):
    """Returns (y, cache).  GEMM operands rounded to bf16, products accumulated in float64 here (the device
    accumulates in fp32: the tests' tolerance covers the difference).  exact=True skips the rounding (the smooth
    function a finite-difference check of backward needs)."""
    rnd = (lambda a: np.asarray(a, np.float32)) if exact else bf16_round
    a0 = rnd(f).astype(np.float64)
    z1 = a0 @ rnd(p.W1).astype(np.float64) + p.b1
    h1 = np.maximum(z1, 0.0).astype(np.float32)
    a1 = rnd(h1).astype(np.float64)
    z2 = a1 @ rnd(p.W2).astype(np.float64) + p.b2
    h2 = np.maximum(z2, 0.0).astype(np.float32)
    y = (h2.astype(np.float64) @ p.w3.astype(np.float64) + p.b3).astype(np.float32)
    return y, (a0, z1, a1, z2, h2)


def loss(p: Params, f: np.ndarray, t: np.ndarray, exact: bool = False) -> float:
    y, _ = forward(p, f, exact)
    return float(np.mean((y.astype(np.float64) - t) ** 2))


def backward(p: Params, f: np.ndarray, t: np.ndarray, exact: bool = False) -> Params:
    """Gradients of the mean squared error, the rounding treated as the identity (straight-through), with the
    device's GEMM operands: dW2 = a1^T bf16(dz2), dh1 = bf16(dz2) bf16(W2)^T, dW1 = a0^T bf16(dz1) (a0, a1 the
    bf16 inputs and hidden activations); bias and last-layer gradients from the fp32 values."""
    y, (a0, z1, a1, z2, h2) = forward(p, f, exact)
    rnd = (lambda a: np.asarray(a, np.float32)) if exact else bf16_round
    m = len(t)
    dy = 2.0 * (y.astype(np.float64) - t) / m
    g = Params.__new__(Params)
    g.w3 = (h2.astype(np.float64).T @ dy).astype(np.float32)
    g.b3 = np.float32(dy.sum())
    dh2 = np.outer(dy, p.w3.astype(np.float64))
    dz2 = (dh2 * (z2 > 0)).astype(np.float32)
    dz2r = rnd(dz2).astype(np.float64)  # the GEMMs' operand
    g.W2 = (a1.T @ dz2r).astype(np.float32)
    g.b2 = dz2.astype(np.float64).sum(axis=0).astype(np.float32)
    dh1 = dz2r @ rnd(p.W2).astype(np.float64).T
    dz1 = (dh1 * (z1 > 0)).astype(np.float32)
    g.W1 = (a0.T @ rnd(dz1).astype(np.float64)).astype(np.float32)
    g.W1[IN:] = 0.0
    g.b1 = dz1.astype(np.float64).sum(axis=0).astype(np.float32)
    return g
