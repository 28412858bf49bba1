"""The learned truncation policy: a network that predicts each term's contribution to the final expectation value,
and a layer pass that keeps the terms it ranks highest (pp_model_*, pp_truncate_learned, pp_propagate_learned).

No counterpart in the reference.  It plugs into the TruncationPolicy seam (crates/paulistrings/src/truncation/
mod.rs:30-113); plain |c| TopN (truncation/builtin.rs:61-150) is both its fallback and its judge.

The network predicts a term's read-out factor: for a term P with k layers still to apply, ln|<psi|P_k|psi>|, P_k
being P propagated through those k layers.  The pass ranks terms by ln|c_P| + that prediction, the predicted ln of
the term's share of the final read-out; a network that predicts a constant ranks exactly as |c| TopN does.

Labels come from the plain engine: each sampled P is propagated on its own (coefficient 1) through its k layers and
read out, floored at ``floor``.  Many such single-term propagations run as one batch (pp_batch_propagate).
"""

from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _native, truncation
from .circuits import Circuit

LABEL_FLOOR = -40.0  # ln of a contribution that is zero, or below e^-40


def lattice_groups(lx: int, ly: int) -> np.ndarray:
    """Row masks then column masks of an lx-by-ly lattice (qubit y * lx + x), padded to 16 groups."""
    if lx > 8 or ly > 8 or lx * ly > 64:
        raise ValueError("the learned features cover lattices of at most 8 x 8 qubits")
    g = np.zeros(16, np.uint64)
    for y in range(ly):
        for x in range(lx):
            q = y * lx + x
            g[y] |= np.uint64(1) << np.uint64(q)
            g[8 + x] |= np.uint64(1) << np.uint64(q)
    return g


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _bind():
    L = _native.lib()
    if getattr(L, "_learned_bound", False):
        return L
    vp, u64, u32, f32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_float
    L.pp_model_param_count.restype = u32
    L.pp_model_param_count.argtypes = []
    L.pp_model_create.argtypes = [vp, vp, f32, u64, ctypes.POINTER(vp)]
    L.pp_model_free.argtypes = [vp]
    L.pp_model_set_params.argtypes = [vp, vp]
    L.pp_model_get_params.argtypes = [vp, vp]
    L.pp_model_predict.argtypes = [vp, vp, vp, vp, vp, u64, vp]
    L.pp_model_gradient.argtypes = [vp, vp, vp, vp, vp, vp, u64, vp, ctypes.POINTER(ctypes.c_double)]
    L.pp_model_train.argtypes = [vp, vp, vp, vp, vp, vp, u64, u32, u32, f32, u64, ctypes.POINTER(ctypes.c_double)]
    L.pp_truncate_learned.argtypes = [vp, vp, u64, f32]
    L.pp_propagate_learned.argtypes = [vp, ctypes.POINTER(_native.Channel), u64, ctypes.c_int,
                                       ctypes.POINTER(_native.Policy), vp, u64, ctypes.POINTER(_native.Options), vp]
    for name in ("pp_model_create", "pp_model_free", "pp_model_set_params", "pp_model_get_params", "pp_model_predict",
                 "pp_model_gradient", "pp_model_train", "pp_truncate_learned", "pp_propagate_learned"):
        getattr(L, name).restype = ctypes.c_int
    L._learned_bound = True
    return L


class _Samples:
    def __init__(self, x, z, c, remaining, labels=None):
        self.x = np.ascontiguousarray(np.asarray(x, np.uint64).reshape(-1))
        self.z = np.ascontiguousarray(np.asarray(z, np.uint64).reshape(-1))
        self.c = np.ascontiguousarray(np.asarray(c, np.float64).reshape(-1))
        n = len(self.x)
        r = np.asarray(remaining, np.float32)
        self.r = np.ascontiguousarray(np.broadcast_to(r, (n,)).astype(np.float32))
        self.l = None if labels is None else np.ascontiguousarray(np.asarray(labels, np.float32).reshape(-1))
        if len(self.z) != n or len(self.c) != n or (self.l is not None and len(self.l) != n):
            raise ValueError("sample arrays differ in length")


class Model:
    """The predictor on one device (pp_model)."""

    def __init__(self, groups, depth_scale: float, seed: int = 0, device: int = 0):
        L = _bind()
        self.ctx = _native.Context.get(device)
        self.groups = np.ascontiguousarray(np.asarray(groups, np.uint64))
        if self.groups.shape != (16,):
            raise ValueError("16 qubit groups")
        self.depth_scale = float(depth_scale)
        self._h = ctypes.c_void_p()
        _native.check(L.pp_model_create(self.ctx.handle, _p(self.groups), self.depth_scale, int(seed),
                                        ctypes.byref(self._h)), self.ctx.handle)

    @staticmethod
    def param_count() -> int:
        return int(_bind().pp_model_param_count())

    def free(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            _bind().pp_model_free(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def get_params(self) -> np.ndarray:
        out = np.zeros(self.param_count(), np.float32)
        _native.check(_bind().pp_model_get_params(self._h, _p(out)), self.ctx.handle)
        return out

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(np.asarray(p, np.float32))
        if p.shape != (self.param_count(),):
            raise ValueError("parameter vector length")
        _native.check(_bind().pp_model_set_params(self._h, _p(p)), self.ctx.handle)

    def predict(self, x, z, c, remaining) -> np.ndarray:
        s = _Samples(x, z, c, remaining)
        out = np.zeros(len(s.x), np.float32)
        _native.check(_bind().pp_model_predict(self._h, _p(s.x), _p(s.z), _p(s.c), _p(s.r), len(s.x), _p(out)),
                      self.ctx.handle)
        return out

    def gradient(self, x, z, c, remaining, labels):
        s = _Samples(x, z, c, remaining, labels)
        g = np.zeros(self.param_count(), np.float32)
        loss = ctypes.c_double()
        _native.check(_bind().pp_model_gradient(self._h, _p(s.x), _p(s.z), _p(s.c), _p(s.r), _p(s.l), len(s.x),
                                                _p(g), ctypes.byref(loss)), self.ctx.handle)
        return g, loss.value

    def train(self, x, z, c, remaining, labels, epochs: int = 20, batch: int = 1024, lr: float = 1e-3,
              seed: int = 1) -> float:
        s = _Samples(x, z, c, remaining, labels)
        loss = ctypes.c_double()
        _native.check(_bind().pp_model_train(self._h, _p(s.x), _p(s.z), _p(s.c), _p(s.r), _p(s.l), len(s.x), epochs,
                                             batch, lr, seed, ctypes.byref(loss)), self.ctx.handle)
        return loss.value


def truncate_learned(sum_, model: Model, n: int, remaining: float) -> None:
    """The learned layer pass on a resident one-word sum: keep the n terms of largest predicted contribution."""
    _native.check(_bind().pp_truncate_learned(sum_._h, model._h, int(n), float(remaining)), sum_.ctx.handle)


def propagate_learned(sum_, circuit: Circuit, model: Model, keep: int, filter_policy=None, direction=None,
                      target_bucket_len: int = 1024, min_buckets: int = 128) -> None:
    """``propagate`` with the learned pass keeping ``keep`` terms after every run of layers."""
    from . import GpuPauliSum, _options, _parse_direction

    chans = circuit.channels()
    arr = (_native.Channel * max(1, len(chans)))()
    for i, row in enumerate(chans):
        arr[i] = GpuPauliSum._channel(row)
    pol = truncation.lower(filter_policy)
    opts = _options(target_bucket_len, min_buckets)
    _native.check(_bind().pp_propagate_learned(sum_._h, arr, len(chans), _parse_direction(direction), ctypes.byref(pol),
                                               model._h, int(keep), ctypes.byref(opts), None), sum_.ctx.handle)


def readout_labels(terms_x, terms_z, circuit_rest: Circuit, state: str, policy=None, direction: str = "heisenberg",
                   floor: float = LABEL_FLOOR, batch: int = 256) -> np.ndarray:
    """ln|<psi| P propagated through circuit_rest |psi>| for each one-word term P (floored at ``floor``), each term
    propagated on its own with coefficient 1 by the plain engine, ``batch`` terms per pp_batch_propagate call."""
    from . import PauliSum, propagate_batch

    x = np.asarray(terms_x, np.uint64).reshape(-1)
    z = np.asarray(terms_z, np.uint64).reshape(-1)
    nq = circuit_rest.num_qubits
    out = np.full(len(x), floor, np.float32)
    for lo in range(0, len(x), batch):
        hi = min(len(x), lo + batch)
        sums = [PauliSum(nq, x[i:i + 1].reshape(1, 1), z[i:i + 1].reshape(1, 1), np.array([1.0])).to_device(0)
                for i in range(lo, hi)]
        propagate_batch(sums, [circuit_rest] * len(sums), [policy] * len(sums), [direction] * len(sums))
        for k, g in enumerate(sums):
            v = abs(g.expectation(state))
            out[lo + k] = max(floor, math.log(v)) if v > 0 else floor
            g.free()
    return out
