"""ctypes binding of the CPU oracle (oracle/libpp_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()`` and by
the ``cpu_baseline`` leg of ``bench.py``, always as the checker or the CPU baseline,
never by the product path.  See pp_oracle.h for what each entry point restates.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libpp_oracle.so"

F64, F32 = 0, 1
NODE_TRUE, NODE_COEFF, NODE_WEIGHT, NODE_AND, NODE_OR = 0, 1, 2, 3, 4
MAX_NODES = 15


class _Node(ctypes.Structure):
    _fields_ = [("op", ctypes.c_int32), ("k", ctypes.c_uint32), ("eps", ctypes.c_double)]


class _Policy(ctypes.Structure):
    _fields_ = [
        ("n_nodes", ctypes.c_int32),
        ("nodes", _Node * MAX_NODES),
        ("has_topn", ctypes.c_int32),
        ("topn", ctypes.c_uint64),
    ]


def build(force: bool = False) -> Path:
    """Compile the oracle with its Makefile (gcc); returns the library path."""
    if force or not LIB_PATH.exists():
        subprocess.run(["make", "-C", str(HERE), "-B" if force else "all"], check=True,
                       stdout=subprocess.DEVNULL)
    return LIB_PATH


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = ctypes.CDLL(str(LIB_PATH))
        vp, u64p, u32p = ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint32)
        L.ppo_sum_new.restype = vp
        L.ppo_sum_new.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_size_t, vp, vp, vp,
                                  ctypes.c_uint64]
        L.ppo_sum_free.argtypes = [vp]
        L.ppo_sum_clone.restype = vp
        L.ppo_sum_clone.argtypes = [vp]
        L.ppo_sum_len.restype = ctypes.c_size_t
        L.ppo_sum_len.argtypes = [vp]
        L.ppo_sum_bits.restype = ctypes.c_int
        L.ppo_sum_bits.argtypes = [vp]
        L.ppo_sum_export.argtypes = [vp, vp, vp, vp, vp]
        L.ppo_propagate.restype = ctypes.c_int
        L.ppo_propagate.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int, ctypes.POINTER(_Policy), ctypes.c_int,
                                    ctypes.c_int, ctypes.c_int, vp]
        L.ppo_topn.argtypes = [vp, ctypes.c_uint64]
        L.ppo_approx_topn.argtypes = [vp, ctypes.c_uint64]
        L.ppo_expectation_product.restype = ctypes.c_double
        L.ppo_expectation_product.argtypes = [vp, vp, vp, vp]
        L.ppo_naive_layer.restype = ctypes.c_size_t
        L.ppo_naive_layer.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_size_t, vp, vp, vp, vp, ctypes.c_int,
                                      ctypes.POINTER(_Policy), vp, vp, vp]
        L.ppo_mix64.restype = ctypes.c_uint64
        L.ppo_mix64.argtypes = [ctypes.c_uint64]
        L.ppo_splitmix64.restype = ctypes.c_uint64
        L.ppo_splitmix64.argtypes = [u64p]
        L.ppo_hash_rows.restype = ctypes.c_int
        L.ppo_hash_rows.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_int, vp, vp]
        L.ppo_desired_bits.restype = ctypes.c_int
        L.ppo_desired_bits.argtypes = [ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t]
        L.ppo_rebucketed_bits.restype = ctypes.c_int
        L.ppo_rebucketed_bits.argtypes = [ctypes.c_size_t, ctypes.c_int, ctypes.c_size_t, ctypes.c_size_t]
        L.ppo_sin_cos.argtypes = [ctypes.c_double, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
        _ = u32p
        _lib = L
    return _lib


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def make_policy(nodes=(), topn=None, approx_topn=None) -> _Policy:
    """``nodes``: postfix list of (op, k, eps); at most one of ``topn`` (TopN) and ``approx_topn`` (ApproxTopN)."""
    p = _Policy()
    if len(nodes) > MAX_NODES:
        raise ValueError("at most 15 keep nodes")
    if topn is not None and approx_topn is not None:
        raise ValueError("one layer pass at most")
    p.n_nodes = len(nodes)
    for i, (op, k, eps) in enumerate(nodes):
        p.nodes[i].op = op
        p.nodes[i].k = k
        p.nodes[i].eps = eps
    p.has_topn = 1 if topn is not None else (2 if approx_topn is not None else 0)
    p.topn = int(topn if topn is not None else (approx_topn if approx_topn is not None else 0))
    return p


def coeff_policy(eps=None, max_weight=None, topn=None, approx_topn=None) -> _Policy:
    """And(CoefficientThreshold(eps), WeightCutoff(max_weight)) [+ TopN or ApproxTopN]."""
    nodes = []
    if eps is not None:
        nodes.append((NODE_COEFF, 0, float(eps)))
    if max_weight is not None:
        nodes.append((NODE_WEIGHT, int(max_weight), 0.0))
        if eps is not None:
            nodes.append((NODE_AND, 0, 0.0))
    return make_policy(nodes, topn, approx_topn)


class OracleSum:
    """A bucketed Pauli sum held by the oracle."""

    def __init__(self, num_qubits: int, x: np.ndarray, z: np.ndarray, coeff: np.ndarray,
                 seed: int = 0x9E3779B97F4A7C15, width: int | None = None):
        x = np.ascontiguousarray(x, dtype=np.uint64)
        z = np.ascontiguousarray(z, dtype=np.uint64)
        if x.ndim == 1:
            x = x[:, None]
            z = z[:, None]
        self.width = width if width is not None else x.shape[1]
        if x.shape[1] != self.width:
            raise ValueError("key width mismatch")
        coeff = np.ascontiguousarray(coeff)
        self.kind = F64 if coeff.dtype == np.float64 else F32
        if coeff.dtype not in (np.float64, np.float32):
            raise TypeError("coefficients must be float64 or float32")
        self.num_qubits = num_qubits
        self._h = lib().ppo_sum_new(self.width, self.kind, num_qubits, len(coeff), _ptr(x), _ptr(z),
                                    _ptr(coeff), seed)
        if not self._h:
            raise ValueError("oracle rejected the sum")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().ppo_sum_free(h)
            self._h = None

    def __len__(self):
        return lib().ppo_sum_len(self._h)

    def clone(self) -> "OracleSum":
        out = OracleSum.__new__(OracleSum)
        out.width, out.kind, out.num_qubits = self.width, self.kind, self.num_qubits
        out._h = lib().ppo_sum_clone(self._h)
        return out

    @property
    def bits(self) -> int:
        return lib().ppo_sum_bits(self._h)

    def export(self):
        n = len(self)
        x = np.zeros((n, self.width), np.uint64)
        z = np.zeros((n, self.width), np.uint64)
        c = np.zeros(n, np.float64 if self.kind == F64 else np.float32)
        b = np.zeros(n, np.uint32)
        if n:
            lib().ppo_sum_export(self._h, _ptr(x), _ptr(z), _ptr(c), _ptr(b))
        return x, z, c, b

    def propagate(self, channels: np.ndarray, direction: str = "forward", policy: _Policy | None = None,
                  target_bucket_len: int = 1024, min_buckets: int = 128, threads: int = 0) -> np.ndarray:
        channels = np.ascontiguousarray(channels)
        lens = np.zeros(len(channels), np.uint64)
        pol = policy if policy is not None else make_policy()
        rc = lib().ppo_propagate(self._h, _ptr(channels), len(channels), 1 if direction == "heisenberg" else 0,
                                 ctypes.byref(pol), target_bucket_len, min_buckets, threads, _ptr(lens))
        if rc != 0:
            raise RuntimeError(f"oracle propagate failed ({rc})")
        return lens

    def topn(self, n: int):
        lib().ppo_topn(self._h, n)

    def approx_topn(self, n: int):
        lib().ppo_approx_topn(self._h, n)

    def expectation(self, axis: str = "z", negative=None) -> float:
        ax = np.zeros(4, np.uint64)
        az = np.zeros(4, np.uint64)
        neg = np.zeros(4, np.uint64)
        allw = np.uint64(0xFFFFFFFFFFFFFFFF)
        if axis in ("x", "y"):
            ax[:] = allw
        if axis in ("z", "y"):
            az[:] = allw
        if negative is not None:
            neg[: len(negative)] = negative
        return lib().ppo_expectation_product(self._h, _ptr(ax), _ptr(az), _ptr(neg))


def naive_layer(width: int, x, z, coeff, channel: np.ndarray, adjoint: bool = False, policy=None):
    x = np.ascontiguousarray(x, dtype=np.uint64).reshape(-1, width)
    z = np.ascontiguousarray(z, dtype=np.uint64).reshape(-1, width)
    coeff = np.ascontiguousarray(coeff)
    kind = F64 if coeff.dtype == np.float64 else F32
    n = len(coeff)
    ox = np.zeros((2 * n + 1, width), np.uint64)
    oz = np.zeros((2 * n + 1, width), np.uint64)
    oc = np.zeros(2 * n + 1, coeff.dtype)
    ch = np.ascontiguousarray(channel).reshape(-1)[:1]
    pol = policy if policy is not None else make_policy()
    m = lib().ppo_naive_layer(width, kind, n, _ptr(x), _ptr(z), _ptr(coeff), _ptr(ch), 1 if adjoint else 0,
                              ctypes.byref(pol), _ptr(ox), _ptr(oz), _ptr(oc))
    return ox[:m], oz[:m], oc[:m]


def hash_rows(num_qubits: int, seed: int, width: int):
    rx = np.zeros(31 * width, np.uint64)
    rz = np.zeros(31 * width, np.uint64)
    n = lib().ppo_hash_rows(num_qubits, seed, width, _ptr(rx), _ptr(rz))
    return rx[: n * width].reshape(n, width), rz[: n * width].reshape(n, width)


def sin_cos(theta: float):
    s = ctypes.c_double()
    c = ctypes.c_double()
    lib().ppo_sin_cos(theta, ctypes.byref(s), ctypes.byref(c))
    return s.value, c.value


def set_threads(n: int) -> None:
    os.environ["OMP_NUM_THREADS"] = str(n)
