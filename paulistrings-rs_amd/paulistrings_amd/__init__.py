"""paulistrings_amd -- Pauli propagation through rotation channels on one MI355X.

The user-facing surface mirrors the reference's Python API (crates/paulistrings-py/src/sum.rs:2132-2290,
python/paulistrings/__init__.py) for the rotation-layer path:

    from paulistrings_amd import PauliSum, Circuit, truncation, propagated
    obs = PauliSum.from_strings({"XIII": 0.25, "IXII": 0.25}, num_qubits=4)
    obs.propagate(circuit, truncation.coeff(1e-4) & truncation.weight(8), direction="heisenberg", device=0)
    obs.expectation("x+")

Every propagation runs on the GPU through libpp_hip.so (include/pp_hip.h); there is no host engine in this
package.  ``PauliSum.to_device`` returns a :class:`GpuPauliSum` that stays resident across calls.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _native, circuits, gates, truncation
from .circuits import Circuit

__all__ = ["PauliSum", "GpuPauliSum", "Circuit", "PropagationStats", "propagated", "propagated_with_stats",
           "propagate_batch",
           "truncation", "gates", "circuits", "DEFAULT_HASH_SEED", "device_available"]

#: pauli_sum/storage.rs:12
DEFAULT_HASH_SEED = 0x9E3779B97F4A7C15

_AXES = {"x": (1, 0), "y": (1, 1), "z": (0, 1)}


def device_available() -> bool:
    """Whether the HIP library loads and sees a device (the reference's ``cuda_available``)."""
    try:
        _native.Context.get(0)
        return True
    except Exception:
        return False


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _parse_direction(direction) -> int:
    if direction is None or direction == "forward":
        return _native.FORWARD
    if direction == "heisenberg":
        return _native.HEISENBERG
    raise ValueError(f"direction must be 'forward' or 'heisenberg', got {direction!r}")


def _options(target_bucket_len: int, min_buckets: int) -> _native.Options:
    # the reference's bounds (crates/paulistrings-py/src/sum.rs, propagate_options)
    if target_bucket_len < 1:
        raise ValueError(f"target_bucket_len must be at least 1, got {target_bucket_len}")
    if min_buckets < 16:
        raise ValueError(f"min_buckets must be at least 16, got {min_buckets}")
    o = _native.Options()
    o.target_bucket_len = int(target_bucket_len)
    o.min_buckets = int(min_buckets)
    return o


def _basis(state: str, num_qubits: int, width: int):
    """A uniform product state label: 'x+', 'z-', ... (ProductBasis::uniform, readout/product_state.rs:45-49)."""
    s = state.lower().strip()
    if len(s) not in (1, 2) or s[0] not in _AXES or (len(s) == 2 and s[1] not in "+-"):
        raise ValueError(f"state must be one of x+, x-, y+, y-, z+, z-; got {state!r}")
    bx, bz = _AXES[s[0]]
    live = [(1 << min(64, max(0, num_qubits - 64 * w))) - 1 for w in range(width)]
    ax = np.array([m if bx else 0 for m in live], np.uint64)
    az = np.array([m if bz else 0 for m in live], np.uint64)
    neg = np.array([m if s.endswith("-") else 0 for m in live], np.uint64)
    return ax, az, neg


def _canonical_terms(x: np.ndarray, z: np.ndarray, c: np.ndarray):
    """Sum repeated keys in input order and drop exact zeros (BuildAccumulator, pauli_sum/accumulator.rs:35-55)."""
    if len(c) == 0:
        return x, z, c
    cols = tuple(z[:, w] for w in reversed(range(z.shape[1]))) + tuple(x[:, w] for w in reversed(range(x.shape[1])))
    order = np.lexsort(cols)
    xs, zs, cs = x[order], z[order], c[order]
    new = np.ones(len(cs), bool)
    new[1:] = np.any(xs[1:] != xs[:-1], axis=1) | np.any(zs[1:] != zs[:-1], axis=1)
    if new.all():
        keep = cs != 0
        return xs[keep], zs[keep], cs[keep]
    starts = np.flatnonzero(new)
    # one addition at a time in input order (the sort is stable), as the reference's hashmap accumulates:
    # np.add.reduceat would sum long runs pairwise
    group = np.cumsum(new) - 1
    sums = np.zeros(len(starts), dtype=c.dtype)
    np.add.at(sums, group, cs)
    keep = sums != 0
    return xs[starts][keep], zs[starts][keep], sums[keep]


@dataclass
class PropagationStats:
    """Per-layer term counts of a propagation (the reference's PropagationStats, sum.rs:900-1031)."""

    terms_in: list[int] = field(default_factory=list)
    terms_out: list[int] = field(default_factory=list)

    @property
    def layers(self) -> int:
        return len(self.terms_out)

    @property
    def peak_terms(self) -> int:
        """``max(terms_in[0], terms_out...)``: the peak resident count between layers."""
        if not self.terms_out:
            return self.terms_in[0] if self.terms_in else 0
        return max([self.terms_in[0]] + self.terms_out)

    @property
    def final_terms(self) -> int:
        return self.terms_out[-1] if self.terms_out else (self.terms_in[0] if self.terms_in else 0)


class PauliSum:
    """A real-coefficient sum of Pauli strings on ``num_qubits`` qubits, held on the host."""

    def __init__(self, num_qubits: int, x=None, z=None, coefficients=None, dtype=np.float64,
                 hash_seed: int = DEFAULT_HASH_SEED):
        self.num_qubits = int(num_qubits)
        self.width = circuits.width_for(self.num_qubits)
        dtype = np.dtype(dtype)
        if dtype not in (np.float64, np.float32):
            raise ValueError(f"dtype must be float64 or float32, got {dtype}")
        if x is None:
            x = np.zeros((0, self.width), np.uint64)
            z = np.zeros((0, self.width), np.uint64)
            coefficients = np.zeros(0, dtype)
        x = np.ascontiguousarray(x, dtype=np.uint64).reshape(-1, self.width)
        z = np.ascontiguousarray(z, dtype=np.uint64).reshape(-1, self.width)
        c = np.ascontiguousarray(coefficients, dtype=dtype).reshape(-1)
        if not (len(x) == len(z) == len(c)):
            raise ValueError("x, z and coefficients must have one row per term")
        live = np.array([(1 << min(64, max(0, self.num_qubits - 64 * w))) - 1 for w in range(self.width)],
                        dtype=np.uint64)
        if len(x) and (np.any(x & ~live) or np.any(z & ~live)):
            raise ValueError(f"a term acts outside the {self.num_qubits}-qubit register")
        self._x, self._z, self._c = _canonical_terms(x, z, c)
        self.hash_seed = int(hash_seed)

    # ---- construction -------------------------------------------------------------------------------------------
    @classmethod
    def from_strings(cls, terms: dict, num_qubits: int | None = None, dtype=np.float64) -> "PauliSum":
        """``{"XIZ": 0.5, ...}``: character i is qubit i (positional labels)."""
        labels = list(terms)
        n = num_qubits if num_qubits is not None else (len(labels[0]) if labels else 0)
        W = circuits.width_for(n)
        x = np.zeros((len(labels), W), np.uint64)
        z = np.zeros((len(labels), W), np.uint64)
        for i, label in enumerate(labels):
            if len(label) != n:
                raise ValueError(f"label {label!r} has {len(label)} characters for {n} qubits")
            for q, ch in enumerate(label.upper()):
                if ch not in "IXYZ":
                    raise ValueError(f"unknown Pauli {ch!r} in {label!r}")
                bit = np.uint64(1) << np.uint64(q % 64)
                if ch in "XY":
                    x[i, q // 64] |= bit
                if ch in "ZY":
                    z[i, q // 64] |= bit
        return cls(n, x, z, np.array([terms[k] for k in labels]), dtype=dtype)

    @classmethod
    def from_arrays(cls, num_qubits: int, x, z, coefficients, dtype=None) -> "PauliSum":
        c = np.asarray(coefficients)
        return cls(num_qubits, x, z, c, dtype=dtype or (c.dtype if c.dtype in (np.float32, np.float64) else np.float64))

    def copy(self) -> "PauliSum":
        out = PauliSum.__new__(PauliSum)
        out.num_qubits, out.width, out.hash_seed = self.num_qubits, self.width, self.hash_seed
        out._x, out._z, out._c = self._x.copy(), self._z.copy(), self._c.copy()
        return out

    # ---- access ---------------------------------------------------------------------------------------------------
    def __len__(self) -> int:
        return len(self._c)

    @property
    def dtype(self):
        return self._c.dtype

    def x_array(self) -> np.ndarray:
        return self._x.copy()

    def z_array(self) -> np.ndarray:
        return self._z.copy()

    def coefficients_array(self) -> np.ndarray:
        return self._c.copy()

    def to_dict(self) -> dict:
        return {(tuple(int(v) for v in self._x[i]), tuple(int(v) for v in self._z[i])): float(self._c[i])
                for i in range(len(self._c))}

    # ---- device ---------------------------------------------------------------------------------------------------
    def to_device(self, device: int = 0) -> "GpuPauliSum":
        """Upload to GPU ``device``, returning a resident :class:`GpuPauliSum`; ``self`` is untouched."""
        return GpuPauliSum._upload(self, device)

    def _take(self, g: "GpuPauliSum") -> None:
        x, z, c = g._download()
        self._x, self._z, self._c = x, z, c

    def propagate(self, circuit: Circuit, policy=None, direction=None, target_bucket_len: int = 1024,
                  min_buckets: int = 128, device=None) -> None:
        """Propagate in place through ``circuit`` on GPU ``device`` (0 when None); the result is in the
        reference's canonical order."""
        g = self.to_device(0 if device is None else int(device))
        g.propagate(circuit, policy, direction, target_bucket_len, min_buckets)
        self._take(g)
        g.free()

    def propagate_with_stats(self, circuit: Circuit, policy=None, direction=None, target_bucket_len: int = 1024,
                             min_buckets: int = 128, device=None) -> PropagationStats:
        g = self.to_device(0 if device is None else int(device))
        stats = g.propagate_with_stats(circuit, policy, direction, target_bucket_len, min_buckets)
        self._take(g)
        g.free()
        return stats

    def expectation(self, state: str = "z+") -> float:
        """``<psi|O|psi>`` in a uniform product state ``'x+'``, ``'z-'``, ... (readout/product_state.rs:84-101)."""
        g = self.to_device(0)
        try:
            return g.expectation(state)
        finally:
            g.free()

    def __repr__(self) -> str:
        return f"PauliSum(num_qubits={self.num_qubits}, terms={len(self)}, dtype={self.dtype})"


class GpuPauliSum:
    """A Pauli sum resident on one GPU, stepped in place (the reference's GpuPauliSum)."""

    def __init__(self):
        raise TypeError("use PauliSum.to_device()")

    @classmethod
    def _upload(cls, host: PauliSum, device: int) -> "GpuPauliSum":
        self = cls.__new__(cls)
        self.ctx = _native.Context.get(device)
        self.num_qubits = host.num_qubits
        self.width = host.width
        self.dtype = host.dtype
        self._h = ctypes.c_void_p()
        kind = _native.COEFF_F64 if host.dtype == np.float64 else _native.COEFF_F32
        x = np.ascontiguousarray(host._x)
        z = np.ascontiguousarray(host._z)
        c = np.ascontiguousarray(host._c)
        _native.check(_native.lib().pp_sum_upload(self.ctx.handle, self.num_qubits, self.width, kind, len(c),
                                                  _ptr(x), _ptr(z), _ptr(c), host.hash_seed,
                                                  ctypes.byref(self._h)), self.ctx.handle)
        return self

    def free(self) -> None:
        if getattr(self, "_h", None) and self._h.value:
            _native.lib().pp_sum_free(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def __len__(self) -> int:
        n = ctypes.c_uint64()
        _native.check(_native.lib().pp_sum_len(self._h, ctypes.byref(n)), self.ctx.handle)
        return n.value

    def clone(self) -> "GpuPauliSum":
        out = GpuPauliSum.__new__(GpuPauliSum)
        out.ctx, out.num_qubits, out.width, out.dtype = self.ctx, self.num_qubits, self.width, self.dtype
        out._h = ctypes.c_void_p()
        _native.check(_native.lib().pp_sum_clone(self._h, ctypes.byref(out._h)), self.ctx.handle)
        return out

    def copy_from(self, src: "GpuPauliSum") -> None:
        """Take ``src``'s terms and layout into this sum's own device memory (pp_sum_copy; asynchronous)."""
        _native.check(_native.lib().pp_sum_copy(self._h, src._h), self.ctx.handle)

    CHECK_NAMES = ("misplaced", "zero", "segments", "duplicates", "unscanned", "overlaps", "length")

    def check_layout(self) -> dict:
        """The resident layout's invariants (pp_sum_check): violations by kind, all zero for a sound layout
        ("unscanned" counts buckets too long for the duplicate scan, not a violation)."""
        out = np.zeros(7, np.uint64)
        _native.check(_native.lib().pp_sum_check(self._h, out.ctypes.data), self.ctx.handle)
        return {k: int(v) for k, v in zip(self.CHECK_NAMES, out)}

    def bucket_bits(self) -> tuple[int, int]:
        r = ctypes.c_int32()
        d = ctypes.c_int32()
        _native.check(_native.lib().pp_sum_bucket_bits(self._h, ctypes.byref(r), ctypes.byref(d)), self.ctx.handle)
        return r.value, d.value

    def _download(self):
        n = len(self)
        x = np.zeros((n, self.width), np.uint64)
        z = np.zeros((n, self.width), np.uint64)
        c = np.zeros(n, self.dtype)
        if n:
            _native.check(_native.lib().pp_sum_download(self._h, _ptr(x), _ptr(z), _ptr(c)), self.ctx.handle)
        return x, z, c

    def to_host(self) -> PauliSum:
        x, z, c = self._download()
        out = PauliSum.__new__(PauliSum)
        out.num_qubits, out.width, out.hash_seed = self.num_qubits, self.width, DEFAULT_HASH_SEED
        out._x, out._z, out._c = x, z, c
        return out

    def _run(self, circuit: Circuit, policy, direction, target_bucket_len, min_buckets) -> np.ndarray:
        if circuit.num_qubits != self.num_qubits:
            raise ValueError(f"the circuit acts on {circuit.num_qubits} qubits, the sum on {self.num_qubits}")
        chans = np.ascontiguousarray(circuit.channels())
        lens = np.zeros(max(1, len(chans)), np.uint64)
        pol = truncation.lower(policy)
        opts = _options(target_bucket_len, min_buckets)
        _native.check(_native.lib().pp_propagate(self._h, _ptr(chans), len(chans), _parse_direction(direction),
                                                 ctypes.byref(pol), ctypes.byref(opts), _ptr(lens)),
                      self.ctx.handle)
        return lens[: len(chans)]

    def propagate(self, circuit: Circuit, policy=None, direction=None, target_bucket_len: int = 1024,
                  min_buckets: int = 128) -> None:
        self._run(circuit, policy, direction, target_bucket_len, min_buckets)

    def propagate_with_stats(self, circuit: Circuit, policy=None, direction=None, target_bucket_len: int = 1024,
                             min_buckets: int = 128) -> PropagationStats:
        start = len(self)
        lens = self._run(circuit, policy, direction, target_bucket_len, min_buckets)
        outs = [int(v) for v in lens]
        return PropagationStats(terms_in=[start] + outs[:-1], terms_out=outs)

    @staticmethod
    def _channel(channel) -> _native.Channel:
        ch = _native.Channel()
        row = np.asarray(channel).reshape(-1)[0]
        for w in range(4):
            ch.gx[w] = int(row["gx"][w])
            ch.gz[w] = int(row["gz"][w])
        ch.theta = float(row["theta"])
        return ch

    def apply_rotation(self, channel, adjoint: bool = False, policy=None) -> None:
        """One rotation layer at the device's current bucketing (no rebucket, no layer pass)."""
        ch = self._channel(channel)
        pol = truncation.lower(policy)
        _native.check(_native.lib().pp_apply_rotation(self._h, ctypes.byref(ch), 1 if adjoint else 0,
                                                      ctypes.byref(pol)), self.ctx.handle)

    def apply_table(self, prepared, policy=None) -> None:
        """One prepared channel (``tables.unitary_table``, ``kraus_table``, ``rotation_table``, ``table``) as a
        layer under the policy's per-term filter, no layer pass (pp_apply_prepared)."""
        pol = truncation.lower(policy)
        _native.check(_native.lib().pp_apply_prepared(self._h, ctypes.byref(prepared), ctypes.byref(pol)),
                      self.ctx.handle)

    def apply_run(self, prepared, policy=None, permutations: bool = False) -> list:
        """A run of prepared channels (rotation tables, or one-to-one steps with ``permutations``) with the
        length after each (pp_apply_run)."""
        pol = truncation.lower(policy)
        arr = (_native.Prepared * len(prepared))(*prepared)
        lens = np.zeros(len(prepared), np.uint64)
        kind = _native.RUN_PERMUTATIONS if permutations else _native.RUN_ROTATIONS
        _native.check(_native.lib().pp_apply_run(self._h, kind, arr, len(prepared), ctypes.byref(pol),
                                                 lens.ctypes.data), self.ctx.handle)
        return [int(v) for v in lens]

    def apply_rotation_to(self, dst: "GpuPauliSum", channel, adjoint: bool = False, policy=None) -> None:
        """``dst`` <- one rotation layer applied to ``self``, which is left unchanged (asynchronous)."""
        ch = self._channel(channel)
        pol = truncation.lower(policy)
        _native.check(_native.lib().pp_apply_rotation_to(self._h, dst._h, ctypes.byref(ch), 1 if adjoint else 0,
                                                         ctypes.byref(pol)), self.ctx.handle)

    def autotune(self, circuit, policy=None) -> "_native.Tuning":
        """Time the device path's candidate sizes on this sum stepped through ``circuit``'s channels (each out of place
        from it), keep the fastest in the context and return it; this sum is left as it was."""
        chans = np.ascontiguousarray(circuit.channels() if isinstance(circuit, Circuit) else circuit)
        arr = (_native.Channel * max(1, len(chans)))()
        for i, row in enumerate(chans):
            arr[i] = self._channel(row)
        pol = truncation.lower(policy)
        out = _native.Tuning()
        _native.check(_native.lib().pp_autotune(self._h, arr, len(chans), ctypes.byref(pol), ctypes.byref(out)),
                      self.ctx.handle)
        return out

    def truncate_topn(self, n: int) -> None:
        _native.check(_native.lib().pp_truncate_topn(self._h, int(n)), self.ctx.handle)

    def truncate_approx_topn(self, n: int) -> None:
        """ApproxTopN(n)'s pass: whole octaves of |c|^2 from the top, the boundary octave filled in canonical order."""
        _native.check(_native.lib().pp_truncate_approx_topn(self._h, int(n)), self.ctx.handle)

    def expectation(self, state: str = "z+") -> float:
        ax, az, neg = _basis(state, self.num_qubits, self.width)
        out = ctypes.c_double()
        _native.check(_native.lib().pp_expectation_product(self._h, _ptr(ax), _ptr(az), _ptr(neg),
                                                           ctypes.byref(out)), self.ctx.handle)
        return out.value

    def __repr__(self) -> str:
        return f"GpuPauliSum(num_qubits={self.num_qubits}, terms={len(self)}, device={self.ctx.device})"


def propagated(initial: PauliSum, circuit: Circuit, policy=None, direction=None, **kw) -> PauliSum:
    out = initial.copy()
    out.propagate(circuit, policy, direction, **kw)
    return out


def propagated_with_stats(initial: PauliSum, circuit: Circuit, policy=None, direction=None, **kw):
    out = initial.copy()
    stats = out.propagate_with_stats(circuit, policy, direction, **kw)
    return out, stats


def propagate_batch(sums, circuits, policies=None, directions=None, target_bucket_len: int = 1024,
                    min_buckets: int = 128) -> None:
    """Propagate many resident sums (``GpuPauliSum``, one context, one width and coefficient type) each through its
    own circuit, stepped together: one launch of each batched kernel per layer step (pp_batch_propagate).  Every
    sum ends exactly as ``propagate`` would leave it."""
    sums = list(sums)
    circuits = list(circuits)
    if len(circuits) != len(sums):
        raise ValueError("one circuit per sum")
    if not sums:
        return
    ctx = sums[0].ctx
    policies = list(policies) if policies is not None else [None] * len(sums)
    directions = list(directions) if directions is not None else [None] * len(sums)
    chans, offsets = [], [0]
    for c in circuits:
        ch = c.channels() if isinstance(c, Circuit) else np.asarray(c)
        chans.extend(ch)
        offsets.append(offsets[-1] + len(ch))
    arr = (_native.Channel * max(1, len(chans)))()
    for i, row in enumerate(chans):
        arr[i] = GpuPauliSum._channel(row)
    offs = (ctypes.c_uint64 * len(offsets))(*offsets)
    dirs = (ctypes.c_int * len(sums))(*[_parse_direction(d) for d in directions])
    pols = (_native.Policy * len(sums))(*[truncation.lower(p) for p in policies])
    handles = (ctypes.c_void_p * len(sums))(*[s._h.value for s in sums])
    b = ctypes.c_void_p()
    _native.check(_native.lib().pp_batch_create(ctx.handle, handles, len(sums), ctypes.byref(b)), ctx.handle)
    try:
        opts = _options(target_bucket_len, min_buckets)
        _native.check(_native.lib().pp_batch_propagate(b, arr, offs, dirs, pols, ctypes.byref(opts)), ctx.handle)
    finally:
        _native.lib().pp_batch_free(b)
