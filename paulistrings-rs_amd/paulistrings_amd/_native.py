"""ctypes binding of libpp_hip.so, the C ABI declared in include/pp_hip.h.

The library is the product path: nothing here falls back to a CPU implementation.  A missing library or a
machine without a HIP device raises.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path

HERE = Path(__file__).resolve().parent
LIB_PATH = Path(os.environ.get("PP_HIP_LIBRARY", HERE / "libpp_hip.so"))

PP_OK = 0
STATUS_NAMES = {
    0: "PP_OK",
    1: "PP_ERR_NO_DEVICE",
    2: "PP_ERR_OUT_OF_MEMORY",
    3: "PP_ERR_UNSUPPORTED",
    4: "PP_ERR_DRIVER",
    5: "PP_ERR_INVALID_ARGUMENT",
}
COEFF_F64, COEFF_F32 = 0, 1
FORWARD, HEISENBERG = 0, 1


class KeepNode(ctypes.Structure):
    _fields_ = [("op", ctypes.c_int32), ("k", ctypes.c_uint32), ("eps", ctypes.c_double)]


class Policy(ctypes.Structure):
    _fields_ = [
        ("n_nodes", ctypes.c_int32),
        ("nodes", KeepNode * 15),
        ("has_topn", ctypes.c_int32),
        ("topn", ctypes.c_uint64),
    ]


class Options(ctypes.Structure):
    _fields_ = [("target_bucket_len", ctypes.c_uint32), ("min_buckets", ctypes.c_uint32)]


class Tuning(ctypes.Structure):
    """pp_tuning: the device path's sizes (bucket_target, stage_rows, groups_per_cu, join_block, join_in_place)."""

    _fields_ = [("bucket_target", ctypes.c_uint32), ("stage_rows", ctypes.c_uint32),
                ("groups_per_cu", ctypes.c_uint32), ("join_block", ctypes.c_uint32),
                ("join_in_place", ctypes.c_uint32)]


ONE_PAIR_PER_GROUP = 0xFFFFFFFF


class Channel(ctypes.Structure):
    _fields_ = [("gx", ctypes.c_uint64 * 4), ("gz", ctypes.c_uint64 * 4), ("theta", ctypes.c_double)]


class DeltaEntry(ctypes.Structure):
    """pp_delta_entry: one key delta of a prepared local table (DeltaEntry, channel/prepared.rs:35-46)."""

    _fields_ = [("local_delta", ctypes.c_uint32), ("amplitude", ctypes.c_double * 16)]


PREPARED_LOCAL, PREPARED_ROTATION = 0, 1
RUN_ROTATIONS, RUN_PERMUTATIONS = 0, 1


class Prepared(ctypes.Structure):
    """pp_prepared: a channel prepared for one layer (Prepared, channel/prepared.rs:343-350)."""

    _fields_ = [
        ("kind", ctypes.c_int32),
        ("k", ctypes.c_uint32),
        ("qubits", ctypes.c_uint32 * 2),
        ("n_deltas", ctypes.c_uint32),
        ("deltas", DeltaEntry * 16),
        ("gx", ctypes.c_uint64 * 4),
        ("gz", ctypes.c_uint64 * 4),
        ("cos", ctypes.c_double),
        ("sin", ctypes.c_double),
    ]


class NativeError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class DeviceMemoryError(NativeError, MemoryError):
    pass


_lib = None


def lib() -> ctypes.CDLL:
    """Load libpp_hip.so; raises ImportError if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                          "(or paulistrings-rs_amd/csrc/build.sh)")
    L = ctypes.CDLL(str(LIB_PATH))
    vp, u64, i32, u32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int32, ctypes.c_uint32
    pvp = ctypes.POINTER(ctypes.c_void_p)
    L.pp_context_create.argtypes = [ctypes.c_int, pvp]
    L.pp_context_destroy.argtypes = [vp]
    L.pp_context_destroy.restype = None
    L.pp_last_error.argtypes = [vp]
    L.pp_last_error.restype = ctypes.c_char_p
    L.pp_sync.argtypes = [vp]
    L.pp_context_stream.argtypes = [vp, pvp]
    L.pp_sum_upload.argtypes = [vp, u32, u32, ctypes.c_int, u64, vp, vp, vp, u64, pvp]
    L.pp_sum_clone.argtypes = [vp, pvp]
    L.pp_sum_copy.argtypes = [vp, vp]
    L.pp_sum_free.argtypes = [vp]
    L.pp_sum_len.argtypes = [vp, ctypes.POINTER(u64)]
    L.pp_sum_download.argtypes = [vp, vp, vp, vp]
    L.pp_sum_bucket_bits.argtypes = [vp, ctypes.POINTER(i32), ctypes.POINTER(i32)]
    L.pp_sum_check.argtypes = [vp, vp]
    L.pp_propagate.argtypes = [vp, vp, u64, ctypes.c_int, ctypes.POINTER(Policy), ctypes.POINTER(Options), vp]
    L.pp_apply_rotation.argtypes = [vp, ctypes.POINTER(Channel), ctypes.c_int, ctypes.POINTER(Policy)]
    L.pp_apply_rotation_to.argtypes = [vp, vp, ctypes.POINTER(Channel), ctypes.c_int, ctypes.POINTER(Policy)]
    L.pp_truncate_topn.argtypes = [vp, u64]
    L.pp_truncate_approx_topn.argtypes = [vp, u64]
    L.pp_batch_create.argtypes = [vp, ctypes.POINTER(vp), u32, pvp]
    L.pp_batch_free.argtypes = [vp]
    L.pp_batch_propagate.argtypes = [vp, ctypes.POINTER(Channel), ctypes.POINTER(u64), ctypes.POINTER(ctypes.c_int),
                                     ctypes.POINTER(Policy), ctypes.POINTER(Options)]
    L.pp_get_tuning.argtypes = [vp, ctypes.POINTER(Tuning)]
    L.pp_set_tuning.argtypes = [vp, ctypes.POINTER(Tuning)]
    L.pp_autotune.argtypes = [vp, ctypes.POINTER(Channel), u64, ctypes.POINTER(Policy), ctypes.POINTER(Tuning)]
    L.pp_graph_capture_begin.argtypes = [vp]
    L.pp_graph_capture_end.argtypes = [vp, pvp]
    L.pp_graph_launch.argtypes = [vp]
    L.pp_graph_free.argtypes = [vp]
    L.pp_expectation_product.argtypes = [vp, vp, vp, vp, ctypes.POINTER(ctypes.c_double)]
    L.pp_profile_enable.argtypes = [vp, ctypes.c_int]
    L.pp_profile_read.argtypes = [vp, ctypes.POINTER(u64), ctypes.POINTER(ctypes.c_double), ctypes.POINTER(u64),
                                  ctypes.POINTER(u64)]
    L.pp_profile_read_score.argtypes = [vp, ctypes.POINTER(u64), ctypes.POINTER(ctypes.c_double), ctypes.POINTER(u64)]
    L.pp_prepare_channel.argtypes = [u32, u32, ctypes.c_int, ctypes.POINTER(Channel), ctypes.c_int,
                                     ctypes.POINTER(Prepared)]
    L.pp_apply_prepared.argtypes = [vp, ctypes.POINTER(Prepared), ctypes.POINTER(Policy)]
    L.pp_apply_run.argtypes = [vp, ctypes.c_int, ctypes.POINTER(Prepared), u32, ctypes.POINTER(Policy), vp]
    L.pp_finalize_layer.argtypes = [vp, ctypes.POINTER(Policy)]
    L.pp_settle.argtypes = [vp]
    L.pp_sum_refine.argtypes = [vp]
    L.pp_sum_coarsen_to.argtypes = [vp, i32]
    L.pp_stream_probe.argtypes = [vp, u64, u32, ctypes.POINTER(ctypes.c_double)]
    L.pp_apply_rotation_to_noalloc.argtypes = [vp, vp, ctypes.POINTER(Channel), ctypes.c_int, ctypes.POINTER(Policy)]
    L.pp_hash_rows.argtypes = [u32, u64, u32, vp, vp]
    L.pp_hash_rows.restype = i32
    for name in ("pp_context_create", "pp_sync", "pp_context_stream", "pp_sum_upload", "pp_sum_clone",
                 "pp_sum_free", "pp_sum_len", "pp_sum_download", "pp_sum_bucket_bits", "pp_propagate",
                 "pp_apply_rotation", "pp_apply_rotation_to", "pp_truncate_topn", "pp_truncate_approx_topn",
                 "pp_graph_capture_begin", "pp_graph_capture_end", "pp_graph_launch", "pp_graph_free",
                 "pp_get_tuning", "pp_set_tuning", "pp_autotune", "pp_batch_create", "pp_batch_free",
                 "pp_batch_propagate",
                 "pp_expectation_product", "pp_profile_enable", "pp_profile_read", "pp_profile_read_score",
                 "pp_prepare_channel", "pp_apply_prepared", "pp_apply_run", "pp_finalize_layer", "pp_settle",
                 "pp_sum_refine", "pp_sum_coarsen_to", "pp_stream_probe", "pp_apply_rotation_to_noalloc",
                 "pp_sum_copy", "pp_sum_check"):
        getattr(L, name).restype = ctypes.c_int
    _lib = L
    return L


def check(status: int, ctx=None) -> None:
    if status == PP_OK:
        return
    msg = lib().pp_last_error(ctx).decode() if ctx else ""
    if status == 2:
        raise DeviceMemoryError(status, msg)
    if status == 5:
        raise ValueError(f"{STATUS_NAMES[status]}: {msg}")
    raise NativeError(status, msg)


class Context:
    """One HIP device and the stream every call of this context runs on."""

    _cache: dict[int, "Context"] = {}

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        st = lib().pp_context_create(int(device), ctypes.byref(h))
        if st == 1:
            raise RuntimeError("no HIP device is visible: this engine runs on an MI355X (gfx950)")
        check(st)
        self.handle = h
        self.device = int(device)

    @classmethod
    def get(cls, device: int = 0) -> "Context":
        if device not in cls._cache:
            cls._cache[device] = cls(device)
        return cls._cache[device]

    def sync(self) -> None:
        check(lib().pp_sync(self.handle), self.handle)

    def stream(self) -> int:
        p = ctypes.c_void_p()
        check(lib().pp_context_stream(self.handle, ctypes.byref(p)), self.handle)
        return p.value or 0

    def profile_enable(self, on: bool = True) -> None:
        check(lib().pp_profile_enable(self.handle, 1 if on else 0), self.handle)

    def profile_read(self):
        n = ctypes.c_uint64()
        ms = ctypes.c_double()
        rin = ctypes.c_uint64()
        rout = ctypes.c_uint64()
        check(lib().pp_profile_read(self.handle, ctypes.byref(n), ctypes.byref(ms), ctypes.byref(rin),
                                    ctypes.byref(rout)), self.handle)
        return n.value, ms.value, rin.value, rout.value

    def profile_read_score(self):
        """(launches, total milliseconds, terms scored) of the learned pass's score kernel since profiling began."""
        n = ctypes.c_uint64()
        ms = ctypes.c_double()
        terms = ctypes.c_uint64()
        check(lib().pp_profile_read_score(self.handle, ctypes.byref(n), ctypes.byref(ms), ctypes.byref(terms)),
              self.handle)
        return n.value, ms.value, terms.value

    def get_tuning(self) -> Tuning:
        t = Tuning()
        check(lib().pp_get_tuning(self.handle, ctypes.byref(t)), self.handle)
        return t

    def set_tuning(self, bucket_target: int, stage_rows: int = 0, groups_per_cu: int = ONE_PAIR_PER_GROUP,
                   join_block: int = 256, join_in_place: int | None = None) -> None:
        if join_in_place is None:
            join_in_place = self.get_tuning().join_in_place
        t = Tuning(bucket_target, stage_rows, groups_per_cu, join_block, join_in_place)
        check(lib().pp_set_tuning(self.handle, ctypes.byref(t)), self.handle)

    def capture(self) -> "GraphCapture":
        """Capture the layers queued inside the ``with`` block into a HIP graph (see pp_graph_capture_begin)."""
        return GraphCapture(self)


class Graph:
    """A captured layer sequence (pp_graph): ``launch()`` replays it on its context's stream."""

    def __init__(self, ctx: "Context", handle):
        self.ctx = ctx
        self._h = handle

    def launch(self) -> None:
        check(lib().pp_graph_launch(self._h), self.ctx.handle)

    def free(self) -> None:
        if self._h:
            check(lib().pp_graph_free(self._h), self.ctx.handle)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class GraphCapture:
    """``with ctx.capture() as cap: ...`` then ``cap.graph``; an exception inside the block discards the capture."""

    def __init__(self, ctx: "Context"):
        self.ctx = ctx
        self.graph: Graph | None = None

    def __enter__(self) -> "GraphCapture":
        check(lib().pp_graph_capture_begin(self.ctx.handle), self.ctx.handle)
        return self

    def __exit__(self, exc_type, exc, tb):
        h = ctypes.c_void_p()
        status = lib().pp_graph_capture_end(self.ctx.handle, ctypes.byref(h))
        if exc_type is not None:
            if status == PP_OK and h.value:
                lib().pp_graph_free(h)
            return False
        check(status, self.ctx.handle)
        self.graph = Graph(self.ctx, h.value)
        return False
