"""MI355X-native RecSplit minimal perfect hash construction.

Python mirror of the reference's GPURecSplit surface over the C ABI of
librecsplit_amd.so (include/gpurecsplit_amd.h):

    rs = GPURecSplit(keys, bucket_size, leaf_size=16)   # GPURecSplit<16>(keys, b, t)
    rs(key) / rs.query(keys)                            # operator()(const hash128_t&)
    rs.serialize() / GPURecSplit.deserialize(...)       # operator<< / operator>>
    rs.get_bits()                                       # getBits()

Keys are 128-bit hashes given as an (n, 2) uint64 array of (first, second)
pairs, the memory layout of the reference's hash128_t (include/util/Hash128.h:14-23).
The GPU build is the only construction path: there is no CPU fallback, and the
import fails loudly when the HIP library is missing.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import weakref
from pathlib import Path

import numpy as np

__all__ = ["GPURecSplit", "RecSplitError", "Context", "Plan", "build_batch", "build_noalloc", "golomb_memo", "hash_strings",
           "LIB_PATH", "STATUS"]

LIB_PATH = Path(__file__).resolve().parent / "librecsplit_amd.so"

STATUS = {
    0: "ok", 1: "invalid argument", 2: "HIP runtime error", 3: "bucket with 3000 or more keys",
    4: "duplicate keys", 5: "seed limit exceeded", 6: "Elias-Fano offset overflow", 7: "out of memory",
    8: "malformed serialized stream", 9: "no device copy", 10: "time limit exceeded", 11: "internal error",
}

KIND_NAMES = ["fanout2", "upper_aggr", "lower_aggr", "leaf_rotation", "leaf_bruteforce"]

# Flags word passed as the C ABI's `rotate` argument.
FLAG_ROTATE = 1    # rotation fitting for full leaves (USE_BIJECTIONS_ROTATE)
FLAG_REMAP64 = 2   # 64-bit remap of the RecSplitRotate variant (USE_64_BIT_REMAP)


def _flags(rotate: bool, remap64: bool) -> int:
    return (FLAG_ROTATE if rotate else 0) | (FLAG_REMAP64 if remap64 else 0)


class RecSplitError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


class BuildStats(C.Structure):
    _fields_ = [
        ("total_ms", C.c_double), ("partition_ms", C.c_double), ("search_ms", C.c_double * 8),
        ("split_ms", C.c_double), ("encode_ms", C.c_double), ("evals", C.c_uint64 * 8),
        ("nodes", C.c_uint64 * 8), ("executed_evals", C.c_uint64 * 8), ("num_buckets", C.c_uint64), ("num_nodes", C.c_uint64),
        ("rice_bits", C.c_uint64), ("max_bucket_size", C.c_uint32), ("search_waves", C.c_uint32),
        ("search_busy", C.c_double * 8), ("search_tail_ms", C.c_double * 8), ("seeds", C.c_uint64 * 8),
        ("kind_waves", C.c_uint32 * 8),
    ]

    def as_dict(self) -> dict:
        return {
            "total_ms": self.total_ms, "partition_ms": self.partition_ms, "split_ms": self.split_ms,
            "encode_ms": self.encode_ms,
            "search_ms": {k: self.search_ms[i] for i, k in enumerate(KIND_NAMES)},
            "evals": {k: int(self.evals[i]) for i, k in enumerate(KIND_NAMES)},
            "nodes": {k: int(self.nodes[i]) for i, k in enumerate(KIND_NAMES)},
            "executed_evals": {k: int(self.executed_evals[i]) for i, k in enumerate(KIND_NAMES)},
            "num_buckets": int(self.num_buckets), "num_nodes": int(self.num_nodes),
            "rice_bits": int(self.rice_bits), "max_bucket_size": int(self.max_bucket_size),
            "search_waves": int(self.search_waves),
            "search_busy": {k: self.search_busy[i] for i, k in enumerate(KIND_NAMES)},
            "search_tail_ms": {k: self.search_tail_ms[i] for i, k in enumerate(KIND_NAMES)},
            "seeds": {k: int(self.seeds[i]) for i, k in enumerate(KIND_NAMES)},
            "kind_waves": {k: int(self.kind_waves[i]) for i, k in enumerate(KIND_NAMES)},
        }


def _load():
    # One HIP runtime per process: torch ships its own libamdhip64 (same SONAME,
    # loaded under another file name by torch's libraries).  Importing torch first
    # makes this library bind to that already-loaded runtime instead of loading a
    # second one from /opt/rocm.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is not built: run `make -C gpurecsplit_amd/csrc` (hipcc, gfx950)")
    lib = C.CDLL(str(LIB_PATH))
    vp, u64p, u32p = C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)
    sig = {
        "rs_ctx_create": (C.c_int, [C.c_int, C.POINTER(vp)]),
        "rs_ctx_destroy": (None, [vp]),
        "rs_ctx_set_time_limit": (C.c_int, [vp, C.c_double]),
        "rs_build": (C.c_int, [vp, u64p, C.c_size_t, C.c_uint32, C.c_size_t, C.c_int, C.POINTER(vp)]),
        "rs_build_device": (C.c_int, [vp, C.c_void_p, C.c_size_t, C.c_uint32, C.c_size_t, C.c_int, C.POINTER(vp)]),
        "rs_mphf_serialized_size": (C.c_size_t, [vp]),
        "rs_mphf_serialize": (C.c_int, [vp, C.c_char_p, C.c_size_t]),
        "rs_mphf_deserialize": (C.c_int, [C.c_char_p, C.c_size_t, C.c_uint32, C.c_int, C.POINTER(vp)]),
        "rs_mphf_query": (C.c_uint64, [vp, C.c_uint64, C.c_uint64]),
        "rs_mphf_query_batch": (None, [vp, u64p, C.c_size_t, u64p]),
        "rs_query_device": (C.c_int, [vp, vp, C.c_void_p, C.c_size_t, C.c_void_p]),
        "rs_autotune": (C.c_int, [vp, C.c_uint32, C.c_size_t, C.c_int, C.c_uint64, C.c_char_p, C.c_size_t]),
        "rs_ctx_set_tune": (C.c_int, [vp, C.c_uint32, C.c_char_p]),
        "rs_ctx_get_tune": (C.c_int, [vp, C.c_uint32, C.c_char_p, C.c_size_t]),
        "rs_ctx_device_bytes": (C.c_uint64, [vp]),
        "rs_ctx_release": (C.c_int, [vp]),
        "rs_build_device_noalloc": (C.c_int, [vp, C.c_void_p, C.c_size_t, C.c_uint32, C.c_size_t, C.c_int,
                                              C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]),
        "rs_build_batch": (C.c_int, [vp, C.c_void_p, u64p, C.c_size_t, C.c_uint32, C.c_size_t, C.c_int,
                                     C.POINTER(vp), C.POINTER(BuildStats)]),
        "rs_mphf_bits": (C.c_uint64, [vp]),
        "rs_mphf_num_keys": (C.c_uint64, [vp]),
        "rs_mphf_stats": (C.c_int, [vp, C.POINTER(BuildStats)]),
        "rs_mphf_free": (None, [vp]),
        "rs_golomb_memo": (C.c_int, [C.c_uint32, u32p]),
        "rs_last_error": (C.c_char_p, []),
        "rs_spooky128": (None, [C.c_char_p, C.c_size_t, u64p, u64p]),
        "rs_first_hash": (None, [C.c_char_p, C.c_size_t, u64p]),
        "rs_hash_strings": (None, [C.c_char_p, u64p, C.c_size_t, u64p]),
        "rs_hash_strings_device": (C.c_int, [vp, C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p]),
        "rs_build_strings": (C.c_int, [vp, C.c_char_p, u64p, C.c_size_t, C.c_uint32, C.c_size_t, C.c_int,
                                       C.POINTER(vp)]),
        "rs_mphf_query_string": (C.c_uint64, [vp, C.c_char_p, C.c_size_t]),
        "rs_plan_create": (C.c_int, [vp, C.c_size_t, C.c_uint32, C.c_size_t, C.c_int, C.POINTER(vp)]),
        "rs_plan_enqueue": (C.c_int, [vp, C.c_void_p, C.c_void_p]),
        "rs_plan_result": (C.c_int, [vp, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]),
        "rs_plan_device_bytes": (C.c_size_t, [vp]),
        "rs_plan_destroy": (None, [vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()

# Every live context, structure and server, released in a fixed order (servers,
# then structures, then contexts) by an exit hook that runs before interpreter
# teardown -- while the HIP runtime and torch are still fully alive -- instead of
# by __del__ calls in whatever order module teardown reaches them.
_live = {"server": weakref.WeakSet(), "plan": weakref.WeakSet(), "mphf": weakref.WeakSet(), "ctx": weakref.WeakSet()}


@atexit.register
def _release_all():
    for kind in ("server", "plan", "mphf", "ctx"):
        for obj in list(_live[kind]):
            obj.close()


def _check(code: int):
    if code != 0:
        raise RecSplitError(code, (lib.rs_last_error() or b"").decode())


def _keys_array(keys) -> np.ndarray:
    k = np.ascontiguousarray(keys, dtype=np.uint64)
    if k.ndim != 2 or k.shape[1] != 2:
        raise ValueError("keys must be an (n, 2) uint64 array of (first, second) pairs")
    return k


def _sync_torch(t):
    """The library runs on its own stream: finish torch's pending work on t first."""
    import torch
    torch.cuda.current_stream(t.device).synchronize()


def _pack_strings(keys) -> tuple[bytes, np.ndarray]:
    bs = [k.encode() if isinstance(k, str) else bytes(k) for k in keys]
    off = np.zeros(len(bs) + 1, dtype=np.uint64)
    if bs:
        off[1:] = np.cumsum([len(b) for b in bs])
    return b"".join(bs), off


def hash_strings(keys) -> np.ndarray:
    """The reference's first_hash (SpookyHash V2, seeds 0) of each string or bytes
    key, on the host: an (n, 2) uint64 array of (first, second)."""
    blob, off = _pack_strings(keys)
    out = np.empty((len(off) - 1, 2), dtype=np.uint64)
    lib.rs_hash_strings(blob, off.ctypes.data_as(C.POINTER(C.c_uint64)), out.shape[0],
                        out.ctypes.data_as(C.POINTER(C.c_uint64)))
    return out


def golomb_memo(leaf_size: int) -> np.ndarray:
    """Golomb-Rice memo of the product (golomb << 27 | nodes << 16 | fixed bits)."""
    out = np.zeros(3000, dtype=np.uint32)
    _check(lib.rs_golomb_memo(leaf_size, out.ctypes.data_as(C.POINTER(C.c_uint32))))
    return out


class Context:
    """A build context on one GPU (device ordinal as seen by HIP)."""

    _default: dict[int, "Context"] = {}

    def __init__(self, device: int = 0, time_limit_s: float | None = None):
        h = C.c_void_p()
        _check(lib.rs_ctx_create(device, C.byref(h)))
        self.h = h
        self.device = device
        _live["ctx"].add(self)
        if time_limit_s is not None:
            self.set_time_limit(time_limit_s)

    def set_time_limit(self, seconds: float):
        """Bounds every later build on this context (0: no limit); a build past
        the limit raises RecSplitError with code 10 and the context stays usable."""
        _check(lib.rs_ctx_set_time_limit(self.h, float(seconds)))

    def autotune(self, leaf_size: int, bucket_size: int, rotate: bool = True, remap64: bool = False,
                 sample_keys: int = 0) -> dict:
        """Times synthetic builds of this leaf size and bucket size under candidate
        launch settings and keeps the fastest as the context's tuning for the leaf
        size; returns the tuner's report."""
        import json
        buf = C.create_string_buffer(1 << 16)
        _check(lib.rs_autotune(self.h, leaf_size, bucket_size, _flags(rotate, remap64), sample_keys, buf, len(buf)))
        return json.loads(buf.value.decode())

    def set_tune(self, tune: str, leaf_size: int = 0):
        """Applies a tuning string (RS_TUNE syntax) to a leaf size (0: all)."""
        _check(lib.rs_ctx_set_tune(self.h, leaf_size, tune.encode()))

    def get_tune(self, leaf_size: int = 0) -> str:
        buf = C.create_string_buffer(1024)
        _check(lib.rs_ctx_get_tune(self.h, leaf_size, buf, len(buf)))
        return buf.value.decode()

    def device_bytes(self) -> int:
        """Device memory the context holds for reuse by later builds."""
        return int(lib.rs_ctx_device_bytes(self.h))

    @classmethod
    def default(cls, device: int = 0) -> "Context":
        if device not in cls._default:
            cls._default[device] = cls(device)
        return cls._default[device]

    def close(self, _destroy=lib.rs_ctx_destroy):
        if getattr(self, "h", None):
            _destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class GPURecSplit:
    """RecSplit MPHF built on the GPU (GPURecSplit.cuh:409-842 surface).

    GPURecSplit(keys, bucket_size, leaf_size=..., rotate=True, num_threads=1)
      keys: (n, 2) uint64 host array, or a torch uint64/int64 CUDA tensor of shape
      (n, 2) already resident on the context's GPU (no host round trip).
    num_threads is accepted for signature parity with the reference (it sizes
    the host sort there); construction here runs entirely on the GPU.
    """

    def __init__(self, keys=None, bucket_size: int = 100, leaf_size: int = 8, rotate: bool = True,
                 num_threads: int = 1, ctx: Context | None = None, _handle=None, remap64: bool = False):
        self.leaf_size = leaf_size
        self.rotate = rotate
        self.remap64 = remap64
        self.bucket_size = bucket_size
        self._ctx = ctx
        if _handle is not None:
            self.h = _handle
            _live["mphf"].add(self)
            return
        self._ctx = ctx or Context.default()
        h = C.c_void_p()
        if hasattr(keys, "data_ptr") and getattr(keys, "is_cuda", False):
            if keys.dim() != 2 or keys.shape[1] != 2 or keys.element_size() != 8 or not keys.is_contiguous():
                raise ValueError("device keys must be a contiguous (n, 2) 64-bit tensor")
            _sync_torch(keys)
            _check(lib.rs_build_device(self._ctx.h, C.c_void_p(keys.data_ptr()), keys.shape[0], leaf_size,
                                       bucket_size, _flags(rotate, remap64), C.byref(h)))
        else:
            k = _keys_array(keys)
            _check(lib.rs_build(self._ctx.h, k.ctypes.data_as(C.POINTER(C.c_uint64)), k.shape[0], leaf_size,
                                bucket_size, _flags(rotate, remap64), C.byref(h)))
        self.h = h
        _live["mphf"].add(self)

    @classmethod
    def from_strings(cls, keys, bucket_size: int, leaf_size: int = 8, rotate: bool = True,
                     ctx: Context | None = None, remap64: bool = False) -> "GPURecSplit":
        """GPURecSplit(const vector<string>&, bucket_size, numThreads) (GPURecSplit.cuh:443-478):
        string or bytes keys, hashed with SpookyHash V2 on the GPU."""
        ctx = ctx or Context.default()
        blob, off = _pack_strings(keys)
        h = C.c_void_p()
        _check(lib.rs_build_strings(ctx.h, blob, off.ctypes.data_as(C.POINTER(C.c_uint64)), len(off) - 1, leaf_size,
                                    bucket_size, _flags(rotate, remap64), C.byref(h)))
        return cls(bucket_size=bucket_size, leaf_size=leaf_size, rotate=rotate, ctx=ctx, _handle=h)

    def query_string(self, key) -> int:
        """operator()(const string&) (AbstractParallelRecSplit.hpp:473)."""
        b = key.encode() if isinstance(key, str) else bytes(key)
        return int(lib.rs_mphf_query_string(self.h, b, len(b)))

    # operator>> (GPURecSplit.cuh:827-841)
    @classmethod
    def deserialize(cls, data: bytes, leaf_size: int, rotate: bool = True, remap64: bool = False) -> "GPURecSplit":
        """The serialized layout does not record the rotation or remap setting: pass
        the ones the structure was built with (the reference fixes them by type)."""
        h = C.c_void_p()
        _check(lib.rs_mphf_deserialize(data, len(data), leaf_size, _flags(rotate, remap64), C.byref(h)))
        import struct
        bucket = struct.unpack_from("<Q", data, 8)[0]
        return cls(bucket_size=bucket, leaf_size=leaf_size, rotate=rotate, _handle=h, remap64=remap64)

    # operator<< (GPURecSplit.cuh:817-825)
    def serialize(self) -> bytes:
        n = lib.rs_mphf_serialized_size(self.h)
        buf = C.create_string_buffer(n)
        _check(lib.rs_mphf_serialize(self.h, buf, n))
        return buf.raw

    # operator()(const hash128_t&) (AbstractParallelRecSplit.hpp:409-466)
    def __call__(self, key) -> int:
        first, second = key
        return int(lib.rs_mphf_query(self.h, int(first), int(second)))

    def query(self, keys) -> np.ndarray:
        """Host queries of many keys."""
        k = _keys_array(keys)
        out = np.empty(k.shape[0], dtype=np.uint64)
        lib.rs_mphf_query_batch(self.h, k.ctypes.data_as(C.POINTER(C.c_uint64)), k.shape[0],
                                out.ctypes.data_as(C.POINTER(C.c_uint64)))
        return out

    def query_device(self, d_keys, d_out=None):
        """Batched GPU queries: d_keys is a contiguous (n, 2) 64-bit CUDA tensor; returns
        (or fills) an (n,) int64 CUDA tensor of results.  A structure without a copy
        on that GPU (deserialized, or from a batch) is uploaded on first use."""
        import torch
        if d_out is None:
            d_out = torch.empty(d_keys.shape[0], dtype=torch.int64, device=d_keys.device)
        _sync_torch(d_keys)
        ctx = self._ctx or Context.default(d_keys.device.index or 0)
        _check(lib.rs_query_device(ctx.h, self.h, C.c_void_p(d_keys.data_ptr()), d_keys.shape[0],
                                   C.c_void_p(d_out.data_ptr())))
        return d_out

    # getBits() (AbstractParallelRecSplit.hpp:476-478), without the object-size term
    def get_bits(self) -> int:
        return int(lib.rs_mphf_bits(self.h))

    @property
    def num_keys(self) -> int:
        return int(lib.rs_mphf_num_keys(self.h))

    def stats(self) -> dict:
        s = BuildStats()
        _check(lib.rs_mphf_stats(self.h, C.byref(s)))
        return s.as_dict()

    def close(self, _free=lib.rs_mphf_free):
        """Frees the structure (host and device copies); later calls do nothing."""
        if getattr(self, "h", None):
            _free(self.h)
            self.h = None

    def __del__(self):
        self.close()


def build_batch(d_keys, set_offsets, bucket_size: int, leaf_size: int, rotate: bool = True,
                ctx: Context | None = None, remap64: bool = False):
    """Builds one MPHF per key set in one batch on the GPU.

    d_keys: contiguous (n, 2) 64-bit CUDA tensor holding all sets back to back;
    set_offsets: num_sets + 1 increasing key offsets (set i = keys[off[i]:off[i+1]]).
    Returns (list of GPURecSplit, batch stats dict).  Each result equals the
    single-set build of its keys."""
    ctx = ctx or Context.default()
    if d_keys.dim() != 2 or d_keys.shape[1] != 2 or d_keys.element_size() != 8 or not d_keys.is_contiguous():
        raise ValueError("d_keys must be a contiguous (n, 2) 64-bit CUDA tensor")
    off = np.ascontiguousarray(set_offsets, dtype=np.uint64)
    S = off.shape[0] - 1
    if off[-1] > d_keys.shape[0]:
        raise ValueError("set offsets exceed the key array")
    _sync_torch(d_keys)
    handles = (C.c_void_p * S)()
    st = BuildStats()
    _check(lib.rs_build_batch(ctx.h, C.c_void_p(d_keys.data_ptr()), off.ctypes.data_as(C.POINTER(C.c_uint64)), S,
                              leaf_size, bucket_size, _flags(rotate, remap64), handles, C.byref(st)))
    out = [GPURecSplit(bucket_size=bucket_size, leaf_size=leaf_size, rotate=rotate, ctx=ctx, _handle=C.c_void_p(h),
                       remap64=remap64) for h in handles]
    return out, st.as_dict()


class Plan:
    """A build plan (rs_plan_*): every device buffer of the construction of one
    MPHF of n keys, allocated once; `enqueue` queues a whole build on a stream with
    no allocation and no host synchronisation, so it can be recorded into a CUDA/HIP
    graph (torch.cuda.graph) and replayed for new keys in the same device buffer.

        plan = Plan(n, bucket_size=100, leaf_size=8)
        plan.enqueue(d_keys)          # current stream; also inside torch.cuda.graph
        torch.cuda.synchronize()
        data = plan.result()          # serialized bytes == GPURecSplit(d_keys, ...).serialize()
    """

    def __init__(self, n: int, bucket_size: int, leaf_size: int, rotate: bool = True, remap64: bool = False,
                 ctx: Context | None = None):
        self._ctx = ctx or Context.default()
        self.n, self.bucket_size, self.leaf_size = n, bucket_size, leaf_size
        self.rotate, self.remap64 = rotate, remap64
        h = C.c_void_p()
        _check(lib.rs_plan_create(self._ctx.h, n, leaf_size, bucket_size, _flags(rotate, remap64), C.byref(h)))
        self.h = h
        _live["plan"].add(self)

    @property
    def device_bytes(self) -> int:
        return int(lib.rs_plan_device_bytes(self.h))

    def enqueue(self, d_keys, stream=None):
        """Queues a build of d_keys (a contiguous (n, 2) 64-bit CUDA tensor) on
        `stream` (a torch.cuda.Stream; None: torch's current stream)."""
        if d_keys.dim() != 2 or d_keys.shape[0] != self.n or d_keys.shape[1] != 2 or d_keys.element_size() != 8 \
                or not d_keys.is_contiguous():
            raise ValueError(f"d_keys must be a contiguous ({self.n}, 2) 64-bit CUDA tensor")
        import torch
        st = stream if stream is not None else torch.cuda.current_stream(d_keys.device)
        _check(lib.rs_plan_enqueue(self.h, C.c_void_p(d_keys.data_ptr()), C.c_void_p(st.cuda_stream)))

    def result(self) -> bytes:
        """The serialized structure of the last build (after its stream's work is
        done); raises RecSplitError for a failed build."""
        n = C.c_size_t()
        _check(lib.rs_plan_result(self.h, None, 0, C.byref(n)))
        buf = C.create_string_buffer(n.value)
        _check(lib.rs_plan_result(self.h, buf, n.value, C.byref(n)))
        return buf.raw[: n.value]

    def build(self, d_keys) -> bytes:
        """enqueue on the current stream, wait, and return result()."""
        import torch
        self.enqueue(d_keys)
        torch.cuda.current_stream(d_keys.device).synchronize()
        return self.result()

    def close(self, _destroy=lib.rs_plan_destroy):
        if getattr(self, "h", None):
            _destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def build_noalloc(d_keys, bucket_size: int, leaf_size: int, rotate: bool = True, ctx: Context | None = None,
                  out_capacity: int | None = None, remap64: bool = False) -> bytes:
    """Allocation-free build (rs_build_device_noalloc): the serialized MPHF of the
    keys in d_keys (a contiguous (n, 2) 64-bit CUDA tensor), built entirely in the
    library's fixed per-GPU scratch array.  out_capacity bounds the host output
    buffer (default: 4 bytes per key + 64 KiB, ample for every leaf size)."""
    ctx = ctx or Context.default()
    if d_keys.dim() != 2 or d_keys.shape[1] != 2 or d_keys.element_size() != 8 or not d_keys.is_contiguous():
        raise ValueError("d_keys must be a contiguous (n, 2) 64-bit CUDA tensor")
    n = d_keys.shape[0]
    cap = out_capacity if out_capacity is not None else 4 * n + (64 << 10)
    buf = C.create_string_buffer(cap)
    ln = C.c_size_t(0)
    _sync_torch(d_keys)
    _check(lib.rs_build_device_noalloc(ctx.h, C.c_void_p(d_keys.data_ptr()), n, leaf_size, bucket_size,
                                       _flags(rotate, remap64),
                                       buf, cap, C.byref(ln)))
    return buf.raw[: ln.value]
