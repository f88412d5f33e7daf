"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module.  The product never does.
"""
import ctypes as C
import os
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
LIB_PATH = ROOT / "oracle" / "liboracle.so"

_u64p = C.POINTER(C.c_uint64)
_u32p = C.POINTER(C.c_uint32)


def _load():
    if not LIB_PATH.exists():
        raise RuntimeError(f"oracle library missing: {LIB_PATH} (run `make -C oracle`)")
    lib = C.CDLL(str(LIB_PATH))
    sig = {
        "orc_remix": (C.c_uint64, [C.c_uint64]),
        "orc_remix32": (C.c_uint32, [C.c_uint32]),
        "orc_remap": (C.c_uint32, [C.c_uint64, C.c_uint32]),
        "orc_remap128": (C.c_uint64, [C.c_uint64, C.c_uint64]),
        "orc_splitmix64_keys": (None, [C.c_uint64, C.c_size_t, _u64p]),
        "orc_golomb_memo": (C.c_int, [C.c_int, _u32p]),
        "orc_aggr_bounds": (C.c_int, [C.c_int, _u32p, _u32p]),
        "orc_search_higher": (C.c_uint64, [_u64p, C.c_uint32, C.c_uint32, C.c_int, _u64p]),
        "orc_search_aggr": (C.c_uint64, [_u64p, C.c_uint32, C.c_uint32, C.c_uint64, _u64p]),
        "orc_search_leaf_rot": (C.c_uint64, [_u64p, C.c_uint32, _u64p]),
        "orc_search_leaf_brute": (C.c_uint64, [_u64p, C.c_uint32, _u64p]),
        "orc_remap64": (C.c_uint32, [C.c_uint64, C.c_uint32]),
        "orc_search_higher_r64": (C.c_uint64, [_u64p, C.c_uint32, C.c_uint32, C.c_int, _u64p]),
        "orc_search_aggr_r64": (C.c_uint64, [_u64p, C.c_uint32, C.c_uint32, C.c_uint64, _u64p]),
        "orc_search_leaf_rot_r64": (C.c_uint64, [_u64p, C.c_uint32, _u64p]),
        "orc_search_leaf_brute_r64": (C.c_uint64, [_u64p, C.c_uint32, _u64p]),
        "orc_packed_aggr_predicate": (C.c_int, [_u32p, C.c_uint32, C.c_uint32]),
        "orc_search_form_aggr_predicate": (C.c_int, [_u32p, C.c_uint32, C.c_uint32]),
        "orc_build": (C.c_void_p, [_u64p, C.c_size_t, C.c_int, C.c_size_t, C.c_int, C.c_int, C.POINTER(C.c_int)]),
        "orc_serialized_size": (C.c_size_t, [C.c_void_p]),
        "orc_serialize": (None, [C.c_void_p, C.c_char_p]),
        "orc_load": (C.c_void_p, [C.c_char_p, C.c_size_t, C.c_int, C.c_int]),
        "orc_query": (C.c_uint64, [C.c_void_p, C.c_uint64, C.c_uint64]),
        "orc_query_batch": (None, [C.c_void_p, _u64p, C.c_size_t, _u64p]),
        "orc_check_bijection": (C.c_uint64, [C.c_void_p, _u64p, C.c_size_t]),
        "orc_ef_roundtrip": (C.c_uint64, [_u64p, _u64p, C.c_uint64]),
        "orc_evals": (C.c_uint64, [C.c_void_p]),
        "orc_nseeds": (C.c_uint64, [C.c_void_p]),
        "orc_rice_bits": (C.c_uint64, [C.c_void_p]),
        "orc_nbuckets": (C.c_uint64, [C.c_void_p]),
        "orc_decode_seeds": (C.c_size_t, [C.c_void_p, _u64p, _u64p, C.c_size_t]),
        "orc_free": (None, [C.c_void_p]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()

START_SEED = [
    0x106393C187CAE21A, 0x6453CEC3F7376937, 0x643E521DDBD2BE98, 0x3740C6412F6572CB,
    0x717D47562F1CE470, 0x4CD6EB4C63BEFB7C, 0x9BFD8C5E18C8DA73, 0x082F20E10092A9A3,
    0x2ADA2CE68D21DEFC, 0xE33CB4F3E7C6466B, 0x3980BE458C509C59, 0xC466FD9584828E8C,
    0x45F0AABE1A61EDE6, 0xF6E7B8B33AD9B98D, 0x4EF95E25F4B4983D, 0x81175195173B92D3,
    0x4E50927D8DD15978, 0x1EA2099D1FAFAE7F, 0x425C8A06FBAAA815, 0xCD4216006C74052A,
]


def ptr64(a: np.ndarray):
    assert a.dtype == np.uint64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_u64p)


def ptr32(a: np.ndarray):
    assert a.dtype == np.uint32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_u32p)


def splitmix_keys(seed: int, n: int) -> np.ndarray:
    """(n, 2) uint64 array of (first, second) pairs."""
    out = np.empty((n, 2), dtype=np.uint64)
    lib.orc_splitmix64_keys(seed, n, ptr64(out))
    return out


def golomb_memo(leaf: int) -> np.ndarray:
    out = np.zeros(3000, dtype=np.uint32)
    assert lib.orc_golomb_memo(leaf, ptr32(out)) == 0
    return out


def aggr_bounds(leaf: int):
    lo, up = C.c_uint32(), C.c_uint32()
    assert lib.orc_aggr_bounds(leaf, C.byref(lo), C.byref(up)) == 0
    return lo.value, up.value


class OracleError(RuntimeError):
    pass


ERRORS = {-1: "bad arguments", -2: "bucket with >= 3000 keys", -3: "equal second inside a bucket",
          -4: "Elias-Fano jump offset overflow"}


# flags of the `rotate` argument: rotation fitting for full leaves, 64-bit remap
# (the RecSplitRotate variant)
FLAG_ROTATE, FLAG_REMAP64 = 1, 2


class OracleMPHF:
    """Oracle MPHF: build from keys, or load from serialized bytes.  `rotate` is a
    bool or a flags word (FLAG_ROTATE | FLAG_REMAP64)."""

    def __init__(self, handle, leaf: int, rotate: bool):
        self.h = handle
        self.leaf = leaf
        self.rotate = rotate

    @classmethod
    def build(cls, keys: np.ndarray, leaf: int, bucket_size: int, rotate: bool = True, threads: int = 1):
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        err = C.c_int(0)
        h = lib.orc_build(ptr64(keys), keys.shape[0], leaf, bucket_size, int(rotate), threads, C.byref(err))
        if not h:
            raise OracleError(ERRORS.get(err.value, f"error {err.value}"))
        return cls(h, leaf, rotate)

    @classmethod
    def load(cls, data: bytes, leaf: int, rotate: bool = True):
        h = lib.orc_load(data, len(data), leaf, int(rotate))
        if not h:
            raise OracleError("malformed serialized MPHF")
        return cls(h, leaf, rotate)

    def serialize(self) -> bytes:
        n = lib.orc_serialized_size(self.h)
        buf = C.create_string_buffer(n)
        lib.orc_serialize(self.h, buf)
        return buf.raw

    def query(self, keys: np.ndarray) -> np.ndarray:
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        out = np.empty(keys.shape[0], dtype=np.uint64)
        lib.orc_query_batch(self.h, ptr64(keys), keys.shape[0], ptr64(out))
        return out

    def bijection_errors(self, keys: np.ndarray) -> int:
        keys = np.ascontiguousarray(keys, dtype=np.uint64)
        return int(lib.orc_check_bijection(self.h, ptr64(keys), keys.shape[0]))

    def seeds(self) -> np.ndarray:
        cap = max(1, int(lib.orc_nseeds(self.h)) if lib.orc_nseeds(self.h) else 1 << 22)
        out = np.zeros(cap, dtype=np.uint64)
        k = lib.orc_decode_seeds(self.h, None, ptr64(out), cap)
        return out[:k]

    @property
    def evals(self) -> int:
        return int(lib.orc_evals(self.h))

    @property
    def nseeds(self) -> int:
        return int(lib.orc_nseeds(self.h))

    def __del__(self):
        if getattr(self, "h", None):
            lib.orc_free(self.h)
            self.h = None
