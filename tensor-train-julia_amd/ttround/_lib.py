"""ctypes binding of libttround.so (the C-ABI declared in include/ttround.h).

The shared library is built in-tree (``make -C tensor-train-julia_amd``) and
loaded from this directory; there is no fallback implementation, so a missing
or unloadable library raises immediately.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libttround.so")

TTR_OK, TTR_INVALID_ARG, TTR_DIM_MISMATCH, TTR_OOM, TTR_NUMERICAL_FAILURE, TTR_DEVICE_ERROR, TTR_NOT_FOUND = range(7)
PREC = {"f64": 0, "fp64": 0, "float64": 0, "f32": 1, "fp32": 1, "float32": 1, "bf16": 2, "bfloat16": 2}
DIRECTION = {"left": 0, "right": 1}


class DimensionMismatch(ValueError):
    """Counterpart of Julia's ``DimensionMismatch`` / dimension ``AssertionError``."""


class NumericalFailure(ArithmeticError):
    """Counterpart of Julia's ``LAPACKException``."""


class DeviceError(RuntimeError):
    pass


_EXC = {
    TTR_INVALID_ARG: ValueError,
    TTR_DIM_MISMATCH: DimensionMismatch,
    TTR_OOM: MemoryError,
    TTR_NUMERICAL_FAILURE: NumericalFailure,
    TTR_DEVICE_ERROR: DeviceError,
    TTR_NOT_FOUND: KeyError,
}

i64 = C.c_int64
i64p = C.POINTER(C.c_int64)
f64p = C.POINTER(C.c_double)
f64pp = C.POINTER(C.POINTER(C.c_double))


class RandItem(C.Structure):
    _fields_ = [("nterms", i64), ("alpha", f64p), ("term_ids", i64p), ("op_id", i64), ("rks", i64p)]


class RandOpts(C.Structure):
    _fields_ = [("orthogonal", C.c_int32), ("sketch_prec", C.c_int32), ("block_rks", i64), ("seed", C.c_uint64),
                ("p", i64p), ("oversampling", C.c_double), ("forward", C.c_int32), ("blocks", f64pp),
                ("blocks_right", f64pp), ("sketch_accum", C.c_int32)]


# name -> (restype, argtypes); every symbol declared in include/ttround.h
SIGNATURES = {
    "ttr_version": (C.c_char_p, []),
    "ttr_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int]),
    "ttr_destroy": (C.c_int, [C.c_void_p]),
    "ttr_last_error": (C.c_char_p, [C.c_void_p]),
    "ttr_tt_upload": (C.c_int, [C.c_void_p, i64, i64p, i64p, i64p, f64pp, i64p]),
    "ttr_tt_info": (C.c_int, [C.c_void_p, i64, i64p, i64p, i64p, i64p]),
    "ttr_tt_download_core": (C.c_int, [C.c_void_p, i64, i64, f64p]),
    "ttr_tt_free": (C.c_int, [C.c_void_p, i64]),
    "ttr_tt_randn": (C.c_int, [C.c_void_p, i64, i64p, i64p, C.c_uint64, i64p]),
    "ttr_tto_upload": (C.c_int, [C.c_void_p, i64, i64p, i64p, f64pp, i64p]),
    "ttr_lincomb": (C.c_int, [C.c_void_p, i64, f64p, i64p, i64p]),
    "ttr_mpo_apply": (C.c_int, [C.c_void_p, i64, i64, i64p]),
    "ttr_hadamard": (C.c_int, [C.c_void_p, i64, i64p, C.c_int32, i64p]),
    "ttr_orthogonalize": (C.c_int, [C.c_void_p, i64, i64p, i64, i64p]),
    "ttr_dot_noalloc": (C.c_int, [C.c_void_p, i64, i64p, i64p, f64p]),
    "ttr_stta_sketch_tt": (C.c_int, [C.c_void_p, i64, i64, i64, f64pp, f64pp]),
    "ttr_round_svd": (C.c_int, [C.c_void_p, i64, i64p, C.c_double, i64, C.c_int, i64p]),
    "ttr_norm": (C.c_int, [C.c_void_p, i64, i64p, f64p]),
    "ttr_dot": (C.c_int, [C.c_void_p, i64, i64p, i64p, f64p]),
    "ttr_partial_contraction": (C.c_int, [C.c_void_p, i64, i64, C.c_int, f64pp]),
    "ttr_dot_operator": (C.c_int, [C.c_void_p, i64, i64p, i64p, i64p, f64p]),
    "ttr_round_rand": (C.c_int, [C.c_void_p, i64, C.POINTER(RandItem), C.POINTER(RandOpts), i64p]),
    "ttr_sketch": (C.c_int, [C.c_void_p, i64, i64p, C.POINTER(RandOpts), f64pp, i64p]),
    "ttr_sketch_op": (C.c_int, [C.c_void_p, i64, i64, i64p, C.POINTER(RandOpts), f64pp, i64p]),
    "ttr_tt_sketch": (C.c_int, [C.c_void_p, i64, i64p, i64p, i64, C.POINTER(RandOpts), f64p, i64, i64p]),
    "ttr_sketch_blocks": (C.c_int, [C.c_void_p, i64, i64p, i64p, C.POINTER(RandOpts), i64, f64p, i64p]),
    "ttr_stta": (C.c_int, [C.c_void_p, i64, C.POINTER(RandItem), C.POINTER(RandOpts), C.c_uint64, i64p]),
    "ttr_stta_sketch": (C.c_int, [C.c_void_p, i64, i64p, C.POINTER(RandOpts), C.c_uint64, f64pp, f64pp, i64p, i64p]),
    "ttr_sketch_draws": (C.c_int, [C.c_void_p, i64, i64, C.c_uint64, f64p]),
    "ttr_profile_enable": (C.c_int, [C.c_void_p, C.c_int]),
    "ttr_graph_enable": (C.c_int, [C.c_void_p, C.c_int]),
    "ttr_round_rand_tuple": (C.c_int, [C.c_void_p, i64, i64p, i64p, i64p, C.POINTER(RandOpts), i64p]),
    "ttr_sketch_tuple": (C.c_int, [C.c_void_p, i64, i64p, i64p, C.POINTER(RandOpts), C.POINTER(f64p), i64p]),
    "ttr_graph_stats": (C.c_int, [C.c_void_p, i64p, i64p, i64p, i64p]),
    "ttr_profile_reset": (C.c_int, [C.c_void_p]),
    "ttr_trim": (C.c_int, [C.c_void_p]),
    "ttr_profile_count": (C.c_int, [C.c_void_p, i64p]),
    "ttr_profile_get": (C.c_int, [C.c_void_p, i64, C.c_char_p, i64, f64p, i64p, f64p, f64p]),
}

_lib = None
_lock = threading.RLock()


def load() -> C.CDLL:
    """Loads libttround.so (raises OSError when it has not been built)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise OSError(f"{LIB_PATH} not found: build it with `make -C tensor-train-julia_amd`")
            lib = C.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _lib = lib
    return _lib


def as_i64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.int64))


def ptr_i64(a: np.ndarray):
    return a.ctypes.data_as(i64p)


def ptr_f64(a: np.ndarray):
    return a.ctypes.data_as(f64p)


class Handle:
    """One libttround handle (one HIP stream and workspace) on one device."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = C.c_void_p()
        st = self.lib.ttr_create(C.byref(h), int(device))
        if st != TTR_OK:
            raise _EXC.get(st, DeviceError)(f"ttr_create failed with status {st}")
        self.h = h
        self.device = device

    def check(self, st: int):
        if st != TTR_OK:
            msg = self.lib.ttr_last_error(self.h).decode(errors="replace")
            raise _EXC.get(st, DeviceError)(msg)

    def profile(self, on: bool = True, level: int = 1):
        """hipEvent timing on the handle's stream: level 1 every timed launch, level 2
        only the GEMM classes and the phases."""
        self.check(self.lib.ttr_profile_enable(self.h, level if on else 0))

    def graphs(self, on: bool = True):
        """ttr_round_rand through cached HIP graphs (off frees them)."""
        self.check(self.lib.ttr_graph_enable(self.h, 1 if on else 0))

    def graph_stats(self) -> dict:
        v = [C.c_int64() for _ in range(4)]
        self.check(self.lib.ttr_graph_stats(self.h, *[C.byref(x) for x in v]))
        return dict(zip(("captures", "replays", "fallbacks", "cached"), (x.value for x in v)))

    def trim(self):
        """Return the idle workspace arena and the freed TT storage to the runtime."""
        self.check(self.lib.ttr_trim(self.h))

    def profile_reset(self):
        self.check(self.lib.ttr_profile_reset(self.h))

    def profile_stats(self) -> dict:
        """{name: {"ms", "launches", "flops", "bytes"}} accumulated since the last reset."""
        n = C.c_int64()
        self.check(self.lib.ttr_profile_count(self.h, C.byref(n)))
        out = {}
        for i in range(n.value):
            name = C.create_string_buffer(128)
            ms, fl, by = C.c_double(), C.c_double(), C.c_double()
            la = C.c_int64()
            self.check(self.lib.ttr_profile_get(self.h, i, name, 128, C.byref(ms), C.byref(la), C.byref(fl), C.byref(by)))
            out[name.value.decode()] = {"ms": ms.value, "launches": la.value, "flops": fl.value, "bytes": by.value}
        return out

    def close(self):
        if getattr(self, "h", None):
            self.lib.ttr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default = None


def default_handle() -> Handle:
    global _default
    with _lock:
        if _default is None:
            _default = Handle(int(os.environ.get("TTROUND_DEVICE", "0")))
    return _default
