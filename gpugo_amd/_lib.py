"""ctypes binding of libgpugo.so (include/gpugo.h).

There is no CPU fallback: if the library is missing or cannot be loaded,
every entry point raises GpugoError.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgpugo.so")

# status codes
OK = 0
E_INVALID, E_DOMAIN, E_NOMEM, E_COMPILE, E_HIP, E_NODEV, E_CAPACITY = -1, -2, -3, -4, -5, -6, -7
_ERR_NAMES = {E_INVALID: "INVALID", E_DOMAIN: "DOMAIN", E_NOMEM: "NOMEM", E_COMPILE: "COMPILE",
              E_HIP: "HIP", E_NODEV: "NODEV", E_CAPACITY: "CAPACITY"}

OPS = {"add": 0, "sub": 1, "mul": 2, "div": 3, "neg": 4, "scale": 5, "addc": 6, "rsubc": 7,
       "divc": 8, "sqr": 9, "pown": 10, "sqrt": 11, "exp": 12, "sin": 13, "cos": 14}
PREC = {"fp64": 0, "fp32": 1}
ROUND = {"exact": 0, "fast": 1, "digits15": 2}
FLAG_MEAN_VALUE = 1
FLAG_SPLIT = 4
FLAG_PERSIST = 8
FLAG_EAGER_COMPACT = 16
FLAG_NO_TABLES = 32
FLAG_NO_BAND = 64
STATUS = {0: "running", 1: "certified", 2: "x_converged", 3: "max_iters", 4: "exhausted", 5: "overflow"}

EXPORTS = [
    "gpugo_init", "gpugo_last_error", "gpugo_version", "gpugo_device_count", "gpugo_ia_eval",
    "gpugo_module_compile", "gpugo_rtc_compile", "gpugo_rtc_code_copy", "gpugo_module_load",
    "gpugo_module_code", "gpugo_module_destroy", "gpugo_eval_boxes", "gpugo_problem_create",
    "gpugo_step", "gpugo_solve", "gpugo_boxes_copy", "gpugo_problem_destroy", "gpugo_problem_reset",
    "gpugo_problem_timing", "gpugo_batch_bytes", "gpugo_mem_info", "gpugo_batch_create",
    "gpugo_batch_admit", "gpugo_batch_run", "gpugo_batch_sync", "gpugo_batch_results",
    "gpugo_batch_counts", "gpugo_batch_destroy", "gpugo_eval_boxes_f32", "gpugo_window_ranges",
    "gpugo_solve_static", "gpugo_batch_advance",
]


class GpugoError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"gpugo error {_ERR_NAMES.get(code, code)}: {msg}")
        self.code = code
        self.msg = msg

    def __reduce__(self):
        # picklable, so an error raised in a worker process reaches its parent
        return (GpugoError, (self.code, self.msg))


class ProblemDesc(ctypes.Structure):
    _fields_ = [
        ("d", ctypes.c_int32), ("wid", ctypes.c_int32), ("cycle", ctypes.c_int32),
        ("k_select", ctypes.c_int32), ("n_samples", ctypes.c_int32), ("precision", ctypes.c_int32),
        ("rounding", ctypes.c_int32), ("flags", ctypes.c_int32),
        ("lo", ctypes.POINTER(ctypes.c_double)), ("hi", ctypes.POINTER(ctypes.c_double)),
        ("params", ctypes.POINTER(ctypes.c_double)), ("n_params", ctypes.c_int64),
        ("tol_f", ctypes.c_double), ("tol_x", ctypes.c_double), ("init_ub", ctypes.c_double),
        ("max_queue", ctypes.c_int64), ("max_parents", ctypes.c_int64), ("max_iters", ctypes.c_int64),
    ]


class BatchDesc(ctypes.Structure):
    _fields_ = [
        ("slots", ctypes.c_int32), ("steps_per_launch", ctypes.c_int32), ("rounding", ctypes.c_int32),
        ("use_graph", ctypes.c_int32), ("max_queue", ctypes.c_int64), ("max_parents", ctypes.c_int64),
        ("bound_blocks", ctypes.c_int32), ("grow", ctypes.c_int32),
        ("promote_age", ctypes.c_int32), ("pad0", ctypes.c_int32),
    ]


class Result(ctypes.Structure):
    _fields_ = [
        ("lb", ctypes.c_double), ("ub", ctypes.c_double), ("width", ctypes.c_double),
        ("status", ctypes.c_int32), ("overflow", ctypes.c_int32),
        ("iterations", ctypes.c_int64), ("boxes_bounded", ctypes.c_int64),
        ("pruned_value", ctypes.c_int64), ("pruned_kkt", ctypes.c_int64),
        ("queue_size", ctypes.c_int64), ("queue_peak", ctypes.c_int64),
        ("n_out", ctypes.c_int64), ("ms", ctypes.c_double),
        ("bound_ms", ctypes.c_double), ("bound_launches", ctypes.c_int64),
    ]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if not k.startswith("reserved")}
        d["status_name"] = STATUS.get(self.status, str(self.status))
        return d


_lib = None


def load():
    """Load libgpugo.so (raises GpugoError if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise GpugoError(E_NODEV, f"{LIB_PATH} is not built; run gpugo_amd.build.build()")
    lib = ctypes.CDLL(LIB_PATH)
    D = ctypes.POINTER(ctypes.c_double)
    I32 = ctypes.POINTER(ctypes.c_int32)
    vp = ctypes.c_void_p
    lib.gpugo_last_error.restype = ctypes.c_char_p
    lib.gpugo_ia_eval.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int64, D, D, D, D,
                                  ctypes.c_double, D, D]
    lib.gpugo_module_compile.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
                                         ctypes.POINTER(vp)]
    lib.gpugo_rtc_compile.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p,
                                      ctypes.POINTER(ctypes.c_size_t)]
    lib.gpugo_rtc_code_copy.argtypes = [ctypes.c_char_p, ctypes.c_size_t]
    lib.gpugo_module_load.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(vp)]
    lib.gpugo_module_code.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_size_t)]
    lib.gpugo_module_destroy.argtypes = [vp]
    lib.gpugo_module_destroy.restype = None
    lib.gpugo_eval_boxes.argtypes = [vp, ctypes.c_int32, D, ctypes.c_int64, ctypes.c_int64, D, D, D, D,
                                     ctypes.c_int32, ctypes.c_int32, D, D, D, D, I32, D]
    lib.gpugo_problem_create.argtypes = [vp, ctypes.POINTER(ProblemDesc), ctypes.POINTER(vp)]
    lib.gpugo_step.argtypes = [vp, ctypes.c_int64, ctypes.POINTER(Result)]
    lib.gpugo_solve.argtypes = [vp, ctypes.POINTER(Result)]
    lib.gpugo_boxes_copy.argtypes = [vp, ctypes.c_int64, D, D, D, ctypes.POINTER(ctypes.c_int64)]
    lib.gpugo_problem_reset.argtypes = [vp]
    lib.gpugo_problem_timing.argtypes = [vp, ctypes.c_int]
    lib.gpugo_problem_destroy.argtypes = [vp]
    lib.gpugo_problem_destroy.restype = None
    I64 = ctypes.POINTER(ctypes.c_int64)
    lib.gpugo_eval_boxes_f32.argtypes = [vp, ctypes.c_int32, D, ctypes.c_int64, ctypes.c_int64, D, D,
                                         ctypes.c_int32, D, D]
    lib.gpugo_batch_bytes.argtypes = [vp, ctypes.POINTER(BatchDesc), I64]
    lib.gpugo_mem_info.argtypes = [I64, I64]
    lib.gpugo_batch_create.argtypes = [vp, ctypes.POINTER(BatchDesc), ctypes.POINTER(vp)]
    lib.gpugo_batch_admit.argtypes = [vp, ctypes.c_int32, ctypes.POINTER(ProblemDesc), ctypes.c_int64, I32]
    lib.gpugo_batch_run.argtypes = [vp, ctypes.c_int64]
    lib.gpugo_batch_advance.argtypes = [vp, I64]
    lib.gpugo_batch_sync.argtypes = [vp]
    lib.gpugo_batch_results.argtypes = [vp, ctypes.c_int64, I64, ctypes.POINTER(Result), I64]
    lib.gpugo_batch_counts.argtypes = [vp, I64, I64, I64]
    lib.gpugo_batch_destroy.argtypes = [vp]
    lib.gpugo_batch_destroy.restype = None
    lib.gpugo_window_ranges.argtypes = [ctypes.c_int64, I32, I32]
    lib.gpugo_solve_static.argtypes = [vp, ctypes.POINTER(ProblemDesc), vp, ctypes.POINTER(Result)]
    _lib = lib
    return lib


def check(rc: int):
    if rc != OK:
        raise GpugoError(rc, load().gpugo_last_error().decode(errors="replace"))


def ptr(a, ctype=ctypes.c_double):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ctype))


_initialized = False


def init(device: int = 0):
    """Select the GPU (idempotent)."""
    global _initialized
    if not _initialized:
        check(load().gpugo_init(device))
        _initialized = True


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)
