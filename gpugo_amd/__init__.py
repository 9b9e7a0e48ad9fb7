"""gpugo_amd -- MI355X-native interval branch-and-bound global optimizer.

Host package over libgpugo.so (hand-written HIP for gfx950, C ABI in
include/gpugo.h).  Objectives are expressions (gpugo_amd.expr) lowered to
per-objective HIP code lists (gpugo_amd.emit) and compiled with hipRTC.
"""
from ._lib import GpugoError, init  # noqa: F401
from .expr import (E, PI, Const, IConst, Objective, P, X, cos, exp, fsum, pown, sin,  # noqa: F401
                   sqr, sqrt)
from .solver import Module, Problem, get_module, ia_eval, minimize  # noqa: F401

__all__ = ["GpugoError", "init", "Objective", "X", "P", "Const", "IConst", "PI", "E", "sqr", "sqrt",
           "exp", "sin", "cos", "pown", "fsum", "Module", "Problem", "get_module", "ia_eval", "minimize"]
