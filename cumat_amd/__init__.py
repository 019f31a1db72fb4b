"""cumat_amd -- MI355X (gfx950) implementation of cuMat's data-parallel hot
path: fused component-wise evaluation, batched row/column/batch reductions,
CSR SpMV and the Jacobi-preconditioned ConjugateGradient, behind the C ABI of
include/cumat_amd.h.

This Python package is a thin host layer over that ABI (the C++ host mirror
of cuMat's API lives in cumat_amd/cuMat/). Every operation runs the HIP
library; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi as capi
from ._capi import CumatArgError, CumatError  # noqa: F401
from .device import DeviceArray, Context, default_context, synchronize  # noqa: F401
from .expr import (Expr, as_expr, assign, cast, select, where_scalar,  # noqa: F401
                   lower, evaluate)
from .sparse import (SparseMatrix, SparsityPattern, ELLMatrix, BlockSparseMatrix, SparseExpressionView,  # noqa: F401
                     sparse_view)
from .reductions import reduce, dot, squared_norm, norm  # noqa: F401
from .solvers import ConjugateGradient, DiagonalPreconditioner, pcg_batched, pcg_ir  # noqa: F401
from . import dense  # noqa: F401

__all__ = [
    "capi", "DeviceArray", "Context", "default_context", "synchronize", "Expr", "as_expr",
    "assign", "cast", "select", "lower", "evaluate", "SparseMatrix", "SparsityPattern",
    "ELLMatrix", "SparseExpressionView", "sparse_view", "reduce", "dot", "squared_norm", "norm", "ConjugateGradient",
    "DiagonalPreconditioner", "pcg_batched", "pcg_ir", "dense", "counters", "reset_counters", "jit_mode",
    "device_info",
]


def counters() -> dict:
    """Profiling counters (cuMat::Profiling, cuMat/src/Profiling.h:23-109)."""
    names = ["DeviceMemAlloc", "DeviceMemFree", "EvalAny", "EvalCwise", "EvalCwiseSparse",
             "EvalReduction", "EvalMatmulSparse", "KernelLaunch", "EvalMatmul", "HostSync",
             "JitCompile", "JitLaunch"]
    out = {}
    for i, n in enumerate(names):
        v = C.c_longlong(0)
        capi.call("cm_counters_get", i, C.byref(v))
        out[n] = v.value
    return out


def reset_counters() -> None:
    capi.call("cm_counters_reset")


def jit_mode(mode: int = -1) -> int:
    """Run-time generated kernels for component-wise programs (cm_jit_mode):
    0 never, 1 from 2^18 elements on (default), 2 always; a negative value
    only queries. Returns the previous mode."""
    return int(capi.lib().cm_jit_mode(int(mode)))


def device_info(device: int = 0) -> dict:
    name = C.create_string_buffer(256)
    cu, hbm, clk, mclk, bw = C.c_int(), C.c_longlong(), C.c_int(), C.c_int(), C.c_int()
    capi.call("cm_device_info", device, name, 256, C.byref(cu), C.byref(hbm), C.byref(clk),
              C.byref(mclk), C.byref(bw))
    return {"name": name.value.decode(), "cu_count": cu.value, "hbm_bytes": hbm.value,
            "clock_khz": clk.value, "mem_clock_khz": mclk.value, "bus_width_bits": bw.value}
