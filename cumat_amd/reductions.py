"""Reductions over any subset of {Row, Column, Batch} and dot / squaredNorm
(cuMat/src/ReductionOpsPlugin.inl:124-450, cuMat/src/ReductionOps.h:167-598)."""
from __future__ import annotations

import math

import numpy as np

from . import _capi as capi
from .device import DeviceArray, ctx_handle

_OPS = {"sum": capi.RED_SUM, "prod": capi.RED_PROD, "min": capi.RED_MIN, "max": capi.RED_MAX,
        "all": capi.RED_ALL, "any": capi.RED_ANY, "bitwise_and": capi.RED_BAND,
        "bitwise_or": capi.RED_BOR}


def reduce(a: DeviceArray, op="sum", axis: int = capi.AXIS_ALL, out: DeviceArray | None = None,
           out_row_major: bool | None = None, ctx=None) -> DeviceArray:
    """out = reduce_op over the axes in 'axis' (bitmask Row=1|Column=2|Batch=4).
    The result keeps the other dimensions (size 1 on reduced axes)."""
    opc = _OPS[op] if isinstance(op, str) else int(op)
    oR = 1 if axis & capi.AXIS_ROW else a.rows
    oC = 1 if axis & capi.AXIS_COL else a.cols
    oB = 1 if axis & capi.AXIS_BATCH else a.batches
    orm = a.row_major if out_row_major is None else out_row_major
    if out is None:
        out = DeviceArray(oR, oC, oB, dtype=a.np_dtype, row_major=orm, ctx=a.ctx)
    elif (out.rows, out.cols, out.batches) != (oR, oC, oB):
        raise capi.CumatArgError(capi.CM_E_DIM, "reduction output has the wrong size")
    capi.call("cm_reduce", ctx_handle(ctx), a.dtype, opc, axis, a.rows, a.cols, a.batches,
              int(a.row_major), a.ptr, int(out.row_major), out.ptr)
    return out


def dot(a: DeviceArray, b: DeviceArray, out: DeviceArray | None = None, ctx=None) -> DeviceArray:
    """Per-batch dot product; a batch-1 operand is broadcast."""
    n = a.rows * a.cols
    if b.rows * b.cols != n:
        raise capi.CumatArgError(capi.CM_E_DIM, "dot: size mismatch")
    batches = max(a.batches, b.batches)
    if out is None:
        out = DeviceArray(1, 1, batches, dtype=a.np_dtype, ctx=a.ctx)
    capi.call("cm_dot", ctx_handle(ctx), a.dtype, n, batches, a.ptr, int(a.batches > 1), b.ptr,
              int(b.batches > 1), out.ptr)
    return out


def squared_norm(a: DeviceArray, out: DeviceArray | None = None, ctx=None) -> DeviceArray:
    if out is None:
        out = DeviceArray(1, 1, a.batches, dtype=a.np_dtype, ctx=a.ctx)
    capi.call("cm_squared_norm", ctx_handle(ctx), a.dtype, a.rows * a.cols, a.batches, a.ptr, out.ptr)
    return out


def norm(a: DeviceArray, ctx=None) -> np.ndarray:
    return np.sqrt(squared_norm(a, ctx=ctx).numpy())
