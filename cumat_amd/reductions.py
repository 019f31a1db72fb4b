"""Reductions over any subset of {Row, Column, Batch} and dot / squaredNorm
(cuMat/src/ReductionOpsPlugin.inl:124-450, cuMat/src/ReductionOps.h:167-598).

A stored array is reduced by cm_reduce / cm_dot. An unevaluated expression is
consumed lazily, as cuMat's input iterator consumes its expression tree
(cuMat/src/Iterator.h:99-105): its program is evaluated inside the reduction
(cm_reduce_expr), one launch and no temporary; dot and squared_norm of
expressions reduce cwiseDot / cwiseAbs2 the same way
(cuMat/src/ReductionOpsPlugin.inl:290-297)."""
from __future__ import annotations

import numpy as np

from . import _capi as capi
from .device import DeviceArray, ctx_handle, dtype_to_np

_OPS = {"sum": capi.RED_SUM, "prod": capi.RED_PROD, "min": capi.RED_MIN, "max": capi.RED_MAX,
        "all": capi.RED_ALL, "any": capi.RED_ANY, "bitwise_and": capi.RED_BAND,
        "bitwise_or": capi.RED_BOR}


def _is_stored(a) -> bool:
    """A DeviceArray, or a leaf expression that is exactly a whole array in
    its own layout."""
    from .expr import LeafNode
    if isinstance(a, DeviceArray):
        return True
    return (isinstance(a, LeafNode) and a.kind == capi.LEAF_STRIDED and a.strides == a.arr.strides()
            and a.shape == (a.arr.rows, a.arr.cols, a.arr.batches))


def _reduce_expr(e, opc, axis, out, out_row_major, ctx):
    from .expr import lower
    R, Cc, B = e.shape
    if e.dtype not in (capi.F32, capi.F64, capi.I32, capi.BOOL):
        raise TypeError("reductions of expressions take float32, float64, int32 or bool expressions")
    oR = 1 if axis & capi.AXIS_ROW else R
    oC = 1 if axis & capi.AXIS_COL else Cc
    oB = 1 if axis & capi.AXIS_BATCH else B
    if out is None:
        out = DeviceArray(oR, oC, oB, dtype=dtype_to_np(e.dtype), row_major=bool(out_row_major), ctx=ctx)
    elif (out.rows, out.cols, out.batches) != (oR, oC, oB) or out.dtype != e.dtype:
        raise capi.CumatArgError(capi.CM_E_DIM, "reduction output has the wrong size or dtype")
    leaves, nl, code, nc = lower(e, (R, Cc, B))
    capi.call("cm_reduce_expr", ctx_handle(ctx), leaves, nl, code, nc, e.dtype, opc, axis, R, Cc, B,
              int(out.row_major), out.ptr)
    return out


def reduce(a, op="sum", axis: int = capi.AXIS_ALL, out: DeviceArray | None = None,
           out_row_major: bool | None = None, ctx=None) -> DeviceArray:
    """out = reduce_op over the axes in 'axis' (bitmask Row=1|Column=2|Batch=4).
    The result keeps the other dimensions (size 1 on reduced axes). 'a' is a
    DeviceArray or an expression (evaluated inside the reduction)."""
    opc = _OPS[op] if isinstance(op, str) else int(op)
    if not _is_stored(a):
        return _reduce_expr(a, opc, axis, out, bool(out_row_major), ctx)
    a = a if isinstance(a, DeviceArray) else a.arr
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


def dot(a, b, out: DeviceArray | None = None, ctx=None) -> DeviceArray:
    """Per-batch dot product; a batch-1 operand is broadcast. Expressions are
    reduced as cwiseDot(a, b).sum<Row | Column>() in one launch."""
    if not (_is_stored(a) and _is_stored(b)):
        from .expr import as_expr
        ea, eb = as_expr(a), as_expr(b)
        if ea.shape[:2] != eb.shape[:2]:
            raise capi.CumatArgError(capi.CM_E_DIM, "dot: size mismatch")
        return _reduce_expr(ea.cwiseDot(eb), capi.RED_SUM, capi.AXIS_ROW | capi.AXIS_COL, out, False, ctx)
    a = a if isinstance(a, DeviceArray) else a.arr
    b = b if isinstance(b, DeviceArray) else b.arr
    n = a.rows * a.cols
    if b.rows * b.cols != n:
        raise capi.CumatArgError(capi.CM_E_DIM, "dot: size mismatch")
    batches = max(a.batches, b.batches)
    if out is None:
        out = DeviceArray(1, 1, batches, dtype=a.np_dtype, ctx=a.ctx)
    capi.call("cm_dot", ctx_handle(ctx), a.dtype, n, batches, a.ptr, int(a.batches > 1), b.ptr,
              int(b.batches > 1), out.ptr)
    return out


def squared_norm(a, out: DeviceArray | None = None, ctx=None) -> DeviceArray:
    """Per-batch squared norm; an expression is reduced as cwiseAbs2().sum()."""
    if not _is_stored(a):
        return _reduce_expr(a.cwiseAbs2(), capi.RED_SUM, capi.AXIS_ROW | capi.AXIS_COL, out, False, ctx)
    a = a if isinstance(a, DeviceArray) else a.arr
    if out is None:
        out = DeviceArray(1, 1, a.batches, dtype=a.np_dtype, ctx=a.ctx)
    capi.call("cm_squared_norm", ctx_handle(ctx), a.dtype, a.rows * a.cols, a.batches, a.ptr, out.ptr)
    return out


def norm(a, ctx=None) -> np.ndarray:
    return np.sqrt(squared_norm(a, ctx=ctx).numpy())
