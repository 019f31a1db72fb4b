"""Dense matrix-vector product and the Jacobi diagonal of a dense matrix.

matvec(A, x) is the matrix-vector case of cuMat's dense ProductOp
(`A * x`, cuMat/src/ProductOp.h:318-420, which cuMat hands to cuBLAS gemm);
here it is one streaming HIP kernel (cm_dense_gemv). A batch-1 operand is
broadcast over the other's batches, as cuMat's batched products do."""
from __future__ import annotations

from . import _capi as capi
from .device import DeviceArray, ctx_handle


def matvec(A: DeviceArray, x: DeviceArray, out: DeviceArray | None = None, mode: int = capi.ASSIGN,
           ctx=None) -> DeviceArray:
    """out = A x (mode ASSIGN), out += A x (ADD) or out -= A x (SUB)."""
    if x.cols != 1 or A.cols != x.rows:
        raise capi.CumatArgError(capi.CM_E_DIM, f"matvec: {A.rows}x{A.cols} times {x.rows}x{x.cols}")
    if A.dtype != x.dtype:
        raise TypeError("matvec: operands of different scalar types (no implicit casting)")
    batches = max(A.batches, x.batches)
    if A.batches not in (1, batches) or x.batches not in (1, batches):
        raise capi.CumatArgError(capi.CM_E_DIM, "matvec: batch counts do not broadcast")
    if out is None:
        if mode != capi.ASSIGN:
            raise capi.CumatArgError(capi.CM_E_ARG, "matvec: += / -= need an output")
        out = DeviceArray(A.rows, 1, batches, dtype=A.np_dtype, ctx=A.ctx)
    if (out.rows, out.cols, out.batches) != (A.rows, 1, batches) or out.dtype != A.dtype:
        raise capi.CumatArgError(capi.CM_E_DIM, "matvec: output has the wrong shape or type")
    capi.call("cm_dense_gemv", ctx_handle(ctx), A.dtype, A.rows, A.cols, batches, A.ptr, int(A.batches > 1),
              int(A.row_major), x.ptr, int(x.batches > 1), out.ptr, int(mode))
    return out


def inverse_diagonal(A: DeviceArray, ctx=None) -> DeviceArray:
    """1 / diag(A), 1 where the diagonal is zero (cwiseInverseCheck,
    cuMat/src/UnaryOps.h:177); one column per batch of A."""
    if A.rows != A.cols:
        raise capi.CumatArgError(capi.CM_E_DIM, "inverse_diagonal: matrix must be square")
    ctx = A.ctx if ctx is None else ctx  # ordered with the other work on A
    out = DeviceArray(A.rows, 1, A.batches, dtype=A.np_dtype, ctx=ctx)
    capi.call("cm_dense_inverse_diagonal", ctx_handle(ctx), A.dtype, A.rows, A.batches, A.ptr,
              int(A.batches > 1), out.ptr)
    return out


def is_dense(m) -> bool:
    return isinstance(m, DeviceArray)

