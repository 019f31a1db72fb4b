"""Sparse matrices: CSR (SparsityPattern<CSR> + SparseMatrix<Scalar,Batches,CSR>,
cuMat/src/SparseMatrixBase.h:89-147, cuMat/src/SparseMatrix.h:48-414) and
ELLPACK (cuMat/src/SparseMatrixBase.h:149-206)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi as capi
from .device import DeviceArray, ctx_handle, np_to_dtype
from .expr import Expr, as_expr, lower


class SparsityPattern:
    """CSR pattern with cuMat's names: JA = row pointer (rows+1), IA = column
    indices (nnz); both int32, shared by every batch of values."""

    def __init__(self, rows, cols, JA, IA, ctx=None):
        JA = np.ascontiguousarray(JA, np.int32)
        IA = np.ascontiguousarray(IA, np.int32)
        self.rows, self.cols, self.nnz = int(rows), int(cols), int(IA.size)
        self.ctx = ctx
        self.assert_valid(JA, IA)
        self.JA = DeviceArray.from_numpy(JA, ctx=ctx)
        self.IA = DeviceArray.from_numpy(IA if IA.size else np.zeros(1, np.int32), ctx=ctx)
        self.host_JA, self.host_IA = JA, IA

    def assert_valid(self, JA, IA):
        """SparsityPattern<CSR>::assertValid (cuMat/src/SparseMatrixBase.h:113-125)."""
        if self.rows < 0 or self.cols < 0:
            raise capi.CumatArgError(capi.CM_E_DIM, "negative dimension")
        if JA.size != self.rows + 1:
            raise capi.CumatArgError(capi.CM_E_DIM, "JA must have rows+1 entries")
        if JA[0] != 0 or JA[-1] != self.nnz or np.any(np.diff(JA) < 0):
            raise capi.CumatArgError(capi.CM_E_DIM, "JA must be non-decreasing from 0 to nnz")
        if IA.size and (IA.min() < 0 or IA.max() >= self.cols):
            raise capi.CumatArgError(capi.CM_E_DIM, "column index out of range")


class SparseMatrix:
    """Batched CSR matrix; values are nnz x 1 x batches (batch stride nnz)."""

    def __init__(self, pattern: SparsityPattern, values=None, batches: int = 1, dtype=np.float32):
        self.pattern = pattern
        self.batches = int(batches)
        if values is None:
            self.values = DeviceArray(max(pattern.nnz, 1), 1, self.batches, dtype=dtype, ctx=pattern.ctx)
        elif isinstance(values, DeviceArray):
            self.values = values
        else:
            v = np.asarray(values)
            self.values = DeviceArray.from_numpy(v.reshape(-1), rows=max(pattern.nnz, 1), cols=1,
                                                 batches=self.batches, ctx=pattern.ctx)
        self.dtype = self.values.dtype

    @classmethod
    def from_csr(cls, rows, cols, JA, IA, vals, batches=1, ctx=None):
        pat = SparsityPattern(rows, cols, JA, IA, ctx=ctx)
        return cls(pat, np.asarray(vals), batches=batches)

    @property
    def rows(self):
        return self.pattern.rows

    @property
    def cols(self):
        return self.pattern.cols

    @property
    def nnz(self):
        return self.pattern.nnz

    def matvec(self, x: DeviceArray, out: DeviceArray | None = None, mode: int = capi.ASSIGN,
               ctx=None) -> DeviceArray:
        """y (mode)= A * x  (ProductAssignment<...SparseMatrix<CSR>...>,
        cuMat/src/SparseProductEvaluation.h:101-139)."""
        p = self.pattern
        if x.rows != p.cols or x.cols != 1:
            raise capi.CumatArgError(capi.CM_E_DIM, f"SpMV: x is {x.rows}x{x.cols}, need {p.cols}x1")
        batches = max(self.batches, x.batches)
        if self.batches not in (1, batches) or x.batches not in (1, batches):
            raise capi.CumatArgError(capi.CM_E_DIM, "SpMV: batch counts do not broadcast")
        vec_dt = x.dtype
        if out is None:
            out = DeviceArray(p.rows, 1, batches, dtype=x.np_dtype, ctx=x.ctx)
        if out.rows != p.rows or out.batches != batches:
            raise capi.CumatArgError(capi.CM_E_DIM, "SpMV: output has the wrong size")
        capi.call("cm_csr_spmv", ctx_handle(ctx), self.dtype, vec_dt, p.rows, p.cols, p.nnz, batches,
                  p.JA.ptr, p.IA.ptr, self.values.ptr, int(self.batches > 1), x.ptr,
                  int(x.batches > 1), out.ptr, mode)
        return out

    __matmul__ = matmul = matvec

    def inverse_diagonal(self, ctx=None) -> DeviceArray:
        """Runs on the matrix's context unless ctx says otherwise, so it is
        ordered with the other work on that matrix."""
        p = self.pattern
        ctx = p.ctx if ctx is None else ctx
        out = DeviceArray(p.rows, 1, self.batches, dtype=self.values.np_dtype, ctx=ctx)
        capi.call("cm_csr_inverse_diagonal", ctx_handle(ctx), self.dtype, p.rows, p.nnz, self.batches,
                  p.JA.ptr, p.IA.ptr, self.values.ptr, int(self.batches > 1), out.ptr)
        return out

    def direct(self):
        """Expression leaf over the stored values, evaluated at each non-zero
        (SparseMatrix::direct(), cuMat/src/SparseMatrix.h:230-233)."""
        from .expr import LeafNode
        return LeafNode(self.values, strides=(-1, -1, self.pattern.nnz),
                        shape=(self.rows, self.cols, self.batches))

    def assign(self, e, mode: int = capi.ASSIGN, ctx=None):
        """Evaluate a cwise expression only at the stored entries
        (CwiseCSREvaluationKernel, cuMat/src/SparseEvaluation.h:14-34,83-138)."""
        e = as_expr(e)
        if e.dtype != self.dtype:
            raise TypeError("no implicit casting into a sparse matrix")
        p = self.pattern
        leaves, nl, code, nc = lower(e, (p.rows, p.cols, self.batches))
        capi.call("cm_csr_cwise_eval", ctx_handle(ctx), leaves, nl, code, nc, p.rows, p.cols, p.nnz,
                  self.batches, p.JA.ptr, p.IA.ptr, self.values.ptr, self.dtype, mode)
        return self

    def to_dense(self):
        """Host dense copy (batches, rows, cols) for checks."""
        v = self.values.numpy().reshape(self.batches, -1)
        out = np.zeros((self.batches, self.rows, self.cols), v.dtype)
        JA, IA = self.pattern.host_JA, self.pattern.host_IA
        rows = np.repeat(np.arange(self.rows), np.diff(JA))
        for b in range(self.batches):
            out[b, rows, IA] = v[b, :self.nnz]
        return out


class ELLMatrix:
    """ELLPACK matrix: column-major rows x nnz_per_row index matrix, -1 padding."""

    def __init__(self, rows, cols, indices_rowsxk, values_rowsxk, ctx=None):
        ind = np.asarray(indices_rowsxk, np.int32)
        self.rows, self.cols = int(rows), int(cols)
        self.npr = ind.shape[1]
        if ind.shape[0] != rows or self.npr > cols:
            raise capi.CumatArgError(capi.CM_E_DIM, "ELLPACK pattern has the wrong size")
        # the kernel gathers x at every index >= 0: reject indices outside
        # [-1, cols) here rather than let one reach the device
        if ind.size and (int(ind.min()) < -1 or int(ind.max()) >= self.cols):
            raise capi.CumatArgError(capi.CM_E_ARG, "ELLPACK index outside [-1, cols)")
        self.indices = DeviceArray.from_numpy(np.asfortranarray(ind).reshape(-1, order="F"), ctx=ctx)
        v = np.asarray(values_rowsxk)
        self.values = DeviceArray.from_numpy(np.asfortranarray(v).reshape(-1, order="F"), ctx=ctx)
        self.dtype = self.values.dtype

    def matvec(self, x: DeviceArray, out: DeviceArray | None = None, mode=capi.ASSIGN, ctx=None):
        if out is None:
            out = DeviceArray(self.rows, 1, x.batches, dtype=x.np_dtype, ctx=x.ctx)
        capi.call("cm_ell_spmv", ctx_handle(ctx), self.dtype, self.rows, self.cols, self.npr,
                  x.batches, self.indices.ptr, self.values.ptr, 0, x.ptr, int(x.batches > 1),
                  out.ptr, mode)
        return out
