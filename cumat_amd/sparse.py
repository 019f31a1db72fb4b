"""Sparse matrices: CSR (SparsityPattern<CSR> + SparseMatrix<Scalar,Batches,CSR>,
cuMat/src/SparseMatrixBase.h:89-147, cuMat/src/SparseMatrix.h:48-414),
ELLPACK (cuMat/src/SparseMatrixBase.h:149-206) and block-CSR over block
scalars (SparseMatrix<real3x3>, tests/TestBlockedConjugateGradient.cu:12-239)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi as capi
from .device import DeviceArray, ctx_handle, np_to_dtype
from .expr import Expr, LeafNode, as_expr, lower


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
        return SparseValueLeaf(self, self.pattern.nnz)

    # A sparse matrix used in an expression reads its stored values; cuMat's
    # coeff() search (cuMat/src/SparseMatrix.h:135-142) gives the same values
    # wherever the view's pattern is the matrix's own, the only pattern the
    # fused product accepts.
    def _leaf(self):
        return self.direct()

    def __mul__(self, o):
        if isinstance(o, DeviceArray):
            return self.matvec(o)
        return self._leaf() * o

    def __rmul__(self, o): return o * self._leaf()
    def __add__(self, o): return self._leaf() + _sparse_operand(o)
    def __radd__(self, o): return _sparse_operand(o) + self._leaf()
    def __sub__(self, o): return self._leaf() - _sparse_operand(o)
    def __rsub__(self, o): return _sparse_operand(o) - self._leaf()
    def __neg__(self): return -self._leaf()

    def sparse_view(self, pattern=None):
        """The matrix as a lazily evaluated sparse operand (its own pattern by
        default)."""
        return SparseExpressionView(self._leaf(), self.pattern if pattern is None else pattern)

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


class SparseValueLeaf(LeafNode):
    """Leaf over the stored values of a CSR or ELLPACK matrix, read at each
    stored entry of the product's own pattern (direct access). Lowered as a
    leaf with row stride -1 and batch stride = stored entries per batch."""

    def __init__(self, matrix, per_batch: int):
        super().__init__(matrix.values, strides=(-1, -1, per_batch),
                         shape=(matrix.rows, matrix.cols, matrix.batches))
        self.matrix = matrix


def _sparse_operand(o):
    return o._leaf() if isinstance(o, (SparseMatrix, ELLMatrix)) else o


def _value_leaves(e):
    """Every sparse value leaf of an expression tree."""
    out = []
    stack = [e]
    while stack:
        n = stack.pop()
        if isinstance(n, SparseValueLeaf):
            out.append(n)
        for k in ("a", "b", "cond"):
            c = getattr(n, k, None)
            if isinstance(c, Expr):
                stack.append(c)
    return out


class SparseExpressionView:
    """expr.sparseView<CSR|ELLPACK>(pattern): a component-wise expression
    read only at the stored entries of 'pattern' (a SparsityPattern for CSR,
    an ELLMatrix for ELLPACK), used as the matrix of a product that
    evaluates the expression inside its value loads -- one launch, no value
    array written (SparseExpressionOp, cuMat/src/SparseExpressionOp.h:32-72;
    ProductAssignment over it, cuMat/src/SparseProductEvaluation.h:144-187,
    261-324)."""

    def __init__(self, e, pattern):
        e = as_expr(e)
        if e.dtype not in (capi.F32, capi.F64, capi.I32):
            raise TypeError("sparse-expression products take float32, float64 or int32 expressions")
        self.expr = e
        self.pattern = pattern
        self.ell = isinstance(pattern, ELLMatrix)
        if not self.ell and not isinstance(pattern, SparsityPattern):
            raise TypeError("sparseView needs a SparsityPattern (CSR) or an ELLMatrix (ELLPACK pattern)")
        R, Cc = pattern.rows, pattern.cols
        if e.shape[0] not in (1, R) or e.shape[1] not in (1, Cc):
            raise capi.CumatArgError(capi.CM_E_DIM, f"expression of shape {e.shape} viewed as {R}x{Cc}")
        for leaf in _value_leaves(e):
            own = leaf.matrix.pattern if not self.ell else leaf.matrix
            if not self._same_pattern(own):
                raise capi.CumatArgError(capi.CM_E_ARG, "a sparse matrix in the expression has another pattern "
                                         "than the view (only the view's own pattern is read directly)")

    def _same_pattern(self, other):
        p = self.pattern
        if other is p:
            return True
        if self.ell:
            return (isinstance(other, ELLMatrix) and other.rows == p.rows and other.npr == p.npr
                    and other.indices.ptr == p.indices.ptr)
        return (isinstance(other, SparsityPattern) and other.rows == p.rows and other.nnz == p.nnz
                and np.array_equal(other.host_JA, p.host_JA) and np.array_equal(other.host_IA, p.host_IA))

    @property
    def rows(self):
        return self.pattern.rows

    @property
    def cols(self):
        return self.pattern.cols

    @property
    def batches(self):
        return self.expr.shape[2]

    def matvec(self, x: DeviceArray, out: DeviceArray | None = None, mode: int = capi.ASSIGN,
               ctx=None) -> DeviceArray:
        p = self.pattern
        if x.dtype != self.expr.dtype:
            raise TypeError("no implicit casting between the expression and the vector")
        if x.rows != p.cols or x.cols != 1:
            raise capi.CumatArgError(capi.CM_E_DIM, f"SpMV: x is {x.rows}x{x.cols}, need {p.cols}x1")
        batches = max(self.batches, x.batches)
        if self.batches not in (1, batches) or x.batches not in (1, batches):
            raise capi.CumatArgError(capi.CM_E_DIM, "SpMV: batch counts do not broadcast")
        if out is None:
            out = DeviceArray(p.rows, 1, batches, dtype=x.np_dtype, ctx=x.ctx)
        if out.dtype != x.dtype or out.rows != p.rows or out.cols != 1 or out.batches != batches:
            raise capi.CumatArgError(capi.CM_E_DIM, "SpMV: output has the wrong size or dtype")
        leaves, nl, code, nc = lower(self.expr, (p.rows, p.cols, batches))
        if self.ell:
            capi.call("cm_ell_expr_spmv", ctx_handle(ctx), leaves, nl, code, nc, x.dtype, p.rows, p.cols,
                      p.npr, batches, p.indices.ptr, x.ptr, int(x.batches > 1), out.ptr, mode)
        else:
            capi.call("cm_csr_expr_spmv", ctx_handle(ctx), leaves, nl, code, nc, x.dtype, p.rows, p.cols,
                      p.nnz, batches, p.JA.ptr, p.IA.ptr, x.ptr, int(x.batches > 1), out.ptr, mode)
        return out

    __matmul__ = matmul = matvec

    def __mul__(self, x):
        return self.matvec(x)


def sparse_view(e, pattern) -> SparseExpressionView:
    """expr.sparseView<CSR|ELLPACK>(pattern) (cuMat/src/SparseExpressionOpPlugin.inl:54-57)."""
    return SparseExpressionView(e, pattern)


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
        # values: rows x npr, or batches x rows x npr; stored as cuMat's
        # DataMatrix (rows x npr x batches, column-major, batch stride rows*npr)
        v = np.asarray(values_rowsxk)
        if v.ndim == 2:
            v = v[None]
        if v.ndim != 3 or v.shape[1:] != ind.shape:
            raise capi.CumatArgError(capi.CM_E_DIM, "ELLPACK values must be rows x npr or batches x rows x npr")
        self.indices = DeviceArray.from_numpy(np.asfortranarray(ind).reshape(-1, order="F"), ctx=ctx)
        self.batches = int(v.shape[0])
        flat = np.concatenate([np.asfortranarray(v[k]).reshape(-1, order="F") for k in range(self.batches)])
        self.values = DeviceArray.from_numpy(flat, rows=max(flat.size // self.batches, 1), cols=1,
                                             batches=self.batches, ctx=ctx)
        self.dtype = self.values.dtype

    def matvec(self, x: DeviceArray, out: DeviceArray | None = None, mode=capi.ASSIGN, ctx=None):
        """y (mode)= A * x (ELLPACKMVKernel_StaticBatches,
        cuMat/src/SparseProductEvaluation.h:194-324). The kernel takes one
        scalar type for values, x and y, so all three must agree; a batched
        matrix (values of shape rows x npr x batches) and a batched x
        broadcast against each other as in the CSR product."""
        if x.dtype != self.dtype:
            raise capi.CumatArgError(capi.CM_E_ARG, f"ELLPACK SpMV: x has dtype {x.np_dtype}, "
                                     f"the matrix {np.dtype(self.values.np_dtype)}")
        if x.rows != self.cols or x.cols != 1:
            raise capi.CumatArgError(capi.CM_E_DIM, f"ELLPACK SpMV: x is {x.rows}x{x.cols}, need {self.cols}x1")
        batches = max(self.batches, x.batches)
        if self.batches not in (1, batches) or x.batches not in (1, batches):
            raise capi.CumatArgError(capi.CM_E_DIM, "ELLPACK SpMV: batch counts do not broadcast")
        if out is None:
            out = DeviceArray(self.rows, 1, batches, dtype=x.np_dtype, ctx=x.ctx)
        if out.dtype != self.dtype or out.rows != self.rows or out.cols != 1 or out.batches != batches:
            raise capi.CumatArgError(capi.CM_E_DIM, "ELLPACK SpMV: output has the wrong size or dtype")
        capi.call("cm_ell_spmv", ctx_handle(ctx), self.dtype, self.rows, self.cols, self.npr,
                  batches, self.indices.ptr, self.values.ptr, int(self.batches > 1), x.ptr,
                  int(x.batches > 1), out.ptr, mode)
        return out

    __matmul__ = matmul = matvec

    def direct(self):
        """The stored values as an expression leaf read at each slot of this
        matrix's own pattern (column-major rows x npr data matrix)."""
        return SparseValueLeaf(self, self.rows * self.npr)

    def _leaf(self):
        return self.direct()

    def __mul__(self, o):
        if isinstance(o, DeviceArray):
            return self.matvec(o)
        return self._leaf() * o

    def __rmul__(self, o): return o * self._leaf()
    def __add__(self, o): return self._leaf() + _sparse_operand(o)
    def __radd__(self, o): return _sparse_operand(o) + self._leaf()
    def __sub__(self, o): return self._leaf() - _sparse_operand(o)
    def __rsub__(self, o): return _sparse_operand(o) - self._leaf()
    def __neg__(self): return -self._leaf()

    def sparse_view(self, pattern=None):
        return SparseExpressionView(self._leaf(), self if pattern is None else pattern)


class BlockSparseMatrix:
    """CSR matrix of block scalars: the reference's SparseMatrix<real3x3>
    with its plug-in types (tests/TestBlockedConjugateGradient.cu:12-239).
    A block is block_dim (2..4) float4 rows (real3x3: three rows, w = 0);
    rows, cols and nnz count blocks. Values are (nnz, D, 4) float32, or
    (batches, nnz, D, 4) for a batched matrix. A vector is a float32
    DeviceArray of 4 * rows entries (one float4 per block row), batched like
    any DeviceArray."""

    def __init__(self, pattern: SparsityPattern, block_dim: int, values):
        D = int(block_dim)
        if not 2 <= D <= 4:
            raise capi.CumatArgError(capi.CM_E_ARG, f"block dimension {D} (2..4)")
        v = np.asarray(values, np.float32)
        if v.ndim == 3:
            v = v[None]
        if v.ndim != 4 or v.shape[1:] != (pattern.nnz, D, 4):
            raise capi.CumatArgError(capi.CM_E_DIM, f"block values must be (nnz, {D}, 4) or "
                                     f"(batches, nnz, {D}, 4), got {np.shape(values)}")
        self.pattern = pattern
        self.block_dim = D
        self.batches = int(v.shape[0])
        per = max(pattern.nnz * D * 4, 4)
        flat = np.zeros(per * self.batches, np.float32)
        for k in range(self.batches):
            flat[k * per:k * per + v[k].size] = v[k].reshape(-1)
        self.values = DeviceArray.from_numpy(flat, rows=per, cols=1, batches=self.batches, ctx=pattern.ctx)
        self.dtype = self.values.dtype

    @classmethod
    def from_bcsr(cls, rows, cols, JA, IA, vals, ctx=None):
        v = np.asarray(vals, np.float32)
        pat = SparsityPattern(rows, cols, JA, IA, ctx=ctx)
        return cls(pat, v.shape[-2], v)

    @property
    def rows(self):
        return self.pattern.rows

    @property
    def cols(self):
        return self.pattern.cols

    @property
    def nnz(self):
        return self.pattern.nnz

    def _check_vec(self, v: DeviceArray, entries: int, what: str):
        if v.dtype != capi.F32:
            raise capi.CumatArgError(capi.CM_E_ARG, f"block SpMV: {what} must be float32")
        if v.rows != 4 * entries or v.cols != 1:
            raise capi.CumatArgError(capi.CM_E_DIM, f"block SpMV: {what} is {v.rows}x{v.cols}, "
                                     f"need {4 * entries}x1 (one float4 per block)")

    def matvec(self, x: DeviceArray, out: DeviceArray | None = None, mode: int = capi.ASSIGN, ctx=None):
        """y (mode)= A * x with the block product functor, rows summed in CSR
        order (cuMat/src/SparseProductEvaluation.h:43-139)."""
        self._check_vec(x, self.cols, "x")
        batches = max(self.batches, x.batches)
        if self.batches not in (1, batches) or x.batches not in (1, batches):
            raise capi.CumatArgError(capi.CM_E_DIM, "block SpMV: batch counts do not broadcast")
        if out is None:
            out = DeviceArray(4 * self.rows, 1, batches, dtype=np.float32, ctx=x.ctx)
        self._check_vec(out, self.rows, "the output")
        if out.batches != batches:
            raise capi.CumatArgError(capi.CM_E_DIM, "block SpMV: output has the wrong batch count")
        p = self.pattern
        capi.call("cm_bcsr_spmv", ctx_handle(ctx), self.block_dim, p.rows, p.cols, p.nnz, batches, p.JA.ptr,
                  p.IA.ptr, self.values.ptr, int(self.batches > 1), x.ptr, int(x.batches > 1), out.ptr, mode)
        return out

    __matmul__ = matmul = matvec

    def inverse_diagonal(self, ctx=None) -> DeviceArray:
        """A.diagonal().cwiseInverseCheck() over block scalars: one float4 per
        block row (components >= block_dim are 0)."""
        p = self.pattern
        out = DeviceArray(4 * p.rows, 1, self.batches, dtype=np.float32, ctx=ctx or p.ctx)
        capi.call("cm_bcsr_inverse_diagonal", ctx_handle(ctx or p.ctx), self.block_dim, p.rows, p.nnz,
                  self.batches, p.JA.ptr, p.IA.ptr, self.values.ptr, int(self.batches > 1), out.ptr)
        return out

