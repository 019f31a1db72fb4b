"""Iterative solvers: ConjugateGradient + DiagonalPreconditioner with cuMat's
API (cuMat/src/ConjugateGradient.h:66-265, cuMat/src/IterativeSolverBase.h:11-135,
cuMat/src/SolverBase.h:34-126) plus compute() (Eigen's name, which cuMat
lacks), and the two solver additions: persistent batched CG and
mixed-precision CG with FP64 iterative refinement."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi as capi
from . import dense
from .device import DeviceArray, ctx_handle
from .sparse import BlockSparseMatrix, SparseMatrix


def _scalar(matrix):
    return matrix.np_dtype if dense.is_dense(matrix) else matrix.values.np_dtype


class DiagonalPreconditioner:
    """Jacobi: entries = A.diagonal().cwiseInverseCheck()
    (cuMat/src/ConjugateGradient.h:218-246); A is a SparseMatrix or a dense
    DeviceArray. The entries are computed on the matrix's context (or ctx)
    and are complete when the constructor returns, so a solve on any other
    context may read them."""

    def __init__(self, matrix, ctx=None):
        if dense.is_dense(matrix):
            self.entries = dense.inverse_diagonal(matrix, ctx=ctx)
        else:
            self.entries = matrix.inverse_diagonal(ctx=ctx)
        capi.call("cm_ctx_sync", ctx_handle(self.entries.ctx))


class IdentityPreconditioner:
    def __init__(self, matrix):
        ones = np.ones(matrix.rows, _scalar(matrix))
        self.entries = DeviceArray.from_numpy(ones)


class ConjugateGradient:
    """x = ConjugateGradient(A).solve(b) with the reference's semantics:
    tolerance defaults to the scalar epsilon, maxIterations to 2*cols,
    error() = sqrt(max_b |r|^2/|b|^2), iterations() excludes the converging
    step, and all right-hand-side batches iterate until all converge.
    The matrix is a CSR SparseMatrix, a dense square DeviceArray
    (ConjugateGradient<MatrixXf>, with the dense product A * p) or a
    BlockSparseMatrix (ConjugateGradient<SparseMatrix<real3x3>>, the
    reference's block-scalar CG, tests/TestBlockedConjugateGradient.cu:244-275;
    vectors of one float4 per block row)."""

    def __init__(self, matrix: SparseMatrix | None = None, preconditioner=None, path: int = capi.CG_AUTO):
        self._tol = None
        self._max_iter = -1
        self._iters = 0
        self._error = 0.0
        self.path = path
        self.last_stats = None
        self.matrix = None
        if matrix is not None:
            self.compute(matrix, preconditioner)

    def compute(self, matrix, preconditioner=None):
        if matrix.rows != matrix.cols:
            raise capi.CumatArgError(capi.CM_E_DIM, "Matrix must be square")
        if matrix.batches != 1:
            raise capi.CumatArgError(capi.CM_E_DIM, "the preconditioner needs a batch-1 matrix "
                                     "(cuMat/src/ConjugateGradient.h:223)")
        self.matrix = matrix
        self.preconditioner = preconditioner or DiagonalPreconditioner(matrix)
        return self

    # ---- IterativeSolverBase accessors (cuMat/src/IterativeSolverBase.h:54-111)
    def tolerance(self):
        if self._tol is None:
            return float(np.finfo(_scalar(self.matrix)).eps)
        return self._tol

    def setTolerance(self, tol):
        self._tol = float(tol)
        return self

    def maxIterations(self):
        return 2 * self.matrix.cols if self._max_iter < 0 else self._max_iter

    def setMaxIterations(self, n):
        self._max_iter = int(n)
        return self

    def iterations(self):
        return self._iters

    def error(self):
        return self._error

    # ---- solve
    def solve(self, b: DeviceArray, x: DeviceArray | None = None, ctx=None) -> DeviceArray:
        return self._solve(b, x, False, ctx)

    def solveWithGuess(self, b: DeviceArray, guess: DeviceArray, ctx=None) -> DeviceArray:
        x = DeviceArray(guess.rows, 1, guess.batches, dtype=guess.np_dtype, ctx=guess.ctx)
        x.copy_from(guess, ctx=ctx)  # on the solve's stream, so the solve sees the copy
        return self._solve(b, x, True, ctx)

    def _solve(self, b, x, use_guess, ctx):
        A = self.matrix
        if isinstance(A, BlockSparseMatrix):
            return self._solve_block(b, x, use_guess, ctx)
        if b.rows != A.cols or b.cols != 1:
            raise capi.CumatArgError(capi.CM_E_DIM, "rhs has the wrong size")
        if x is None:
            x = DeviceArray(b.rows, 1, b.batches, dtype=b.np_dtype, ctx=b.ctx)
        it = C.c_int(0)
        err = C.c_double(0)
        st = capi.CgStats()
        if dense.is_dense(A):
            capi.call("cm_pcg_dense_solve", ctx_handle(ctx), A.dtype, A.rows, b.batches, A.ptr, 0,
                      int(A.row_major), self.preconditioner.entries.ptr, b.ptr, x.ptr, int(use_guess),
                      float(self.tolerance()), int(self.maxIterations()), C.byref(it), C.byref(err),
                      C.byref(st))
            self._iters, self._error = it.value, err.value
            self.last_stats = {"path": st.path_used, "launches": st.launches, "seconds": st.seconds}
            return x
        p = A.pattern
        capi.call("cm_pcg_solve", ctx_handle(ctx), A.dtype, p.rows, p.nnz, b.batches, p.JA.ptr, p.IA.ptr,
                  A.values.ptr, self.preconditioner.entries.ptr, b.ptr, x.ptr, int(use_guess),
                  float(self.tolerance()), int(self.maxIterations()), int(self.path), C.byref(it),
                  C.byref(err), C.byref(st))
        self._iters, self._error = it.value, err.value
        self.last_stats = {"path": st.path_used, "launches": st.launches, "seconds": st.seconds}
        return x


    def _solve_block(self, b, x, use_guess, ctx):
        A = self.matrix
        if b.dtype != capi.F32 or b.rows != 4 * A.cols or b.cols != 1:
            raise capi.CumatArgError(capi.CM_E_DIM, "rhs must be float32 with one float4 per block row")
        if x is None:
            x = DeviceArray(b.rows, 1, b.batches, dtype=np.float32, ctx=b.ctx)
        if x.dtype != capi.F32 or x.rows != b.rows or x.batches != b.batches:
            raise capi.CumatArgError(capi.CM_E_DIM, "solution has the wrong size")
        it = C.c_int(0)
        err = C.c_double(0)
        st = capi.CgStats()
        p = A.pattern
        capi.call("cm_pcg_bcsr_solve", ctx_handle(ctx), A.block_dim, p.rows, p.nnz, b.batches, p.JA.ptr,
                  p.IA.ptr, A.values.ptr, 0, self.preconditioner.entries.ptr, b.ptr, x.ptr, int(use_guess),
                  float(self.tolerance()), int(self.maxIterations()), C.byref(it), C.byref(err), C.byref(st))
        self._iters, self._error = it.value, err.value
        self.last_stats = {"path": st.path_used, "launches": st.launches, "seconds": st.seconds}
        return x


def pcg_batched(row_off: DeviceArray, JA: DeviceArray, IA: DeviceArray, vals: DeviceArray,
                b: DeviceArray, x: DeviceArray, tol: float, max_iter: int, iters: DeviceArray,
                errors: DeviceArray, block_size: int = 0, use_graph: bool = False, ctx=None):
    """Persistent batched CG over independent CSR systems (block-diagonal
    concatenation, row_off gives each system's rows)."""
    nsys = row_off.size - 1
    capi.call("cm_pcg_batched_solve", ctx_handle(ctx), nsys, row_off.ptr, x.size, IA.size, JA.ptr,
              IA.ptr, vals.ptr, b.ptr, x.ptr, C.c_float(tol), int(max_iter), iters.ptr, errors.ptr,
              int(block_size), int(use_graph))


def pcg_batched_autotune(row_off, JA, IA, vals, b, tol, max_iter, ctx=None, with_groups=False):
    """Times the launch shapes of the persistent batched CG on this batch
    (threads per system x workgroups per CU); a later pcg_batched of the same
    batch shape with block_size=0 uses the best. Returns (block size, ms), or
    (block size, workgroups per CU, ms) with with_groups=True."""
    bs = C.c_int(0)
    gpc = C.c_int(0)
    ms = C.c_float(0)
    capi.call("cm_pcg_batched_autotune", ctx_handle(ctx), row_off.size - 1, row_off.ptr, b.size,
              IA.size, JA.ptr, IA.ptr, vals.ptr, b.ptr, C.c_float(tol), int(max_iter), C.byref(bs),
              C.byref(gpc), C.byref(ms))
    if with_groups:
        return bs.value, gpc.value, ms.value
    return bs.value, ms.value


def pcg_ir(JA: DeviceArray, IA: DeviceArray, vals64: DeviceArray, b: DeviceArray, x: DeviceArray,
           storage: str = "f32", tol: float = 1e-12, inner_tol: float = 1e-4, max_outer: int = 50,
           max_inner: int = -1, ctx=None) -> dict:
    """Mixed-precision CG with FP64 iterative refinement; returns stats."""
    st = capi.IrStats()
    vdt = capi.BF16 if storage == "bf16" else capi.F32
    n = b.rows
    capi.call("cm_pcg_ir_solve", ctx_handle(ctx), n, IA.size, JA.ptr, IA.ptr, vals64.ptr, vdt, b.ptr,
              x.ptr, float(tol), float(inner_tol), int(max_outer), int(max_inner), C.byref(st))
    return {"outer": st.outer_iterations, "inner": st.inner_iterations,
            "rel_residual": st.final_rel_residual, "seconds": st.seconds}
