"""CPU oracle for the cuMat hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this package, and only as the checker. The product
package ``cumat_amd`` never imports it.

Two back ends:

* ``liboracle.so`` -- the plain-C restatement in ``cumat_oracle.c`` (every
  function cites the reference file:line it follows);
* ``_ref/libeigen_ref.so`` -- the Eigen 3.3.90 headers vendored by the
  reference (``third-party/Eigen``) behind a small driver, built only where
  the reference tree exists; the tests use it to pin the restatement.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libeigen_ref.so")

I32, F32, F64, BOOL = 0, 1, 2, 3
_DT = {np.dtype(np.int32): I32, np.dtype(np.float32): F32, np.dtype(np.float64): F64,
       np.dtype(np.bool_): BOOL, np.dtype(np.uint8): BOOL}

_lib = None
_ref = None


def build(with_ref: bool = True) -> None:
    """Compile the oracle (and the Eigen reference driver if the reference
    tree is present)."""
    subprocess.check_call(["make", "-s", "-C", HERE, "all"])
    if with_ref and os.path.isdir("/root/reference/third-party/Eigen"):
        subprocess.check_call(["make", "-s", "-C", HERE, "ref"])


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle/liboracle.so is not built (run oracle.build())")
        L = C.CDLL(LIB_PATH)
        vp, i32, i64, f64 = C.c_void_p, C.c_int, C.c_int64, C.c_double
        L.orc_poisson2d.argtypes = [i32, f64, vp, vp, vp]
        L.orc_poisson2d.restype = i64
        for name in ("orc_csr_spmv_f32", "orc_csr_spmv_f64", "orc_csr_spmv_i32"):
            getattr(L, name).argtypes = [i32, i32, i32, i32, vp, vp, vp, i32, vp, i32, vp, i32]
        for name in ("orc_ell_spmv_f32", "orc_ell_spmv_f64", "orc_ell_spmv_i32"):
            getattr(L, name).argtypes = [i32, i32, i32, i32, vp, vp, i32, vp, i32, vp, i32]
        L.orc_reduce.argtypes = [i32, i32, i32, i32, i32, i32, i32, vp, i32, vp]
        L.orc_dot.argtypes = [i32, i64, i32, vp, i32, vp, i32, vp]
        L.orc_cwise_binary.argtypes = [i32, i32, i64, vp, vp, vp]
        L.orc_cwise_unary.argtypes = [i32, i32, i64, vp, vp]
        L.orc_assign.argtypes = [i32, i32, i64, vp, vp]
        L.orc_csr_inverse_diagonal_f64.argtypes = [i32, vp, vp, vp, vp]
        L.orc_bcsr_spmv_f32.argtypes = [i32, i32, i32, i32, i32, vp, vp, vp, i32, vp, i32, vp, i32]
        L.orc_bcsr_inverse_diagonal_f32.argtypes = [i32, i32, i32, i32, vp, vp, vp, i32, vp]
        for name in ("orc_pcg_f32", "orc_pcg_f64"):
            getattr(L, name).argtypes = [i32, i32, i32, vp, vp, vp, vp, vp, i32, f64, i32,
                                         C.POINTER(C.c_double)]
            getattr(L, name).restype = i32
        L.orc_edge_grad_f64.argtypes = [i32, vp, vp, vp, vp, vp]
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref():
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError("oracle/_ref/libeigen_ref.so is not built")
        L = C.CDLL(REF_PATH)
        vp, i32, i64, f64 = C.c_void_p, C.c_int, C.c_int64, C.c_double
        for name in ("eref_pcg_f32", "eref_pcg_f64"):
            getattr(L, name).argtypes = [i32, i32, vp, vp, vp, vp, vp, i32, f64, i32, i32,
                                         C.POINTER(C.c_double)]
            getattr(L, name).restype = i32
        L.eref_csr_create_f32.argtypes = [i32, i32, i32, vp, vp, vp]
        L.eref_csr_create_f32.restype = vp
        L.eref_csr_destroy_f32.argtypes = [vp]
        L.eref_spmv_f32.argtypes = [vp, vp, vp, i32]
        L.eref_pcg_handle_f32.argtypes = [vp, vp, vp, f64, i32, i32, C.POINTER(C.c_double)]
        L.eref_pcg_handle_f32.restype = i32
        L.eref_dot_f32.argtypes = [i64, vp, vp]
        L.eref_dot_f32.restype = C.c_float
        L.eref_dot_f64.argtypes = [i64, vp, vp]
        L.eref_dot_f64.restype = C.c_double
        L.eref_lincomb8_f32.argtypes = [i64, vp, vp, vp]
        L.eref_axis_sum_f32.argtypes = [i32, i32, i32, i32, vp, vp]
        L.eref_threads_max.restype = i32
        L.eref_batched_pcg_f32.argtypes = [i32, vp, vp, vp, vp, vp, vp, f64, i32, i32, i32, vp]
        L.eref_batched_pcg_f32.restype = i32
        _ref = L
    return _ref


# ---------------------------------------------------------------- helpers
def poisson2d(g: int, shift: float = 0.0):
    """CSR (JA, IA, vals float64) of the 2D 5-point Poisson matrix."""
    n = g * g
    JA = np.zeros(n + 1, np.int32)
    IA = np.zeros(5 * n, np.int32)
    V = np.zeros(5 * n, np.float64)
    nnz = lib().orc_poisson2d(g, float(shift), _p(JA), _p(IA), _p(V))
    return JA, IA[:nnz].copy(), V[:nnz].copy()


def csr_spmv(JA, IA, vals, x, rows, cols, batches=1, vals_batched=False, x_batched=False,
             y=None, mode=0):
    dt = vals.dtype
    if y is None:
        y = np.zeros(rows * batches, dt)
    nnz = len(IA)
    fn = {np.dtype(np.float32): "orc_csr_spmv_f32", np.dtype(np.float64): "orc_csr_spmv_f64",
          np.dtype(np.int32): "orc_csr_spmv_i32"}[dt]
    JA = np.ascontiguousarray(JA, np.int32)
    IA = np.ascontiguousarray(IA, np.int32)
    getattr(lib(), fn)(rows, cols, nnz, batches, _p(JA), _p(IA), _p(vals), int(vals_batched),
                       _p(x), int(x_batched), _p(y), mode)
    return y


def ell_spmv(indices_colmajor, vals, x, rows, cols, nnz_per_row, batches=1,
             vals_batched=False, x_batched=False, y=None, mode=0):
    """ELLPACK product in the dtype of vals (float32, float64 or int32)."""
    dt = np.dtype(vals.dtype)
    fn = {np.dtype(np.float32): "orc_ell_spmv_f32", np.dtype(np.float64): "orc_ell_spmv_f64",
          np.dtype(np.int32): "orc_ell_spmv_i32"}[dt]
    if y is None:
        y = np.zeros(rows * batches, dt)
    ind = np.ascontiguousarray(indices_colmajor, np.int32)
    vals = np.ascontiguousarray(vals)
    x = np.ascontiguousarray(x, dt)
    getattr(lib(), fn)(rows, cols, nnz_per_row, batches, _p(ind), _p(vals),
                       int(vals_batched), _p(x), int(x_batched), _p(y), mode)
    return y


def ell_spmv_f32(indices_colmajor, vals, x, rows, cols, nnz_per_row, batches=1,
                 vals_batched=False, x_batched=False, y=None, mode=0):
    return ell_spmv(indices_colmajor, np.asarray(vals, np.float32), x, rows, cols, nnz_per_row, batches,
                    vals_batched, x_batched, y, mode)


def bcsr_spmv(D, JA, IA, vals, x, rows, cols, batches=1, vals_batched=False, x_batched=False,
              y=None, mode=0):
    """Block-CSR product: vals (nnz, D, 4) float32 blocks of D float4 rows,
    x (cols*4 per batch) float32; returns rows*4 per batch."""
    if y is None:
        y = np.zeros(rows * 4 * batches, np.float32)
    JA = np.ascontiguousarray(JA, np.int32)
    IA = np.ascontiguousarray(IA, np.int32)
    vals = np.ascontiguousarray(vals, np.float32)
    x = np.ascontiguousarray(x, np.float32)
    lib().orc_bcsr_spmv_f32(D, rows, cols, len(IA), batches, _p(JA), _p(IA), _p(vals), int(vals_batched),
                            _p(x), int(x_batched), _p(y), mode)
    return y


def bcsr_inverse_diagonal(D, JA, IA, vals, rows, batches=1, vals_batched=False):
    out = np.zeros(rows * 4 * (batches if vals_batched else 1), np.float32)
    JA = np.ascontiguousarray(JA, np.int32)
    IA = np.ascontiguousarray(IA, np.int32)
    vals = np.ascontiguousarray(vals, np.float32)
    lib().orc_bcsr_inverse_diagonal_f32(D, rows, len(IA), batches, _p(JA), _p(IA), _p(vals),
                                        int(vals_batched), _p(out))
    return out


def bcsr_expand(D, JA, IA, vals, rows, dtype=np.float64):
    """The scalar CSR matrix of size 4*rows that a block matrix acts as on
    float4 vectors: row 4i+q (q < D) holds row q of each block (i, j) at
    columns 4j..4j+3, in block order; rows 4i+q with q >= D are empty."""
    JA = np.asarray(JA, np.int64)
    IA = np.asarray(IA, np.int64)
    nnz = IA.size
    V = np.asarray(vals, np.float64).reshape(nnz, D, 4)
    brow = np.repeat(np.arange(rows, dtype=np.int64), np.diff(JA))
    q = np.arange(D, dtype=np.int64)
    c = np.arange(4, dtype=np.int64)
    srow = np.broadcast_to(4 * brow[:, None, None] + q[None, :, None], (nnz, D, 4)).reshape(-1)
    scol = np.broadcast_to(4 * IA[:, None, None] + c[None, None, :], (nnz, D, 4)).reshape(-1)
    order = np.argsort(srow, kind="stable")  # (block, q, c) order -> by row, blocks then c
    counts = np.bincount(srow, minlength=4 * rows)
    sJA = np.zeros(4 * rows + 1, np.int64)
    sJA[1:] = np.cumsum(counts)
    return (sJA.astype(np.int32), scol[order].astype(np.int32), V.reshape(-1)[order].astype(dtype))


def bcsr_pcg(D, JA, IA, vals, b, rows, batches=1, x0=None, tol=1e-6, max_iter=1000, dtype=np.float64):
    """Jacobi-PCG over block scalars (cuMat/src/ConjugateGradient.h:123-206
    with the plug-ins of tests/TestBlockedConjugateGradient.cu:111-239):
    every vector operation is per component, dot sums the four components,
    squaredNorm the first D (the rest stay 0), and the Jacobi entries are
    the diagonal blocks' diagonals. That is the scalar PCG of the expanded
    matrix (bcsr_expand) on b with components >= D zeroed: the empty rows'
    Jacobi entry 1 then only ever multiplies 0. Returns (x, iterations,
    error)."""
    sJA, sIA, sV = bcsr_expand(D, JA, IA, vals, rows, dtype)
    bz = np.array(b, dtype).reshape(batches, rows, 4)
    bz[:, :, D:] = 0
    return pcg(sJA, sIA, sV, bz.reshape(-1), 4 * rows, batches=batches,
               x0=None if x0 is None else np.asarray(x0, dtype), tol=tol, max_iter=max_iter)


def reduce(a, op, axis, rows, cols, batches, in_row_major, out_row_major):
    dt = _DT[a.dtype]
    oR = 1 if axis & 1 else rows
    oC = 1 if axis & 2 else cols
    oB = 1 if axis & 4 else batches
    out = np.zeros(oR * oC * oB, a.dtype)
    a = np.ascontiguousarray(a)
    lib().orc_reduce(dt, op, axis, rows, cols, batches, int(in_row_major), _p(a),
                     int(out_row_major), _p(out))
    return out


def dot(a, b, n, batches=1, a_batched=True, b_batched=True):
    out = np.zeros(batches, a.dtype)
    lib().orc_dot(_DT[a.dtype], n, batches, _p(a), int(a_batched), _p(b), int(b_batched),
                  _p(out))
    return out


def cwise_binary(fn, a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b, a.dtype)
    # logical ops and comparisons (cm_binary_fn >= LAND) produce bool bytes
    outdt = np.uint8 if fn >= 11 else a.dtype
    out = np.zeros(a.shape, outdt)
    lib().orc_cwise_binary(_DT[a.dtype], fn, a.size, _p(a), _p(b), _p(out))
    return out


def cwise_unary(fn, a):
    a = np.ascontiguousarray(a)
    out = np.zeros_like(a)
    lib().orc_cwise_unary(_DT[a.dtype], fn, a.size, _p(a), _p(out))
    return out


def assign(mode, dst, src):
    dst = np.array(dst, copy=True)
    src = np.ascontiguousarray(src, dst.dtype)
    lib().orc_assign(_DT[dst.dtype], mode, dst.size, _p(dst), _p(src))
    return dst


def inverse_diagonal(JA, IA, vals, rows):
    out = np.zeros(rows, np.float64)
    lib().orc_csr_inverse_diagonal_f64(rows, _p(JA), _p(IA),
                                       _p(np.ascontiguousarray(vals, np.float64)), _p(out))
    return out


def pcg(JA, IA, vals, b, n, batches=1, x0=None, tol=1e-6, max_iter=1000):
    """Jacobi-PCG with cuMat/Eigen semantics; returns (x, iterations, error)."""
    dt = vals.dtype
    JA = np.ascontiguousarray(JA, np.int32)
    IA = np.ascontiguousarray(IA, np.int32)
    b = np.ascontiguousarray(b, dt)
    x = np.zeros(n * batches, dt) if x0 is None else np.array(x0, dt, copy=True)
    err = C.c_double(0)
    fn = lib().orc_pcg_f32 if dt == np.float32 else lib().orc_pcg_f64
    it = fn(n, len(IA), batches, _p(JA), _p(IA), _p(vals), _p(b), _p(x), int(x0 is not None),
            float(tol), int(max_iter), C.byref(err))
    return x, it, err.value


def edge_grad(ei, ej, lam, x):
    out = np.zeros(len(ei), np.float64)
    lib().orc_edge_grad_f64(len(ei), _p(np.ascontiguousarray(ei, np.int32)),
                            _p(np.ascontiguousarray(ej, np.int32)),
                            _p(np.ascontiguousarray(lam, np.float64)),
                            _p(np.ascontiguousarray(x, np.float64)), _p(out))
    return out


# ------------------------------------------------------- Eigen reference
def eigen_pcg(JA, IA, vals, b, n, x0=None, tol=1e-6, max_iter=1000, nthreads=1):
    dt = vals.dtype
    JA = np.ascontiguousarray(JA, np.int32)
    IA = np.ascontiguousarray(IA, np.int32)
    b = np.ascontiguousarray(b, dt)
    x = np.zeros(n, dt) if x0 is None else np.array(x0, dt, copy=True)
    err = C.c_double(0)
    fn = ref().eref_pcg_f32 if dt == np.float32 else ref().eref_pcg_f64
    it = fn(n, len(IA), _p(JA), _p(IA), _p(vals), _p(b), _p(x), int(x0 is not None),
            float(tol), int(max_iter), int(nthreads), C.byref(err))
    return x, it, err.value
