"""Differentiable CG for PyTorch (no reference counterpart).

x = solve(A(theta), b) on a family of variable-coefficient edge Laplacians
(A_ii = sum of incident theta_e, A_ij = -theta_e, Dirichlet edges on the
diagonal only). The backward pass needs one more CG solve with A^T = A and one
fused per-edge kernel; only x and the assembled values are kept, not the
iterates:
    lam          = A^{-1} dL/dx
    dL/db        = lam
    dL/dtheta_e  = -lam^T (dA/dtheta_e) x = -(lam_i - lam_j)(x_i - x_j)
                   (boundary edge: -lam_i x_i)
Everything runs in the HIP library through ctypes on the caller's torch
stream; torch only provides device memory and the stream.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _capi as capi

_DT = {torch.float32: capi.F32, torch.float64: capi.F64}


class _StreamCtx:
    """One library context per device whose stream follows torch's current
    stream."""
    _by_dev: dict = {}

    @classmethod
    def get(cls, device: torch.device):
        idx = device.index if device.index is not None else torch.cuda.current_device()
        c = cls._by_dev.get(idx)
        if c is None:
            h = C.c_void_p()
            capi.call("cm_ctx_create", idx, C.byref(h))
            c = {"h": h.value, "stream": None}
            cls._by_dev[idx] = c
        s = torch.cuda.current_stream(idx).cuda_stream
        if c["stream"] != s:
            capi.call("cm_ctx_set_stream", c["h"], C.c_void_p(s))
            c["stream"] = s
        return c["h"]


def _p(t: torch.Tensor):
    return C.c_void_p(t.data_ptr())


class EdgeLaplacian:
    """Pattern and edge maps of A(theta) on one device (from
    cumat_amd.problems.grid_laplacian)."""

    def __init__(self, P: dict, device="cuda"):
        dev = torch.device(device)
        t = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.int32), device=dev)  # noqa: E731
        self.n, self.nnz, self.nedges = int(P["n"]), int(P["nnz"]), int(P["nedges"])
        self.JA, self.IA = t(P["JA"]), t(P["IA"])
        self.eptr, self.eidx = t(P["eptr"]), t(P["eidx"])
        self.ei, self.ej = t(P["ei"]), t(P["ej"])
        self.device = dev

    def _check(self, name: str, t: torch.Tensor, last: int, batches: int | None = None,
               dtype: torch.dtype | None = None) -> None:
        """Every tensor handed to the kernels must be a (B, last) tensor of a
        supported dtype on this operator's device: the kernels receive raw
        pointers and derive every extent from n, nnz and nedges."""
        if t.dtype not in _DT:
            raise capi.CumatArgError(capi.CM_E_ARG, f"{name}: dtype {t.dtype} (need float32 or float64)")
        if dtype is not None and t.dtype != dtype:
            raise capi.CumatArgError(capi.CM_E_ARG, f"{name}: dtype {t.dtype}, expected {dtype}")
        if t.device.type != "cuda" or (self.device.index is not None and t.device.index != self.device.index):
            raise capi.CumatArgError(capi.CM_E_ARG, f"{name}: on {t.device}, the operator is on {self.device}")
        if t.dim() != 2 or t.shape[1] != last or (batches is not None and t.shape[0] != batches):
            want = f"({batches if batches is not None else 'B'}, {last})"
            raise capi.CumatArgError(capi.CM_E_DIM, f"{name}: shape {tuple(t.shape)}, need {want}")
        if t.shape[0] < 1:
            raise capi.CumatArgError(capi.CM_E_DIM, f"{name}: no batches")

    def assemble(self, theta: torch.Tensor) -> torch.Tensor:
        """CSR values of A(theta) per batch: (B, nnz)."""
        self._check("theta", theta, self.nedges)
        theta = theta.contiguous()
        B = theta.shape[0]
        vals = torch.empty(B, self.nnz, dtype=theta.dtype, device=theta.device)
        h = _StreamCtx.get(theta.device)
        capi.call("cm_edge_assemble", h, _DT[theta.dtype], self.nnz, B, self.nedges, _p(self.eptr),
                  _p(self.eidx), _p(theta), _p(vals))
        return vals

    def solve(self, vals: torch.Tensor, rhs: torch.Tensor, tol: float, max_iter: int):
        """Per-batch Jacobi-PCG with batch-specific values; returns
        (x, iterations per batch, error per batch)."""
        self._check("rhs", rhs, self.n)
        self._check("vals", vals, self.nnz, batches=rhs.shape[0], dtype=rhs.dtype)
        vals, rhs = vals.contiguous(), rhs.contiguous()
        B = rhs.shape[0]
        x = torch.empty_like(rhs)
        its = (C.c_int * B)()
        errs = (C.c_double * B)()
        st = capi.CgStats()
        h = _StreamCtx.get(rhs.device)
        capi.call("cm_pcg_solve_ex", h, _DT[rhs.dtype], self.n, self.nnz, B, _p(self.JA), _p(self.IA),
                  _p(vals), 1, None, _p(rhs), _p(x), 0, float(tol), int(max_iter), 1, its, errs,
                  C.byref(st))
        return x, list(its), list(errs)

    def solve_mixed(self, vals: torch.Tensor, rhs: torch.Tensor, tol: float, inner_tol: float = 1e-5,
                    storage: str = "f32", max_outer: int = 50):
        """Per-batch mixed-precision CG with FP64 refinement (cm_pcg_ir_solve):
        FP64 values and right-hand sides, FP32 (or BF16) inner solves, the
        residual and the solution kept in FP64 until the FP64 relative
        residual is at most tol. Returns (x, inner iterations per batch,
        final relative residual per batch)."""
        self._check("rhs", rhs, self.n, dtype=torch.float64)
        self._check("vals", vals, self.nnz, batches=rhs.shape[0], dtype=torch.float64)
        vals, rhs = vals.contiguous(), rhs.contiguous()
        B = rhs.shape[0]
        x = torch.empty_like(rhs)
        h = _StreamCtx.get(rhs.device)
        vdt = capi.BF16 if storage == "bf16" else capi.F32
        its, res = [], []
        for k in range(B):
            st = capi.IrStats()
            capi.call("cm_pcg_ir_solve", h, self.n, self.nnz, _p(self.JA), _p(self.IA),
                      C.c_void_p(vals.data_ptr() + 8 * k * self.nnz), vdt,
                      C.c_void_p(rhs.data_ptr() + 8 * k * self.n), C.c_void_p(x.data_ptr() + 8 * k * self.n),
                      float(tol), float(inner_tol), int(max_outer), -1, C.byref(st))
            its.append(st.inner_iterations)
            res.append(st.final_rel_residual)
        return x, its, res

    def edge_grad(self, lam: torch.Tensor, x: torch.Tensor) -> torch.Tensor:
        self._check("x", x, self.n)
        self._check("lam", lam, self.n, batches=x.shape[0], dtype=x.dtype)
        lam, x = lam.contiguous(), x.contiguous()
        B = x.shape[0]
        g = torch.empty(B, self.nedges, dtype=x.dtype, device=x.device)
        h = _StreamCtx.get(x.device)
        capi.call("cm_edge_grad", h, _DT[x.dtype], self.nedges, B, self.n, _p(self.ei), _p(self.ej),
                  _p(lam), _p(x), _p(g))
        return g


class CGSolve(torch.autograd.Function):
    """x = A(theta)^{-1} b, differentiable in theta and b.

    precision "native": Jacobi-PCG in the dtype of theta and b (batched, one
    launch sequence for all right-hand sides; stops at the relative error
    tol). "mixed": float64 theta and b, each solve by FP32 (or, with
    "mixed-bf16", BF16-valued) inner CG inside FP64 refinement to the FP64
    relative residual tol -- FP64 accuracy of x, lambda and the gradient
    with the inner iterations in single precision."""

    @staticmethod
    def forward(ctx, theta, b, lap: EdgeLaplacian, tol: float, max_iter: int, precision: str = "native"):
        # theta and b must agree in dtype and batch count: the solver reads
        # one (B, nnz) block of values per right-hand side
        lap._check("b", b, lap.n)
        lap._check("theta", theta, lap.nedges, batches=b.shape[0], dtype=b.dtype)
        if precision not in ("native", "mixed", "mixed-bf16"):
            raise capi.CumatArgError(capi.CM_E_ARG, f"precision {precision!r} (native, mixed, mixed-bf16)")
        if precision != "native" and b.dtype != torch.float64:
            raise capi.CumatArgError(capi.CM_E_ARG, "mixed precision takes float64 theta and b")
        vals = lap.assemble(theta)
        if precision == "native":
            x, its, errs = lap.solve(vals, b, tol, max_iter)
        else:
            x, its, errs = lap.solve_mixed(vals, b, tol, storage="bf16" if precision == "mixed-bf16" else "f32")
        ctx.save_for_backward(vals, x)
        ctx.lap, ctx.tol, ctx.max_iter, ctx.precision = lap, tol, max_iter, precision
        ctx.forward_stats = (its, errs)
        CGSolve.last_forward_stats = (its, errs)
        return x

    @staticmethod
    def backward(ctx, gx):
        vals, x = ctx.saved_tensors
        lap = ctx.lap
        if ctx.precision == "native":
            lam, its, errs = lap.solve(vals, gx, ctx.tol, ctx.max_iter)  # A^T = A
        else:
            lam, its, errs = lap.solve_mixed(vals, gx, ctx.tol,
                                             storage="bf16" if ctx.precision == "mixed-bf16" else "f32")
        CGSolve.last_backward_stats = (its, errs)
        gtheta = lap.edge_grad(lam, x)
        return gtheta, lam, None, None, None, None


def cg_solve(theta: torch.Tensor, b: torch.Tensor, lap: EdgeLaplacian, tol: float = 1e-7,
             max_iter: int = 20000, precision: str = "native") -> torch.Tensor:
    return CGSolve.apply(theta, b, lap, tol, max_iter, precision)
