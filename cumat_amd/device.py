"""Device memory and contexts (cuMat Context / DevicePointer / Matrix storage,
cuMat/src/Context.h:132-446, cuMat/src/DevicePointer.h:11-148,
cuMat/src/Matrix.h:30-436)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi as capi

_NP2DT = {np.dtype(np.int32): capi.I32, np.dtype(np.float32): capi.F32,
          np.dtype(np.float64): capi.F64, np.dtype(np.bool_): capi.BOOL,
          np.dtype(np.uint8): capi.BOOL, np.dtype(np.uint16): capi.BF16}
_DT2NP = {capi.I32: np.int32, capi.F32: np.float32, capi.F64: np.float64, capi.BOOL: np.bool_,
          capi.BF16: np.uint16}
_DTSIZE = {capi.I32: 4, capi.F32: 4, capi.F64: 8, capi.BOOL: 1, capi.BF16: 2}


def np_to_dtype(dt) -> int:
    return _NP2DT[np.dtype(dt)]


def dtype_to_np(dt: int):
    return _DT2NP[dt]


class Context:
    """A device + HIP stream (cuMat::Context). ``None`` everywhere means the
    calling thread's default context."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        capi.call("cm_ctx_create", device, C.byref(h))
        self.handle = h.value
        self.device = device

    def stream(self) -> int:
        s = C.c_void_p()
        capi.call("cm_ctx_get_stream", self.handle, C.byref(s))
        return s.value or 0

    def set_stream(self, stream: int) -> None:
        capi.call("cm_ctx_set_stream", self.handle, C.c_void_p(stream))

    def sync(self) -> None:
        capi.call("cm_ctx_sync", self.handle)

    def close(self) -> None:
        if self.handle:
            capi.call("cm_ctx_destroy", self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def default_context():
    return None


def ctx_handle(ctx) -> int | None:
    if ctx is None:
        return None
    return ctx.handle if isinstance(ctx, Context) else ctx


def synchronize(ctx=None) -> None:
    capi.call("cm_ctx_sync", ctx_handle(ctx))


class DeviceArray:
    """A dense rows x cols x batches array in device memory with cuMat's
    layouts: column major idx = r + R*(c + C*b), row major idx = c + C*(r + R*b)
    (cuMat/src/Matrix.h:613-627). Memory comes from cm_malloc (caching) unless
    an external pointer is wrapped."""

    def __init__(self, rows: int, cols: int = 1, batches: int = 1, dtype=np.float32,
                 row_major: bool = False, ctx=None, ptr: int | None = None, owner=None):
        self.rows, self.cols, self.batches = int(rows), int(cols), int(batches)
        self.dtype = capi.BOOL if np.dtype(dtype) == np.bool_ else np_to_dtype(dtype)
        self.row_major = bool(row_major)
        self.ctx = ctx
        self._owner = owner
        self._own = ptr is None
        nbytes = self.size * _DTSIZE[self.dtype]
        if ptr is None:
            p = C.c_void_p()
            capi.call("cm_malloc", ctx_handle(ctx), max(nbytes, 1), C.byref(p))
            self.ptr = p.value
        else:
            self.ptr = ptr

    # ---- geometry
    @property
    def size(self) -> int:
        return self.rows * self.cols * self.batches

    @property
    def nbytes(self) -> int:
        return self.size * _DTSIZE[self.dtype]

    @property
    def np_dtype(self):
        return dtype_to_np(self.dtype)

    def strides(self):
        """Element strides (sr, sc, sb)."""
        R, Cc = self.rows, self.cols
        if self.row_major:
            return (Cc, 1, R * Cc)
        return (1, R, R * Cc)

    # ---- host transfers
    @classmethod
    def from_numpy(cls, a, rows=None, cols=None, batches=None, row_major=False, ctx=None):
        """Flat array in the given layout, or a (batches, rows, cols) / (rows, cols)
        / (rows,) array interpreted logically (like Matrix::fromArray)."""
        a = np.asarray(a)
        if rows is None:
            if a.ndim == 1:
                rows, cols, batches = a.shape[0], 1, 1
                flat = a
            elif a.ndim == 2:
                rows, cols, batches = a.shape[0], a.shape[1], 1
                flat = _pack(a[None], row_major)
            else:
                batches, rows, cols = a.shape
                flat = _pack(a, row_major)
        else:
            cols = 1 if cols is None else cols
            batches = 1 if batches is None else batches
            flat = a.reshape(-1)
        dt = a.dtype
        if dt == np.bool_:
            flat = flat.astype(np.uint8)
        out = cls(rows, cols, batches, dtype=dt, row_major=row_major, ctx=ctx)
        flat = np.ascontiguousarray(flat)
        assert flat.size == out.size, (flat.size, out.size)
        capi.call("cm_memcpy_h2d", ctx_handle(ctx), out.ptr, flat.ctypes.data, out.nbytes)
        return out

    def numpy(self, logical: bool = False):
        """Copy to host. Flat in storage order, or (batches, rows, cols) if logical."""
        npdt = np.uint8 if self.dtype == capi.BOOL else self.np_dtype
        h = np.empty(self.size, npdt)
        if self.size:
            capi.call("cm_memcpy_d2h", ctx_handle(self.ctx), h.ctypes.data, self.ptr, self.nbytes)
        if self.dtype == capi.BOOL:
            h = h.astype(np.bool_)
        if logical:
            return _unpack(h, self.rows, self.cols, self.batches, self.row_major)
        return h

    def copy_from(self, other: "DeviceArray", ctx=None) -> None:
        """Device copy, ordered on ctx (default: this array's context)."""
        assert other.nbytes == self.nbytes
        ctx = self.ctx if ctx is None else ctx
        capi.call("cm_memcpy_d2d", ctx_handle(ctx), self.ptr, other.ptr, self.nbytes)

    def fill_zero(self) -> None:
        """Matrix::setZero (cuMat/src/Matrix.h:1232-1238)."""
        capi.call("cm_fill_bytes", ctx_handle(self.ctx), self.ptr, 0, self.nbytes)

    def free(self) -> None:
        if self._own and self.ptr:
            capi.call("cm_free", ctx_handle(self.ctx), self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            if getattr(self, "_own", False) and getattr(self, "ptr", None):
                capi.lib().cm_free(ctx_handle(self.ctx), self.ptr)
                self.ptr = None
        except Exception:
            pass

    # ---- expression sugar (see expr.py)
    def expr(self):
        from .expr import LeafNode
        return LeafNode(self)

    def __repr__(self):
        return (f"DeviceArray({self.rows}x{self.cols}x{self.batches}, dtype={self.dtype}, "
                f"{'row' if self.row_major else 'col'}-major)")


def _pack(a3, row_major):
    # a3: (batches, rows, cols) logical -> flat storage order
    if row_major:
        return np.ascontiguousarray(a3).reshape(-1)
    return np.ascontiguousarray(np.transpose(a3, (0, 2, 1))).reshape(-1)


def _unpack(flat, R, Cc, B, row_major):
    if row_major:
        return flat.reshape(B, R, Cc)
    return np.transpose(flat.reshape(B, Cc, R), (0, 2, 1))
