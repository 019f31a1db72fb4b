"""Lazy component-wise expressions lowered to the accumulator program of
cm_cwise_eval (one kernel launch per assignment).

Mirrors cuMat's expression templates on the hot path: BinaryOp with
broadcasting of size-1 axes (cuMat/src/BinaryOps.h:69-167), HostScalar
(cuMat/src/NullaryOps.h:78-108), UnaryOp (cuMat/src/UnaryOps.h:128-225),
CastingOp (cuMat/src/UnaryOps.h:248-312), ExtractDiagonal / AsDiagonal
(cuMat/src/UnaryOps.h:361-487), transposed views, and the assignment modes
of CwiseAssignmentHandler (cuMat/src/CwiseOp.h:15-62). As in cuMat there is
no implicit casting between scalar types.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi as capi
from .device import DeviceArray, ctx_handle

_BOOL_RESULT = {capi.BIN_LAND, capi.BIN_LOR, capi.BIN_EQ, capi.BIN_NE, capi.BIN_LT, capi.BIN_GT,
                capi.BIN_LE, capi.BIN_GE}
_INT_ONLY = {capi.BIN_MOD, capi.BIN_AND, capi.BIN_OR, capi.BIN_XOR}
_FLOAT_ONLY = {capi.BIN_POW}


class Expr:
    """Base class; shape = (rows, cols, batches), 1 on broadcast axes."""
    dtype: int
    shape: tuple

    # arithmetic (cuMat/src/BinaryOps.h:286-293)
    def __add__(self, o): return Binary(capi.BIN_ADD, self, as_expr(o, self))
    def __radd__(self, o): return Binary(capi.BIN_ADD, as_expr(o, self), self)
    def __sub__(self, o): return Binary(capi.BIN_SUB, self, as_expr(o, self))
    def __rsub__(self, o): return Binary(capi.BIN_SUB, as_expr(o, self), self)
    def __mul__(self, o): return Binary(capi.BIN_MUL, self, as_expr(o, self))
    def __rmul__(self, o): return Binary(capi.BIN_MUL, as_expr(o, self), self)
    def __truediv__(self, o): return Binary(capi.BIN_DIV, self, as_expr(o, self))
    def __rtruediv__(self, o): return Binary(capi.BIN_DIV, as_expr(o, self), self)
    def __mod__(self, o): return Binary(capi.BIN_MOD, self, as_expr(o, self))
    def __and__(self, o): return Binary(capi.BIN_AND, self, as_expr(o, self))
    def __or__(self, o): return Binary(capi.BIN_OR, self, as_expr(o, self))
    def __xor__(self, o): return Binary(capi.BIN_XOR, self, as_expr(o, self))
    def __neg__(self): return Unary(capi.UN_NEG, self)
    def __invert__(self): return Unary(capi.UN_BITNOT, self)
    # comparisons (cuMat/src/BinaryOps.h:336-341)
    def __lt__(self, o): return Binary(capi.BIN_LT, self, as_expr(o, self))
    def __gt__(self, o): return Binary(capi.BIN_GT, self, as_expr(o, self))
    def __le__(self, o): return Binary(capi.BIN_LE, self, as_expr(o, self))
    def __ge__(self, o): return Binary(capi.BIN_GE, self, as_expr(o, self))
    def eq(self, o): return Binary(capi.BIN_EQ, self, as_expr(o, self))
    def ne(self, o): return Binary(capi.BIN_NE, self, as_expr(o, self))

    # named cwise functors (cuMat plugins)
    def cwiseMul(self, o): return Binary(capi.BIN_MUL, self, as_expr(o, self))
    def cwiseDot(self, o): return Binary(capi.BIN_MUL, self, as_expr(o, self))
    def cwiseDiv(self, o): return Binary(capi.BIN_DIV, self, as_expr(o, self))
    def cwisePow(self, o): return Binary(capi.BIN_POW, self, as_expr(o, self))
    def cwiseMin(self, o): return Binary(capi.BIN_MIN, self, as_expr(o, self))
    def cwiseMax(self, o): return Binary(capi.BIN_MAX, self, as_expr(o, self))
    def cwiseBinaryAnd(self, o): return Binary(capi.BIN_AND, self, as_expr(o, self))
    def cwiseBinaryOr(self, o): return Binary(capi.BIN_OR, self, as_expr(o, self))
    def cwiseBinaryXor(self, o): return Binary(capi.BIN_XOR, self, as_expr(o, self))
    def cwiseLogicalAnd(self, o): return Binary(capi.BIN_LAND, self, as_expr(o, self))
    def cwiseLogicalOr(self, o): return Binary(capi.BIN_LOR, self, as_expr(o, self))
    def cwiseAbs(self): return Unary(capi.UN_ABS, self)
    def cwiseAbs2(self): return Unary(capi.UN_ABS2, self)
    def cwiseInverse(self): return Unary(capi.UN_INVERSE, self)
    def cwiseInverseCheck(self): return Unary(capi.UN_INVERSE_CHECK, self)
    def cwiseNegate(self): return Unary(capi.UN_NEG, self)
    def cwiseSqrt(self): return Unary(capi.UN_SQRT, self)
    def cwiseRsqrt(self): return Unary(capi.UN_RSQRT, self)
    def cwiseCbrt(self): return Unary(capi.UN_CBRT, self)
    def cwiseExp(self): return Unary(capi.UN_EXP, self)
    def cwiseLog(self): return Unary(capi.UN_LOG, self)
    def cwiseLog1p(self): return Unary(capi.UN_LOG1P, self)
    def cwiseLog10(self): return Unary(capi.UN_LOG10, self)
    def cwiseSin(self): return Unary(capi.UN_SIN, self)
    def cwiseCos(self): return Unary(capi.UN_COS, self)
    def cwiseTan(self): return Unary(capi.UN_TAN, self)
    def cwiseTanh(self): return Unary(capi.UN_TANH, self)
    def cwiseFloor(self): return Unary(capi.UN_FLOOR, self)
    def cwiseCeil(self): return Unary(capi.UN_CEIL, self)
    def cwiseRound(self): return Unary(capi.UN_ROUND, self)
    def cwiseLogicalNot(self): return Unary(capi.UN_NOT, self)
    def cwiseBinaryNot(self): return Unary(capi.UN_BITNOT, self)
    def cast(self, dtype): return cast(self, dtype)


def _dt_of_py(v, like: Expr | None) -> int:
    if like is not None:
        return like.dtype
    if isinstance(v, (bool, np.bool_)):
        return capi.BOOL
    if isinstance(v, (int, np.integer)):
        return capi.I32
    return capi.F64 if isinstance(v, np.float64) else capi.F32


class LeafNode(Expr):
    def __init__(self, arr: DeviceArray, strides=None, shape=None, kind=capi.LEAF_STRIDED):
        self.arr = arr
        self.dtype = arr.dtype
        self.shape = shape or (arr.rows, arr.cols, arr.batches)
        self.strides = strides or arr.strides()
        self.kind = kind

    @property
    def T(self):
        """transpose view (cuMat TransposeOp in cwise context)"""
        sr, sc, sb = self.strides
        R, Cc, B = self.shape
        return LeafNode(self.arr, (sc, sr, sb), (Cc, R, B), self.kind)

    def diagonal(self):
        """ExtractDiagonalOp (cuMat/src/UnaryOps.h:458-487) as a strided view."""
        sr, sc, sb = self.strides
        R, Cc, B = self.shape
        return LeafNode(self.arr, (sr + sc, 0, sb), (min(R, Cc), 1, B))

    def asDiagonal(self):
        """AsDiagonalOp (cuMat/src/UnaryOps.h:361-437): vector -> diagonal matrix."""
        sr, sc, sb = self.strides
        R, Cc, B = self.shape
        n = R if Cc == 1 else Cc
        s = sr if Cc == 1 else sc
        return LeafNode(self.arr, (s, 0, sb), (n, n, B), capi.LEAF_AS_DIAGONAL)


class Scalar(Expr):
    """HostScalar: a 1x1x1 constant (cuMat/src/NullaryOps.h:105-108)."""

    def __init__(self, value, dtype):
        self.value = value
        self.dtype = dtype
        self.shape = (1, 1, 1)


class Binary(Expr):
    def __init__(self, fn, a: Expr, b: Expr):
        if a.dtype != b.dtype:
            raise TypeError(f"no implicit casting between dtypes {a.dtype} and {b.dtype} "
                            "(cuMat/src/BinaryOps.h:114-115); use .cast()")
        if fn in _INT_ONLY and a.dtype not in (capi.I32, capi.BOOL):
            raise TypeError("integer-only functor on a floating type")
        self.fn, self.a, self.b = fn, a, b
        self.dtype = capi.BOOL if fn in _BOOL_RESULT else a.dtype
        self.shape = _broadcast(a.shape, b.shape)


class Unary(Expr):
    def __init__(self, fn, a: Expr):
        self.fn, self.a = fn, a
        self.dtype = capi.BOOL if fn == capi.UN_NOT else a.dtype
        self.shape = a.shape


class Cast(Expr):
    def __init__(self, a: Expr, dtype: int):
        self.a, self.dtype, self.shape = a, dtype, a.shape


class Select(Expr):
    """cond ? a : b with a scalar or leaf b (cuMat cwiseSelect)."""

    def __init__(self, cond: Expr, a: Expr, b: Expr):
        if a.dtype != b.dtype:
            raise TypeError("select branches must have the same dtype")
        self.cond, self.a, self.b = cond, a, b
        self.dtype = a.dtype
        self.shape = _broadcast(_broadcast(cond.shape, a.shape), b.shape)


def _broadcast(s1, s2):
    out = []
    for x, y in zip(s1, s2):
        if x == y or y == 1:
            out.append(x)
        elif x == 1:
            out.append(y)
        else:
            raise capi.CumatArgError(capi.CM_E_DIM, f"dimension mismatch {s1} vs {s2} "
                                     "(cuMat/src/BinaryOps.h:117-130)")
    return tuple(out)


def as_expr(v, like: Expr | None = None) -> Expr:
    if isinstance(v, Expr):
        return v
    if isinstance(v, DeviceArray):
        return LeafNode(v)
    dt = _dt_of_py(v, like)
    return Scalar(v, dt)


def cast(e, dtype) -> Expr:
    e = as_expr(e)
    from .device import np_to_dtype
    dt = dtype if isinstance(dtype, int) else np_to_dtype(dtype)
    return e if dt == e.dtype else Cast(e, dt)


def select(cond, a, b) -> Expr:
    a = as_expr(a)
    return Select(as_expr(cond), a, as_expr(b, a))


def where_scalar(cond, a, b):
    return select(cond, a, b)


# ------------------------------------------------------------- lowering
class Program:
    def __init__(self):
        self.leaves = []   # (LeafNode)
        self.code = []     # Instr


def _is_simple(e):
    return isinstance(e, (LeafNode, Scalar))


def lower(e: Expr, dst_shape):
    """Lower an expression to (leaves, code) for cm_cwise_eval."""
    prog = Program()

    def leaf_index(node: LeafNode):
        for i, l in enumerate(prog.leaves):
            if l.arr.ptr == node.arr.ptr and l.strides == node.strides and l.kind == node.kind \
                    and l.shape == node.shape:
                return i
        prog.leaves.append(node)
        if len(prog.leaves) > capi.MAX_LEAVES:
            raise capi.CumatArgError(capi.CM_E_ARG, "expression has too many distinct leaves")
        return len(prog.leaves) - 1

    def src_of(node, ins):
        if isinstance(node, LeafNode):
            ins.src, ins.arg = capi.SRC_LEAF, leaf_index(node)
        else:
            ins.src = capi.SRC_IMM
            if node.dtype in (capi.F32, capi.F64):
                ins.fimm = float(node.value)
            else:
                ins.iimm = int(node.value)

    def emit(node, depth):
        if depth >= capi.MAX_STACK:
            raise capi.CumatArgError(capi.CM_E_ARG, "expression too deep for one kernel")
        ins = capi.Instr()
        if _is_simple(node):
            ins.opcode, ins.dtype = capi.OP_LOAD, node.dtype
            src_of(node, ins)
            prog.code.append(ins)
        elif isinstance(node, Unary):
            emit(node.a, depth)
            ins.opcode, ins.dtype, ins.fn = capi.OP_UNARY, node.a.dtype, node.fn
            prog.code.append(ins)
            if node.dtype != node.a.dtype:  # logical not yields bool
                c = capi.Instr()
                c.opcode, c.dtype, c.arg = capi.OP_CAST, node.dtype, capi.I32
                prog.code.append(c)
        elif isinstance(node, Cast):
            emit(node.a, depth)
            ins.opcode, ins.dtype, ins.arg = capi.OP_CAST, node.dtype, node.a.dtype
            prog.code.append(ins)
        elif isinstance(node, Select):
            emit(node.a, depth)
            push = capi.Instr()
            push.opcode, push.arg, push.dtype = capi.OP_PUSH, depth, node.a.dtype
            prog.code.append(push)
            emit(node.cond, depth + 1)
            ins.opcode, ins.dtype, ins.arg = capi.OP_SELECT, node.dtype, depth
            if isinstance(node.b, Scalar):
                src_of(node.b, ins)
            elif isinstance(node.b, LeafNode):
                ins.src = capi.SRC_LEAF
                ins.iimm = leaf_index(node.b)
            else:
                raise capi.CumatArgError(capi.CM_E_ARG, "select: else-branch must be a leaf or scalar")
            prog.code.append(ins)
        elif isinstance(node, Binary):
            a, b = node.a, node.b
            # axpy peephole: X + leaf*scalar -> FMA
            if (node.fn == capi.BIN_ADD and isinstance(b, Binary) and b.fn == capi.BIN_MUL
                    and isinstance(b.a, LeafNode) and isinstance(b.b, Scalar)
                    and b.a.dtype == node.dtype):
                emit(a, depth)
                ins.opcode, ins.dtype, ins.arg = capi.OP_FMA, node.dtype, leaf_index(b.a)
                if node.dtype in (capi.F32, capi.F64):
                    ins.fimm = float(b.b.value)
                else:
                    ins.iimm = int(b.b.value)
                prog.code.append(ins)
                return
            ins.opcode, ins.dtype, ins.fn = capi.OP_BINARY, a.dtype, node.fn
            if _is_simple(b):
                emit(a, depth)
                src_of(b, ins)
            elif _is_simple(a):
                emit(b, depth)
                src_of(a, ins)
                ins.flags = 1
            else:
                emit(a, depth)
                push = capi.Instr()
                push.opcode, push.arg, push.dtype = capi.OP_PUSH, depth, a.dtype
                prog.code.append(push)
                emit(b, depth + 1)
                ins.src, ins.arg, ins.flags = capi.SRC_STACK, depth, 1
            prog.code.append(ins)
            if node.dtype != a.dtype and node.fn not in _BOOL_RESULT:
                raise AssertionError("unexpected dtype change")
        else:
            raise TypeError(f"cannot lower {type(node)}")

    emit(e, 0)
    if len(prog.code) > capi.MAX_CODE:
        raise capi.CumatArgError(capi.CM_E_ARG, "expression program too long for one kernel")
    # leaf strides against the destination shape (broadcast axes -> 0)
    leaves = (capi.Leaf * max(1, len(prog.leaves)))()
    for i, l in enumerate(prog.leaves):
        sr, sc, sb = l.strides
        R, Cc, B = l.shape
        if R == 1 and dst_shape[0] != 1:
            sr = 0
        if Cc == 1 and dst_shape[1] != 1:
            sc = 0
        if B == 1 and dst_shape[2] != 1:
            sb = 0
        leaves[i].ptr, leaves[i].dtype, leaves[i].kind = l.arr.ptr, l.arr.dtype, l.kind
        leaves[i].sr, leaves[i].sc, leaves[i].sb = sr, sc, sb
    code = (capi.Instr * len(prog.code))(*prog.code)
    return leaves, len(prog.leaves), code, len(prog.code)


def assign(dst: DeviceArray, e, mode: int = capi.ASSIGN, ctx=None) -> DeviceArray:
    """dst (mode)= e in ONE kernel launch (Assignment<Dense,Cwise>::assign,
    cuMat/src/CwiseOp.h:130-158)."""
    e = as_expr(e)
    shape = (dst.rows, dst.cols, dst.batches)
    for a, b in zip(e.shape, shape):
        if a != b and a != 1:
            raise capi.CumatArgError(capi.CM_E_DIM, f"assignment of {e.shape} into {shape}")
    if e.dtype != dst.dtype:
        raise TypeError(f"no implicit casting: expression dtype {e.dtype} into {dst.dtype}")
    leaves, nl, code, nc = lower(e, shape)
    if any(leaves[i].sr == -1 for i in range(nl)):
        raise TypeError("a sparse matrix's stored values are read only at its own entries: assign into "
                        "the sparse matrix or use it through a sparse view, not in a dense assignment")
    sr, sc, sb = dst.strides()
    capi.call("cm_cwise_eval", ctx_handle(ctx if ctx is not None else dst.ctx), leaves, nl, code,
              nc, dst.ptr, dst.dtype, dst.rows, dst.cols, dst.batches, sr, sc, sb, mode)
    return dst


def evaluate(e, row_major=False, ctx=None) -> DeviceArray:
    """Evaluate an expression into a new DeviceArray (Matrix(const MatrixBase&))."""
    e = as_expr(e)
    R, Cc, B = e.shape
    from .device import dtype_to_np
    out = DeviceArray(R, Cc, B, dtype=dtype_to_np(e.dtype), row_major=row_major, ctx=ctx)
    return assign(out, e, ctx=ctx)
