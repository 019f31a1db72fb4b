"""ctypes binding of the C ABI declared in include/cumat_amd.h.

The shared library is built in-tree (cumat_amd/lib/libcumat_amd.so). There is
no fallback: if the library is missing, importing this module raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libcumat_amd.so")

# ---- enums (mirror include/cumat_amd.h) ----
CM_OK, CM_E_ARG, CM_E_DIM, CM_E_HIP, CM_E_NOMEM, CM_E_UNSUPPORTED, CM_E_STATE = range(7)
I32, F32, F64, BOOL, BF16 = 0, 1, 2, 3, 4
ASSIGN, ADD, SUB, MUL, DIV, MOD, AND, OR = range(8)
AXIS_ROW, AXIS_COL, AXIS_BATCH, AXIS_ALL = 1, 2, 4, 7
RED_SUM, RED_PROD, RED_MIN, RED_MAX, RED_ALL, RED_ANY, RED_BAND, RED_BOR = range(8)
(CNT_DEVICE_MEM_ALLOC, CNT_DEVICE_MEM_FREE, CNT_EVAL_ANY, CNT_EVAL_CWISE, CNT_EVAL_CWISE_SPARSE,
 CNT_EVAL_REDUCTION, CNT_EVAL_MATMUL_SPARSE, CNT_KERNEL_LAUNCH) = range(8)
LEAF_STRIDED, LEAF_AS_DIAGONAL = 0, 1
OP_LOAD, OP_PUSH, OP_BINARY, OP_UNARY, OP_CAST, OP_FMA, OP_SELECT = range(7)
SRC_NONE, SRC_LEAF, SRC_IMM, SRC_STACK = range(4)
(BIN_ADD, BIN_SUB, BIN_MUL, BIN_DIV, BIN_MOD, BIN_POW, BIN_MIN, BIN_MAX, BIN_AND, BIN_OR,
 BIN_XOR, BIN_LAND, BIN_LOR, BIN_EQ, BIN_NE, BIN_LT, BIN_GT, BIN_LE, BIN_GE) = range(19)
(UN_NEG, UN_ABS, UN_ABS2, UN_INVERSE, UN_INVERSE_CHECK, UN_SQRT, UN_RSQRT, UN_CBRT, UN_EXP,
 UN_LOG, UN_LOG1P, UN_LOG10, UN_SIN, UN_COS, UN_TAN, UN_ASIN, UN_ACOS, UN_ATAN, UN_SINH,
 UN_COSH, UN_TANH, UN_FLOOR, UN_CEIL, UN_ROUND, UN_NOT, UN_BITNOT, UN_SIGN) = range(27)
CG_AUTO, CG_MULTIKERNEL, CG_RESIDENT, CG_FUSED = 0, 1, 2, 3
MAX_LEAVES, MAX_CODE, MAX_STACK = 12, 48, 4


class Leaf(C.Structure):
    _fields_ = [("ptr", C.c_void_p), ("dtype", C.c_int32), ("kind", C.c_int32),
                ("sr", C.c_int64), ("sc", C.c_int64), ("sb", C.c_int64)]


class Instr(C.Structure):
    _fields_ = [("opcode", C.c_int32), ("dtype", C.c_int32), ("fn", C.c_int32),
                ("src", C.c_int32), ("arg", C.c_int32), ("flags", C.c_int32),
                ("fimm", C.c_double), ("iimm", C.c_int64)]


class CgStats(C.Structure):
    _fields_ = [("path_used", C.c_int32), ("launches", C.c_int32), ("seconds", C.c_double)]


class IrStats(C.Structure):
    _fields_ = [("outer_iterations", C.c_int32), ("inner_iterations", C.c_int32),
                ("final_rel_residual", C.c_double), ("seconds", C.c_double)]


class CumatError(RuntimeError):
    """A non-zero status from the C ABI (the reference throws C++ exceptions:
    std::invalid_argument for argument/dimension errors, cuda_error for device
    errors)."""

    def __init__(self, code, msg):
        super().__init__(f"cumat_amd error {code}: {msg}")
        self.code = code


class CumatArgError(CumatError, ValueError):
    pass


_VP, _I32, _I64, _F64, _SZ = C.c_void_p, C.c_int, C.c_int64, C.c_double, C.c_size_t
_PROTOS = {
    "cm_abi_version": ([], _I32),
    "cm_last_error": ([], C.c_char_p),
    "cm_device_count": ([C.POINTER(C.c_int)], _I32),
    "cm_device_info": ([_I32, C.c_char_p, _I32, C.POINTER(C.c_int), C.POINTER(C.c_longlong),
                        C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)], _I32),
    "cm_ctx_create": ([_I32, C.POINTER(_VP)], _I32),
    "cm_ctx_destroy": ([_VP], _I32),
    "cm_ctx_get_stream": ([_VP, C.POINTER(_VP)], _I32),
    "cm_ctx_set_stream": ([_VP, _VP], _I32),
    "cm_ctx_sync": ([_VP], _I32),
    "cm_malloc": ([_VP, _SZ, C.POINTER(_VP)], _I32),
    "cm_free": ([_VP, _VP], _I32),
    "cm_memcpy_h2d": ([_VP, _VP, _VP, _SZ], _I32),
    "cm_memcpy_d2h": ([_VP, _VP, _VP, _SZ], _I32),
    "cm_memcpy_d2d": ([_VP, _VP, _VP, _SZ], _I32),
    "cm_fill_bytes": ([_VP, _VP, _I32, _SZ], _I32),
    "cm_counters_get": ([_I32, C.POINTER(C.c_longlong)], _I32),
    "cm_counters_reset": ([], _I32),
    "cm_cwise_eval": ([_VP, C.POINTER(Leaf), _I32, C.POINTER(Instr), _I32, _VP, _I32, _I32, _I32,
                       _I32, _I64, _I64, _I64, _I32], _I32),
    "cm_lincomb_f32": ([_VP, _I32, C.POINTER(_VP), C.POINTER(C.c_float), _VP, _I64], _I32),
    "cm_jit_mode": ([_I32], _I32),
    "cm_reduce": ([_VP, _I32, _I32, _I32, _I32, _I32, _I32, _I32, _VP, _I32, _VP], _I32),
    "cm_reduce_expr": ([_VP, C.POINTER(Leaf), _I32, C.POINTER(Instr), _I32, _I32, _I32, _I32, _I32, _I32, _I32,
                        _I32, _VP], _I32),
    "cm_dot": ([_VP, _I32, _I64, _I32, _VP, _I32, _VP, _I32, _VP], _I32),
    "cm_squared_norm": ([_VP, _I32, _I64, _I32, _VP, _VP], _I32),
    "cm_dot_noalloc": ([_VP, _I32, _I64, _I32, _VP, _I32, _VP, _I32, _VP], _I32),
    "cm_csr_spmv": ([_VP, _I32, _I32, _I32, _I32, _I32, _I32, _VP, _VP, _VP, _I32, _VP, _I32,
                     _VP, _I32], _I32),
    "cm_ell_spmv": ([_VP, _I32, _I32, _I32, _I32, _I32, _VP, _VP, _I32, _VP, _I32, _VP, _I32],
                    _I32),
    "cm_csr_cwise_eval": ([_VP, C.POINTER(Leaf), _I32, C.POINTER(Instr), _I32, _I32, _I32, _I32,
                           _I32, _VP, _VP, _VP, _I32, _I32], _I32),
    "cm_csr_expr_spmv": ([_VP, C.POINTER(Leaf), _I32, C.POINTER(Instr), _I32, _I32, _I32, _I32, _I32,
                          _I32, _VP, _VP, _VP, _I32, _VP, _I32], _I32),
    "cm_ell_expr_spmv": ([_VP, C.POINTER(Leaf), _I32, C.POINTER(Instr), _I32, _I32, _I32, _I32, _I32,
                          _I32, _VP, _VP, _I32, _VP, _I32], _I32),
    "cm_csr_inverse_diagonal": ([_VP, _I32, _I32, _I32, _I32, _VP, _VP, _VP, _I32, _VP], _I32),
    "cm_bcsr_spmv": ([_VP, _I32, _I32, _I32, _I32, _I32, _VP, _VP, _VP, _I32, _VP, _I32, _VP, _I32], _I32),
    "cm_bcsr_inverse_diagonal": ([_VP, _I32, _I32, _I32, _I32, _VP, _VP, _VP, _I32, _VP], _I32),
    "cm_pcg_bcsr_solve": ([_VP, _I32, _I32, _I32, _I32, _VP, _VP, _VP, _I32, _VP, _VP, _VP, _I32, _F64, _I32,
                           C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(CgStats)], _I32),
    "cm_dense_gemv": ([_VP, _I32, _I32, _I32, _I32, _VP, _I32, _I32, _VP, _I32, _VP, _I32], _I32),
    "cm_dense_inverse_diagonal": ([_VP, _I32, _I32, _I32, _VP, _I32, _VP], _I32),
    "cm_pcg_dense_solve": ([_VP, _I32, _I32, _I32, _VP, _I32, _I32, _VP, _VP, _VP, _I32, _F64, _I32,
                            C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(CgStats)], _I32),
    "cm_pcg_solve": ([_VP, _I32, _I32, _I32, _I32, _VP, _VP, _VP, _VP, _VP, _VP, _I32, _F64, _I32,
                      _I32, C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(CgStats)], _I32),
    "cm_pcg_solve_ex": ([_VP, _I32, _I32, _I32, _I32, _VP, _VP, _VP, _I32, _VP, _VP, _VP, _I32, _F64,
                         _I32, _I32, C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(CgStats)], _I32),
    "cm_pcg_batched_solve": ([_VP, _I32, _VP, _I32, _I32, _VP, _VP, _VP, _VP, _VP, C.c_float, _I32,
                              _VP, _VP, _I32, _I32], _I32),
    "cm_pcg_batched_autotune": ([_VP, _I32, _VP, _I32, _I32, _VP, _VP, _VP, _VP, C.c_float, _I32,
                                 C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_float)], _I32),
    "cm_pcg_ir_solve": ([_VP, _I32, _I32, _VP, _VP, _VP, _I32, _VP, _VP, _F64, _F64, _I32, _I32,
                         C.POINTER(IrStats)], _I32),
    "cm_edge_assemble": ([_VP, _I32, _I32, _I32, _I32, _VP, _VP, _VP, _VP], _I32),
    "cm_edge_grad": ([_VP, _I32, _I32, _I32, _I32, _VP, _VP, _VP, _VP, _VP], _I32),
}

_lib = None


def exported_symbols():
    return list(_PROTOS)


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        L = C.CDLL(LIB_PATH, mode=C.RTLD_GLOBAL)
        for name, (args, res) in _PROTOS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
    return _lib


def last_error() -> str:
    return lib().cm_last_error().decode(errors="replace")


def check(status: int) -> None:
    if status != CM_OK:
        msg = last_error()
        if status in (CM_E_ARG, CM_E_DIM):
            raise CumatArgError(status, msg)
        raise CumatError(status, msg)


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args))
