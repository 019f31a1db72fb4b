// cwise_core.hpp -- the accumulator-program interpreter shared by the dense
// component-wise kernels (cwise.hip) and the sparse-expression products
// (spexpr.hip): program/leaf descriptors, value registers, the functors of
// cuMat/src/BinaryOps.h:230-243 and cuMat/src/UnaryOps.h:128-312, casts,
// and the leaf addressing of cuMat's Matrix layouts
// (cuMat/src/Matrix.h:587-627).
#pragma once
#include "cm_common.hpp"
#include "cwise_ops.hpp"

#include <cmath>
#include <type_traits>

namespace cm {

constexpr int CW_BS = 256;
// elements per thread and pass: 8 for 32-bit programs (the interpretation
// cost is shared by more elements: measured +8 % to +47 % on 2^26-element
// programs against 4), 4 for 64-bit ones (register banks twice as wide)
template <bool W> struct CwVec { static constexpr int v = W ? 4 : 8; };

struct CwProgram {
    cm_leaf leaves[CM_CWISE_MAX_LEAVES];
    cm_instr code[CM_CWISE_MAX_CODE];
    int nleaves;
    int ncode;
    unsigned leaf_direct;  // bit l: leaf l has the destination's strides
    unsigned leaf_scalar;  // bit l: leaf l is a single broadcast value
    void* dst;
    int dst_dtype;
    int mode;
    long long dst_sr, dst_sc, dst_sb;
    long long rows, cols, batches, n;
    int row_order;  // iterate elements in row-major logical order
    int dst_direct; // destination address == logical element index
};

// ----------------------------------------------------- value registers
// W: 64-bit lanes (any double in the program) or 32-bit lanes.
template <bool W> struct Reg;
template <> struct Reg<false> {
    uint32_t v[CwVec<false>::v];
};
template <> struct Reg<true> {
    uint64_t v[CwVec<true>::v];
};

template <bool W> __device__ __forceinline__ float getf(const Reg<W>& r, int k) { return __uint_as_float((uint32_t)r.v[k]); }
template <bool W> __device__ __forceinline__ int geti(const Reg<W>& r, int k) { return (int)(uint32_t)r.v[k]; }
template <bool W> __device__ __forceinline__ double getd(const Reg<W>& r, int k)
{
    if constexpr (W) return __longlong_as_double((long long)r.v[k]);
    else return 0.0;
}
template <bool W> __device__ __forceinline__ void setf(Reg<W>& r, int k, float x) { r.v[k] = __float_as_uint(x); }
template <bool W> __device__ __forceinline__ void seti(Reg<W>& r, int k, int x) { r.v[k] = (uint32_t)x; }
template <bool W> __device__ __forceinline__ void setd(Reg<W>& r, int k, double x)
{
    if constexpr (W) r.v[k] = (uint64_t)__double_as_longlong(x);
}

// Load one element of a leaf at element offset 'off' (dtype conversions are
// explicit CAST instructions; the raw bits are kept).
template <bool W>
__device__ __forceinline__ void load_raw(const cm_leaf& L, long long off, Reg<W>& r, int k)
{
    switch (L.dtype) {
    case CM_F32: r.v[k] = __float_as_uint(((const float*)L.ptr)[off]); break;
    case CM_I32: r.v[k] = (uint32_t)((const int*)L.ptr)[off]; break;
    case CM_BOOL: r.v[k] = (uint32_t)((const unsigned char*)L.ptr)[off]; break;
    case CM_F64:
        if constexpr (W) r.v[k] = (uint64_t)((const long long*)L.ptr)[off];
        break;
    default: r.v[k] = 0; break;
    }
}

// The per-element semantics live in cwise_ops.hpp (shared with the
// run-time generated kernels); the functions below dispatch the runtime
// dtype / functor codes of an instruction onto them, one wave-uniform switch
// per instruction and a loop over the register's elements inside each case.
#define CW_EACH(...) _Pragma("unroll") for (int k = 0; k < CwVec<W>::v; ++k) { __VA_ARGS__; } break

// Convert register contents between dtypes (CastingOp, cuMat/src/UnaryOps.h:248-312).
template <bool W, int FROM>
__device__ __forceinline__ void convert_from(Reg<W>& r, int to)
{
    switch (to) {
    case CM_F32: CW_EACH(r.v[k] = ops::convert1<FROM, CM_F32>(r.v[k]));
    case CM_F64: CW_EACH(r.v[k] = ops::convert1<FROM, CM_F64>(r.v[k]));
    case CM_I32: CW_EACH(r.v[k] = ops::convert1<FROM, CM_I32>(r.v[k]));
    case CM_BOOL: CW_EACH(r.v[k] = ops::convert1<FROM, CM_BOOL>(r.v[k]));
    default: break;
    }
}
template <bool W>
__device__ __forceinline__ void convert(Reg<W>& r, int from, int to)
{
    if (from == to) return;
    switch (from) {
    case CM_F32: convert_from<W, CM_F32>(r, to); break;
    case CM_F64: convert_from<W, CM_F64>(r, to); break;
    case CM_I32: convert_from<W, CM_I32>(r, to); break;
    case CM_BOOL: convert_from<W, CM_BOOL>(r, to); break;
    default: convert_from<W, -1>(r, to); break;  // unknown source: converts 0
    }
}

// ------------------------------------------------------- binary functors
// cuMat/src/BinaryOps.h:230-243 and the comparison functors
// (cuMat/src/BinaryOps.h:336-341); comparisons and logical ops yield bool.
template <bool W, int DT>
__device__ __forceinline__ void binary_dt(int fn, Reg<W>& a, const Reg<W>& x, const Reg<W>& y)
{
#define CW_BIN(FN) case FN: CW_EACH(a.v[k] = ops::binary1<DT, FN>(a.v[k], x.v[k], y.v[k]))
    switch (fn) {
    CW_BIN(CM_BIN_ADD); CW_BIN(CM_BIN_SUB); CW_BIN(CM_BIN_MUL); CW_BIN(CM_BIN_DIV);
    CW_BIN(CM_BIN_MOD); CW_BIN(CM_BIN_POW); CW_BIN(CM_BIN_MIN); CW_BIN(CM_BIN_MAX);
    CW_BIN(CM_BIN_AND); CW_BIN(CM_BIN_OR); CW_BIN(CM_BIN_XOR); CW_BIN(CM_BIN_LAND);
    CW_BIN(CM_BIN_LOR); CW_BIN(CM_BIN_EQ); CW_BIN(CM_BIN_NE); CW_BIN(CM_BIN_LT);
    CW_BIN(CM_BIN_GT); CW_BIN(CM_BIN_LE); CW_BIN(CM_BIN_GE);
    default: break;
    }
#undef CW_BIN
}
template <bool W>
__device__ __forceinline__ void binary(int fn, int dt, Reg<W>& a, const Reg<W>& b, bool rev)
{
    // operand order by element-wise selects (a select between two register
    // structs' addresses would force them into scratch memory)
    Reg<W> x, y;
#pragma unroll
    for (int k = 0; k < CwVec<W>::v; ++k) {
        x.v[k] = rev ? b.v[k] : a.v[k];
        y.v[k] = rev ? a.v[k] : b.v[k];
    }
    if (dt == CM_F32) binary_dt<W, CM_F32>(fn, a, x, y);
    else if (dt == CM_F64) binary_dt<W, CM_F64>(fn, a, x, y);
    else binary_dt<W, CM_I32>(fn, a, x, y);  // I32 and BOOL: 32-bit integer arithmetic
}

// -------------------------------------------------------- unary functors
// cuMat/src/UnaryOps.h:128-225 (device math is HIP's own).
template <bool W, int DT>
__device__ __forceinline__ void unary_dt(int fn, Reg<W>& a)
{
#define CW_UN(FN) case FN: CW_EACH(a.v[k] = ops::unary1<DT, FN>(a.v[k]))
    switch (fn) {
    CW_UN(CM_UN_NEG); CW_UN(CM_UN_ABS); CW_UN(CM_UN_ABS2); CW_UN(CM_UN_INVERSE);
    CW_UN(CM_UN_INVERSE_CHECK); CW_UN(CM_UN_SQRT); CW_UN(CM_UN_RSQRT); CW_UN(CM_UN_CBRT);
    CW_UN(CM_UN_EXP); CW_UN(CM_UN_LOG); CW_UN(CM_UN_LOG1P); CW_UN(CM_UN_LOG10);
    CW_UN(CM_UN_SIN); CW_UN(CM_UN_COS); CW_UN(CM_UN_TAN); CW_UN(CM_UN_ASIN);
    CW_UN(CM_UN_ACOS); CW_UN(CM_UN_ATAN); CW_UN(CM_UN_SINH); CW_UN(CM_UN_COSH);
    CW_UN(CM_UN_TANH); CW_UN(CM_UN_FLOOR); CW_UN(CM_UN_CEIL); CW_UN(CM_UN_ROUND);
    CW_UN(CM_UN_NOT); CW_UN(CM_UN_BITNOT); CW_UN(CM_UN_SIGN);
    default: break;
    }
#undef CW_UN
}
template <bool W>
__device__ __forceinline__ void unary(int fn, int dt, Reg<W>& a)
{
    if (dt == CM_F32) unary_dt<W, CM_F32>(fn, a);
    else if (dt == CM_F64) unary_dt<W, CM_F64>(fn, a);
    else unary_dt<W, CM_I32>(fn, a);
}
#undef CW_EACH

// ------------------------------------------------------- program runner
// Leaf values and the stack live in register vectors that are indexed with a
// wave-uniform index; the backend lowers such an index to GPR-indexing mode
// (s_set_gpr_idx_on + v_mov), one move per dword and no scratch memory.
template <bool W, int N> struct Bank {
    using E = typename std::conditional<W, uint64_t, uint32_t>::type;
    typedef E V __attribute__((ext_vector_type(N)));
    V k[CwVec<W>::v];
};

template <bool W, int NL>
__device__ __forceinline__ void fetch_leaf(const Bank<W, NL>& lv, int l, Reg<W>& out)
{
    l = __builtin_amdgcn_readfirstlane(l);
#pragma unroll
    for (int k = 0; k < CwVec<W>::v; ++k) out.v[k] = lv.k[k][l];
}

template <bool W>
__device__ __forceinline__ void imm(const cm_instr& I, Reg<W>& out)
{
#pragma unroll
    for (int k = 0; k < CwVec<W>::v; ++k) {
        if (I.dtype == CM_F32) setf(out, k, (float)I.fimm);
        else if (I.dtype == CM_F64) setd(out, k, I.fimm);
        else seti(out, k, (int)I.iimm);
    }
}

template <bool W>
using Stack = Bank<W, CM_CWISE_MAX_STACK>;

template <bool W>
__device__ __forceinline__ void stack_get(const Stack<W>& st, int s, Reg<W>& out)
{
    s = __builtin_amdgcn_readfirstlane(s);
#pragma unroll
    for (int k = 0; k < CwVec<W>::v; ++k) out.v[k] = st.k[k][s];
}
template <bool W>
__device__ __forceinline__ void stack_put(Stack<W>& st, int s, const Reg<W>& v)
{
    s = __builtin_amdgcn_readfirstlane(s);
#pragma unroll
    for (int k = 0; k < CwVec<W>::v; ++k) st.k[k][s] = v.v[k];
}

template <bool W, int NL>
__device__ __forceinline__ void fetch_src(const CwProgram& P, const cm_instr& I, const Bank<W, NL>& lv,
                                          const Stack<W>& st, Reg<W>& t)
{
    if (I.src == CM_SRC_LEAF) {
        fetch_leaf<W, NL>(lv, I.arg, t);
        const int ldt = P.leaves[I.arg].dtype;
        if (ldt != I.dtype) convert<W>(t, ldt, I.dtype);
    } else if (I.src == CM_SRC_IMM) {
        imm<W>(I, t);
    } else if (I.src == CM_SRC_STACK) {
        stack_get<W>(st, I.arg, t);
    }
}

template <bool W, int NL>
__device__ __forceinline__ int run_program(const CwProgram& P, const Bank<W, NL>& lv, Reg<W>& acc)
{
    Stack<W> st;
    int dt = CM_F32;
    for (int pc = 0; pc < P.ncode; ++pc) {
        const cm_instr I = P.code[pc];
        Reg<W> t;
        switch (I.opcode) {
        case CM_OP_LOAD:
            fetch_src<W, NL>(P, I, lv, st, acc);
            break;
        case CM_OP_PUSH:
            stack_put<W>(st, I.arg, acc);
            break;
        case CM_OP_BINARY:
            fetch_src<W, NL>(P, I, lv, st, t);
            binary<W>(I.fn, I.dtype, acc, t, (I.flags & 1) != 0);
            break;
        case CM_OP_UNARY:
            unary<W>(I.fn, I.dtype, acc);
            break;
        case CM_OP_CAST:
            convert<W>(acc, I.arg, I.dtype);
            break;
        case CM_OP_FMA: {
            // acc = acc + leaf * imm  (one linear-combination term)
            fetch_leaf<W, NL>(lv, I.arg, t);
            if (I.dtype == CM_F32) {
#pragma unroll
                for (int k = 0; k < CwVec<W>::v; ++k) acc.v[k] = ops::fma1<CM_F32>(acc.v[k], t.v[k], I.fimm, I.iimm);
            } else if (I.dtype == CM_F64) {
#pragma unroll
                for (int k = 0; k < CwVec<W>::v; ++k) acc.v[k] = ops::fma1<CM_F64>(acc.v[k], t.v[k], I.fimm, I.iimm);
            } else {
#pragma unroll
                for (int k = 0; k < CwVec<W>::v; ++k) acc.v[k] = ops::fma1<CM_I32>(acc.v[k], t.v[k], I.fimm, I.iimm);
            }
            break;
        }
        case CM_OP_SELECT: {
            // acc = acc ? stack[arg] : src   (cwiseSelect)
            Reg<W> a, b;
            stack_get<W>(st, I.arg, a);
            I.src == CM_SRC_IMM ? imm<W>(I, b) : fetch_leaf<W, NL>(lv, (int)I.iimm, b);
#pragma unroll
            for (int k = 0; k < CwVec<W>::v; ++k) acc.v[k] = ((uint32_t)acc.v[k] != 0) ? a.v[k] : b.v[k];
            break;
        }
        default:
            break;
        }
        dt = I.dtype;
    }
    return dt;
}

// Store with assignment mode (CwiseAssignmentHandler, cuMat/src/CwiseOp.h:15-62).
template <bool W>
__device__ __forceinline__ void store_elem(const CwProgram& P, long long off, const Reg<W>& acc, int k)
{
    void* d = P.dst;
    switch (P.dst_dtype) {
    case CM_F32: {
        float* p = (float*)d + off;
        float v = getf(acc, k);
        *p = (P.mode == CM_ASSIGN) ? v : assign_mode<float>(P.mode, *p, v);
        break;
    }
    case CM_F64: {
        if constexpr (W) {
            double* p = (double*)d + off;
            double v = getd(acc, k);
            *p = (P.mode == CM_ASSIGN) ? v : assign_mode<double>(P.mode, *p, v);
        }
        break;
    }
    case CM_I32: {
        int* p = (int*)d + off;
        int v = geti(acc, k);
        *p = (P.mode == CM_ASSIGN) ? v : assign_mode<int>(P.mode, *p, v);
        break;
    }
    case CM_BOOL: {
        unsigned char* p = (unsigned char*)d + off;
        int v = geti(acc, k) != 0;
        if (P.mode == CM_AND) v = v && *p;
        else if (P.mode == CM_OR) v = v || *p;
        *p = (unsigned char)v;
        break;
    }
    default: break;
    }
}

struct Coords {
    long long r, c, b;
};

__device__ __forceinline__ Coords decompose(const CwProgram& P, long long e)
{
    Coords q;
    const long long rc = P.rows * P.cols;
    q.b = e / rc;
    long long w = e - q.b * rc;
    if (P.row_order) { q.r = w / P.cols; q.c = w - q.r * P.cols; }
    else { q.c = w / P.rows; q.r = w - q.c * P.rows; }
    return q;
}

__device__ __forceinline__ long long leaf_offset(const cm_leaf& L, const Coords& q, bool& zero)
{
    zero = false;
    if (L.kind == CM_LEAF_AS_DIAGONAL) {
        zero = (q.r != q.c);
        return q.r * L.sr + q.b * L.sb;
    }
    return q.r * L.sr + q.c * L.sc + q.b * L.sb;
}

// --------------------------------------------------------- host side
inline int build_program(const cm_leaf* leaves, int nleaves, const cm_instr* code, int ncode, CwProgram& P,
                         bool& wide)
{
    CM_REQUIRE(nleaves >= 0 && nleaves <= CM_CWISE_MAX_LEAVES, CM_E_ARG, "cwise: %d leaves (max %d)", nleaves,
               CM_CWISE_MAX_LEAVES);
    CM_REQUIRE(ncode >= 1 && ncode <= CM_CWISE_MAX_CODE, CM_E_ARG, "cwise: %d instructions (max %d)", ncode,
               CM_CWISE_MAX_CODE);
    memset(&P, 0, sizeof(P));
    wide = false;
    for (int l = 0; l < nleaves; ++l) {
        P.leaves[l] = leaves[l];
        CM_REQUIRE(leaves[l].ptr != nullptr, CM_E_ARG, "cwise: leaf %d has a NULL pointer", l);
        CM_REQUIRE(dtype_size(leaves[l].dtype) > 0 && leaves[l].dtype != CM_BF16, CM_E_UNSUPPORTED,
                   "cwise: leaf %d dtype %d", l, leaves[l].dtype);
        if (leaves[l].dtype == CM_F64) wide = true;
    }
    int depth_ok = 1;
    for (int i = 0; i < ncode; ++i) {
        const cm_instr& I = code[i];
        P.code[i] = I;
        CM_REQUIRE(I.opcode >= CM_OP_LOAD && I.opcode <= CM_OP_SELECT, CM_E_ARG, "cwise: bad opcode %d", I.opcode);
        if (I.dtype == CM_F64) wide = true;
        if (I.opcode == CM_OP_CAST && I.arg == CM_F64) wide = true;
        if ((I.opcode == CM_OP_LOAD || I.opcode == CM_OP_BINARY) && I.src == CM_SRC_LEAF)
            CM_REQUIRE(I.arg >= 0 && I.arg < nleaves, CM_E_ARG, "cwise: leaf index %d out of range", I.arg);
        if (I.opcode == CM_OP_FMA)
            CM_REQUIRE(I.arg >= 0 && I.arg < nleaves, CM_E_ARG, "cwise: leaf index %d out of range", I.arg);
        if (I.opcode == CM_OP_PUSH || (I.src == CM_SRC_STACK && I.opcode != CM_OP_CAST) || I.opcode == CM_OP_SELECT)
            depth_ok &= (I.arg >= 0 && I.arg < CM_CWISE_MAX_STACK);
    }
    CM_REQUIRE(depth_ok, CM_E_ARG, "cwise: stack slot out of range (max %d)", CM_CWISE_MAX_STACK);
    P.nleaves = nleaves;
    P.ncode = ncode;
    return CM_OK;
}

}  // namespace cm
