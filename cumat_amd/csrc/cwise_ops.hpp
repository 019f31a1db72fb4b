// cwise_ops.hpp -- per-element semantics of the accumulator machine:
// dtype conversion (CastingOp, cuMat/src/UnaryOps.h:248-312), the binary
// functors (cuMat/src/BinaryOps.h:230-243, comparisons :336-341) and the
// unary functors (cuMat/src/UnaryOps.h:128-225) on one raw value register.
//
// A register R is uint32_t (programs without 64-bit values) or uint64_t;
// 32-bit values sit in its low half, bool is 0/1 in an int. Every function
// takes its dtype and functor as template arguments, so each instantiation
// is one expression.
//
// This header is self-contained on purpose (no #include): the library
// compiles it into the program interpreter, and the run-time compiler gets
// the same text ahead of each generated kernel (jit.hip), so both paths
// evaluate every instruction with the same expressions. The includer
// provides uint32_t / uint64_t, the CM_* dtype, functor and assignment-mode
// constants, and the device math functions.
#pragma once

namespace cm {
namespace ops {

template <typename R> __device__ __forceinline__ float getf(R r) { return __uint_as_float((uint32_t)r); }
template <typename R> __device__ __forceinline__ int geti(R r) { return (int)(uint32_t)r; }
template <typename R> __device__ __forceinline__ double getd(R r)
{
    if constexpr (sizeof(R) == 8) return __longlong_as_double((long long)r);
    else return 0.0;
}
template <typename R> __device__ __forceinline__ R mkf(float x) { return (R)__float_as_uint(x); }
template <typename R> __device__ __forceinline__ R mki(int x) { return (R)(uint32_t)x; }
template <typename R> __device__ __forceinline__ R mkd(double x)
{
    if constexpr (sizeof(R) == 8) return (R)__double_as_longlong(x);
    else return (R)0;
}

// Conversion between dtypes; an unknown target leaves the register as is.
template <int FROM, int TO, typename R>
__device__ __forceinline__ R convert1(R r)
{
    if constexpr (FROM == TO) {
        return r;
    } else {
        double d = 0.0;
        long long i = 0;
        bool isf = false;
        if constexpr (FROM == CM_F32) { d = getf(r); isf = true; }
        else if constexpr (FROM == CM_F64) { d = getd(r); isf = true; }
        else if constexpr (FROM == CM_I32) { i = geti(r); }
        else if constexpr (FROM == CM_BOOL) { i = geti(r) != 0; }
        if constexpr (TO == CM_F32) {
            if constexpr (FROM == CM_F64) return mkf<R>((float)d);
            else return mkf<R>(isf ? (float)d : (float)i);
        } else if constexpr (TO == CM_F64) {
            return mkd<R>(isf ? d : (double)i);
        } else if constexpr (TO == CM_I32) {
            return mki<R>(isf ? (int)d : (int)i);
        } else if constexpr (TO == CM_BOOL) {
            return mki<R>(isf ? (d != 0.0) : (i != 0));
        } else {
            return r;
        }
    }
}

// Binary functor on ordered operands (x fn y); 'a' is returned for a functor
// the dtype does not define (the accumulator is left unchanged).
template <int DT, int FN, typename R>
__device__ __forceinline__ R binary1(R a, R x, R y)
{
    if constexpr (DT == CM_F32) {
        const float p = getf(x), q = getf(y);
        if constexpr (FN == CM_BIN_ADD) return mkf<R>(p + q);
        else if constexpr (FN == CM_BIN_SUB) return mkf<R>(p - q);
        else if constexpr (FN == CM_BIN_MUL) return mkf<R>(p * q);
        else if constexpr (FN == CM_BIN_DIV) return mkf<R>(p / q);
        else if constexpr (FN == CM_BIN_POW) return mkf<R>(powf(p, q));
        else if constexpr (FN == CM_BIN_MIN) return mkf<R>(p < q ? p : q);
        else if constexpr (FN == CM_BIN_MAX) return mkf<R>(p > q ? p : q);
        else if constexpr (FN == CM_BIN_EQ) return mki<R>(p == q);
        else if constexpr (FN == CM_BIN_NE) return mki<R>(p != q);
        else if constexpr (FN == CM_BIN_LT) return mki<R>(p < q);
        else if constexpr (FN == CM_BIN_GT) return mki<R>(p > q);
        else if constexpr (FN == CM_BIN_LE) return mki<R>(p <= q);
        else if constexpr (FN == CM_BIN_GE) return mki<R>(p >= q);
        else return a;
    } else if constexpr (DT == CM_F64) {
        if constexpr (sizeof(R) == 8) {
            const double p = getd(x), q = getd(y);
            if constexpr (FN == CM_BIN_ADD) return mkd<R>(p + q);
            else if constexpr (FN == CM_BIN_SUB) return mkd<R>(p - q);
            else if constexpr (FN == CM_BIN_MUL) return mkd<R>(p * q);
            else if constexpr (FN == CM_BIN_DIV) return mkd<R>(p / q);
            else if constexpr (FN == CM_BIN_POW) return mkd<R>(pow(p, q));
            else if constexpr (FN == CM_BIN_MIN) return mkd<R>(p < q ? p : q);
            else if constexpr (FN == CM_BIN_MAX) return mkd<R>(p > q ? p : q);
            else if constexpr (FN == CM_BIN_EQ) return (R)(uint32_t)(p == q);
            else if constexpr (FN == CM_BIN_NE) return (R)(uint32_t)(p != q);
            else if constexpr (FN == CM_BIN_LT) return (R)(uint32_t)(p < q);
            else if constexpr (FN == CM_BIN_GT) return (R)(uint32_t)(p > q);
            else if constexpr (FN == CM_BIN_LE) return (R)(uint32_t)(p <= q);
            else if constexpr (FN == CM_BIN_GE) return (R)(uint32_t)(p >= q);
            else return a;
        } else {
            return a;
        }
    } else {
        // I32 and BOOL: 32-bit integer arithmetic with wrap-around; division
        // and remainder by zero give 0
        const uint32_t u = (uint32_t)x, v = (uint32_t)y;
        const int p = geti(x), q = geti(y);
        if constexpr (FN == CM_BIN_ADD) return (R)(u + v);
        else if constexpr (FN == CM_BIN_SUB) return (R)(u - v);
        else if constexpr (FN == CM_BIN_MUL) return (R)(u * v);
        else if constexpr (FN == CM_BIN_DIV) return mki<R>(q != 0 ? p / q : 0);
        else if constexpr (FN == CM_BIN_MOD) return mki<R>(q != 0 ? p % q : 0);
        else if constexpr (FN == CM_BIN_MIN) return mki<R>(p < q ? p : q);
        else if constexpr (FN == CM_BIN_MAX) return mki<R>(p > q ? p : q);
        else if constexpr (FN == CM_BIN_AND) return (R)(u & v);
        else if constexpr (FN == CM_BIN_OR) return (R)(u | v);
        else if constexpr (FN == CM_BIN_XOR) return (R)(u ^ v);
        else if constexpr (FN == CM_BIN_LAND) return (R)(uint32_t)(p != 0 && q != 0);
        else if constexpr (FN == CM_BIN_LOR) return (R)(uint32_t)(p != 0 || q != 0);
        else if constexpr (FN == CM_BIN_EQ) return (R)(uint32_t)(p == q);
        else if constexpr (FN == CM_BIN_NE) return (R)(uint32_t)(p != q);
        else if constexpr (FN == CM_BIN_LT) return (R)(uint32_t)(p < q);
        else if constexpr (FN == CM_BIN_GT) return (R)(uint32_t)(p > q);
        else if constexpr (FN == CM_BIN_LE) return (R)(uint32_t)(p <= q);
        else if constexpr (FN == CM_BIN_GE) return (R)(uint32_t)(p >= q);
        else return a;
    }
}

// Unary functor (device math is HIP's own); 'a' is returned unchanged for a
// functor the dtype does not define.
template <int DT, int FN, typename R>
__device__ __forceinline__ R unary1(R a)
{
    if constexpr (DT == CM_F32) {
        const float p = getf(a);
        if constexpr (FN == CM_UN_NEG) return mkf<R>(-p);
        else if constexpr (FN == CM_UN_ABS) return mkf<R>(fabsf(p));
        else if constexpr (FN == CM_UN_ABS2) return mkf<R>(p * p);
        else if constexpr (FN == CM_UN_INVERSE) return mkf<R>(1.0f / p);
        else if constexpr (FN == CM_UN_INVERSE_CHECK) return mkf<R>(p == 0.0f ? 1.0f : 1.0f / p);
        else if constexpr (FN == CM_UN_SQRT) return mkf<R>(sqrtf(p));
        else if constexpr (FN == CM_UN_RSQRT) return mkf<R>(1.0f / sqrtf(p));
        else if constexpr (FN == CM_UN_CBRT) return mkf<R>(cbrtf(p));
        else if constexpr (FN == CM_UN_EXP) return mkf<R>(expf(p));
        else if constexpr (FN == CM_UN_LOG) return mkf<R>(logf(p));
        else if constexpr (FN == CM_UN_LOG1P) return mkf<R>(log1pf(p));
        else if constexpr (FN == CM_UN_LOG10) return mkf<R>(log10f(p));
        else if constexpr (FN == CM_UN_SIN) return mkf<R>(sinf(p));
        else if constexpr (FN == CM_UN_COS) return mkf<R>(cosf(p));
        else if constexpr (FN == CM_UN_TAN) return mkf<R>(tanf(p));
        else if constexpr (FN == CM_UN_ASIN) return mkf<R>(asinf(p));
        else if constexpr (FN == CM_UN_ACOS) return mkf<R>(acosf(p));
        else if constexpr (FN == CM_UN_ATAN) return mkf<R>(atanf(p));
        else if constexpr (FN == CM_UN_SINH) return mkf<R>(sinhf(p));
        else if constexpr (FN == CM_UN_COSH) return mkf<R>(coshf(p));
        else if constexpr (FN == CM_UN_TANH) return mkf<R>(tanhf(p));
        else if constexpr (FN == CM_UN_FLOOR) return mkf<R>(floorf(p));
        else if constexpr (FN == CM_UN_CEIL) return mkf<R>(ceilf(p));
        else if constexpr (FN == CM_UN_ROUND) return mkf<R>(roundf(p));
        else if constexpr (FN == CM_UN_SIGN) return mkf<R>((float)((p > 0.f) - (p < 0.f)));
        else if constexpr (FN == CM_UN_NOT) return mki<R>(p == 0.0f);
        else return a;
    } else if constexpr (DT == CM_F64) {
        if constexpr (sizeof(R) == 8) {
            const double p = getd(a);
            if constexpr (FN == CM_UN_NEG) return mkd<R>(-p);
            else if constexpr (FN == CM_UN_ABS) return mkd<R>(fabs(p));
            else if constexpr (FN == CM_UN_ABS2) return mkd<R>(p * p);
            else if constexpr (FN == CM_UN_INVERSE) return mkd<R>(1.0 / p);
            else if constexpr (FN == CM_UN_INVERSE_CHECK) return mkd<R>(p == 0.0 ? 1.0 : 1.0 / p);
            else if constexpr (FN == CM_UN_SQRT) return mkd<R>(sqrt(p));
            else if constexpr (FN == CM_UN_RSQRT) return mkd<R>(1.0 / sqrt(p));
            else if constexpr (FN == CM_UN_CBRT) return mkd<R>(cbrt(p));
            else if constexpr (FN == CM_UN_EXP) return mkd<R>(exp(p));
            else if constexpr (FN == CM_UN_LOG) return mkd<R>(log(p));
            else if constexpr (FN == CM_UN_LOG1P) return mkd<R>(log1p(p));
            else if constexpr (FN == CM_UN_LOG10) return mkd<R>(log10(p));
            else if constexpr (FN == CM_UN_SIN) return mkd<R>(sin(p));
            else if constexpr (FN == CM_UN_COS) return mkd<R>(cos(p));
            else if constexpr (FN == CM_UN_TAN) return mkd<R>(tan(p));
            else if constexpr (FN == CM_UN_ASIN) return mkd<R>(asin(p));
            else if constexpr (FN == CM_UN_ACOS) return mkd<R>(acos(p));
            else if constexpr (FN == CM_UN_ATAN) return mkd<R>(atan(p));
            else if constexpr (FN == CM_UN_SINH) return mkd<R>(sinh(p));
            else if constexpr (FN == CM_UN_COSH) return mkd<R>(cosh(p));
            else if constexpr (FN == CM_UN_TANH) return mkd<R>(tanh(p));
            else if constexpr (FN == CM_UN_FLOOR) return mkd<R>(floor(p));
            else if constexpr (FN == CM_UN_CEIL) return mkd<R>(ceil(p));
            else if constexpr (FN == CM_UN_ROUND) return mkd<R>(round(p));
            else if constexpr (FN == CM_UN_SIGN) return mkd<R>((double)((p > 0.0) - (p < 0.0)));
            else if constexpr (FN == CM_UN_NOT) return (R)(uint32_t)(p == 0.0);
            else return a;
        } else {
            return a;
        }
    } else {
        const int p = geti(a);
        if constexpr (FN == CM_UN_NEG) return (R)(uint32_t)(0u - (uint32_t)a);
        else if constexpr (FN == CM_UN_ABS) return mki<R>(p < 0 ? (int)(0u - (uint32_t)a) : p);
        else if constexpr (FN == CM_UN_ABS2) return (R)((uint32_t)a * (uint32_t)a);
        else if constexpr (FN == CM_UN_INVERSE) return mki<R>(p != 0 ? 1 / p : 0);
        else if constexpr (FN == CM_UN_INVERSE_CHECK) return mki<R>(p == 0 ? 1 : 1 / p);
        else if constexpr (FN == CM_UN_NOT) return (R)(uint32_t)(p == 0);
        else if constexpr (FN == CM_UN_BITNOT) return (R)(~(uint32_t)a);
        else if constexpr (FN == CM_UN_SIGN) return mki<R>((p > 0) - (p < 0));
        else return a;
    }
}

// One linear-combination term acc + t * c (CM_OP_FMA), contracted to a
// fused multiply-add for floating point as the library's compiler does.
template <int DT, typename R>
__device__ __forceinline__ R fma1(R acc, R t, double fimm, long long iimm)
{
    if constexpr (DT == CM_F32) {
        return mkf<R>(__builtin_fmaf(getf(t), (float)fimm, getf(acc)));
    } else if constexpr (DT == CM_F64) {
        return mkd<R>(__builtin_fma(getd(t), fimm, getd(acc)));
    } else {
        return (R)((uint32_t)acc + (uint32_t)t * (uint32_t)iimm);
    }
}

// Assignment mode applied to the destination's old value
// (CwiseAssignmentHandler, cuMat/src/CwiseOp.h:15-62); int32 arithmetic wraps.
template <typename T>
__device__ __forceinline__ T assign1(int mode, T old, T v)
{
    switch (mode) {
    case CM_ADD: return old + v;
    case CM_SUB: return old - v;
    case CM_MUL: return old * v;
    case CM_DIV: return old / v;
    default: return v;
    }
}
template <>
__device__ __forceinline__ int assign1<int>(int mode, int old, int v)
{
    switch (mode) {
    case CM_ADD: return (int)((unsigned)old + (unsigned)v);
    case CM_SUB: return (int)((unsigned)old - (unsigned)v);
    case CM_MUL: return (int)((unsigned)old * (unsigned)v);
    case CM_DIV: return old / v;
    case CM_MOD: return old % v;
    case CM_AND: return old & v;
    case CM_OR: return old | v;
    default: return v;
    }
}

}  // namespace ops
}  // namespace cm
