#ifndef CUMAT_AMD_REDUCTIONOPS_H
#define CUMAT_AMD_REDUCTIONOPS_H

// Reductions over any combination of the row / column / batch axes, dot,
// squaredNorm and norm (cuMat/src/ReductionOpsPlugin.inl:124-450 with the
// functors and initial values of cuMat/src/ReductionOps.h:859-968). The
// result keeps the non-reduced dimensions and has size 1 on the reduced ones.
// Unlike cuMat the reduction is evaluated when it is called (one cm_reduce
// launch; an input that is an expression is evaluated first), and the result
// is a Matrix, which every cuMat use of a reduction accepts.

#include "Matrix.h"

CUMAT_NAMESPACE_BEGIN

namespace internal {

// the input of a reduction as stored data: matrices are used as they are,
// expressions are evaluated once
template <typename S, int R, int C, int B, int F>
Matrix<S, R, C, B, F> materialize(const Matrix<S, R, C, B, F>& m) { return m; }
template <typename S, int F>
Matrix<S, Dynamic, Dynamic, Dynamic, F> materialize(const DenseRef<S, F>& m)
{
    return Matrix<S, Dynamic, Dynamic, Dynamic, F>(m.dataPointer(), m.rows(), m.cols(), m.batches());
}
template <typename E>
auto materialize(const MatrixBase<E>& e) { return e.eval(); }

template <typename M>
struct MatrixTraits;
template <typename S, int R, int C, int B, int F>
struct MatrixTraits<Matrix<S, R, C, B, F>> {
    enum { Rows = R, Columns = C, Batches = B, Flags = F };
};

template <int _Axis, typename E>
auto reduceStatic(const MatrixBase<E>& expr, int op)
{
    static_assert(_Axis >= 0 && _Axis <= 7, "axis must be a combination of Axis::Row, Axis::Column, Axis::Batch");
    auto in = materialize(expr.derived());
    typedef decltype(in) In;
    typedef typename In::Scalar S;
    typedef MatrixTraits<In> T;
    typedef Matrix<S, (_Axis & Axis::Row) ? 1 : (int)T::Rows, (_Axis & Axis::Column) ? 1 : (int)T::Columns,
                   (_Axis & Axis::Batch) ? 1 : (int)T::Batches, T::Flags>
        Out;
    Out out((_Axis & Axis::Row) ? 1 : in.rows(), (_Axis & Axis::Column) ? 1 : in.cols(),
            (_Axis & Axis::Batch) ? 1 : in.batches());
    CUMAT_AMD_CALL(cm_reduce(Context::current().handle(), DType<S>::value, op, _Axis, (int)in.rows(), (int)in.cols(),
                             (int)in.batches(), (T::Flags & RowMajor) ? 1 : 0, in.data(),
                             (T::Flags & RowMajor) ? 1 : 0, out.data()));
    return out;
}

template <typename E>
auto reduceDynamic(const MatrixBase<E>& expr, int axis, int op)
{
    CUMAT_ASSERT_ARGUMENT(axis >= 0 && axis <= 7);
    auto in = materialize(expr.derived());
    typedef decltype(in) In;
    typedef typename In::Scalar S;
    typedef MatrixTraits<In> T;
    typedef Matrix<S, Dynamic, Dynamic, Dynamic, T::Flags> Out;
    Out out((axis & Axis::Row) ? 1 : in.rows(), (axis & Axis::Column) ? 1 : in.cols(),
            (axis & Axis::Batch) ? 1 : in.batches());
    CUMAT_AMD_CALL(cm_reduce(Context::current().handle(), DType<S>::value, op, axis, (int)in.rows(), (int)in.cols(),
                             (int)in.batches(), (T::Flags & RowMajor) ? 1 : 0, in.data(),
                             (T::Flags & RowMajor) ? 1 : 0, out.data()));
    return out;
}

}  // namespace internal

#define CUMAT_AMD_REDUCTION(name, op) \
    template <typename _Derived> \
    template <int _Axis> \
    auto MatrixBase<_Derived>::name() const \
    { \
        return internal::reduceStatic<_Axis>(*this, op); \
    }
CUMAT_AMD_REDUCTION(sum, CM_RED_SUM)
CUMAT_AMD_REDUCTION(prod, CM_RED_PROD)
CUMAT_AMD_REDUCTION(minCoeff, CM_RED_MIN)
CUMAT_AMD_REDUCTION(maxCoeff, CM_RED_MAX)
CUMAT_AMD_REDUCTION(all, CM_RED_ALL)
CUMAT_AMD_REDUCTION(any, CM_RED_ANY)
CUMAT_AMD_REDUCTION(bitwiseAnd, CM_RED_BAND)
CUMAT_AMD_REDUCTION(bitwiseOr, CM_RED_BOR)
#undef CUMAT_AMD_REDUCTION

// without an axis: reduce everything (Axis::All)
template <typename _Derived> auto MatrixBase<_Derived>::sum() const { return internal::reduceStatic<Axis::All>(*this, CM_RED_SUM); }
template <typename _Derived> auto MatrixBase<_Derived>::prod() const { return internal::reduceStatic<Axis::All>(*this, CM_RED_PROD); }
template <typename _Derived> auto MatrixBase<_Derived>::minCoeff() const { return internal::reduceStatic<Axis::All>(*this, CM_RED_MIN); }
template <typename _Derived> auto MatrixBase<_Derived>::maxCoeff() const { return internal::reduceStatic<Axis::All>(*this, CM_RED_MAX); }
template <typename _Derived> auto MatrixBase<_Derived>::all() const { return internal::reduceStatic<Axis::All>(*this, CM_RED_ALL); }
template <typename _Derived> auto MatrixBase<_Derived>::any() const { return internal::reduceStatic<Axis::All>(*this, CM_RED_ANY); }

// run-time axis (ReductionOp_DynamicSwitched, cuMat/src/ReductionOps.h:776-857)
template <typename _Derived> auto MatrixBase<_Derived>::sum(int axis) const { return internal::reduceDynamic(*this, axis, CM_RED_SUM); }
template <typename _Derived> auto MatrixBase<_Derived>::prod(int axis) const { return internal::reduceDynamic(*this, axis, CM_RED_PROD); }
template <typename _Derived> auto MatrixBase<_Derived>::minCoeff(int axis) const { return internal::reduceDynamic(*this, axis, CM_RED_MIN); }
template <typename _Derived> auto MatrixBase<_Derived>::maxCoeff(int axis) const { return internal::reduceDynamic(*this, axis, CM_RED_MAX); }

// dot product per batch of two vectors (cuMat/src/ReductionOpsPlugin.inl:395-431);
// a batch-1 operand is broadcast over the other's batches
template <typename _Derived>
template <typename _Other>
auto MatrixBase<_Derived>::dot(const MatrixBase<_Other>& rhs) const
{
    static_assert(std::is_same<typename _Derived::Scalar, typename _Other::Scalar>::value,
                  "No implicit casting is allowed in dot products");
    typedef typename _Derived::Scalar S;
    auto a = internal::materialize(derived());
    auto b = internal::materialize(rhs.derived());
    CUMAT_ASSERT_DIMENSION(a.rows() * a.cols() == b.rows() * b.cols());
    CUMAT_ASSERT_DIMENSION(a.batches() == b.batches() || a.batches() == 1 || b.batches() == 1);
    Index batches = a.batches() > b.batches() ? a.batches() : b.batches();
    Matrix<S, 1, 1, Dynamic, ColumnMajor> out(1, 1, batches);
    CUMAT_AMD_CALL(cm_dot(Context::current().handle(), internal::DType<S>::value, (int64_t)(a.rows() * a.cols()),
                          (int)batches, a.data(), a.batches() > 1 ? 1 : 0, b.data(), b.batches() > 1 ? 1 : 0,
                          out.data()));
    return out;
}

template <typename _Derived>
auto MatrixBase<_Derived>::squaredNorm() const
{
    typedef typename _Derived::Scalar S;
    auto a = internal::materialize(derived());
    Matrix<S, 1, 1, Dynamic, ColumnMajor> out(1, 1, a.batches());
    CUMAT_AMD_CALL(cm_squared_norm(Context::current().handle(), internal::DType<S>::value,
                                   (int64_t)(a.rows() * a.cols()), (int)a.batches(), a.data(), out.data()));
    return out;
}

template <typename _Derived>
auto MatrixBase<_Derived>::norm() const
{
    return Matrix<typename _Derived::Scalar, 1, 1, Dynamic, ColumnMajor>(squaredNorm().cwiseSqrt());
}

CUMAT_NAMESPACE_END

#endif
