#ifndef CUMAT_AMD_REDUCTIONOPS_H
#define CUMAT_AMD_REDUCTIONOPS_H

// Reductions over any combination of the row / column / batch axes, dot,
// squaredNorm and norm (cuMat/src/ReductionOpsPlugin.inl:124-450 with the
// functors and initial values of cuMat/src/ReductionOps.h:859-968). The
// result keeps the non-reduced dimensions and has size 1 on the reduced ones.
// Unlike cuMat the reduction is evaluated when it is called, and the result
// is a Matrix, which every cuMat use of a reduction accepts. A stored matrix
// runs cm_reduce / cm_dot; an expression input is consumed lazily, as
// cuMat's input iterator does (cuMat/src/Iterator.h:99-105): its program is
// evaluated inside the reduction (cm_reduce_expr), so no temporary is
// written; dot / squaredNorm of expressions reduce cwiseDot / cwiseAbs2 the
// same way (cuMat/src/ReductionOpsPlugin.inl:290-297).

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

// stored inputs (a Matrix or a strided reference to one)
template <typename T> struct is_stored : std::false_type {};
template <typename S, int R, int C, int B, int F> struct is_stored<Matrix<S, R, C, B, F>> : std::true_type {};
template <typename S, int F> struct is_stored<DenseRef<S, F>> : std::true_type {};

// reduce_op over 'axis' of an expression, evaluated inside the reduction;
// out has the kept dimensions in the given layout
template <typename E>
void reduceExprInto(const E& e, int axis, int op, typename E::Scalar* out, bool out_row_major)
{
    typedef typename E::Scalar S;
    ProgramBuilder b;
    e.emit(b, 0);
    const Index R = e.rows(), C = e.cols(), B = e.batches();
    std::vector<cm_leaf> leaves(b.leaves.size() ? b.leaves.size() : 1);
    for (size_t k = 0; k < b.leaves.size(); ++k) {
        const LeafDesc& l = b.leaves[k];
        if (l.sr < 0 || l.sc < 0)
            throw std::invalid_argument("cuMat: SparseMatrix::direct() is only valid in sparse assignments and views");
        leaves[k].ptr = l.ptr;
        leaves[k].dtype = l.dtype;
        leaves[k].kind = l.kind;
        leaves[k].sr = (l.rows == 1 && R != 1) ? 0 : l.sr;
        leaves[k].sc = (l.cols == 1 && C != 1) ? 0 : l.sc;
        leaves[k].sb = (l.batches == 1 && B != 1) ? 0 : l.sb;
    }
    CUMAT_AMD_CALL(cm_reduce_expr(Context::current().handle(), leaves.data(), (int)b.leaves.size(), b.code.data(),
                                  (int)b.code.size(), DType<S>::value, op, axis, (int)R, (int)C, (int)B,
                                  out_row_major ? 1 : 0, out));
}

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
    if constexpr (!is_stored<E>::value) {
        const E& e = expr.derived();
        typedef typename E::Scalar S;
        typedef Matrix<S, (_Axis & Axis::Row) ? 1 : (int)E::Rows, (_Axis & Axis::Column) ? 1 : (int)E::Columns,
                       (_Axis & Axis::Batch) ? 1 : (int)E::Batches, E::Flags>
            Out;
        Out out((_Axis & Axis::Row) ? 1 : e.rows(), (_Axis & Axis::Column) ? 1 : e.cols(),
                (_Axis & Axis::Batch) ? 1 : e.batches());
        reduceExprInto(e, _Axis, op, out.data(), (E::Flags & RowMajor) != 0);
        return out;
    } else {
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
}

template <typename E>
auto reduceDynamic(const MatrixBase<E>& expr, int axis, int op)
{
    CUMAT_ASSERT_ARGUMENT(axis >= 0 && axis <= 7);
    if constexpr (!is_stored<E>::value) {
        const E& e = expr.derived();
        typedef Matrix<typename E::Scalar, Dynamic, Dynamic, Dynamic, E::Flags> Out;
        Out out((axis & Axis::Row) ? 1 : e.rows(), (axis & Axis::Column) ? 1 : e.cols(),
                (axis & Axis::Batch) ? 1 : e.batches());
        reduceExprInto(e, axis, op, out.data(), (E::Flags & RowMajor) != 0);
        return out;
    } else {
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
    if constexpr (!internal::is_stored<_Derived>::value || !internal::is_stored<_Other>::value) {
        // cwiseDot(rhs).sum<Row | Column>() in one reduction
        CUMAT_ASSERT_DIMENSION(derived().rows() == rhs.derived().rows() && derived().cols() == rhs.derived().cols());
        const auto prod = derived().cwiseDot(rhs.derived());
        Matrix<S, 1, 1, Dynamic, ColumnMajor> out(1, 1, prod.batches());
        internal::reduceExprInto(prod, Axis::Row | Axis::Column, CM_RED_SUM, out.data(), false);
        return out;
    } else {
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
}

template <typename _Derived>
auto MatrixBase<_Derived>::squaredNorm() const
{
    typedef typename _Derived::Scalar S;
    if constexpr (!internal::is_stored<_Derived>::value) {
        // cwiseAbs2().sum<Row | Column>() in one reduction
        const auto sq = derived().cwiseAbs2();
        Matrix<S, 1, 1, Dynamic, ColumnMajor> out(1, 1, sq.batches());
        internal::reduceExprInto(sq, Axis::Row | Axis::Column, CM_RED_SUM, out.data(), false);
        return out;
    } else {
    auto a = internal::materialize(derived());
    Matrix<S, 1, 1, Dynamic, ColumnMajor> out(1, 1, a.batches());
    CUMAT_AMD_CALL(cm_squared_norm(Context::current().handle(), internal::DType<S>::value,
                                   (int64_t)(a.rows() * a.cols()), (int)a.batches(), a.data(), out.data()));
    return out;
    }
}

template <typename _Derived>
auto MatrixBase<_Derived>::norm() const
{
    return Matrix<typename _Derived::Scalar, 1, 1, Dynamic, ColumnMajor>(squaredNorm().cwiseSqrt());
}

CUMAT_NAMESPACE_END

#endif
