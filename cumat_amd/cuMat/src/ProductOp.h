#ifndef CUMAT_AMD_PRODUCTOP_H
#define CUMAT_AMD_PRODUCTOP_H

// Matrix-vector products y = A * x (cuMat's ProductOp,
// cuMat/src/ProductOp.h:126-305, with operator* between two matrices at
// cuMat/src/ProductOp.h:520-523): a lazy node evaluated when it is assigned
// to a dense matrix (=, +=, -=). A dense left operand runs cm_dense_gemv
// (cuMat hands it to cuBLAS gemm, cuMat/src/ProductOp.h:318-420); a sparse
// one runs the CSR / ELLPACK SpMV (productEvalTo in SparseMatrix.h). The
// right operand is a column vector, batched or not; a batch-1 operand is
// broadcast over the other's batches. Matrix-matrix products (right operand
// with more than one column) are not on this library's path and are
// rejected.

#include "BlockScalar.h"
#include "ReductionOps.h"

CUMAT_NAMESPACE_BEGIN

template <typename _Left, typename _Right>
class ProductOp {
    _Left A_;
    _Right x_;

public:
    // the result's entry type: the vector entry type of a block-scalar
    // matrix (BlockScalar.h), else the common scalar type
    typedef typename internal::vector_scalar<typename _Left::Scalar>::type Scalar;
    static_assert(std::is_same<Scalar, typename _Right::Scalar>::value,
                  "No implicit casting between the factors of a product");

    ProductOp(const _Left& A, const _Right& x) : A_(A), x_(x)
    {
        CUMAT_ASSERT_DIMENSION(x.cols() == 1);
        CUMAT_ASSERT_DIMENSION(x.rows() == A.cols());
        CUMAT_ASSERT_DIMENSION(A.batches() == x.batches() || A.batches() == 1 || x.batches() == 1);
    }
    Index rows() const { return A_.rows(); }
    Index cols() const { return 1; }
    Index batches() const { return A_.batches() > x_.batches() ? A_.batches() : x_.batches(); }

    void evalTo(Scalar* y, Index y_rows, Index y_batches, AssignmentMode mode) const
    {
        CUMAT_ASSERT_DIMENSION(y_rows == rows() && y_batches == batches());
        auto x = internal::materialize(x_);
        productEvalTo(A_, x, y, batches(), mode);
    }
};

// dense left operand (a matrix or any expression, evaluated first)
template <typename E, typename X>
void productEvalTo(const MatrixBase<E>& Aexpr, const X& x, typename E::Scalar* y, Index batches, AssignmentMode mode)
{
    auto A = internal::materialize(Aexpr.derived());
    typedef decltype(A) M;
    CUMAT_AMD_CALL(cm_dense_gemv(Context::current().handle(), internal::DType<typename E::Scalar>::value,
                                 (int)A.rows(), (int)A.cols(), (int)batches, A.data(), A.batches() > 1 ? 1 : 0,
                                 ((int)M::Flags & (int)Flags::RowMajor) ? 1 : 0, x.data(), x.batches() > 1 ? 1 : 0,
                                 y, (int)mode));
}

// A * x between two dense matrices / expressions
template <typename L, typename R>
ProductOp<L, R> operator*(const MatrixBase<L>& l, const MatrixBase<R>& r)
{
    return ProductOp<L, R>(l.derived(), r.derived());
}

CUMAT_NAMESPACE_END

#endif
