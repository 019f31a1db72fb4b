#ifndef CUMAT_AMD_EIGEN_INTEROP_H
#define CUMAT_AMD_EIGEN_INTEROP_H

// Type mapping between cuMat matrices and Eigen matrices, active when the
// user defines CUMAT_EIGEN_SUPPORT=1 and has Eigen's headers on the include
// path (cuMat/src/EigenInteropHelpers.h:8-150, cuMat/src/Macros.h:65-71).
// The conversions themselves are Matrix::toEigen()/fromEigen() and, for CSR
// matrices, SparsityPattern<CSR>::fromEigen() and
// SparseMatrix<..., CSR>::fromEigen()/toEigen(); all of them copy
// synchronously through host memory and are meant for data in and out of a
// computation, not for its inner loop.
//
// Storage: a cuMat matrix of one row maps to a row-major Eigen matrix and
// one of one column to a column-major one (Eigen requires that of vectors;
// the bytes are the same either way), otherwise the storage flag carries
// over. The Eigen side is DontAlign so that its data() can be the target of
// a plain copy whatever the size.

#include "Macros.h"

#if CUMAT_EIGEN_SUPPORT == 1
#include <Eigen/Core>

CUMAT_NAMESPACE_BEGIN

template <typename _Scalar, int _Rows, int _Columns, int _Batches, int _Flags>
class Matrix;

namespace eigen {

template <int _Size>
struct SizeCuMatToEigen {
    enum { size = _Size == Dynamic ? int(::Eigen::Dynamic) : _Size };
};
template <int _Size>
struct SizeEigenToCuMat {
    enum { size = _Size == int(::Eigen::Dynamic) ? int(Dynamic) : _Size };
};

template <int _Rows, int _Columns, int _Flags>
struct StorageCuMatToEigen {
    enum {
        value = _Rows == 1 ? int(::Eigen::RowMajor)
                : _Columns == 1 ? int(::Eigen::ColMajor)
                : (_Flags & Flags::RowMajor) ? int(::Eigen::RowMajor)
                                             : int(::Eigen::ColMajor)
    };
};

// the Eigen type of Matrix<_Scalar, _Rows, _Columns, *, _Flags>
template <typename _Scalar, int _Rows, int _Columns, int _Flags>
struct EigenTypeOf {
    typedef ::Eigen::Matrix<_Scalar, SizeCuMatToEigen<_Rows>::size, SizeCuMatToEigen<_Columns>::size,
                            StorageCuMatToEigen<_Rows, _Columns, _Flags>::value | ::Eigen::DontAlign>
        type;
};

template <typename _CuMatMatrixType>
struct MatrixCuMatToEigen {
    typedef typename EigenTypeOf<typename _CuMatMatrixType::Scalar, _CuMatMatrixType::Rows,
                                 _CuMatMatrixType::Columns, _CuMatMatrixType::Flags>::type type;
};

// the cuMat matrix (one batch) holding an Eigen matrix type
template <typename _EigenMatrixType>
struct MatrixEigenToCuMat {
    typedef Matrix<typename _EigenMatrixType::Scalar, SizeEigenToCuMat<_EigenMatrixType::RowsAtCompileTime>::size,
                   SizeEigenToCuMat<_EigenMatrixType::ColsAtCompileTime>::size, 1,
                   _EigenMatrixType::IsRowMajor ? Flags::RowMajor : Flags::ColumnMajor>
        type;
};

}  // namespace eigen

CUMAT_NAMESPACE_END

#endif  // CUMAT_EIGEN_SUPPORT

#endif
