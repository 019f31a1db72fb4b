#ifndef CUMAT_AMD_SPARSEMATRIX_H
#define CUMAT_AMD_SPARSEMATRIX_H

// Sparse matrices of the host layer: SparsityPattern<CSR> with cuMat's
// names JA (row pointer, rows+1) and IA (column index, nnz)
// (cuMat/src/SparseMatrixBase.h:89-147), SparsityPattern<ELLPACK>
// (cuMat/src/SparseMatrixBase.h:149-206), and SparseMatrix<Scalar, Batches,
// Flags> whose values are an nnz x 1 x Batches matrix sharing one pattern
// (cuMat/src/SparseMatrix.h:48-414). A * x with a dense vector x is the
// sparse product of cuMat/src/SparseProductEvaluation.h:43-139 (CSR) and
// 194-324 (ELLPACK), evaluated by one cm_csr_spmv / cm_ell_spmv launch.
// Assigning a dense component-wise expression to a sparse matrix evaluates it
// at the stored entries only (cuMat/src/SparseEvaluation.h:14-34,83-138).

#include <vector>

#include "ProductOp.h"

#if CUMAT_EIGEN_SUPPORT == 1
#include <Eigen/SparseCore>
#endif

CUMAT_NAMESPACE_BEGIN

template <int _SparseFlags> struct SparsityPattern;

template <>
struct SparsityPattern<SparseFlags::CSR> {
    typedef int StorageIndex;
    typedef Matrix<StorageIndex, Dynamic, 1, 1, ColumnMajor> IndexVector;
    template <typename _Scalar, int _Batches>
    using DataMatrix = Matrix<_Scalar, Dynamic, 1, _Batches, ColumnMajor>;

    Index nnz = 0;
    Index rows = 0;
    Index cols = 0;
    IndexVector IA;
    IndexVector JA;

    // cuMat's dimension checks plus the content checks every kernel relies on
    // (JA non-decreasing from 0 to nnz, IA within [0, cols)): a malformed
    // pattern is rejected on the host instead of reaching a kernel.
    void assertValid() const
    {
        CUMAT_ASSERT_DIMENSION(JA.size() == rows + 1);
        CUMAT_ASSERT_DIMENSION(rows > 0);
        CUMAT_ASSERT_DIMENSION(cols > 0);
        CUMAT_ASSERT_DIMENSION(IA.size() == nnz);
        std::vector<int> ja((size_t)JA.size()), ia((size_t)IA.size());
        JA.copyToHost(ja.data());
        IA.copyToHost(ia.data());
        CUMAT_ASSERT_ARGUMENT(ja[0] == 0 && ja[(size_t)rows] == nnz);
        for (Index r = 0; r < rows; ++r) CUMAT_ASSERT_ARGUMENT(ja[(size_t)r] <= ja[(size_t)r + 1]);
        for (int c : ia) CUMAT_ASSERT_ARGUMENT(c >= 0 && c < cols);
    }
    Index getNNZ() const { return nnz; }
    template <typename _Scalar, int _Batches>
    DataMatrix<_Scalar, _Batches> allocateDataMatrix(Index batches) const
    {
        return DataMatrix<_Scalar, _Batches>(nnz, 1, batches);
    }
    SparsityPattern deepClone() const
    {
        SparsityPattern c;
        c.nnz = nnz;
        c.rows = rows;
        c.cols = cols;
        c.IA = IA.deepClone();
        c.JA = JA.deepClone();
        return c;
    }

#if CUMAT_EIGEN_SUPPORT == 1
    // The pattern of an Eigen sparse matrix: the rows of a row-major Eigen
    // matrix are CSR rows (outer index = JA, inner index = IA), as the
    // reference's CSR benchmark builds it by hand
    // (benchmarks/csrmv/Implementation_cuMat_CSR.cu:47-57). A column-major
    // or uncompressed matrix is converted to a compressed row-major copy
    // first.
    template <typename _EigenScalar, int _EigenOptions, typename _EigenIndex>
    static SparsityPattern fromEigen(const ::Eigen::SparseMatrix<_EigenScalar, _EigenOptions, _EigenIndex>& m)
    {
        if (!(_EigenOptions & ::Eigen::RowMajor) || !m.isCompressed() ||
            !std::is_same<_EigenIndex, StorageIndex>::value) {
            ::Eigen::SparseMatrix<_EigenScalar, ::Eigen::RowMajor, StorageIndex> c(m);
            c.makeCompressed();
            return fromEigen(c);
        }
        SparsityPattern p;
        p.rows = m.rows();
        p.cols = m.cols();
        p.nnz = m.nonZeros();
        p.JA = IndexVector(p.rows + 1);
        p.JA.copyFromHost(reinterpret_cast<const StorageIndex*>(m.outerIndexPtr()));
        p.IA = IndexVector(p.nnz);
        p.IA.copyFromHost(reinterpret_cast<const StorageIndex*>(m.innerIndexPtr()));
        p.assertValid();
        return p;
    }
#endif
};

template <>
struct SparsityPattern<SparseFlags::ELLPACK> {
    typedef int StorageIndex;
    // rows x nnzPerRow, column major; -1 marks padding
    typedef Matrix<StorageIndex, Dynamic, Dynamic, 1, ColumnMajor> IndexMatrix;
    template <typename _Scalar, int _Batches>
    using DataMatrix = Matrix<_Scalar, Dynamic, Dynamic, _Batches, ColumnMajor>;

    Index rows = 0;
    Index cols = 0;
    Index nnzPerRow = 0;
    IndexMatrix indices;

    void assertValid() const
    {
        CUMAT_ASSERT_DIMENSION(rows > 0);
        CUMAT_ASSERT_DIMENSION(cols > 0);
        CUMAT_ASSERT_DIMENSION(indices.rows() == rows && indices.cols() == nnzPerRow);
        std::vector<int> ind((size_t)indices.size());
        indices.copyToHost(ind.data());
        for (int c : ind) CUMAT_ASSERT_ARGUMENT(c >= -1 && c < cols);
    }
    Index getNNZ() const { return rows * nnzPerRow; }
    template <typename _Scalar, int _Batches>
    DataMatrix<_Scalar, _Batches> allocateDataMatrix(Index batches) const
    {
        return DataMatrix<_Scalar, _Batches>(rows, nnzPerRow, batches);
    }
};

template <typename _Scalar, int _Batches, int _SparseFlags>
class SparseMatrix {
    static_assert(_SparseFlags == SparseFlags::CSR || _SparseFlags == SparseFlags::ELLPACK,
                  "SparseFlags must be CSR or ELLPACK");

public:
    typedef _Scalar Scalar;
    typedef SparsityPattern<_SparseFlags> Pattern;
    typedef typename Pattern::template DataMatrix<_Scalar, _Batches> DataMatrix;
    enum { Batches = _Batches, SFlags = _SparseFlags };

private:
    Pattern sparsity_;
    Index batches_ = 0;
    DataMatrix A_;

public:
    SparseMatrix() = default;
    explicit SparseMatrix(const Pattern& pattern, Index batches = (_Batches == Dynamic ? 1 : _Batches))
        : sparsity_(pattern), batches_(batches), A_(pattern.template allocateDataMatrix<_Scalar, _Batches>(batches))
    {
        sparsity_.assertValid();
    }

    DataMatrix& getData() { return A_; }
    const DataMatrix& getData() const { return A_; }
    const Pattern& getSparsityPattern() const { return sparsity_; }
    Index rows() const { return sparsity_.rows; }
    Index cols() const { return sparsity_.cols; }
    Index batches() const { return batches_; }
    Index nnz() const { return sparsity_.getNNZ(); }
    Index size() const { return rows() * cols() * batches_; }
    bool isExclusiveUse() const { return A_.isExclusiveUse(); }
    void makeExclusiveUse() { A_.makeExclusiveUse(); }
    SparseMatrix deepClone() const
    {
        SparseMatrix m;
        m.sparsity_ = sparsity_;
        m.batches_ = batches_;
        m.A_ = A_.deepClone();
        return m;
    }

#if CUMAT_EIGEN_SUPPORT == 1
    // CSR matrix of one batch from / to an Eigen sparse matrix (compressed,
    // row-major, int indices on the way out); synchronous copies through
    // host memory
    typedef ::Eigen::SparseMatrix<_Scalar, ::Eigen::RowMajor, int> EigenSparseMatrix_t;
    template <int _EigenOptions, typename _EigenIndex>
    static SparseMatrix fromEigen(const ::Eigen::SparseMatrix<_Scalar, _EigenOptions, _EigenIndex>& m)
    {
        static_assert(_SparseFlags == SparseFlags::CSR, "Eigen interop is CSR only");
        static_assert(_Batches == 1 || _Batches == Dynamic, "Eigen matrices have one batch");
        EigenSparseMatrix_t c(m);  // row-major, compressed copy: values in CSR order
        c.makeCompressed();
        SparseMatrix out(Pattern::fromEigen(c), 1);
        out.A_.copyFromHost(c.valuePtr());
        return out;
    }
    EigenSparseMatrix_t toEigen() const
    {
        static_assert(_SparseFlags == SparseFlags::CSR, "Eigen interop is CSR only");
        CUMAT_ASSERT_ARGUMENT(batches_ == 1);
        const Index n = nnz();
        EigenSparseMatrix_t m(rows(), cols());
        m.resizeNonZeros(n);
        sparsity_.JA.copyToHost(m.outerIndexPtr());
        sparsity_.IA.copyToHost(m.innerIndexPtr());
        A_.copyToHost(m.valuePtr());
        return m;
    }
#endif

    // sparse * dense column vector (batched: matrix and/or vector batches)
    template <typename V>
    ProductOp<SparseMatrix, V> operator*(const MatrixBase<V>& v) const
    {
        return ProductOp<SparseMatrix, V>(*this, v.derived());
    }

    // dense expression evaluated at the stored entries (CSR)
    template <typename E>
    void assignCwise(const MatrixBase<E>& e, AssignmentMode mode)
    {
        static_assert(_SparseFlags == SparseFlags::CSR, "component-wise evaluation into sparse matrices is CSR only");
        static_assert(std::is_same<typename E::Scalar, _Scalar>::value, "No implicit casting in assignments");
        const E& x = e.derived();
        CUMAT_ASSERT_DIMENSION(x.rows() == rows() || x.rows() == 1);
        CUMAT_ASSERT_DIMENSION(x.cols() == cols() || x.cols() == 1);
        CUMAT_ASSERT_DIMENSION(x.batches() == batches_ || x.batches() == 1);
        internal::ProgramBuilder b;
        x.emit(b, 0);
        std::vector<cm_leaf> leaves(b.leaves.size() ? b.leaves.size() : 1);
        for (size_t k = 0; k < b.leaves.size(); ++k) {
            const internal::LeafDesc& l = b.leaves[k];
            leaves[k].ptr = l.ptr;
            leaves[k].dtype = l.dtype;
            leaves[k].kind = l.kind;
            leaves[k].sr = (l.rows == 1 && rows() != 1) ? 0 : l.sr;
            leaves[k].sc = (l.cols == 1 && cols() != 1) ? 0 : l.sc;
            leaves[k].sb = (l.batches == 1 && batches_ != 1) ? 0 : l.sb;
        }
        CUMAT_AMD_CALL(cm_csr_cwise_eval(Context::current().handle(), leaves.data(), (int)b.leaves.size(),
                                         b.code.data(), (int)b.code.size(), (int)rows(), (int)cols(), (int)nnz(),
                                         (int)batches_, sparsity_.JA.data(), sparsity_.IA.data(), A_.data(),
                                         internal::DType<_Scalar>::value, (int)mode));
    }
    template <typename E> SparseMatrix& operator=(const MatrixBase<E>& e) { assignCwise(e, AssignmentMode::ASSIGN); return *this; }
    template <typename E> SparseMatrix& operator+=(const MatrixBase<E>& e) { assignCwise(e, AssignmentMode::ADD); return *this; }
    template <typename E> SparseMatrix& operator-=(const MatrixBase<E>& e) { assignCwise(e, AssignmentMode::SUB); return *this; }
    template <typename E> SparseMatrix& operator*=(const MatrixBase<E>& e) { assignCwise(e, AssignmentMode::MUL); return *this; }
    template <typename E> SparseMatrix& operator/=(const MatrixBase<E>& e) { assignCwise(e, AssignmentMode::DIV); return *this; }
    // compound operations with another matrix of the same pattern, or a scalar
    // (cuMat/src/SparseMatrix.h:334-386)
    SparseMatrix& operator+=(const SparseMatrix& o) { assignCwise(o.direct(), AssignmentMode::ADD); return *this; }
    SparseMatrix& operator-=(const SparseMatrix& o) { assignCwise(o.direct(), AssignmentMode::SUB); return *this; }
    SparseMatrix& operator*=(const SparseMatrix& o) { assignCwise(o.direct(), AssignmentMode::MUL); return *this; }
    template <typename S, typename = internal::if_arith<S>>
    SparseMatrix& operator*=(const S& s) { assignCwise(HostScalar<_Scalar>((_Scalar)s), AssignmentMode::MUL); return *this; }
    template <typename S, typename = internal::if_arith<S>>
    SparseMatrix& operator+=(const S& s) { assignCwise(HostScalar<_Scalar>((_Scalar)s), AssignmentMode::ADD); return *this; }

    // the stored values as an expression leaf evaluated at each stored entry
    // (SparseMatrix::direct(), cuMat/src/SparseMatrix.h:230-233); valid in
    // assignments to a sparse matrix of the same pattern and in sparse views
    // of that pattern (sparseView(...) * x)
    StridedView<_Scalar> direct() const
    {
        internal::LeafDesc d{};
        d.ptr = A_.data();
        d.dtype = internal::DType<_Scalar>::value;
        d.kind = CM_LEAF_STRIDED;
        d.sr = -1;  // marks the value array of the destination pattern
        d.sc = -1;
        d.sb = nnz();  // entries per batch (CSR: nnz; ELLPACK: rows x nnzPerRow)
        d.rows = rows();
        d.cols = cols();
        d.batches = batches_;
        return StridedView<_Scalar>(A_.dataPointer(), d);
    }

    // identity pattern values: 1 on the diagonal, 0 elsewhere, as an
    // expression (evaluated at the stored entries when assigned)
    static StridedView<_Scalar> Identity(Index rows, Index cols)
    {
        Matrix<_Scalar, 1, 1, 1, ColumnMajor> one;
        one.inplace() = HostScalar<_Scalar>(_Scalar(1));
        internal::LeafDesc d{};
        d.ptr = one.data();
        d.dtype = internal::DType<_Scalar>::value;
        d.kind = CM_LEAF_AS_DIAGONAL;
        d.sr = d.sc = d.sb = 0;
        d.rows = rows;
        d.cols = cols;
        d.batches = 1;
        return StridedView<_Scalar>(one.dataPointer(), d);
    }

    // host copy of the values as a dense rows x cols x batches array
    // ([batch][row][col] order), for tests and printing
    std::vector<_Scalar> toDenseHost() const
    {
        static_assert(_SparseFlags == SparseFlags::CSR, "CSR only");
        std::vector<int> ja((size_t)(rows() + 1)), ia((size_t)nnz());
        std::vector<_Scalar> v((size_t)(nnz() * batches_)), out((size_t)size(), _Scalar(0));
        sparsity_.JA.copyToHost(ja.data());
        sparsity_.IA.copyToHost(ia.data());
        A_.copyToHost(v.data());
        for (Index b = 0; b < batches_; ++b)
            for (Index r = 0; r < rows(); ++r)
                for (int k = ja[(size_t)r]; k < ja[(size_t)r + 1]; ++k)
                    out[(size_t)((b * rows() + r) * cols() + ia[(size_t)k])] = v[(size_t)(k + nnz() * b)];
        return out;
    }
};

// A sparse matrix in an expression reads its stored values; on its own
// pattern this equals cuMat's coeff() search (cuMat/src/SparseMatrix.h:
// 135-142), the only pattern a sparse view of it may be evaluated on:
// (2 * A).sparseView<CSR>(A.getSparsityPattern()) * b.
template <typename S, typename _Scalar, int _Batches, int _SparseFlags, typename = internal::if_arith<S>>
auto operator*(const S& s, const SparseMatrix<_Scalar, _Batches, _SparseFlags>& A)
{
    return (_Scalar)s * A.direct();
}
template <typename S, typename _Scalar, int _Batches, int _SparseFlags, typename = internal::if_arith<S>>
auto operator*(const SparseMatrix<_Scalar, _Batches, _SparseFlags>& A, const S& s)
{
    return A.direct() * (_Scalar)s;
}

// expr.sparseView<CSR|ELLPACK>(pattern): the expression read only at the
// stored entries of 'pattern' (SparseExpressionOp,
// cuMat/src/SparseExpressionOp.h:32-72). Used as the left factor of a
// product with a dense vector, evaluated by ONE cm_csr_expr_spmv /
// cm_ell_expr_spmv launch that evaluates the expression inside the value
// loads (no value array is written).
template <typename _Child, int _SparseFlags>
class SparseExpressionOp {
    _Child e_;
    SparsityPattern<_SparseFlags> pattern_;

public:
    typedef typename _Child::Scalar Scalar;
    enum { SFlags = _SparseFlags };
    SparseExpressionOp(const _Child& e, const SparsityPattern<_SparseFlags>& p) : e_(e), pattern_(p)
    {
        CUMAT_ASSERT_DIMENSION(e.rows() == p.rows || e.rows() == 1);
        CUMAT_ASSERT_DIMENSION(e.cols() == p.cols || e.cols() == 1);
    }
    Index rows() const { return pattern_.rows; }
    Index cols() const { return pattern_.cols; }
    Index batches() const { return e_.batches(); }
    const _Child& expression() const { return e_; }
    const SparsityPattern<_SparseFlags>& getSparsityPattern() const { return pattern_; }

    template <typename V>
    ProductOp<SparseExpressionOp, V> operator*(const MatrixBase<V>& v) const
    {
        return ProductOp<SparseExpressionOp, V>(*this, v.derived());
    }
};

template <typename _Derived>
template <int _SparseFlags>
SparseExpressionOp<_Derived, _SparseFlags> MatrixBase<_Derived>::sparseView(
    const SparsityPattern<_SparseFlags>& pattern) const
{
    return SparseExpressionOp<_Derived, _SparseFlags>(derived(), pattern);
}

// y (mode)= view * x
template <typename C, int F, typename X>
void productEvalTo(const SparseExpressionOp<C, F>& A, const X& x, typename C::Scalar* y, Index batches,
                   AssignmentMode mode)
{
    typedef typename C::Scalar S;
    const auto& P = A.getSparsityPattern();
    internal::ProgramBuilder b;
    A.expression().emit(b, 0);
    const Index R = P.rows, Cc = P.cols;
    std::vector<cm_leaf> leaves(b.leaves.size() ? b.leaves.size() : 1);
    for (size_t k = 0; k < b.leaves.size(); ++k) {
        const internal::LeafDesc& l = b.leaves[k];
        leaves[k].ptr = l.ptr;
        leaves[k].dtype = l.dtype;
        leaves[k].kind = l.kind;
        leaves[k].sr = (l.rows == 1 && R != 1) ? 0 : l.sr;
        leaves[k].sc = (l.cols == 1 && Cc != 1) ? 0 : l.sc;
        leaves[k].sb = (l.batches == 1 && batches != 1) ? 0 : l.sb;
    }
    const int dt = internal::DType<S>::value;
    if constexpr (F == SparseFlags::CSR) {
        CUMAT_AMD_CALL(cm_csr_expr_spmv(Context::current().handle(), leaves.data(), (int)b.leaves.size(),
                                        b.code.data(), (int)b.code.size(), dt, (int)R, (int)Cc, (int)P.nnz,
                                        (int)batches, P.JA.data(), P.IA.data(), x.data(), x.batches() > 1 ? 1 : 0,
                                        y, (int)mode));
    } else {
        CUMAT_AMD_CALL(cm_ell_expr_spmv(Context::current().handle(), leaves.data(), (int)b.leaves.size(),
                                        b.code.data(), (int)b.code.size(), dt, (int)R, (int)Cc, (int)P.nnzPerRow,
                                        (int)batches, P.indices.data(), x.data(), x.batches() > 1 ? 1 : 0, y,
                                        (int)mode));
    }
}

namespace internal {
template <typename T>
struct is_sparse_matrix : std::false_type {};
template <typename S, int B, int F>
struct is_sparse_matrix<SparseMatrix<S, B, F>> : std::true_type {};
}  // namespace internal

// y (mode)= A * x with a sparse A (ProductOp<SparseMatrix, V>, evaluated
// when assigned to a dense matrix): one SpMV launch
// (ProductAssignment over CSR / ELLPACK, cuMat/src/SparseProductEvaluation.h:43-139,194-324)
template <typename S, int B, int F, typename X>
void productEvalTo(const SparseMatrix<S, B, F>& A, const X& x, typename internal::vector_scalar<S>::type* y,
                   Index batches, AssignmentMode mode)
{
    const auto& P = A.getSparsityPattern();
    if constexpr (internal::is_block_scalar<S>::value) {
        // block scalars: the block product of each stored block with x's
        // entry, rows summed in CSR order (cm_bcsr_spmv)
        static_assert(F == SparseFlags::CSR, "block scalars are stored in CSR");
        CUMAT_AMD_CALL(cm_bcsr_spmv(Context::current().handle(), BlockScalarTraits<S>::rows, (int)P.rows,
                                    (int)P.cols, (int)P.nnz, (int)batches, P.JA.data(), P.IA.data(),
                                    (const float*)A.getData().data(), A.batches() > 1 ? 1 : 0,
                                    (const float*)x.data(), x.batches() > 1 ? 1 : 0, (float*)y, (int)mode));
    } else if constexpr (F == SparseFlags::CSR) {
        const int dt = internal::DType<S>::value;
        CUMAT_AMD_CALL(cm_csr_spmv(Context::current().handle(), dt, dt, (int)P.rows, (int)P.cols, (int)P.nnz,
                                   (int)batches, P.JA.data(), P.IA.data(), A.getData().data(),
                                   A.batches() > 1 ? 1 : 0, x.data(), x.batches() > 1 ? 1 : 0, y, (int)mode));
    } else {
        CUMAT_AMD_CALL(cm_ell_spmv(Context::current().handle(), internal::DType<S>::value, (int)P.rows,
                                   (int)P.cols, (int)P.nnzPerRow, (int)batches, P.indices.data(),
                                   A.getData().data(), A.batches() > 1 ? 1 : 0, x.data(),
                                   x.batches() > 1 ? 1 : 0, y, (int)mode));
    }
}

// typedefs (cuMat/src/SparseMatrix.h:606-623)
#define CUMAT_AMD_SPARSE_TYPEDEFS(S, s) \
    typedef SparseMatrix<S, 1, SparseFlags::CSR> SMatrixX##s; \
    typedef SparseMatrix<S, Dynamic, SparseFlags::CSR> BSMatrixX##s; \
    typedef SparseMatrix<S, 1, SparseFlags::CSR> SMatrixX##s##_CSR; \
    typedef SparseMatrix<S, Dynamic, SparseFlags::CSR> BSMatrixX##s##_CSR; \
    typedef SparseMatrix<S, 1, SparseFlags::ELLPACK> SMatrixX##s##_ELLPACK; \
    typedef SparseMatrix<S, Dynamic, SparseFlags::ELLPACK> BSMatrixX##s##_ELLPACK;
CUMAT_AMD_SPARSE_TYPEDEFS(int, i)
CUMAT_AMD_SPARSE_TYPEDEFS(float, f)
CUMAT_AMD_SPARSE_TYPEDEFS(double, d)
#undef CUMAT_AMD_SPARSE_TYPEDEFS

CUMAT_NAMESPACE_END

#endif
