#ifndef CUMAT_AMD_CONJUGATEGRADIENT_H
#define CUMAT_AMD_CONJUGATEGRADIENT_H

// Jacobi-preconditioned Conjugate Gradient on CSR or dense matrices
// (ConjugateGradient<SMatrixXf> / <MatrixXf>), with the API of
// cuMat's ConjugateGradient / IterativeSolverBase / SolverBase
// (cuMat/src/ConjugateGradient.h:66-265, cuMat/src/IterativeSolverBase.h:11-227,
// cuMat/src/SolverBase.h:34-126): default tolerance epsilon, default
// maxIterations 2*cols, iterations() and error() of the last solve, batched
// right-hand sides. compute(A) (Eigen's name) re-targets the solver to another
// matrix. The whole loop runs on the device (one cm_pcg_solve call); the host
// only checks convergence every few iterations.
//
// Block scalars (BlockScalar.h: SparseMatrix<BlockRows<3>> with Float4
// vectors, or the user's own types declared through BlockScalarTraits) run
// the block-scalar CG of the reference's blocked test.
//
// Two extensions without a cuMat counterpart sit beside it:
// BatchedConjugateGradient (many independent uneven systems, one persistent
// launch) and MixedPrecisionConjugateGradient (FP32 / BF16 inner CG inside
// FP64 iterative refinement).

#include <limits>

#include "SparseMatrix.h"

CUMAT_NAMESPACE_BEGIN

// invdiag = 1 / diag(A), 1 where the diagonal is zero (cwiseInverseCheck)
// (block scalars: the diagonal blocks' diagonals, one vector entry per
// block row, with the zero guard per component)
template <typename _MatrixType>
class DiagonalPreconditioner {
public:
    typedef typename _MatrixType::Scalar Scalar;
    typedef Matrix<typename internal::vector_scalar<Scalar>::type, Dynamic, 1, 1, ColumnMajor> Vector;

private:
    Vector entries_;

public:
    DiagonalPreconditioner() = default;
    explicit DiagonalPreconditioner(const _MatrixType& A) : entries_(A.rows())
    {
        if constexpr (internal::is_block_scalar<Scalar>::value) {
            const auto& P = A.getSparsityPattern();
            CUMAT_AMD_CALL(cm_bcsr_inverse_diagonal(Context::current().handle(), BlockScalarTraits<Scalar>::rows,
                                                    (int)A.rows(), (int)A.nnz(), 1, P.JA.data(), P.IA.data(),
                                                    (const float*)A.getData().data(), 0, (float*)entries_.data()));
        } else if constexpr (internal::is_sparse_matrix<_MatrixType>::value) {
            static_assert((int)_MatrixType::SFlags == (int)SparseFlags::CSR, "the CG solver takes CSR matrices");
            const auto& P = A.getSparsityPattern();
            CUMAT_AMD_CALL(cm_csr_inverse_diagonal(Context::current().handle(), internal::DType<Scalar>::value,
                                                   (int)A.rows(), (int)A.nnz(), 1, P.JA.data(), P.IA.data(),
                                                   A.getData().data(), 0, entries_.data()));
        } else {
            CUMAT_ASSERT_DIMENSION(A.rows() == A.cols());
            CUMAT_AMD_CALL(cm_dense_inverse_diagonal(Context::current().handle(), internal::DType<Scalar>::value,
                                                     (int)A.rows(), 1, A.data(), 0, entries_.data()));
        }
    }
    const Vector& entries() const { return entries_; }
    template <typename R>
    auto solve(const MatrixBase<R>& b) const { return entries_.cwiseMul(b.derived()); }
};

// no preconditioning: an inverse diagonal of ones
template <typename _MatrixType>
class IdentityPreconditioner {
public:
    typedef typename _MatrixType::Scalar Scalar;
    typedef Matrix<Scalar, Dynamic, 1, 1, ColumnMajor> Vector;

private:
    Vector entries_;

public:
    IdentityPreconditioner() = default;
    explicit IdentityPreconditioner(const _MatrixType& A) : entries_(A.rows())
    {
        entries_.inplace() = Vector::Constant(A.rows(), Scalar(1));
    }
    const Vector& entries() const { return entries_; }
    template <typename R>
    const R& solve(const MatrixBase<R>& b) const { return b.derived(); }
};

namespace internal {
template <typename M, bool Sparse = is_sparse_matrix<M>::value>
struct cg_matrix_ok : std::integral_constant<bool, (int)M::SFlags == (int)SparseFlags::CSR> {};
template <typename M>
struct cg_matrix_ok<M, false> : std::true_type {};
}  // namespace internal

template <typename _MatrixType, typename _Preconditioner = DiagonalPreconditioner<_MatrixType>>
class ConjugateGradient {
    static_assert(internal::cg_matrix_ok<_MatrixType>::value, "the CG solver takes CSR or dense matrices");
    static_assert((int)_MatrixType::Batches == 1, "the matrix of the CG solver has one batch (cuMat/src/ConjugateGradient.h:223)");
    static constexpr bool IsSparse = internal::is_sparse_matrix<_MatrixType>::value;

public:
    typedef _MatrixType MatrixType;
    typedef _Preconditioner Preconditioner;
    typedef typename _MatrixType::Scalar Scalar;
    // tolerances and errors in the elemental type (float for block scalars)
    typedef typename internal::elemental_scalar<Scalar>::type RealScalar;
    typedef Matrix<typename internal::vector_scalar<Scalar>::type, Dynamic, 1, Dynamic, ColumnMajor> Vector;
    static constexpr bool IsBlock = internal::is_block_scalar<Scalar>::value;

private:
    MatrixType matrix_;
    Preconditioner preconditioner_;
    bool initialized_ = false;
    RealScalar tolerance_ = std::numeric_limits<RealScalar>::epsilon();
    Index maxIterations_ = -1;
    mutable Index iterations_ = 0;
    mutable RealScalar error_ = 0;
    int path_ = CM_CG_AUTO;
    mutable cm_cg_stats stats_{};

public:
    ConjugateGradient() = default;
    explicit ConjugateGradient(const MatrixType& A) { compute(A); }
    ConjugateGradient(const MatrixType& A, const Preconditioner& P) : matrix_(A), preconditioner_(P), initialized_(true)
    {
        CUMAT_ASSERT_DIMENSION(A.rows() == A.cols());
    }

    ConjugateGradient& compute(const MatrixType& A)
    {
        CUMAT_ASSERT_DIMENSION(A.rows() == A.cols());
        matrix_ = A;
        preconditioner_ = Preconditioner(A);
        initialized_ = true;
        return *this;
    }

    const MatrixType& matrix() const { return matrix_; }
    const Preconditioner& preconditioner() const { return preconditioner_; }
    RealScalar tolerance() const { return tolerance_; }
    ConjugateGradient& setTolerance(RealScalar t)
    {
        tolerance_ = t;
        return *this;
    }
    Index maxIterations() const { return maxIterations_ < 0 ? 2 * matrix_.cols() : maxIterations_; }
    ConjugateGradient& setMaxIterations(Index n)
    {
        maxIterations_ = n;
        return *this;
    }
    Index iterations() const { return iterations_; }
    RealScalar error() const { return error_; }
    // launch shape: CM_CG_AUTO, CM_CG_MULTIKERNEL or CM_CG_RESIDENT
    ConjugateGradient& setPath(int path)
    {
        path_ = path;
        return *this;
    }
    const cm_cg_stats& lastStats() const { return stats_; }

    template <typename B>
    Vector solve(const MatrixBase<B>& b) const
    {
        auto rhs = internal::materialize(b.derived());
        Vector x(rhs.rows(), 1, rhs.batches());
        run(rhs, x, false);
        return x;
    }
    template <typename B, typename G>
    Vector solveWithGuess(const MatrixBase<B>& b, const MatrixBase<G>& guess) const
    {
        auto rhs = internal::materialize(b.derived());
        CUMAT_ASSERT_DIMENSION(guess.rows() == rhs.rows() && guess.cols() == 1);
        Vector x(rhs.rows(), 1, rhs.batches());
        if constexpr (IsBlock) {
            // block vector entries are copied as bytes (one guess per batch)
            auto g = internal::materialize(guess.derived());
            CUMAT_ASSERT_DIMENSION(g.batches() == rhs.batches());
            CUMAT_AMD_CALL(cm_memcpy_d2d(Context::current().handle(), x.data(), g.data(),
                                         sizeof(typename Vector::Scalar) * x.size()));
        } else {
            x.inplace() = guess.derived();  // broadcasts a batch-1 guess
        }
        run(rhs, x, true);
        return x;
    }

private:
    template <typename R>
    void run(const R& rhs, Vector& x, bool use_guess) const
    {
        CUMAT_ASSERT_ERROR(initialized_);
        CUMAT_ASSERT_DIMENSION(rhs.rows() == matrix_.cols() && rhs.cols() == 1);
        int it = 0;
        double err = 0;
        if constexpr (IsBlock) {
            // the block-scalar CG of the reference's blocked test
            // (tests/TestBlockedConjugateGradient.cu:244-275)
            const auto& P = matrix_.getSparsityPattern();
            CUMAT_AMD_CALL(cm_pcg_bcsr_solve(Context::current().handle(), BlockScalarTraits<Scalar>::rows,
                                             (int)matrix_.rows(), (int)matrix_.nnz(), (int)rhs.batches(),
                                             P.JA.data(), P.IA.data(), (const float*)matrix_.getData().data(), 0,
                                             (const float*)preconditioner_.entries().data(),
                                             (const float*)rhs.data(), (float*)x.data(), use_guess ? 1 : 0,
                                             (double)tolerance_, (int)maxIterations(), &it, &err, &stats_));
        } else if constexpr (IsSparse) {
            const auto& P = matrix_.getSparsityPattern();
            CUMAT_AMD_CALL(cm_pcg_solve(Context::current().handle(), internal::DType<Scalar>::value,
                                        (int)matrix_.rows(), (int)matrix_.nnz(), (int)rhs.batches(), P.JA.data(),
                                        P.IA.data(), matrix_.getData().data(), preconditioner_.entries().data(),
                                        rhs.data(), x.data(), use_guess ? 1 : 0, (double)tolerance_,
                                        (int)maxIterations(), path_, &it, &err, &stats_));
        } else {
            // the dense product A * p (cuMat/src/ProductOp.h:318-420) in place of the SpMV
            CUMAT_AMD_CALL(cm_pcg_dense_solve(Context::current().handle(), internal::DType<Scalar>::value,
                                              (int)matrix_.rows(), (int)rhs.batches(), matrix_.data(), 0,
                                              ((int)MatrixType::Flags & (int)Flags::RowMajor) ? 1 : 0,
                                              preconditioner_.entries().data(), rhs.data(), x.data(),
                                              use_guess ? 1 : 0, (double)tolerance_, (int)maxIterations(), &it,
                                              &err, &stats_));
        }
        iterations_ = it;
        error_ = (RealScalar)err;
    }
};

// Many independent SPD systems of uneven size (block-diagonal CSR with global
// column indices, system s owning rows [row_off[s], row_off[s+1])), solved by
// one persistent launch, each system stopping at its own convergence.
class BatchedConjugateGradient {
    Matrix<int, Dynamic, 1, 1, ColumnMajor> rowOff_;
    SMatrixXf A_;
    float tol_ = 1e-6f;
    int maxIter_ = 1000;
    int blockSize_ = 0;
    int tunedBlockSize_ = 0;
    bool useGraph_ = true;

public:
    BatchedConjugateGradient(const Matrix<int, Dynamic, 1, 1, ColumnMajor>& row_off, const SMatrixXf& A)
        : rowOff_(row_off), A_(A)
    {
        CUMAT_ASSERT_DIMENSION(row_off.rows() >= 2);
    }
    BatchedConjugateGradient& setTolerance(float t) { tol_ = t; return *this; }
    BatchedConjugateGradient& setMaxIterations(int n) { maxIter_ = n; return *this; }
    BatchedConjugateGradient& setBlockSize(int bs) { blockSize_ = bs; return *this; }
    BatchedConjugateGradient& setUseGraph(bool g) { useGraph_ = g; return *this; }
    Index systems() const { return rowOff_.rows() - 1; }

    // picks the launch shape (threads per system x workgroups per CU) by
    // timing the candidates on this batch; later solves use it
    int autotune(const VectorXf& b)
    {
        const auto& P = A_.getSparsityPattern();
        int bs = 0, gpc = 0;
        float ms = 0;
        CUMAT_AMD_CALL(cm_pcg_batched_autotune(Context::current().handle(), (int)systems(), rowOff_.data(),
                                               (int)A_.rows(), (int)A_.nnz(), P.JA.data(), P.IA.data(),
                                               A_.getData().data(), b.data(), tol_, maxIter_, &bs, &gpc, &ms));
        blockSize_ = 0;  // the tuned shape, block size and grid together
        tunedBlockSize_ = bs;
        return bs;
    }
    int tunedBlockSize() const { return tunedBlockSize_; }

    // x (zero initial guess) and per-system iteration counts / errors
    void solve(const VectorXf& b, VectorXf& x, VectorXi& iterations, VectorXf& errors) const
    {
        CUMAT_ASSERT_DIMENSION(b.rows() == A_.rows());
        if (x.rows() != b.rows()) x = VectorXf(b.rows());
        if (iterations.rows() != systems()) iterations = VectorXi(systems());
        if (errors.rows() != systems()) errors = VectorXf(systems());
        const auto& P = A_.getSparsityPattern();
        CUMAT_AMD_CALL(cm_pcg_batched_solve(Context::current().handle(), (int)systems(), rowOff_.data(),
                                            (int)A_.rows(), (int)A_.nnz(), P.JA.data(), P.IA.data(),
                                            A_.getData().data(), b.data(), x.data(), tol_, maxIter_,
                                            iterations.data(), errors.data(), blockSize_, useGraph_ ? 1 : 0));
    }
};

// FP64 iterative refinement around an FP32 inner CG whose matrix values are
// stored as FP32 (CM_F32) or BF16 (CM_BF16).
class MixedPrecisionConjugateGradient {
    SMatrixXd A_;
    int storage_ = CM_F32;
    double tol_ = 1e-12, innerTol_ = 1e-4;
    int maxOuter_ = 50, maxInner_ = -1;
    cm_ir_stats stats_{};

public:
    explicit MixedPrecisionConjugateGradient(const SMatrixXd& A, int storage = CM_F32) : A_(A), storage_(storage)
    {
        CUMAT_ASSERT_ARGUMENT(storage == CM_F32 || storage == CM_BF16);
        CUMAT_ASSERT_DIMENSION(A.rows() == A.cols());
    }
    MixedPrecisionConjugateGradient& setTolerance(double t) { tol_ = t; return *this; }
    MixedPrecisionConjugateGradient& setInnerTolerance(double t) { innerTol_ = t; return *this; }
    MixedPrecisionConjugateGradient& setMaxOuterIterations(int n) { maxOuter_ = n; return *this; }
    MixedPrecisionConjugateGradient& setMaxInnerIterations(int n) { maxInner_ = n; return *this; }
    const cm_ir_stats& stats() const { return stats_; }
    Index outerIterations() const { return stats_.outer_iterations; }
    Index innerIterations() const { return stats_.inner_iterations; }
    double error() const { return stats_.final_rel_residual; }

    VectorXd solve(const VectorXd& b)
    {
        CUMAT_ASSERT_DIMENSION(b.rows() == A_.rows());
        VectorXd x(b.rows());
        const auto& P = A_.getSparsityPattern();
        CUMAT_AMD_CALL(cm_pcg_ir_solve(Context::current().handle(), (int)A_.rows(), (int)A_.nnz(), P.JA.data(),
                                       P.IA.data(), A_.getData().data(), storage_, b.data(), x.data(), tol_,
                                       innerTol_, maxOuter_, maxInner_, &stats_));
        return x;
    }
};

CUMAT_NAMESPACE_END

#endif
