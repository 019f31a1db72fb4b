// eigen_ref.cpp -- thin extern "C" driver over the Eigen 3.3.90 headers that
// the reference vendors (third-party/Eigen). TEST INFRASTRUCTURE ONLY: the
// oracle's reference build, used by tests/ to pin the CPU restatement and by
// bench.py as the CPU baseline. Built by oracle/Makefile ("make ref") from
// the headers where they lie in the reference tree; output goes to _ref/.
//
// Each function runs the Eigen routine the reference benchmarks or ports:
//  - ConjugateGradient + DiagonalPreconditioner:
//      third-party/Eigen/src/IterativeLinearSolvers/ConjugateGradient.h:28-91
//      (cuMat's loop is "ported from Eigen", cuMat/src/ConjugateGradient.h:139)
//  - row-major SpMV y = A*x (OpenMP-parallel):
//      third-party/Eigen/src/SparseCore/SparseDenseProduct.h:39-51
//  - dot, linear combination and axis sums as in
//      benchmarks/{dotproduct,linear_combination,batched_reduction}.
#include <Eigen/Core>
#include <Eigen/Sparse>
#include <Eigen/IterativeLinearSolvers>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {
template <typename S>
using SpRow = Eigen::SparseMatrix<S, Eigen::RowMajor, int>;

template <typename S>
SpRow<S> make_matrix(int rows, int cols, int nnz, const int* JA, const int* IA, const S* vals)
{
    Eigen::Map<const SpRow<S>> m(rows, cols, nnz, JA, IA, vals);
    return SpRow<S>(m);
}

template <typename S>
int pcg(int n, int nnz, const int* JA, const int* IA, const S* vals, const S* b, S* x,
        int use_guess, double tol, int max_iter, int nthreads, double* err)
{
    Eigen::setNbThreads(nthreads);
    SpRow<S> A = make_matrix<S>(n, n, nnz, JA, IA, vals);
    Eigen::ConjugateGradient<SpRow<S>, Eigen::Lower | Eigen::Upper, Eigen::DiagonalPreconditioner<S>> cg;
    cg.setTolerance((S)tol);
    cg.setMaxIterations(max_iter);
    cg.compute(A);
    Eigen::Map<const Eigen::Matrix<S, Eigen::Dynamic, 1>> bb(b, n);
    Eigen::Map<Eigen::Matrix<S, Eigen::Dynamic, 1>> xx(x, n);
    Eigen::Matrix<S, Eigen::Dynamic, 1> sol;
    if (use_guess) {
        Eigen::Matrix<S, Eigen::Dynamic, 1> g = xx;
        sol = cg.solveWithGuess(bb, g);
    } else {
        sol = cg.solve(bb);
    }
    xx = sol;
    *err = (double)cg.error();
    return (int)cg.iterations();
}
}  // namespace

extern "C" {

int eref_pcg_f32(int n, int nnz, const int* JA, const int* IA, const float* vals, const float* b,
                 float* x, int use_guess, double tol, int max_iter, int nthreads, double* err)
{
    return pcg<float>(n, nnz, JA, IA, vals, b, x, use_guess, tol, max_iter, nthreads, err);
}

int eref_pcg_f64(int n, int nnz, const int* JA, const int* IA, const double* vals, const double* b,
                 double* x, int use_guess, double tol, int max_iter, int nthreads, double* err)
{
    return pcg<double>(n, nnz, JA, IA, vals, b, x, use_guess, tol, max_iter, nthreads, err);
}

// Opaque handle so that a timing loop measures only y = A*x.
void* eref_csr_create_f32(int rows, int cols, int nnz, const int* JA, const int* IA, const float* vals)
{
    return new SpRow<float>(make_matrix<float>(rows, cols, nnz, JA, IA, vals));
}
void eref_csr_destroy_f32(void* h) { delete static_cast<SpRow<float>*>(h); }

void eref_spmv_f32(void* h, const float* x, float* y, int nthreads)
{
    Eigen::setNbThreads(nthreads);
    const SpRow<float>& A = *static_cast<SpRow<float>*>(h);
    Eigen::Map<const Eigen::VectorXf> xx(x, A.cols());
    Eigen::Map<Eigen::VectorXf> yy(y, A.rows());
    yy.noalias() = A * xx;
}

// CG on a prebuilt handle (matrix construction excluded from timing).
int eref_pcg_handle_f32(void* h, const float* b, float* x, double tol, int max_iter, int nthreads,
                        double* err)
{
    Eigen::setNbThreads(nthreads);
    const SpRow<float>& A = *static_cast<SpRow<float>*>(h);
    Eigen::ConjugateGradient<SpRow<float>, Eigen::Lower | Eigen::Upper,
                             Eigen::DiagonalPreconditioner<float>> cg;
    cg.setTolerance((float)tol);
    cg.setMaxIterations(max_iter);
    cg.compute(A);
    Eigen::Map<const Eigen::VectorXf> bb(b, A.rows());
    Eigen::Map<Eigen::VectorXf> xx(x, A.rows());
    xx = cg.solve(bb);
    *err = (double)cg.error();
    return (int)cg.iterations();
}

float eref_dot_f32(int64_t n, const float* a, const float* b)
{
    Eigen::Map<const Eigen::VectorXf> aa(a, n), bb(b, n);
    return aa.dot(bb);
}

double eref_dot_f64(int64_t n, const double* a, const double* b)
{
    Eigen::Map<const Eigen::VectorXd> aa(a, n), bb(b, n);
    return aa.dot(bb);
}

// Linear combination as one Eigen expression (benchmarks/linear_combination/
// Implementation_Eigen.cpp:49-62 builds the same sum for a fixed count).
void eref_lincomb8_f32(int64_t n, const float* const* v, const float* c, float* out)
{
    using M = Eigen::Map<const Eigen::VectorXf>;
    M v0(v[0], n), v1(v[1], n), v2(v[2], n), v3(v[3], n), v4(v[4], n), v5(v[5], n), v6(v[6], n),
        v7(v[7], n);
    Eigen::Map<Eigen::VectorXf> o(out, n);
    o = (v0 * c[0] + v1 * c[1] + v2 * c[2] + v3 * c[3] + v4 * c[4] + v5 * c[5] + v6 * c[6] +
         v7 * c[7]).eval();
}

// Axis sums of a column-major rows x cols x batches float tensor.
// axis 1 (Row): out (1 x cols x batches) = colwise sums;
// axis 2 (Column): out (rows x 1 x batches) = rowwise sums;
// axis 4 (Batch): out (rows x cols) = sum over batches.
void eref_axis_sum_f32(int axis, int rows, int cols, int batches, const float* in, float* out)
{
    using MM = Eigen::Map<const Eigen::MatrixXf>;
    const int64_t slice = (int64_t)rows * cols;
    if (axis == 1) {
        for (int b = 0; b < batches; ++b) {
            MM m(in + slice * b, rows, cols);
            Eigen::Map<Eigen::RowVectorXf> o(out + (int64_t)cols * b, cols);
            o = m.colwise().sum();
        }
    } else if (axis == 2) {
        for (int b = 0; b < batches; ++b) {
            MM m(in + slice * b, rows, cols);
            Eigen::Map<Eigen::VectorXf> o(out + (int64_t)rows * b, rows);
            o = m.rowwise().sum();
        }
    } else if (axis == 4) {
        Eigen::Map<Eigen::MatrixXf> o(out, rows, cols);
        o = MM(in, rows, cols);
        for (int b = 1; b < batches; ++b) o += MM(in + slice * b, rows, cols);
    }
}

int eref_threads_max(void) { return Eigen::nbThreads(); }

// The batched workload as the CPU would run it: Eigen's ConjugateGradient on
// every system of a block-diagonal CSR (system s owns rows
// [row_off[s], row_off[s+1]), global column indices), the systems spread over
// 'nthreads' OpenMP threads (each solve single-threaded). 'stride' > 1 solves
// only every stride-th system (a bounded sample). Returns the systems solved.
int eref_batched_pcg_f32(int nsys, const int* row_off, const int* JA, const int* IA, const float* vals,
                         const float* b, float* x, double tol, int max_iter, int nthreads, int stride,
                         int* iters)
{
    if (stride < 1) stride = 1;
    const int nsolve = (nsys + stride - 1) / stride;
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 4)
    for (int q = 0; q < nsolve; ++q) {
        const int s = q * stride;
        const int lo = row_off[s], hi = row_off[s + 1], n = hi - lo;
        const int k0 = JA[lo], nnz = JA[hi] - k0;
        std::vector<int> ja(n + 1), ia(nnz);
        for (int i = 0; i <= n; ++i) ja[i] = JA[lo + i] - k0;
        for (int k = 0; k < nnz; ++k) ia[k] = IA[k0 + k] - lo;
        SpRow<float> A = make_matrix<float>(n, n, nnz, ja.data(), ia.data(), vals + k0);
        Eigen::ConjugateGradient<SpRow<float>, Eigen::Lower | Eigen::Upper, Eigen::DiagonalPreconditioner<float>> cg;
        cg.setTolerance((float)tol);
        cg.setMaxIterations(max_iter);
        cg.compute(A);
        Eigen::Map<const Eigen::VectorXf> bb(b + lo, n);
        Eigen::Map<Eigen::VectorXf> xx(x + lo, n);
        xx = cg.solve(bb);
        if (iters) iters[s] = (int)cg.iterations();
    }
    return nsolve;
}

}  // extern "C"
