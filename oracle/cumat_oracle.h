/*
 * cumat_oracle.h -- CPU restatement of the reference algorithms on the cuMat
 * hot path. TEST INFRASTRUCTURE ONLY: used by tests/, by the smoke check and
 * by bench.py's CPU-baseline leg as the checker, never by the product path.
 *
 * Every function names the reference file:line it restates. Each is pinned
 * by tests/test_oracle.py against the reference's own test vectors
 * (tests/golden/reference_tests.json), the Eigen build of oracle/_ref or the
 * committed Eigen runs (tests/golden/eigen_cg.json), and closed forms:
 *   orc_csr_spmv_*    4x5 fixture (plain, batched, +=/-=), Dirichlet A*1
 *   orc_ell_spmv_*    ELLPACK fixture in every dtype, equal to the CSR form
 *   orc_bcsr_*        the dense expansion (numpy), blocks holding one scalar
 *                     equal to orc_csr_spmv_f32, the blocked CG fixture
 *   orc_reduce        integer fixtures of every op / axis / layout, sum 1..N
 *   orc_dot           batched dot and squared-norm fixtures
 *   orc_cwise_*, orc_assign  compound-assignment fixtures, C / and %
 *   orc_csr_inverse_diagonal_f64  diagonal fixture, zero guard
 *   orc_pcg_*         Eigen's ConjugateGradient (same iterations), closed forms
 *   orc_edge_grad_f64 central differences of lambda^T A(theta) x (bilinear)
 */
#ifndef CUMAT_ORACLE_H
#define CUMAT_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* dtypes and enums mirror include/cumat_amd.h */
enum { ORC_I32 = 0, ORC_F32 = 1, ORC_F64 = 2, ORC_BOOL = 3 };

/* ---- problem generators (benchmarks/conjugate_gradient/Implementation_cuMat.cu:29-41) ---- */
/* 2D 5-point Poisson, IDX(x,y) = y + x*g, columns ascending, diag 4+shift,
 * off-diagonal -1, Dirichlet boundary. rowptr: n+1, colind/vals: 5n max.
 * Returns nnz. */
int64_t orc_poisson2d(int g, double shift, int32_t* rowptr, int32_t* colind, double* vals);

/* ---- CSR SpMV (cuMat/src/SparseProductEvaluation.h:43-88): sequential
 * accumulation over the row in CSR order, in the value type. ---- */
void orc_csr_spmv_f32(int rows, int cols, int nnz, int batches,
                      const int32_t* JA, const int32_t* IA,
                      const float* vals, int vals_batched,
                      const float* x, int x_batched, float* y, int mode);
void orc_csr_spmv_f64(int rows, int cols, int nnz, int batches,
                      const int32_t* JA, const int32_t* IA,
                      const double* vals, int vals_batched,
                      const double* x, int x_batched, double* y, int mode);
void orc_csr_spmv_i32(int rows, int cols, int nnz, int batches,
                      const int32_t* JA, const int32_t* IA,
                      const int32_t* vals, int vals_batched,
                      const int32_t* x, int x_batched, int32_t* y, int mode);
/* Block CSR with D x D float blocks (D = 2..4) stored as D float4 rows and
 * float4 vector entries, the reference's block-scalar plug-in types
 * (tests/TestBlockedConjugateGradient.cu:12-239); rows/cols count blocks. */
void orc_bcsr_spmv_f32(int D, int rows, int cols, int nnz, int batches,
                       const int32_t* JA, const int32_t* IA,
                       const float* vals, int vals_batched,
                       const float* x, int x_batched, float* y, int mode);
void orc_bcsr_inverse_diagonal_f32(int D, int rows, int nnz, int batches, const int32_t* JA, const int32_t* IA,
                                   const float* vals, int vals_batched, float* out);
/* ELLPACK (cuMat/src/SparseProductEvaluation.h:194-226) */
void orc_ell_spmv_f32(int rows, int cols, int nnz_per_row, int batches,
                      const int32_t* indices, const float* vals, int vals_batched,
                      const float* x, int x_batched, float* y, int mode);
void orc_ell_spmv_f64(int rows, int cols, int nnz_per_row, int batches,
                      const int32_t* indices, const double* vals, int vals_batched,
                      const double* x, int x_batched, double* y, int mode);
void orc_ell_spmv_i32(int rows, int cols, int nnz_per_row, int batches,
                      const int32_t* indices, const int32_t* vals, int vals_batched,
                      const int32_t* x, int x_batched, int32_t* y, int mode);

/* ---- reductions (cuMat/src/ReductionOps.h:57-160 axis semantics,
 * functors cuMat/src/ReductionOps.h:859-968, initial values
 * cuMat/src/ReductionOpsPlugin.inl:9-213). Floating sums/products are
 * accumulated in long double and rounded once. ---- */
void orc_reduce(int dtype, int op, int axis, int rows, int cols, int batches,
                int in_row_major, const void* in, int out_row_major, void* out);
/* dot / squaredNorm (cuMat/src/ReductionOpsPlugin.inl:395-431) */
void orc_dot(int dtype, int64_t n, int batches, const void* a, int a_batched,
             const void* b, int b_batched, void* out);

/* ---- component-wise functors (cuMat/src/BinaryOps.h:230-243,
 * cuMat/src/UnaryOps.h:128-225) and assignment modes
 * (cuMat/src/CwiseOp.h:15-62) on flat arrays ---- */
void orc_cwise_binary(int dtype, int fn, int64_t n, const void* a, const void* b, void* out);
void orc_cwise_unary(int dtype, int fn, int64_t n, const void* a, void* out);
void orc_assign(int dtype, int mode, int64_t n, void* dst, const void* src);

/* ---- Jacobi (cuMat/src/ConjugateGradient.h:218-246, UnaryOps.h:177) ---- */
void orc_csr_inverse_diagonal_f64(int rows, const int32_t* JA, const int32_t* IA,
                                  const double* vals, double* invdiag);

/* ---- Jacobi-PCG (cuMat/src/ConjugateGradient.h:123-206, identical to
 * third-party/Eigen/src/IterativeLinearSolvers/ConjugateGradient.h:28-91).
 * Scalars of the f32 variant are float; each dot is accumulated in double
 * and rounded to float. Matrix batch 1, rhs/solution 'batches' wide; all
 * batches iterate until all converge. Returns iterations; *err = error(). */
int orc_pcg_f32(int n, int nnz, int batches, const int32_t* JA, const int32_t* IA,
                const float* vals, const float* b, float* x, int use_guess,
                double tol, int max_iter, double* err);
int orc_pcg_f64(int n, int nnz, int batches, const int32_t* JA, const int32_t* IA,
                const double* vals, const double* b, double* x, int use_guess,
                double tol, int max_iter, double* err);

/* ---- edge Laplacian helpers for the adjoint addition ---- */
/* dL/dtheta_e = -(lam_i - lam_j)(x_i - x_j), boundary edge (ej<0): -lam_i x_i */
void orc_edge_grad_f64(int nedges, const int32_t* ei, const int32_t* ej,
                       const double* lam, const double* x, double* grad);

#ifdef __cplusplus
}
#endif
#endif
