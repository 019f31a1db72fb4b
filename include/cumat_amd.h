/*
 * cumat_amd.h -- C ABI of the MI355X (gfx950) hot path of cuMat:
 * fused component-wise evaluation, batched row/column/batch reductions,
 * CSR SpMV, and the Jacobi-preconditioned ConjugateGradient loop, plus three
 * additions with no reference counterpart (persistent batched CG,
 * mixed-precision CG with FP64 iterative refinement, adjoint CG gradient).
 *
 * The reference (cuMat) is a header-only CUDA C++ template library with no
 * binary interface; each entry point below replaces one template "plugin"
 * surface of it, cited as file:line of the reference tree.
 *
 * Conventions
 *  - Every function returns an int status (CM_OK == 0). On failure the
 *    thread-local message is available from cm_last_error(). Nothing throws.
 *  - All device pointers are plain HIP device pointers owned by the caller
 *    (cm_malloc is offered for callers without their own allocator).
 *  - All work is stream-ordered on the stream of the cm_ctx_t argument; a
 *    NULL context means the calling thread's default context (device 0, its
 *    own non-blocking stream), like cuMat's Context::current()
 *    (cuMat/src/Context.h:204-214).
 *  - Dense layout follows cuMat's Matrix (cuMat/src/Matrix.h:613-627):
 *    column major  idx = row + rows*(col + cols*batch),
 *    row major     idx = col + cols*(row + rows*batch).
 *  - Sparse CSR keeps cuMat's names (cuMat/src/SparseMatrixBase.h:89-147):
 *    JA = row pointer (rows+1 entries), IA = column index (nnz entries),
 *    both int32; batched values are stored batch after batch (stride nnz).
 */
#ifndef CUMAT_AMD_H
#define CUMAT_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CM_ABI_VERSION 1

/* ---- status codes (replace cuMat's exceptions, cuMat/src/Errors.h:17-30,
 *      cuMat/src/Macros.h:158-192) ---- */
enum cm_status {
    CM_OK = 0,
    CM_E_ARG = 1,          /* invalid argument (std::invalid_argument)       */
    CM_E_DIM = 2,          /* dimension mismatch (std::invalid_argument)     */
    CM_E_HIP = 3,          /* HIP runtime error (cuMat::cuda_error)          */
    CM_E_NOMEM = 4,        /* allocation failure                             */
    CM_E_UNSUPPORTED = 5,  /* combination not implemented                    */
    CM_E_STATE = 6         /* object used in the wrong state                 */
};

/* ---- scalar types ---- */
enum cm_dtype {
    CM_I32 = 0,
    CM_F32 = 1,
    CM_F64 = 2,
    CM_BOOL = 3,   /* stored as one byte, 0 or 1 (cuMat's bool matrices)   */
    CM_BF16 = 4    /* storage only: sparse matrix values in mixed CG        */
};

/* ---- assignment modes: order of cuMat::AssignmentMode
 *      (cuMat/src/Constants.h:153-163), semantics of
 *      CwiseAssignmentHandler (cuMat/src/CwiseOp.h:15-62) ---- */
enum cm_assign_mode {
    CM_ASSIGN = 0, CM_ADD = 1, CM_SUB = 2, CM_MUL = 3,
    CM_DIV = 4, CM_MOD = 5, CM_AND = 6, CM_OR = 7
};

/* ---- reduction axes: cuMat::Axis bits (cuMat/src/Constants.h Axis) ---- */
enum cm_axis { CM_AXIS_ROW = 1, CM_AXIS_COL = 2, CM_AXIS_BATCH = 4, CM_AXIS_ALL = 7 };

/* ---- reduction operators: cuMat functor::Sum..BitwiseOr
 *      (cuMat/src/ReductionOps.h:859-968); the initial values are those of
 *      cuMat/src/ReductionOpsPlugin.inl:9-213 ---- */
enum cm_reduce_op {
    CM_RED_SUM = 0, CM_RED_PROD = 1, CM_RED_MIN = 2, CM_RED_MAX = 3,
    CM_RED_ALL = 4, CM_RED_ANY = 5, CM_RED_BAND = 6, CM_RED_BOR = 7
};

/* ---- profiling counters: cuMat::Profiling::Counter
 *      (cuMat/src/Profiling.h:23-109); used by tests to prove fusion ---- */
enum cm_counter {
    CM_CNT_DEVICE_MEM_ALLOC = 0,
    CM_CNT_DEVICE_MEM_FREE = 1,
    CM_CNT_EVAL_ANY = 2,
    CM_CNT_EVAL_CWISE = 3,
    CM_CNT_EVAL_CWISE_SPARSE = 4,
    CM_CNT_EVAL_REDUCTION = 5,
    CM_CNT_EVAL_MATMUL_SPARSE = 6,
    CM_CNT_KERNEL_LAUNCH = 7,
    CM_CNT_EVAL_MATMUL = 8,
    /* no reference counterpart: host waits on a stream or event issued by the
     * library (what a stream-ordered caller wants to see stay at zero) */
    CM_CNT_HOST_SYNC = 9,
    /* no reference counterpart: kernels compiled at run time for a
     * component-wise program, and launches of such kernels (cm_jit_mode) */
    CM_CNT_JIT_COMPILE = 10,
    CM_CNT_JIT_LAUNCH = 11,
    CM_CNT_COUNT = 12
};

typedef struct cm_context* cm_ctx_t;

/* ======================================================================
 * Runtime: replaces cuMat::Context, DevicePointer, Profiling
 * (cuMat/src/Context.h:132-446, cuMat/src/DevicePointer.h:11-148,
 *  cuMat/src/Profiling.h:23-145)
 * ====================================================================== */
int         cm_abi_version(void);
const char* cm_last_error(void);
int cm_device_count(int* count);
int cm_device_info(int device, char* name, int name_len, int* cu_count,
                   long long* hbm_bytes, int* clock_khz, int* mem_clock_khz,
                   int* bus_width_bits);
int cm_ctx_create(int device, cm_ctx_t* out);
int cm_ctx_destroy(cm_ctx_t ctx);
int cm_ctx_get_stream(cm_ctx_t ctx, void** stream);
int cm_ctx_set_stream(cm_ctx_t ctx, void* stream);
int cm_ctx_sync(cm_ctx_t ctx);
/* Caching allocation, like cuMat's use of cub::CachingDeviceAllocator
 * (cuMat/src/Context.h:223-229,271-286,311-325): cm_free is stream-ordered
 * and never waits. A freed block returns to the cache with an event recorded
 * on the freeing context's stream; work on that stream may reuse it at once
 * (stream order), and another stream reuses it only after that event (a
 * device-side wait, or a block whose event has completed). */
int cm_malloc(cm_ctx_t ctx, size_t bytes, void** ptr);
int cm_free(cm_ctx_t ctx, void* ptr);
/* d2d is stream-ordered; h2d is stream-ordered from pinned host memory and
 * waits for the copy from pageable memory (the caller may reuse the buffer
 * on return); d2h synchronizes the stream like Matrix::copyToHost
 * (cuMat/src/Matrix.h:847-858). */
int cm_memcpy_h2d(cm_ctx_t ctx, void* dst, const void* src, size_t bytes);
int cm_memcpy_d2h(cm_ctx_t ctx, void* dst, const void* src, size_t bytes);
int cm_memcpy_d2d(cm_ctx_t ctx, void* dst, const void* src, size_t bytes);
/* byte fill done by a kernel (safe under stream capture); Matrix::setZero,
 * cuMat/src/Matrix.h:1232-1238 */
int cm_fill_bytes(cm_ctx_t ctx, void* ptr, int value, size_t bytes);
int cm_counters_get(int counter, long long* value);
int cm_counters_reset(void);

/* ======================================================================
 * Fused component-wise evaluation.
 * Replaces CwiseEvaluationKernel + Assignment<Dense,Cwise>
 * (cuMat/src/CwiseOp.h:66-85,130-158) with the functors of
 * cuMat/src/BinaryOps.h:230-243, cuMat/src/UnaryOps.h:128-225 and the
 * casting op cuMat/src/UnaryOps.h:248-312.
 *
 * An expression tree is lowered by the host (C++ header layer or Python) to
 * a short accumulator program over up to CM_CWISE_MAX_LEAVES strided leaves.
 * ONE kernel launch evaluates the whole program for every destination
 * element, so no intermediate is written (the fusion of cuMat's kernel
 * merging, README.md:13,16).
 * ====================================================================== */
#define CM_CWISE_MAX_LEAVES 12
#define CM_CWISE_MAX_CODE   48
#define CM_CWISE_MAX_STACK  4

enum cm_leaf_kind {
    CM_LEAF_STRIDED = 0,    /* value = ptr[r*sr + c*sc + b*sb]                */
    CM_LEAF_AS_DIAGONAL = 1 /* value = (r==c) ? ptr[r*sr + b*sb] : 0
                               (AsDiagonalOp, cuMat/src/UnaryOps.h:361-437)  */
};

typedef struct cm_leaf {
    const void* ptr;
    int32_t dtype;          /* cm_dtype of the stored values                  */
    int32_t kind;           /* cm_leaf_kind                                   */
    int64_t sr, sc, sb;     /* element strides; 0 on a broadcast axis         */
} cm_leaf;

/* Opcodes of the accumulator machine. 'acc' is the running value; every
 * instruction states the scalar type (cm_dtype) its result has. */
enum cm_opcode {
    CM_OP_LOAD = 0,       /* acc = src                                        */
    CM_OP_PUSH = 1,       /* stack[arg] = acc                                 */
    CM_OP_BINARY = 2,     /* acc = acc (fn) src, or src (fn) acc if flags&1   */
    CM_OP_UNARY = 3,      /* acc = fn(acc)                                    */
    CM_OP_CAST = 4,       /* acc = (dtype) acc; 'arg' is the source dtype     */
    CM_OP_FMA = 5,        /* acc = acc + leaf[arg] * imm  (axpy term)         */
    CM_OP_SELECT = 6      /* acc = acc ? stack[arg] : src (cwiseSelect)       */
};
enum cm_src_kind { CM_SRC_NONE = 0, CM_SRC_LEAF = 1, CM_SRC_IMM = 2, CM_SRC_STACK = 3 };

/* binary functors, cuMat/src/BinaryOps.h:230-243 + BinaryOpsPlugin.inl */
enum cm_binary_fn {
    CM_BIN_ADD = 0, CM_BIN_SUB, CM_BIN_MUL, CM_BIN_DIV, CM_BIN_MOD,
    CM_BIN_POW, CM_BIN_MIN, CM_BIN_MAX, CM_BIN_AND, CM_BIN_OR, CM_BIN_XOR,
    CM_BIN_LAND, CM_BIN_LOR,
    CM_BIN_EQ, CM_BIN_NE, CM_BIN_LT, CM_BIN_GT, CM_BIN_LE, CM_BIN_GE,
    CM_BIN_COUNT
};
/* unary functors, cuMat/src/UnaryOps.h:128-225 */
enum cm_unary_fn {
    CM_UN_NEG = 0, CM_UN_ABS, CM_UN_ABS2, CM_UN_INVERSE, CM_UN_INVERSE_CHECK,
    CM_UN_SQRT, CM_UN_RSQRT, CM_UN_CBRT, CM_UN_EXP, CM_UN_LOG, CM_UN_LOG1P,
    CM_UN_LOG10, CM_UN_SIN, CM_UN_COS, CM_UN_TAN, CM_UN_ASIN, CM_UN_ACOS,
    CM_UN_ATAN, CM_UN_SINH, CM_UN_COSH, CM_UN_TANH, CM_UN_FLOOR, CM_UN_CEIL,
    CM_UN_ROUND, CM_UN_NOT, CM_UN_BITNOT, CM_UN_SIGN,
    CM_UN_COUNT
};

typedef struct cm_instr {
    int32_t opcode;     /* cm_opcode                                          */
    int32_t dtype;      /* result type                                        */
    int32_t fn;         /* cm_binary_fn / cm_unary_fn                         */
    int32_t src;        /* cm_src_kind                                        */
    int32_t arg;        /* leaf index / stack slot / source dtype             */
    int32_t flags;      /* bit0: reversed operands                            */
    double  fimm;       /* floating immediate (HostScalar, NullaryOps.h:105)  */
    int64_t iimm;       /* integer immediate                                  */
} cm_instr;

/* dst_sr/sc/sb are the destination's element strides (a column-major dense
 * destination has 1, rows, rows*cols). */
int cm_cwise_eval(cm_ctx_t ctx,
                  const cm_leaf* leaves, int nleaves,
                  const cm_instr* code, int ncode,
                  void* dst, int dst_dtype,
                  int rows, int cols, int batches,
                  int64_t dst_sr, int64_t dst_sc, int64_t dst_sb,
                  int mode);

/* No counterpart in the reference (cuMat compiles one kernel per expression
 * type at build time). Large component-wise evaluations (cm_cwise_eval on a
 * destination and leaves of one contiguous layout) run a kernel generated for
 * the program and compiled at run time with hipRTC, cached per device by the
 * program's signature; results are bit-identical to the built-in
 * interpreter. mode: 0 never, 1 from 2^18 elements on (default; the
 * environment variable CUMAT_JIT sets the initial value), 2 always; any other
 * value only queries. Returns the previous mode. */
int cm_jit_mode(int mode);

/* Pre-instantiated fused linear combination out = sum_i coeff[i]*vec[i]
 * (the expression of benchmarks/linear_combination/Implementation_cuMat.cu,
 * evaluated by CwiseEvaluationKernel in the reference). k <= 16. */
int cm_lincomb_f32(cm_ctx_t ctx, int k, const float* const* vecs,
                   const float* coeffs, float* out, int64_t n);

/* ======================================================================
 * Reductions. Replace ReductionEvaluator<...>::eval / Auto selection
 * (cuMat/src/ReductionOps.h:45-49,167-534) and the full reduction
 * (cuMat/src/ReductionOps.h:564-598). 'axis' is a cm_axis bitmask (0..7);
 * the output keeps the non-reduced dimensions and has size 1 on reduced
 * ones, in the layout given by out_row_major. Reductions are deterministic:
 * the same input always gives bit-identical output.
 * ====================================================================== */
int cm_reduce(cm_ctx_t ctx, int dtype, int op, int axis,
              int rows, int cols, int batches, int in_row_major,
              const void* in, int out_row_major, void* out);

/* Reduction whose input is an unevaluated component-wise expression: the
 * program of cm_cwise_eval (result dtype 'dtype': F32, F64, I32 or BOOL;
 * bitwise operators take I32 / BOOL) is evaluated inside the reduction, so no temporary is written -- cuMat's
 * lazy input iterator (cuMat/src/Iterator.h:99-105) and the fused
 * cwiseDot(...).sum() of dot (cuMat/src/ReductionOpsPlugin.inl:290-297,
 * cuMat/src/ReductionOps.h:564-598). The expression's logical size is
 * rows x cols x batches (leaf strides as for cm_cwise_eval, 0 on broadcast
 * axes); 'op' and 'axis' as in cm_reduce; the output keeps the non-reduced
 * dimensions in the layout given by out_row_major. Deterministic. */
int cm_reduce_expr(cm_ctx_t ctx, const cm_leaf* leaves, int nleaves,
                   const cm_instr* code, int ncode, int dtype, int op, int axis,
                   int rows, int cols, int batches, int out_row_major,
                   void* out);

/* dot / squaredNorm per batch (cuMat/src/ReductionOpsPlugin.inl:395-431):
 * out[b] = sum_i a[i + n*b] * b[i + n*b]; a_batched/b_batched = 0 broadcasts
 * a batch-1 operand. Result stays on the device (batches values). */
int cm_dot(cm_ctx_t ctx, int dtype, int64_t n, int batches,
           const void* a, int a_batched, const void* b, int b_batched,
           void* out);
int cm_squared_norm(cm_ctx_t ctx, int dtype, int64_t n, int batches,
                    const void* a, void* out);

/* No counterpart in the reference. cm_dot for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): same arguments, kernels, summation order
 * and results as cm_dot, but the partial sums of a split reduction live in a
 * fixed 64 KiB device buffer of the library (one per GPU, kept until the
 * process ends) instead of the caching allocator; it calls no hipMalloc,
 * hipFree or allocating routine. ctx must come from cm_ctx_create (NULL is
 * rejected: the thread's default context would be created on first use).
 * Every call on one GPU shares that buffer: calls that overlap on a GPU (two
 * streams at once) race and must be serialized by the caller. A reduction
 * needing more than 64 KiB of partials is rejected (CM_E_DIM); the dot
 * product of one batch needs at most 4096 x 8 bytes. */
int cm_dot_noalloc(cm_ctx_t ctx, int dtype, int64_t n, int batches,
                   const void* a, int a_batched, const void* b, int b_batched,
                   void* out);

/* ======================================================================
 * Sparse CSR matrix-vector product. Replaces CSRMVKernel_StaticBatches +
 * ProductAssignment<...SparseMatrix<CSR>...>
 * (cuMat/src/SparseProductEvaluation.h:43-139) and the SparseExpressionOp
 * variant (cuMat/src/SparseProductEvaluation.h:144-187).
 * y[row + rows*b] (mode)= sum_k vals[k + nnz*b'] * x[IA[k] + cols*b''] with
 * b' = vals_batched ? b : 0 and b'' = x_batched ? b : 0.
 * (val_dtype, vec_dtype) in {(I32,I32), (F32,F32), (F64,F64), (BF16,F32)}.
 * An empty row assigns 0 (the reference leaves such rows untouched,
 * cuMat/src/SparseProductEvaluation.h:57).
 * ====================================================================== */
int cm_csr_spmv(cm_ctx_t ctx, int val_dtype, int vec_dtype,
                int rows, int cols, int nnz, int batches,
                const int* JA, const int* IA,
                const void* vals, int vals_batched,
                const void* x, int x_batched,
                void* y, int mode);

/* ELLPACK SpMV (cuMat/src/SparseProductEvaluation.h:194-324; pattern
 * cuMat/src/SparseMatrixBase.h:149-206): 'indices' is a column-major
 * rows x nnz_per_row int32 matrix, -1 marks padding. */
int cm_ell_spmv(cm_ctx_t ctx, int dtype, int rows, int cols, int nnz_per_row,
                int batches, const int* indices, const void* vals,
                int vals_batched, const void* x, int x_batched, void* y,
                int mode);

/* Block-CSR SpMV over block scalars: ProductAssignment of
 * SparseMatrix<real3x3> * Matrix<real3> (cuMat/src/SparseProductEvaluation.h:
 * 43-139) with the block plug-ins of tests/TestBlockedConjugateGradient.cu
 * (ProductElementFunctor :144-157, dot :17-20, types :12-64). A block is
 * block_dim (2..4) float4 rows (16 * block_dim bytes; real3x3 is three),
 * a vector entry one float4; rows, cols and nnz count blocks. For block
 * row i: y_i (mode)= sum over k of A_k x_{IA[k]} in CSR order, component q
 * of A_k x_j = a_q.x*x.x + a_q.y*x.y + a_q.z*x.z + a_q.w*x.w (not fused),
 * components >= block_dim = 0. Batches as cm_csr_spmv (value stride nnz
 * blocks, vector stride cols entries); mode CM_ASSIGN .. CM_DIV per
 * component; an empty block row assigns 0. vals, x and y 16-byte aligned. */
int cm_bcsr_spmv(cm_ctx_t ctx, int block_dim, int rows, int cols, int nnz,
                 int batches, const int* JA, const int* IA,
                 const float* vals, int vals_batched,
                 const float* x, int x_batched, float* y, int mode);

/* Jacobi entries of a block matrix, A.diagonal().cwiseInverseCheck() with
 * the block diagonal extraction and checked inverse
 * (cuMat/src/ConjugateGradient.h:218-246; tests/TestBlockedConjugateGradient.cu:
 * 174-210): one float4 per block row, component q < block_dim = 1/d_q of the
 * row's diagonal block (1 where d_q == 0 or no diagonal block is stored),
 * the rest 0; one row of entries per batch when vals_batched. */
int cm_bcsr_inverse_diagonal(cm_ctx_t ctx, int block_dim, int rows, int nnz,
                             int batches, const int* JA, const int* IA,
                             const float* vals, int vals_batched,
                             float* invdiag);

/* Product whose sparse matrix is a component-wise expression, evaluated
 * lazily inside the product's value loads: ONE launch, no value array
 * written. Replaces ProductAssignment over SparseExpressionOp<CSR>
 * (cuMat/src/SparseProductEvaluation.h:144-187; SparseExpressionOp,
 * cuMat/src/SparseExpressionOp.h:32-72; sparseView,
 * cuMat/src/SparseExpressionOpPlugin.inl:54-57).
 *   y[row + rows*b] (mode)= sum_k v(row, IA[k], b, k) * x[IA[k] + cols*b'']
 * where v is the program of cm_cwise_eval evaluated at stored entry k of
 * row 'row' in batch b; a leaf with sr == -1 is a value array of this
 * pattern (SparseMatrix::direct(), cuMat/src/SparseMatrix.h:230-233), read
 * at element k + b*sb; other leaves are dense strided leaves at
 * (row, IA[k], b). The program's result dtype must equal 'dtype' (the type
 * of x and y: F32, F64 or I32); mode is CM_ASSIGN, CM_ADD or CM_SUB.
 * Rows are summed in CSR order. */
int cm_csr_expr_spmv(cm_ctx_t ctx, const cm_leaf* leaves, int nleaves,
                     const cm_instr* code, int ncode, int dtype,
                     int rows, int cols, int nnz, int batches,
                     const int* JA, const int* IA,
                     const void* x, int x_batched, void* y, int mode);

/* The ELLPACK form (ProductAssignment over SparseExpressionOp<ELLPACK>,
 * cuMat/src/SparseProductEvaluation.h:261-324): slot (row, ci) has column
 * indices[row + rows*ci] (< 0: padding, skipped); a value leaf (sr == -1)
 * is read at row + rows*ci + b*sb (the column-major rows x nnz_per_row
 * data matrix); slots are added in order ci = 0, 1, ... */
int cm_ell_expr_spmv(cm_ctx_t ctx, const cm_leaf* leaves, int nleaves,
                     const cm_instr* code, int ncode, int dtype,
                     int rows, int cols, int nnz_per_row, int batches,
                     const int* indices,
                     const void* x, int x_batched, void* y, int mode);

/* Dense matrix-vector product y = A x per batch, the matrix-vector case of
 * the dense ProductOp, which cuMat evaluates with cuBLAS gemm
 * (ProductAssignment, cuMat/src/ProductOp.h:318-420). A is rows x cols,
 * column-major (cuMat's default) or row-major (a_row_major = 1), batch
 * stride rows*cols if a_batched, else shared by all batches; x has cols
 * entries per batch (shared if x_batched = 0); y has rows per batch.
 * mode: CM_ASSIGN, CM_ADD or CM_SUB. */
int cm_dense_gemv(cm_ctx_t ctx, int dtype, int rows, int cols, int batches,
                  const void* A, int a_batched, int a_row_major,
                  const void* x, int x_batched, void* y, int mode);

/* Jacobi inverse diagonal of a dense n x n matrix: inv[i] = A(i,i) == 0 ?
 * 1 : 1 / A(i,i) (ExtractDiagonalOp + cwiseInverseCheck,
 * cuMat/src/ConjugateGradient.h:228-230, cuMat/src/UnaryOps.h:177);
 * 'batches' rows of n if a_batched, else one. */
int cm_dense_inverse_diagonal(cm_ctx_t ctx, int dtype, int n, int batches,
                              const void* A, int a_batched, void* inv);

/* Component-wise evaluation into the non-zeros of a CSR matrix
 * (CwiseCSREvaluationKernel, cuMat/src/SparseEvaluation.h:14-34,83-138):
 * the program is evaluated at (row, IA[k], batch) for every stored k and
 * written to vals[k + nnz*batch] with 'mode'. */
int cm_csr_cwise_eval(cm_ctx_t ctx,
                      const cm_leaf* leaves, int nleaves,
                      const cm_instr* code, int ncode,
                      int rows, int cols, int nnz, int batches,
                      const int* JA, const int* IA,
                      void* vals, int val_dtype, int mode);

/* Jacobi setup: invdiag[i] = A_ii == 0 ? 1 : 1/A_ii
 * (ExtractDiagonalOp over SparseMatrix::coeff + cwiseInverseCheck,
 * cuMat/src/ConjugateGradient.h:228-230, cuMat/src/UnaryOps.h:177,
 * cuMat/src/SparseMatrixBase.h:258-271). Works for F32/F64 values. */
int cm_csr_inverse_diagonal(cm_ctx_t ctx, int dtype, int rows, int nnz,
                            int batches, const int* JA, const int* IA,
                            const void* vals, int vals_batched, void* invdiag);

/* ======================================================================
 * Jacobi-preconditioned Conjugate Gradient on a CSR matrix.
 * Replaces ConjugateGradient<SparseMatrix<S,1,CSR>,
 * DiagonalPreconditioner>::_solve_with_guess_impl
 * (cuMat/src/ConjugateGradient.h:123-206) and _solve_impl (114-121).
 * Semantics are the reference's: threshold tol^2*|b|^2 per batch, early outs
 * (zero rhs -> x = 0, 0 iterations, error 0; guess already converged ->
 * 0 iterations), all batches iterate until every batch has converged, the
 * converging iteration is not counted, error = sqrt(max_b |r|^2/|b|^2).
 * The matrix has batch 1 (broadcast), as the reference's preconditioner
 * requires (cuMat/src/ConjugateGradient.h:223). invdiag may be NULL (it is
 * then computed). x holds the guess on entry if use_guess, else is zeroed.
 * iterations_out / error_out are host pointers.
 * ====================================================================== */
/* Launch shapes: AUTO picks RESIDENT for one small system (n <= 6144),
 * FUSED up to 2^17 unknowns (n x batches, where an iteration is bound by
 * the launches) and MULTIKERNEL beyond. MULTIKERNEL: three launches per iteration (SpMV,
 * x/r update, p update). FUSED: two launches per iteration (the p update is
 * folded into the SpMV launch, each gathered p value being recomputed from
 * r, M^-1 and the previous p). RESIDENT: the whole solve in one workgroup
 * (n <= 16384, one right-hand side). All give the reference's semantics. */
enum cm_cg_path { CM_CG_AUTO = 0, CM_CG_MULTIKERNEL = 1, CM_CG_RESIDENT = 2, CM_CG_FUSED = 3 };

typedef struct cm_cg_stats {
    int32_t path_used;        /* cm_cg_path that ran                         */
    int32_t launches;         /* kernels issued for the solve                */
    double  seconds;          /* host wall time of the solve                 */
} cm_cg_stats;

int cm_pcg_solve(cm_ctx_t ctx, int dtype,
                 int n, int nnz, int batches,
                 const int* JA, const int* IA, const void* vals,
                 const void* invdiag,
                 const void* b, void* x, int use_guess,
                 double tol, int max_iter, int path,
                 int* iterations_out, double* error_out,
                 cm_cg_stats* stats);

/* Extension of cm_pcg_solve (multi-kernel path, batches <= 64) with no
 * reference counterpart: vals_batched = 1 gives every batch its own matrix
 * values (stride nnz, same pattern; invdiag, if given, has stride n), and
 * per_batch = 1 lets each batch stop at its own convergence (each then has
 * the plain solver's semantics on its own; iterations_each / error_each,
 * host arrays of 'batches' entries, receive them). With per_batch = 0 the
 * reference's rule holds: all batches iterate until all have converged. */
int cm_pcg_solve_ex(cm_ctx_t ctx, int dtype,
                    int n, int nnz, int batches,
                    const int* JA, const int* IA, const void* vals,
                    int vals_batched, const void* invdiag,
                    const void* b, void* x, int use_guess,
                    double tol, int max_iter, int per_batch,
                    int* iterations_each, double* error_each,
                    cm_cg_stats* stats);

/* Jacobi-PCG with a dense n x n matrix: ConjugateGradient<MatrixXf> /
 * <MatrixXd> (cuMat/src/ConjugateGradient.h:66-207 with the dense product
 * A * p, cuMat/src/ProductOp.h:318-420), same semantics and outputs as
 * cm_pcg_solve (multi-kernel path, batches <= 64). A is column-major or
 * row-major (a_row_major = 1), one matrix for all batches unless a_batched
 * (stride n*n); invdiag (n, or n per batch if a_batched) may be NULL, the
 * solver then derives it from A. */
int cm_pcg_dense_solve(cm_ctx_t ctx, int dtype, int n, int batches,
                       const void* A, int a_batched, int a_row_major,
                       const void* invdiag, const void* b, void* x,
                       int use_guess, double tol, int max_iter,
                       int* iterations_out, double* error_out,
                       cm_cg_stats* stats);

/* Jacobi-PCG over block scalars: ConjugateGradient<SparseMatrix<real3x3>>
 * (cuMat/src/ConjugateGradient.h:47-61,66-207 with the plug-ins of
 * tests/TestBlockedConjugateGradient.cu:111-239; the reference's test
 * :244-275). Blocks and vectors as cm_bcsr_spmv; b and x hold 'rows'
 * float4 entries per batch. Vector operations act per component, dot
 * products sum all four components, norms the first block_dim (components
 * >= block_dim of b are ignored; those of x keep the guess's values, 0
 * without a guess), and the Jacobi entries are cm_bcsr_inverse_diagonal's
 * (invdiag may be NULL: the solver derives them). Semantics and outputs as
 * cm_pcg_solve (multi-kernel path, batches <= 64, all batches iterate until
 * all converge; max_iter < 0: 2 * rows); vals_batched as cm_pcg_solve_ex. */
int cm_pcg_bcsr_solve(cm_ctx_t ctx, int block_dim, int rows, int nnz,
                      int batches, const int* JA, const int* IA,
                      const float* vals, int vals_batched,
                      const float* invdiag, const float* b, float* x,
                      int use_guess, double tol, int max_iter,
                      int* iterations_out, double* error_out,
                      cm_cg_stats* stats);

/* ======================================================================
 * Addition 2 (no reference counterpart): persistent batched CG over many
 * independent, uneven CSR systems with per-system convergence. System s
 * owns rows [row_off[s], row_off[s+1]) and non-zeros starting at
 * JA[row_off[s]] of one concatenated block-diagonal CSR (column indices are
 * global). Each system follows the plain solver's semantics on its own and
 * stops when it converges; iterations[s] and errors[s] are device arrays.
 * block_size == 0 takes the shape cm_pcg_batched_autotune chose for this
 * batch shape (or a default when it was not tuned); otherwise the block
 * size is block_size and the grid as many workgroups as fit.
 * ====================================================================== */
int cm_pcg_batched_solve(cm_ctx_t ctx, int nsys, const int* row_off,
                         int total_rows, int total_nnz,
                         const int* JA, const int* IA, const float* vals,
                         const float* b, float* x, float tol, int max_iter,
                         int* iterations, float* errors,
                         int block_size, int use_graph);
/* Autotuner: times the candidate launch shapes on the given batch -- threads
 * per system (256, 512, 1024; the register tile follows from the largest
 * system) x workgroups per CU of the persistent grid (1, 2, 4) -- and
 * returns the best (host ints; either pointer may be NULL). A later
 * cm_pcg_batched_solve of the same batch shape with block_size == 0 uses
 * that shape. */
int cm_pcg_batched_autotune(cm_ctx_t ctx, int nsys, const int* row_off,
                            int total_rows, int total_nnz, const int* JA,
                            const int* IA, const float* vals, const float* b,
                            float tol, int max_iter, int* best_block_size,
                            int* best_groups_per_cu, float* best_ms);

/* ======================================================================
 * Addition 1 (no reference counterpart): mixed-precision CG with FP64
 * iterative refinement. The outer loop keeps x, r = b - A x and the
 * convergence test in FP64 on the device; each inner solve is FP32 Jacobi
 * CG on A e = r with FP32 vectors and FP32 (val_dtype CM_F32) or BF16
 * (CM_BF16) matrix storage. BF16 storage keeps the diagonal in FP32 and
 * folds the rounding of each row's off-diagonals into it, so every row sum
 * of the stored matrix equals the row sum of A. The outer loop stops at the
 * FP64 relative residual tol, after max_outer steps, or when a step no
 * longer halves the FP64 residual (the rounding floor of the residual
 * itself; final_rel_residual then reports where it stopped).
 * ====================================================================== */
typedef struct cm_ir_stats {
    int32_t outer_iterations;
    int32_t inner_iterations;     /* summed over outer steps                 */
    double  final_rel_residual;   /* |b - A x|_2 / |b|_2 evaluated in FP64   */
    double  seconds;
} cm_ir_stats;

int cm_pcg_ir_solve(cm_ctx_t ctx, int n, int nnz,
                    const int* JA, const int* IA, const double* vals64,
                    int val_dtype, const double* b, double* x,
                    double tol, double inner_tol, int max_outer,
                    int max_inner, cm_ir_stats* stats);

/* ======================================================================
 * Addition 3 (no reference counterpart): adjoint of a CG solve on a
 * variable-coefficient edge Laplacian. Edge e joins nodes ei[e] and ej[e]
 * (ej[e] < 0 marks a Dirichlet boundary edge touching only ei[e]).
 *  - cm_edge_assemble writes the CSR values of A(theta) by a deterministic
 *    gather: non-zero k receives sum_{t in [eptr[k], eptr[k+1])}
 *    sign(eidx[t]) * theta[|eidx[t]| - 1], where eidx = e+1 adds theta_e
 *    (diagonal entries) and -(e+1) subtracts it (off-diagonal entries), so
 *    A_ii = sum of incident theta_e and A_ij = -theta_e.
 *  - cm_edge_grad writes dL/dtheta_e = -(lam_i - lam_j) * (x_i - x_j)
 *    (boundary edge: -lam_i * x_i) for every edge and batch.
 * Batched arrays are stored batch after batch (theta: nedges, values: nnz,
 * lam/x: n per batch).
 * ====================================================================== */
int cm_edge_assemble(cm_ctx_t ctx, int dtype, int nnz, int batches, int nedges,
                     const int* eptr, const int* eidx, const void* theta,
                     void* vals);
int cm_edge_grad(cm_ctx_t ctx, int dtype, int nedges, int batches, int n,
                 const int* ei, const int* ej, const void* lam,
                 const void* x, void* grad);

#ifdef __cplusplus
}
#endif
#endif /* CUMAT_AMD_H */
