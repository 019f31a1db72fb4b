// spmv.hip -- CSR and ELLPACK sparse matrix-vector products and the Jacobi
// diagonal extraction.
//   CSR:     replaces CSRMVKernel_StaticBatches + ProductAssignment
//            (cuMat/src/SparseProductEvaluation.h:43-139)
//   ELLPACK: replaces ELLPACKMVKernel_StaticBatches
//            (cuMat/src/SparseProductEvaluation.h:194-324)
//   Jacobi:  replaces ExtractDiagonalOp over SparseMatrix::coeff +
//            cwiseInverseCheck (cuMat/src/ConjugateGradient.h:228-230)
#include "bcsr.hpp"
#include "spmv_core.hpp"
#include "spmv_wave.hpp"

namespace cm {

constexpr int SPMV_BS = 256;
constexpr int SPMV_EPT = 8;   // staged non-zeros per lane (8 per row on average)

template <typename VT, typename T>
__global__ __launch_bounds__(SPMV_BS) void csr_spmv_kernel(int rows, int cols, int nnz,
                                                           const int* __restrict__ JA,
                                                           const int* __restrict__ IA,
                                                           const VT* __restrict__ vals, int vals_batched,
                                                           const T* __restrict__ x, int x_batched,
                                                           T* __restrict__ y, int mode)
{
    constexpr int WPB = SPMV_BS / 64;
    __shared__ T lds[WPB][SPMV_EPT * 64];
    T* slice = lds[threadIdx.x >> 6];
    const int lane = threadIdx.x & 63;
    const int b = blockIdx.y;
    const VT* v = vals + (vals_batched ? (long long)b * nnz : 0);
    const T* xb = x + (x_batched ? (long long)b * cols : 0);
    T* yb = y + (long long)b * rows;
    ChunkWalk w(div_up_dev(rows, 64), WPB);
    if (!w.valid()) return;  // wave-uniform; no block barrier below
    WaveChunk<VT, SPMV_EPT> cur, nxt;
    chunk_load(cur, JA, IA, v, w.cur * 64, rows);
    for (;;) {
        const int nc = w.peek_next();
        const bool more = nc < w.hi;
        if (more) chunk_load(nxt, JA, IA, v, nc * 64, rows);  // in flight during the gathers below
        const T acc = chunk_compute<VT, T, SPMV_EPT>(cur, JA, IA, v, [xb](int c) { return xb[c]; }, slice);
        if (lane < cur.nr) {
            T* yy = yb + cur.r0 + lane;
            if (mode == CM_ASSIGN) __builtin_nontemporal_store(acc, yy);
            else *yy = assign_mode<T>(mode, *yy, acc);
        }
        if (!more) break;
        cur = nxt;
        w.next();
    }
}

// Rows of at most MAXW slots: every slot's index and value load is issued
// at once, then the gathers, then the adds in slot order (the same sums as
// ell_spmv_kernel; 2048^2 width 5: 36.3 -> 34.6 us).
template <typename T, int MAXW>
__global__ void ell_spmv_w_kernel(int rows, int cols, int npr, const int* __restrict__ idx,
                                  const T* __restrict__ vals, int vals_batched, const T* __restrict__ x,
                                  int x_batched, T* __restrict__ y, int mode)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const int b = blockIdx.y;
    const T* v = vals + (vals_batched ? (long long)b * rows * npr : 0);
    const T* xb = x + (x_batched ? (long long)b * cols : 0);
    int c[MAXW];
    T a[MAXW], g[MAXW];
#pragma unroll
    for (int j = 0; j < MAXW; ++j) {
        const bool ok = j < npr;  // uniform
        c[j] = ok ? __builtin_nontemporal_load(idx + r + (long long)rows * j) : -1;
        a[j] = ok ? __builtin_nontemporal_load(v + r + (long long)rows * j) : T(0);
    }
#pragma unroll
    for (int j = 0; j < MAXW; ++j) g[j] = c[j] >= 0 ? xb[c[j]] : T(0);
    T acc = T(0);
#pragma unroll
    for (int j = 0; j < MAXW; ++j)
        if (c[j] >= 0) acc += a[j] * g[j];
    T* yy = y + r + (long long)rows * b;
    *yy = (mode == CM_ASSIGN) ? acc : assign_mode<T>(mode, *yy, acc);
}

template <typename T>
__global__ void ell_spmv_kernel(int rows, int cols, int npr, const int* __restrict__ idx,
                                const T* __restrict__ vals, int vals_batched,
                                const T* __restrict__ x, int x_batched, T* __restrict__ y, int mode)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const int b = blockIdx.y;
    const T* v = vals + (vals_batched ? (long long)b * rows * npr : 0);
    const T* xb = x + (x_batched ? (long long)b * cols : 0);
    T acc = T(0);
    // column-major: the k-th entries of consecutive rows are contiguous.
    // Four entries' index and value loads are issued together (streamed
    // once: non-temporal), then gathered and added in k order, skipping
    // padding (index < 0)
    int k = 0;
    for (; k + 4 <= npr; k += 4) {
        int c[4];
        T a[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            c[j] = __builtin_nontemporal_load(idx + r + (long long)rows * (k + j));
            a[j] = __builtin_nontemporal_load(v + r + (long long)rows * (k + j));
        }
        T g[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) g[j] = c[j] >= 0 ? xb[c[j]] : T(0);
#pragma unroll
        for (int j = 0; j < 4; ++j)
            if (c[j] >= 0) acc += a[j] * g[j];
    }
    for (; k < npr; ++k) {
        const int c = idx[r + (long long)rows * k];
        if (c < 0) continue;
        acc += v[r + (long long)rows * k] * xb[c];
    }
    T* yy = y + r + (long long)rows * b;
    *yy = (mode == CM_ASSIGN) ? acc : assign_mode<T>(mode, *yy, acc);
}

template <typename T>
__global__ void csr_inverse_diagonal_kernel(int rows, int nnz, const int* __restrict__ JA,
                                            const int* __restrict__ IA, const T* __restrict__ vals,
                                            int vals_batched, T* __restrict__ inv)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const int b = blockIdx.y;
    const T* v = vals + (vals_batched ? (long long)b * nnz : 0);
    T d = T(0);
    for (int k = JA[r]; k < JA[r + 1]; ++k)
        if (IA[k] == r) { d = v[k]; break; }
    inv[r + (long long)rows * b] = (d == T(0)) ? T(1) : T(1) / d;
}

template <typename VT, typename T>
static int launch_csr(cm_ctx_t ctx, int rows, int cols, int nnz, int batches, const int* JA,
                      const int* IA, const void* vals, int vb, const void* x, int xb, void* y, int mode)
{
    // grid cap: 8 resident blocks per CU (the wave limit at 256 threads);
    // CUMAT_SPMV_GRID overrides it for tuning sweeps
    static const int gmax = [] {
        const char* s = getenv("CUMAT_SPMV_GRID");
        return (s && atoi(s) >= 8) ? atoi(s) : 2048;
    }();
    dim3 grid(walk_grid(div_up(div_up(rows, 64), SPMV_BS / 64), gmax), batches);
    csr_spmv_kernel<VT, T><<<grid, SPMV_BS, 0, ctx->stream>>>(rows, cols, nnz, JA, IA, (const VT*)vals, vb,
                                                              (const T*)x, xb, (T*)y, mode);
    CM_LAUNCH_CHECK();
    return CM_OK;
}

}  // namespace cm

extern "C" {

int cm_csr_spmv(cm_ctx_t ctx, int val_dtype, int vec_dtype, int rows, int cols, int nnz, int batches,
                const int* JA, const int* IA, const void* vals, int vals_batched, const void* x,
                int x_batched, void* y, int mode)
{
    CM_API_BEGIN
    CM_TRY(cm::resolve(ctx, &ctx));
    CM_REQUIRE(rows >= 0 && cols >= 0 && nnz >= 0 && batches >= 1, CM_E_DIM,
               "cm_csr_spmv: bad sizes rows=%d cols=%d nnz=%d batches=%d", rows, cols, nnz, batches);
    CM_REQUIRE(mode >= CM_ASSIGN && mode <= CM_OR, CM_E_ARG, "cm_csr_spmv: bad mode %d", mode);
    cm::count(CM_CNT_EVAL_MATMUL_SPARSE);
    cm::count(CM_CNT_EVAL_ANY);
    if (rows == 0) return CM_OK;
    CM_REQUIRE(JA && y && (nnz == 0 || (IA && vals && x)), CM_E_ARG, "cm_csr_spmv: NULL pointer");
    if (val_dtype == CM_F32 && vec_dtype == CM_F32)
        return cm::launch_csr<float, float>(ctx, rows, cols, nnz, batches, JA, IA, vals, vals_batched, x,
                                            x_batched, y, mode);
    if (val_dtype == CM_F64 && vec_dtype == CM_F64)
        return cm::launch_csr<double, double>(ctx, rows, cols, nnz, batches, JA, IA, vals, vals_batched, x,
                                              x_batched, y, mode);
    if (val_dtype == CM_I32 && vec_dtype == CM_I32)
        return cm::launch_csr<int, int>(ctx, rows, cols, nnz, batches, JA, IA, vals, vals_batched, x,
                                        x_batched, y, mode);
    if (val_dtype == CM_BF16 && vec_dtype == CM_F32)
        return cm::launch_csr<unsigned short, float>(ctx, rows, cols, nnz, batches, JA, IA, vals,
                                                     vals_batched, x, x_batched, y, mode);
    return cm::fail(CM_E_UNSUPPORTED, "cm_csr_spmv: unsupported dtype pair (%d,%d)", val_dtype, vec_dtype);
    CM_API_END
}

int cm_ell_spmv(cm_ctx_t ctx, int dtype, int rows, int cols, int nnz_per_row, int batches,
                const int* indices, const void* vals, int vals_batched, const void* x, int x_batched,
                void* y, int mode)
{
    CM_API_BEGIN
    CM_TRY(cm::resolve(ctx, &ctx));
    CM_REQUIRE(rows >= 0 && cols >= 0 && nnz_per_row >= 0 && batches >= 1, CM_E_DIM,
               "cm_ell_spmv: bad sizes");
    CM_REQUIRE(mode >= CM_ASSIGN && mode <= CM_OR, CM_E_ARG, "cm_ell_spmv: bad mode %d", mode);
    cm::count(CM_CNT_EVAL_MATMUL_SPARSE);
    cm::count(CM_CNT_EVAL_ANY);
    if (rows == 0) return CM_OK;
    dim3 grid(cm::div_up(rows, 256), batches);
    if (dtype == CM_F32 && nnz_per_row <= 8) {
        cm::ell_spmv_w_kernel<float, 8><<<grid, 256, 0, ctx->stream>>>(rows, cols, nnz_per_row, indices,
                                                                        (const float*)vals, vals_batched,
                                                                        (const float*)x, x_batched, (float*)y, mode);
    } else if (dtype == CM_F32) {
        cm::ell_spmv_kernel<float><<<grid, 256, 0, ctx->stream>>>(rows, cols, nnz_per_row, indices,
                                                                   (const float*)vals, vals_batched,
                                                                   (const float*)x, x_batched, (float*)y, mode);
    } else if (dtype == CM_F64) {
        cm::ell_spmv_kernel<double><<<grid, 256, 0, ctx->stream>>>(rows, cols, nnz_per_row, indices,
                                                                    (const double*)vals, vals_batched,
                                                                    (const double*)x, x_batched, (double*)y,
                                                                    mode);
    } else if (dtype == CM_I32) {
        cm::ell_spmv_kernel<int><<<grid, 256, 0, ctx->stream>>>(rows, cols, nnz_per_row, indices,
                                                                 (const int*)vals, vals_batched, (const int*)x,
                                                                 x_batched, (int*)y, mode);
    } else {
        return cm::fail(CM_E_UNSUPPORTED, "cm_ell_spmv: unsupported dtype %d", dtype);
    }
    CM_LAUNCH_CHECK();
    return CM_OK;
    CM_API_END
}

int cm_csr_inverse_diagonal(cm_ctx_t ctx, int dtype, int rows, int nnz, int batches, const int* JA,
                            const int* IA, const void* vals, int vals_batched, void* invdiag)
{
    CM_API_BEGIN
    CM_TRY(cm::resolve(ctx, &ctx));
    CM_REQUIRE(rows >= 0 && batches >= 1, CM_E_DIM, "cm_csr_inverse_diagonal: bad sizes");
    cm::count(CM_CNT_EVAL_CWISE);
    cm::count(CM_CNT_EVAL_ANY);
    if (rows == 0) return CM_OK;
    dim3 grid(cm::div_up(rows, 256), vals_batched ? batches : 1);
    if (dtype == CM_F32)
        cm::csr_inverse_diagonal_kernel<float><<<grid, 256, 0, ctx->stream>>>(
            rows, nnz, JA, IA, (const float*)vals, vals_batched, (float*)invdiag);
    else if (dtype == CM_F64)
        cm::csr_inverse_diagonal_kernel<double><<<grid, 256, 0, ctx->stream>>>(
            rows, nnz, JA, IA, (const double*)vals, vals_batched, (double*)invdiag);
    else
        return cm::fail(CM_E_UNSUPPORTED, "cm_csr_inverse_diagonal: unsupported dtype %d", dtype);
    CM_LAUNCH_CHECK();
    return CM_OK;
    CM_API_END
}

int cm_bcsr_spmv(cm_ctx_t ctx, int block_dim, int rows, int cols, int nnz, int batches, const int* JA,
                 const int* IA, const float* vals, int vals_batched, const float* x, int x_batched, float* y,
                 int mode)
{
    CM_API_BEGIN
    CM_TRY(cm::resolve(ctx, &ctx));
    CM_REQUIRE(block_dim >= 2 && block_dim <= 4, CM_E_UNSUPPORTED, "cm_bcsr_spmv: block dimension %d (2..4)",
               block_dim);
    CM_REQUIRE(rows >= 0 && cols >= 0 && nnz >= 0 && batches >= 1 && batches <= 65535, CM_E_DIM,
               "cm_bcsr_spmv: bad sizes rows=%d cols=%d nnz=%d batches=%d (1..65535)", rows, cols, nnz, batches);
    CM_REQUIRE(mode >= CM_ASSIGN && mode <= CM_DIV, CM_E_ARG, "cm_bcsr_spmv: bad mode %d", mode);
    cm::count(CM_CNT_EVAL_MATMUL_SPARSE);
    cm::count(CM_CNT_EVAL_ANY);
    if (rows == 0) return CM_OK;
    CM_REQUIRE(JA && y && (nnz == 0 || (IA && vals && x)), CM_E_ARG, "cm_bcsr_spmv: NULL pointer");
    CM_REQUIRE((((uintptr_t)vals | (uintptr_t)x | (uintptr_t)y) & 15) == 0, CM_E_ARG,
               "cm_bcsr_spmv: blocks and vectors must be 16-byte aligned (float4 entries)");
    cm::launch_bcsr_spmv(ctx->stream, block_dim, rows, batches, JA, IA, (const cm::bc_f4*)vals,
                         vals_batched ? (long long)nnz * block_dim : 0, (const cm::bc_f4*)x,
                         x_batched ? (long long)cols : 0, (cm::bc_f4*)y, mode);
    CM_LAUNCH_CHECK();
    return CM_OK;
    CM_API_END
}

int cm_bcsr_inverse_diagonal(cm_ctx_t ctx, int block_dim, int rows, int nnz, int batches, const int* JA,
                             const int* IA, const float* vals, int vals_batched, float* invdiag)
{
    CM_API_BEGIN
    CM_TRY(cm::resolve(ctx, &ctx));
    CM_REQUIRE(block_dim >= 2 && block_dim <= 4, CM_E_UNSUPPORTED,
               "cm_bcsr_inverse_diagonal: block dimension %d (2..4)", block_dim);
    CM_REQUIRE(rows >= 0 && nnz >= 0 && batches >= 1 && batches <= 65535, CM_E_DIM,
               "cm_bcsr_inverse_diagonal: bad sizes");
    cm::count(CM_CNT_EVAL_CWISE);
    cm::count(CM_CNT_EVAL_ANY);
    if (rows == 0) return CM_OK;
    CM_REQUIRE(JA && invdiag && (nnz == 0 || (IA && vals)), CM_E_ARG, "cm_bcsr_inverse_diagonal: NULL pointer");
    CM_REQUIRE((((uintptr_t)vals | (uintptr_t)invdiag) & 15) == 0, CM_E_ARG,
               "cm_bcsr_inverse_diagonal: blocks and entries must be 16-byte aligned");
    cm::launch_bcsr_invdiag(ctx->stream, block_dim, rows, vals_batched ? batches : 1, JA, IA,
                            (const cm::bc_f4*)vals, vals_batched ? (long long)nnz * block_dim : 0,
                            (cm::bc_f4*)invdiag);
    CM_LAUNCH_CHECK();
    return CM_OK;
    CM_API_END
}

}  // extern "C"
