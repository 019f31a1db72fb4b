// dense.hip -- C entry points of the dense matrix-vector product and the
// Jacobi diagonal of a dense matrix (kernels in dense.hpp).
#include "dense.hpp"

namespace cm {

template <typename T>
static int launch_dense_gemv(cm_ctx_t ctx, int rows, int cols, int batches, const void* A, int a_batched,
                             int a_row_major, const void* x, int x_batched, void* y, int mode)
{
    const long long bstride = a_batched ? (long long)rows * cols : 0;
    // column-major with rows a multiple of the vector width and a 16-byte
    // aligned base: 16-byte loads (16384^2 FP32: see dense_gemv_v_kernel)
    constexpr int VW = DenseVec<T>::w;
    const bool vec = !a_row_major && rows % VW == 0 && (((uintptr_t)A & 15) == 0);
    const int tile = vec ? 64 * VW : DN_TILE;
    int nsplit = 1, csz = cols;
    dense_split(rows, cols, nsplit, csz, tile);
    // per-range partial vectors in the context's stream-ordered scratch
    void* part = nullptr;
    if (nsplit > 1) CM_TRY(scratch(ctx, 1, sizeof(T) * (size_t)nsplit * rows * batches, &part));
    dim3 grid(dense_grid(rows, nsplit, tile), batches);
    if (vec)
        dense_gemv_v_kernel<T><<<grid, DN_BS, 0, ctx->stream>>>(rows, cols, (const T*)A, bstride, (const T*)x,
                                                                x_batched, (T*)y, mode, nsplit, csz, (T*)part);
    else if (a_row_major)
        dense_gemv_kernel<T, true><<<grid, DN_BS, 0, ctx->stream>>>(rows, cols, (const T*)A, bstride, (const T*)x,
                                                                    x_batched, (T*)y, mode, nsplit, csz, (T*)part);
    else
        dense_gemv_kernel<T, false><<<grid, DN_BS, 0, ctx->stream>>>(rows, cols, (const T*)A, bstride, (const T*)x,
                                                                     x_batched, (T*)y, mode, nsplit, csz, (T*)part);
    CM_LAUNCH_CHECK();
    if (nsplit > 1) {
        dense_gemv_combine_kernel<T><<<dim3(div_up(rows, 256), batches), 256, 0, ctx->stream>>>(
            rows, nsplit, (const T*)part, (T*)y, mode);
        CM_LAUNCH_CHECK();
    }
    return CM_OK;
}

}  // namespace cm

extern "C" {

int cm_dense_gemv(cm_ctx_t ctx, int dtype, int rows, int cols, int batches, const void* A, int a_batched,
                  int a_row_major, const void* x, int x_batched, void* y, int mode)
{
    CM_API_BEGIN
    CM_TRY(cm::resolve(ctx, &ctx));
    CM_REQUIRE(rows >= 0 && cols >= 0 && batches >= 1, CM_E_DIM, "cm_dense_gemv: bad sizes rows=%d cols=%d batches=%d",
               rows, cols, batches);
    CM_REQUIRE(mode == CM_ASSIGN || mode == CM_ADD || mode == CM_SUB, CM_E_ARG,
               "cm_dense_gemv: mode %d (=, += and -= only)", mode);
    cm::count(CM_CNT_EVAL_MATMUL);
    cm::count(CM_CNT_EVAL_ANY);
    if (rows == 0) return CM_OK;
    CM_REQUIRE(y && (cols == 0 || (A && x)), CM_E_ARG, "cm_dense_gemv: NULL pointer");
    if (dtype == CM_F32)
        return cm::launch_dense_gemv<float>(ctx, rows, cols, batches, A, a_batched, a_row_major, x, x_batched, y,
                                            mode);
    if (dtype == CM_F64)
        return cm::launch_dense_gemv<double>(ctx, rows, cols, batches, A, a_batched, a_row_major, x, x_batched, y,
                                             mode);
    if (dtype == CM_I32)
        return cm::launch_dense_gemv<int>(ctx, rows, cols, batches, A, a_batched, a_row_major, x, x_batched, y,
                                          mode);
    return cm::fail(CM_E_UNSUPPORTED, "cm_dense_gemv: dtype %d", dtype);
    CM_API_END
}

int cm_dense_inverse_diagonal(cm_ctx_t ctx, int dtype, int n, int batches, const void* A, int a_batched, void* inv)
{
    CM_API_BEGIN
    CM_TRY(cm::resolve(ctx, &ctx));
    CM_REQUIRE(n >= 0 && batches >= 1, CM_E_DIM, "cm_dense_inverse_diagonal: bad sizes");
    if (n == 0) return CM_OK;
    CM_REQUIRE(A && inv, CM_E_ARG, "cm_dense_inverse_diagonal: NULL pointer");
    const int nb = a_batched ? batches : 1;
    const long long bstride = a_batched ? (long long)n * n : 0;
    dim3 grid(cm::div_up(n, 256), nb);
    if (dtype == CM_F32)
        cm::dense_inverse_diagonal_kernel<float><<<grid, 256, 0, ctx->stream>>>(n, (const float*)A, bstride,
                                                                                (float*)inv);
    else if (dtype == CM_F64)
        cm::dense_inverse_diagonal_kernel<double><<<grid, 256, 0, ctx->stream>>>(n, (const double*)A, bstride,
                                                                                 (double*)inv);
    else
        return cm::fail(CM_E_UNSUPPORTED, "cm_dense_inverse_diagonal: dtype %d", dtype);
    CM_LAUNCH_CHECK();
    return CM_OK;
    CM_API_END
}

}  // extern "C"
