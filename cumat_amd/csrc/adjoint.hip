// adjoint.hip -- kernels of the differentiable CG (no reference counterpart):
// assembly of a variable-coefficient edge Laplacian A(theta) into CSR values
// by a deterministic gather, and the fused per-edge gradient
// dL/dtheta_e = -lam^T (dA/dtheta_e) x, where dA/dtheta_e =
// (u_i - u_j)(u_i - u_j)^T for an interior edge and u_i u_i^T for a
// Dirichlet boundary edge.
#include "cm_common.hpp"

namespace cm {

template <typename T>
__global__ void edge_assemble_kernel(int nnz, int nedges, const int* __restrict__ eptr, const int* __restrict__ eidx,
                                     const T* __restrict__ theta, T* __restrict__ vals)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nnz) return;
    const long long b = blockIdx.y;
    const T* th = theta + b * nedges;
    T acc = T(0);
    for (int t = eptr[k]; t < eptr[k + 1]; ++t) {
        int c = eidx[t];
        acc += (c > 0) ? th[c - 1] : -th[-c - 1];
    }
    vals[k + b * nnz] = acc;
}

template <typename T>
__global__ void edge_grad_kernel(int nedges, int n, const int* __restrict__ ei, const int* __restrict__ ej,
                                 const T* __restrict__ lam, const T* __restrict__ x, T* __restrict__ grad)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nedges) return;
    const long long b = blockIdx.y;
    const T* l = lam + b * n;
    const T* xx = x + b * n;
    const int i = ei[e], j = ej[e];
    T g;
    if (j >= 0) g = -(l[i] - l[j]) * (xx[i] - xx[j]);
    else g = -l[i] * xx[i];
    grad[e + b * nedges] = g;
}

}  // namespace cm

extern "C" {

int cm_edge_assemble(cm_ctx_t ctx, int dtype, int nnz, int batches, int nedges, const int* eptr, const int* eidx,
                     const void* theta, void* vals)
{
    CM_API_BEGIN
    CM_TRY(cm::resolve(ctx, &ctx));
    CM_REQUIRE(nnz >= 0 && batches >= 1 && nedges >= 0, CM_E_DIM, "cm_edge_assemble: bad sizes");
    if (nnz == 0) return CM_OK;
    CM_REQUIRE(eptr && eidx && theta && vals, CM_E_ARG, "cm_edge_assemble: NULL pointer");
    dim3 grid(cm::div_up(nnz, 256), batches);
    if (dtype == CM_F32)
        cm::edge_assemble_kernel<float><<<grid, 256, 0, ctx->stream>>>(nnz, nedges, eptr, eidx, (const float*)theta,
                                                                       (float*)vals);
    else if (dtype == CM_F64)
        cm::edge_assemble_kernel<double><<<grid, 256, 0, ctx->stream>>>(nnz, nedges, eptr, eidx,
                                                                        (const double*)theta, (double*)vals);
    else
        return cm::fail(CM_E_UNSUPPORTED, "cm_edge_assemble: dtype %d", dtype);
    CM_LAUNCH_CHECK();
    return CM_OK;
    CM_API_END
}

int cm_edge_grad(cm_ctx_t ctx, int dtype, int nedges, int batches, int n, const int* ei, const int* ej,
                 const void* lam, const void* x, void* grad)
{
    CM_API_BEGIN
    CM_TRY(cm::resolve(ctx, &ctx));
    CM_REQUIRE(nedges >= 0 && batches >= 1 && n >= 0, CM_E_DIM, "cm_edge_grad: bad sizes");
    if (nedges == 0) return CM_OK;
    CM_REQUIRE(ei && ej && lam && x && grad, CM_E_ARG, "cm_edge_grad: NULL pointer");
    dim3 grid(cm::div_up(nedges, 256), batches);
    if (dtype == CM_F32)
        cm::edge_grad_kernel<float><<<grid, 256, 0, ctx->stream>>>(nedges, n, ei, ej, (const float*)lam,
                                                                   (const float*)x, (float*)grad);
    else if (dtype == CM_F64)
        cm::edge_grad_kernel<double><<<grid, 256, 0, ctx->stream>>>(nedges, n, ei, ej, (const double*)lam,
                                                                    (const double*)x, (double*)grad);
    else
        return cm::fail(CM_E_UNSUPPORTED, "cm_edge_grad: dtype %d", dtype);
    CM_LAUNCH_CHECK();
    return CM_OK;
    CM_API_END
}

}  // extern "C"
