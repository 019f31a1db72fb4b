// bcsr.hpp -- block-CSR products over block scalars: D x D float blocks
// (D = 2..4) stored as D float4 rows, vector entries float4. These are the
// reference's block-scalar plug-in types (real3x3 = three real3 rows,
// real3 = float4 with w = 0; tests/TestBlockedConjugateGradient.cu:12-64),
// whose product functor gives component q of A_ij x_j as the four-term dot
// of block row q with x_j (tests/TestBlockedConjugateGradient.cu:144-157,
// dot :17-20); components >= D of a product are 0.
//
// Wave-chunk walk: a wavefront owns 64 consecutive block rows. Its lanes
// take the chunk's blocks 16 at a time, four lanes per block: lane 4g + q
// loads row q of block g (16 bytes: the 16 blocks of a step are one
// contiguous run of 16*D float4), gathers the block's x entry (16 bytes,
// one address for the four lanes), and parks its dot -- component q of the
// block's product -- in the wavefront's LDS slice, where the block's four
// components form one float4. Then lane r sums block row r's products in
// CSR order with 16-byte LDS reads: the first product, then += (the
// reference's row loop, cuMat/src/SparseProductEvaluation.h:55-87). Chunks
// with more than BC_WIN blocks are taken in windows of BC_WIN blocks, each
// lane carrying its row's sum from one window to the next.
// The dots are not contracted into fused multiply-adds, so each product
// (and so each row sum) has the same bits as the sequential CPU loop.
#pragma once
#include "cm_common.hpp"
#include "spmv_core.hpp"
#include "spmv_wave.hpp"

namespace cm {

typedef float bc_f4 __attribute__((ext_vector_type(4)));

constexpr int BC_WIN = 256;  // blocks per LDS window (4 KB per wavefront)
constexpr int BC_U = 4;      // 16-block steps whose loads are issued together

__device__ __forceinline__ float bc_dot(bc_f4 a, bc_f4 x)
{
#pragma clang fp contract(off)
    return a.x * x.x + a.y * x.y + a.z * x.z + a.w * x.w;
}

template <bool NT>
__device__ __forceinline__ bc_f4 bc_load(const bc_f4* p)
{
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}

// (A x) of block row r0 + lane for lane < nr (0 elsewhere); A's block k is
// vals[k*D .. k*D + D-1]; xg(c) yields x's entry c. slice: BC_WIN float4 of
// this wavefront's LDS. NT: stream the matrix (a single product) rather
// than keep it cached (an iterative solver re-reads it).
template <int D, bool NT, typename XG>
__device__ __forceinline__ bc_f4 bcsr_chunk_product(const int* __restrict__ JA, const int* __restrict__ IA,
                                                    const bc_f4* __restrict__ vals, int r0, int nr, XG xg,
                                                    bc_f4* slice)
{
    static_assert(D >= 2 && D <= 4, "blocks of 2 to 4 float4 rows");
    const int lane = threadIdx.x & 63;
    const int q = lane & 3, g = lane >> 2;
    const int s = JA[r0], e = JA[r0 + nr];
    const int rb = lane < nr ? JA[r0 + lane] : 0;
    const int re = lane < nr ? JA[r0 + lane + 1] : 0;
    float* sf = reinterpret_cast<float*>(slice);
    bc_f4 acc = bc_f4(0.0f);
    for (int w0 = s; w0 < e; w0 += BC_WIN) {
        const int w1 = min(e, w0 + BC_WIN);
        for (int k0 = w0; k0 < w1; k0 += 16 * BC_U) {
            int col[BC_U];
            bc_f4 a[BC_U];
#pragma unroll
            for (int u = 0; u < BC_U; ++u) {
                const int k = k0 + 16 * u + g;
                const bool ok = k < w1;
                col[u] = ok ? IA[k] : 0;
                a[u] = (ok && q < D) ? bc_load<NT>(vals + (long long)k * D + q) : bc_f4(0.0f);
            }
            bc_f4 xv[BC_U];
#pragma unroll
            for (int u = 0; u < BC_U; ++u) {
                const int k = k0 + 16 * u + g;
                xv[u] = (k < w1) ? xg(col[u]) : bc_f4(0.0f);
            }
#pragma unroll
            for (int u = 0; u < BC_U; ++u) {
                const int k = k0 + 16 * u + g;
                if (k < w1) sf[(k - w0) * 4 + q] = (q < D) ? bc_dot(a[u], xv[u]) : 0.0f;
            }
        }
        // LDS operations of one wavefront complete in issue order: the reads
        // below see every lane's write, and the next window's writes come
        // after these reads. The wave barriers only keep the compiler from
        // moving LDS accesses across these points.
        __builtin_amdgcn_wave_barrier();
        const int lo = max(rb, w0), hi = min(re, w1);
        for (int k = lo; k < hi; ++k) {
            const bc_f4 v = slice[k - w0];
            acc = (k == rb) ? v : acc + v;
        }
        __builtin_amdgcn_wave_barrier();
    }
    return acc;
}

// y (mode)= A x for 'rows' block rows; grid (G, batches), G a multiple of 8,
// 256 threads. Strides in float4 entries (0: one array for every batch).
template <int D>
__global__ __launch_bounds__(256) void bcsr_spmv_kernel(int rows, const int* __restrict__ JA,
                                                        const int* __restrict__ IA, const bc_f4* __restrict__ vals,
                                                        long long vstride, const bc_f4* __restrict__ x,
                                                        long long xstride, bc_f4* __restrict__ y, int mode)
{
    __shared__ bc_f4 lds[4][BC_WIN];
    bc_f4* slice = lds[threadIdx.x >> 6];
    const int lane = threadIdx.x & 63;
    const int bb = blockIdx.y;
    const bc_f4* vb = vals + bb * vstride;
    const bc_f4* xb = x + bb * xstride;
    bc_f4* yb = y + (long long)bb * rows;
    for (ChunkWalk w(div_up_dev(rows, 64), 4); w.valid(); w.next()) {
        const int r0 = w.cur * 64, nr = min(64, rows - r0);
        const bc_f4 v = bcsr_chunk_product<D, true>(JA, IA, vb, r0, nr, [xb](int c) { return xb[c]; }, slice);
        if (lane < nr) {
            bc_f4* o = yb + r0 + lane;
            if (mode == CM_ASSIGN) {
                *o = v;
            } else {
                bc_f4 old = *o;
#pragma unroll
                for (int c = 0; c < 4; ++c) old[c] = assign_mode<float>(mode, old[c], v[c]);
                *o = old;
            }
        }
    }
}

// Jacobi entries of a block matrix, A.diagonal().cwiseInverseCheck() over
// block scalars (cuMat/src/ConjugateGradient.h:218-246 with the diagonal
// extraction and checked inverse of tests/TestBlockedConjugateGradient.cu:
// 174-210): the diagonal of the row's first diagonal block (a zero block
// where none is stored), 1/d per component or 1 where d == 0; components
// >= D are 0. Thread per block row, grid.y = value batches.
template <int D>
__global__ __launch_bounds__(256) void bcsr_invdiag_kernel(int rows, const int* __restrict__ JA,
                                                           const int* __restrict__ IA,
                                                           const bc_f4* __restrict__ vals, long long vstride,
                                                           bc_f4* __restrict__ out)
{
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= rows) return;
    const bc_f4* vb = vals + blockIdx.y * vstride;
    int kd = -1;
    for (int k = JA[r]; k < JA[r + 1]; ++k)
        if (IA[k] == r) {
            kd = k;
            break;
        }
    bc_f4 o = bc_f4(0.0f);
#pragma unroll
    for (int c = 0; c < D; ++c) {
        const float d = kd < 0 ? 0.0f : vb[(long long)kd * D + c][c];
        o[c] = d == 0.0f ? 1.0f : 1.0f / d;
    }
    out[(long long)blockIdx.y * rows + r] = o;
}

// dst = src with components >= D set to 0 (right-hand sides of the block
// solver: the reference's norms and preconditioner never see them)
template <int D>
__global__ __launch_bounds__(256) void bcsr_mask_kernel(long long n, const bc_f4* __restrict__ src,
                                                        bc_f4* __restrict__ dst)
{
    const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    bc_f4 v = src[i];
#pragma unroll
    for (int c = D; c < 4; ++c) v[c] = 0.0f;
    dst[i] = v;
}

inline void launch_bcsr_spmv(hipStream_t stream, int D, int rows, int batches, const int* JA, const int* IA,
                             const bc_f4* vals, long long vstride, const bc_f4* x, long long xstride, bc_f4* y,
                             int mode)
{
    const dim3 grid(walk_grid(div_up(div_up((long long)rows, 64), 4), 2048), batches);
    if (D == 2)
        bcsr_spmv_kernel<2><<<grid, 256, 0, stream>>>(rows, JA, IA, vals, vstride, x, xstride, y, mode);
    else if (D == 3)
        bcsr_spmv_kernel<3><<<grid, 256, 0, stream>>>(rows, JA, IA, vals, vstride, x, xstride, y, mode);
    else
        bcsr_spmv_kernel<4><<<grid, 256, 0, stream>>>(rows, JA, IA, vals, vstride, x, xstride, y, mode);
}

inline void launch_bcsr_invdiag(hipStream_t stream, int D, int rows, int vbatches, const int* JA, const int* IA,
                                const bc_f4* vals, long long vstride, bc_f4* out)
{
    const dim3 grid(div_up((long long)rows, 256), vbatches);
    if (D == 2)
        bcsr_invdiag_kernel<2><<<grid, 256, 0, stream>>>(rows, JA, IA, vals, vstride, out);
    else if (D == 3)
        bcsr_invdiag_kernel<3><<<grid, 256, 0, stream>>>(rows, JA, IA, vals, vstride, out);
    else
        bcsr_invdiag_kernel<4><<<grid, 256, 0, stream>>>(rows, JA, IA, vals, vstride, out);
}

}  // namespace cm
