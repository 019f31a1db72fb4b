// dense.hpp -- dense matrix-vector product y = A x and the Jacobi diagonal of
// a dense matrix.
//
// Replaces the matrix-vector case of cuMat's dense ProductOp, which cuMat
// evaluates with cuBLAS gemm (cuMat/src/ProductOp.h:318-420), and the
// ExtractDiagonalOp + cwiseInverseCheck preconditioner setup on a dense
// matrix (cuMat/src/ConjugateGradient.h:218-246, cuMat/src/UnaryOps.h:177).
// A matrix-vector product reads every matrix element once and does one
// multiply-add per element: it is HBM-bound, so it is a streaming kernel, not
// a GEMM on the matrix cores.
//
// Work unit: a tile of 64 consecutive rows times a range of columns per
// 256-thread block.
//   column-major A (cuMat's default): lane l owns row r0 + l, so each column
//     of the tile is one 256-byte contiguous load per wavefront; wavefront w
//     sums the columns c0 + w, c0 + w + 4, ... with 8 loads in flight, and
//     the four partial sums are added in wavefront order through LDS;
//   row-major A: each wavefront takes 16 rows of the tile, lanes stride the
//     row with contiguous loads, fixed-order wavefront reduction.
// Few row tiles cannot fill 256 CUs, so the standalone product splits the
// columns into ranges of at least 512 (about 2048 blocks in all), writes one
// partial vector per range and adds the ranges in order in a second pass.
// The summation order depends only on the sizes, so results are
// deterministic from run to run.
#pragma once
#include <algorithm>

#include "cm_common.hpp"

namespace cm {

constexpr int DN_BS = 256;    // threads per block: 4 wavefronts
constexpr int DN_TILE = 64;   // rows per tile
constexpr int DN_UNROLL = 8;  // column loads in flight per wavefront

// Value of row r0 + lane of (A x) restricted to columns [c0, c1), on the
// lanes of wavefront 0 (0 elsewhere and beyond the last row). A is rows x
// cols. red: 4*64 elements of LDS. Called by all threads of the block.
template <typename T, bool ROW_MAJOR>
__device__ __forceinline__ T dense_tile(int rows, int cols, const T* __restrict__ A, const T* __restrict__ x,
                                        int r0, int c0, int c1, T* red)
{
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    if constexpr (!ROW_MAJOR) {
        const int r = r0 + lane;
        const bool ok = r < rows;
        const T* a = A + (ok ? r : 0);
        const long long lda = rows;
        T acc = T(0);
        int j = c0 + wave;
        for (; j + 4 * (DN_UNROLL - 1) < c1; j += 4 * DN_UNROLL) {
            T v[DN_UNROLL];
#pragma unroll
            for (int u = 0; u < DN_UNROLL; ++u) v[u] = a[(long long)(j + 4 * u) * lda];
#pragma unroll
            for (int u = 0; u < DN_UNROLL; ++u) acc += v[u] * x[j + 4 * u];
        }
        for (; j < c1; j += 4) acc += a[(long long)j * lda] * x[j];
        red[wave * 64 + lane] = ok ? acc : T(0);
    } else {
        for (int i = wave; i < DN_TILE; i += 4) {
            const int r = r0 + i;
            T s = T(0);
            if (r < rows) {  // uniform over the wavefront
                const T* a = A + (long long)r * cols;
                int c = c0 + lane;
                for (; c + 64 * 3 < c1; c += 64 * 4) {
                    const T a0 = a[c], a1 = a[c + 64], a2 = a[c + 128], a3 = a[c + 192];
                    s += a0 * x[c];
                    s += a1 * x[c + 64];
                    s += a2 * x[c + 128];
                    s += a3 * x[c + 192];
                }
                for (; c < c1; c += 64) s += a[c] * x[c];
            }
            s = wave_reduce<OpSum>(s);
            if (lane == 0) red[i] = s;
        }
    }
    __syncthreads();
    T y = T(0);
    if (wave == 0) {
        if constexpr (!ROW_MAJOR) y = ((red[lane] + red[64 + lane]) + red[128 + lane]) + red[192 + lane];
        else y = red[lane];
    }
    __syncthreads();  // red is reused by the next tile
    return y;
}

// y_b = A_b x_b (assignment mode as cm_assign_mode); A_b = A + b*a_bstride.
// Block (t, s) of the nsplit * ntiles grid takes row tile t and column range
// s (csz columns each); with nsplit = 1 it writes y, otherwise the partial
// part[(b * nsplit + s) * rows + r], which dense_gemv_combine_kernel adds.
template <typename T, bool ROW_MAJOR>
__global__ __launch_bounds__(DN_BS) void dense_gemv_kernel(int rows, int cols, const T* __restrict__ A,
                                                           long long a_bstride, const T* __restrict__ x,
                                                           int x_batched, T* __restrict__ y, int mode,
                                                           int nsplit, int csz, T* __restrict__ part)
{
    __shared__ T red[4 * 64];
    const int bb = blockIdx.y;
    const T* Ab = A + bb * a_bstride;
    const T* xb = x + (x_batched ? (long long)bb * cols : 0);
    T* yb = y + (long long)bb * rows;
    const int ntiles = div_up_dev(rows, DN_TILE);
    const long long units = (long long)ntiles * nsplit;
    for (long long u = blockIdx.x; u < units; u += gridDim.x) {
        const int t = (int)(u / nsplit), s = (int)(u % nsplit);
        const int r0 = t * DN_TILE;
        const int c0 = s * csz, c1 = min(cols, c0 + csz);
        const T v = dense_tile<T, ROW_MAJOR>(rows, cols, Ab, xb, r0, c0, c1, red);
        const int r = r0 + (int)threadIdx.x;
        if (threadIdx.x < 64 && r < rows) {
            if (nsplit > 1) part[((long long)bb * nsplit + s) * rows + r] = v;
            else yb[r] = (mode == CM_ASSIGN) ? v : assign_mode<T>(mode, yb[r], v);
        }
    }
}

// Column-major A whose rows and base allow 16-byte loads: the same product
// with lane l of a (64*VW)-row tile owning the VW = 16/sizeof(T) rows
// r0 + VW*l ..., so each column of the tile is one 16-byte load per lane
// (1 KB per wavefront for FP32, against 256 B in dense_tile); wavefront w
// sums the columns c0 + w, c0 + w + 4, ... with DN_UNROLL loads in flight
// and the four wavefront sums are added in order, as in dense_tile. Same
// grid and partial layout as dense_gemv_kernel.
template <typename T>
struct DenseVec {
    static constexpr int w = 16 / (int)sizeof(T);
    typedef T type __attribute__((ext_vector_type(16 / sizeof(T))));
};

template <typename T>
__global__ __launch_bounds__(DN_BS) void dense_gemv_v_kernel(int rows, int cols, const T* __restrict__ A,
                                                             long long a_bstride, const T* __restrict__ x,
                                                             int x_batched, T* __restrict__ y, int mode,
                                                             int nsplit, int csz, T* __restrict__ part)
{
    typedef typename DenseVec<T>::type V;
    constexpr int VW = DenseVec<T>::w;
    constexpr int TILE = 64 * VW;
    __shared__ V red[4 * 64];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int bb = blockIdx.y;
    const T* Ab = A + bb * a_bstride;
    const T* xb = x + (x_batched ? (long long)bb * cols : 0);
    T* yb = y + (long long)bb * rows;
    const int ntiles = div_up_dev(rows, TILE);
    const long long units = (long long)ntiles * nsplit;
    const long long ldv = rows / VW;  // column stride in vectors
    for (long long u = blockIdx.x; u < units; u += gridDim.x) {
        const int t = (int)(u / nsplit), s = (int)(u % nsplit);
        const int r = t * TILE + VW * lane;  // first of this lane's rows (rows % VW == 0)
        const bool ok = r < rows;
        const int c0 = s * csz, c1 = min(cols, c0 + csz);
        const V* a = reinterpret_cast<const V*>(Ab + (ok ? r : 0));
        V acc = V(T(0));
        int j = c0 + wave;
        for (; j + 4 * (DN_UNROLL - 1) < c1; j += 4 * DN_UNROLL) {
            V v[DN_UNROLL];
#pragma unroll
            for (int k = 0; k < DN_UNROLL; ++k) v[k] = __builtin_nontemporal_load(a + (long long)(j + 4 * k) * ldv);
#pragma unroll
            for (int k = 0; k < DN_UNROLL; ++k) acc += v[k] * xb[j + 4 * k];
        }
        for (; j < c1; j += 4) acc += __builtin_nontemporal_load(a + (long long)j * ldv) * xb[j];
        red[wave * 64 + lane] = acc;
        __syncthreads();
        if (wave == 0 && ok) {
            const V v = ((red[lane] + red[64 + lane]) + red[128 + lane]) + red[192 + lane];
#pragma unroll
            for (int k = 0; k < VW; ++k) {
                if (nsplit > 1) part[((long long)bb * nsplit + s) * rows + r + k] = v[k];
                else yb[r + k] = (mode == CM_ASSIGN) ? v[k] : assign_mode<T>(mode, yb[r + k], v[k]);
            }
        }
        __syncthreads();  // red is reused by the next unit
    }
}

// y_b[r] (mode)= sum over s of part[(b * nsplit + s) * rows + r], in s order
template <typename T>
__global__ void dense_gemv_combine_kernel(int rows, int nsplit, const T* __restrict__ part, T* __restrict__ y,
                                          int mode)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const int bb = blockIdx.y;
    const T* p = part + (long long)bb * nsplit * rows + r;
    T v = p[0];
    for (int s = 1; s < nsplit; ++s) v += p[(long long)s * rows];
    T* yy = y + (long long)bb * rows + r;
    *yy = (mode == CM_ASSIGN) ? v : assign_mode<T>(mode, *yy, v);
}

// column split of the standalone product: ranges of >= 512 columns, about
// 2048 blocks over all row tiles
inline void dense_split(int rows, int cols, int& nsplit, int& csz, int tile = DN_TILE)
{
    const long long tiles = div_up(rows, tile);
    long long ns = std::max<long long>(1, std::min<long long>(div_up(2048, tiles), cols / 512));
    csz = (int)div_up(cols, ns);
    nsplit = (int)div_up(cols, std::max(csz, 1));
    if (nsplit < 1) nsplit = 1;
}

// inv_b[i] = A_b(i,i) == 0 ? 1 : 1 / A_b(i,i)
template <typename T>
__global__ void dense_inverse_diagonal_kernel(int n, const T* __restrict__ A, long long a_bstride,
                                              T* __restrict__ inv)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const T d = A[blockIdx.y * a_bstride + (long long)i * n + i];
    inv[(long long)blockIdx.y * n + i] = (d == T(0)) ? T(1) : T(1) / d;
}

inline int dense_grid(int rows, int nsplit = 1, int tile = DN_TILE)
{
    return (int)std::min<long long>((long long)div_up(rows, tile) * nsplit, 4096);
}

}  // namespace cm
