// spmv_core.hpp -- block-level CSR SpMV used by the standalone SpMV entry
// point and by the fused CG kernels.
//
// Layout of the work: a workgroup owns RPB consecutive rows. When the rows'
// non-zeros fit in the LDS product buffer (always for stencil matrices),
// the block streams IA/vals/x-gathers in CSR order with fully coalesced
// loads, parks the products in LDS and then each thread sums its own row
// sequentially in CSR order (the reference's order,
// cuMat/src/SparseProductEvaluation.h:55-87). Blocks whose rows hold more
// non-zeros than fit fall back to one wavefront per row with a fixed-order
// DPP reduction.
#pragma once
#include "cm_common.hpp"

namespace cm {

// Map the hardware block index to a logical row block so that consecutive
// row blocks run on the same XCD (hardware dispatches block i to XCD i%8):
// stencil neighbours of a row block then hit the same L2.
__device__ __forceinline__ unsigned xcd_remap(unsigned hw, unsigned nblocks)
{
    constexpr unsigned NXCD = 8;
    if (nblocks < NXCD * 2) return hw;
    unsigned per = nblocks / NXCD;          // blocks per XCD in the regular part
    unsigned regular = per * NXCD;
    if (hw >= regular) return hw;           // tail keeps its index
    return (hw % NXCD) * per + hw / NXCD;
}

// Row blocks [0, NB) are split into 8 contiguous ranges, one per XCD;
// hardware block h (dispatched to XCD h % 8) walks its XCD's range with
// stride G/8 (G: grid size, a multiple of 8), so stencil neighbours of a row
// block share an L2.
struct RowBlockWalk {
    int lo, hi, step, cur;
    __device__ RowBlockWalk(int h, int G, int NB)
    {
        const int x = h & 7, l = h >> 3;
        lo = (int)(((long long)NB * x) >> 3);
        hi = (int)(((long long)NB * (x + 1)) >> 3);
        step = G >> 3;
        cur = lo + l;
    }
    __device__ bool valid() const { return cur < hi; }
    __device__ void next() { cur += step; }
};

// Grid for a walk over NB row blocks: a multiple of 8, at most gmax.
inline int walk_grid(long long NB, int gmax)
{
    long long g = ((NB + 7) / 8) * 8;
    if (g > gmax) g = gmax;
    if (g < 8) g = 8;
    return (int)g;
}

template <typename T, int RPB, int CAP>
struct SpmvSmem {
    T prod[CAP];
};

template <typename VT, typename T, int BS, int RPB, int CAP>
__device__ __forceinline__ void block_csr_spmv_nodiag(const int* __restrict__ JA, const int* __restrict__ IA,
                                                      const VT* __restrict__ vals,
                                                      const T* __restrict__ x, int r0, int nr,
                                                      SpmvSmem<T, RPB, CAP>& sm, T (&out)[RPB / BS]);

// Computes y_local[i] = (A x)[r0 + i] for i < nr into 'out' (registers of the
// calling thread: row r0 + tid + j*BS -> out[j]). VT: stored value type.
// With 'diag' != NULL the matrix is diag[i] (kept in T) plus the stored
// entries (whose diagonal slots hold 0): the mixed-precision matrices keep
// their diagonal in FP32. Assumes blockDim.x == BS.
template <typename VT, typename T, int BS, int RPB, int CAP>
__device__ __forceinline__ void block_csr_spmv(const int* __restrict__ JA, const int* __restrict__ IA,
                                               const VT* __restrict__ vals,
                                               const T* __restrict__ x, int r0, int nr,
                                               SpmvSmem<T, RPB, CAP>& sm, T (&out)[RPB / BS],
                                               const T* __restrict__ diag = nullptr)
{
    block_csr_spmv_nodiag<VT, T, BS, RPB, CAP>(JA, IA, vals, x, r0, nr, sm, out);
    if (diag) {
#pragma unroll
        for (int j = 0; j < RPB / BS; ++j) {
            int i = threadIdx.x + j * BS;
            if (i < nr) out[j] += diag[r0 + i] * x[r0 + i];
        }
    }
}

template <typename VT, typename T, int BS, int RPB, int CAP>
__device__ __forceinline__ void block_csr_spmv_nodiag(const int* __restrict__ JA, const int* __restrict__ IA,
                                                      const VT* __restrict__ vals,
                                                      const T* __restrict__ x, int r0, int nr,
                                                      SpmvSmem<T, RPB, CAP>& sm, T (&out)[RPB / BS])
{
    constexpr int RPT = RPB / BS;
    constexpr int EPT = CAP / BS;  // staged non-zeros per thread
    const int tid = threadIdx.x;
    // block range (uniform) and this thread's row bounds, all issued at once
    const int s = JA[r0];
    const int e = JA[r0 + nr];
    int rb[RPT], re[RPT];
#pragma unroll
    for (int j = 0; j < RPT; ++j) {
        int i = tid + j * BS;
        rb[j] = (i < nr) ? JA[r0 + i] : 0;
        re[j] = (i < nr) ? JA[r0 + i + 1] : 0;
    }
    if (e - s <= CAP) {
        // every index/value load of the block is in flight before the first
        // x gather; then every gather before the first product
        int col[EPT];
        VT v[EPT];
#pragma unroll
        for (int j = 0; j < EPT; ++j) {
            int k = s + tid + j * BS;
            bool ok = k < e;
            col[j] = ok ? __builtin_nontemporal_load(IA + k) : 0;
            v[j] = ok ? __builtin_nontemporal_load(vals + k) : VT(0);
        }
        T xv[EPT];
#pragma unroll
        for (int j = 0; j < EPT; ++j) {
            int k = s + tid + j * BS;
            xv[j] = (k < e) ? x[col[j]] : T(0);
        }
#pragma unroll
        for (int j = 0; j < EPT; ++j) {
            int k = s + tid + j * BS;
            if (k < e) sm.prod[k - s] = Load<VT, T>::get(v, j) * xv[j];
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < RPT; ++j) {
            T acc = T(0);
            int a = rb[j] - s, b = re[j] - s;
            if (a < b) {
                acc = sm.prod[a];
                for (int k = a + 1; k < b; ++k) acc += sm.prod[k];
            }
            out[j] = acc;
        }
    } else {
        // long rows: one wavefront per row, lanes stride the row, fixed-order
        // wave reduction; results parked in prod[] (RPB <= CAP).
        const int lane = tid & 63, wave = tid >> 6;
        for (int i = wave; i < nr; i += BS / 64) {
            int a = JA[r0 + i], b = JA[r0 + i + 1];
            T acc = T(0);
            for (int k = a + lane; k < b; k += 64) acc += Load<VT, T>::get(vals, k) * x[IA[k]];
            acc = wave_reduce<OpSum>(acc);
            if (lane == 0) sm.prod[i] = acc;
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < RPT; ++j) {
            int i = tid + j * BS;
            out[j] = (i < nr) ? sm.prod[i] : T(0);
        }
    }
}

}  // namespace cm
