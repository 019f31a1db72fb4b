// spmv_wave.hpp -- wavefront-level, software-pipelined CSR SpMV core.
//
// The unit of work is a chunk of 64 consecutive rows owned by one wavefront
// (lane l sums row r0 + l). For a chunk whose non-zeros fit the lane budget
// (EPT per lane, i.e. EPT average non-zeros per row), the wavefront
//   1. loads the chunk's column indices and values with fully coalesced
//      loads, EPT per lane, plus each lane's row bounds;
//   2. gathers x at those columns and parks the products in its private LDS
//      slice (no workgroup barrier: only this wavefront touches the slice);
//   3. sums each lane's row sequentially in CSR order (the reference's
//      summation order, cuMat/src/SparseProductEvaluation.h:55-87).
// Step 1 of the NEXT chunk is issued before steps 2-3 of the current one, so
// index/value loads of one chunk overlap the gathers of the previous one and
// every wavefront keeps two chunks of loads in flight.
// Chunks with more non-zeros fall back to one row at a time per wavefront
// (lanes stride the row, fixed-order DPP reduction).
// Chunks are assigned so that each XCD owns a contiguous range of rows:
// stencil neighbours of a chunk are then resident in the same L2.
#pragma once
#include "cm_common.hpp"

namespace cm {

template <typename VT, int EPT>
struct WaveChunk {
    int r0, nr, s, e;
    int rb, re;  // this lane's row bounds
    int col[EPT];
    VT v[EPT];
};

// Chunk walk of one wavefront over NC chunks: XCD x (= block % 8) owns chunks
// [x*NC/8, (x+1)*NC/8); its GW/8 wavefronts stride through that range.
struct ChunkWalk {
    int cur, hi, step;
    __device__ ChunkWalk(int NC, int waves_per_block)
    {
        const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        const int G = gridDim.x;                  // multiple of 8
        const int x = blockIdx.x & 7;
        const int lw = (blockIdx.x >> 3) * waves_per_block + wave;
        const int lo = (int)(((long long)NC * x) >> 3);
        hi = (int)(((long long)NC * (x + 1)) >> 3);
        step = (G >> 3) * waves_per_block;
        cur = lo + lw;
    }
    __device__ bool valid() const { return cur < hi; }
    __device__ int peek_next() const { return cur + step; }
    __device__ void next() { cur += step; }
};

template <bool NT, typename X>
__device__ __forceinline__ X ld_matrix(const X* p)
{
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}

// NT: which matrix streams are loaded non-temporally (streaming: bit 0 the
// column indices, bit 1 the values). A single product streams a matrix that
// is not read again soon; an iterative solver reads the same matrix every
// iteration, and plain loads let the 256 MB last-level cache keep part of it
// from one iteration to the next.
template <typename VT, int EPT, int NT = 3>
__device__ __forceinline__ void chunk_load(WaveChunk<VT, EPT>& C, const int* __restrict__ JA,
                                           const int* __restrict__ IA, const VT* __restrict__ vals, int r0,
                                           int nrows)
{
    const int lane = threadIdx.x & 63;
    C.r0 = r0;
    C.nr = min(64, nrows - r0);
    C.s = JA[r0];
    C.e = JA[r0 + C.nr];
    C.rb = (lane < C.nr) ? JA[r0 + lane] : 0;
    C.re = (lane < C.nr) ? JA[r0 + lane + 1] : 0;
    if (C.e - C.s <= EPT * 64) {
#pragma unroll
        for (int j = 0; j < EPT; ++j) {
            const int k = C.s + lane + 64 * j;
            const bool ok = k < C.e;
            C.col[j] = ok ? ld_matrix<(NT & 1) != 0>(IA + k) : 0;
            C.v[j] = ok ? ld_matrix<(NT & 2) != 0>(vals + k) : VT(0);
        }
    }
}

// Returns (A x)[r0 + lane] (0 for lanes beyond the chunk). 'slice' is this
// wavefront's LDS slice of EPT*64 elements. XG(col) yields x[col].
template <typename VT, typename T, int EPT, typename XG>
__device__ __forceinline__ T chunk_compute(const WaveChunk<VT, EPT>& C, const int* __restrict__ JA,
                                           const int* __restrict__ IA, const VT* __restrict__ vals, XG xg,
                                           T* slice)
{
    const int lane = threadIdx.x & 63;
    T acc = T(0);
    if (C.e - C.s <= EPT * 64) {
        T xv[EPT];
#pragma unroll
        for (int j = 0; j < EPT; ++j) {
            const int k = C.s + lane + 64 * j;
            xv[j] = (k < C.e) ? xg(C.col[j]) : T(0);
        }
#pragma unroll
        for (int j = 0; j < EPT; ++j) {
            const int k = C.s + lane + 64 * j;
            if (k < C.e) slice[k - C.s] = Load<VT, T>::get(C.v, j) * xv[j];
        }
        // LDS operations of one wavefront complete in issue order, so the
        // reads below see every lane's write; the wave barrier only keeps
        // the compiler from moving code across this point.
        __builtin_amdgcn_wave_barrier();
        const int a = C.rb - C.s, b = C.re - C.s;
        if (lane < C.nr && a < b) {
            acc = slice[a];
            for (int k = a + 1; k < b; ++k) acc += slice[k];
        }
        __builtin_amdgcn_wave_barrier();
    } else {
        // dense chunk: one row at a time, lanes stride the row
        for (int i = 0; i < C.nr; ++i) {
            const int a = JA[C.r0 + i], b = JA[C.r0 + i + 1];
            T s = T(0);
            for (int k = a + lane; k < b; k += 64) s += Load<VT, T>::get(vals, k) * xg(IA[k]);
            s = wave_reduce<OpSum>(s);
            if (lane == i) acc = s;
        }
    }
    return acc;
}

}  // namespace cm
