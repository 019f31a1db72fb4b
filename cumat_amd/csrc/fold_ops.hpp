// fold_ops.hpp -- wave64 building blocks of every reduction: DPP lane
// exchanges, the reduction operators (cuMat functor::Sum..BitwiseOr,
// cuMat/src/ReductionOps.h:859-968) and the fixed-order wave and block
// reductions.
//
// Self-contained on purpose (no #include), like cwise_ops.hpp: the library
// compiles it, and the run-time compiler gets the same text ahead of each
// generated reduction kernel (jit.hip), so both combine partial results in
// the same order and give the same bits.
#pragma once

namespace cm {

// DPP control words (GCN/CDNA ISA).
enum : int {
    DPP_QUAD_XOR1 = 0xB1,        // quad_perm [1,0,3,2]
    DPP_QUAD_XOR2 = 0x4E,        // quad_perm [2,3,0,1]
    DPP_ROW_HALF_MIRROR = 0x141, // lane i <- lane 7-i within 8
    DPP_ROW_MIRROR = 0x140,      // lane i <- lane 15-i within 16
};

template <int CTRL>
__device__ __forceinline__ int dpp_i32(int v)
{
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ float dpp(float v)
{
    return __int_as_float(dpp_i32<CTRL>(__float_as_int(v)));
}
template <int CTRL>
__device__ __forceinline__ int dpp(int v)
{
    return dpp_i32<CTRL>(v);
}
template <int CTRL>
__device__ __forceinline__ double dpp(double v)
{
    long long b = __double_as_longlong(v);
    int lo = dpp_i32<CTRL>((int)(b & 0xffffffffll));
    int hi = dpp_i32<CTRL>((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ float readlane(float v, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ int readlane(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ double readlane(double v, int l)
{
    long long b = __double_as_longlong(v);
    int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), l);
    int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Reduction operators (cuMat functor::Sum..BitwiseOr semantics).
struct OpSum {
    template <typename T> __device__ __forceinline__ static T apply(T a, T b) { return a + b; }
};
struct OpProd {
    template <typename T> __device__ __forceinline__ static T apply(T a, T b) { return a * b; }
};
struct OpMin {
    template <typename T> __device__ __forceinline__ static T apply(T a, T b) { return (a < b) ? a : b; }
};
struct OpMax {
    template <typename T> __device__ __forceinline__ static T apply(T a, T b) { return (a > b) ? a : b; }
};
struct OpAll {
    template <typename T> __device__ __forceinline__ static T apply(T a, T b) { return (T)((a != T(0)) && (b != T(0))); }
};
struct OpAny {
    template <typename T> __device__ __forceinline__ static T apply(T a, T b) { return (T)((a != T(0)) || (b != T(0))); }
};
struct OpBand {
    template <typename T> __device__ __forceinline__ static T apply(T a, T b) { return a & b; }
};
struct OpBor {
    template <typename T> __device__ __forceinline__ static T apply(T a, T b) { return a | b; }
};

// Wrapping integer arithmetic: int32 sums/products are done on uint32.
template <> __device__ __forceinline__ int OpSum::apply<int>(int a, int b)
{
    return (int)((unsigned)a + (unsigned)b);
}
template <> __device__ __forceinline__ int OpProd::apply<int>(int a, int b)
{
    return (int)((unsigned)a * (unsigned)b);
}

// Reduce within each row of 16 lanes; every lane of the row ends with the
// row's result. The order of combination is fixed, so the result is
// bit-reproducible.
template <typename Op, typename T>
__device__ __forceinline__ T row16_reduce(T v)
{
    v = Op::apply(v, dpp<DPP_QUAD_XOR1>(v));
    v = Op::apply(v, dpp<DPP_QUAD_XOR2>(v));
    v = Op::apply(v, dpp<DPP_ROW_HALF_MIRROR>(v));
    v = Op::apply(v, dpp<DPP_ROW_MIRROR>(v));
    return v;
}

// Full wave64 reduction; the result is returned (uniform) in every lane.
// All 64 lanes must be active.
template <typename Op, typename T>
__device__ __forceinline__ T wave_reduce(T v)
{
    v = row16_reduce<Op>(v);
    T a = readlane(v, 0), b = readlane(v, 16), c = readlane(v, 32), d = readlane(v, 48);
    return Op::apply(Op::apply(a, b), Op::apply(c, d));
}

// Block reduction for blockDim.x == BS (multiple of 64). 'lds' needs BS/64
// entries. Result valid in every thread. Contains two barriers.
template <typename Op, int BS, typename T>
__device__ __forceinline__ T block_reduce(T v, T* lds)
{
    constexpr int NW = BS / 64;
    v = wave_reduce<Op>(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) lds[w] = v;
    __syncthreads();
    T r = lds[0];
#pragma unroll
    for (int i = 1; i < NW; ++i) r = Op::apply(r, lds[i]);
    __syncthreads();
    return r;
}

}  // namespace cm
