// reduce.hip -- batched reductions over any subset of {Row, Column, Batch}
// and the fused dot / squaredNorm.
//
// Replaces ReductionEvaluator<...>::eval with its Segmented / Thread / Warp /
// Block / Device algorithms and the RTX-tuned Auto table
// (cuMat/src/ReductionOps.h:167-534, cuMat/src/ReductionAlgorithmSelection.h)
// and the CUB full reduction used by dot/squaredNorm
// (cuMat/src/ReductionOps.h:564-598, cuMat/src/ReductionOpsPlugin.inl:395-431).
//
// The axis set is first simplified (size-1 axes may be treated as reduced,
// as cuMat does at compile time, cuMat/src/ReductionOps.h:840-845), then
// adjacent axes with the same status are merged in memory order. Only two
// access shapes remain, each with a layout-driven kernel:
//   * contiguous segments  (reduced axes are the fastest ones) -> seg kernels,
//     wave64 lanes stride a segment, DPP combine;
//   * strided columns      (a kept axis is the fastest one)     -> col kernel,
//     lanes own consecutive kept elements, loop over the reduced rows.
// Every combination order is fixed, so results are bit-reproducible.
#include "cm_common.hpp"

#include <limits>

// Fixed device buffer of cm_dot_noalloc: the per-chunk partial sums of one
// batch (a split dot product has at most 4096 chunks, 8 bytes each for
// FP64). One copy per GPU, alive until the process ends; every call on that
// GPU shares it. In host code the name is a host-side stand-in: its device
// address comes from hipGetSymbolAddress (noalloc_scratch_init below) and
// only that address is handed to kernels.
constexpr size_t NOALLOC_SCRATCH_BYTES = 64 * 1024;
static __device__ __align__(256) unsigned char noalloc_scratch[NOALLOC_SCRATCH_BYTES];

namespace cm {

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_noalloc_mutex;
static void* g_noalloc_ptr[NOALLOC_MAX_DEVICES];

// Looks up this device's copy of noalloc_scratch (the device must be
// current); called once per device when a context is created there.
int noalloc_scratch_init(int device)
{
    std::lock_guard<std::mutex> lk(g_noalloc_mutex);
    CM_REQUIRE(device >= 0 && device < NOALLOC_MAX_DEVICES, CM_E_UNSUPPORTED, "device %d: at most %d devices",
               device, NOALLOC_MAX_DEVICES);
    if (g_noalloc_ptr[device]) return CM_OK;
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(noalloc_scratch));
    if (e != hipSuccess || p == nullptr)
        return fail(CM_E_HIP, "hipGetSymbolAddress(noalloc_scratch) on device %d failed: %s", device,
                    hipGetErrorString(e));
    g_noalloc_ptr[device] = p;
    return CM_OK;
}

static void* noalloc_scratch_ptr(int device)
{
    std::lock_guard<std::mutex> lk(g_noalloc_mutex);
    return (device >= 0 && device < NOALLOC_MAX_DEVICES) ? g_noalloc_ptr[device] : nullptr;
}

// ---------------------------------------------------------------- types
template <typename Op, typename T> struct Identity;
template <typename T> struct Identity<OpSum, T> { __device__ static T v() { return T(0); } };
template <typename T> struct Identity<OpProd, T> { __device__ static T v() { return T(1); } };
template <typename T> struct Identity<OpMin, T> { __device__ static T v() { return std::numeric_limits<T>::max(); } };
template <typename T> struct Identity<OpMax, T> { __device__ static T v() { return std::numeric_limits<T>::lowest(); } };
template <typename T> struct Identity<OpAll, T> { __device__ static T v() { return T(1); } };
template <typename T> struct Identity<OpAny, T> { __device__ static T v() { return T(0); } };
template <typename T> struct Identity<OpBand, T> { __device__ static T v() { return ~T(0); } };
template <typename T> struct Identity<OpBor, T> { __device__ static T v() { return T(0); } };

// Element sources: plain load, product of two streams, square.
template <typename S, typename T> struct SrcPlain {
    const S* a;
    __device__ __forceinline__ T operator()(long long i) const { return (T)a[i]; }
};
template <typename S, typename T> struct SrcDot {
    const S* a;
    const S* b;
    __device__ __forceinline__ T operator()(long long i) const { return (T)a[i] * (T)b[i]; }
};

constexpr int RBS = 256;  // threads per block of every reduction kernel

// ------------------------------------------------- contiguous segments
// Thread per segment (short segments, L <= 16).
template <typename Op, typename T, typename S, typename Src>
__global__ void seg_thread_kernel(Src src, long long L, long long nseg, long long seg_stride, S* out)
{
    long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nseg) return;
    T acc = Identity<Op, T>::v();
    for (long long j = 0; j < L; ++j) acc = Op::apply(acc, src(s * seg_stride + j));
    out[s] = (S)acc;
}

// Wave per segment (medium segments): lanes stride the segment.
template <typename Op, typename T, typename S, typename Src>
__global__ __launch_bounds__(RBS) void seg_wave_kernel(Src src, long long L, long long nseg,
                                                       long long seg_stride, S* out)
{
    const int lane = threadIdx.x & 63;
    long long s = (long long)blockIdx.x * (RBS / 64) + (threadIdx.x >> 6);
    if (s >= nseg) return;  // uniform per wave
    const long long base = s * seg_stride;
    T acc = Identity<Op, T>::v();
    long long j = lane;
    // four independent loads in flight per lane
    for (; j + 192 < L; j += 256) {
        T v0 = src(base + j), v1 = src(base + j + 64), v2 = src(base + j + 128), v3 = src(base + j + 192);
        acc = Op::apply(acc, Op::apply(Op::apply(v0, v1), Op::apply(v2, v3)));
    }
    for (; j < L; j += 64) acc = Op::apply(acc, src(base + j));
    acc = wave_reduce<Op>(acc);
    if (lane == 0) out[s] = (S)acc;
}

// float4 / int4 variant of the wave kernel for 4-byte plain sources whose
// segments are 16-byte aligned and a multiple of 4 long.
template <typename Op, typename T>
__global__ __launch_bounds__(RBS) void seg_wave_vec4_kernel(const T* __restrict__ a, long long L4,
                                                            long long nseg, T* out)
{
    using V = typename std::conditional<std::is_same<T, float>::value, float4, int4>::type;
    const int lane = threadIdx.x & 63;
    long long s = (long long)blockIdx.x * (RBS / 64) + (threadIdx.x >> 6);
    if (s >= nseg) return;
    const V* p = reinterpret_cast<const V*>(a) + s * L4;
    typedef T T4 __attribute__((ext_vector_type(4)));
    const T4* p4 = reinterpret_cast<const T4*>(p);
    T acc = Identity<Op, T>::v();
    long long j = lane;
    // four 16-byte loads in flight per lane (a 4 KB segment in one round),
    // folded in the same order as the two-load loop below
    for (; j + 192 < L4; j += 256) {
        const T4 u0 = __builtin_nontemporal_load(p4 + j), u1 = __builtin_nontemporal_load(p4 + j + 64),
                 u2 = __builtin_nontemporal_load(p4 + j + 128), u3 = __builtin_nontemporal_load(p4 + j + 192);
        acc = Op::apply(acc, Op::apply(Op::apply(u0.x, u0.y), Op::apply(u0.z, u0.w)));
        acc = Op::apply(acc, Op::apply(Op::apply(u1.x, u1.y), Op::apply(u1.z, u1.w)));
        acc = Op::apply(acc, Op::apply(Op::apply(u2.x, u2.y), Op::apply(u2.z, u2.w)));
        acc = Op::apply(acc, Op::apply(Op::apply(u3.x, u3.y), Op::apply(u3.z, u3.w)));
    }
    for (; j + 64 < L4; j += 128) {
        V u = p[j], w = p[j + 64];
        acc = Op::apply(acc, Op::apply(Op::apply(u.x, u.y), Op::apply(u.z, u.w)));
        acc = Op::apply(acc, Op::apply(Op::apply(w.x, w.y), Op::apply(w.z, w.w)));
    }
    for (; j < L4; j += 64) {
        V u = p[j];
        acc = Op::apply(acc, Op::apply(Op::apply(u.x, u.y), Op::apply(u.z, u.w)));
    }
    acc = wave_reduce<Op>(acc);
    if (lane == 0) out[s] = acc;
}

// Block per chunk (long segments / full reductions). nchunk chunks of
// 'chunk' elements per segment; with nchunk > 1 each block writes one
// partial and seg_combine_kernel folds them in chunk order (a kernel
// boundary instead of device-scope fences, which would write back the L2
// of the XCD once per block).
template <typename Op, typename T, typename S, typename Src>
__global__ __launch_bounds__(RBS) void seg_block_kernel(Src src, long long L, long long seg_stride,
                                                        int nchunk, long long chunk, T* partials, S* out)
{
    __shared__ T lds[RBS / 64];
    const long long s = blockIdx.x / nchunk;
    const int c = blockIdx.x % nchunk;
    const long long beg = c * chunk;
    const long long end = min(L, beg + chunk);
    const long long base = s * seg_stride;
    T acc = Identity<Op, T>::v();
    long long j = beg + threadIdx.x;
    for (; j + 3 * RBS < end; j += 4 * RBS) {
        T v0 = src(base + j), v1 = src(base + j + RBS), v2 = src(base + j + 2 * RBS), v3 = src(base + j + 3 * RBS);
        acc = Op::apply(acc, Op::apply(Op::apply(v0, v1), Op::apply(v2, v3)));
    }
    for (; j < end; j += RBS) acc = Op::apply(acc, src(base + j));
    acc = block_reduce<Op, RBS>(acc, lds);
    if (nchunk == 1) {
        if (threadIdx.x == 0) out[s] = (S)acc;
        return;
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = acc;
}

// One block per segment folds that segment's nchunk partials (fixed order).
template <typename Op, typename T, typename S>
__global__ __launch_bounds__(RBS) void seg_combine_kernel(const T* __restrict__ partials, int nchunk, S* out)
{
    __shared__ T lds[RBS / 64];
    const long long s = blockIdx.x;
    T r = Identity<Op, T>::v();
    for (int k = threadIdx.x; k < nchunk; k += RBS) r = Op::apply(r, partials[s * nchunk + k]);
    r = block_reduce<Op, RBS>(r, lds);
    if (threadIdx.x == 0) out[s] = (S)r;
}

// ------------------------------------------------- strided columns
// out[p*Q + q] = reduce_k in[p*PS + k*st + q], k < K, q < Q, p < P.
// Lanes own consecutive q; the block's waves split K; 'ksplit' blocks per
// (p, q-tile) split K further and combine through partials.
template <typename Op, typename T, typename S>
__global__ __launch_bounds__(RBS) void col_kernel(const S* __restrict__ in, long long P, long long Q,
                                                  long long K, long long st, long long PS, int ksplit,
                                                  T* partials, S* out)
{
    __shared__ T lds[RBS];
    constexpr int NW = RBS / 64;
    const long long qtiles = (Q + 63) / 64;
    const long long tile = blockIdx.x / ksplit;  // (p, qtile)
    const int ks = blockIdx.x % ksplit;
    const long long p = tile / qtiles;
    const long long q = (tile % qtiles) * 64 + (threadIdx.x & 63);
    const int w = threadIdx.x >> 6;
    const long long kchunk = (K + ksplit - 1) / ksplit;
    const long long kb = ks * kchunk, ke = min(K, kb + kchunk);
    T acc = Identity<Op, T>::v();
    if (q < Q) {
        const S* col = in + p * PS + q;
        long long k = kb + w;
        for (; k + 3 * NW < ke; k += 4 * NW) {
            T v0 = (T)col[k * st], v1 = (T)col[(k + NW) * st], v2 = (T)col[(k + 2 * NW) * st],
              v3 = (T)col[(k + 3 * NW) * st];
            acc = Op::apply(acc, Op::apply(Op::apply(v0, v1), Op::apply(v2, v3)));
        }
        for (; k < ke; k += NW) acc = Op::apply(acc, (T)col[k * st]);
    }
    lds[threadIdx.x] = acc;
    __syncthreads();
    if (w == 0) {
        T r = lds[threadIdx.x];
#pragma unroll
        for (int i = 1; i < NW; ++i) r = Op::apply(r, lds[threadIdx.x + 64 * i]);
        acc = r;
    }
    if (ksplit == 1) {
        if (w == 0 && q < Q) out[p * Q + q] = (S)acc;
        return;
    }
    if (w == 0) partials[(long long)blockIdx.x * 64 + threadIdx.x] = acc;
}

// Folds the ksplit partial rows of every (p, q-tile) in split order.
template <typename Op, typename T, typename S>
__global__ __launch_bounds__(64) void col_combine_kernel(const T* __restrict__ partials, long long Q, int ksplit,
                                                         S* out)
{
    const long long qtiles = (Q + 63) / 64;
    const long long tile = blockIdx.x;
    const long long p = tile / qtiles;
    const long long q = (tile % qtiles) * 64 + threadIdx.x;
    T r = Identity<Op, T>::v();
    for (int k2 = 0; k2 < ksplit; ++k2) r = Op::apply(r, partials[(tile * ksplit + k2) * 64 + threadIdx.x]);
    if (q < Q) out[p * Q + q] = (S)r;
}

// col_kernel with 16-byte loads (4-byte types, Q, st and PS multiples of 4,
// 16-byte aligned input): lane l owns q = 4*(qtile*64 + l) + c, c < 4, so
// a wavefront's load covers 1 KB of a row instead of 256 B. BS threads per
// block (BS/64 wavefronts split k). With BS = 256 each component accumulates
// over k in exactly the order of col_kernel, so the results are
// bit-identical to it; BS = 1024 folds in another fixed order.
template <typename Op, typename T, int BS = RBS>
__global__ __launch_bounds__(BS) void col_kernel_v4(const T* __restrict__ in, long long P, long long Q,
                                                    long long K, long long st, long long PS, int ksplit,
                                                    T* partials, T* out)
{
    typedef T T4 __attribute__((ext_vector_type(4)));
    __shared__ T4 lds[BS];
    constexpr int NW = BS / 64;
    const long long Q4 = Q / 4;
    const long long qtiles = (Q4 + 63) / 64;
    const long long tile = blockIdx.x / ksplit;
    const int ks = blockIdx.x % ksplit;
    const long long p = tile / qtiles;
    const long long q4 = (tile % qtiles) * 64 + (threadIdx.x & 63);
    const int w = threadIdx.x >> 6;
    const long long kchunk = (K + ksplit - 1) / ksplit;
    const long long kb = ks * kchunk, ke = min(K, kb + kchunk);
    const T id = Identity<Op, T>::v();
    T4 acc = {id, id, id, id};
    auto app = [](T4 a, T4 b) {
        T4 r;
        r.x = Op::apply(a.x, b.x);
        r.y = Op::apply(a.y, b.y);
        r.z = Op::apply(a.z, b.z);
        r.w = Op::apply(a.w, b.w);
        return r;
    };
    if (q4 < Q4) {
        const T4* col = reinterpret_cast<const T4*>(in + p * PS) + q4;
        const long long st4 = st / 4;
        long long k = kb + w;
        for (; k + 3 * NW < ke; k += 4 * NW) {
            const T4 v0 = __builtin_nontemporal_load(col + k * st4), v1 = __builtin_nontemporal_load(col + (k + NW) * st4),
                     v2 = __builtin_nontemporal_load(col + (k + 2 * NW) * st4),
                     v3 = __builtin_nontemporal_load(col + (k + 3 * NW) * st4);
            acc = app(acc, app(app(v0, v1), app(v2, v3)));
        }
        for (; k < ke; k += NW) acc = app(acc, __builtin_nontemporal_load(col + k * st4));
    }
    lds[threadIdx.x] = acc;
    __syncthreads();
    if (w == 0) {
        T4 r = lds[threadIdx.x];
#pragma unroll
        for (int i = 1; i < NW; ++i) r = app(r, lds[threadIdx.x + 64 * i]);
        acc = r;
    }
    if (ksplit == 1) {
        if (w == 0 && q4 < Q4) *(reinterpret_cast<T4*>(out + p * Q) + q4) = acc;
        return;
    }
    if (w == 0) reinterpret_cast<T4*>(partials)[(long long)blockIdx.x * 64 + threadIdx.x] = acc;
}

// Short folds over many columns (the batch sum of 64 rows of 2^20 values):
// each lane owns U groups of four consecutive q (RBS*4 apart) and folds all
// K rows itself, KU rows' 16-byte loads in flight per group, so a block
// reads runs of 4 KB of each row and no LDS combine is needed. Each output
// folds k in ascending order (a sequential fold). Requirements as
// col_kernel_v4.
template <typename Op, typename T, int KU, int U>
__global__ __launch_bounds__(RBS) void col_stream_v4_kernel(const T* __restrict__ in, long long Q, long long K,
                                                            long long st, long long PS, T* out)
{
    typedef T T4 __attribute__((ext_vector_type(4)));
    const long long Q4 = Q / 4;
    const long long qtiles = (Q4 + RBS * U - 1) / (RBS * U);
    const long long p = blockIdx.x / qtiles;
    const long long q40 = (blockIdx.x % qtiles) * (RBS * U) + threadIdx.x;
    auto app = [](T4 a, T4 b) {
        T4 r;
        r.x = Op::apply(a.x, b.x);
        r.y = Op::apply(a.y, b.y);
        r.z = Op::apply(a.z, b.z);
        r.w = Op::apply(a.w, b.w);
        return r;
    };
    const T id = Identity<Op, T>::v();
    T4 acc[U];
    bool ok[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
        acc[u] = T4{id, id, id, id};
        ok[u] = q40 + u * RBS < Q4;
    }
    const T4* col = reinterpret_cast<const T4*>(in + p * PS) + q40;
    const long long st4 = st / 4;
    const T4 zero = {T(0), T(0), T(0), T(0)};
    long long k = 0;
    for (; k + KU <= K; k += KU) {
        T4 v[KU][U];
#pragma unroll
        for (int j = 0; j < KU; ++j)
#pragma unroll
            for (int u = 0; u < U; ++u)
                v[j][u] = ok[u] ? __builtin_nontemporal_load(col + (k + j) * st4 + u * RBS) : zero;
#pragma unroll
        for (int j = 0; j < KU; ++j)
#pragma unroll
            for (int u = 0; u < U; ++u) acc[u] = app(acc[u], v[j][u]);
    }
    for (; k < K; ++k)
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (ok[u]) acc[u] = app(acc[u], __builtin_nontemporal_load(col + k * st4 + u * RBS));
#pragma unroll
    for (int u = 0; u < U; ++u)
        if (ok[u]) *(reinterpret_cast<T4*>(out + p * Q) + q40 + u * RBS) = acc[u];
}

// Folds the ksplit partial tiles of col_kernel_v4 in split order.
template <typename Op, typename T>
__global__ __launch_bounds__(64) void col_combine_v4_kernel(const T* __restrict__ partials, long long Q, int ksplit,
                                                            T* out)
{
    typedef T T4 __attribute__((ext_vector_type(4)));
    const long long Q4 = Q / 4;
    const long long qtiles = (Q4 + 63) / 64;
    const long long tile = blockIdx.x;
    const long long p = tile / qtiles;
    const long long q4 = (tile % qtiles) * 64 + threadIdx.x;
    const T4* pt = reinterpret_cast<const T4*>(partials);
    T4 r = pt[(tile * ksplit) * 64 + threadIdx.x];
    for (int k2 = 1; k2 < ksplit; ++k2) {
        const T4 v = pt[(tile * ksplit + k2) * 64 + threadIdx.x];
        r.x = Op::apply(r.x, v.x);
        r.y = Op::apply(r.y, v.y);
        r.z = Op::apply(r.z, v.z);
        r.w = Op::apply(r.w, v.w);
    }
    if (q4 < Q4) *(reinterpret_cast<T4*>(out + p * Q) + q4) = r;
}

// ------------------------------------------ 16-byte block-per-chunk path
// Sources of four consecutive elements (aligned, 4-byte types).
template <typename T> struct Src4Plain {
    const T* a;
    typedef T T4 __attribute__((ext_vector_type(4)));
    __device__ __forceinline__ T4 operator()(long long i4) const
    {
        return __builtin_nontemporal_load(reinterpret_cast<const T4*>(a) + i4);
    }
};
template <typename T> struct Src4Dot {
    const T* a;
    const T* b;
    typedef T T4 __attribute__((ext_vector_type(4)));
    __device__ __forceinline__ T4 operator()(long long i4) const
    {
        return __builtin_nontemporal_load(reinterpret_cast<const T4*>(a) + i4) *
               __builtin_nontemporal_load(reinterpret_cast<const T4*>(b) + i4);
    }
};

// seg_block_kernel with 16-byte loads: segment length L4 groups of four;
// each lane folds its groups per component, the four components are folded
// (x op y) op (z op w), then the block reduction as before.
template <typename Op, typename T, typename Src4>
__global__ __launch_bounds__(RBS) void seg_block_v4_kernel(Src4 src, long long L4, int nchunk, long long chunk4,
                                                           T* partials, T* out)
{
    typedef T T4 __attribute__((ext_vector_type(4)));
    __shared__ T lds[RBS / 64];
    auto app = [](T4 a, T4 b) {
        T4 r;
        r.x = Op::apply(a.x, b.x);
        r.y = Op::apply(a.y, b.y);
        r.z = Op::apply(a.z, b.z);
        r.w = Op::apply(a.w, b.w);
        return r;
    };
    const long long s = blockIdx.x / nchunk;
    const int c = blockIdx.x % nchunk;
    const long long beg = c * chunk4;
    const long long end = min(L4, beg + chunk4);
    const long long base = s * L4;
    const T id = Identity<Op, T>::v();
    T4 acc = {id, id, id, id};
    long long j = beg + threadIdx.x;
    for (; j + 3 * RBS < end; j += 4 * RBS) {
        const T4 v0 = src(base + j), v1 = src(base + j + RBS), v2 = src(base + j + 2 * RBS),
                 v3 = src(base + j + 3 * RBS);
        acc = app(acc, app(app(v0, v1), app(v2, v3)));
    }
    for (; j < end; j += RBS) acc = app(acc, src(base + j));
    T a = Op::apply(Op::apply(acc.x, acc.y), Op::apply(acc.z, acc.w));
    a = block_reduce<Op, RBS>(a, lds);
    if (nchunk == 1) {
        if (threadIdx.x == 0) out[s] = a;
        return;
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = a;
}

// Long contiguous segments (L % 4 == 0, 16-byte aligned) through the 16-byte
// block kernel; same chunking as run_segments' block path.
// Where the per-chunk partials of a split reduction live: the context's
// scratch (grown by the caching allocator) unless a fixed buffer is given.
struct PartBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
};

static int partials(cm_ctx_t ctx, const PartBuf& pb, size_t bytes, void** out)
{
    if (pb.ptr) {
        CM_REQUIRE(bytes <= pb.bytes, CM_E_DIM,
                   "reduction partials of %zu bytes exceed the fixed %zu-byte device buffer", bytes, pb.bytes);
        *out = pb.ptr;
        return CM_OK;
    }
    return scratch(ctx, 1, bytes, out);
}

template <typename Op, typename T, typename Src4>
static int run_segments_v4(cm_ctx_t ctx, Src4 src, long long L, long long nseg, T* out,
                           const PartBuf& pb = PartBuf())
{
    const long long cus = ctx->cu_count > 0 ? ctx->cu_count : 256;
    const long long L4 = L / 4;
    long long want_blocks = cus * 8;
    long long nchunk = (want_blocks + nseg - 1) / nseg;
    long long maxchunks = (L4 + 1023) / 1024;  // >= 4096 elements per chunk
    if (nchunk > maxchunks) nchunk = maxchunks;
    if (nchunk < 1) nchunk = 1;
    if (nchunk > 4096) nchunk = 4096;
    long long chunk4 = (L4 + nchunk - 1) / nchunk;
    nchunk = (L4 + chunk4 - 1) / chunk4;
    void* part = nullptr;
    if (nchunk > 1) CM_TRY(partials(ctx, pb, sizeof(T) * nseg * nchunk, &part));
    const long long blocks = nseg * nchunk;
    CM_REQUIRE(blocks < (1ll << 31), CM_E_DIM, "reduction too large");
    seg_block_v4_kernel<Op, T, Src4><<<(unsigned)blocks, RBS, 0, ctx->stream>>>(src, L4, (int)nchunk, chunk4,
                                                                                (T*)part, out);
    CM_LAUNCH_CHECK();
    if (nchunk > 1) {
        seg_combine_kernel<Op, T, T><<<(unsigned)nseg, RBS, 0, ctx->stream>>>((const T*)part, (int)nchunk, out);
        CM_LAUNCH_CHECK();
    }
    return CM_OK;
}

// ---------------------------------------------------------- host side
struct Shape {
    long long I, M, O;  // inner, middle, outer(batch) sizes in memory order
    bool ri, rm, ro;    // reduced flags
};

template <typename Op, typename T, typename S, typename Src>
static int run_segments(cm_ctx_t ctx, Src src, long long L, long long nseg, S* out, bool plain_vec4,
                        const void* plain_ptr, const PartBuf& pb = PartBuf())
{
    if (nseg == 0) return CM_OK;
    if (L == 0) {
        // empty reduction: every output is the identity
        return fail(CM_E_DIM, "reduction over an empty axis");
    }
    if (L <= 16 && nseg >= 64) {
        seg_thread_kernel<Op, T, S, Src><<<div_up(nseg, RBS), RBS, 0, ctx->stream>>>(src, L, nseg, L, out);
        CM_LAUNCH_CHECK();
        return CM_OK;
    }
    const long long cus = ctx->cu_count > 0 ? ctx->cu_count : 256;
    if (L <= 16384 && nseg >= cus * 4) {
        if constexpr (std::is_same<S, T>::value && (std::is_same<T, float>::value || std::is_same<T, int>::value)) {
            if (plain_vec4) {
                seg_wave_vec4_kernel<Op, T><<<div_up(nseg, RBS / 64), RBS, 0, ctx->stream>>>(
                    (const T*)plain_ptr, L / 4, nseg, (T*)out);
                CM_LAUNCH_CHECK();
                return CM_OK;
            }
        }
        {
            seg_wave_kernel<Op, T, S, Src><<<div_up(nseg, RBS / 64), RBS, 0, ctx->stream>>>(src, L, nseg, L, out);
        }
        CM_LAUNCH_CHECK();
        return CM_OK;
    }
    // block per chunk: aim for >= 8 blocks per CU overall, chunks >= 4096
    if constexpr (std::is_same<S, T>::value && sizeof(T) == 4) {
        if (plain_vec4 && L % 4 == 0)
            return run_segments_v4<Op, T>(ctx, Src4Plain<T>{(const T*)plain_ptr}, L, nseg, (T*)out, pb);
    }
    long long want_blocks = cus * 8;
    long long nchunk = (want_blocks + nseg - 1) / nseg;
    long long maxchunks = (L + 4095) / 4096;
    if (nchunk > maxchunks) nchunk = maxchunks;
    if (nchunk < 1) nchunk = 1;
    if (nchunk > 4096) nchunk = 4096;
    long long chunk = (L + nchunk - 1) / nchunk;
    nchunk = (L + chunk - 1) / chunk;
    void* part = nullptr;
    if (nchunk > 1) CM_TRY(partials(ctx, pb, sizeof(T) * nseg * nchunk, &part));
    long long blocks = nseg * nchunk;
    CM_REQUIRE(blocks < (1ll << 31), CM_E_DIM, "reduction too large");
    seg_block_kernel<Op, T, S, Src><<<(unsigned)blocks, RBS, 0, ctx->stream>>>(src, L, L, (int)nchunk, chunk,
                                                                               (T*)part, out);
    CM_LAUNCH_CHECK();
    if (nchunk > 1) {
        seg_combine_kernel<Op, T, S><<<(unsigned)nseg, RBS, 0, ctx->stream>>>((const T*)part, (int)nchunk, out);
        CM_LAUNCH_CHECK();
    }
    return CM_OK;
}

template <typename Op, typename T, typename S>
static int run_columns(cm_ctx_t ctx, const S* in, long long P, long long Q, long long K, long long st,
                       long long PS, S* out)
{
    if (P * Q == 0) return CM_OK;
    if (K == 0) return fail(CM_E_DIM, "reduction over an empty axis");
    const long long cus = ctx->cu_count > 0 ? ctx->cu_count : 256;
    if constexpr (std::is_same<S, T>::value && sizeof(T) == 4) {
        const bool aligned = Q % 4 == 0 && st % 4 == 0 && PS % 4 == 0 && (((uintptr_t)in & 15) == 0) &&
                             (((uintptr_t)out & 15) == 0);
        // long folds: column sum of 1024 x 1024 x 64 51.7 -> 44.9 us against
        // 4-byte loads
        const bool v4 = aligned && K >= 256;
        // short folds (K < 256) over many columns: the streaming kernel with
        // two rows x two column groups of 16-byte loads in flight per lane.
        // Batch sum of 1024 x 1024 x 64 (same box): 40.0 us against 52.1 us
        // for col_kernel; rows x groups 2x4 40.4, 4x4 42.9, 8x1 54.2, 1x4
        // 54.2, 1x8 61.0 us: loads spread over fewer rows (4 MiB apart here)
        // at a time stream faster, but one row at a time leaves too little
        // in flight
        constexpr int SKU = 2, SU = 2;
        const long long stream_blocks = P * ((Q / 4 + RBS * SU - 1) / (RBS * SU));
        if (aligned && !v4 && stream_blocks >= cus) {
            CM_REQUIRE(stream_blocks < (1ll << 31), CM_E_DIM, "reduction too large");
            col_stream_v4_kernel<Op, T, SKU, SU><<<(unsigned)stream_blocks, RBS, 0, ctx->stream>>>(
                (const T*)in, Q, K, st, PS, (T*)out);
            CM_LAUNCH_CHECK();
            return CM_OK;
        }
        if (v4) {
            const long long tiles = P * ((Q / 4 + 63) / 64);
            // at least one column tile per CU: one 1024-thread block per tile
            // folds all K rows (16 wavefronts split them, combined in LDS),
            // so no partials and no second launch
            if (tiles >= cus) {
                CM_REQUIRE(tiles < (1ll << 31), CM_E_DIM, "reduction too large");
                col_kernel_v4<Op, T, 1024><<<(unsigned)tiles, 1024, 0, ctx->stream>>>(
                    (const T*)in, P, Q, K, st, PS, 1, nullptr, (T*)out);
                CM_LAUNCH_CHECK();
                return CM_OK;
            }
            long long ksplit = 1;
            if (tiles < cus * 4) {
                ksplit = (cus * 8 + tiles - 1) / tiles;
                long long maxsplit = (K + 63) / 64;  // at least 64 rows per block
                if (ksplit > maxsplit) ksplit = maxsplit;
                if (ksplit < 1) ksplit = 1;
                if (ksplit > 1024) ksplit = 1024;
            }
            void* part = nullptr;
            if (ksplit > 1) CM_TRY(scratch(ctx, 1, sizeof(T) * tiles * ksplit * 256, &part));
            const long long blocks = tiles * ksplit;
            CM_REQUIRE(blocks < (1ll << 31), CM_E_DIM, "reduction too large");
            col_kernel_v4<Op, T><<<(unsigned)blocks, RBS, 0, ctx->stream>>>((const T*)in, P, Q, K, st, PS,
                                                                             (int)ksplit, (T*)part, (T*)out);
            CM_LAUNCH_CHECK();
            if (ksplit > 1) {
                col_combine_v4_kernel<Op, T><<<(unsigned)tiles, 64, 0, ctx->stream>>>((const T*)part, Q, (int)ksplit,
                                                                                      (T*)out);
                CM_LAUNCH_CHECK();
            }
            return CM_OK;
        }
    }
    long long tiles = P * ((Q + 63) / 64);
    long long ksplit = 1;
    if (tiles < cus * 4) {
        ksplit = (cus * 8 + tiles - 1) / tiles;
        long long maxsplit = (K + 63) / 64;  // at least 64 rows per block
        if (ksplit > maxsplit) ksplit = maxsplit;
        if (ksplit < 1) ksplit = 1;
        if (ksplit > 1024) ksplit = 1024;
    }
    void* part = nullptr;
    if (ksplit > 1) CM_TRY(scratch(ctx, 1, sizeof(T) * tiles * ksplit * 64, &part));
    long long blocks = tiles * ksplit;
    CM_REQUIRE(blocks < (1ll << 31), CM_E_DIM, "reduction too large");
    col_kernel<Op, T, S><<<(unsigned)blocks, RBS, 0, ctx->stream>>>(in, P, Q, K, st, PS, (int)ksplit, (T*)part,
                                                                      out);
    CM_LAUNCH_CHECK();
    if (ksplit > 1) {
        col_combine_kernel<Op, T, S><<<(unsigned)tiles, 64, 0, ctx->stream>>>((const T*)part, Q, (int)ksplit, out);
        CM_LAUNCH_CHECK();
    }
    return CM_OK;
}

// Reduce the simplified shape; output in kept-coordinate memory order.
template <typename Op, typename T, typename S>
static int reduce_shape(cm_ctx_t ctx, const S* in, Shape sh, S* out)
{
    // merge runs: list of (size, reduced) in memory order, skipping size 1
    long long sz[3] = {sh.I, sh.M, sh.O};
    bool rd[3] = {sh.ri, sh.rm, sh.ro};
    long long rs[3];
    bool rr[3];
    int n = 0;
    for (int a = 0; a < 3; ++a) {
        if (sz[a] == 1) continue;
        if (n > 0 && rr[n - 1] == rd[a]) rs[n - 1] *= sz[a];
        else { rs[n] = sz[a]; rr[n] = rd[a]; ++n; }
    }
    const bool vec_ok = (sizeof(S) == 4) && std::is_same<S, T>::value && (((uintptr_t)in & 15) == 0);
    if (n == 0) {  // a single element
        return run_segments<Op, T, S>(ctx, SrcPlain<S, T>{in}, 1, 1, out, false, in);
    }
    if (n == 1) {
        if (!rr[0]) {  // nothing reduced: copy (through a trivial segment reduce of length 1)
            return run_segments<Op, T, S>(ctx, SrcPlain<S, T>{in}, 1, rs[0], out, false, in);
        }
        return run_segments<Op, T, S>(ctx, SrcPlain<S, T>{in}, rs[0], 1, out, vec_ok && rs[0] % 4 == 0, in);
    }
    if (n == 2) {
        if (rr[0]) {  // R K: contiguous segments
            bool v4 = vec_ok && (rs[0] % 4 == 0);
            return run_segments<Op, T, S>(ctx, SrcPlain<S, T>{in}, rs[0], rs[1], out, v4, in);
        }
        // K R: columns
        return run_columns<Op, T, S>(ctx, in, 1, rs[0], rs[1], rs[0], 0, out);
    }
    // n == 3
    if (!rr[0]) {  // K R K
        return run_columns<Op, T, S>(ctx, in, rs[2], rs[0], rs[1], rs[0], rs[0] * rs[1], out);
    }
    // R K R: segments into a temporary (kept-major order), then columns
    void* tmp = nullptr;
    CM_TRY(scratch(ctx, 2, sizeof(S) * rs[1] * rs[2], &tmp));
    bool v4 = vec_ok && (rs[0] % 4 == 0);
    CM_TRY((run_segments<Op, T, S>(ctx, SrcPlain<S, T>{in}, rs[0], rs[1] * rs[2], (S*)tmp, v4, in)));
    return run_columns<Op, T, S>(ctx, (const S*)tmp, 1, rs[1], rs[2], rs[1], 0, out);
}

template <typename T, typename S>
static int dispatch_op(cm_ctx_t ctx, int op, const S* in, Shape sh, S* out)
{
    switch (op) {
    case CM_RED_SUM: return reduce_shape<OpSum, T, S>(ctx, in, sh, out);
    case CM_RED_PROD: return reduce_shape<OpProd, T, S>(ctx, in, sh, out);
    case CM_RED_MIN: return reduce_shape<OpMin, T, S>(ctx, in, sh, out);
    case CM_RED_MAX: return reduce_shape<OpMax, T, S>(ctx, in, sh, out);
    case CM_RED_ALL: return reduce_shape<OpAll, T, S>(ctx, in, sh, out);
    case CM_RED_ANY: return reduce_shape<OpAny, T, S>(ctx, in, sh, out);
    default: return fail(CM_E_UNSUPPORTED, "reduction op %d not available for this type", op);
    }
}
template <typename T, typename S>
static int dispatch_op_int(cm_ctx_t ctx, int op, const S* in, Shape sh, S* out)
{
    if (op == CM_RED_BAND) return reduce_shape<OpBand, T, S>(ctx, in, sh, out);
    if (op == CM_RED_BOR) return reduce_shape<OpBor, T, S>(ctx, in, sh, out);
    return dispatch_op<T, S>(ctx, op, in, sh, out);
}

// Generic relayout copy for outputs whose layout differs from the input's:
// dst[r,c,b] (layout 'orm') = src[r,c,b] (layout 'irm').
template <typename S>
__global__ void relayout_kernel(const S* src, S* dst, long long R, long long C, long long B, int irm, int orm)
{
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long n = R * C * B;
    if (i >= n) return;
    long long r, c, b;
    if (orm) { c = i % C; r = (i / C) % R; b = i / (R * C); }
    else { r = i % R; c = (i / R) % C; b = i / (R * C); }
    long long si = irm ? (c + C * (r + R * b)) : (r + R * (c + C * b));
    dst[i] = src[si];
}

template <typename S>
static int reduce_entry(cm_ctx_t ctx, int dtype, int op, int axis, int rows, int cols, int batches, int irm,
                        const S* in, int orm, S* out)
{
    Shape sh;
    // memory order: column major (rows, cols, batch); row major (cols, rows, batch)
    if (irm) { sh.I = cols; sh.M = rows; sh.ri = axis & CM_AXIS_COL; sh.rm = axis & CM_AXIS_ROW; }
    else { sh.I = rows; sh.M = cols; sh.ri = axis & CM_AXIS_ROW; sh.rm = axis & CM_AXIS_COL; }
    sh.O = batches;
    sh.ro = axis & CM_AXIS_BATCH;
    long long oR = (axis & CM_AXIS_ROW) ? 1 : rows, oC = (axis & CM_AXIS_COL) ? 1 : cols;
    long long oB = (axis & CM_AXIS_BATCH) ? 1 : batches;
    bool same_order = (irm == orm) || oR == 1 || oC == 1;
    S* dst = out;
    if (!same_order) {
        void* tmp = nullptr;
        CM_TRY(scratch(ctx, 3, sizeof(S) * oR * oC * oB, &tmp));
        dst = (S*)tmp;
    }
    int st;
    if (dtype == CM_F32) st = dispatch_op<float, S>(ctx, op, in, sh, dst);
    else if (dtype == CM_F64) st = dispatch_op<double, S>(ctx, op, in, sh, dst);
    else st = dispatch_op_int<int, S>(ctx, op, in, sh, dst);
    CM_TRY(st);
    if (!same_order) {
        long long n = oR * oC * oB;
        relayout_kernel<S><<<div_up(n, 256), 256, 0, ctx->stream>>>(dst, out, oR, oC, oB, irm, orm);
        CM_LAUNCH_CHECK();
    }
    return CM_OK;
}

// dot: per batch sum a*b (a/b optionally broadcast)
template <typename T>
static int dot_entry(cm_ctx_t ctx, long long n, int batches, const T* a, int ab, const T* b, int bb, T* out,
                     const PartBuf& part = PartBuf())
{
    for (int k = 0; k < batches; ++k) {
        // batch-by-batch keeps broadcasting trivial; launches stay few for the
        // batch counts of the solver (<= a handful)
        const T* pa = a + (ab ? (long long)k * n : 0);
        const T* pb = b + (bb ? (long long)k * n : 0);
        if constexpr (sizeof(T) == 4) {
            // 16-byte loads for long aligned vectors (2^26 FP32: see run_segments_v4)
            if (n % 4 == 0 && n >= (1 << 16) && (((uintptr_t)pa | (uintptr_t)pb) & 15) == 0) {
                CM_TRY((run_segments_v4<OpSum, T>(ctx, Src4Dot<T>{pa, pb}, n, 1, out + k, part)));
                continue;
            }
        }
        CM_TRY((run_segments<OpSum, T, T>(ctx, SrcDot<T, T>{pa, pb}, n, 1, out + k, false, nullptr, part)));
    }
    return CM_OK;
}

}  // namespace cm

extern "C" {

int cm_reduce(cm_ctx_t ctx, int dtype, int op, int axis, int rows, int cols, int batches, int in_row_major,
              const void* in, int out_row_major, void* out)
{
    CM_API_BEGIN
    CM_TRY(cm::resolve(ctx, &ctx));
    CM_REQUIRE(rows >= 0 && cols >= 0 && batches >= 0, CM_E_DIM, "cm_reduce: negative size");
    CM_REQUIRE(axis >= 0 && axis <= 7, CM_E_ARG, "cm_reduce: axis %d out of range", axis);
    CM_REQUIRE(op >= CM_RED_SUM && op <= CM_RED_BOR, CM_E_ARG, "cm_reduce: op %d out of range", op);
    cm::count(CM_CNT_EVAL_REDUCTION);
    cm::count(CM_CNT_EVAL_ANY);
    if ((long long)rows * cols * batches == 0) return CM_OK;
    switch (dtype) {
    case CM_F32:
        CM_REQUIRE(op <= CM_RED_ANY, CM_E_UNSUPPORTED, "bitwise reduction on float");
        return cm::reduce_entry<float>(ctx, dtype, op, axis, rows, cols, batches, in_row_major,
                                       (const float*)in, out_row_major, (float*)out);
    case CM_F64:
        CM_REQUIRE(op <= CM_RED_ANY, CM_E_UNSUPPORTED, "bitwise reduction on double");
        return cm::reduce_entry<double>(ctx, dtype, op, axis, rows, cols, batches, in_row_major,
                                        (const double*)in, out_row_major, (double*)out);
    case CM_I32:
        return cm::reduce_entry<int>(ctx, dtype, op, axis, rows, cols, batches, in_row_major, (const int*)in,
                                     out_row_major, (int*)out);
    case CM_BOOL:
        return cm::reduce_entry<unsigned char>(ctx, dtype, op, axis, rows, cols, batches, in_row_major,
                                               (const unsigned char*)in, out_row_major, (unsigned char*)out);
    default:
        return cm::fail(CM_E_UNSUPPORTED, "cm_reduce: dtype %d", dtype);
    }
    CM_API_END
}

int cm_dot(cm_ctx_t ctx, int dtype, int64_t n, int batches, const void* a, int a_batched, const void* b,
           int b_batched, void* out)
{
    CM_API_BEGIN
    CM_TRY(cm::resolve(ctx, &ctx));
    CM_REQUIRE(n >= 0 && batches >= 1, CM_E_DIM, "cm_dot: bad sizes");
    cm::count(CM_CNT_EVAL_REDUCTION);
    cm::count(CM_CNT_EVAL_ANY);
    if (n == 0) return cm_fill_bytes(ctx, out, 0, cm::dtype_size(dtype) * batches);
    switch (dtype) {
    case CM_F32: return cm::dot_entry<float>(ctx, n, batches, (const float*)a, a_batched, (const float*)b, b_batched, (float*)out);
    case CM_F64: return cm::dot_entry<double>(ctx, n, batches, (const double*)a, a_batched, (const double*)b, b_batched, (double*)out);
    case CM_I32: return cm::dot_entry<int>(ctx, n, batches, (const int*)a, a_batched, (const int*)b, b_batched, (int*)out);
    default: return cm::fail(CM_E_UNSUPPORTED, "cm_dot: dtype %d", dtype);
    }
    CM_API_END
}

int cm_squared_norm(cm_ctx_t ctx, int dtype, int64_t n, int batches, const void* a, void* out)
{
    return cm_dot(ctx, dtype, n, batches, a, 1, a, 1, out);
}

int cm_dot_noalloc(cm_ctx_t ctx, int dtype, int64_t n, int batches, const void* a, int a_batched, const void* b,
                   int b_batched, void* out)
{
    CM_API_BEGIN
    // the thread's default context would be created (a stream) on first use
    CM_REQUIRE(ctx != nullptr, CM_E_ARG, "cm_dot_noalloc: needs a context from cm_ctx_create");
    CM_TRY(cm::resolve(ctx, &ctx));
    CM_REQUIRE(n >= 0 && batches >= 1, CM_E_DIM, "cm_dot_noalloc: bad sizes");
    CM_REQUIRE(out && (n == 0 || (a && b)), CM_E_ARG, "cm_dot_noalloc: NULL pointer");
    void* buf = cm::noalloc_scratch_ptr(ctx->device);
    CM_REQUIRE(buf != nullptr, CM_E_STATE, "cm_dot_noalloc: device %d has no fixed buffer (context set-up failed)",
               ctx->device);
    cm::count(CM_CNT_EVAL_REDUCTION);
    cm::count(CM_CNT_EVAL_ANY);
    if (n == 0) return cm_fill_bytes(ctx, out, 0, cm::dtype_size(dtype) * batches);
    const cm::PartBuf pb{buf, NOALLOC_SCRATCH_BYTES};
    switch (dtype) {
    case CM_F32:
        return cm::dot_entry<float>(ctx, n, batches, (const float*)a, a_batched, (const float*)b, b_batched,
                                    (float*)out, pb);
    case CM_F64:
        return cm::dot_entry<double>(ctx, n, batches, (const double*)a, a_batched, (const double*)b, b_batched,
                                     (double*)out, pb);
    case CM_I32:
        return cm::dot_entry<int>(ctx, n, batches, (const int*)a, a_batched, (const int*)b, b_batched, (int*)out,
                                  pb);
    default: return cm::fail(CM_E_UNSUPPORTED, "cm_dot_noalloc: dtype %d", dtype);
    }
    CM_API_END
}

}  // extern "C"
