// cm_common.hpp -- shared pieces of the gfx950 library: error plumbing of the
// C ABI, the context object, scratch memory, launch helpers, and wave64
// DPP reductions used by every reduction-shaped kernel.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_bf16.h>

#include <atomic>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "cumat_amd.h"
#include "cwise_ops.hpp"
#include "fold_ops.hpp"

namespace cm {

// ------------------------------------------------------------------ errors
int fail(int code, const char* fmt, ...);
void clear_error();

#define CM_HIP(call)                                                                  \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess)                                                         \
            return ::cm::fail(CM_E_HIP, "%s failed: %s (%s:%d)", #call,              \
                              hipGetErrorString(e_), __FILE__, __LINE__);             \
    } while (0)

#define CM_LAUNCH_CHECK()                                                             \
    do {                                                                              \
        hipError_t e_ = hipGetLastError();                                            \
        if (e_ != hipSuccess)                                                         \
            return ::cm::fail(CM_E_HIP, "kernel launch failed: %s (%s:%d)",          \
                              hipGetErrorString(e_), __FILE__, __LINE__);             \
        ::cm::count(CM_CNT_KERNEL_LAUNCH);                                            \
    } while (0)

#define CM_REQUIRE(cond, code, ...)                                                   \
    do {                                                                              \
        if (!(cond)) return ::cm::fail(code, __VA_ARGS__);                            \
    } while (0)

#define CM_TRY(expr)                                                                  \
    do {                                                                              \
        int s_ = (expr);                                                              \
        if (s_ != CM_OK) return s_;                                                   \
    } while (0)

// Entry-point guard: the C ABI never lets an exception escape.
#define CM_API_BEGIN try {
#define CM_API_END                                                                    \
    }                                                                                 \
    catch (const std::exception& ex) { return ::cm::fail(CM_E_STATE, "%s", ex.what()); } \
    catch (...) { return ::cm::fail(CM_E_STATE, "unknown exception"); }

// ---------------------------------------------------------------- counters
void count(int counter, long long v = 1);

// ----------------------------------------------------------------- context
struct Scratch {
    void* ptr = nullptr;
    size_t bytes = 0;
};

}  // namespace cm

struct cm_context {
    int device = 0;
    hipStream_t stream = nullptr;
    bool owns_stream = false;
    // Device scratch (grow-only) for reduction partials and solver state.
    cm::Scratch scratch[4];
    // Pinned host word(s) for flag read-back.
    void* pinned = nullptr;      // host-mapped, coherent page (CM_PINNED_BYTES)
    void* pinned_dev = nullptr;  // its device address
    // Cached device properties.
    int cu_count = 0;
};

namespace cm {

// Resolve a possibly-NULL context to the thread's default one and make its
// device current.
int resolve(cm_ctx_t in, cm_ctx_t* out);
// Zero device memory with a fill kernel (capture-safe).
int fill_zero(cm_ctx_t ctx, void* p, size_t bytes);
// Grow-only scratch slot 'slot' of at least 'bytes' bytes (256-aligned).
int scratch(cm_ctx_t ctx, int slot, size_t bytes, void** out);
// Pinned host memory of the context (at least 256 bytes).
// A per-context page of host memory that kernels write and the host reads
// while they run (coherent: device stores reach it without a cache flush).
constexpr size_t CM_PINNED_BYTES = 4096;
int pinned(cm_ctx_t ctx, void** host, void** dev);
// Caching device allocator (the same pool as cm_malloc / cm_free).
int dalloc(cm_ctx_t ctx, size_t bytes, void** out);
// Stream-ordered on ctx's stream; ctx == nullptr: no queued work uses p.
int dfree(cm_ctx_t ctx, void* p);
// Drops the captured batched-CG graphs of a context (before it is destroyed).
void graph_cache_drop(cm_ctx_t ctx);
// Looks up the device address of cm_dot_noalloc's fixed buffer on 'device'
// (current); once per device, at context creation.
int noalloc_scratch_init(int device);

// Host wait for the context's stream (counted as CM_CNT_HOST_SYNC).
int host_sync(cm_ctx_t ctx);

// RAII holder for temporaries of one entry point; the blocks return to the
// cache when the holder dies, stream-ordered after the work queued with them.
struct Temps {
    cm_ctx_t ctx;
    std::vector<void*> ptrs;
    // set once the caller has synchronised the stream after the last use:
    // the blocks then return to the cache without a recorded event
    bool drained = false;
    explicit Temps(cm_ctx_t c) : ctx(c) {}
    int get(size_t bytes, void** out)
    {
        int s = dalloc(ctx, bytes, out);
        if (s == CM_OK) ptrs.push_back(*out);
        return s;
    }
    ~Temps()
    {
        for (void* p : ptrs) (void)dfree(drained ? nullptr : ctx, p);
    }
};

inline size_t dtype_size(int dt)
{
    switch (dt) {
    case CM_I32: return 4;
    case CM_F32: return 4;
    case CM_F64: return 8;
    case CM_BOOL: return 1;
    case CM_BF16: return 2;
    default: return 0;
    }
}

inline unsigned div_up(long long a, long long b) { return (unsigned)((a + b - 1) / b); }
__device__ __forceinline__ int div_up_dev(int a, int b) { return (a + b - 1) / b; }

// -------------------------------------------------------------- device side
// DPP helpers, the reduction operators and the wave / block reductions
// live in fold_ops.hpp (shared with the run-time generated kernels).

// Two simultaneous block sums (shares the barriers). lds needs 2*BS/64.
template <int BS, typename T>
__device__ __forceinline__ void block_sum2(T& a, T& b, T* lds)
{
    constexpr int NW = BS / 64;
    a = wave_reduce<OpSum>(a);
    b = wave_reduce<OpSum>(b);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { lds[w] = a; lds[NW + w] = b; }
    __syncthreads();
    T ra = lds[0], rb = lds[NW];
#pragma unroll
    for (int i = 1; i < NW; ++i) { ra += lds[i]; rb += lds[NW + i]; }
    __syncthreads();
    a = ra;
    b = rb;
}

// bf16 storage helpers (round-to-nearest-even on conversion from float).
__device__ __forceinline__ float bf16_to_float(unsigned short h)
{
    return __uint_as_float(((unsigned)h) << 16);
}

// Value loader: converts stored type S to compute type T.
template <typename S, typename T> struct Load {
    __device__ __forceinline__ static T get(const S* p, long long i) { return (T)p[i]; }
};
template <typename T> struct Load<unsigned short, T> {
    __device__ __forceinline__ static T get(const unsigned short* p, long long i)
    {
        return (T)bf16_to_float(p[i]);
    }
};

// Apply an assignment mode (CwiseAssignmentHandler, cuMat/src/CwiseOp.h:15-62).
template <typename T>
__device__ __forceinline__ T assign_mode(int mode, T old, T v)
{
    return ops::assign1<T>(mode, old, v);
}

}  // namespace cm
