#ifndef CUMAT_AMD_CONTEXT_H
#define CUMAT_AMD_CONTEXT_H

// Runtime pieces of the host layer: the per-thread Context
// (cuMat/src/Context.h:132-446), the Profiling counters
// (cuMat/src/Profiling.h:23-145) and the reference-counted DevicePointer
// (cuMat/src/DevicePointer.h:11-148). Device memory comes from the library's
// caching allocator (cm_malloc / cm_free); all work is ordered on the
// context's stream.

#include <memory>

#include "Macros.h"

CUMAT_NAMESPACE_BEGIN

class Context {
    cm_ctx_t handle_;
    bool owned_;

    explicit Context(cm_ctx_t h, bool owned) : handle_(h), owned_(owned) {}

public:
    // An explicit context on 'device' with its own stream.
    explicit Context(int device) : handle_(nullptr), owned_(true)
    {
        CUMAT_AMD_CALL(cm_ctx_create(device, &handle_));
    }
    ~Context()
    {
        if (owned_ && handle_) cm_ctx_destroy(handle_);
    }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;

    // The calling thread's context (device 0, its own stream), created on
    // first use like cuMat's Context::current().
    static Context& current()
    {
        static thread_local Context ctx(nullptr, false);
        return ctx;
    }

    cm_ctx_t handle() const { return handle_; }
    void* stream() const
    {
        void* s = nullptr;
        CUMAT_AMD_CALL(cm_ctx_get_stream(handle_, &s));
        return s;
    }
    void setStream(void* stream) { CUMAT_AMD_CALL(cm_ctx_set_stream(handle_, stream)); }
    void synchronize() const { CUMAT_AMD_CALL(cm_ctx_sync(handle_)); }

    void* mallocDevice(size_t bytes) const
    {
        void* p = nullptr;
        CUMAT_AMD_CALL(cm_malloc(handle_, bytes, &p));
        return p;
    }
    void freeDevice(void* p) const { cm_free(handle_, p); }
};

class Profiling {
public:
    enum Counter {
        DeviceMemAlloc = CM_CNT_DEVICE_MEM_ALLOC,
        DeviceMemFree = CM_CNT_DEVICE_MEM_FREE,
        EvalAny = CM_CNT_EVAL_ANY,
        EvalCwise = CM_CNT_EVAL_CWISE,
        EvalCwiseSparse = CM_CNT_EVAL_CWISE_SPARSE,
        EvalReduction = CM_CNT_EVAL_REDUCTION,
        EvalMatmulSparse = CM_CNT_EVAL_MATMUL_SPARSE,
        KernelLaunch = CM_CNT_KERNEL_LAUNCH,
        // no counterpart in cuMat: host waits issued by the library, and
        // kernels compiled / launched for component-wise programs at run time
        HostSync = CM_CNT_HOST_SYNC,
        JitCompile = CM_CNT_JIT_COMPILE,
        JitLaunch = CM_CNT_JIT_LAUNCH
    };
    static Profiling& instance()
    {
        static Profiling p;
        return p;
    }
    void resetAll() { CUMAT_AMD_CALL(cm_counters_reset()); }
    size_t get(Counter c) const
    {
        long long v = 0;
        CUMAT_AMD_CALL(cm_counters_get(c, &v));
        return (size_t)v;
    }
};

// Shared ownership of one device allocation; copies share it, the last owner
// returns it to the caching allocator. An offset view (slice of a batch)
// shares ownership of the whole allocation.
template <typename T>
class DevicePointer {
    std::shared_ptr<T> p_;

public:
    DevicePointer() = default;
    explicit DevicePointer(Index count)
    {
        if (count <= 0) return;
        cm_ctx_t h = Context::current().handle();
        T* raw = static_cast<T*>(Context::current().mallocDevice(sizeof(T) * (size_t)count));
        p_ = std::shared_ptr<T>(raw, [h](T* q) { cm_free(h, q); });
    }
    DevicePointer(const DevicePointer& owner, Index offset) : p_(owner.p_, owner.p_.get() + offset) {}

    T* pointer() const { return p_.get(); }
    long getCounter() const { return p_.use_count(); }
    explicit operator bool() const { return (bool)p_; }
};

CUMAT_NAMESPACE_END

#endif
