// runtime.hip -- contexts, device memory, error strings, profiling counters.
// Replaces cuMat's Context / DevicePointer / Profiling
// (cuMat/src/Context.h:132-446, cuMat/src/DevicePointer.h:11-148,
//  cuMat/src/Profiling.h:23-145) behind the C ABI.
#include "cm_common.hpp"

#include <map>
#include <memory>

namespace cm {

static thread_local std::string g_last_error;
static std::atomic<long long> g_counters[CM_CNT_COUNT];

int fail(int code, const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

void clear_error() { g_last_error.clear(); }

void count(int counter, long long v)
{
    if (counter >= 0 && counter < CM_CNT_COUNT) g_counters[counter].fetch_add(v, std::memory_order_relaxed);
}

// ---- caching allocator: freed blocks are kept per (device, size class) and
// reused, like cuMat's use of cub::CachingDeviceAllocator
// (cuMat/src/Context.h:223-229,271-286,311-325). Frees are stream-ordered:
// a freed block carries the stream that freed it and an event recorded there
// after the work queued so far. The same stream reuses it at once (its later
// work runs after that work); another stream either takes a block whose event
// has completed or makes its stream wait for the event on the device. The
// host never waits in cm_free. Thread-safe.
namespace {
struct Block {
    void* ptr;
    size_t bytes;
    int device;
};
struct FreeBlock {
    void* ptr;
    hipStream_t stream;   // stream of the last free (nullptr: no pending work)
    hipEvent_t event;     // recorded on 'stream' at the free (nullptr: none)
};
std::mutex g_alloc_mutex;
std::multimap<std::pair<int, size_t>, FreeBlock> g_free_blocks;
std::map<void*, Block> g_live_blocks;
std::vector<hipEvent_t> g_event_pool;  // idle events (disable-timing)

size_t size_class(size_t bytes)
{
    if (bytes < 256) return 256;
    size_t c = 256;
    while (c < bytes && c < (size_t(1) << 20)) c <<= 1;
    if (c >= bytes) return c;
    // above 1 MiB: round up to 1 MiB granules
    return (bytes + (size_t(1) << 20) - 1) & ~((size_t(1) << 20) - 1);
}

// caller holds g_alloc_mutex
void event_release(hipEvent_t e)
{
    if (e) g_event_pool.push_back(e);
}
hipEvent_t event_acquire()
{
    if (!g_event_pool.empty()) {
        hipEvent_t e = g_event_pool.back();
        g_event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
    return e;
}
}  // namespace

// Picks a cached block of class 'cls' for 'stream' (caller holds the mutex).
// Returns nullptr when none is usable without a host wait.
static void* take_cached(int device, size_t cls, hipStream_t stream)
{
    auto range = g_free_blocks.equal_range({device, cls});
    if (range.first == range.second) return nullptr;
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    (void)hipStreamIsCapturing(stream, &cap);
    const bool capturing = cap != hipStreamCaptureStatusNone;
    auto pick = range.second;
    // 1. freed on this stream (or nothing pending): stream order suffices
    for (auto it = range.first; it != range.second; ++it)
        if (!it->second.event || it->second.stream == stream) { pick = it; break; }
    // 2. its event has completed
    if (pick == range.second)
        for (auto it = range.first; it != range.second; ++it)
            if (hipEventQuery(it->second.event) == hipSuccess) { pick = it; break; }
    // 3. make this stream wait for the free on the device (not while the
    //    stream is captured: the wait would tie the graph to outside work)
    if (pick == range.second && !capturing) {
        for (auto it = range.first; it != range.second; ++it)
            if (hipStreamWaitEvent(stream, it->second.event, 0) == hipSuccess) { pick = it; break; }
    }
    (void)hipGetLastError();
    if (pick == range.second) return nullptr;
    void* p = pick->second.ptr;
    event_release(pick->second.event);
    g_free_blocks.erase(pick);
    return p;
}

static int device_alloc(int device, size_t bytes, hipStream_t stream, void** out)
{
    size_t cls = size_class(bytes);
    {
        std::lock_guard<std::mutex> lk(g_alloc_mutex);
        void* p = take_cached(device, cls, stream);
        if (p) {
            *out = p;
            g_live_blocks[p] = Block{p, cls, device};
            return CM_OK;
        }
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, cls);
    if (e != hipSuccess) {
        // release the cache of this device and retry once (hipFree waits for
        // the device, so pending work on the cached blocks has finished)
        std::lock_guard<std::mutex> lk(g_alloc_mutex);
        for (auto it = g_free_blocks.begin(); it != g_free_blocks.end();) {
            if (it->first.first == device) {
                (void)hipFree(it->second.ptr);
                event_release(it->second.event);
                it = g_free_blocks.erase(it);
            } else {
                ++it;
            }
        }
        (void)hipGetLastError();
        e = hipMalloc(&p, cls);
        if (e != hipSuccess) return fail(CM_E_NOMEM, "hipMalloc(%zu) failed: %s", cls, hipGetErrorString(e));
    }
    count(CM_CNT_DEVICE_MEM_ALLOC);
    std::lock_guard<std::mutex> lk(g_alloc_mutex);
    g_live_blocks[p] = Block{p, cls, device};
    *out = p;
    return CM_OK;
}

// Returns a block to the cache. ordered: after the work queued on 'stream'
// (which may be the null stream); otherwise no work may use it any more.
static int device_free(void* p, hipStream_t stream, bool ordered)
{
    if (!p) return CM_OK;
    std::lock_guard<std::mutex> lk(g_alloc_mutex);
    auto it = g_live_blocks.find(p);
    if (it == g_live_blocks.end()) return fail(CM_E_ARG, "cm_free: pointer %p was not allocated by cm_malloc", p);
    FreeBlock fb{p, nullptr, nullptr};
    if (ordered) {
        hipEvent_t ev = event_acquire();
        if (ev && hipEventRecord(ev, stream) == hipSuccess) {
            fb.stream = stream;
            fb.event = ev;
        } else {
            // no event: wait here rather than hand out a block still in use
            event_release(ev);
            (void)hipGetLastError();
            count(CM_CNT_HOST_SYNC);
            if (hipStreamSynchronize(stream) != hipSuccess)
                return fail(CM_E_HIP, "cm_free: stream synchronization failed");
        }
    }
    g_free_blocks.insert({{it->second.device, it->second.bytes}, fb});
    g_live_blocks.erase(it);
    count(CM_CNT_DEVICE_MEM_FREE);
    return CM_OK;
}

int host_sync(cm_ctx_t ctx)
{
    count(CM_CNT_HOST_SYNC);
    CM_HIP(hipStreamSynchronize(ctx->stream));
    return CM_OK;
}

static int init_context(cm_context* c, int device)
{
    int n = 0;
    CM_HIP(hipGetDeviceCount(&n));
    CM_REQUIRE(device >= 0 && device < n, CM_E_ARG, "device %d out of range (%d devices)", device, n);
    c->device = device;
    CM_HIP(hipSetDevice(device));
    CM_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    c->owns_stream = true;
    hipDeviceProp_t prop;
    CM_HIP(hipGetDeviceProperties(&prop, device));
    c->cu_count = prop.multiProcessorCount;
    CM_TRY(noalloc_scratch_init(device));
    return CM_OK;
}

static void release_context(cm_context* c)
{
    if (!c) return;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    graph_cache_drop(c);
    for (auto& s : c->scratch)
        if (s.ptr) device_free(s.ptr, nullptr, false);
    if (c->pinned) hipHostFree(c->pinned);
    if (c->owns_stream && c->stream) hipStreamDestroy(c->stream);
}

namespace {
struct DefaultCtx {
    std::unique_ptr<cm_context> ctx;
    ~DefaultCtx()
    {
        // The runtime may already be unloading at thread exit; errors ignored.
        if (ctx) release_context(ctx.get());
    }
};
thread_local DefaultCtx g_default;
}  // namespace

int resolve(cm_ctx_t in, cm_ctx_t* out)
{
    if (in) {
        CM_HIP(hipSetDevice(in->device));
        *out = in;
        return CM_OK;
    }
    if (!g_default.ctx) {
        std::unique_ptr<cm_context> c(new cm_context());
        CM_TRY(init_context(c.get(), 0));
        g_default.ctx = std::move(c);
    }
    CM_HIP(hipSetDevice(g_default.ctx->device));
    *out = g_default.ctx.get();
    return CM_OK;
}

int fill_zero(cm_ctx_t ctx, void* p, size_t bytes);

int scratch(cm_ctx_t ctx, int slot, size_t bytes, void** out)
{
    Scratch& s = ctx->scratch[slot];
    if (s.bytes < bytes) {
        if (s.ptr) {
            // the old block may still be read by queued work: a stream-ordered
            // free keeps it from other streams until that work is done
            CM_TRY(device_free(s.ptr, ctx->stream, true));
            s.ptr = nullptr;
            s.bytes = 0;
        }
        void* p = nullptr;
        CM_TRY(device_alloc(ctx->device, bytes, ctx->stream, &p));
        // zero once so that "last block" counters start armed (a fill kernel,
        // not a memset node, so that the call is safe under stream capture)
        CM_TRY(fill_zero(ctx, p, bytes));
        s.ptr = p;
        s.bytes = bytes;
    }
    *out = s.ptr;
    return CM_OK;
}

int dalloc(cm_ctx_t ctx, size_t bytes, void** out) { return device_alloc(ctx->device, bytes, ctx->stream, out); }
int dfree(cm_ctx_t ctx, void* p) { return device_free(p, ctx ? ctx->stream : nullptr, ctx != nullptr); }

int pinned(cm_ctx_t ctx, void** host, void** dev)
{
    if (!ctx->pinned) {
        void* h = nullptr;
        CM_HIP(hipHostMalloc(&h, CM_PINNED_BYTES, hipHostMallocMapped | hipHostMallocCoherent));
        void* d = nullptr;
        const hipError_t e = hipHostGetDevicePointer(&d, h, 0);
        if (e != hipSuccess) {
            hipHostFree(h);
            return fail(CM_E_HIP, "hipHostGetDevicePointer: %s", hipGetErrorString(e));
        }
        memset(h, 0, CM_PINNED_BYTES);
        ctx->pinned = h;
        ctx->pinned_dev = d;
    }
    *host = ctx->pinned;
    *dev = ctx->pinned_dev;
    return CM_OK;
}

}  // namespace cm

// ---------------------------------------------------------------- kernels
__global__ void cm_fill_bytes_kernel(unsigned char* p, unsigned char v, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    // 16-byte body when aligned, byte tail otherwise
    if (((uintptr_t)p & 15) == 0) {
        uint4 w;
        unsigned int b4 = v * 0x01010101u;
        w.x = w.y = w.z = w.w = b4;
        size_t n16 = n / 16;
        for (size_t k = i; k < n16; k += stride) reinterpret_cast<uint4*>(p)[k] = w;
        for (size_t k = n16 * 16 + i; k < n; k += stride) p[k] = v;
    } else {
        for (size_t k = i; k < n; k += stride) p[k] = v;
    }
}

namespace cm {
int fill_zero(cm_ctx_t ctx, void* p, size_t bytes)
{
    unsigned blocks = (unsigned)std::min<size_t>(div_up((long long)(bytes / 16 + 1), 256), 4096);
    cm_fill_bytes_kernel<<<blocks, 256, 0, ctx->stream>>>((unsigned char*)p, 0, bytes);
    CM_LAUNCH_CHECK();
    return CM_OK;
}
}  // namespace cm

// ---------------------------------------------------------------- C ABI
extern "C" {

int cm_abi_version(void) { return CM_ABI_VERSION; }

const char* cm_last_error(void) { return cm::g_last_error.c_str(); }

int cm_device_count(int* count)
{
    CM_API_BEGIN
    CM_REQUIRE(count, CM_E_ARG, "count is NULL");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *count = n;
    return CM_OK;
    CM_API_END
}

int cm_device_info(int device, char* name, int name_len, int* cu_count, long long* hbm_bytes,
                   int* clock_khz, int* mem_clock_khz, int* bus_width_bits)
{
    CM_API_BEGIN
    hipDeviceProp_t p;
    CM_HIP(hipGetDeviceProperties(&p, device));
    if (name && name_len > 0) {
        snprintf(name, (size_t)name_len, "%s (%s)", p.name, p.gcnArchName);
    }
    if (cu_count) *cu_count = p.multiProcessorCount;
    if (hbm_bytes) *hbm_bytes = (long long)p.totalGlobalMem;
    if (clock_khz) *clock_khz = p.clockRate;
    if (mem_clock_khz) *mem_clock_khz = p.memoryClockRate;
    if (bus_width_bits) *bus_width_bits = p.memoryBusWidth;
    return CM_OK;
    CM_API_END
}

int cm_ctx_create(int device, cm_ctx_t* out)
{
    CM_API_BEGIN
    CM_REQUIRE(out, CM_E_ARG, "out is NULL");
    std::unique_ptr<cm_context> c(new cm_context());
    CM_TRY(cm::init_context(c.get(), device));
    *out = c.release();
    return CM_OK;
    CM_API_END
}

int cm_ctx_destroy(cm_ctx_t ctx)
{
    CM_API_BEGIN
    if (!ctx) return CM_OK;
    cm::release_context(ctx);
    delete ctx;
    return CM_OK;
    CM_API_END
}

int cm_ctx_get_stream(cm_ctx_t ctx, void** stream)
{
    CM_API_BEGIN
    CM_TRY(cm::resolve(ctx, &ctx));
    *stream = (void*)ctx->stream;
    return CM_OK;
    CM_API_END
}

int cm_ctx_set_stream(cm_ctx_t ctx, void* stream)
{
    CM_API_BEGIN
    CM_TRY(cm::resolve(ctx, &ctx));
    CM_TRY(::cm::host_sync(ctx));
    if (ctx->owns_stream && ctx->stream) CM_HIP(hipStreamDestroy(ctx->stream));
    ctx->stream = (hipStream_t)stream;
    ctx->owns_stream = false;
    return CM_OK;
    CM_API_END
}

int cm_ctx_sync(cm_ctx_t ctx)
{
    CM_API_BEGIN
    CM_TRY(cm::resolve(ctx, &ctx));
    CM_TRY(cm::host_sync(ctx));
    return CM_OK;
    CM_API_END
}

int cm_malloc(cm_ctx_t ctx, size_t bytes, void** ptr)
{
    CM_API_BEGIN
    CM_REQUIRE(ptr, CM_E_ARG, "ptr is NULL");
    CM_TRY(cm::resolve(ctx, &ctx));
    if (bytes == 0) { *ptr = nullptr; return CM_OK; }
    return cm::device_alloc(ctx->device, bytes, ctx->stream, ptr);
    CM_API_END
}

int cm_free(cm_ctx_t ctx, void* ptr)
{
    CM_API_BEGIN
    if (!ptr) return CM_OK;
    CM_TRY(cm::resolve(ctx, &ctx));
    // stream-ordered: queued work on this stream may still read the block
    return cm::device_free(ptr, ctx->stream, true);
    CM_API_END
}

int cm_memcpy_h2d(cm_ctx_t ctx, void* dst, const void* src, size_t bytes)
{
    CM_API_BEGIN
    CM_TRY(cm::resolve(ctx, &ctx));
    if (!bytes) return CM_OK;
    CM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    // a pageable source may be reused by the caller on return, so that copy
    // is waited for; pinned host memory stays stream-ordered
    hipPointerAttribute_t attr;
    const bool pinned_src = hipPointerGetAttributes(&attr, src) == hipSuccess && attr.type == hipMemoryTypeHost;
    (void)hipGetLastError();
    if (!pinned_src) CM_TRY(cm::host_sync(ctx));
    return CM_OK;
    CM_API_END
}

int cm_memcpy_d2h(cm_ctx_t ctx, void* dst, const void* src, size_t bytes)
{
    CM_API_BEGIN
    CM_TRY(cm::resolve(ctx, &ctx));
    if (!bytes) return CM_OK;
    CM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    CM_TRY(cm::host_sync(ctx));
    return CM_OK;
    CM_API_END
}

int cm_memcpy_d2d(cm_ctx_t ctx, void* dst, const void* src, size_t bytes)
{
    CM_API_BEGIN
    CM_TRY(cm::resolve(ctx, &ctx));
    if (!bytes) return CM_OK;
    CM_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    return CM_OK;
    CM_API_END
}

int cm_fill_bytes(cm_ctx_t ctx, void* ptr, int value, size_t bytes)
{
    CM_API_BEGIN
    CM_TRY(cm::resolve(ctx, &ctx));
    if (!bytes) return CM_OK;
    CM_REQUIRE(ptr, CM_E_ARG, "ptr is NULL");
    unsigned blocks = (unsigned)std::min<size_t>(cm::div_up((long long)(bytes / 16 + 1), 256), 4096);
    cm_fill_bytes_kernel<<<blocks, 256, 0, ctx->stream>>>((unsigned char*)ptr, (unsigned char)value, bytes);
    CM_LAUNCH_CHECK();
    return CM_OK;
    CM_API_END
}

int cm_counters_get(int counter, long long* value)
{
    CM_API_BEGIN
    CM_REQUIRE(counter >= 0 && counter < CM_CNT_COUNT, CM_E_ARG, "counter %d out of range", counter);
    CM_REQUIRE(value, CM_E_ARG, "value is NULL");
    *value = cm::g_counters[counter].load();
    return CM_OK;
    CM_API_END
}

int cm_counters_reset(void)
{
    for (auto& c : cm::g_counters) c.store(0);
    return CM_OK;
}

}  // extern "C"
