// cg.hip -- Jacobi-preconditioned Conjugate Gradient on CSR matrices.
//
// Replaces ConjugateGradient<SparseMatrix<S,1,CSR>, DiagonalPreconditioner>
// ::_solve_with_guess_impl (cuMat/src/ConjugateGradient.h:123-206). The
// reference issues ~11 launches and one device->host copy + stream sync per
// iteration. Here every scalar (alpha, beta, |r|^2, r.z, the convergence
// test, the iteration count) stays on the device:
//
//  MULTIKERNEL path (any n, up to 16 right-hand sides; cg_multi.hpp):
//    setup   r = b - A x, p = M^-1 r, |b|^2, |r|^2, r.p, early outs
//    per iteration, three launches (K1 SpMV + p.Ap partials, K2 x/r update
//    + |r|^2, r.z partials, K3 convergence test + p update) or, on the
//    FUSED path, two (KA: convergence test of the previous iteration, p
//    update fused into the SpMV, p.Ap partials; KB: as K2); consumers
//    re-reduce the producers' partials in a fixed order, so no device-scope
//    fence is needed. Kernels launched after convergence return at once;
//    the kernels publish the status to host-mapped memory every few
//    iterations, where the host reads it without a copy and stops
//    launching.
//  RESIDENT path (n <= 16384, one right-hand side): one workgroup runs the
//    whole solve; p lives in LDS, x / r / M^-1 in registers, the reductions
//    are DPP wave reductions + LDS, and no kernel boundary is crossed.
//
// Both follow the reference's semantics exactly: threshold tol^2 |b|^2,
// early outs, the converging iteration is not counted, error =
// sqrt(max_b |r|^2 / |b|^2), all batches iterate until all have converged.
#include "cg_multi.hpp"

#include <algorithm>
#include <climits>
#include <chrono>
#include <cmath>
#include <map>
#include <tuple>

namespace cm {

// ================================================================= resident
// One workgroup of BS threads solves one system with n <= BS*RPT rows; p in
// LDS, everything else in registers. Used for single small systems and, over
// many systems, by the persistent batched solver (sys_rows/row_off != NULL).
template <typename T>
struct ResidentArgs {
    int nsys;
    const int* row_off;   // nsys+1 (NULL: a single system of n rows)
    const int* order;     // processing order of systems (NULL: identity)
    int n_single;
    const int* JA;
    const int* IA;
    const T* vals;
    const T* invd;        // may be NULL: computed from the diagonal
    const T* b;
    T* x;
    int use_guess;
    T tol;
    int max_iter;
    int* iters_out;       // per system
    T* err_out;           // per system
    unsigned* queue;      // queue[0]: work counter, queue[1]: invalid-pattern
                          // flag (both zeroed before launch)
    int nnz_total;        // entries of IA / vals (every JA value must lie in
                          // [0, nnz_total])
};

template <typename T, int BS, int RPT>
__global__ __launch_bounds__(BS) void cg_resident_kernel(const ResidentArgs<T> A)
{
    extern __shared__ unsigned char smem_raw[];
    T* pl = reinterpret_cast<T*>(smem_raw);
    __shared__ T red[2 * BS / 64];
    __shared__ int s_sys;
    __shared__ int s_bad;
    const int tid = threadIdx.x;
    for (;;) {
        // ---- fetch the next system (dynamic queue balances uneven sizes)
        if (tid == 0) {
            s_sys = (int)atomicAdd(A.queue, 1u);
            s_bad = 0;
        }
        __syncthreads();
        const int q = s_sys;
        __syncthreads();
        if (q >= A.nsys) break;
        const int sys = A.order ? A.order[q] : q;
        const int row0 = A.row_off ? A.row_off[sys] : 0;
        const int n = A.row_off ? (A.row_off[sys + 1] - row0) : A.n_single;
        // ---- the pattern must keep every row inside its own system: p lives
        // in LDS indexed by IA - row0, so a column outside [row0, row0 + n)
        // or a row pointer outside [0, nnz_total] would leave the tile. Such a
        // system is not solved: iterations -1, error NaN, queue[1] set.
        bool bad = false;
#pragma unroll
        for (int j = 0; j < RPT; ++j) {
            const int i = tid + j * BS;
            if (i < n) {
                const int s = A.JA[row0 + i], e = A.JA[row0 + i + 1];
                if (s < 0 || e < s || e > A.nnz_total) {
                    bad = true;
                } else {
                    for (int k = s; k < e; ++k) {
                        const int c = A.IA[k] - row0;
                        bad |= (c < 0) | (c >= n);
                    }
                }
            }
        }
        // one flag for the workgroup, read into a scalar register so that the
        // branches on it are uniform
        if (bad) s_bad = 1;
        __syncthreads();
        // Every path ends at the barrier at the end of the loop body: the
        // loop has no early "continue". (With one, the compiled loop sent the
        // lanes other than thread 0 straight back to the barrier at its top
        // while thread 0 waited to fetch the next system, so a workgroup that
        // met an invalid system spun on it for ever.)
        const bool invalid = __builtin_amdgcn_readfirstlane(s_bad) != 0;
        if (invalid && tid == 0) {
            if (A.iters_out) A.iters_out[sys] = -1;
            if (A.err_out) A.err_out[sys] = T(NAN);
            atomicOr(A.queue + 1, 1u);
        }
        if (!invalid) {
            T xr[RPT], rr[RPT], id[RPT];
            // ---- setup: invdiag, x, r = b - A x
#pragma unroll
            for (int j = 0; j < RPT; ++j) {
                int i = tid + j * BS;
                xr[j] = T(0);
                id[j] = T(1);
                if (i < n) {
                    if (A.invd) id[j] = A.invd[row0 + i];
                    else {
                        T d = T(0);
                        for (int k = A.JA[row0 + i]; k < A.JA[row0 + i + 1]; ++k)
                            if (A.IA[k] - row0 == i) { d = A.vals[k]; break; }
                        id[j] = (d == T(0)) ? T(1) : T(1) / d;
                    }
                    if (A.use_guess) xr[j] = A.x[row0 + i];
                    pl[i] = xr[j];
                }
            }
            __syncthreads();
            T bb = 0, r2 = 0;
#pragma unroll
            for (int j = 0; j < RPT; ++j) {
                int i = tid + j * BS;
                rr[j] = T(0);
                if (i < n) {
                    T ax = T(0);
                    if (A.use_guess) {
                        int s = A.JA[row0 + i], e = A.JA[row0 + i + 1];
                        if (s < e) {
                            ax = A.vals[s] * pl[A.IA[s] - row0];
                            for (int k = s + 1; k < e; ++k) ax += A.vals[k] * pl[A.IA[k] - row0];
                        }
                    }
                    T bv = A.b[row0 + i];
                    rr[j] = bv - ax;
                    bb += bv * bv;
                    r2 += rr[j] * rr[j];
                }
            }
            block_sum2<BS>(bb, r2, red);
            const T thr = A.tol * A.tol * bb;
            int iters = 0;
            T err = T(0);
            bool done = false;
            if (bb < A.tol * A.tol) {  // zero right-hand side
#pragma unroll
                for (int j = 0; j < RPT; ++j) xr[j] = T(0);
                done = true;
            } else if (r2 < thr) {  // the guess is already converged
                err = sqrt(r2 / bb);
                done = true;
            }
            if (!done) {
                // p = M^-1 r -> LDS, absNew = r.p
                T absNew = 0, dummy = 0;
#pragma unroll
                for (int j = 0; j < RPT; ++j) {
                    int i = tid + j * BS;
                    if (i < n) {
                        T pv = id[j] * rr[j];
                        pl[i] = pv;
                        absNew += rr[j] * pv;
                    }
                }
                block_sum2<BS>(absNew, dummy, red);  // also orders the LDS writes
                T rsq = r2;
                int it = 0;
                while (it < A.max_iter) {
                    // tmp = A p ; p.tmp
                    T tmp[RPT];
                    T pap = 0;
#pragma unroll
                    for (int j = 0; j < RPT; ++j) {
                        int i = tid + j * BS;
                        tmp[j] = T(0);
                        if (i < n) {
                            int s = A.JA[row0 + i], e = A.JA[row0 + i + 1];
                            T acc = T(0);
                            if (s < e) {
                                acc = A.vals[s] * pl[A.IA[s] - row0];
                                for (int k = s + 1; k < e; ++k) acc += A.vals[k] * pl[A.IA[k] - row0];
                            }
                            tmp[j] = acc;
                            pap += pl[i] * acc;
                        }
                    }
                    T z0 = 0;
                    block_sum2<BS>(pap, z0, red);
                    const T alpha = absNew / pap;
                    T sr = 0, sz = 0;
#pragma unroll
                    for (int j = 0; j < RPT; ++j) {
                        int i = tid + j * BS;
                        if (i < n) {
                            xr[j] += alpha * pl[i];
                            rr[j] -= alpha * tmp[j];
                            sr += rr[j] * rr[j];
                            sz += rr[j] * (id[j] * rr[j]);
                        }
                    }
                    block_sum2<BS>(sr, sz, red);
                    rsq = sr;
                    if (sr < thr) break;
                    const T beta = sz / absNew;
                    absNew = sz;
#pragma unroll
                    for (int j = 0; j < RPT; ++j) {
                        int i = tid + j * BS;
                        if (i < n) pl[i] = id[j] * rr[j] + beta * pl[i];
                    }
                    __syncthreads();
                    ++it;
                }
                iters = it;
                err = sqrt(rsq / bb);
            }
#pragma unroll
            for (int j = 0; j < RPT; ++j) {
                int i = tid + j * BS;
                if (i < n) A.x[row0 + i] = xr[j];
            }
            if (tid == 0) {
                if (A.iters_out) A.iters_out[sys] = iters;
                if (A.err_out) A.err_out[sys] = err;
            }
        }
        __syncthreads();
    }
}

// --------------------------------------------------------------- host side
// grid <= 0: a persistent grid of 'gpc' workgroups per CU (gpc <= 0: as
// many as the occupancy calculator allows; the workgroups pop systems from
// the queue, so more than can be resident only wait their turn).
template <typename T, int BS, int RPT>
static int launch_resident_one(cm_ctx_t ctx, const ResidentArgs<T>& A, int grid, size_t lds, int gpc)
{
    auto k = cg_resident_kernel<T, BS, RPT>;
    CM_HIP(hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    if (grid <= 0) {
        int per_cu = 0;
        CM_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k, BS, lds));
        if (per_cu < 1) return fail(CM_E_UNSUPPORTED, "resident CG: tile does not fit on a CU");
        if (gpc > 0) per_cu = gpc;
        grid = std::min(A.nsys, per_cu * (ctx->cu_count > 0 ? ctx->cu_count : 256));
    }
    if (grid <= 0) return CM_OK;  // nothing to solve (attribute set up only)
    k<<<grid, BS, lds, ctx->stream>>>(A);
    CM_LAUNCH_CHECK();
    return CM_OK;
}

// Picks the register tile RPT for a block size and the largest system.
template <typename T, int BS>
static int launch_resident_bs(cm_ctx_t ctx, const ResidentArgs<T>& A, int nmax, int grid, int gpc)
{
    const size_t lds = sizeof(T) * (size_t)nmax;
    const int rpt = (nmax + BS - 1) / BS;
    if (rpt <= 1) return launch_resident_one<T, BS, 1>(ctx, A, grid, lds, gpc);
    if (rpt <= 2) return launch_resident_one<T, BS, 2>(ctx, A, grid, lds, gpc);
    if (rpt <= 4) return launch_resident_one<T, BS, 4>(ctx, A, grid, lds, gpc);
    if (rpt <= 8) return launch_resident_one<T, BS, 8>(ctx, A, grid, lds, gpc);
    if (rpt <= 12) return launch_resident_one<T, BS, 12>(ctx, A, grid, lds, gpc);
    if (rpt <= 16) return launch_resident_one<T, BS, 16>(ctx, A, grid, lds, gpc);
    if constexpr (BS <= 512) {
        if (rpt <= 24) return launch_resident_one<T, BS, 24>(ctx, A, grid, lds, gpc);
        if (rpt <= 32) return launch_resident_one<T, BS, 32>(ctx, A, grid, lds, gpc);
    }
    return fail(CM_E_UNSUPPORTED, "resident CG: %d rows exceed %d threads x tile", nmax, BS);
}

template <typename T>
static int launch_resident(cm_ctx_t ctx, const ResidentArgs<T>& A, int nmax, int bs, int grid, int gpc = 0)
{
    switch (bs) {
    case 256: return launch_resident_bs<T, 256>(ctx, A, nmax, grid, gpc);
    case 512: return launch_resident_bs<T, 512>(ctx, A, nmax, grid, gpc);
    case 1024: return launch_resident_bs<T, 1024>(ctx, A, nmax, grid, gpc);
    default: return fail(CM_E_ARG, "resident CG: block size %d (256, 512 or 1024)", bs);
    }
}

template <typename T>
static int pcg_resident(cm_ctx_t ctx, int n, int nnz, const int* JA, const int* IA, const T* vals, const T* invd,
                        const T* b, T* x, int use_guess, double tol, int max_iter, int* iters_out, double* err_out,
                        cm_cg_stats* stats)
{
    Temps tmp(ctx);
    void* q;
    void* io;
    void* eo;
    CM_TRY(tmp.get(256, &q));
    CM_TRY(tmp.get(256, &io));
    CM_TRY(tmp.get(256, &eo));
    CM_TRY(fill_zero(ctx, q, 256));
    ResidentArgs<T> A{};
    A.nsys = 1;
    A.n_single = n;
    A.JA = JA;
    A.IA = IA;
    A.vals = vals;
    A.invd = invd;
    A.b = b;
    A.x = x;
    A.use_guess = use_guess;
    A.tol = (T)tol;
    A.max_iter = max_iter;
    A.iters_out = (int*)io;
    A.err_out = (T*)eo;
    A.queue = (unsigned*)q;
    A.nnz_total = nnz;
    auto t0 = std::chrono::steady_clock::now();
    CM_TRY(launch_resident<T>(ctx, A, n, 1024, 1));
    int hi = 0;
    T he = 0;
    unsigned bad = 0;
    CM_HIP(hipMemcpyAsync(&hi, io, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    CM_HIP(hipMemcpyAsync(&he, eo, sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
    CM_HIP(hipMemcpyAsync(&bad, (unsigned*)q + 1, sizeof(unsigned), hipMemcpyDeviceToHost, ctx->stream));
    CM_TRY(::cm::host_sync(ctx));
    CM_REQUIRE(!bad, CM_E_ARG, "resident CG: a row pointer or column index lies outside the %d x %d matrix", n, n);
    auto t1 = std::chrono::steady_clock::now();
    if (iters_out) *iters_out = hi;
    if (err_out) *err_out = (double)he;
    if (stats) {
        stats->path_used = CM_CG_RESIDENT;
        stats->launches = 2;
        stats->seconds = std::chrono::duration<double>(t1 - t0).count();
    }
    return CM_OK;
}

// ------------------------------------------------------ batched systems
struct BatchedPlan {
    int nmax = 0;
    std::vector<int> order;  // systems by decreasing nnz (longest first)
    uint64_t offsets_hash = 0;  // FNV-1a of row_off: identifies the solve order
};

static int make_plan(cm_ctx_t ctx, int nsys, const int* row_off_dev, int total_rows, BatchedPlan& plan)
{
    std::vector<int> ro(nsys + 1);
    CM_HIP(hipMemcpyAsync(ro.data(), row_off_dev, sizeof(int) * (nsys + 1), hipMemcpyDeviceToHost, ctx->stream));
    CM_TRY(::cm::host_sync(ctx));
    CM_REQUIRE(ro[0] >= 0 && ro[nsys] <= total_rows, CM_E_ARG,
               "batched CG: row offsets [%d, %d] outside the %d rows", ro[0], ro[nsys], total_rows);
    plan.nmax = 0;
    uint64_t h = 1469598103934665603ull;
    for (int v : ro) h = (h ^ (uint32_t)v) * 1099511628211ull;
    plan.offsets_hash = h;
    std::vector<std::pair<long long, int>> cost(nsys);
    for (int s = 0; s < nsys; ++s) {
        int n = ro[s + 1] - ro[s];
        CM_REQUIRE(n >= 0, CM_E_ARG, "batched CG: row offsets decrease at system %d", s);
        plan.nmax = std::max(plan.nmax, n);
        // work per iteration ~ n and CG iterations ~ sqrt(n) for grid
        // Laplacians: schedule the heaviest systems first (LPT order)
        cost[s] = {-(long long)((double)n * std::sqrt((double)n)), s};
    }
    std::sort(cost.begin(), cost.end());
    plan.order.resize(nsys);
    for (int s = 0; s < nsys; ++s) plan.order[s] = cost[s].second;
    return CM_OK;
}

// A captured solve fixes the launch shape (block size, register tile and LDS
// size, both from the largest system), the solve order and every pointer, so
// all of them are part of the key.
struct GraphKey {
    cm_ctx_t ctx;
    const void* ptrs[9];
    int nsys, bs, gpc, max_iter, nmax, total_rows, total_nnz;
    float tol;
    uint64_t offsets_hash;
    bool operator<(const GraphKey& o) const { return memcmp(this, &o, sizeof(GraphKey)) < 0; }
};
struct GraphEntry {
    hipGraphExec_t exec = nullptr;
    void* order = nullptr;
    void* queue = nullptr;
    unsigned long long last_use = 0;
};
// Bounded caches: at most this many captured solves (least recently used
// first out) and autotuned shapes.
constexpr size_t GRAPH_CACHE_MAX = 8;
constexpr size_t TUNE_CACHE_MAX = 64;
static std::mutex g_graph_mutex;
static std::map<GraphKey, GraphEntry> g_graphs;
static unsigned long long g_graph_clock = 0;
// autotuned launch shape per batch shape: (block size, workgroups per CU)
static std::map<std::tuple<int, int, int, int>, std::pair<int, int>> g_tuned_bs;

// Destroys a cached graph and returns its blocks (caller holds the mutex; the
// stream is idle after the synchronous solve that last used it).
static void graph_release(GraphEntry& ge)
{
    if (ge.exec) (void)hipGraphExecDestroy(ge.exec);
    if (ge.order) (void)dfree(nullptr, ge.order);
    if (ge.queue) (void)dfree(nullptr, ge.queue);
    ge = GraphEntry();
}

// Removes every cached graph of one context (called before it is destroyed).
void graph_cache_drop(cm_ctx_t ctx)
{
    std::lock_guard<std::mutex> lk(g_graph_mutex);
    for (auto it = g_graphs.begin(); it != g_graphs.end();) {
        if (it->first.ctx == ctx) {
            graph_release(it->second);
            it = g_graphs.erase(it);
        } else {
            ++it;
        }
    }
}

static int batched_launch(cm_ctx_t ctx, int nsys, const int* row_off, int total_nnz, const int* JA, const int* IA,
                          const float* vals, const float* b, float* x, float tol, int max_iter, int* iters,
                          float* errors, int bs, const BatchedPlan& plan, void* order_dev, void* queue_dev,
                          int gpc = 0)
{
    CM_TRY(fill_zero(ctx, queue_dev, 256));
    ResidentArgs<float> A{};
    A.nsys = nsys;
    A.row_off = row_off;
    A.order = (const int*)order_dev;
    A.JA = JA;
    A.IA = IA;
    A.vals = vals;
    A.invd = nullptr;
    A.b = b;
    A.x = x;
    A.use_guess = 0;
    A.tol = tol;
    A.max_iter = max_iter;
    A.iters_out = iters;
    A.err_out = errors;
    A.queue = (unsigned*)queue_dev;
    A.nnz_total = total_nnz;
    return launch_resident<float>(ctx, A, plan.nmax, bs, 0, gpc);
}

static int default_bs(int nmax)
{
    if (nmax <= 2048) return 256;
    if (nmax <= 8192) return 512;
    return 1024;
}

// ============================================ mixed precision + FP64 refinement
// FP64 CSR -> inner storage. F32: every value rounded to float. BF16: the
// off-diagonals rounded to bf16 (round to nearest even), the diagonal kept
// in FP32 and shifted by the row's rounding errors so that each row sum of
// the stored matrix equals the row sum of A (for an M-matrix / weighted
// Laplacian this keeps the stored matrix spectrally within 1 +- 2^-9 of A).
__device__ __forceinline__ unsigned short f64_to_bf16(double v)
{
    float f = (float)v;
    unsigned u = __float_as_uint(f);
    unsigned lsb = (u >> 16) & 1u;
    u += 0x7fffu + lsb;  // round to nearest, ties to even (finite inputs)
    return (unsigned short)(u >> 16);
}

template <bool BF16>
__global__ void ir_convert_kernel(int n, const int* __restrict__ JA, const int* __restrict__ IA,
                                  const double* __restrict__ v64, void* __restrict__ vlow, float* __restrict__ diag,
                                  float* __restrict__ invd)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double d = 0.0;
    for (int k = JA[i]; k < JA[i + 1]; ++k) {
        double a = v64[k];
        if (IA[k] == i) {
            d += a;
            if (BF16) ((unsigned short*)vlow)[k] = 0;
            else ((float*)vlow)[k] = (float)a;
        } else if (BF16) {
            unsigned short h = f64_to_bf16(a);
            ((unsigned short*)vlow)[k] = h;
            d += a - (double)bf16_to_float(h);
        } else {
            ((float*)vlow)[k] = (float)a;
        }
    }
    float df = (float)d;
    if (BF16) diag[i] = df;
    else {
        // diagonal as stored in the FP32 matrix
        df = 0.f;
        for (int k = JA[i]; k < JA[i + 1]; ++k)
            if (IA[k] == i) { df = ((float*)vlow)[k]; break; }
    }
    invd[i] = (df == 0.f) ? 1.f : 1.f / df;
}

// r = b - A x in FP64 (FP64 values, FP64 vectors)
__global__ __launch_bounds__(CG_BS) void ir_residual_kernel(int n, const int* __restrict__ JA, const int* __restrict__ IA,
                                                            const double* __restrict__ v64, const double* __restrict__ b,
                                                            const double* __restrict__ x, double* __restrict__ r)
{
    __shared__ SpmvSmem<double, CG_RPB, CG_CAP> sm;
    const int blk = xcd_remap(blockIdx.x, gridDim.x);
    const int r0 = blk * CG_RPB;
    if (r0 >= n) return;
    const int nr = min(CG_RPB, n - r0);
    double ax[CG_RPB / CG_BS];
    block_csr_spmv<double, double, CG_BS, CG_RPB, CG_CAP>(JA, IA, v64, x, r0, nr, sm, ax);
#pragma unroll
    for (int j = 0; j < CG_RPB / CG_BS; ++j) {
        int i = threadIdx.x + j * CG_BS;
        if (i < nr) r[r0 + i] = b[r0 + i] - ax[j];
    }
}

__global__ void ir_scale_kernel(int n, const double* __restrict__ r, const double* __restrict__ rn2, float* __restrict__ r32)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 1.0 / sqrt(*rn2);
    r32[i] = (float)(r[i] * s);
}

__global__ void ir_update_kernel(int n, double* __restrict__ x, const float* __restrict__ e, const double* __restrict__ rn2)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    x[i] += sqrt(*rn2) * (double)e[i];
}

static int pcg_ir(cm_ctx_t ctx, int n, int nnz, const int* JA, const int* IA, const double* v64, int vdt,
                  const double* b, double* x, double tol, double inner_tol, int max_outer, int max_inner,
                  cm_ir_stats* stats)
{
    Temps tmp(ctx);
    void *vlow, *diag, *invd, *r, *r32, *e, *dd;
    CM_TRY(tmp.get((vdt == CM_BF16 ? 2 : 4) * (size_t)std::max(nnz, 1), &vlow));
    CM_TRY(tmp.get(sizeof(float) * n, &diag));
    CM_TRY(tmp.get(sizeof(float) * n, &invd));
    CM_TRY(tmp.get(sizeof(double) * n, &r));
    CM_TRY(tmp.get(sizeof(float) * n, &r32));
    CM_TRY(tmp.get(sizeof(float) * n, &e));
    CM_TRY(tmp.get(sizeof(double) * 2, &dd));
    double* bn2 = (double*)dd;
    double* rn2 = bn2 + 1;
    auto t0 = std::chrono::steady_clock::now();
    const unsigned g = div_up(n, 256);
    if (vdt == CM_BF16)
        ir_convert_kernel<true><<<g, 256, 0, ctx->stream>>>(n, JA, IA, v64, vlow, (float*)diag, (float*)invd);
    else
        ir_convert_kernel<false><<<g, 256, 0, ctx->stream>>>(n, JA, IA, v64, vlow, (float*)diag, (float*)invd);
    CM_LAUNCH_CHECK();
    CM_TRY(fill_zero(ctx, x, sizeof(double) * n));
    CM_HIP(hipMemcpyAsync(r, b, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));
    CM_TRY(cm_dot(ctx, CM_F64, n, 1, b, 1, b, 1, bn2));
    double hbn2 = 0, hrn2 = 0;
    CM_HIP(hipMemcpyAsync(&hbn2, bn2, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    CM_TRY(::cm::host_sync(ctx));
    int outer = 0, inner_total = 0;
    double rel = (hbn2 > 0) ? 1.0 : 0.0;
    double prev = 2.0;
    while (hbn2 > 0) {
        CM_TRY(cm_dot(ctx, CM_F64, n, 1, r, 1, r, 1, rn2));
        CM_HIP(hipMemcpyAsync(&hrn2, rn2, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        CM_TRY(::cm::host_sync(ctx));
        rel = std::sqrt(hrn2 / hbn2);
        if (rel <= tol || outer >= max_outer || hrn2 == 0.0) break;
        // stagnation: an outer step that no longer halves the FP64 residual
        // has reached the rounding floor of the FP64 residual itself
        if (rel > 0.5 * prev) break;
        prev = rel;
        ir_scale_kernel<<<g, 256, 0, ctx->stream>>>(n, (const double*)r, rn2, (float*)r32);
        CM_LAUNCH_CHECK();
        int it = 0;
        double err = 0;
        if (vdt == CM_BF16)
            CM_TRY((pcg_multikernel<unsigned short, float>(ctx, n, nnz, 1, JA, IA, (const unsigned short*)vlow,
                                                           (const float*)invd, (const float*)r32, (float*)e, 0,
                                                           inner_tol, max_inner, &it, &err, nullptr,
                                                           (const float*)diag)));
        else
            CM_TRY((pcg_multikernel<float, float>(ctx, n, nnz, 1, JA, IA, (const float*)vlow, (const float*)invd,
                                                  (const float*)r32, (float*)e, 0, inner_tol, max_inner, &it, &err,
                                                  nullptr)));
        inner_total += it;
        ir_update_kernel<<<g, 256, 0, ctx->stream>>>(n, x, (const float*)e, rn2);
        CM_LAUNCH_CHECK();
        ir_residual_kernel<<<div_up(n, CG_RPB), CG_BS, 0, ctx->stream>>>(n, JA, IA, v64, b, x, (double*)r);
        CM_LAUNCH_CHECK();
        ++outer;
    }
    auto t1 = std::chrono::steady_clock::now();
    if (stats) {
        stats->outer_iterations = outer;
        stats->inner_iterations = inner_total;
        stats->final_rel_residual = rel;
        stats->seconds = std::chrono::duration<double>(t1 - t0).count();
    }
    return CM_OK;
}

}  // namespace cm

extern "C" {

int cm_pcg_batched_solve(cm_ctx_t ctx, int nsys, const int* row_off, int total_rows, int total_nnz, const int* JA,
                         const int* IA, const float* vals, const float* b, float* x, float tol, int max_iter,
                         int* iterations, float* errors, int block_size, int use_graph)
{
    CM_API_BEGIN
    CM_TRY(cm::resolve(ctx, &ctx));
    CM_REQUIRE(nsys >= 0 && total_rows >= 0 && total_nnz >= 0, CM_E_DIM, "cm_pcg_batched_solve: bad sizes");
    if (nsys == 0) return CM_OK;
    CM_REQUIRE(row_off && JA && IA && vals && b && x, CM_E_ARG, "cm_pcg_batched_solve: NULL pointer");
    cm::BatchedPlan plan;
    CM_TRY(cm::make_plan(ctx, nsys, row_off, total_rows, plan));
    CM_REQUIRE(plan.nmax <= 16384, CM_E_UNSUPPORTED, "batched CG: system of %d rows (max 16384)", plan.nmax);
    int gpc = 0;
    if (block_size == 0) {
        std::lock_guard<std::mutex> lk(cm::g_graph_mutex);
        auto it = cm::g_tuned_bs.find({nsys, total_rows, total_nnz, plan.nmax});
        if (it != cm::g_tuned_bs.end()) {
            block_size = it->second.first;
            gpc = it->second.second;
        } else {
            block_size = cm::default_bs(plan.nmax);
        }
    }
    unsigned bad = 0;
    if (!use_graph) {
        cm::Temps tmp(ctx);
        void *ord, *q;
        CM_TRY(tmp.get(sizeof(int) * nsys, &ord));
        CM_TRY(tmp.get(256, &q));
        CM_HIP(hipMemcpyAsync(ord, plan.order.data(), sizeof(int) * nsys, hipMemcpyHostToDevice, ctx->stream));
        CM_TRY(cm::batched_launch(ctx, nsys, row_off, total_nnz, JA, IA, vals, b, x, tol, max_iter, iterations,
                                  errors, block_size, plan, ord, q, gpc));
        CM_HIP(hipMemcpyAsync(&bad, (unsigned*)q + 1, sizeof(unsigned), hipMemcpyDeviceToHost, ctx->stream));
        CM_TRY(::cm::host_sync(ctx));
    } else {
        // graph path: the (queue reset + persistent solve) pair is captured
        // once per argument set and replayed
        cm::GraphKey key;
        memset(&key, 0, sizeof(key));
        key.ctx = ctx;
        const void* ps[9] = {row_off, JA, IA, vals, b, x, iterations, errors, nullptr};
        memcpy(key.ptrs, ps, sizeof(ps));
        key.nsys = nsys;
        key.bs = block_size;
        key.gpc = gpc;
        key.max_iter = max_iter;
        key.nmax = plan.nmax;
        key.total_rows = total_rows;
        key.total_nnz = total_nnz;
        key.tol = tol;
        key.offsets_hash = plan.offsets_hash;
        std::lock_guard<std::mutex> lk(cm::g_graph_mutex);
        auto it = cm::g_graphs.find(key);
        if (it == cm::g_graphs.end()) {
            if (cm::g_graphs.size() >= cm::GRAPH_CACHE_MAX) {
                auto lru = cm::g_graphs.begin();
                for (auto j = cm::g_graphs.begin(); j != cm::g_graphs.end(); ++j)
                    if (j->second.last_use < lru->second.last_use) lru = j;
                // the evicted graph may belong to another context's stream
                CM_HIP(hipStreamSynchronize(lru->first.ctx->stream));
                cm::graph_release(lru->second);
                cm::g_graphs.erase(lru);
            }
            cm::GraphEntry ge;
            CM_TRY(cm::dalloc(ctx, sizeof(int) * nsys, &ge.order));
            int s0 = cm::dalloc(ctx, 256, &ge.queue);
            if (s0 != CM_OK) {
                cm::graph_release(ge);
                return s0;
            }
            hipGraph_t g = nullptr;
            int s = CM_OK;
            if (hipMemcpyAsync(ge.order, plan.order.data(), sizeof(int) * nsys, hipMemcpyHostToDevice,
                               ctx->stream) != hipSuccess)
                s = cm::fail(CM_E_HIP, "batched CG: upload of the solve order failed");
            // the kernel attribute must be set outside capture
            if (s == CM_OK)
                s = cm::batched_launch(ctx, 0, row_off, total_nnz, JA, IA, vals, b, x, tol, max_iter, iterations,
                                       errors, block_size, plan, ge.order, ge.queue, gpc);
            if (s == CM_OK && ::cm::host_sync(ctx) != CM_OK)
                s = cm::fail(CM_E_HIP, "batched CG: stream synchronization failed");
            if (s == CM_OK) {
                if (hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
                    s = cm::fail(CM_E_HIP, "batched CG: stream capture failed to start");
                } else {
                    s = cm::batched_launch(ctx, nsys, row_off, total_nnz, JA, IA, vals, b, x, tol, max_iter,
                                           iterations, errors, block_size, plan, ge.order, ge.queue, gpc);
                    hipError_t e = hipStreamEndCapture(ctx->stream, &g);
                    if (s == CM_OK && e != hipSuccess)
                        s = cm::fail(CM_E_HIP, "batched CG: capture failed: %s", hipGetErrorString(e));
                }
            }
            if (s == CM_OK && hipGraphInstantiate(&ge.exec, g, nullptr, nullptr, 0) != hipSuccess)
                s = cm::fail(CM_E_HIP, "batched CG: graph instantiation failed");
            if (g) (void)hipGraphDestroy(g);
            if (s != CM_OK) {
                cm::graph_release(ge);
                return s;
            }
            it = cm::g_graphs.emplace(key, ge).first;
        }
        it->second.last_use = ++cm::g_graph_clock;
        CM_HIP(hipGraphLaunch(it->second.exec, ctx->stream));
        cm::count(CM_CNT_KERNEL_LAUNCH, 2);
        CM_HIP(hipMemcpyAsync(&bad, (unsigned*)it->second.queue + 1, sizeof(unsigned), hipMemcpyDeviceToHost,
                              ctx->stream));
        CM_TRY(::cm::host_sync(ctx));
    }
    CM_REQUIRE(!bad, CM_E_ARG,
               "batched CG: a system's row pointers or column indices leave its own rows (those systems report "
               "iterations -1 and error NaN)");
    return CM_OK;
    CM_API_END
}

int cm_pcg_batched_autotune(cm_ctx_t ctx, int nsys, const int* row_off, int total_rows, int total_nnz, const int* JA,
                            const int* IA, const float* vals, const float* b, float tol, int max_iter,
                            int* best_block_size, int* best_groups_per_cu, float* best_ms)
{
    CM_API_BEGIN
    CM_TRY(cm::resolve(ctx, &ctx));
    CM_REQUIRE(nsys > 0, CM_E_DIM, "autotune: no systems");
    CM_REQUIRE(row_off && JA && IA && vals && b, CM_E_ARG, "autotune: NULL pointer");
    cm::BatchedPlan plan;
    CM_TRY(cm::make_plan(ctx, nsys, row_off, total_rows, plan));
    cm::Temps tmp(ctx);
    void *xs, *it, *er, *ord, *q;
    CM_TRY(tmp.get(sizeof(float) * (size_t)total_rows, &xs));
    CM_TRY(tmp.get(sizeof(int) * nsys, &it));
    CM_TRY(tmp.get(sizeof(float) * nsys, &er));
    CM_TRY(tmp.get(sizeof(int) * nsys, &ord));
    CM_TRY(tmp.get(256, &q));
    CM_HIP(hipMemcpyAsync(ord, plan.order.data(), sizeof(int) * nsys, hipMemcpyHostToDevice, ctx->stream));
    hipEvent_t e0, e1;
    CM_HIP(hipEventCreate(&e0));
    CM_HIP(hipEventCreate(&e1));
    int best = 0, best_gpc = 0;
    float bestms = 1e30f;
    // launch shapes: threads per system (the register tile follows from the
    // largest system) x workgroups per CU of the persistent grid
    const int cands[3] = {256, 512, 1024};
    const int gpcs[3] = {1, 2, 4};
    int status = CM_OK;
    for (int c = 0; c < 9 && status == CM_OK; ++c) {
        const int bs = cands[c / 3], gpc = gpcs[c % 3];
        if ((plan.nmax + bs - 1) / bs > (bs <= 512 ? 32 : 16)) continue;  // register tile too large
        float ms_min = 1e30f;
        for (int rep = 0; rep < 2 && status == CM_OK; ++rep) {
            if (hipEventRecord(e0, ctx->stream) != hipSuccess) { status = cm::fail(CM_E_HIP, "autotune: event"); break; }
            status = cm::batched_launch(ctx, nsys, row_off, total_nnz, JA, IA, vals, b, (float*)xs, tol, max_iter,
                                        (int*)it, (float*)er, bs, plan, ord, q, gpc);
            if (status != CM_OK) break;
            float ms = 0;
            if (hipEventRecord(e1, ctx->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                status = cm::fail(CM_E_HIP, "autotune: event timing failed");
                break;
            }
            ms_min = std::min(ms_min, ms);
        }
        if (status != CM_OK) break;
        if (ms_min < bestms) { bestms = ms_min; best = bs; best_gpc = gpc; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    CM_TRY(status);
    unsigned bad = 0;
    CM_HIP(hipMemcpyAsync(&bad, (unsigned*)q + 1, sizeof(unsigned), hipMemcpyDeviceToHost, ctx->stream));
    CM_TRY(::cm::host_sync(ctx));
    CM_REQUIRE(!bad, CM_E_ARG, "autotune: a system's row pointers or column indices leave its own rows");
    CM_REQUIRE(best > 0, CM_E_UNSUPPORTED, "autotune: no launch shape fits systems of %d rows", plan.nmax);
    {
        std::lock_guard<std::mutex> lk(cm::g_graph_mutex);
        if (cm::g_tuned_bs.size() >= cm::TUNE_CACHE_MAX) cm::g_tuned_bs.clear();
        cm::g_tuned_bs[{nsys, total_rows, total_nnz, plan.nmax}] = {best, best_gpc};
    }
    if (best_block_size) *best_block_size = best;
    if (best_groups_per_cu) *best_groups_per_cu = best_gpc;
    if (best_ms) *best_ms = bestms;
    return CM_OK;
    CM_API_END
}

int cm_pcg_solve(cm_ctx_t ctx, int dtype, int n, int nnz, int batches, const int* JA, const int* IA,
                 const void* vals, const void* invdiag, const void* b, void* x, int use_guess, double tol,
                 int max_iter, int path, int* iterations_out, double* error_out, cm_cg_stats* stats)
{
    CM_API_BEGIN
    CM_TRY(cm::resolve(ctx, &ctx));
    CM_REQUIRE(n >= 0 && nnz >= 0 && batches >= 1, CM_E_DIM, "cm_pcg_solve: bad sizes");
    CM_REQUIRE(tol >= 0, CM_E_ARG, "cm_pcg_solve: negative tolerance");
    if (max_iter < 0) max_iter = 2 * n;  // IterativeSolverBase::maxIterations (cuMat/src/IterativeSolverBase.h:85)
    if (n == 0) {
        if (iterations_out) *iterations_out = 0;
        if (error_out) *error_out = 0;
        return CM_OK;
    }
    CM_REQUIRE(JA && IA && vals && b && x, CM_E_ARG, "cm_pcg_solve: NULL pointer");
    const bool resident_ok = batches == 1 && n <= 16384;
    // measured on MI355X (FP32 2D Poisson, iterations/s): the one-workgroup
    // solve wins up to 64^2 = 4096 rows (138 k vs 97 k), the launches from
    // 96^2 = 9216 rows (96 k vs 69 k); the crossover is ~6000 rows. Up to
    // 2^17 unknowns (n x batches) an iteration is launch-bound and the
    // two-launch form wins despite its tripled gathers (us per iteration,
    // three / two launches: 128^2 12.6 / 10.2, 256^2 13.0 / 10.2, 512^2
    // 13.8 / 13.7, 1024^2 26.9 / 33.1)
    if (path == CM_CG_AUTO)
        path = (resident_ok && n <= 6144) ? CM_CG_RESIDENT
                                          : ((long long)n * batches <= (1 << 17) ? CM_CG_FUSED : CM_CG_MULTIKERNEL);
    if (path == CM_CG_RESIDENT) {
        CM_REQUIRE(resident_ok, CM_E_UNSUPPORTED, "resident CG needs one rhs and n <= 16384");
        if (dtype == CM_F32)
            return cm::pcg_resident<float>(ctx, n, nnz, JA, IA, (const float*)vals, (const float*)invdiag, (const float*)b,
                                           (float*)x, use_guess, tol, max_iter, iterations_out, error_out, stats);
        if (dtype == CM_F64)
            return cm::pcg_resident<double>(ctx, n, nnz, JA, IA, (const double*)vals, (const double*)invdiag,
                                            (const double*)b, (double*)x, use_guess, tol, max_iter, iterations_out,
                                            error_out, stats);
        return cm::fail(CM_E_UNSUPPORTED, "cm_pcg_solve: dtype %d", dtype);
    }
    CM_REQUIRE(path == CM_CG_MULTIKERNEL || path == CM_CG_FUSED, CM_E_ARG, "cm_pcg_solve: unknown path %d", path);
    cm::CgBatchOpts opt;
    opt.fused = path == CM_CG_FUSED;
    if (dtype == CM_F32)
        return cm::pcg_multikernel<float, float>(ctx, n, nnz, batches, JA, IA, (const float*)vals,
                                                 (const float*)invdiag, (const float*)b, (float*)x, use_guess, tol,
                                                 max_iter, iterations_out, error_out, stats, nullptr, opt);
    if (dtype == CM_F64)
        return cm::pcg_multikernel<double, double>(ctx, n, nnz, batches, JA, IA, (const double*)vals,
                                                   (const double*)invdiag, (const double*)b, (double*)x, use_guess,
                                                   tol, max_iter, iterations_out, error_out, stats, nullptr, opt);
    return cm::fail(CM_E_UNSUPPORTED, "cm_pcg_solve: dtype %d", dtype);
    CM_API_END
}

int cm_pcg_dense_solve(cm_ctx_t ctx, int dtype, int n, int batches, const void* A, int a_batched, int a_row_major,
                       const void* invdiag, const void* b, void* x, int use_guess, double tol, int max_iter,
                       int* iterations_out, double* error_out, cm_cg_stats* stats)
{
    CM_API_BEGIN
    CM_TRY(cm::resolve(ctx, &ctx));
    CM_REQUIRE(n >= 0 && batches >= 1 && batches <= cm::CG_MAXB, CM_E_DIM,
               "cm_pcg_dense_solve: bad sizes (batches 1..%d)", cm::CG_MAXB);
    CM_REQUIRE(tol >= 0, CM_E_ARG, "cm_pcg_dense_solve: negative tolerance");
    if (max_iter < 0) max_iter = 2 * n;  // IterativeSolverBase::maxIterations (cuMat/src/IterativeSolverBase.h:85)
    if (n == 0) {
        if (iterations_out) *iterations_out = 0;
        if (error_out) *error_out = 0;
        return CM_OK;
    }
    CM_REQUIRE(A && b && x, CM_E_ARG, "cm_pcg_dense_solve: NULL pointer");
    CM_REQUIRE(dtype == CM_F32 || dtype == CM_F64, CM_E_UNSUPPORTED, "cm_pcg_dense_solve: dtype %d", dtype);
    cm::Temps tmp(ctx);
    const int nb = a_batched ? batches : 1;
    if (!invdiag) {
        void* id;
        CM_TRY(tmp.get((dtype == CM_F32 ? 4 : 8) * (size_t)n * nb, &id));
        CM_TRY(cm_dense_inverse_diagonal(ctx, dtype, n, batches, A, a_batched, id));
        invdiag = id;
    }
    cm::CgBatchOpts opt;
    opt.vals_batched = a_batched;  // the inverse diagonal then has one row of n per batch
    const long long bstride = a_batched ? (long long)n * n : 0;
    if (dtype == CM_F32) {
        cm::DenseOp<float> op;
        op.A = (const float*)A;
        op.bstride = bstride;
        op.row_major = a_row_major != 0;
        opt.dense = &op;
        return cm::pcg_multikernel<float, float>(ctx, n, 0, batches, nullptr, nullptr, nullptr,
                                                 (const float*)invdiag, (const float*)b, (float*)x, use_guess, tol,
                                                 max_iter, iterations_out, error_out, stats, nullptr, opt);
    }
    cm::DenseOp<double> op;
    op.A = (const double*)A;
    op.bstride = bstride;
    op.row_major = a_row_major != 0;
    opt.dense = &op;
    return cm::pcg_multikernel<double, double>(ctx, n, 0, batches, nullptr, nullptr, nullptr, (const double*)invdiag,
                                               (const double*)b, (double*)x, use_guess, tol, max_iter,
                                               iterations_out, error_out, stats, nullptr, opt);
    CM_API_END
}

int cm_pcg_bcsr_solve(cm_ctx_t ctx, int block_dim, int rows, int nnz, int batches, const int* JA, const int* IA,
                      const float* vals, int vals_batched, const float* invdiag, const float* b, float* x,
                      int use_guess, double tol, int max_iter, int* iterations_out, double* error_out,
                      cm_cg_stats* stats)
{
    CM_API_BEGIN
    CM_TRY(cm::resolve(ctx, &ctx));
    CM_REQUIRE(block_dim >= 2 && block_dim <= 4, CM_E_UNSUPPORTED, "cm_pcg_bcsr_solve: block dimension %d (2..4)",
               block_dim);
    CM_REQUIRE(rows >= 0 && nnz >= 0 && batches >= 1 && batches <= cm::CG_MAXB, CM_E_DIM,
               "cm_pcg_bcsr_solve: bad sizes (batches 1..%d)", cm::CG_MAXB);
    CM_REQUIRE(rows <= INT_MAX / 4, CM_E_DIM, "cm_pcg_bcsr_solve: %d block rows (max %d: 4 floats per row)", rows,
               INT_MAX / 4);
    CM_REQUIRE(tol >= 0, CM_E_ARG, "cm_pcg_bcsr_solve: negative tolerance");
    if (max_iter < 0) max_iter = 2 * rows;  // IterativeSolverBase::maxIterations: 2 * cols (in blocks)
    if (rows == 0) {
        if (iterations_out) *iterations_out = 0;
        if (error_out) *error_out = 0;
        return CM_OK;
    }
    CM_REQUIRE(JA && b && x && (nnz == 0 || (IA && vals)), CM_E_ARG, "cm_pcg_bcsr_solve: NULL pointer");
    CM_REQUIRE((((uintptr_t)vals | (uintptr_t)invdiag | (uintptr_t)b | (uintptr_t)x) & 15) == 0, CM_E_ARG,
               "cm_pcg_bcsr_solve: blocks and vectors must be 16-byte aligned (float4 entries)");
    cm::Temps tmp(ctx);
    const long long N = (long long)rows * batches;  // float4 entries of b and x
    const int nbi = vals_batched ? batches : 1;
    // the right-hand side with components >= D cleared (the reference's
    // squaredNorm and Jacobi entries never see them), then the Jacobi entries
    float* bz;
    CM_TRY(tmp.get(16 * (size_t)N + (invdiag ? 0 : 16 * (size_t)rows * nbi), (void**)&bz));
    const dim3 mg(cm::div_up(N, 256));
    if (block_dim == 2)
        cm::bcsr_mask_kernel<2><<<mg, 256, 0, ctx->stream>>>(N, (const cm::bc_f4*)b, (cm::bc_f4*)bz);
    else if (block_dim == 3)
        cm::bcsr_mask_kernel<3><<<mg, 256, 0, ctx->stream>>>(N, (const cm::bc_f4*)b, (cm::bc_f4*)bz);
    else
        cm::bcsr_mask_kernel<4><<<mg, 256, 0, ctx->stream>>>(N, (const cm::bc_f4*)b, (cm::bc_f4*)bz);
    CM_LAUNCH_CHECK();
    const long long vstride = vals_batched ? (long long)nnz * block_dim : 0;
    if (!invdiag) {
        float* id = bz + 4 * N;
        cm::launch_bcsr_invdiag(ctx->stream, block_dim, rows, nbi, JA, IA, (const cm::bc_f4*)vals, vstride,
                                (cm::bc_f4*)id);
        CM_LAUNCH_CHECK();
        invdiag = id;
    }
    cm::BlockOp op;
    op.JA = JA;
    op.IA = IA;
    op.vals = (const cm::bc_f4*)vals;
    op.vstride = vstride;
    op.rows = rows;
    op.D = block_dim;
    op.nnz = nnz;
    cm::CgBatchOpts opt;
    opt.vals_batched = vals_batched;  // the Jacobi entries then have one row of 4*rows per batch
    opt.block = &op;
    return cm::pcg_multikernel<float, float>(ctx, 4 * rows, 0, batches, nullptr, nullptr, nullptr, invdiag, bz, x,
                                             use_guess, tol, max_iter, iterations_out, error_out, stats, nullptr,
                                             opt);
    CM_API_END
}

int cm_pcg_solve_ex(cm_ctx_t ctx, int dtype, int n, int nnz, int batches, const int* JA, const int* IA,
                    const void* vals, int vals_batched, const void* invdiag, const void* b, void* x, int use_guess,
                    double tol, int max_iter, int per_batch, int* iterations_each, double* error_each,
                    cm_cg_stats* stats)
{
    CM_API_BEGIN
    CM_TRY(cm::resolve(ctx, &ctx));
    CM_REQUIRE(n >= 0 && nnz >= 0 && batches >= 1 && batches <= cm::CG_MAXB, CM_E_DIM,
               "cm_pcg_solve_ex: bad sizes (batches 1..%d)", cm::CG_MAXB);
    CM_REQUIRE(tol >= 0, CM_E_ARG, "cm_pcg_solve_ex: negative tolerance");
    if (max_iter < 0) max_iter = 2 * n;
    if (n == 0) {
        for (int k = 0; k < batches; ++k) {
            if (iterations_each) iterations_each[k] = 0;
            if (error_each) error_each[k] = 0;
        }
        return CM_OK;
    }
    CM_REQUIRE(JA && IA && vals && b && x, CM_E_ARG, "cm_pcg_solve_ex: NULL pointer");
    cm::CgBatchOpts opt;
    opt.vals_batched = vals_batched;
    opt.per_batch = per_batch;
    std::vector<int> its(batches, 0);
    std::vector<double> errs(batches, 0.0);
    opt.biters = its.data();
    opt.berr = errs.data();
    int it = 0;
    double err = 0;
    int s;
    if (dtype == CM_F32)
        s = cm::pcg_multikernel<float, float>(ctx, n, nnz, batches, JA, IA, (const float*)vals, (const float*)invdiag,
                                              (const float*)b, (float*)x, use_guess, tol, max_iter, &it, &err, stats,
                                              nullptr, opt);
    else if (dtype == CM_F64)
        s = cm::pcg_multikernel<double, double>(ctx, n, nnz, batches, JA, IA, (const double*)vals,
                                                (const double*)invdiag, (const double*)b, (double*)x, use_guess, tol,
                                                max_iter, &it, &err, stats, nullptr, opt);
    else
        return cm::fail(CM_E_UNSUPPORTED, "cm_pcg_solve_ex: dtype %d", dtype);
    CM_TRY(s);
    for (int k = 0; k < batches; ++k) {
        if (iterations_each) iterations_each[k] = per_batch ? its[k] : it;
        if (error_each) error_each[k] = per_batch ? errs[k] : err;
    }
    return CM_OK;
    CM_API_END
}

int cm_pcg_ir_solve(cm_ctx_t ctx, int n, int nnz, const int* JA, const int* IA, const double* vals64, int val_dtype,
                    const double* b, double* x, double tol, double inner_tol, int max_outer, int max_inner,
                    cm_ir_stats* stats)
{
    CM_API_BEGIN
    CM_TRY(cm::resolve(ctx, &ctx));
    CM_REQUIRE(n >= 1 && nnz >= 0, CM_E_DIM, "cm_pcg_ir_solve: bad sizes");
    CM_REQUIRE(val_dtype == CM_F32 || val_dtype == CM_BF16, CM_E_ARG, "cm_pcg_ir_solve: inner storage %d",
               val_dtype);
    CM_REQUIRE(JA && IA && vals64 && b && x, CM_E_ARG, "cm_pcg_ir_solve: NULL pointer");
    CM_REQUIRE(tol > 0 && inner_tol > 0 && max_outer >= 0, CM_E_ARG, "cm_pcg_ir_solve: bad tolerances");
    if (max_inner < 0) max_inner = 2 * n;
    return cm::pcg_ir(ctx, n, nnz, JA, IA, vals64, val_dtype, b, x, tol, inner_tol, max_outer, max_inner, stats);
    CM_API_END
}

}  // extern "C"
