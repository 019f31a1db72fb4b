// cg_multi.hpp -- the multi-kernel Jacobi-PCG (any n, up to CG_MAXB
// right-hand sides), fence-free.
//
// Per iteration i, three launches (the default form):
//   K1  tmp = A p, per-block partial p.tmp
//   K2  every block reduces K1's partials of its batch in the same fixed
//       order -> the same alpha = absNew[i&1] / p.tmp; r -= alpha tmp
//       (and x += alpha p for small systems); per-block partials of |r|^2
//       and r.(M^-1 r)
//   K3  every block reduces K2's partials of all batches -> convergence test
//       and beta = rz / absNew[i&1]; p = M^-1 r + beta p; for large systems
//       also x += alpha p (alpha re-derived from K1's partials), which saves
//       K2 the x and p passes; block 0 records absNew[(i+1)&1] = rz, the
//       status, the iteration count and the error. In that form a batch's
//       last x update is recorded and applied by the finalize kernel, so no
//       kernel depends on a decision another block of the same launch writes.
// The FUSED form has two launches per iteration (p alternates between two
// buffers); it saves the p re-read and a launch but triples the gathers of
// the SpMV, which costs more on MI355X than it saves (94 us for KA against
// 48 + 18 us for K1 + K3 at 2048^2), so it is not the default:
//   KA  (i >= 1) every block reduces KB's partials of iteration i-1 for all
//       batches in the same fixed order -> convergence test of iteration
//       i-1 and beta = rz / absNew[(i-1)&1]; block 0 records absNew[i&1] = rz,
//       the status, the iteration count and the error. Then
//       p_i = M^-1 r + beta p_{i-1} and tmp = A p_i in one pass (gathered
//       p_i values are recomputed from r, M^-1, p_{i-1}), partial p_i.tmp.
//       (i == 0: p_0 from the setup, no test.)
//   KB  every block reduces KA's partials of its batch -> the same
//       alpha = absNew[i&1] / p.tmp; x += alpha p_i, r -= alpha tmp;
//       per-block partials of |r|^2 and r.(M^-1 r)
// After the last iteration one more KA completes its convergence test.
// Cross-block data only flows through kernel boundaries (no device-scope
// fences, which on a multi-XCD part write back the L2 each time), and every
// reduction is done in one fixed order, so runs are bit-reproducible.
//
// Batches: the matrix is either shared (the reference's case: a batch-1
// matrix with batched right-hand sides, cuMat/src/ConjugateGradient.h:223)
// or has its own values per batch (vstride = nnz). Convergence is either the
// reference's -- every batch iterates until all have converged
// (cuMat/src/ConjugateGradient.h:43-45,190-193) -- or per batch: a converged
// batch stops updating and keeps its own iteration count and error.
// The grid shrinks as the batch count grows so that each reduction folds at
// most CG_GMAX partials.
#pragma once
#include <cstdlib>
#include <map>
#include <type_traits>

#include "dense.hpp"
#include "spmv_core.hpp"
#include "spmv_wave.hpp"

namespace cm {

constexpr int CG_MAXB = 64;
constexpr int CG_BS = 256;
constexpr int CG_RPB = 512;    // rows per block of the setup SpMV
constexpr int CG_CAP = 4096;   // staged non-zeros per setup row block
constexpr int CG_EPT = 8;      // staged non-zeros per lane in K1
constexpr int CG_VPT = 4;      // vector elements per thread per sweep
constexpr int CG_GMAX = 2048;  // max blocks (= max partials) per kernel
// max blocks of K2 / K3: each of their blocks first folds all of K1's
// partials, so fewer, longer blocks pay off (measured at 2048^2, us per
// iteration: K2/K3 grid 2048 -> 72.8, 1024 -> 69.1, 768 -> 70.2, 512 ->
// 70.1, 256 -> 73.6; K1 itself keeps 2048 blocks: 1536 -> 71.6, 1024 -> 78.8)
constexpr int CG_G2MAX = 1024;
// from this many unknowns (n x batches) on, x is updated in K3 instead of K2
// (measured on MI355X: +6 % at 1024^2, -14 % at 512^2 where latency rules)
constexpr long long CG_XK3_MIN = 1 << 20;

enum CgStatus : int { CG_RUNNING = 0, CG_CONVERGED = 1, CG_EARLY_ZERO = 2, CG_EARLY_GUESS = 3, CG_MAXIT = 4 };

template <typename T>
struct CgState {
    T rhs2[CG_MAXB];
    T thr[CG_MAXB];
    T absNew[2][CG_MAXB];
    int bdone[CG_MAXB];   // per-batch mode: CgStatus of each batch
    int biters[CG_MAXB];
    double berr[CG_MAXB];
    T alpha_fin[CG_MAXB]; // alpha of a batch's last iteration ...
    int xpend[CG_MAXB];   // ... whose x update the finalize kernel applies
    int status;
    int iters;
    int it;        // index of the running iteration (three-launch form)
    int max_iter;
    int batches;
    int per_batch;
    double err;
};

// Sum of G partials (stride 'stride' between consecutive ones) in a fixed
// order; identical in every block. lds: BS/64 entries.
// All of a thread's loads (G <= CG_GMAX = 8 * CG_BS) are issued before the
// first add, so the partials written by the previous launch (often in another
// XCD's cache) cost one memory latency, not eight; the adds keep the order
// i = tid, tid + BS, ... so the result does not depend on the unrolling.
constexpr int CG_PPT = CG_GMAX / CG_BS;  // partials per thread
static_assert(CG_PPT * CG_BS == CG_GMAX, "partials must fit one unrolled pass");

template <typename T>
__device__ __forceinline__ T reduce_partials(const T* __restrict__ part, int G, int stride, T* lds)
{
    T v[CG_PPT];
#pragma unroll
    for (int j = 0; j < CG_PPT; ++j) {
        const int i = threadIdx.x + j * CG_BS;
        v[j] = (i < G) ? part[(long long)i * stride] : T(0);
    }
    T a = T(0);
#pragma unroll
    for (int j = 0; j < CG_PPT; ++j)
        if (threadIdx.x + j * CG_BS < G) a += v[j];
    return block_reduce<OpSum, CG_BS>(a, lds);
}

// Two interleaved partial sets (part[2i], part[2i+1]) reduced together
// (part must be 2*sizeof(T)-aligned); lds needs 2*CG_BS/64 entries. Same order and results as two calls of
// reduce_partials with stride 2.
template <typename T>
__device__ __forceinline__ void reduce_partials_pair(const T* __restrict__ part, int G, T& a, T& b, T* lds)
{
    typedef T T2 __attribute__((ext_vector_type(2)));
    T va[CG_PPT], vb[CG_PPT];
#pragma unroll
    for (int j = 0; j < CG_PPT; ++j) {
        const int i = threadIdx.x + j * CG_BS;
        T2 v = {T(0), T(0)};
        if (i < G) v = *reinterpret_cast<const T2*>(part + 2ll * i);  // one load per pair
        va[j] = v.x;
        vb[j] = v.y;
    }
    a = T(0);
    b = T(0);
#pragma unroll
    for (int j = 0; j < CG_PPT; ++j)
        if (threadIdx.x + j * CG_BS < G) {
            a += va[j];
            b += vb[j];
        }
    block_sum2<CG_BS>(a, b, lds);
}

// One right-hand side: the p.Ap partials (stride 1) and the (|r|^2, r.z)
// pairs loaded together and summed with shared barriers; lds needs
// 3*CG_BS/64 entries. Same order and results as reduce_partials followed by
// reduce_partials_pair, with one memory latency and one barrier round less.
template <typename T>
__device__ __forceinline__ void reduce_partials_three(const T* __restrict__ part1, int G1,
                                                      const T* __restrict__ part2, int G2, T& pap, T& a, T& b,
                                                      T* lds)
{
    typedef T T2 __attribute__((ext_vector_type(2)));
    T v1[CG_PPT], va[CG_PPT], vb[CG_PPT];
#pragma unroll
    for (int j = 0; j < CG_PPT; ++j) {
        const int i = threadIdx.x + j * CG_BS;
        v1[j] = (i < G1) ? part1[i] : T(0);
        T2 v = {T(0), T(0)};
        if (i < G2) v = *reinterpret_cast<const T2*>(part2 + 2ll * i);
        va[j] = v.x;
        vb[j] = v.y;
    }
    pap = T(0);
    a = T(0);
    b = T(0);
#pragma unroll
    for (int j = 0; j < CG_PPT; ++j) {
        const int i = threadIdx.x + j * CG_BS;
        if (i < G1) pap += v1[j];
        if (i < G2) {
            a += va[j];
            b += vb[j];
        }
    }
    constexpr int NW = CG_BS / 64;
    pap = wave_reduce<OpSum>(pap);
    a = wave_reduce<OpSum>(a);
    b = wave_reduce<OpSum>(b);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        lds[w] = pap;
        lds[NW + w] = a;
        lds[2 * NW + w] = b;
    }
    __syncthreads();
    T r0 = lds[0], r1 = lds[NW], r2 = lds[2 * NW];
#pragma unroll
    for (int i = 1; i < NW; ++i) {
        r0 += lds[i];
        r1 += lds[NW + i];
        r2 += lds[2 * NW + i];
    }
    __syncthreads();
    pap = r0;
    a = r1;
    b = r2;
}

template <typename T>
__device__ __forceinline__ bool batch_inactive(const CgState<T>* st, int bb)
{
    return st->status != CG_RUNNING || (st->per_batch && st->bdone[bb] != CG_RUNNING);
}

// ---------------------------------------------------------------- setup
template <typename VT, typename T>
__global__ __launch_bounds__(CG_BS) void cg_setup_kernel(int n, const int* __restrict__ JA, const int* __restrict__ IA,
                                                         const VT* __restrict__ vals, long long vstride,
                                                         const T* __restrict__ invd, int istride,
                                                         const T* __restrict__ b, T* __restrict__ x, T* __restrict__ r,
                                                         T* __restrict__ p, int use_guess, T* __restrict__ part,
                                                         const T* __restrict__ diag, const T* __restrict__ ax_pre)
{
    // ax_pre: A x of the guess computed beforehand (dense operators); JA, IA
    // and vals are then unused
    __shared__ SpmvSmem<T, CG_RPB, CG_CAP> sm;
    __shared__ T red[2 * CG_BS / 64];
    const int G = gridDim.x;
    const int NB = div_up_dev(n, CG_RPB);
    const int bb = blockIdx.y;
    const long long off = (long long)bb * n;
    const VT* vb = vals + bb * vstride;
    const T* ib = invd + (long long)bb * istride;
    T sbb = 0, srr = 0, srp = 0;
    for (RowBlockWalk w(blockIdx.x, G, NB); w.valid(); w.next()) {
        const int r0 = w.cur * CG_RPB;
        const int nr = min(CG_RPB, n - r0);
        T ax[CG_RPB / CG_BS];
        if (use_guess && ax_pre) {
#pragma unroll
            for (int j = 0; j < CG_RPB / CG_BS; ++j) {
                const int i = threadIdx.x + j * CG_BS;
                ax[j] = (i < nr) ? ax_pre[off + r0 + i] : T(0);
            }
        } else if (use_guess) {
            block_csr_spmv<VT, T, CG_BS, CG_RPB, CG_CAP>(JA, IA, vb, x + off, r0, nr, sm, ax, diag);
            __syncthreads();  // sm is reused by the next row block
        } else {
#pragma unroll
            for (int j = 0; j < CG_RPB / CG_BS; ++j) ax[j] = T(0);
        }
#pragma unroll
        for (int j = 0; j < CG_RPB / CG_BS; ++j) {
            int i = threadIdx.x + j * CG_BS;
            if (i < nr) {
                long long row = r0 + i;
                T bv = b[off + row];
                T rv = bv - ax[j];
                T pv = ib[row] * rv;
                if (!use_guess) x[off + row] = T(0);
                r[off + row] = rv;
                p[off + row] = pv;
                sbb += bv * bv;
                srr += rv * rv;
                srp += rv * pv;
            }
        }
    }
    block_sum2<CG_BS>(sbb, srr, red);
    srp = block_reduce<OpSum, CG_BS>(srp, red);
    if (threadIdx.x == 0) {
        long long base = 3ll * ((long long)bb * G + blockIdx.x);
        part[base + 0] = sbb;
        part[base + 1] = srr;
        part[base + 2] = srp;
    }
}

template <typename T>
__global__ __launch_bounds__(CG_BS) void cg_setup_finalize_kernel(int G, int B, T tol, const T* __restrict__ part,
                                                                  CgState<T>* st)
{
    __shared__ T red[CG_BS / 64];
    __shared__ T tot[3][CG_MAXB];
    for (int k = 0; k < B; ++k)
        for (int c = 0; c < 3; ++c) {
            T v = reduce_partials(part + 3ll * k * G + c, G, 3, red);
            if (threadIdx.x == 0) tot[c][k] = v;
        }
    __syncthreads();
    if (threadIdx.x == 0) {
        bool all_zero = true, all_conv = true, all_done = true;
        T m = 0;
        for (int k = 0; k < B; ++k) {
            st->rhs2[k] = tot[0][k];
            st->thr[k] = tol * tol * tot[0][k];
            st->absNew[0][k] = tot[2][k];
            const bool zero = tot[0][k] < tol * tol;
            const bool conv = tot[1][k] < st->thr[k];
            all_zero = all_zero && zero;
            all_conv = all_conv && conv;
            T q = tot[1][k] / tot[0][k];
            m = (k == 0 || q > m) ? q : m;
            // per-batch early outs (the plain method's, applied per system)
            st->bdone[k] = zero ? CG_EARLY_ZERO : (conv ? CG_EARLY_GUESS : CG_RUNNING);
            st->biters[k] = 0;
            st->berr[k] = zero ? 0.0 : sqrt((double)q);
            all_done = all_done && st->bdone[k] != CG_RUNNING;
        }
        st->iters = 0;
        st->batches = B;
        if (st->per_batch ? (all_done && all_zero) : all_zero) { st->status = CG_EARLY_ZERO; st->err = 0.0; }
        else if (st->per_batch ? all_done : all_conv) { st->status = CG_EARLY_GUESS; st->err = sqrt((double)m); }
        else if (st->max_iter <= 0) { st->status = CG_MAXIT; st->err = sqrt((double)m); }
        else { st->status = CG_RUNNING; st->err = sqrt((double)m); }
    }
}

// ---------------------------------------------------------------- K1 (three-launch form)
// tmp = A p, per-block partial p.tmp
template <typename VT, typename T>
__global__ __launch_bounds__(CG_BS) void cg_spmv_kernel(int n, const int* __restrict__ JA, const int* __restrict__ IA,
                                                        const VT* __restrict__ vals, long long vstride,
                                                        const T* __restrict__ p, T* __restrict__ tmp,
                                                        T* __restrict__ part, CgState<T>* st,
                                                        const T* __restrict__ diag)
{
    const int bb = blockIdx.y;
    // K1 owns the iteration counter that K2 and K3 read: no launch argument
    // changes from one iteration to the next, so iterations can be replayed
    // from a captured graph (K1 itself never reads it, K2 / K3 read it after
    // this launch has completed)
    if (blockIdx.x == 0 && bb == 0 && threadIdx.x == 0 && st->status == CG_RUNNING) st->it = st->it + 1;
    if (batch_inactive(st, bb)) return;
    constexpr int WPB = CG_BS / 64;
    __shared__ T lds[WPB][CG_EPT * 64];
    __shared__ T red[CG_BS / 64];
    T* slice = lds[threadIdx.x >> 6];
    const int lane = threadIdx.x & 63;
    const long long off = (long long)bb * n;
    const T* pb = p + off;
    const VT* vb = vals + bb * vstride;
    T s = 0;
    ChunkWalk w(div_up_dev(n, 64), WPB);
    if (w.valid()) {
        WaveChunk<VT, CG_EPT> cur, nxt;
        chunk_load<VT, CG_EPT, false>(cur, JA, IA, vb, w.cur * 64, n);
        for (;;) {
            const int nc = w.peek_next();
            const bool more = nc < w.hi;
            if (more) chunk_load<VT, CG_EPT, false>(nxt, JA, IA, vb, nc * 64, n);
            T ap = chunk_compute<VT, T, CG_EPT>(cur, JA, IA, vb, [pb](int c) { return pb[c]; }, slice);
            if (lane < cur.nr) {
                const long long row = cur.r0 + lane;
                const T pv = pb[row];
                if (diag) ap += diag[row] * pv;
                tmp[off + row] = ap;
                s += pv * ap;
            }
            if (!more) break;
            cur = nxt;
            w.next();
        }
    }
    s = block_reduce<OpSum, CG_BS>(s, red);
    if (threadIdx.x == 0) part[(long long)bb * gridDim.x + blockIdx.x] = s;
}

// K1 for a dense operator: tmp = A p, per-block partial p.tmp (same grid,
// partial layout and iteration counter as cg_spmv_kernel)
template <typename T, bool ROW_MAJOR>
__global__ __launch_bounds__(CG_BS) void cg_dense_kernel(int n, const T* __restrict__ A, long long a_bstride,
                                                         const T* __restrict__ p, T* __restrict__ tmp,
                                                         T* __restrict__ part, CgState<T>* st)
{
    static_assert(CG_BS == DN_BS, "dense tiles use the CG block size");
    const int bb = blockIdx.y;
    if (blockIdx.x == 0 && bb == 0 && threadIdx.x == 0 && st->status == CG_RUNNING) st->it = st->it + 1;
    if (batch_inactive(st, bb)) return;
    __shared__ T red[4 * 64];
    __shared__ T red2[CG_BS / 64];
    const long long off = (long long)bb * n;
    const T* Ab = A + bb * a_bstride;
    const T* pb = p + off;
    T s = T(0);
    const int ntiles = div_up_dev(n, DN_TILE);
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const int r0 = t * DN_TILE;
        const T v = dense_tile<T, ROW_MAJOR>(n, n, Ab, pb, r0, 0, n, red);
        const int r = r0 + (int)threadIdx.x;
        if (threadIdx.x < 64 && r < n) {
            tmp[off + r] = v;
            s += pb[r] * v;
        }
    }
    s = block_reduce<OpSum, CG_BS>(s, red2);
    if (threadIdx.x == 0) part[(long long)bb * gridDim.x + blockIdx.x] = s;
}

// p_{i} = M^-1 r_i + beta p_{i-1} with one explicit rounding pattern, so that
// the value a lane writes for its own row and the value other lanes
// recompute for that column are bit-identical
__device__ __forceinline__ float cg_pnext(float beta, float pold, float d, float r)
{
    return __builtin_fmaf(beta, pold, d * r);
}
__device__ __forceinline__ double cg_pnext(double beta, double pold, double d, double r)
{
    return __builtin_fma(beta, pold, d * r);
}

// ---------------------------------------------------------------- KA
// Iteration 'it' (>= 1) first completes iteration it-1: every block folds
// KB's partials of |r|^2 and r.(M^-1 r) of all batches in the same order ->
// convergence test and beta; block (0, b) records batch b's state and block
// (0, 0) the global status (the launch with it == max_iter only evaluates).
// Then p_it = M^-1 r + beta p_{it-1} (written for own rows into the other p
// buffer) and Ap_it: the gathered p_it[c] are recomputed from r, M^-1 and
// p_{it-1} at column c, so the p update and the SpMV share one launch.
// FIRST (it == 0): p_0 comes from the setup, no evaluation.
// A block that returns early because block (0,0) already flagged the status
// in this same launch would have returned anyway: within it < max_iter the
// status only changes to "converged", which every block computes alike.
template <typename VT, typename T, bool FIRST>
__global__ __launch_bounds__(CG_BS) void cg_iter_kernel(int n, int it, int B, int G2, const int* __restrict__ JA,
                                                        const int* __restrict__ IA, const VT* __restrict__ vals,
                                                        long long vstride, const T* __restrict__ invd, int istride,
                                                        const T* __restrict__ r, const T* __restrict__ pprev,
                                                        T* __restrict__ pnext, T* __restrict__ tmp,
                                                        T* __restrict__ part1, const T* __restrict__ part2,
                                                        CgState<T>* st, const T* __restrict__ diag)
{
    const int bb = blockIdx.y;
    if (st->status != CG_RUNNING) return;
    const bool per_batch = st->per_batch != 0;
    // a batch that finished earlier has nothing to do; block (0,0) still
    // evaluates, it owns the global status
    if (per_batch && st->bdone[bb] != CG_RUNNING && !(FIRST == false && blockIdx.x == 0 && bb == 0)) return;
    constexpr int WPB = CG_BS / 64;
    __shared__ T lds[WPB][CG_EPT * 64];
    __shared__ T red[2 * CG_BS / 64];
    __shared__ T rr_s[CG_MAXB], rz_s[CG_MAXB];
    T* slice = lds[threadIdx.x >> 6];
    const int lane = threadIdx.x & 63;
    const long long off = (long long)bb * n;
    const VT* vb = vals + bb * vstride;
    const T* ib = invd + (long long)bb * istride;
    const T* rb = r + off;
    const T* pp = pprev + off;
    T* pn = pnext + off;
    ChunkWalk w(div_up_dev(n, 64), WPB);
    WaveChunk<VT, CG_EPT> cur, nxt;
    // the first chunk's matrix loads overlap the partial reductions below
    if (w.valid()) chunk_load<VT, CG_EPT, false>(cur, JA, IA, vb, w.cur * 64, n);
    T beta = T(0);
    if (!FIRST) {
        const int prev = it - 1;
        for (int k = 0; k < B; ++k) {
            T a, c;
            reduce_partials_pair(part2 + 2ll * k * G2, G2, a, c, red);
            if (threadIdx.x == 0) { rr_s[k] = a; rz_s[k] = c; }
        }
        __syncthreads();
        bool all = true;
        T m = 0;
        bool have_m = false;
        for (int k = 0; k < B; ++k) {
            const bool done = per_batch && st->bdone[k] != CG_RUNNING;
            all = all && (done || rr_s[k] < st->thr[k]);
            if (!done) {
                T q = rr_s[k] / st->rhs2[k];
                m = (!have_m || q > m) ? q : m;
                have_m = true;
            }
        }
        const bool own_prev = per_batch && st->bdone[bb] != CG_RUNNING;
        const bool own_conv = own_prev || rr_s[bb] < st->thr[bb];
        const T absOld = st->absNew[prev & 1][bb];
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            if (per_batch && !own_prev) {
                st->berr[bb] = sqrt((double)(rr_s[bb] / st->rhs2[bb]));
                if (own_conv) {
                    st->bdone[bb] = CG_CONVERGED;
                    st->biters[bb] = prev;  // the converging step is not counted
                } else {
                    st->biters[bb] = it;
                }
            }
            if (bb == 0) {
                if (!per_batch) st->err = sqrt((double)m);
                if (all) {
                    st->status = CG_CONVERGED;
                    st->iters = prev;
                } else {
                    st->iters = it;
                    if (it >= st->max_iter) st->status = CG_MAXIT;
                }
            }
            // with the reference's rule a converged batch keeps iterating until
            // all have converged, so its r.z is recorded like any other
            if (!(per_batch && own_conv)) st->absNew[it & 1][bb] = rz_s[bb];
        }
        if (per_batch ? own_conv : all) return;
        if (it >= st->max_iter) return;
        beta = rz_s[bb] / absOld;
    }
    T s = 0;
    if (w.valid()) {
        auto pval = [=](int c) -> T { return FIRST ? pp[c] : cg_pnext(beta, pp[c], ib[c], rb[c]); };
        for (;;) {
            const int nc = w.peek_next();
            const bool more = nc < w.hi;
            if (more) chunk_load<VT, CG_EPT, false>(nxt, JA, IA, vb, nc * 64, n);
            T ap = chunk_compute<VT, T, CG_EPT>(cur, JA, IA, vb, pval, slice);
            if (lane < cur.nr) {
                const int row = cur.r0 + lane;
                const T pv = pval(row);
                if (!FIRST) pn[row] = pv;
                if (diag) ap += diag[row] * pv;
                tmp[off + row] = ap;
                s += pv * ap;
            }
            if (!more) break;
            cur = nxt;
            w.next();
        }
    }
    s = block_reduce<OpSum, CG_BS>(s, red);
    if (threadIdx.x == 0) part1[(long long)bb * gridDim.x + blockIdx.x] = s;
}

// Vector width of K2 / K3: VW elements per load (VW = 16 / sizeof(T) when n
// is a multiple of it and every vector is 16-byte aligned, else 1).
template <typename T, int VW> struct CgVec { typedef T type __attribute__((ext_vector_type(VW))); };
template <typename T> struct CgVec<T, 1> { typedef T type; };
template <typename T, int VW>
__device__ __forceinline__ T cg_hsum(const typename CgVec<T, VW>::type& v)
{
    if constexpr (VW == 1) return v;
    else if constexpr (VW == 2) return v.x + v.y;
    else return (v.x + v.y) + (v.z + v.w);
}

// ---------------------------------------------------------------- K2
// The first sweep's operands are loaded before the partials of K1 are
// reduced, so the streaming loads overlap that reduction's latency.
// UPDATE_X = false (three-launch form): x += alpha p is left to K3, which
// reads p anyway, so this pass reads r, Ap, M^-1 and writes r only.
template <typename T, bool UPDATE_X, int VW = 1>
__global__ __launch_bounds__(CG_BS) void cg_update_kernel(int n, int it_arg, int G1, const T* __restrict__ part1,
                                                          const T* __restrict__ invd, int istride,
                                                          const T* __restrict__ p, const T* __restrict__ tmp,
                                                          T* __restrict__ x, T* __restrict__ r,
                                                          T* __restrict__ part2, const CgState<T>* st)
{
    typedef typename CgVec<T, VW>::type V;
    const int bb = blockIdx.y;
    if (batch_inactive(st, bb)) return;
    const int it = it_arg >= 0 ? it_arg : st->it;  // < 0: the counter kept by K1
    __shared__ T red[2 * CG_BS / 64];
    const long long off = (long long)bb * n;
    // vectors of VW elements (n is a multiple of VW)
    const long long nv = n / VW;
    const V* xb = reinterpret_cast<const V*>(x + off);
    V* xo = reinterpret_cast<V*>(x + off);
    const V* pb = reinterpret_cast<const V*>(p + off);
    const V* tb = reinterpret_cast<const V*>(tmp + off);
    V* rb = reinterpret_cast<V*>(r + off);
    const V* ib = reinterpret_cast<const V*>(invd + (long long)bb * istride);
    const long long stride = (long long)gridDim.x * CG_BS;
    long long i0 = (long long)blockIdx.x * CG_BS + threadIdx.x;
    V xv[CG_VPT], pv[CG_VPT], rv[CG_VPT], tv[CG_VPT], dv[CG_VPT];
    auto load = [&](long long base) {
#pragma unroll
        for (int j = 0; j < CG_VPT; ++j) {
            const long long i = base + j * stride;
            const bool ok = i < nv;
            if (UPDATE_X) {
                xv[j] = ok ? xb[i] : V(0);
                pv[j] = ok ? pb[i] : V(0);
            }
            rv[j] = ok ? rb[i] : V(0);
            tv[j] = ok ? tb[i] : V(0);
            dv[j] = ok ? ib[i] : V(0);
        }
    };
    const T rz = st->absNew[it & 1][bb];
    load(i0);
    const T pap = reduce_partials(part1 + (long long)bb * G1, G1, 1, red);
    const T alpha = rz / pap;
    V vrr = V(0), vrz = V(0);
    while (i0 < nv) {
#pragma unroll
        for (int j = 0; j < CG_VPT; ++j) {
            const long long i = i0 + j * stride;
            if (i < nv) {
                const V rn = rv[j] - alpha * tv[j];
                if (UPDATE_X) xo[i] = xv[j] + alpha * pv[j];
                rb[i] = rn;
                vrr += rn * rn;
                vrz += rn * (dv[j] * rn);
            }
        }
        i0 += stride * CG_VPT;
        if (i0 < nv) load(i0);
    }
    T srr = cg_hsum<T, VW>(vrr), srz = cg_hsum<T, VW>(vrz);
    block_sum2<CG_BS>(srr, srz, red);
    if (threadIdx.x == 0) {
        long long base = 2ll * ((long long)bb * gridDim.x + blockIdx.x);
        part2[base] = srr;
        part2[base + 1] = srz;
    }
}

// ---------------------------------------------------------------- K3 (three-launch form)
// convergence test of iteration it and p = M^-1 r + beta p
// XK3: K3 also performs x += alpha p (K2 then skips x and p: one vector pass
// less per iteration; pays off for large systems, where the pass costs more
// than the extra partial reduction K3 needs for alpha).
template <typename T, bool XK3, int VW = 1>
__global__ __launch_bounds__(CG_BS) void cg_pupdate_kernel(int n, int it_arg, int B, int G2, const T* __restrict__ part2,
                                                           const T* __restrict__ invd, int istride,
                                                           const T* __restrict__ r, T* __restrict__ p,
                                                           CgState<T>* st, int G1, const T* __restrict__ part1,
                                                           T* __restrict__ x)
{
    const int bb = blockIdx.y;
    if (st->status != CG_RUNNING) return;
    const int it = it_arg >= 0 ? it_arg : st->it;  // < 0: the counter kept by K1
    const bool per_batch = st->per_batch != 0;
    // a batch that finished in an earlier iteration has nothing to update;
    // block (0,0) still runs, it owns the global status
    if (per_batch && st->bdone[bb] != CG_RUNNING && !(blockIdx.x == 0 && bb == 0)) return;
    __shared__ T red[3 * CG_BS / 64];
    __shared__ T rr_s[CG_MAXB], rz_s[CG_MAXB], thr_s[CG_MAXB], rhs_s[CG_MAXB];
    __shared__ int done_s[CG_MAXB];
    // the state the decisions below read, fetched before the reductions so
    // that its latency overlaps theirs (and every block sees the values as
    // they were when this launch started)
    if (threadIdx.x < B) {
        thr_s[threadIdx.x] = st->thr[threadIdx.x];
        rhs_s[threadIdx.x] = st->rhs2[threadIdx.x];
        done_s[threadIdx.x] = per_batch ? st->bdone[threadIdx.x] : (int)CG_RUNNING;
    }
    const T absOld = st->absNew[it & 1][bb];
    const int max_iter = st->max_iter;
    typedef typename CgVec<T, VW>::type V;
    const long long off = (long long)bb * n;
    const long long nv = n / VW;  // vectors of VW elements (n is a multiple of VW)
    V* pb = reinterpret_cast<V*>(p + off);
    const V* rb = reinterpret_cast<const V*>(r + off);
    V* xb = reinterpret_cast<V*>(x + off);
    const V* ib = reinterpret_cast<const V*>(invd + (long long)bb * istride);
    const long long stride = (long long)gridDim.x * CG_BS;
    long long i0 = (long long)blockIdx.x * CG_BS + threadIdx.x;
    V pv[CG_VPT], rv[CG_VPT], dv[CG_VPT], xv[CG_VPT];
    auto load = [&](long long base) {
#pragma unroll
        for (int j = 0; j < CG_VPT; ++j) {
            const long long i = base + j * stride;
            const bool ok = i < nv;
            pv[j] = ok ? pb[i] : V(0);
            rv[j] = ok ? rb[i] : V(0);
            dv[j] = ok ? ib[i] : V(0);
            if (XK3) xv[j] = ok ? xb[i] : V(0);
        }
    };
    load(i0);
    // alpha of this iteration, recomputed from K1's partials exactly as K2
    // computed it (same partials, same order); every block computes the same
    // totals (fixed order); the partials of a batch that finished earlier are
    // stale and are not used
    T pap = T(1);
    if (XK3 && B == 1) {
        T a, c;
        reduce_partials_three(part1, G1, part2, G2, pap, a, c, red);
        if (threadIdx.x == 0) { rr_s[0] = a; rz_s[0] = c; }
    } else {
        if (XK3) pap = reduce_partials(part1 + (long long)bb * G1, G1, 1, red);
        for (int k = 0; k < B; ++k) {
            T a, c;
            reduce_partials_pair(part2 + 2ll * k * G2, G2, a, c, red);
            if (threadIdx.x == 0) { rr_s[k] = a; rz_s[k] = c; }
        }
    }
    __syncthreads();
    const T alpha = absOld / pap;
    bool all = true;
    T m = 0;
    bool have_m = false;
    for (int k = 0; k < B; ++k) {
        const bool prev = done_s[k] != CG_RUNNING;
        all = all && (prev || rr_s[k] < thr_s[k]);
        if (!prev) {
            T q = rr_s[k] / rhs_s[k];
            m = (!have_m || q > m) ? q : m;
            have_m = true;
        }
    }
    const bool own_prev = done_s[bb] != CG_RUNNING;
    const bool own_conv = own_prev || rr_s[bb] < thr_s[bb];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (per_batch && !own_prev) {
            st->berr[bb] = sqrt((double)(rr_s[bb] / rhs_s[bb]));
            if (own_conv) {
                st->bdone[bb] = CG_CONVERGED;
                st->biters[bb] = it;  // the converging step is not counted
            } else {
                st->biters[bb] = it + 1;
            }
        }
        if (bb == 0) {
            if (!per_batch) st->err = sqrt((double)m);
            if (all) {
                st->status = CG_CONVERGED;
                st->iters = it;
            } else {
                st->iters = it + 1;
                if (it + 1 >= max_iter) st->status = CG_MAXIT;
            }
        }
        // with the reference's rule a converged batch keeps iterating until
        // all have converged, so its r.z is recorded like any other
        if (!(per_batch && own_conv)) st->absNew[(it + 1) & 1][bb] = rz_s[bb];
    }
    // The last iteration of this batch (converged, or the iteration limit):
    // its x update is left to cg_finalize_kernel, which applies the recorded
    // alpha to the (then unchanged) p. Every block reaches the same decision
    // from the same partials, so a block that returned at the top because
    // block 0 already flagged the end in this launch skips nothing needed.
    const bool last = (per_batch ? own_conv : all) || it + 1 >= max_iter;
    if (last) {
        if (XK3 && blockIdx.x == 0 && threadIdx.x == 0 && !own_prev) {
            st->alpha_fin[bb] = alpha;
            st->xpend[bb] = 1;
        }
        return;
    }
    const T beta = rz_s[bb] / absOld;
    while (i0 < nv) {
#pragma unroll
        for (int j = 0; j < CG_VPT; ++j) {
            const long long i = i0 + j * stride;
            if (i < nv) {
                if (XK3) xb[i] = xv[j] + alpha * pv[j];
                pb[i] = dv[j] * rv[j] + beta * pv[j];
            }
        }
        i0 += stride * CG_VPT;
        if (i0 < nv) load(i0);
    }
}

// End of a solve, per batch: x = 0 for a zero right-hand side, or the
// pending x += alpha p of the batch's last iteration (three-launch form).
template <typename T>
__global__ void cg_finalize_kernel(int n, T* x, const T* p, const CgState<T>* st)
{
    const int bb = blockIdx.y;
    const bool zero = st->status == CG_EARLY_ZERO || (st->per_batch && st->bdone[bb] == CG_EARLY_ZERO);
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long long k = (long long)bb * n + i;
    if (zero) x[k] = T(0);
    else if (st->xpend[bb]) x[k] = x[k] + st->alpha_fin[bb] * p[k];
}

template <typename T>
__global__ void cg_state_init_kernel(CgState<T>* st, int max_iter, int batches, int per_batch)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        st->status = CG_RUNNING;
        st->iters = 0;
        st->it = -1;
        st->max_iter = max_iter;
        st->batches = batches;
        st->per_batch = per_batch;
        st->err = 0.0;
        for (int k = 0; k < batches && k < CG_MAXB; ++k) st->xpend[k] = 0;
    }
}

template <typename T>
__global__ void invdiag_kernel(int n, long long nnz, const int* __restrict__ JA, const int* __restrict__ IA,
                               const T* __restrict__ vals, T* __restrict__ inv)
{
    int rr = blockIdx.x * blockDim.x + threadIdx.x;
    if (rr >= n) return;
    const T* v = vals + blockIdx.y * nnz;
    T d = T(0);
    for (int k = JA[rr]; k < JA[rr + 1]; ++k)
        if (IA[k] == rr) { d = v[k]; break; }
    inv[(long long)blockIdx.y * n + rr] = (d == T(0)) ? T(1) : T(1) / d;
}

// ------------------------------------------------------------- host driver
// One three-launch iteration. The iteration index lives on the device (K1
// advances it), so consecutive iterations have identical launch arguments
// (replaying chunks of them from a captured hipGraph was measured: 72.8
// against 70.3 us per iteration with direct launches at 2048^2, so the
// solver launches directly).
// A dense operator in place of the CSR matrix (cm_pcg_dense_solve): n x n,
// batch stride bstride (0: one matrix for all right-hand sides)
template <typename T>
struct DenseOp {
    const T* A = nullptr;
    long long bstride = 0;
    bool row_major = false;
};

template <typename VT, typename T>
static void launch_cg_iteration(cm_ctx_t ctx, int n, int B, int G1, int G2, const int* JA, const int* IA,
                                const VT* vals, long long vstride, const T* invd, int istride, T* p, T* t, T* x, T* r,
                                T* part1, T* part2, CgState<T>* st, const T* diag, const DenseOp<T>* dense = nullptr)
{
    if (dense && dense->row_major)
        cg_dense_kernel<T, true><<<dim3(G1, B), CG_BS, 0, ctx->stream>>>(n, dense->A, dense->bstride, p, t, part1, st);
    else if (dense)
        cg_dense_kernel<T, false><<<dim3(G1, B), CG_BS, 0, ctx->stream>>>(n, dense->A, dense->bstride, p, t, part1,
                                                                          st);
    else
        cg_spmv_kernel<VT, T><<<dim3(G1, B), CG_BS, 0, ctx->stream>>>(n, JA, IA, vals, vstride, p, t, part1, st, diag);
    // 16-byte loads and stores when every vector allows them
    constexpr int VW = 16 / (int)sizeof(T);
    const bool wide = n % VW == 0 && ((((uintptr_t)invd | (uintptr_t)p | (uintptr_t)t | (uintptr_t)x |
                                        (uintptr_t)r) & 15) == 0);
    if ((long long)n * B >= CG_XK3_MIN && wide) {
        cg_update_kernel<T, false, VW><<<dim3(G2, B), CG_BS, 0, ctx->stream>>>(n, -1, G1, part1, invd, istride, p,
                                                                               t, x, r, part2, st);
        cg_pupdate_kernel<T, true, VW><<<dim3(G2, B), CG_BS, 0, ctx->stream>>>(n, -1, B, G2, part2, invd, istride,
                                                                               r, p, st, G1, part1, x);
    } else if ((long long)n * B >= CG_XK3_MIN) {
        cg_update_kernel<T, false><<<dim3(G2, B), CG_BS, 0, ctx->stream>>>(n, -1, G1, part1, invd, istride, p, t, x,
                                                                           r, part2, st);
        cg_pupdate_kernel<T, true><<<dim3(G2, B), CG_BS, 0, ctx->stream>>>(n, -1, B, G2, part2, invd, istride, r, p,
                                                                           st, G1, part1, x);
    } else {
        cg_update_kernel<T, true><<<dim3(G2, B), CG_BS, 0, ctx->stream>>>(n, -1, G1, part1, invd, istride, p, t, x,
                                                                          r, part2, st);
        cg_pupdate_kernel<T, false><<<dim3(G2, B), CG_BS, 0, ctx->stream>>>(n, -1, B, G2, part2, invd, istride, r,
                                                                            p, st, G1, part1, x);
    }
}

struct CgBatchOpts {
    int vals_batched = 0;   // matrix values per batch (stride nnz)
    int per_batch = 0;      // per-batch convergence
    int* biters = nullptr;  // host, B entries (per-batch mode)
    double* berr = nullptr; // host, B entries (per-batch mode)
    // two launches per iteration (p update fused into the SpMV, gathered p
    // recomputed from r, M^-1, p_prev) instead of three
    bool fused = false;
    const void* dense = nullptr;  // DenseOp<T>: a dense operator replaces the CSR matrix
};

template <typename VT, typename T>
static int pcg_multikernel(cm_ctx_t ctx, int n, int nnz, int B, const int* JA, const int* IA, const VT* vals,
                           const T* invd, const T* b, T* x, int use_guess, double tol, int max_iter, int* iters_out,
                           double* err_out, cm_cg_stats* stats, const T* diag = nullptr,
                           CgBatchOpts opt = CgBatchOpts())
{
    constexpr bool same_type = std::is_same<VT, T>::value;
    CM_REQUIRE(invd || same_type, CM_E_ARG, "CG: mixed storage needs an explicit inverse diagonal");
    const DenseOp<T>* dense = (const DenseOp<T>*)opt.dense;
    CM_REQUIRE(!dense || (invd && !opt.fused), CM_E_ARG,
               "CG: a dense operator needs an inverse diagonal and the three-launch form");
    CM_REQUIRE(B >= 1 && B <= CG_MAXB, CM_E_UNSUPPORTED, "CG: %d right-hand sides (max %d)", B, CG_MAXB);
    Temps tmp(ctx);
    const long long N = (long long)n * B;
    void *r_, *p_, *t_, *st_, *p1_, *p2_;
    CM_TRY(tmp.get(sizeof(T) * N, &r_));
    CM_TRY(tmp.get(sizeof(T) * 2 * N, &p_));  // p of even / odd iterations
    CM_TRY(tmp.get(sizeof(T) * N, &t_));
    CM_TRY(tmp.get(sizeof(CgState<T>), &st_));
    // grids: multiples of 8 (one share per XCD); with B batches each gets at
    // most CG_GMAX/B blocks so that every reduction folds <= CG_GMAX partials.
    // K1 walks 64-row chunks with 4 wavefronts per block; the setup walks
    // 512-row blocks with the same grid.
    const int gcap = std::max(8, CG_GMAX / B);
    const int G1 = walk_grid(div_up(div_up(n, 64), CG_BS / 64), gcap);
    const int G2 = walk_grid(div_up(n, (long long)CG_BS * CG_VPT), std::min(gcap, CG_G2MAX));
    CM_TRY(tmp.get(sizeof(T) * 3 * (size_t)B * std::max(G1, G2), &p1_));
    CM_TRY(tmp.get(sizeof(T) * 2 * (size_t)B * G2, &p2_));
    T* r = (T*)r_;
    T* pbuf[2] = {(T*)p_, (T*)p_ + N};
    T* p = pbuf[0];
    T* t = (T*)t_;
    T* part1 = (T*)p1_;
    T* part2 = (T*)p2_;
    CgState<T>* st = (CgState<T>*)st_;
    const long long vstride = opt.vals_batched ? (long long)nnz : 0;
    // a caller-given inverse diagonal has one row of n per batch when the
    // values are batched
    int istride = opt.vals_batched ? n : 0;
    if (!invd) {
        void* id_;
        const int nb = opt.vals_batched ? B : 1;
        CM_TRY(tmp.get(sizeof(T) * (size_t)n * nb, &id_));
        invdiag_kernel<T><<<dim3(div_up(n, 256), nb), 256, 0, ctx->stream>>>(n, vstride, JA, IA, (const T*)vals,
                                                                             (T*)id_);
        CM_LAUNCH_CHECK();
        invd = (const T*)id_;
        istride = opt.vals_batched ? n : 0;
    }
    const bool fused = opt.fused;
    int launches = 0;
    auto t0 = std::chrono::steady_clock::now();
    cg_state_init_kernel<T><<<1, 64, 0, ctx->stream>>>(st, max_iter, B, opt.per_batch);
    CM_LAUNCH_CHECK();
    const T* ax_pre = nullptr;
    if (dense && use_guess) {
        // t = A x_0, read by the setup kernel
        // one column range: the same summation order as the iteration kernel
        if (dense->row_major)
            dense_gemv_kernel<T, true><<<dim3(dense_grid(n), B), DN_BS, 0, ctx->stream>>>(
                n, n, dense->A, dense->bstride, x, 1, t, CM_ASSIGN, 1, n, nullptr);
        else
            dense_gemv_kernel<T, false><<<dim3(dense_grid(n), B), DN_BS, 0, ctx->stream>>>(
                n, n, dense->A, dense->bstride, x, 1, t, CM_ASSIGN, 1, n, nullptr);
        CM_LAUNCH_CHECK();
        ax_pre = t;
        launches += 1;
    }
    cg_setup_kernel<VT, T><<<dim3(G1, B), CG_BS, 0, ctx->stream>>>(n, JA, IA, vals, vstride, invd, istride, b, x, r,
                                                                   p, use_guess, part1, diag, ax_pre);
    CM_LAUNCH_CHECK();
    cg_setup_finalize_kernel<T><<<1, CG_BS, 0, ctx->stream>>>(G1, B, (T)tol, part1, st);
    CM_LAUNCH_CHECK();
    launches += 3;
    int* hstat = nullptr;
    CM_TRY(pinned(ctx, (void**)&hstat));
    hstat[0] = 0;
    hipEvent_t ev;
    CM_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    const int CHUNK = 8;
    bool pending = false;
    int it = 0;
    while (it < max_iter) {
        int todo = std::min(CHUNK, max_iter - it);
        for (int k = 0; k < todo; ++k, ++it) {
            if (!fused) {
                launch_cg_iteration<VT, T>(ctx, n, B, G1, G2, JA, IA, vals, vstride, invd, istride, p, t, x, r, part1,
                                           part2, st, diag, dense);
                continue;
            }
            T* pcur = pbuf[it & 1];
            if (it == 0)
                cg_iter_kernel<VT, T, true><<<dim3(G1, B), CG_BS, 0, ctx->stream>>>(
                    n, it, B, G2, JA, IA, vals, vstride, invd, istride, r, pcur, pcur, t, part1, part2, st, diag);
            else
                cg_iter_kernel<VT, T, false><<<dim3(G1, B), CG_BS, 0, ctx->stream>>>(
                    n, it, B, G2, JA, IA, vals, vstride, invd, istride, r, pbuf[(it - 1) & 1], pcur, t, part1,
                    part2, st, diag);
            cg_update_kernel<T, true><<<dim3(G2, B), CG_BS, 0, ctx->stream>>>(n, it, G1, part1, invd, istride, pcur,
                                                                              t, x, r, part2, st);
        }
        CM_LAUNCH_CHECK();
        launches += (fused ? 2 : 3) * todo;
        // lagged status check: read the previous chunk's word without draining
        // the queue, then request the current one
        if (pending) {
            count(CM_CNT_HOST_SYNC);
            CM_HIP(hipEventSynchronize(ev));
            if (hstat[0] != CG_RUNNING) break;
        }
        CM_HIP(hipMemcpyAsync(hstat, &st->status, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        CM_HIP(hipEventRecord(ev, ctx->stream));
        pending = true;
    }
    // completes the last launched iteration (convergence test / iteration
    // limit); a no-op once the status is final
    if (fused && it >= 1) {
        cg_iter_kernel<VT, T, false><<<dim3(G1, B), CG_BS, 0, ctx->stream>>>(
            n, it, B, G2, JA, IA, vals, vstride, invd, istride, r, pbuf[(it - 1) & 1], pbuf[it & 1], t, part1, part2,
            st, diag);
        CM_LAUNCH_CHECK();
        launches += 1;
    }
    cg_finalize_kernel<T><<<dim3(div_up(n, 256), B), 256, 0, ctx->stream>>>(n, x, p, st);
    CM_LAUNCH_CHECK();
    launches += 1;
    CgState<T> hs;
    CM_HIP(hipMemcpyAsync(&hs, st, sizeof(hs), hipMemcpyDeviceToHost, ctx->stream));
    CM_TRY(::cm::host_sync(ctx));
    CM_HIP(hipEventDestroy(ev));
    auto t1 = std::chrono::steady_clock::now();
    if (opt.per_batch) {
        double m = 0;
        int mi = 0;
        for (int k = 0; k < B; ++k) {
            // a batch still running at the end stopped at the iteration limit
            int itk = (hs.bdone[k] == CG_RUNNING) ? hs.iters : hs.biters[k];
            if (opt.biters) opt.biters[k] = itk;
            if (opt.berr) opt.berr[k] = hs.berr[k];
            m = std::max(m, hs.berr[k]);
            mi = std::max(mi, itk);
        }
        if (iters_out) *iters_out = mi;
        if (err_out) *err_out = m;
    } else {
        if (iters_out) *iters_out = hs.iters;
        if (err_out) *err_out = hs.err;
    }
    if (stats) {
        stats->path_used = fused ? CM_CG_FUSED : CM_CG_MULTIKERNEL;
        stats->launches = launches;
        stats->seconds = std::chrono::duration<double>(t1 - t0).count();
    }
    return CM_OK;
}

}  // namespace cm
