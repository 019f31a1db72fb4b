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
//       status, the iteration count and the error. K3 decides from the
//       state K1 recorded when the iteration started, so every block of a
//       batch takes the same decision, and in a batch's last iteration it
//       applies the x update only.
// Setup: one vector pass (reads b and M^-1 without a guess; r_0 = b and
// x_0 = 0 are not written where K2 / K3 can use b and zero instead) and a
// one-block fold that initialises the state. The host learns the state from
// a host-mapped mirror the kernels publish to (no copies): it keeps at most
// two chunks of CG_CHUNK iterations in flight and reads the result after
// the final synchronisation.
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
#include <cstring>
#include <map>
#include <type_traits>

#include "bcsr.hpp"
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
    // status and per-batch status as K1 of the running iteration found them:
    // K3 decides from these, since its block (0, 0) changes the live ones
    // while the other blocks of the same launch still read
    int status_k1;
    int bdone_k1[CG_MAXB];
    int status;
    int iters;
    int it;        // index of the running iteration (three-launch form)
    int max_iter;
    int batches;
    int per_batch;
    double err;
};

// The host's view of a solve, in host-mapped coherent memory: the kernels
// that decide the state publish it here at the end of every CG_CHUNK-th
// iteration and when the solve ends, so the host can throttle its launches
// and read the result without a copy.
constexpr int CG_CHUNK = 8;
struct CgMirror {
    int status;
    int iters;
    int it_done;  // last iteration whose convergence test has run (-1: none)
    int pad;
    double err;
    int bdone[CG_MAXB];
    int biters[CG_MAXB];
    double berr[CG_MAXB];
    int xset[CG_MAXB];  // batch's x holds a value (guess, zeros or an update)
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

// Several right-hand sides (B > 1): the (|r|^2, r.z) pair sums of all
// batches in one pass instead of B block reductions. CG_BS / B threads per
// batch each fold a fixed stride of the batch's G pairs in order (G <=
// CG_GMAX / B, so at most CG_PPT + 1 loads each), then thread k adds its
// batch's sums in thread order: the same fixed order, and so the same
// result, in every block. lds: 2 * CG_BS entries; rr / rz get B entries
// (valid after a barrier).
template <typename T>
__device__ __forceinline__ void reduce_pairs_batched(const T* __restrict__ part, int G, int B, T* rr, T* rz, T* lds)
{
    typedef T T2 __attribute__((ext_vector_type(2)));
    const int tpb = CG_BS / B;
    const int k = threadIdx.x / tpb, sub = threadIdx.x - k * tpb;
    T a = T(0), c = T(0);
    if (k < B) {
        const T* pb = part + 2ll * k * G;
        T2 v[CG_PPT + 1];
#pragma unroll
        for (int j = 0; j <= CG_PPT; ++j) {
            const int i = sub + j * tpb;
            v[j] = (i < G) ? *reinterpret_cast<const T2*>(pb + 2ll * i) : T2{T(0), T(0)};
        }
#pragma unroll
        for (int j = 0; j <= CG_PPT; ++j)
            if (sub + j * tpb < G) {
                a += v[j].x;
                c += v[j].y;
            }
    }
    lds[threadIdx.x] = a;
    lds[CG_BS + threadIdx.x] = c;
    __syncthreads();
    if ((int)threadIdx.x < B) {
        const int b0 = threadIdx.x * tpb;
        T sa = lds[b0], sc = lds[CG_BS + b0];
        for (int j = 1; j < tpb; ++j) {
            sa += lds[b0 + j];
            sc += lds[CG_BS + b0 + j];
        }
        rr[threadIdx.x] = sa;
        rz[threadIdx.x] = sc;
    }
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

// Setup without a guess (x_0 = 0, r_0 = b, p_0 = M^-1 b), vectors of VW
// elements: reads b and M^-1 only. r_0 and x_0 are written only when asked
// (the three-launch form reads b for r_0 and writes x first in K3). Same
// partial layout as cg_setup_kernel: (|b|^2, |r|^2, r.p) per block.
template <typename T, int VW>
__global__ __launch_bounds__(CG_BS) void cg_setup_zero_kernel(int n, const T* __restrict__ invd, int istride,
                                                              const T* __restrict__ b, T* __restrict__ x,
                                                              T* __restrict__ r, T* __restrict__ p, int write_x,
                                                              int write_r, T* __restrict__ part)
{
    typedef T V __attribute__((ext_vector_type(VW)));
    __shared__ T red[2 * CG_BS / 64];
    const int bb = blockIdx.y;
    const long long off = (long long)bb * n;
    const long long nv = n / VW;
    const V* bv = reinterpret_cast<const V*>(b + off);
    const V* iv = reinterpret_cast<const V*>(invd + (long long)bb * istride);
    V* xv = reinterpret_cast<V*>(x + off);
    V* rv = reinterpret_cast<V*>(r + off);
    V* pv = reinterpret_cast<V*>(p + off);
    const long long stride = (long long)gridDim.x * CG_BS;
    V vb[CG_VPT], vd[CG_VPT];
    V sbb = V(0), srp = V(0);
    for (long long i0 = (long long)blockIdx.x * CG_BS + threadIdx.x; i0 < nv; i0 += stride * CG_VPT) {
#pragma unroll
        for (int j = 0; j < CG_VPT; ++j) {
            const long long i = min(i0 + j * stride, nv - 1);  // unconditional loads
            vb[j] = bv[i];
            vd[j] = iv[i];
        }
#pragma unroll
        for (int j = 0; j < CG_VPT; ++j) {
            const long long i = i0 + j * stride;
            if (i < nv) {
                const V q = vd[j] * vb[j];
                pv[i] = q;
                if (write_r) rv[i] = vb[j];
                if (write_x) xv[i] = V(0);
                sbb += vb[j] * vb[j];
                srp += vb[j] * q;
            }
        }
    }
    T a = sbb[0], c = srp[0];
#pragma unroll
    for (int k = 1; k < VW; ++k) {
        a += sbb[k];
        c += srp[k];
    }
    block_sum2<CG_BS>(a, c, red);
    if (threadIdx.x == 0) {
        const long long base = 3ll * ((long long)bb * gridDim.x + blockIdx.x);
        part[base + 0] = a;
        part[base + 1] = a;  // r_0 = b
        part[base + 2] = c;
    }
}

// Folds the setup's partials and initialises the whole state (and the
// host's mirror of it): early outs for a zero right-hand side or a guess
// that already converged, the thresholds and r.z of iteration 0.
// x_set: x holds a value already (the guess, or zeros written by the setup).
template <typename T>
__global__ __launch_bounds__(CG_BS) void cg_setup_finalize_kernel(int G, int B, T tol, const T* __restrict__ part,
                                                                  CgState<T>* st, int max_iter, int per_batch,
                                                                  CgMirror* mirror, int x_set)
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
            mirror->bdone[k] = st->bdone[k];
            mirror->biters[k] = 0;
            mirror->berr[k] = st->berr[k];
            mirror->xset[k] = x_set;
        }
        st->iters = 0;
        st->it = -1;
        st->max_iter = max_iter;
        st->batches = B;
        st->per_batch = per_batch;
        int status;
        if (per_batch ? (all_done && all_zero) : all_zero) { status = CG_EARLY_ZERO; st->err = 0.0; }
        else if (per_batch ? all_done : all_conv) { status = CG_EARLY_GUESS; st->err = sqrt((double)m); }
        else if (max_iter <= 0) { status = CG_MAXIT; st->err = sqrt((double)m); }
        else { status = CG_RUNNING; st->err = sqrt((double)m); }
        st->status = status;
        st->status_k1 = status;
        mirror->iters = 0;
        mirror->it_done = -1;
        mirror->err = st->err;
        mirror->status = status;
    }
}

// K1 of every form records the state it starts from (see CgState) and owns
// the iteration counter that K2 and K3 read: no launch argument changes
// from one iteration to the next, so iterations can be replayed from a
// captured graph (K1 itself never reads the counter, K2 / K3 read it after
// this launch has completed)
template <typename T>
__device__ __forceinline__ void cg_k1_snapshot(CgState<T>* st, int bb)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (bb == 0) {
            const int s0 = st->status;
            st->status_k1 = s0;
            if (s0 == CG_RUNNING) st->it = st->it + 1;
        }
        st->bdone_k1[bb] = st->bdone[bb];
    }
}

// ---------------------------------------------------------------- K1 (three-launch form)
// tmp = A p, per-block partial p.tmp
template <typename VT, typename T, int MNT = 0>
__global__ __launch_bounds__(CG_BS) void cg_spmv_kernel(int n, const int* __restrict__ JA, const int* __restrict__ IA,
                                                        const VT* __restrict__ vals, long long vstride,
                                                        const T* __restrict__ p, T* __restrict__ tmp,
                                                        T* __restrict__ part, CgState<T>* st,
                                                        const T* __restrict__ diag)
{
    const int bb = blockIdx.y;
    cg_k1_snapshot(st, bb);
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
        chunk_load<VT, CG_EPT, MNT>(cur, JA, IA, vb, w.cur * 64, n);
        for (;;) {
            const int nc = w.peek_next();
            const bool more = nc < w.hi;
            if (more) chunk_load<VT, CG_EPT, MNT>(nxt, JA, IA, vb, nc * 64, n);
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
    cg_k1_snapshot(st, bb);
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

// K1 for a block-CSR operator of D x D blocks (bcsr.hpp): the vectors of
// the solve hold one float4 per block row (n = 4 * rows floats per batch);
// tmp = A p and the per-block partial p.tmp, same grid rule, partial layout
// and iteration counter as cg_spmv_kernel. NT: stream the blocks past the
// caches (a matrix larger than the last-level cache gains nothing from
// keeping it: 96^3 3x3 blocks, 319 MB, 78.9 -> 71 us per product)
template <int D, bool NT>
__global__ __launch_bounds__(CG_BS) void cg_bcsr_kernel(int rows, const int* __restrict__ JA,
                                                        const int* __restrict__ IA, const bc_f4* __restrict__ vals,
                                                        long long vstride, const float* __restrict__ p,
                                                        float* __restrict__ tmp, float* __restrict__ part,
                                                        CgState<float>* st)
{
    const int bb = blockIdx.y;
    cg_k1_snapshot(st, bb);
    if (batch_inactive(st, bb)) return;
    constexpr int WPB = CG_BS / 64;
    __shared__ bc_f4 lds[WPB][BC_WIN];
    __shared__ float red[WPB];
    bc_f4* slice = lds[threadIdx.x >> 6];
    const int lane = threadIdx.x & 63;
    const long long off = (long long)bb * rows;
    const bc_f4* pb = reinterpret_cast<const bc_f4*>(p) + off;
    bc_f4* tb = reinterpret_cast<bc_f4*>(tmp) + off;
    const bc_f4* vb = vals + bb * vstride;
    float s = 0.0f;
    for (ChunkWalk w(div_up_dev(rows, 64), WPB); w.valid(); w.next()) {
        const int r0 = w.cur * 64, nr = min(64, rows - r0);
        const bc_f4 ap = bcsr_chunk_product<D, NT>(JA, IA, vb, r0, nr, [pb](int c) { return pb[c]; }, slice);
        if (lane < nr) {
            const bc_f4 pv = pb[r0 + lane];
            tb[r0 + lane] = ap;
            s += pv.x * ap.x + pv.y * ap.y + pv.z * ap.z + pv.w * ap.w;
        }
    }
    s = block_reduce<OpSum, CG_BS>(s, red);
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
                                                        CgState<T>* st, const T* __restrict__ diag,
                                                        CgMirror* __restrict__ mirror)
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
    __shared__ T pairs[2 * CG_BS];
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
        if (B == 1) {
            T a, c;
            reduce_partials_pair(part2, G2, a, c, red);
            if (threadIdx.x == 0) { rr_s[0] = a; rz_s[0] = c; }
        } else {
            reduce_pairs_batched(part2, G2, B, rr_s, rz_s, pairs);
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
            const bool pub = all || it >= st->max_iter || (prev % CG_CHUNK) == CG_CHUNK - 1;
            if (per_batch && !own_prev && (pub || own_conv)) {
                mirror->berr[bb] = st->berr[bb];
                mirror->biters[bb] = st->biters[bb];
                mirror->bdone[bb] = st->bdone[bb];
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
                if (pub) {
                    mirror->err = st->err;
                    mirror->iters = st->iters;
                    mirror->it_done = prev;
                    mirror->status = st->status;
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
// loads / stores of a vector stream, non-temporal (streaming) when NT
template <bool NT, typename V>
__device__ __forceinline__ V cg_ld(const V* p)
{
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}
template <bool NT, typename V>
__device__ __forceinline__ void cg_st(V* p, V v)
{
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

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
// VNT: streams loaded / stored non-temporally (bit 0 r, 1 tmp, 2 M^-1,
// 3 the r store)
template <typename T, bool UPDATE_X, int VW = 1, int VNT = 0>
__global__ __launch_bounds__(CG_BS) void cg_update_kernel(int n, int it_arg, int G1, const T* __restrict__ part1,
                                                          const T* __restrict__ invd, int istride,
                                                          const T* __restrict__ p, const T* __restrict__ tmp,
                                                          T* __restrict__ x, T* __restrict__ r,
                                                          T* __restrict__ part2, const CgState<T>* st,
                                                          const T* __restrict__ r0)
{
    // r0: when set, iteration 0 reads r_0 from it (r_0 = b without a guess:
    // the setup then does not write r)
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
    const V* rs = (it == 0 && r0) ? reinterpret_cast<const V*>(r0 + off) : rb;
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
            rv[j] = ok ? cg_ld<(VNT & 1) != 0>(rs + i) : V(0);
            tv[j] = ok ? cg_ld<(VNT & 2) != 0>(tb + i) : V(0);
            dv[j] = ok ? cg_ld<(VNT & 4) != 0>(ib + i) : V(0);
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
                cg_st<(VNT & 8) != 0>(rb + i, rn);
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
// than the extra partial reduction K3 needs for alpha). In a batch's last
// iteration (converged, or the iteration limit) K3 applies only that x
// update. x0_zero: x_0 = 0 was not written, so iteration 0 stores alpha p
// without reading x.
// The early outs and the decisions read the state as K1 of this iteration
// found it (status_k1, bdone_k1), never what block (0, 0) of this launch
// writes, so every block of a batch reaches the same decision.
// Block (0, 0) publishes the global state and block (0, b) batch b's to
// the host mirror at the end of every CG_CHUNK-th iteration and at the end.
// VNT: streams loaded / stored non-temporally (bit 0 p, 1 r, 2 M^-1, 3 x,
// 4 the x store, 5 the p store)
template <typename T, bool XK3, int VW = 1, int VNT = 0>
__global__ __launch_bounds__(CG_BS) void cg_pupdate_kernel(int n, int it_arg, int B, int G2, const T* __restrict__ part2,
                                                           const T* __restrict__ invd, int istride,
                                                           const T* __restrict__ r, T* __restrict__ p,
                                                           CgState<T>* st, int G1, const T* __restrict__ part1,
                                                           T* __restrict__ x, CgMirror* __restrict__ mirror,
                                                           int x0_zero)
{
    const int bb = blockIdx.y;
    if (st->status_k1 != CG_RUNNING) return;
    const int it = it_arg >= 0 ? it_arg : st->it;  // < 0: the counter kept by K1
    const bool per_batch = st->per_batch != 0;
    // a batch that finished in an earlier iteration has nothing to update;
    // block (0,0) still runs, it owns the global status
    if (per_batch && st->bdone_k1[bb] != CG_RUNNING && !(blockIdx.x == 0 && bb == 0)) return;
    __shared__ T red[3 * CG_BS / 64];
    __shared__ T rr_s[CG_MAXB], rz_s[CG_MAXB], thr_s[CG_MAXB], rhs_s[CG_MAXB];
    __shared__ int done_s[CG_MAXB];
    __shared__ T pairs[2 * CG_BS];
    // the state the decisions below read, fetched before the reductions so
    // that its latency overlaps theirs
    if (threadIdx.x < B) {
        thr_s[threadIdx.x] = st->thr[threadIdx.x];
        rhs_s[threadIdx.x] = st->rhs2[threadIdx.x];
        done_s[threadIdx.x] = per_batch ? st->bdone_k1[threadIdx.x] : (int)CG_RUNNING;
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
    // x_0 = 0 unwritten: iteration 0 stores alpha p without reading x
    const bool read_x = XK3 && !(x0_zero && it == 0);
    V pv[CG_VPT], rv[CG_VPT], dv[CG_VPT], xv[CG_VPT];
    auto load = [&](long long base) {
#pragma unroll
        for (int j = 0; j < CG_VPT; ++j) {
            const long long i = base + j * stride;
            const bool ok = i < nv;
            pv[j] = ok ? cg_ld<(VNT & 1) != 0>(pb + i) : V(0);
            rv[j] = ok ? cg_ld<(VNT & 2) != 0>(rb + i) : V(0);
            dv[j] = ok ? cg_ld<(VNT & 4) != 0>(ib + i) : V(0);
            if (XK3) xv[j] = (ok && read_x) ? cg_ld<(VNT & 8) != 0>(xb + i) : V(0);
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
        if (B == 1) {
            T a, c;
            reduce_partials_pair(part2, G2, a, c, red);
            if (threadIdx.x == 0) { rr_s[0] = a; rz_s[0] = c; }
        } else {
            reduce_pairs_batched(part2, G2, B, rr_s, rz_s, pairs);
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
    // this batch's last iteration: converged, or the iteration limit
    const bool last = (per_batch ? own_conv : all) || it + 1 >= max_iter;
    const bool publish = last || (it % CG_CHUNK) == CG_CHUNK - 1 || (per_batch && own_conv);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (per_batch && !own_prev) {
            const double e = sqrt((double)(rr_s[bb] / rhs_s[bb]));
            const int bi = own_conv ? it : it + 1;  // the converging step is not counted
            st->berr[bb] = e;
            st->biters[bb] = bi;
            if (own_conv) st->bdone[bb] = CG_CONVERGED;
            if (publish) {
                mirror->berr[bb] = e;
                mirror->biters[bb] = bi;
                mirror->bdone[bb] = own_conv ? (int)CG_CONVERGED : (int)CG_RUNNING;
            }
        }
        if (XK3 && !own_prev && read_x == false) mirror->xset[bb] = 1;
        if (bb == 0) {
            const int status = all ? (int)CG_CONVERGED : (it + 1 >= max_iter ? (int)CG_MAXIT : (int)CG_RUNNING);
            const int iters = all ? it : it + 1;
            const double err = per_batch ? st->err : sqrt((double)m);
            st->err = err;
            st->iters = iters;
            st->status = status;
            const bool pub_all = publish || status != CG_RUNNING;
            if (pub_all) {
                mirror->err = err;
                mirror->iters = iters;
                mirror->it_done = it;
                mirror->status = status;
            }
        }
        // with the reference's rule a converged batch keeps iterating until
        // all have converged, so its r.z is recorded like any other
        if (!(per_batch && own_conv)) st->absNew[(it + 1) & 1][bb] = rz_s[bb];
    }
    // block (0, 0) of a batch that finished earlier only evaluates; without
    // XK3 the last iteration's x update was K2's
    if (own_prev || (last && !XK3)) return;
    const T beta = rz_s[bb] / absOld;
    while (i0 < nv) {
#pragma unroll
        for (int j = 0; j < CG_VPT; ++j) {
            const long long i = i0 + j * stride;
            if (i < nv) {
                if (XK3) cg_st<(VNT & 16) != 0>(xb + i, xv[j] + alpha * pv[j]);
                if (!last) cg_st<(VNT & 32) != 0>(pb + i, dv[j] * rv[j] + beta * pv[j]);
            }
        }
        i0 += stride * CG_VPT;
        if (i0 < nv) load(i0);
    }
}

// After a solve: x = 0 for the batches in 'mask' (a zero right-hand side,
// or no iteration ran on an x_0 = 0 that was never written)
template <typename T>
__global__ void cg_zero_batches_kernel(int n, T* x, unsigned long long mask)
{
    const int bb = blockIdx.y;
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && ((mask >> bb) & 1ull)) x[(long long)bb * n + i] = T(0);
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
// K3 updates x (instead of K2) from this many unknowns on
inline bool cg_xk3(int n, int B) { return (long long)n * B >= CG_XK3_MIN; }

template <typename T>
struct DenseOp {
    const T* A = nullptr;
    long long bstride = 0;
    bool row_major = false;
};

// a block-CSR operator (FP32 blocks of D float4 rows, bcsr.hpp)
struct BlockOp {
    const int* JA = nullptr;
    const int* IA = nullptr;
    const bc_f4* vals = nullptr;
    long long vstride = 0;  // float4 entries between the batches' values (0: shared)
    int rows = 0;           // block rows; the solve's vectors hold 4 * rows floats per batch
    int D = 3;
    long long nnz = 0;      // stored blocks
};

// blocks above this many bytes are streamed (non-temporal loads) in K1: half
// of the 256 MB last-level cache
constexpr long long CG_BCSR_STREAM_BYTES = 128ll << 20;

inline void launch_cg_bcsr(hipStream_t stream, const BlockOp& op, int B, int G1, const float* p, float* t,
                           float* part1, CgState<float>* st)
{
    const dim3 grid(G1, B);
    const long long bytes = 16ll * op.D * op.nnz * (op.vstride ? B : 1);
#define CG_BCSR_LAUNCH(DD, NT)                                                                                 \
    cg_bcsr_kernel<DD, NT><<<grid, CG_BS, 0, stream>>>(op.rows, op.JA, op.IA, op.vals, op.vstride, p, t, part1, st)
    if (bytes > CG_BCSR_STREAM_BYTES) {
        if (op.D == 2) CG_BCSR_LAUNCH(2, true);
        else if (op.D == 3) CG_BCSR_LAUNCH(3, true);
        else CG_BCSR_LAUNCH(4, true);
    } else {
        if (op.D == 2) CG_BCSR_LAUNCH(2, false);
        else if (op.D == 3) CG_BCSR_LAUNCH(3, false);
        else CG_BCSR_LAUNCH(4, false);
    }
#undef CG_BCSR_LAUNCH
}

template <typename VT, typename T>
static void launch_cg_iteration(cm_ctx_t ctx, int n, int B, int G1, int G2, const int* JA, const int* IA,
                                const VT* vals, long long vstride, const T* invd, int istride, T* p, T* t, T* x, T* r,
                                T* part1, T* part2, CgState<T>* st, const T* diag, const DenseOp<T>* dense,
                                const T* r0, CgMirror* mirror, int x0_zero, const BlockOp* block = nullptr)
{
    if (block) {
        if constexpr (std::is_same<T, float>::value) launch_cg_bcsr(ctx->stream, *block, B, G1, p, t, part1, st);
    } else if (dense && dense->row_major)
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
    if (cg_xk3(n, B) && wide && (((uintptr_t)r0 & 15) == 0)) {
        // Ap is dead after this pass: streaming loads keep it out of the caches
        cg_update_kernel<T, false, VW, 2><<<dim3(G2, B), CG_BS, 0, ctx->stream>>>(n, -1, G1, part1, invd, istride,
                                                                                  p, t, x, r, part2, st, r0);
        cg_pupdate_kernel<T, true, VW><<<dim3(G2, B), CG_BS, 0, ctx->stream>>>(
            n, -1, B, G2, part2, invd, istride, r, p, st, G1, part1, x, mirror, x0_zero);
    } else if (cg_xk3(n, B)) {
        cg_update_kernel<T, false><<<dim3(G2, B), CG_BS, 0, ctx->stream>>>(n, -1, G1, part1, invd, istride, p, t, x,
                                                                           r, part2, st, r0);
        cg_pupdate_kernel<T, true><<<dim3(G2, B), CG_BS, 0, ctx->stream>>>(n, -1, B, G2, part2, invd, istride, r, p,
                                                                           st, G1, part1, x, mirror, x0_zero);
    } else {
        cg_update_kernel<T, true><<<dim3(G2, B), CG_BS, 0, ctx->stream>>>(n, -1, G1, part1, invd, istride, p, t, x,
                                                                          r, part2, st, r0);
        cg_pupdate_kernel<T, false><<<dim3(G2, B), CG_BS, 0, ctx->stream>>>(n, -1, B, G2, part2, invd, istride, r,
                                                                            p, st, G1, part1, x, mirror, 0);
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
    const BlockOp* block = nullptr;  // a block-CSR operator replaces the CSR matrix (FP32)
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
    const BlockOp* block = opt.block;
    CM_REQUIRE(!block || (invd && !opt.fused && !dense && std::is_same<T, float>::value && n == 4 * block->rows),
               CM_E_ARG, "CG: a block operator needs FP32, an inverse diagonal and the three-launch form");
    CM_REQUIRE(B >= 1 && B <= CG_MAXB, CM_E_UNSUPPORTED, "CG: %d right-hand sides (max %d)", B, CG_MAXB);
    Temps tmp(ctx);
    const long long N = (long long)n * B;
    // grids: multiples of 8 (one share per XCD); with B batches each gets at
    // most CG_GMAX/B blocks so that every reduction folds <= CG_GMAX partials.
    // K1 walks 64-row chunks with 4 wavefronts per block; the setup walks
    // 512-row blocks with the same grid.
    const int gcap = std::max(8, CG_GMAX / B);
    // (a block operator walks 64-block-row chunks of its rows)
    const int G1 = walk_grid(div_up(div_up(block ? block->rows : n, 64), CG_BS / 64), gcap);
    const int G2 = walk_grid(div_up(n, (long long)CG_BS * CG_VPT), std::min(gcap, CG_G2MAX));
    // reduce_pairs_batched folds at most CG_PPT + 1 pairs per thread
    CM_REQUIRE(B == 1 || G2 <= (CG_PPT + 1) * (CG_BS / B), CM_E_UNSUPPORTED, "CG: grid %d too large for %d batches",
               G2, B);
    // one block for all work vectors (one allocator round trip per solve):
    // r, p of even / odd iterations, Ap, the state, the two partial sets,
    // and the inverse diagonal when the caller gives none; 256-byte aligned
    const bool own_invd = invd == nullptr;
    const int nbi = opt.vals_batched ? B : 1;
    auto al = [](size_t v) { return (v + 255) & ~(size_t)255; };
    const size_t sz[7] = {al(sizeof(T) * N), al(sizeof(T) * 2 * N), al(sizeof(T) * N), al(sizeof(CgState<T>)),
                          al(sizeof(T) * 3 * (size_t)B * std::max(G1, G2)), al(sizeof(T) * 2 * (size_t)B * G2),
                          own_invd ? al(sizeof(T) * (size_t)n * nbi) : 0};
    size_t total = 0;
    for (size_t v : sz) total += v;
    char* blk = nullptr;
    CM_TRY(tmp.get(total, (void**)&blk));
    char* cur = blk;
    auto take = [&](int k) {
        char* q = cur;
        cur += sz[k];
        return q;
    };
    T* r = (T*)take(0);
    T* p_ = (T*)take(1);
    T* pbuf[2] = {p_, p_ + N};
    T* p = pbuf[0];
    T* t = (T*)take(2);
    CgState<T>* st = (CgState<T>*)take(3);
    T* part1 = (T*)take(4);
    T* part2 = (T*)take(5);
    T* id_ = (T*)take(6);
    const long long vstride = opt.vals_batched ? (long long)nnz : 0;
    // a caller-given inverse diagonal has one row of n per batch when the
    // values are batched
    int istride = opt.vals_batched ? n : 0;
    if (own_invd) {
        invdiag_kernel<T><<<dim3(div_up(n, 256), nbi), 256, 0, ctx->stream>>>(n, vstride, JA, IA, (const T*)vals,
                                                                              id_);
        CM_LAUNCH_CHECK();
        invd = id_;
        istride = opt.vals_batched ? n : 0;
    }
    const bool fused = opt.fused;
    // without a guess, x_0 = 0 and r_0 = b are not written: K2 reads b for
    // r_0, and in the three-launch form with the x update in K3, iteration 0
    // stores alpha p without reading x (fused and small forms get x_0 = 0
    // from the setup)
    const bool xk3 = !fused && cg_xk3(n, B);
    const int x0_zero = (!use_guess && xk3) ? 1 : 0;
    const T* r0 = use_guess ? nullptr : b;
    void* mirror_host = nullptr;
    void* mirror_dev = nullptr;
    CM_TRY(pinned(ctx, &mirror_host, &mirror_dev));
    static_assert(sizeof(CgMirror) <= CM_PINNED_BYTES, "the solve mirror must fit the pinned page");
    volatile CgMirror* hm = (volatile CgMirror*)mirror_host;
    CgMirror* mirror = (CgMirror*)mirror_dev;
    // the previous solve on this context left its final state here
    hm->status = CG_RUNNING;
    hm->it_done = -1;
    int launches = 0;
    auto t0 = std::chrono::steady_clock::now();
    int Gs = G1;
    if (use_guess) {
        const T* ax_pre = nullptr;
        if (dense) {
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
        if (block) {
            // t = A x_0 (float4 entries), read by the setup kernel
            if constexpr (std::is_same<T, float>::value)
                launch_bcsr_spmv(ctx->stream, block->D, block->rows, B, block->JA, block->IA, block->vals,
                                 block->vstride, (const bc_f4*)x, block->rows, (bc_f4*)t, CM_ASSIGN);
            CM_LAUNCH_CHECK();
            ax_pre = t;
            launches += 1;
        }
        cg_setup_kernel<VT, T><<<dim3(G1, B), CG_BS, 0, ctx->stream>>>(n, JA, IA, vals, vstride, invd, istride, b,
                                                                       x, r, p, 1, part1, diag, ax_pre);
    } else {
        constexpr int VW = 16 / (int)sizeof(T);
        Gs = G2;
        if (n % VW == 0 && ((((uintptr_t)invd | (uintptr_t)b | (uintptr_t)x | (uintptr_t)r | (uintptr_t)p) & 15) == 0))
            cg_setup_zero_kernel<T, VW><<<dim3(Gs, B), CG_BS, 0, ctx->stream>>>(n, invd, istride, b, x, r, p,
                                                                                !x0_zero, 0, part1);
        else
            cg_setup_zero_kernel<T, 1><<<dim3(Gs, B), CG_BS, 0, ctx->stream>>>(n, invd, istride, b, x, r, p,
                                                                               !x0_zero, 0, part1);
    }
    CM_LAUNCH_CHECK();
    cg_setup_finalize_kernel<T><<<1, CG_BS, 0, ctx->stream>>>(Gs, B, (T)tol, part1, st, max_iter,
                                                              opt.per_batch ? 1 : 0, mirror, x0_zero ? 0 : 1);
    CM_LAUNCH_CHECK();
    launches += 2;
    // At most two chunks of iterations in flight: before issuing more, wait
    // until the mirror shows the chunk before the last one tested (kernels
    // publish every CG_CHUNK-th iteration), and stop once the solve ended.
    // If the stream drains first, every launched iteration has run.
    auto wait_mirror = [&](int need) {
        count(CM_CNT_HOST_SYNC);
        for (unsigned k = 1;; ++k) {
            if (hm->status != CG_RUNNING || hm->it_done >= need) return;
            if ((k & 255) == 0 && hipStreamQuery(ctx->stream) == hipSuccess) return;
            __builtin_ia32_pause();
        }
    };
    int it = 0;
    while (it < max_iter) {
        int todo = std::min(CG_CHUNK, max_iter - it);
        for (int k = 0; k < todo; ++k, ++it) {
            if (!fused) {
                launch_cg_iteration<VT, T>(ctx, n, B, G1, G2, JA, IA, vals, vstride, invd, istride, p, t, x, r, part1,
                                           part2, st, diag, dense, r0, mirror, x0_zero, block);
                continue;
            }
            T* pcur = pbuf[it & 1];
            if (it == 0)
                cg_iter_kernel<VT, T, true><<<dim3(G1, B), CG_BS, 0, ctx->stream>>>(
                    n, it, B, G2, JA, IA, vals, vstride, invd, istride, r, pcur, pcur, t, part1, part2, st, diag,
                    mirror);
            else
                cg_iter_kernel<VT, T, false><<<dim3(G1, B), CG_BS, 0, ctx->stream>>>(
                    n, it, B, G2, JA, IA, vals, vstride, invd, istride, r, pbuf[(it - 1) & 1], pcur, t, part1,
                    part2, st, diag, mirror);
            cg_update_kernel<T, true><<<dim3(G2, B), CG_BS, 0, ctx->stream>>>(n, it, G1, part1, invd, istride, pcur,
                                                                              t, x, r, part2, st, r0);
        }
        CM_LAUNCH_CHECK();
        launches += (fused ? 2 : 3) * todo;
        if (it > CG_CHUNK && it < max_iter) {
            wait_mirror(it - CG_CHUNK - 1);
            if (hm->status != CG_RUNNING) break;
        }
    }
    // completes the last launched iteration (convergence test / iteration
    // limit); a no-op once the status is final
    if (fused && it >= 1) {
        cg_iter_kernel<VT, T, false><<<dim3(G1, B), CG_BS, 0, ctx->stream>>>(
            n, it, B, G2, JA, IA, vals, vstride, invd, istride, r, pbuf[(it - 1) & 1], pbuf[it & 1], t, part1, part2,
            st, diag, mirror);
        CM_LAUNCH_CHECK();
        launches += 1;
    }
    CM_TRY(::cm::host_sync(ctx));
    // x = 0 where the right-hand side was zero, or where x_0 = 0 was never
    // written because no iteration ran for the batch
    unsigned long long zmask = 0;
    for (int k = 0; k < B; ++k)
        if (hm->status == CG_EARLY_ZERO || (opt.per_batch && hm->bdone[k] == CG_EARLY_ZERO) || !hm->xset[k])
            zmask |= 1ull << k;
    if (zmask) {
        cg_zero_batches_kernel<T><<<dim3(div_up(n, 256), B), 256, 0, ctx->stream>>>(n, x, zmask);
        CM_LAUNCH_CHECK();
        launches += 1;
        CM_TRY(::cm::host_sync(ctx));
    }
    tmp.drained = true;  // the stream is idle: the work block returns to the cache at once
    auto t1 = std::chrono::steady_clock::now();
    CgMirror hs;
    memcpy(&hs, (const void*)hm, sizeof(hs));
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
