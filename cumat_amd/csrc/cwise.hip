// cwise.hip -- fused component-wise evaluation.
//
// Replaces CwiseEvaluationKernel<T,M,Mode> + Assignment<Dense,Cwise>
// (cuMat/src/CwiseOp.h:66-85,130-158). cuMat instantiates one kernel per
// expression type in the caller's translation unit; here the host lowers the
// expression tree to a short accumulator program (cm_instr) over strided
// leaves (cm_leaf), and one pre-compiled gfx950 kernel evaluates it:
//   * every leaf value of a thread's 4 elements is fetched before the program
//     runs, so all loads are in flight together (memory-level parallelism is
//     what a bandwidth-bound kernel needs);
//   * the program lives in the kernel-argument segment and every branch on it
//     is wave-uniform (scalar), so dispatch costs SALU slots only;
//   * leaves whose strides equal the destination's, and scalar leaves, need
//     no index math; only broadcast/strided leaves decompose the index.
// Also here: the pre-instantiated linear combination (k <= 16 streams) and
// evaluation into the non-zeros of a CSR matrix
// (CwiseCSREvaluationKernel, cuMat/src/SparseEvaluation.h:14-34).
#include "cwise_core.hpp"
#include "jit.hpp"

#include <cinttypes>

namespace cm {

// FLAT: every leaf is direct or scalar and the destination is direct, so no
// index decomposition is needed anywhere.
// Element of slot k of a thread whose first element is e0. 4-byte programs:
// a thread's 8 elements are two groups of 4 that lie 256 elements apart, so
// that one 16-byte load of the wavefront covers 1 KB contiguously; 8-byte
// programs keep 4 consecutive elements.
constexpr bool CW_SPREAD = true;
template <bool W>
__device__ __forceinline__ long long cw_elem(long long e0, int k)
{
    return (CW_SPREAD && !W) ? e0 + (k >> 2) * 256 + (k & 3) : e0 + k;
}
template <bool W>
__device__ __forceinline__ long long cw_first(long long tid)
{
    return (CW_SPREAD && !W) ? (tid >> 6) * (64 * CwVec<W>::v) + (tid & 63) * 4 : tid * CwVec<W>::v;
}

template <bool W, int NL, bool FLAT>
__global__ __launch_bounds__(CW_BS) void cwise_kernel(const CwProgram P)
{
    const long long n = P.n;
    const long long stride = (long long)gridDim.x * CW_BS * CwVec<W>::v;
    for (long long e0 = cw_first<W>((long long)blockIdx.x * CW_BS + threadIdx.x); e0 < n; e0 += stride) {
        Bank<W, NL> lv;
        const bool full = cw_elem<W>(e0, CwVec<W>::v - 1) < n;
        Coords q[CwVec<W>::v];
        if (!FLAT) {
#pragma unroll
            for (int k = 0; k < CwVec<W>::v; ++k) q[k] = decompose(P, min(cw_elem<W>(e0, k), n - 1));
        }
        // ---- fetch every leaf for the 4 elements before computing
#pragma unroll
        for (int l = 0; l < NL; ++l) {
            const cm_leaf& L = P.leaves[l];
            Reg<W> t;
#pragma unroll
            for (int k = 0; k < CwVec<W>::v; ++k) t.v[k] = 0;
            if (l >= P.nleaves) {
                // unused leaf slot
            } else if ((P.leaf_scalar >> l) & 1) {
                load_raw<W>(L, 0, t, 0);
#pragma unroll
                for (int k = 1; k < CwVec<W>::v; ++k) t.v[k] = t.v[0];
            } else if (FLAT || ((P.leaf_direct >> l) & 1)) {
                // destination-shaped leaf: contiguous, vectorizable
                const int sz = (L.dtype == CM_F64) ? 8 : (L.dtype == CM_BOOL ? 1 : 4);
                if (full && sz == 4 && (((uintptr_t)L.ptr & 15) == 0) && (FLAT || P.dst_direct)) {
                    // streamed once: non-temporal (measured +3-7 % on 2^26-element programs)
                    typedef unsigned int u4v __attribute__((ext_vector_type(4)));
#pragma unroll
                    for (int q = 0; q < CwVec<W>::v / 4; ++q) {
                        const u4v u = __builtin_nontemporal_load(
                            reinterpret_cast<const u4v*>((const uint32_t*)L.ptr + cw_elem<W>(e0, 4 * q)));
                        t.v[4 * q] = u.x; t.v[4 * q + 1] = u.y; t.v[4 * q + 2] = u.z; t.v[4 * q + 3] = u.w;
                    }
                } else {
#pragma unroll
                    for (int k = 0; k < CwVec<W>::v; ++k) {
                        long long e = min(cw_elem<W>(e0, k), n - 1);
                        long long off = (FLAT || P.dst_direct) ? e
                                         : (q[k].r * P.dst_sr + q[k].c * P.dst_sc + q[k].b * P.dst_sb);
                        load_raw<W>(L, off, t, k);
                    }
                }
            } else {
#pragma unroll
                for (int k = 0; k < CwVec<W>::v; ++k) {
                    bool z;
                    long long off = leaf_offset(L, q[k], z);
                    load_raw<W>(L, z ? 0 : off, t, k);
                    if (z) t.v[k] = 0;
                }
            }
#pragma unroll
            for (int k = 0; k < CwVec<W>::v; ++k) lv.k[k][l] = t.v[k];
        }
        Reg<W> acc;
        run_program<W, NL>(P, lv, acc);
        // ---- store
        if (FLAT && full && (P.dst_dtype == CM_F32 || P.dst_dtype == CM_I32) && (((uintptr_t)P.dst & 15) == 0)) {
            // 16-byte stores; a compound mode reads the old values with 16-byte loads
            typedef unsigned int u4v __attribute__((ext_vector_type(4)));
            const bool f32 = P.dst_dtype == CM_F32;
            if (P.mode == CM_ASSIGN) {
#pragma unroll
                for (int q = 0; q < CwVec<W>::v / 4; ++q) {
                    u4v u;
                    u.x = (uint32_t)acc.v[4 * q]; u.y = (uint32_t)acc.v[4 * q + 1];
                    u.z = (uint32_t)acc.v[4 * q + 2]; u.w = (uint32_t)acc.v[4 * q + 3];
                    __builtin_nontemporal_store(u, reinterpret_cast<u4v*>((uint32_t*)P.dst + cw_elem<W>(e0, 4 * q)));
                }
            } else {
#pragma unroll
                for (int q = 0; q < CwVec<W>::v / 4; ++q) {
                    u4v* d = reinterpret_cast<u4v*>((uint32_t*)P.dst + cw_elem<W>(e0, 4 * q));
                    const u4v old = *d;
                    u4v u;
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const uint32_t v = (uint32_t)acc.v[4 * q + j];
                        u[j] = f32 ? __float_as_uint(assign_mode<float>(P.mode, __uint_as_float(old[j]), __uint_as_float(v)))
                                   : (uint32_t)assign_mode<int>(P.mode, (int)old[j], (int)v);
                    }
                    __builtin_nontemporal_store(u, d);
                }
            }
        } else {
#pragma unroll
            for (int k = 0; k < CwVec<W>::v; ++k) {
                const long long e = cw_elem<W>(e0, k);
                if (e < n) {
                    long long off = (FLAT || P.dst_direct) ? e
                                     : (q[k].r * P.dst_sr + q[k].c * P.dst_sc + q[k].b * P.dst_sb);
                    store_elem<W>(P, off, acc, k);
                }
            }
        }
    }
}

// ------------------------------------------------ CSR destination
// The program evaluated at stored entry k of batch q.b, coordinates
// q = (row, column, batch), one entry at a time (lanes 1..3 of the
// registers idle). dst layout: vals[k + nnz*b].
template <bool W, int NL>
__device__ __forceinline__ void csr_entry_eval(const CwProgram& P, const Coords& q, long long k, long long nnz)
{
    const long long b = q.b;
    {
        Bank<W, NL> lv;
#pragma unroll
        for (int l = 0; l < NL; ++l) {
            Reg<W> t;
#pragma unroll
            for (int j = 0; j < CwVec<W>::v; ++j) t.v[j] = 0;
            if (l < P.nleaves) {
                bool z;
                long long off = leaf_offset(P.leaves[l], q, z);
                // a leaf marked sr == -1 is a value array of this pattern
                // (SparseMatrix::direct()): entry k of batch b, batch stride
                // nnz, or 0 when one batch is broadcast over several
                if (P.leaves[l].sr == -1) off = k + P.leaves[l].sb * b, z = false;
                load_raw<W>(P.leaves[l], z ? 0 : off, t, 0);
                if (z) t.v[0] = 0;
            }
#pragma unroll
            for (int j = 0; j < CwVec<W>::v; ++j) lv.k[j][l] = t.v[j];
        }
        Reg<W> acc;
        run_program<W, NL>(P, lv, acc);
        store_elem<W>(P, k + nnz * b, acc, 0);
    }
}

// Thread per row, walking the row's stored entries (the loop of
// cuMat/src/SparseEvaluation.h:14-34): any program, including leaves that
// depend on the row.
template <bool W, int NL>
__global__ void cwise_csr_kernel(const CwProgram P, const int* __restrict__ JA, const int* __restrict__ IA,
                                 long long nnz)
{
    const long long row = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long b = blockIdx.y;
    if (row >= P.rows) return;
    for (int k = JA[row]; k < JA[row + 1]; ++k) csr_entry_eval<W, NL>(P, Coords{row, (long long)IA[k], b}, k, nnz);
}

// Thread per stored entry (coalesced over the value array) for programs
// whose leaves need no row coordinate: every leaf is the value array itself
// or a plain strided leaf with row stride 0; the column is read from IA only
// when some leaf depends on it.
template <bool W, int NL>
__global__ void cwise_csr_entry_kernel(const CwProgram P, const int* __restrict__ IA, long long nnz, int need_col)
{
    const long long b = blockIdx.y;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < nnz; k += stride)
        csr_entry_eval<W, NL>(P, Coords{0, need_col ? (long long)IA[k] : 0, b}, k, nnz);
}

template <bool W, int NL>
static int launch_dense(cm_ctx_t ctx, const CwProgram& P, bool flat)
{
    const long long chunks = (P.n + CwVec<W>::v - 1) / CwVec<W>::v;
    long long blocks = (chunks + CW_BS - 1) / CW_BS;
    // grid-stride beyond 64 blocks per CU (2^26-element programs: 2-8 %
    // faster than 16 per CU; one block per 2048 elements with no loop loses
    // 6 % on a copy)
    const long long cap = (long long)(ctx->cu_count > 0 ? ctx->cu_count : 256) * 64;
    if (blocks > cap) blocks = cap;
    if (blocks < 1) blocks = 1;
    if (flat) cwise_kernel<W, NL, true><<<(unsigned)blocks, CW_BS, 0, ctx->stream>>>(P);
    else cwise_kernel<W, NL, false><<<(unsigned)blocks, CW_BS, 0, ctx->stream>>>(P);
    CM_LAUNCH_CHECK();
    return CM_OK;
}

template <bool W>
static int launch_dense_w(cm_ctx_t ctx, const CwProgram& P, bool flat)
{
    if (P.nleaves <= 4) return launch_dense<W, 4>(ctx, P, flat);
    if (P.nleaves <= 8) return launch_dense<W, 8>(ctx, P, flat);
    return launch_dense<W, CM_CWISE_MAX_LEAVES>(ctx, P, flat);
}

// dst[k] = src[IA[k] * sc]: a leaf that depends on the column, laid out
// along the stored entries (1-, 4- or 8-byte elements)
template <typename E>
__global__ void csr_gather_kernel(const E* __restrict__ src, long long sc, const int* __restrict__ IA, long long nnz,
                                  E* __restrict__ dst)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < nnz; k += stride)
        dst[k] = src[(long long)IA[k] * sc];
}

template <bool W>
static int launch_csr_w(cm_ctx_t ctx, const CwProgram& P, const int* JA, const int* IA, long long nnz)
{
    bool row_free = true, need_col = false;
    for (int l = 0; l < P.nleaves; ++l) {
        const cm_leaf& L = P.leaves[l];
        if (L.sr == -1) continue;  // the sparse value array
        if (L.kind != CM_LEAF_STRIDED || L.sr != 0) row_free = false;
        if (L.sc != 0) need_col = true;
    }
    if (row_free) {
        const long long cap = (long long)(ctx->cu_count > 0 ? ctx->cu_count : 256) * 64;
        const long long blocks = std::max<long long>(1, std::min<long long>(div_up(nnz, 256), cap));
        dim3 g((unsigned)blocks, (unsigned)P.batches);
        if (P.nleaves <= 4)
            cwise_csr_entry_kernel<W, 4><<<g, 256, 0, ctx->stream>>>(P, IA, nnz, need_col ? 1 : 0);
        else
            cwise_csr_entry_kernel<W, CM_CWISE_MAX_LEAVES><<<g, 256, 0, ctx->stream>>>(P, IA, nnz, need_col ? 1 : 0);
        CM_LAUNCH_CHECK();
        return CM_OK;
    }
    dim3 grid(div_up(P.rows, 128), (unsigned)P.batches);
    if (P.nleaves <= 4) cwise_csr_kernel<W, 4><<<grid, 128, 0, ctx->stream>>>(P, JA, IA, nnz);
    else cwise_csr_kernel<W, CM_CWISE_MAX_LEAVES><<<grid, 128, 0, ctx->stream>>>(P, JA, IA, nnz);
    CM_LAUNCH_CHECK();
    return CM_OK;
}


// ------------------------------------------------ generated kernels
// The flat case of cwise_kernel<false, NL, true> (32-bit programs; every
// leaf is direct or a broadcast scalar and the destination is contiguous)
// as straight-line code for one program: the same element mapping (8 per
// thread, two 16-byte groups 256 elements apart), the same 16-byte
// non-temporal loads and stores where the interpreter takes them, the
// same per-element semantics (cwise_ops.hpp). Returns CM_OK when launched,
// JIT_DECLINED when the interpreter should run the program.
constexpr int JIT_DECLINED = 1;

struct JitCwiseArgs {
    const void* leaf[CM_CWISE_MAX_LEAVES];
    long long gsc[CM_CWISE_MAX_LEAVES];  // gathered leaves: element stride of the index
    const int* idx;                       // gathered leaves: leaf[idx[e] * gsc]
    void* dst;
    long long n;
};

static const char* jit_raw_load(int dtype)
{
    // raw register bits of one stored element (the interpreter's load_raw;
    // 32-bit values are zero-extended in a 64-bit register)
    if (dtype == CM_BOOL) return "(R)((const unsigned char*)A.leaf[%d])[%s]";
    if (dtype == CM_F64) return "(R)((const uint64_t*)A.leaf[%d])[%s]";
    return "(R)((const uint32_t*)A.leaf[%d])[%s]";
}

// Source and cache key of the generated kernel cm_jit_cwise for P. Leaves in
// 'gmask' are gathered: element e reads leaf[idx[e] * gsc] (a leaf that
// depends on the column of a stored CSR entry, idx = the column indices).
// wide: a program with 64-bit values (FP64 leaves, instructions or
// destination), evaluated in 64-bit registers as the interpreter's wide
// path does; FP64 leaves and destinations then move two elements per
// 16-byte access. Gathered leaves are 32-bit programs only.
int jit_cwise_flat_source(const CwProgram& P, std::string& key, std::string& body, unsigned gmask = 0,
                          const int* idx = nullptr, bool wide = false)
{
    const int dst_dt = P.dst_dtype;
    if (dst_dt != CM_F32 && dst_dt != CM_I32 && dst_dt != CM_BOOL && !(wide && dst_dt == CM_F64))
        return JIT_DECLINED;
    if (wide && gmask) return JIT_DECLINED;
    const bool dst_al = ((uintptr_t)P.dst & 15) == 0;
    const bool idx_al = ((uintptr_t)idx & 15) == 0;
    key.clear();
    body.clear();
    char buf[256];
    snprintf(buf, sizeof(buf), "%s|d%d%c|m%d|", wide ? "cwflat64" : "cwflat32", dst_dt, dst_al ? 'a' : 'u', P.mode);
    key += buf;
    if (gmask) {
        snprintf(buf, sizeof(buf), "i%c|", idx_al ? 'a' : 'u');
        key += buf;
    }
    // leaf value expressions and the per-leaf load code
    std::string decls, loads_full, loads_tail, scalars, gathers;
    if (gmask) {
        // the index of a thread's 8 elements, loaded like a direct leaf
        decls += "        int ix[8];\n";
        if (idx_al)
            loads_full += "            { const u4v w = __builtin_nontemporal_load((const u4v*)(A.idx + e0));\n"
                          "              ix[0] = w.x; ix[1] = w.y; ix[2] = w.z; ix[3] = w.w; }\n"
                          "            { const u4v w = __builtin_nontemporal_load((const u4v*)(A.idx + e0 + 256));\n"
                          "              ix[4] = w.x; ix[5] = w.y; ix[6] = w.z; ix[7] = w.w; }\n";
        else
            loads_full += "#pragma unroll\n            for (int k = 0; k < 8; ++k) ix[k] = A.idx[EL(e0, k)];\n";
        loads_tail += "#pragma unroll\n            for (int k = 0; k < 8; ++k) ix[k] = A.idx[min(EL(e0, k), n - 1)];\n";
    }
    std::string names[CM_CWISE_MAX_LEAVES];
    const char* namep[CM_CWISE_MAX_LEAVES];
    // strided leaves (dense route only): neither a broadcast scalar nor in
    // the destination's layout; element e reads the leaf at the coordinates
    // (r, c, b) of e, with the shape and strides compiled in as constants
    std::string strided_loads;
    bool any_strided = false;
    for (int l = 0; l < P.nleaves; ++l) {
        const cm_leaf& L = P.leaves[l];
        if (L.dtype != CM_F32 && L.dtype != CM_I32 && L.dtype != CM_BOOL && !(wide && L.dtype == CM_F64))
            return JIT_DECLINED;
        const bool scalar = (P.leaf_scalar >> l) & 1;
        const bool gathered = (gmask >> l) & 1;
        const bool strided = !gmask && !scalar && !((P.leaf_direct >> l) & 1);
        const bool al = L.dtype != CM_BOOL && ((uintptr_t)L.ptr & 15) == 0;
        snprintf(buf, sizeof(buf), "L%d%c%c|", L.dtype, gathered ? 'g' : (scalar ? 's' : (strided ? 'x' : 'd')),
                 al ? 'a' : 'u');
        key += buf;
        if (strided) {
            if (L.kind != CM_LEAF_STRIDED || !P.dst_direct) return JIT_DECLINED;
            any_strided = true;
            snprintf(buf, sizeof(buf), "S%lld,%lld,%lld|", (long long)L.sr, (long long)L.sc, (long long)L.sb);
            key += buf;
            snprintf(buf, sizeof(buf), "v%d[k]", l);
            names[l] = buf;
            decls += "        R v" + std::to_string(l) + "[8];\n";
            std::string off;
            const long long st[3] = {L.sr, L.sc, L.sb};
            const char* cv[3] = {"cr_", "cc_", "cb_"};
            for (int a = 0; a < 3; ++a)
                if (st[a] != 0) {
                    char t[96];
                    snprintf(t, sizeof(t), "%s(long long)%s[k] * %lldll", off.empty() ? "" : " + ", cv[a], st[a]);
                    off += t;
                }
            if (off.empty()) off = "0";
            char ld[256];
            snprintf(ld, sizeof(ld), jit_raw_load(L.dtype), l, off.c_str());
            strided_loads += std::string("#pragma unroll\n        for (int k = 0; k < 8; ++k) v") +
                             std::to_string(l) + "[k] = " + ld + ";\n";
            namep[l] = names[l].c_str();
            continue;
        }
        if (gathered) {
            snprintf(buf, sizeof(buf), "v%d[k]", l);
            names[l] = buf;
            decls += "        R v" + std::to_string(l) + "[8];\n";
            char ld[160], at[64];
            snprintf(at, sizeof(at), "(long long)ix[k] * A.gsc[%d]", l);
            snprintf(ld, sizeof(ld), jit_raw_load(L.dtype), l, at);
            gathers += std::string("#pragma unroll\n        for (int k = 0; k < 8; ++k) v") + std::to_string(l) +
                       "[k] = " + ld + ";\n";
        } else if (scalar) {
            snprintf(buf, sizeof(buf), "c%d", l);
            names[l] = buf;
            char ld[160];
            snprintf(ld, sizeof(ld), jit_raw_load(L.dtype), l, "0");
            scalars += "    const R c" + std::to_string(l) + " = " + ld + ";\n";
        } else {
            snprintf(buf, sizeof(buf), "v%d[k]", l);
            names[l] = buf;
            decls += "        R v" + std::to_string(l) + "[8];\n";
            char ld[160];
            if (al && L.dtype == CM_F64) {
                // two doubles per 16-byte non-temporal load, both groups
                for (int h = 0; h < 4; ++h) {
                    const int off = (h >> 1) * 256 + (h & 1) * 2;
                    snprintf(buf, sizeof(buf),
                             "            w = __builtin_nontemporal_load((const u2v*)((const uint64_t*)A.leaf[%d] + e0 + %d));\n"
                             "            v%d[%d] = w.x; v%d[%d] = w.y;\n",
                             l, off, l, 2 * h, l, 2 * h + 1);
                    loads_full += buf;
                }
            } else if (al) {
                // streamed once: 16-byte non-temporal loads of both groups
                snprintf(buf, sizeof(buf),
                         "            u = __builtin_nontemporal_load((const u4v*)((const uint32_t*)A.leaf[%d] + e0));\n"
                         "            v%d[0] = u.x; v%d[1] = u.y; v%d[2] = u.z; v%d[3] = u.w;\n", l, l, l, l, l);
                loads_full += buf;
                snprintf(buf, sizeof(buf),
                         "            u = __builtin_nontemporal_load((const u4v*)((const uint32_t*)A.leaf[%d] + e0 + 256));\n"
                         "            v%d[4] = u.x; v%d[5] = u.y; v%d[6] = u.z; v%d[7] = u.w;\n", l, l, l, l, l);
                loads_full += buf;
            } else {
                snprintf(ld, sizeof(ld), jit_raw_load(L.dtype), l, "EL(e0, k)");
                loads_full += std::string("#pragma unroll\n            for (int k = 0; k < 8; ++k) v") +
                              std::to_string(l) + "[k] = " + ld + ";\n";
            }
            snprintf(ld, sizeof(ld), jit_raw_load(L.dtype), l, "min(EL(e0, k), n - 1)");
            loads_tail += std::string("#pragma unroll\n            for (int k = 0; k < 8; ++k) v") +
                          std::to_string(l) + "[k] = " + ld + ";\n";
        }
        namep[l] = names[l].c_str();
    }
    for (int i = 0; i < P.ncode; ++i) {
        const cm_instr& I = P.code[i];
        uint64_t fb;
        memcpy(&fb, &I.fimm, 8);
        snprintf(buf, sizeof(buf), "%d,%d,%d,%d,%d,%d,%" PRIx64 ",%" PRIx64 ";", I.opcode, I.dtype, I.fn, I.src,
                 I.arg, I.flags, fb, (uint64_t)I.iimm);
        key += buf;
    }
    std::string prog;
    if (!jit::program_body(P, namep, prog)) return JIT_DECLINED;
    // coordinates of a thread's 8 elements (the interpreter's decompose(),
    // with the shape as constants: divisions become multiplications)
    std::string coords, shape;
    if (any_strided) {
        const bool narrow = P.n <= 0xffffffffll;
        snprintf(buf, sizeof(buf), "sh%lld,%lld,%lld,%d|", (long long)P.rows, (long long)P.cols, (long long)P.batches,
                 P.row_order);
        key += buf;
        snprintf(buf, sizeof(buf),
                 "typedef %s IX;\nconstexpr IX SH_R = %lld, SH_C = %lld, SH_RC = %lld;\n",
                 narrow ? "unsigned" : "unsigned long long", (long long)P.rows, (long long)P.cols,
                 (long long)P.rows * P.cols);
        shape = buf;
        coords = "        IX cr_[8], cc_[8], cb_[8];\n"
                 "#pragma unroll\n"
                 "        for (int k = 0; k < 8; ++k) {\n"
                 "            const IX e_ = (IX)min(EL(e0, k), n - 1);\n"
                 "            cb_[k] = e_ / SH_RC;\n"
                 "            const IX w_ = e_ - cb_[k] * SH_RC;\n";
        coords += P.row_order ? "            cr_[k] = w_ / SH_C;\n            cc_[k] = w_ - cr_[k] * SH_C;\n"
                              : "            cc_[k] = w_ / SH_R;\n            cr_[k] = w_ - cc_[k] * SH_R;\n";
        coords += "        }\n";
    }
    // store of element k at e (the interpreter's store_elem)
    std::string store_one;
    if (dst_dt == CM_F64) {
        store_one = "{ double* p = (double*)A.dst + e; const double v = getd(res[k]); "
                    "*p = (MODE == CM_ASSIGN) ? v : assign1<double>(MODE, *p, v); }";
    } else if (dst_dt == CM_F32) {
        store_one = "{ float* p = (float*)A.dst + e; const float v = getf(res[k]); "
                    "*p = (MODE == CM_ASSIGN) ? v : assign1<float>(MODE, *p, v); }";
    } else if (dst_dt == CM_I32) {
        store_one = "{ int* p = (int*)A.dst + e; const int v = geti(res[k]); "
                    "*p = (MODE == CM_ASSIGN) ? v : assign1<int>(MODE, *p, v); }";
    } else {
        store_one = "{ unsigned char* p = (unsigned char*)A.dst + e; int v = geti(res[k]) != 0; "
                    "if (MODE == CM_AND) v = v && *p; else if (MODE == CM_OR) v = v || *p; *p = (unsigned char)v; }";
    }
    std::string store_vec;
    if (dst_al && dst_dt == CM_F64) {
        store_vec =
            "        if (full) {\n"
            "#pragma unroll\n"
            "            for (int h = 0; h < 4; ++h) {\n"
            "                u2v* d = (u2v*)((uint64_t*)A.dst + e0 + (h >> 1) * 256 + (h & 1) * 2);\n"
            "                u2v u;\n"
            "                if (MODE == CM_ASSIGN) {\n"
            "                    u.x = res[2 * h]; u.y = res[2 * h + 1];\n"
            "                } else {\n"
            "                    const u2v old = *d;\n"
            "#pragma unroll\n"
            "                    for (int j = 0; j < 2; ++j) u[j] = (unsigned long long)__double_as_longlong(\n"
            "                        assign1<double>(MODE, __longlong_as_double((long long)old[j]), getd(res[2 * h + j])));\n"
            "                }\n"
            "                __builtin_nontemporal_store(u, d);\n"
            "            }\n"
            "            continue;\n"
            "        }\n";
    } else if (dst_al && dst_dt != CM_BOOL) {
        const char* comb = dst_dt == CM_F32
                               ? "__float_as_uint(assign1<float>(MODE, __uint_as_float(old[j]), getf(res[4 * q + j])))"
                               : "(uint32_t)assign1<int>(MODE, (int)old[j], geti(res[4 * q + j]))";
        store_vec = std::string(
                        "        if (full) {\n"
                        "#pragma unroll\n"
                        "            for (int q = 0; q < 2; ++q) {\n"
                        "                u4v* d = (u4v*)((uint32_t*)A.dst + e0 + 256 * q);\n"
                        "                u4v u;\n"
                        "                if (MODE == CM_ASSIGN) {\n"
                        "                    u.x = (uint32_t)res[4 * q]; u.y = (uint32_t)res[4 * q + 1];\n"
                        "                    u.z = (uint32_t)res[4 * q + 2]; u.w = (uint32_t)res[4 * q + 3];\n"
                        "                } else {\n"
                        "                    const u4v old = *d;\n"
                        "#pragma unroll\n"
                        "                    for (int j = 0; j < 4; ++j) u[j] = ") +
                    comb +
                    ";\n"
                    "                }\n"
                    "                __builtin_nontemporal_store(u, d);\n"
                    "            }\n"
                    "            continue;\n"
                    "        }\n";
    }
    snprintf(buf, sizeof(buf), "constexpr int MODE = %d;\n", P.mode);
    body = std::string("typedef unsigned int u4v __attribute__((ext_vector_type(4)));\n"
                       "typedef unsigned long long u2v __attribute__((ext_vector_type(2)));\n") + shape + buf +
           "struct JitCwiseArgs { const void* leaf[" + std::to_string(CM_CWISE_MAX_LEAVES) + "]; long long gsc[" +
           std::to_string(CM_CWISE_MAX_LEAVES) + "]; const int* idx; void* dst; long long n; };\n"
           "#define EL(e0, k) ((e0) + ((k) >> 2) * 256 + ((k) & 3))\n"
           "extern \"C\" __global__ __launch_bounds__(256) void cm_jit_cwise(const JitCwiseArgs A)\n{\n"
           + std::string(wide ? "    typedef uint64_t R;\n" : "    typedef uint32_t R;\n") +
           "    const long long n = A.n;\n"
           "    const long long stride = (long long)gridDim.x * (256 * 8);\n" +
           scalars +
           "    const long long tid = (long long)blockIdx.x * 256 + threadIdx.x;\n"
           "    for (long long e0 = (tid >> 6) * 512 + (tid & 63) * 4; e0 < n; e0 += stride) {\n"
           "        const bool full = EL(e0, 7) < n;\n" +
           decls +
           "        if (full) {\n"
           "            u4v u;\n            (void)u;\n"
           "            u2v w;\n            (void)w;\n" +
           loads_full +
           "        } else {\n" +
           loads_tail +
           "        }\n" +
           coords + strided_loads +
           gathers +
           "        R res[8];\n"
           "#pragma unroll\n"
           "        for (int k = 0; k < 8; ++k) {\n"
           "            R acc = 0, t = 0, s0 = 0, s1 = 0, s2 = 0, s3 = 0;\n"
           "            (void)t; (void)s0; (void)s1; (void)s2; (void)s3;\n" +
           prog +
           "            res[k] = acc;\n"
           "        }\n" +
           store_vec +
           "#pragma unroll\n"
           "        for (int k = 0; k < 8; ++k) {\n"
           "            const long long e = EL(e0, k);\n"
           "            if (e < n) " + store_one + "\n"
           "        }\n"
           "    }\n"
           "}\n";
    return CM_OK;
}

static int jit_cwise_flat(cm_ctx_t ctx, const CwProgram& P, unsigned gmask = 0, const int* idx = nullptr,
                          bool wide = false)
{
    // everything jit_cwise_flat_source reads: flags, destination, shape,
    // leaves (strides and 16-byte alignment, not addresses) and the code
    std::string sig("cw");
    jit::sig_add(sig, (int)wide);
    jit::sig_add(sig, gmask);
    jit::sig_add(sig, (int)(((uintptr_t)idx & 15) == 0));
    jit::sig_add(sig, P.dst_dtype);
    jit::sig_add(sig, (int)(((uintptr_t)P.dst & 15) == 0));
    jit::sig_add(sig, P.mode);
    jit::sig_add(sig, P.dst_direct);
    jit::sig_add(sig, P.row_order);
    jit::sig_add(sig, P.rows);
    jit::sig_add(sig, P.cols);
    jit::sig_add(sig, P.batches);
    jit::sig_add(sig, (int)(P.n <= 0xffffffffll));
    jit::sig_add(sig, P.leaf_scalar);
    jit::sig_add(sig, P.leaf_direct);
    jit::sig_add(sig, P.nleaves);
    for (int l = 0; l < P.nleaves; ++l) {
        const cm_leaf& L = P.leaves[l];
        jit::sig_add(sig, L.dtype);
        jit::sig_add(sig, L.kind);
        jit::sig_add(sig, L.sr);
        jit::sig_add(sig, L.sc);
        jit::sig_add(sig, L.sb);
        jit::sig_add(sig, (int)(((uintptr_t)L.ptr & 15) == 0));
    }
    jit::sig_add(sig, P.ncode);
    for (int i = 0; i < P.ncode; ++i) {
        const cm_instr& I = P.code[i];
        jit::sig_add(sig, I.opcode);
        jit::sig_add(sig, I.dtype);
        jit::sig_add(sig, I.fn);
        jit::sig_add(sig, I.src);
        jit::sig_add(sig, I.arg);
        jit::sig_add(sig, I.flags);
        jit::sig_add(sig, I.fimm);
        jit::sig_add(sig, I.iimm);
    }
    jit::Kernel k;
    if (jit::find_sig(sig, &k) != 0) {
        std::string key, body;
        if (jit_cwise_flat_source(P, key, body, gmask, idx, wide) != CM_OK) return JIT_DECLINED;
        if (jit::get(key, body, "cm_jit_cwise", &k) != 0) return JIT_DECLINED;
        jit::put_sig(sig, k);
    }
    JitCwiseArgs A;
    memset(&A, 0, sizeof(A));
    for (int l = 0; l < P.nleaves; ++l) {
        A.leaf[l] = P.leaves[l].ptr;
        A.gsc[l] = P.leaves[l].sc;
    }
    A.idx = idx;
    A.dst = P.dst;
    A.n = P.n;
    // the interpreter's grid: one block per 2048 elements, at most 64 per CU
    long long blocks = (P.n + 2047) / 2048;
    const long long cap = (long long)(ctx->cu_count > 0 ? ctx->cu_count : 256) * 64;
    if (blocks > cap) blocks = cap;
    if (blocks < 1) blocks = 1;
    if (jit::launch(k, (unsigned)blocks, 256, ctx->stream, &A, sizeof(A)) != 0)
        return fail(CM_E_HIP, "cm_cwise_eval: launch of the generated kernel failed");
    count(CM_CNT_KERNEL_LAUNCH);
    return CM_OK;
}

// ------------------------------------------------ linear combination
constexpr int LC_MAX = 16;
struct LincombArgs {
    const float* v[LC_MAX];
    float c[LC_MAX];
};

// out = sum_i c_i v_i, 16-byte vector loads of all k streams issued before
// the first FMA; the grid is sized to the element count (one float4 per
// thread), so the loop body runs once per thread at the benchmark size.
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <int K>
__global__ __launch_bounds__(256) void lincomb_kernel(const LincombArgs A, float* __restrict__ out, long long n4)
{
    const long long stride = (long long)gridDim.x * 256;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n4; i += stride) {
        f32x4 v[K];
#pragma unroll
        for (int j = 0; j < K; ++j) v[j] = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(A.v[j]) + i);
        f32x4 acc = v[0] * A.c[0];
#pragma unroll
        for (int j = 1; j < K; ++j) acc += v[j] * A.c[j];
        __builtin_nontemporal_store(acc, reinterpret_cast<f32x4*>(out) + i);
    }
}

__global__ void lincomb_tail_kernel(const LincombArgs A, int k, float* out, long long start, long long n)
{
    long long i = start + (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float acc = A.v[0][i] * A.c[0];
    for (int j = 1; j < k; ++j) acc += A.v[j][i] * A.c[j];
    out[i] = acc;
}

template <int K>
static void launch_lincomb_k(cm_ctx_t ctx, const LincombArgs& A, float* out, long long n4)
{
    long long blocks = (n4 + 255) / 256;
    const long long cap = 1ll << 30;
    if (blocks > cap) blocks = cap;
    lincomb_kernel<K><<<(unsigned)blocks, 256, 0, ctx->stream>>>(A, out, n4);
}

// out[0..n) = sum_j A.c[j] * A.v[j][0..n): 16-byte path for the aligned
// prefix, scalar tail
static int lincomb_run(cm_ctx_t ctx, const LincombArgs& A, int k, float* out, long long n)
{
    bool aligned = ((uintptr_t)out & 15) == 0;
    for (int j = 0; j < k; ++j) aligned = aligned && (((uintptr_t)A.v[j] & 15) == 0);
    long long n4 = aligned ? n / 4 : 0;
    if (n4 > 0) {
        switch (k) {
#define LC_CASE(K) case K: launch_lincomb_k<K>(ctx, A, out, n4); break;
            LC_CASE(1) LC_CASE(2) LC_CASE(3) LC_CASE(4) LC_CASE(5) LC_CASE(6) LC_CASE(7) LC_CASE(8)
            LC_CASE(9) LC_CASE(10) LC_CASE(11) LC_CASE(12) LC_CASE(13) LC_CASE(14) LC_CASE(15) LC_CASE(16)
#undef LC_CASE
        default: break;
        }
        CM_LAUNCH_CHECK();
    }
    long long start = n4 * 4;
    if (start < n) {
        lincomb_tail_kernel<<<div_up(n - start, 256), 256, 0, ctx->stream>>>(A, k, out, start, n);
        CM_LAUNCH_CHECK();
    }
    return CM_OK;
}

// A flat FP32 program of the form  v0 [* c0] + v1*c1 + ... (one LOAD, an
// optional multiply by an immediate, then FMA terms) assigned with '=' is a
// linear combination: it runs on the dedicated streaming kernel.
static bool match_lincomb(const CwProgram& P, LincombArgs& A, int& k)
{
    auto vec_leaf = [&](int l) {
        return l >= 0 && l < P.nleaves && P.leaves[l].dtype == CM_F32 && P.leaves[l].kind == CM_LEAF_STRIDED &&
               ((P.leaf_direct >> l) & 1) && !((P.leaf_scalar >> l) & 1);
    };
    if (P.ncode < 1 || P.ncode > LC_MAX + 1) return false;
    const cm_instr& I0 = P.code[0];
    if (I0.opcode != CM_OP_LOAD || I0.src != CM_SRC_LEAF || I0.dtype != CM_F32 || !vec_leaf(I0.arg)) return false;
    memset(&A, 0, sizeof(A));
    A.v[0] = (const float*)P.leaves[I0.arg].ptr;
    A.c[0] = 1.0f;
    k = 1;
    int pc = 1;
    if (pc < P.ncode && P.code[pc].opcode == CM_OP_BINARY) {
        const cm_instr& I = P.code[pc];
        if (I.fn != CM_BIN_MUL || I.src != CM_SRC_IMM || I.dtype != CM_F32) return false;
        A.c[0] = (float)I.fimm;
        ++pc;
    }
    for (; pc < P.ncode; ++pc) {
        const cm_instr& I = P.code[pc];
        if (I.opcode != CM_OP_FMA || I.dtype != CM_F32 || !vec_leaf(I.arg) || k >= LC_MAX) return false;
        A.v[k] = (const float*)P.leaves[I.arg].ptr;
        A.c[k] = (float)I.fimm;
        ++k;
    }
    return k >= 2;
}

}  // namespace cm

namespace cm {
static int cwise_dense(cm_ctx_t ctx, const cm_leaf* leaves, int nleaves, const cm_instr* code, int ncode, void* dst,
                       int dst_dtype, int rows, int cols, int batches, int64_t dst_sr, int64_t dst_sc, int64_t dst_sb,
                       int mode);
}  // namespace cm

extern "C" {

int cm_cwise_eval(cm_ctx_t ctx, const cm_leaf* leaves, int nleaves, const cm_instr* code, int ncode, void* dst,
                  int dst_dtype, int rows, int cols, int batches, int64_t dst_sr, int64_t dst_sc, int64_t dst_sb,
                  int mode)
{
    CM_API_BEGIN
    CM_TRY(cm::resolve(ctx, &ctx));
    CM_REQUIRE(rows >= 0 && cols >= 0 && batches >= 0, CM_E_DIM, "cm_cwise_eval: negative size");
    CM_REQUIRE(mode >= CM_ASSIGN && mode <= CM_OR, CM_E_ARG, "cm_cwise_eval: bad mode %d", mode);
    CM_REQUIRE(dst_dtype != CM_BF16 && cm::dtype_size(dst_dtype) > 0, CM_E_UNSUPPORTED,
               "cm_cwise_eval: dst dtype %d", dst_dtype);
    cm::count(CM_CNT_EVAL_CWISE);
    cm::count(CM_CNT_EVAL_ANY);
    return cm::cwise_dense(ctx, leaves, nleaves, code, ncode, dst, dst_dtype, rows, cols, batches, dst_sr, dst_sc,
                           dst_sb, mode);
    CM_API_END
}

}  // extern "C"

namespace cm {

// the dense evaluation behind cm_cwise_eval (arguments checked there)
static int cwise_dense(cm_ctx_t ctx, const cm_leaf* leaves, int nleaves, const cm_instr* code, int ncode, void* dst,
                       int dst_dtype, int rows, int cols, int batches, int64_t dst_sr, int64_t dst_sc, int64_t dst_sb,
                       int mode)
{
    const long long n = (long long)rows * cols * batches;
    if (n == 0) return CM_OK;
    CM_REQUIRE(dst, CM_E_ARG, "cm_cwise_eval: dst is NULL");
    cm::CwProgram P;
    bool wide = false;
    CM_TRY(cm::build_program(leaves, nleaves, code, ncode, P, wide));
    if (dst_dtype == CM_F64) wide = true;
    P.dst = dst;
    P.dst_dtype = dst_dtype;
    P.mode = mode;
    P.dst_sr = dst_sr;
    P.dst_sc = dst_sc;
    P.dst_sb = dst_sb;
    P.rows = rows;
    P.cols = cols;
    P.batches = batches;
    P.n = n;
    const long long R = rows, Cc = cols;
    const bool colmajor = (dst_sr == 1 || R == 1) && (dst_sc == R || Cc == 1) && (dst_sb == R * Cc || batches == 1);
    const bool rowmajor = (dst_sc == 1 || Cc == 1) && (dst_sr == Cc || R == 1) && (dst_sb == R * Cc || batches == 1);
    P.row_order = (!colmajor && rowmajor) ? 1 : 0;
    P.dst_direct = (colmajor || rowmajor) ? 1 : 0;
    bool flat = P.dst_direct;
    for (int l = 0; l < nleaves; ++l) {
        const cm_leaf& L = leaves[l];
        bool scalar = (L.kind == CM_LEAF_STRIDED) && (L.sr == 0 || R == 1) && (L.sc == 0 || Cc == 1) &&
                      (L.sb == 0 || batches == 1);
        bool direct = (L.kind == CM_LEAF_STRIDED) && (L.sr == dst_sr || R == 1) && (L.sc == dst_sc || Cc == 1) &&
                      (L.sb == dst_sb || batches == 1);
        if (scalar) P.leaf_scalar |= 1u << l;
        else if (direct) P.leaf_direct |= 1u << l;
        else flat = false;
    }
    if (!wide && flat && P.dst_direct && mode == CM_ASSIGN && dst_dtype == CM_F32) {
        cm::LincombArgs A;
        int k = 0;
        if (cm::match_lincomb(P, A, k)) return cm::lincomb_run(ctx, A, k, (float*)dst, n);
    }
    // (also programs with strided or broadcast leaves, coordinates compiled in)
    if (P.dst_direct && cm::jit::want(n)) {
        const int s = cm::jit_cwise_flat(ctx, P, 0, nullptr, wide);
        if (s != cm::JIT_DECLINED) return s;
    }
    if (wide) return cm::launch_dense_w<true>(ctx, P, flat);
    return cm::launch_dense_w<false>(ctx, P, flat);
}

}  // namespace cm

extern "C" {

int cm_lincomb_f32(cm_ctx_t ctx, int k, const float* const* vecs, const float* coeffs, float* out, int64_t n)
{
    CM_API_BEGIN
    CM_TRY(cm::resolve(ctx, &ctx));
    CM_REQUIRE(k >= 1 && k <= cm::LC_MAX, CM_E_ARG, "cm_lincomb_f32: k=%d out of range 1..%d", k, cm::LC_MAX);
    CM_REQUIRE(n >= 0, CM_E_DIM, "cm_lincomb_f32: n < 0");
    cm::count(CM_CNT_EVAL_CWISE);
    cm::count(CM_CNT_EVAL_ANY);
    if (n == 0) return CM_OK;
    cm::LincombArgs A;
    memset(&A, 0, sizeof(A));
    for (int j = 0; j < k; ++j) {
        CM_REQUIRE(vecs[j], CM_E_ARG, "cm_lincomb_f32: vector %d is NULL", j);
        A.v[j] = vecs[j];
        A.c[j] = coeffs[j];
    }
    return cm::lincomb_run(ctx, A, k, out, n);
    CM_API_END
}

int cm_csr_cwise_eval(cm_ctx_t ctx, const cm_leaf* leaves, int nleaves, const cm_instr* code, int ncode, int rows,
                      int cols, int nnz, int batches, const int* JA, const int* IA, void* vals, int val_dtype,
                      int mode)
{
    CM_API_BEGIN
    CM_TRY(cm::resolve(ctx, &ctx));
    CM_REQUIRE(rows >= 0 && cols >= 0 && nnz >= 0 && batches >= 1, CM_E_DIM, "cm_csr_cwise_eval: bad sizes");
    CM_REQUIRE(mode >= CM_ASSIGN && mode <= CM_OR, CM_E_ARG, "cm_csr_cwise_eval: bad mode %d", mode);
    cm::count(CM_CNT_EVAL_CWISE_SPARSE);
    cm::count(CM_CNT_EVAL_ANY);
    if (rows == 0 || nnz == 0) return CM_OK;
    CM_REQUIRE(JA && IA && vals, CM_E_ARG, "cm_csr_cwise_eval: NULL pointer");
    CM_REQUIRE(val_dtype != CM_BF16 && cm::dtype_size(val_dtype) > 0, CM_E_UNSUPPORTED,
               "cm_csr_cwise_eval: value dtype %d", val_dtype);
    // the program is checked before anything is launched
    CM_REQUIRE((leaves || nleaves == 0) && code, CM_E_ARG, "cm_csr_cwise_eval: NULL leaves or program");
    cm::CwProgram P;
    bool wide = false;
    CM_TRY(cm::build_program(leaves, nleaves, code, ncode, P, wide));
    if (val_dtype == CM_F64) wide = true;
    // a program whose leaves are the value arrays, scalars and leaves that
    // depend on the column only (row stride 0; one batch or batch stride 0)
    // needs no row coordinate
    bool row_free = true;
    unsigned value_mask = 0, gather_mask = 0, scalar_mask = 0;
    size_t gbytes = 0;
    for (int l = 0; row_free && l < nleaves; ++l) {
        const cm_leaf& L = leaves[l];
        if (L.sr == -1) {  // a value array
            value_mask |= 1u << l;
            continue;
        }
        const int es = cm::dtype_size(L.dtype);
        row_free = L.kind == CM_LEAF_STRIDED && L.sr == 0 && (es == 1 || es == 4 || es == 8);
        if (row_free && L.sc != 0) {
            row_free = L.sb == 0 || batches == 1;
            gather_mask |= 1u << l;
            gbytes += ((size_t)es * nnz + 255) / 256 * 256;
        } else if (row_free) {
            scalar_mask |= 1u << l;
        }
    }
    if (row_free && batches == 1 && !wide && cm::jit::want(nnz)) {
        // one generated kernel over the stored entries: value arrays read
        // directly, column leaves gathered through IA, scalars broadcast
        cm::CwProgram G = P;
        for (int l = 0; l < nleaves; ++l)
            if ((value_mask >> l) & 1) G.leaves[l].sr = 1, G.leaves[l].sc = 0, G.leaves[l].sb = 0;
        G.leaf_direct = value_mask;
        G.leaf_scalar = scalar_mask;
        G.dst = vals;
        G.dst_dtype = val_dtype;
        G.mode = mode;
        G.rows = nnz;
        G.cols = 1;
        G.batches = 1;
        G.n = nnz;
        G.dst_direct = 1;
        const int s = cm::jit_cwise_flat(ctx, G, gather_mask, IA);
        if (s != cm::JIT_DECLINED) return s;
    }
    if (row_free) {
        // each column leaf is first gathered along IA into an nnz-long
        // temporary, then the program is a dense program over the nnz x 1 x
        // batches value array (value leaves become leaves with row stride 1
        // and their own batch stride, the same offsets as below), run by the
        // streaming dense kernels
        void* tmp = nullptr;
        if (gbytes) CM_TRY(cm::scratch(ctx, 0, gbytes, &tmp));
        cm_leaf dl[CM_CWISE_MAX_LEAVES];
        size_t off = 0;
        const long long cus = ctx->cu_count > 0 ? ctx->cu_count : 256;
        const unsigned gblocks = (unsigned)std::max<long long>(1, std::min<long long>(cm::div_up(nnz, 256), cus * 16));
        for (int l = 0; l < nleaves; ++l) {
            dl[l] = leaves[l];
            if ((value_mask >> l) & 1) {
                dl[l].sr = 1;
                dl[l].sc = 0;
            } else if ((gather_mask >> l) & 1) {
                const int es = cm::dtype_size(dl[l].dtype);
                void* g = (char*)tmp + off;
                if (es == 4)
                    cm::csr_gather_kernel<uint32_t><<<gblocks, 256, 0, ctx->stream>>>(
                        (const uint32_t*)dl[l].ptr, dl[l].sc, IA, nnz, (uint32_t*)g);
                else if (es == 8)
                    cm::csr_gather_kernel<uint64_t><<<gblocks, 256, 0, ctx->stream>>>(
                        (const uint64_t*)dl[l].ptr, dl[l].sc, IA, nnz, (uint64_t*)g);
                else
                    cm::csr_gather_kernel<uint8_t><<<gblocks, 256, 0, ctx->stream>>>(
                        (const uint8_t*)dl[l].ptr, dl[l].sc, IA, nnz, (uint8_t*)g);
                CM_LAUNCH_CHECK();
                dl[l].ptr = g;
                dl[l].sr = 1;
                dl[l].sc = 0;
                dl[l].sb = 0;
                off += ((size_t)es * nnz + 255) / 256 * 256;
            }
        }
        return cm::cwise_dense(ctx, dl, nleaves, code, ncode, vals, val_dtype, nnz, 1, batches, 1, nnz, nnz, mode);
    }
    P.dst = vals;
    P.dst_dtype = val_dtype;
    P.mode = mode;
    P.rows = rows;
    P.cols = cols;
    P.batches = batches;
    P.n = (long long)nnz * batches;
    if (wide) return cm::launch_csr_w<true>(ctx, P, JA, IA, nnz);
    return cm::launch_csr_w<false>(ctx, P, JA, IA, nnz);
    CM_API_END
}

}  // extern "C"
