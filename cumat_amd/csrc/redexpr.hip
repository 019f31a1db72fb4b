// redexpr.hip -- reductions whose input is an unevaluated component-wise
// expression (one launch, plus one combine launch for long reductions).
//
// cuMat's reductions consume the expression lazily: the input iterator calls
// coeff() on the expression tree (StridedMatrixInputIterator,
// cuMat/src/Iterator.h:99-105), and dot() is cwiseDot(rhs).sum<Row|Column>()
// fused into the reduction (cuMat/src/ReductionOpsPlugin.inl:290-297,
// cuMat/src/ReductionOps.h:564-598). Here the expression is the accumulator
// program of cm_cwise_eval, evaluated at the reduction's element loads.
//
// The expression has no memory layout of its own, so the reduction walks a
// virtual order in which the reduced axes are the inner ones: element j of
// segment s has its kept coordinates from s (in the output's layout order)
// and its reduced coordinates from j (row, column, batch, innermost first).
// Every reduction is then a set of 'nseg' contiguous segments of length L:
// a block per chunk of a segment evaluates the program for 8 (4 for 64-bit
// programs) elements per thread and pass, folds them in a fixed order, and
// long segments fold their per-chunk partials in a second launch, the same
// order every run (results are bit-reproducible).
#include "cwise_core.hpp"
#include "jit.hpp"

#include <limits>

namespace cm {

constexpr int RX_BS = 256;

template <typename Op, typename T> struct RxIdentity;
template <typename T> struct RxIdentity<OpSum, T> { __device__ static T v() { return T(0); } };
template <typename T> struct RxIdentity<OpProd, T> { __device__ static T v() { return T(1); } };
template <typename T> struct RxIdentity<OpMin, T> { __device__ static T v() { return std::numeric_limits<T>::max(); } };
template <typename T> struct RxIdentity<OpMax, T> { __device__ static T v() { return std::numeric_limits<T>::lowest(); } };
template <typename T> struct RxIdentity<OpAll, T> { __device__ static T v() { return T(1); } };
template <typename T> struct RxIdentity<OpAny, T> { __device__ static T v() { return T(0); } };
template <typename T> struct RxIdentity<OpBand, T> { __device__ static T v() { return ~T(0); } };
template <typename T> struct RxIdentity<OpBor, T> { __device__ static T v() { return T(0); } };

// Virtual order: up to three kept axes (kdim / kaxis, innermost first, in
// the output layout's order) and up to three reduced axes (rdim / raxis).
// Axis ids: 0 row, 1 column, 2 batch.
struct RxShape {
    long long kdim[3], rdim[3];
    int kaxis[3], raxis[3];
    int nk, nr;
    long long L, nseg;
};

template <typename T> __device__ __forceinline__ T rx_get(const Reg<false>& r, int k);
template <> __device__ __forceinline__ float rx_get<float>(const Reg<false>& r, int k) { return getf(r, k); }
template <> __device__ __forceinline__ int rx_get<int>(const Reg<false>& r, int k) { return geti(r, k); }
template <> __device__ __forceinline__ double rx_get<double>(const Reg<false>&, int) { return 0.0; }
template <typename T> __device__ __forceinline__ T rx_get(const Reg<true>& r, int k);
template <> __device__ __forceinline__ float rx_get<float>(const Reg<true>& r, int k) { return getf(r, k); }
template <> __device__ __forceinline__ int rx_get<int>(const Reg<true>& r, int k) { return geti(r, k); }
template <> __device__ __forceinline__ double rx_get<double>(const Reg<true>& r, int k) { return getd(r, k); }

__device__ __forceinline__ Coords rx_coords(const RxShape& S, long long s, long long j)
{
    long long c[3] = {0, 0, 0};
#pragma unroll
    for (int a = 0; a < 3; ++a)
        if (a < S.nk) {
            const long long q = s / S.kdim[a];
            c[S.kaxis[a]] = s - q * S.kdim[a];
            s = q;
        }
#pragma unroll
    for (int a = 0; a < 3; ++a)
        if (a < S.nr) {
            const long long q = j / S.rdim[a];
            c[S.raxis[a]] = j - q * S.rdim[a];
            j = q;
        }
    Coords q;
    q.r = c[0];
    q.c = c[1];
    q.b = c[2];
    return q;
}

// FLAT: every leaf is a broadcast scalar or lies in memory in the virtual
// order (element j of segment s at s*L + j), so no coordinates are formed.
template <typename Op, typename T, typename S, bool W, int NL, bool FLAT>
__global__ __launch_bounds__(RX_BS) void reduce_expr_kernel(const CwProgram P, const RxShape SH, int nchunk,
                                                            long long chunk, T* __restrict__ partials,
                                                            S* __restrict__ out, int sp)
{
    constexpr int V = CwVec<W>::v;
    __shared__ T lds[RX_BS / 64];
    const long long s = blockIdx.x / nchunk;
    const int c = blockIdx.x % nchunk;
    const long long beg = c * chunk;
    const long long end = min(SH.L, beg + chunk);
    T acc = RxIdentity<Op, T>::v();
    // element of slot k in the pass starting at p0: 32-bit reductions with a
    // leaf in the virtual order (sp) give a thread two groups of 4 that lie
    // 256 elements apart (one 16-byte load of a wavefront then covers 1 KB
    // of that leaf); others V consecutive elements
    const bool SPREAD = (FLAT || sp) && !W;
    const long long t0 = SPREAD ? (long long)(threadIdx.x >> 6) * 512 + (threadIdx.x & 63) * 4
                                : (long long)threadIdx.x * V;
    for (long long p0 = beg; p0 < end; p0 += (long long)RX_BS * V) {
        const long long j0 = p0 + t0;
        if (!SPREAD && j0 >= end) break;
        Bank<W, NL> lv;
#pragma unroll
        for (int k = 0; k < V; ++k) {
            const long long j = SPREAD ? j0 + (k >> 2) * 256 + (k & 3) : j0 + k;
            const bool ok = j < end;
            if constexpr (FLAT) {
                const long long e = s * SH.L + (ok ? j : beg);
#pragma unroll
                for (int l = 0; l < NL; ++l) {
                    Reg<W> t;
                    t.v[0] = 0;
                    if (l < P.nleaves) load_raw<W>(P.leaves[l], ((P.leaf_scalar >> l) & 1) ? 0 : e, t, 0);
                    lv.k[k][l] = t.v[0];
                }
            } else {
                const Coords q = rx_coords(SH, s, ok ? j : beg);
#pragma unroll
                for (int l = 0; l < NL; ++l) {
                    Reg<W> t;
                    t.v[0] = 0;
                    if (l < P.nleaves) {
                        bool z;
                        const long long off = leaf_offset(P.leaves[l], q, z);
                        load_raw<W>(P.leaves[l], z ? 0 : off, t, 0);
                        if (z) t.v[0] = 0;
                    }
                    lv.k[k][l] = t.v[0];
                }
            }
        }
        Reg<W> val;
        run_program<W, NL>(P, lv, val);
#pragma unroll
        for (int k = 0; k < V; ++k) {
            const long long j = SPREAD ? j0 + (k >> 2) * 256 + (k & 3) : j0 + k;
            if (j < end) acc = Op::apply(acc, rx_get<T>(val, k));
        }
    }
    acc = block_reduce<Op, RX_BS>(acc, lds);
    if (threadIdx.x == 0) {
        if (nchunk == 1) out[s] = (S)acc;
        else partials[blockIdx.x] = acc;
    }
}

template <typename Op, typename T, typename S>
__global__ __launch_bounds__(RX_BS) void reduce_expr_combine_kernel(const T* __restrict__ partials, int nchunk,
                                                                    S* __restrict__ out)
{
    __shared__ T lds[RX_BS / 64];
    const long long s = blockIdx.x;
    T r = RxIdentity<Op, T>::v();
    for (int k = threadIdx.x; k < nchunk; k += RX_BS) r = Op::apply(r, partials[s * nchunk + k]);
    r = block_reduce<Op, RX_BS>(r, lds);
    if (threadIdx.x == 0) out[s] = (S)r;
}

// Whether every leaf of P is a broadcast scalar or lies in the virtual order
// of SH (marks the scalars in P.leaf_scalar).
static bool rx_flat(CwProgram& P, const RxShape& SH)
{
    const long long dim[3] = {P.rows, P.cols, P.batches};
    long long es[3] = {0, 0, 0};
    long long st = 1;
    for (int i = 0; i < SH.nr; ++i) es[SH.raxis[i]] = st, st *= SH.rdim[i];
    for (int i = 0; i < SH.nk; ++i) es[SH.kaxis[i]] = st, st *= SH.kdim[i];
    P.leaf_scalar = 0;
    for (int l = 0; l < P.nleaves; ++l) {
        const cm_leaf& L = P.leaves[l];
        if (L.kind != CM_LEAF_STRIDED) return false;
        const long long ls[3] = {L.sr, L.sc, L.sb};
        bool scalar = true, flat = true;
        for (int a = 0; a < 3; ++a) {
            if (dim[a] <= 1) continue;
            scalar = scalar && ls[a] == 0;
            flat = flat && ls[a] == es[a];
        }
        if (scalar) P.leaf_scalar |= 1u << l;
        else if (!flat) return false;
    }
    return true;
}

// Whether some leaf of P that is not a broadcast scalar lies in the virtual
// order of SH.
static bool rx_any_virtual(const CwProgram& P, const RxShape& SH)
{
    const long long dim[3] = {P.rows, P.cols, P.batches};
    long long es[3] = {0, 0, 0};
    long long st = 1;
    for (int i = 0; i < SH.nr; ++i) es[SH.raxis[i]] = st, st *= SH.rdim[i];
    for (int i = 0; i < SH.nk; ++i) es[SH.kaxis[i]] = st, st *= SH.kdim[i];
    for (int l = 0; l < P.nleaves; ++l) {
        const cm_leaf& L = P.leaves[l];
        if (L.kind != CM_LEAF_STRIDED) continue;
        const long long ls[3] = {L.sr, L.sc, L.sb};
        bool scalar = true, virt = true;
        for (int a = 0; a < 3; ++a) {
            if (dim[a] <= 1) continue;
            scalar = scalar && ls[a] == 0;
            virt = virt && ls[a] == es[a];
        }
        if (virt && !scalar) return true;
    }
    return false;
}

constexpr int RX_JIT_DECLINED = 1;

// Arguments of the generated reduction kernel cm_jit_reduce.
struct JitRxArgs {
    const void* leaf[CM_CWISE_MAX_LEAVES];
    void* partials;
    void* out;
    long long L, chunk;
    int nchunk, pad;
};

template <typename Op> struct RxOpName;
template <> struct RxOpName<OpSum> { static const char* get() { return "cm::OpSum"; } };
template <> struct RxOpName<OpProd> { static const char* get() { return "cm::OpProd"; } };
template <> struct RxOpName<OpMin> { static const char* get() { return "cm::OpMin"; } };
template <> struct RxOpName<OpMax> { static const char* get() { return "cm::OpMax"; } };
template <> struct RxOpName<OpAll> { static const char* get() { return "cm::OpAll"; } };
template <> struct RxOpName<OpAny> { static const char* get() { return "cm::OpAny"; } };
template <> struct RxOpName<OpBand> { static const char* get() { return "cm::OpBand"; } };
template <> struct RxOpName<OpBor> { static const char* get() { return "cm::OpBor"; } };

// The identity of Op in T as a source literal (the values of RxIdentity).
template <typename Op, typename T>
static std::string rx_identity_src()
{
    char buf[64];
    const bool f = std::is_same<T, float>::value;
    if (std::is_same<Op, OpMin>::value)
        snprintf(buf, sizeof(buf), f ? "__uint_as_float(0x7f7fffffu)" : "2147483647");
    else if (std::is_same<Op, OpMax>::value)
        snprintf(buf, sizeof(buf), f ? "__uint_as_float(0xff7fffffu)" : "(-2147483647 - 1)");
    else if (std::is_same<Op, OpProd>::value || std::is_same<Op, OpAll>::value)
        snprintf(buf, sizeof(buf), f ? "1.0f" : "1");
    else if (std::is_same<Op, OpBand>::value)
        snprintf(buf, sizeof(buf), "(~0)");
    else
        snprintf(buf, sizeof(buf), f ? "0.0f" : "0");
    return buf;
}

// Source and key of the generated FLAT reduction (32-bit programs): the
// kernel of reduce_expr_kernel<Op, T, S, false, NL, true> with the program
// inlined, 16-byte loads where a thread's two groups of 4 elements are whole
// and aligned, the same elements per thread, the same per-thread fold order
// and the same block reduction (fold_ops.hpp), so both give the same bits.
template <typename Op, typename T, typename S>
int jit_reduce_flat_source(const CwProgram& P, std::string& key, std::string& body)
{
    if (!std::is_same<T, float>::value && !std::is_same<T, int>::value) return RX_JIT_DECLINED;
    char buf[256];
    const char* tname = std::is_same<T, float>::value ? "float" : "int";
    const char* sname = std::is_same<S, float>::value ? "float" : (std::is_same<S, int>::value ? "int" : "unsigned char");
    snprintf(buf, sizeof(buf), "rxflat32|%s|%s|%s|", RxOpName<Op>::get(), tname, sname);
    key = buf;
    std::string decls, vload, sload, scalars, names[CM_CWISE_MAX_LEAVES];
    const char* namep[CM_CWISE_MAX_LEAVES];
    bool all_al = true;
    for (int l = 0; l < P.nleaves; ++l) {
        const cm_leaf& L = P.leaves[l];
        if (L.dtype != CM_F32 && L.dtype != CM_I32 && L.dtype != CM_BOOL) return RX_JIT_DECLINED;
        const bool scalar = (P.leaf_scalar >> l) & 1;
        const bool al = L.dtype != CM_BOOL && ((uintptr_t)L.ptr & 15) == 0;
        snprintf(buf, sizeof(buf), "L%d%c%c|", L.dtype, scalar ? 's' : 'd', al ? 'a' : 'u');
        key += buf;
        const std::string ls = std::to_string(l);
        const char* rd = L.dtype == CM_BOOL ? "(R)((const unsigned char*)A.leaf[" : "((const uint32_t*)A.leaf[";
        if (scalar) {
            names[l] = "c" + ls;
            scalars += "    const R c" + ls + " = " + rd + ls + "])[0];\n";
        } else {
            names[l] = "v" + ls + "[k]";
            decls += "        R v" + ls + "[8];\n";
            if (!al) all_al = false;
            if (al)
                vload += "            { const u4v a = __builtin_nontemporal_load((const u4v*)((const uint32_t*)A.leaf[" + ls +
                         "] + e0)); const u4v b = __builtin_nontemporal_load((const u4v*)((const uint32_t*)A.leaf[" + ls +
                         "] + e0 + 256));\n              v" + ls + "[0] = a.x; v" + ls + "[1] = a.y; v" + ls + "[2] = a.z; v" +
                         ls + "[3] = a.w; v" + ls + "[4] = b.x; v" + ls + "[5] = b.y; v" + ls + "[6] = b.z; v" + ls +
                         "[7] = b.w; }\n";
            sload += "#pragma unroll\n            for (int k = 0; k < 8; ++k) v" + ls + "[k] = " + rd + ls +
                     "])[s * A.L + (EL(j0, k) < end ? EL(j0, k) : beg)];\n";
        }
        namep[l] = names[l].c_str();
    }
    for (int i = 0; i < P.ncode; ++i) {
        const cm_instr& I = P.code[i];
        uint64_t fb;
        memcpy(&fb, &I.fimm, 8);
        snprintf(buf, sizeof(buf), "%d,%d,%d,%d,%d,%d,%llx,%llx;", I.opcode, I.dtype, I.fn, I.src, I.arg, I.flags,
                 (unsigned long long)fb, (unsigned long long)I.iimm);
        key += buf;
    }
    std::string prog;
    if (!jit::program_body(P, namep, prog)) return RX_JIT_DECLINED;
    const char* get = std::is_same<T, float>::value ? "getf" : "geti";
    body = std::string("typedef unsigned int u4v __attribute__((ext_vector_type(4)));\n") + "typedef " + tname +
           " TT;\ntypedef " + sname + " SS;\ntypedef " + RxOpName<Op>::get() +
           " OP;\n"
           "struct JitRxArgs { const void* leaf[" + std::to_string(CM_CWISE_MAX_LEAVES) +
           "]; void* partials; void* out; long long L, chunk; int nchunk, pad; };\n"
           "#define EL(j0, k) ((j0) + ((k) >> 2) * 256 + ((k) & 3))\n"
           "extern \"C\" __global__ __launch_bounds__(256) void cm_jit_reduce(const JitRxArgs A)\n{\n"
           "    typedef uint32_t R;\n"
           "    __shared__ TT lds[4];\n"
           "    const long long s = blockIdx.x / A.nchunk;\n"
           "    const int c = blockIdx.x % A.nchunk;\n"
           "    const long long beg = c * A.chunk;\n"
           "    const long long end = min(A.L, beg + A.chunk);\n" +
           scalars +
           "    TT acc = " + rx_identity_src<Op, T>() + ";\n"
           "    const long long t0 = (long long)(threadIdx.x >> 6) * 512 + (threadIdx.x & 63) * 4;\n"
           "    for (long long p0 = beg; p0 < end; p0 += 256 * 8) {\n"
           "        const long long j0 = p0 + t0;\n" + decls +
           "        const long long e0 = s * A.L + j0;\n" +
           (all_al ? "        if (j0 + 259 < end && (e0 & 3) == 0) {\n" + vload + "        } else {\n"
                   : std::string("        {\n")) +
           sload +
           "        }\n"
           "#pragma unroll\n"
           "        for (int k = 0; k < 8; ++k) {\n"
           "            R acc_r = 0, t = 0, s0 = 0, s1 = 0, s2 = 0, s3 = 0;\n"
           "            (void)t; (void)s0; (void)s1; (void)s2; (void)s3;\n"
           "#define acc acc_r\n" +
           prog +
           "#undef acc\n"
           "            if (EL(j0, k) < end) acc = OP::apply(acc, (TT)" + get + "(acc_r));\n"
           "        }\n"
           "    }\n"
           "    acc = cm::block_reduce<OP, 256>(acc, lds);\n"
           "    if (threadIdx.x == 0) {\n"
           "        if (A.nchunk == 1) ((SS*)A.out)[s] = (SS)acc;\n"
           "        else ((TT*)A.partials)[blockIdx.x] = acc;\n"
           "    }\n"
           "}\n";
    return CM_OK;
}

// Source and key of the generated reduction for programs with strided or
// broadcast leaves (32-bit programs): the kernel of
// reduce_expr_kernel<Op, T, S, false, NL, false> -- a thread's two groups of
// 4 elements 256 apart, the same fold order and block reduction -- with the
// program inlined. Leaves in the virtual order are loaded as in the flat
// kernel (16-byte loads where a thread's groups are whole and aligned); the
// others at the element's coordinates, with the virtual order's shape and
// their strides compiled in as constants (rx_coords() without run-time
// divisions).
template <typename Op, typename T, typename S>
int jit_reduce_strided_source(const CwProgram& P, const RxShape& SH, std::string& key, std::string& body,
                              bool spread)
{
    if (!std::is_same<T, float>::value && !std::is_same<T, int>::value) return RX_JIT_DECLINED;
    char buf[256];
    const char* tname = std::is_same<T, float>::value ? "float" : "int";
    const char* sname = std::is_same<S, float>::value ? "float" : (std::is_same<S, int>::value ? "int" : "unsigned char");
    snprintf(buf, sizeof(buf), "rxst32%c|%s|%s|%s|k", spread ? 's' : 'c', RxOpName<Op>::get(), tname, sname);
    key = buf;
    const bool narrow = (double)SH.L * (double)SH.nseg < 4294967295.0;
    const char* sfx = narrow ? "u" : "ull";
    const char* ix = narrow ? "unsigned" : "unsigned long long";
    std::string coords = std::string("            ") + ix + " s_ = (" + ix + ")s, j_ = (" + ix +
                         ")(EL(j0, k) < end ? EL(j0, k) : beg);\n"
                         "            long long q3[3] = {0, 0, 0};\n";
    for (int a = 0; a < SH.nk; ++a) {
        snprintf(buf, sizeof(buf), "%lld@%d,", SH.kdim[a], SH.kaxis[a]);
        key += buf;
        snprintf(buf, sizeof(buf), "            { const auto q = s_ / %lld%s; q3[%d] = (long long)(s_ - q * %lld%s); s_ = q; }\n",
                 SH.kdim[a], sfx, SH.kaxis[a], SH.kdim[a], sfx);
        coords += buf;
    }
    key += "r";
    for (int a = 0; a < SH.nr; ++a) {
        snprintf(buf, sizeof(buf), "%lld@%d,", SH.rdim[a], SH.raxis[a]);
        key += buf;
        snprintf(buf, sizeof(buf), "            { const auto q = j_ / %lld%s; q3[%d] = (long long)(j_ - q * %lld%s); j_ = q; }\n",
                 SH.rdim[a], sfx, SH.raxis[a], SH.rdim[a], sfx);
        coords += buf;
    }
    key += "|";
    // element strides of the virtual order (as rx_flat)
    const long long dim[3] = {P.rows, P.cols, P.batches};
    long long es[3] = {0, 0, 0};
    long long stv = 1;
    for (int i = 0; i < SH.nr; ++i) es[SH.raxis[i]] = stv, stv *= SH.rdim[i];
    for (int i = 0; i < SH.nk; ++i) es[SH.kaxis[i]] = stv, stv *= SH.kdim[i];
    std::string decls, vload, sload, xload, scalars, names[CM_CWISE_MAX_LEAVES];
    const char* namep[CM_CWISE_MAX_LEAVES];
    bool all_al = true, any_v = false;
    for (int l = 0; l < P.nleaves; ++l) {
        const cm_leaf& L = P.leaves[l];
        if (L.dtype != CM_F32 && L.dtype != CM_I32 && L.dtype != CM_BOOL) return RX_JIT_DECLINED;
        if (L.kind != CM_LEAF_STRIDED) return RX_JIT_DECLINED;
        const long long ls[3] = {L.sr, L.sc, L.sb};
        bool scalar = true, virt = true;
        for (int a = 0; a < 3; ++a) {
            if (dim[a] <= 1) continue;
            scalar = scalar && ls[a] == 0;
            virt = virt && ls[a] == es[a];
        }
        const bool al = L.dtype != CM_BOOL && ((uintptr_t)L.ptr & 15) == 0;
        snprintf(buf, sizeof(buf), "L%d%c%c%lld,%lld,%lld|", L.dtype, scalar ? 's' : (virt ? 'v' : 'x'), al ? 'a' : 'u',
                 ls[0], ls[1], ls[2]);
        key += buf;
        const std::string lsn = std::to_string(l);
        const char* rd = L.dtype == CM_BOOL ? "(R)((const unsigned char*)A.leaf[" : "((const uint32_t*)A.leaf[";
        if (scalar) {
            names[l] = "c" + lsn;
            scalars += "    const R c" + lsn + " = " + rd + lsn + "])[0];\n";
        } else if (virt) {
            any_v = true;
            names[l] = "v" + lsn + "[k]";
            decls += "        R v" + lsn + "[8];\n";
            if (!al) all_al = false;
            vload += "            { const u4v a = __builtin_nontemporal_load((const u4v*)((const uint32_t*)A.leaf[" + lsn +
                     "] + e0)); const u4v b = __builtin_nontemporal_load((const u4v*)((const uint32_t*)A.leaf[" + lsn +
                     "] + e0 + 256));\n              v" + lsn + "[0] = a.x; v" + lsn + "[1] = a.y; v" + lsn + "[2] = a.z; v" +
                     lsn + "[3] = a.w; v" + lsn + "[4] = b.x; v" + lsn + "[5] = b.y; v" + lsn + "[6] = b.z; v" + lsn +
                     "[7] = b.w; }\n";
            sload += "#pragma unroll\n            for (int k = 0; k < 8; ++k) v" + lsn + "[k] = " + rd + lsn +
                     "])[s * A.L + (EL(j0, k) < end ? EL(j0, k) : beg)];\n";
        } else {
            names[l] = "v" + lsn + "[k]";
            decls += "        R v" + lsn + "[8];\n";
            std::string off;
            for (int a = 0; a < 3; ++a)
                if (ls[a] != 0 && dim[a] > 1) {
                    snprintf(buf, sizeof(buf), "%sq3[%d] * %lldll", off.empty() ? "" : " + ", a, ls[a]);
                    off += buf;
                }
            if (off.empty()) off = "0";
            xload += "            v" + lsn + "[k] = " + rd + lsn + "])[" + off + "];\n";
        }
        namep[l] = names[l].c_str();
    }
    for (int i = 0; i < P.ncode; ++i) {
        const cm_instr& I = P.code[i];
        uint64_t fb;
        memcpy(&fb, &I.fimm, 8);
        snprintf(buf, sizeof(buf), "%d,%d,%d,%d,%d,%d,%llx,%llx;", I.opcode, I.dtype, I.fn, I.src, I.arg, I.flags,
                 (unsigned long long)fb, (unsigned long long)I.iimm);
        key += buf;
    }
    std::string prog;
    if (!jit::program_body(P, namep, prog)) return RX_JIT_DECLINED;
    const char* get = std::is_same<T, float>::value ? "getf" : "geti";
    std::string vpart;
    if (any_v)
        vpart = std::string("        const long long e0 = s * A.L + j0;\n") +
                (all_al && spread ? "        if (j0 + 259 < end && (e0 & 3) == 0) {\n" + vload + "        } else {\n"
                                  : std::string("        {\n")) +
                sload + "        }\n";
    std::string xpart;
    if (!xload.empty())
        xpart = "#pragma unroll\n"
                "        for (int k = 0; k < 8; ++k) {\n" + coords + xload + "        }\n";
    body = std::string("typedef unsigned int u4v __attribute__((ext_vector_type(4)));\n") + "typedef " + tname +
           " TT;\ntypedef " + sname + " SS;\ntypedef " + RxOpName<Op>::get() +
           " OP;\n"
           "struct JitRxArgs { const void* leaf[" + std::to_string(CM_CWISE_MAX_LEAVES) +
           "]; void* partials; void* out; long long L, chunk; int nchunk, pad; };\n"
           + (spread ? "#define EL(j0, k) ((j0) + ((k) >> 2) * 256 + ((k) & 3))\n" : "#define EL(j0, k) ((j0) + (k))\n") +
           "extern \"C\" __global__ __launch_bounds__(256) void cm_jit_reduce(const JitRxArgs A)\n{\n"
           "    typedef uint32_t R;\n"
           "    __shared__ TT lds[4];\n"
           "    const long long s = blockIdx.x / A.nchunk;\n"
           "    const int c = blockIdx.x % A.nchunk;\n"
           "    const long long beg = c * A.chunk;\n"
           "    const long long end = min(A.L, beg + A.chunk);\n" +
           scalars +
           "    TT acc = " + rx_identity_src<Op, T>() + ";\n" +
           (spread ? "    const long long t0 = (long long)(threadIdx.x >> 6) * 512 + (threadIdx.x & 63) * 4;\n"
                   : "    const long long t0 = (long long)threadIdx.x * 8;\n") +
           "    for (long long p0 = beg; p0 < end; p0 += 256 * 8) {\n"
           "        const long long j0 = p0 + t0;\n" + (spread ? "" : "        if (j0 >= end) break;\n") + decls +
           vpart + xpart +
           "#pragma unroll\n"
           "        for (int k = 0; k < 8; ++k) {\n"
           "            R acc_r = 0, t = 0, s0 = 0, s1 = 0, s2 = 0, s3 = 0;\n"
           "            (void)t; (void)s0; (void)s1; (void)s2; (void)s3;\n"
           "#define acc acc_r\n" +
           prog +
           "#undef acc\n"
           "            if (EL(j0, k) < end) acc = OP::apply(acc, (TT)" + get + "(acc_r));\n"
           "        }\n"
           "    }\n"
           "    acc = cm::block_reduce<OP, 256>(acc, lds);\n"
           "    if (threadIdx.x == 0) {\n"
           "        if (A.nchunk == 1) ((SS*)A.out)[s] = (SS)acc;\n"
           "        else ((TT*)A.partials)[blockIdx.x] = acc;\n"
           "    }\n"
           "}\n";
    return CM_OK;
}

// Source and key of the generated reduction of a 64-bit program (FP64
// leaves or instructions): the kernel of reduce_expr_kernel<Op, T, S, true,
// NL, FLAT> -- 4 consecutive elements of the virtual order per thread and
// pass in 64-bit registers, the same fold order and block reduction -- with
// the program inlined; leaves in the virtual order are read at s*L + j, the
// others at the element's coordinates (shape and strides compiled in).
template <typename Op, typename T, typename S>
int jit_reduce_wide_source(const CwProgram& P, const RxShape& SH, std::string& key, std::string& body)
{
    char buf[256];
    const char* tname = std::is_same<T, float>::value ? "float" : (std::is_same<T, int>::value ? "int" : "double");
    const char* sname = std::is_same<S, float>::value
                            ? "float"
                            : (std::is_same<S, int>::value ? "int" : (std::is_same<S, double>::value ? "double" : "unsigned char"));
    snprintf(buf, sizeof(buf), "rxw64|%s|%s|%s|k", RxOpName<Op>::get(), tname, sname);
    key = buf;
    const bool narrow = (double)SH.L * (double)SH.nseg < 4294967295.0;
    const char* sfx = narrow ? "u" : "ull";
    const char* ix = narrow ? "unsigned" : "unsigned long long";
    std::string coords = std::string("            ") + ix + " s_ = (" + ix + ")s, j_ = (" + ix +
                         ")(j0 + k < end ? j0 + k : beg);\n"
                         "            long long q3[3] = {0, 0, 0};\n";
    for (int a = 0; a < SH.nk; ++a) {
        snprintf(buf, sizeof(buf), "%lld@%d,", SH.kdim[a], SH.kaxis[a]);
        key += buf;
        snprintf(buf, sizeof(buf), "            { const auto q = s_ / %lld%s; q3[%d] = (long long)(s_ - q * %lld%s); s_ = q; }\n",
                 SH.kdim[a], sfx, SH.kaxis[a], SH.kdim[a], sfx);
        coords += buf;
    }
    key += "r";
    for (int a = 0; a < SH.nr; ++a) {
        snprintf(buf, sizeof(buf), "%lld@%d,", SH.rdim[a], SH.raxis[a]);
        key += buf;
        snprintf(buf, sizeof(buf), "            { const auto q = j_ / %lld%s; q3[%d] = (long long)(j_ - q * %lld%s); j_ = q; }\n",
                 SH.rdim[a], sfx, SH.raxis[a], SH.rdim[a], sfx);
        coords += buf;
    }
    key += "|";
    const long long dim[3] = {P.rows, P.cols, P.batches};
    long long es[3] = {0, 0, 0};
    long long stv = 1;
    for (int i = 0; i < SH.nr; ++i) es[SH.raxis[i]] = stv, stv *= SH.rdim[i];
    for (int i = 0; i < SH.nk; ++i) es[SH.kaxis[i]] = stv, stv *= SH.kdim[i];
    std::string decls, loads, scalars, names[CM_CWISE_MAX_LEAVES];
    const char* namep[CM_CWISE_MAX_LEAVES];
    bool any_x = false;
    for (int l = 0; l < P.nleaves; ++l) {
        const cm_leaf& L = P.leaves[l];
        if (L.dtype != CM_F32 && L.dtype != CM_I32 && L.dtype != CM_BOOL && L.dtype != CM_F64) return RX_JIT_DECLINED;
        if (L.kind != CM_LEAF_STRIDED) return RX_JIT_DECLINED;
        const long long ls[3] = {L.sr, L.sc, L.sb};
        bool scalar = true, virt = true;
        for (int a = 0; a < 3; ++a) {
            if (dim[a] <= 1) continue;
            scalar = scalar && ls[a] == 0;
            virt = virt && ls[a] == es[a];
        }
        snprintf(buf, sizeof(buf), "L%d%c%lld,%lld,%lld|", L.dtype, scalar ? 's' : (virt ? 'v' : 'x'), ls[0], ls[1], ls[2]);
        key += buf;
        const std::string lsn = std::to_string(l);
        const char* rd = L.dtype == CM_BOOL ? "(R)((const unsigned char*)A.leaf["
                                            : (L.dtype == CM_F64 ? "(R)((const uint64_t*)A.leaf[" : "(R)((const uint32_t*)A.leaf[");
        if (scalar) {
            names[l] = "c" + lsn;
            scalars += "    const R c" + lsn + " = " + rd + lsn + "])[0];\n";
        } else {
            names[l] = "v" + lsn + "[k]";
            decls += "        R v" + lsn + "[4];\n";
            std::string off;
            if (virt) {
                off = "s * A.L + (j0 + k < end ? j0 + k : beg)";
            } else {
                any_x = true;
                for (int a = 0; a < 3; ++a)
                    if (ls[a] != 0 && dim[a] > 1) {
                        snprintf(buf, sizeof(buf), "%sq3[%d] * %lldll", off.empty() ? "" : " + ", a, ls[a]);
                        off += buf;
                    }
                if (off.empty()) off = "0";
            }
            loads += "            v" + lsn + "[k] = " + rd + lsn + "])[" + off + "];\n";
        }
        namep[l] = names[l].c_str();
    }
    for (int i = 0; i < P.ncode; ++i) {
        const cm_instr& I = P.code[i];
        uint64_t fb;
        memcpy(&fb, &I.fimm, 8);
        snprintf(buf, sizeof(buf), "%d,%d,%d,%d,%d,%d,%llx,%llx;", I.opcode, I.dtype, I.fn, I.src, I.arg, I.flags,
                 (unsigned long long)fb, (unsigned long long)I.iimm);
        key += buf;
    }
    std::string prog;
    if (!jit::program_body(P, namep, prog)) return RX_JIT_DECLINED;
    const char* get = std::is_same<T, float>::value ? "getf" : (std::is_same<T, int>::value ? "geti" : "getd");
    std::string ident;
    if (std::is_same<T, double>::value) {
        if (std::is_same<Op, OpMin>::value) ident = "__longlong_as_double(0x7fefffffffffffffll)";
        else if (std::is_same<Op, OpMax>::value) ident = "__longlong_as_double((long long)0xffefffffffffffffull)";
        else if (std::is_same<Op, OpProd>::value || std::is_same<Op, OpAll>::value) ident = "1.0";
        else ident = "0.0";
    } else {
        ident = rx_identity_src<Op, T>();
    }
    body = std::string("typedef ") + tname + " TT;\ntypedef " + sname + " SS;\ntypedef " + RxOpName<Op>::get() +
           " OP;\n"
           "struct JitRxArgs { const void* leaf[" + std::to_string(CM_CWISE_MAX_LEAVES) +
           "]; void* partials; void* out; long long L, chunk; int nchunk, pad; };\n"
           "extern \"C\" __global__ __launch_bounds__(256) void cm_jit_reduce(const JitRxArgs A)\n{\n"
           "    typedef uint64_t R;\n"
           "    __shared__ TT lds[4];\n"
           "    const long long s = blockIdx.x / A.nchunk;\n"
           "    const int c = blockIdx.x % A.nchunk;\n"
           "    const long long beg = c * A.chunk;\n"
           "    const long long end = min(A.L, beg + A.chunk);\n" +
           scalars +
           "    TT acc = " + ident + ";\n"
           "    const long long t0 = (long long)threadIdx.x * 4;\n"
           "    for (long long p0 = beg; p0 < end; p0 += 256 * 4) {\n"
           "        const long long j0 = p0 + t0;\n"
           "        if (j0 >= end) break;\n" + decls +
           "#pragma unroll\n"
           "        for (int k = 0; k < 4; ++k) {\n" + (any_x ? coords : std::string()) + loads +
           "        }\n"
           "#pragma unroll\n"
           "        for (int k = 0; k < 4; ++k) {\n"
           "            R acc_r = 0, t = 0, s0 = 0, s1 = 0, s2 = 0, s3 = 0;\n"
           "            (void)t; (void)s0; (void)s1; (void)s2; (void)s3;\n"
           "#define acc acc_r\n" +
           prog +
           "#undef acc\n"
           "            if (j0 + k < end) acc = OP::apply(acc, (TT)" + get + "(acc_r));\n"
           "        }\n"
           "    }\n"
           "    acc = cm::block_reduce<OP, 256>(acc, lds);\n"
           "    if (threadIdx.x == 0) {\n"
           "        if (A.nchunk == 1) ((SS*)A.out)[s] = (SS)acc;\n"
           "        else ((TT*)A.partials)[blockIdx.x] = acc;\n"
           "    }\n"
           "}\n";
    return CM_OK;
}

// instances with external linkage, so the generator can be exercised on a
// host without a GPU (source generation and compilation only)
template int jit_reduce_flat_source<OpSum, float, float>(const CwProgram&, std::string&, std::string&);
template int jit_reduce_flat_source<OpAll, int, unsigned char>(const CwProgram&, std::string&, std::string&);
template int jit_reduce_strided_source<OpSum, float, float>(const CwProgram&, const RxShape&, std::string&,
                                                            std::string&, bool);
template int jit_reduce_wide_source<OpSum, double, double>(const CwProgram&, const RxShape&, std::string&,
                                                           std::string&);

template <typename Op, typename T, typename S, bool W, int NL>
static int launch_reduce_expr(cm_ctx_t ctx, const CwProgram& P0, const RxShape& SH, S* out)
{
    CwProgram P = P0;
    const bool flat = rx_flat(P, SH);
    // the spread element mapping pays where a leaf lies in the virtual order
    // (its loads are then 16 bytes wide); a program whose leaves are all
    // strided keeps consecutive elements per thread
    const bool spread = flat || rx_any_virtual(P0, SH);
    constexpr int V = CwVec<W>::v;
    const long long cus = ctx->cu_count > 0 ? ctx->cu_count : 256;
    // chunks of at least one pass of the block; about 8 blocks per CU overall
    long long nchunk = (cus * 8 + SH.nseg - 1) / SH.nseg;
    const long long maxchunks = (SH.L + (long long)RX_BS * V - 1) / ((long long)RX_BS * V);
    if (nchunk > maxchunks) nchunk = maxchunks;
    if (nchunk > 4096) nchunk = 4096;
    if (nchunk < 1) nchunk = 1;
    long long chunk = (SH.L + nchunk - 1) / nchunk;
    chunk = (chunk + V - 1) / V * V;
    nchunk = (SH.L + chunk - 1) / chunk;
    const long long blocks = SH.nseg * nchunk;
    CM_REQUIRE(blocks < (1ll << 31), CM_E_DIM, "cm_reduce_expr: reduction too large");
    void* part = nullptr;
    if (nchunk > 1) CM_TRY(scratch(ctx, 1, sizeof(T) * (size_t)blocks, &part));
    bool done = false;
    {
        if (jit::want(SH.L * SH.nseg)) {
            // everything the generators read (the front cache's signature)
            std::string sig(W ? "rxw" : "rx");
            sig += RxOpName<Op>::get();
            jit::sig_add(sig, (int)sizeof(T));
            jit::sig_add(sig, (int)std::is_same<T, float>::value);
            jit::sig_add(sig, (int)sizeof(S));
            jit::sig_add(sig, (int)std::is_same<S, float>::value);
            jit::sig_add(sig, (int)flat);
            jit::sig_add(sig, (int)spread);
            jit::sig_add(sig, SH.nk);
            jit::sig_add(sig, SH.nr);
            for (int a = 0; a < 3; ++a) {
                jit::sig_add(sig, a < SH.nk ? SH.kdim[a] : 0ll);
                jit::sig_add(sig, a < SH.nk ? SH.kaxis[a] : -1);
                jit::sig_add(sig, a < SH.nr ? SH.rdim[a] : 0ll);
                jit::sig_add(sig, a < SH.nr ? SH.raxis[a] : -1);
            }
            jit::sig_add(sig, (int)((double)SH.L * (double)SH.nseg < 4294967295.0));
            jit::sig_add(sig, P.rows);
            jit::sig_add(sig, P.cols);
            jit::sig_add(sig, P.batches);
            jit::sig_add(sig, P.leaf_scalar);
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
            int gen = CM_OK;
            if (jit::find_sig(sig, &k) != 0) {
                std::string key, body;
                if constexpr (W)
                    gen = jit_reduce_wide_source<Op, T, S>(P, SH, key, body);
                else
                    gen = flat ? jit_reduce_flat_source<Op, T, S>(P, key, body)
                               : jit_reduce_strided_source<Op, T, S>(P, SH, key, body, spread);
                if (gen == CM_OK) gen = jit::get(key, body, "cm_jit_reduce", &k) == 0 ? CM_OK : RX_JIT_DECLINED;
                if (gen == CM_OK) jit::put_sig(sig, k);
            }
            if (gen == CM_OK) {
                JitRxArgs A;
                memset(&A, 0, sizeof(A));
                for (int l = 0; l < P.nleaves; ++l) A.leaf[l] = P.leaves[l].ptr;
                A.partials = part;
                A.out = out;
                A.L = SH.L;
                A.chunk = chunk;
                A.nchunk = (int)nchunk;
                if (jit::launch(k, (unsigned)blocks, RX_BS, ctx->stream, &A, sizeof(A)) != 0)
                    return fail(CM_E_HIP, "cm_reduce_expr: launch of the generated kernel failed");
                count(CM_CNT_KERNEL_LAUNCH);
                done = true;
            }
        }
    }
    if (!done) {
        if (flat)
            reduce_expr_kernel<Op, T, S, W, NL, true><<<(unsigned)blocks, RX_BS, 0, ctx->stream>>>(
                P, SH, (int)nchunk, chunk, (T*)part, out, 1);
        else
            reduce_expr_kernel<Op, T, S, W, NL, false><<<(unsigned)blocks, RX_BS, 0, ctx->stream>>>(
                P, SH, (int)nchunk, chunk, (T*)part, out, spread ? 1 : 0);
        CM_LAUNCH_CHECK();
    }
    if (nchunk > 1) {
        reduce_expr_combine_kernel<Op, T, S><<<(unsigned)SH.nseg, RX_BS, 0, ctx->stream>>>((const T*)part,
                                                                                          (int)nchunk, out);
        CM_LAUNCH_CHECK();
    }
    return CM_OK;
}

// T is the accumulation type, S the stored result (unsigned char for bool
// results, whose 0 / 1 values fold in int as the stored-bool reduction folds
// them in unsigned char: sums wrap modulo 256 in both)
template <typename T, typename S, bool W, int NL>
static int rx_dispatch_op(cm_ctx_t ctx, int op, const CwProgram& P, const RxShape& SH, void* out)
{
    switch (op) {
    case CM_RED_SUM: return launch_reduce_expr<OpSum, T, S, W, NL>(ctx, P, SH, (S*)out);
    case CM_RED_PROD: return launch_reduce_expr<OpProd, T, S, W, NL>(ctx, P, SH, (S*)out);
    case CM_RED_MIN: return launch_reduce_expr<OpMin, T, S, W, NL>(ctx, P, SH, (S*)out);
    case CM_RED_MAX: return launch_reduce_expr<OpMax, T, S, W, NL>(ctx, P, SH, (S*)out);
    case CM_RED_ALL: return launch_reduce_expr<OpAll, T, S, W, NL>(ctx, P, SH, (S*)out);
    case CM_RED_ANY: return launch_reduce_expr<OpAny, T, S, W, NL>(ctx, P, SH, (S*)out);
    default: break;
    }
    if constexpr (std::is_same<T, int>::value) {
        switch (op) {
        case CM_RED_BAND: return launch_reduce_expr<OpBand, int, S, W, NL>(ctx, P, SH, (S*)out);
        case CM_RED_BOR: return launch_reduce_expr<OpBor, int, S, W, NL>(ctx, P, SH, (S*)out);
        default: break;
        }
    }
    return fail(CM_E_UNSUPPORTED, "cm_reduce_expr: bitwise reduction of a floating expression");
}

}  // namespace cm

extern "C" {

int cm_reduce_expr(cm_ctx_t ctx, const cm_leaf* leaves, int nleaves, const cm_instr* code, int ncode, int dtype,
                   int op, int axis, int rows, int cols, int batches, int out_row_major, void* out)
{
    CM_API_BEGIN
    CM_TRY(cm::resolve(ctx, &ctx));
    CM_REQUIRE(rows >= 1 && cols >= 1 && batches >= 1, CM_E_DIM, "cm_reduce_expr: bad sizes %d x %d x %d", rows,
               cols, batches);
    CM_REQUIRE(axis >= 0 && axis <= 7, CM_E_ARG, "cm_reduce_expr: axis %d out of range", axis);
    CM_REQUIRE(op >= CM_RED_SUM && op <= CM_RED_BOR, CM_E_ARG, "cm_reduce_expr: op %d out of range", op);
    CM_REQUIRE(dtype == CM_F32 || dtype == CM_F64 || dtype == CM_I32 || dtype == CM_BOOL, CM_E_UNSUPPORTED,
               "cm_reduce_expr: dtype %d (F32, F64, I32 or BOOL)", dtype);
    CM_REQUIRE(out != nullptr && code != nullptr && (leaves != nullptr || nleaves == 0), CM_E_ARG,
               "cm_reduce_expr: NULL pointer");
    cm::CwProgram P;
    bool wide = false;
    CM_TRY(cm::build_program(leaves, nleaves, code, ncode, P, wide));
    // an instruction's dtype is its operands' type; logical operators and
    // comparisons yield bool (0 / 1 in the low 32 bits of the lane)
    const cm_instr& last = code[ncode - 1];
    const int yields = (last.opcode == CM_OP_BINARY && last.fn >= CM_BIN_LAND) ? (int)CM_BOOL : last.dtype;
    CM_REQUIRE(yields == dtype, CM_E_ARG, "cm_reduce_expr: the program yields dtype %d, the reduction is %d", yields,
               dtype);
    if (dtype == CM_F64) wide = true;
    for (int l = 0; l < nleaves; ++l)
        CM_REQUIRE(leaves[l].sr != -1, CM_E_ARG, "cm_reduce_expr: leaf %d reads sparse values directly", l);
    cm::count(CM_CNT_EVAL_REDUCTION);
    cm::count(CM_CNT_EVAL_ANY);
    P.rows = rows;
    P.cols = cols;
    P.batches = batches;
    cm::RxShape SH;
    memset(&SH, 0, sizeof(SH));
    const long long dim[3] = {rows, cols, batches};
    const int red_bit[3] = {CM_AXIS_ROW, CM_AXIS_COL, CM_AXIS_BATCH};
    // kept axes in the output's order (column major: row, col, batch; row
    // major: col, row, batch); reduced axes row, col, batch
    const int korder[3] = {out_row_major ? 1 : 0, out_row_major ? 0 : 1, 2};
    SH.L = 1;
    SH.nseg = 1;
    for (int a = 0; a < 3; ++a) {
        const int ax = korder[a];
        if (!(axis & red_bit[ax])) {
            SH.kdim[SH.nk] = dim[ax];
            SH.kaxis[SH.nk++] = ax;
            SH.nseg *= dim[ax];
        }
        if (axis & red_bit[a]) {
            SH.rdim[SH.nr] = dim[a];
            SH.raxis[SH.nr++] = a;
            SH.L *= dim[a];
        }
    }
    const bool few = nleaves <= 4;
    if (dtype == CM_F64)
        return few ? cm::rx_dispatch_op<double, double, true, 4>(ctx, op, P, SH, out)
                   : cm::rx_dispatch_op<double, double, true, CM_CWISE_MAX_LEAVES>(ctx, op, P, SH, out);
    if (dtype == CM_F32 && wide)
        return few ? cm::rx_dispatch_op<float, float, true, 4>(ctx, op, P, SH, out)
                   : cm::rx_dispatch_op<float, float, true, CM_CWISE_MAX_LEAVES>(ctx, op, P, SH, out);
    if (dtype == CM_F32)
        return few ? cm::rx_dispatch_op<float, float, false, 4>(ctx, op, P, SH, out)
                   : cm::rx_dispatch_op<float, float, false, CM_CWISE_MAX_LEAVES>(ctx, op, P, SH, out);
    // int32 and bool results: a program with 64-bit values keeps them in the
    // low half of its 64-bit lanes
    if (dtype == CM_BOOL && wide)
        return few ? cm::rx_dispatch_op<int, unsigned char, true, 4>(ctx, op, P, SH, out)
                   : cm::rx_dispatch_op<int, unsigned char, true, CM_CWISE_MAX_LEAVES>(ctx, op, P, SH, out);
    if (dtype == CM_BOOL)
        return few ? cm::rx_dispatch_op<int, unsigned char, false, 4>(ctx, op, P, SH, out)
                   : cm::rx_dispatch_op<int, unsigned char, false, CM_CWISE_MAX_LEAVES>(ctx, op, P, SH, out);
    if (wide)
        return few ? cm::rx_dispatch_op<int, int, true, 4>(ctx, op, P, SH, out)
                   : cm::rx_dispatch_op<int, int, true, CM_CWISE_MAX_LEAVES>(ctx, op, P, SH, out);
    return few ? cm::rx_dispatch_op<int, int, false, 4>(ctx, op, P, SH, out)
               : cm::rx_dispatch_op<int, int, false, CM_CWISE_MAX_LEAVES>(ctx, op, P, SH, out);
    CM_API_END
}

}  // extern "C"
