// spexpr.hip -- products whose sparse matrix is a component-wise expression.
//
// cuMat evaluates (expr).sparseView<CSR|ELLPACK>(pattern) * x in ONE kernel:
// the SpMV kernel reads each stored value through getSparseCoeff of the
// lazy expression (SparseExpressionOp, cuMat/src/SparseExpressionOp.h:32-72;
// ProductAssignment over it, cuMat/src/SparseProductEvaluation.h:144-187 for
// CSR and :261-324 for ELLPACK). Here the expression is the accumulator
// program of cm_cwise_eval, evaluated inside the product's value loads:
//  * CSR: the wave-chunk product of spmv_wave.hpp (64 rows per wavefront,
//    a lane owns 8 stored entries at stride 64, coalesced index loads): the
//    program runs once per lane over its 8 entries (one value-register lane
//    per entry), each product goes to the wavefront's LDS slice and every
//    row is summed in CSR order by its own lane. No value array is written.
//  * ELLPACK: one thread per row (cuMat's layout), the program over 8 (4 for
//    64-bit programs) consecutive slots of the row at a time, products added
//    in slot order, padding (index < 0) skipped.
// Leaf addressing at stored entry k of row r, column c, batch b: a leaf with
// sr == -1 is a value array of the product's own pattern (direct access,
// SparseMatrix::direct(), cuMat/src/SparseMatrix.h:230-233): element
// k + b*sb; any other leaf is a dense strided leaf at (r, c, b).
#include "cwise_core.hpp"
#include "jit.hpp"
#include "spmv_core.hpp"
#include "spmv_wave.hpp"

namespace cm {

constexpr int SX_BS = 256;

template <typename T> __device__ __forceinline__ T reg_get(const Reg<false>& r, int k);
template <> __device__ __forceinline__ float reg_get<float>(const Reg<false>& r, int k) { return getf(r, k); }
template <> __device__ __forceinline__ int reg_get<int>(const Reg<false>& r, int k) { return geti(r, k); }
template <> __device__ __forceinline__ double reg_get<double>(const Reg<false>&, int) { return 0.0; }
template <typename T> __device__ __forceinline__ T reg_get(const Reg<true>& r, int k);
template <> __device__ __forceinline__ float reg_get<float>(const Reg<true>& r, int k) { return getf(r, k); }
template <> __device__ __forceinline__ int reg_get<int>(const Reg<true>& r, int k) { return geti(r, k); }
template <> __device__ __forceinline__ double reg_get<double>(const Reg<true>& r, int k) { return getd(r, k); }

// Fills value-register lane j of every leaf for the entry (q, direct index
// kd); lanes whose entry is absent (ok == false) hold 0.
template <bool W, int NL>
__device__ __forceinline__ void sx_fetch(const CwProgram& P, Bank<W, NL>& lv, int j, bool ok, const Coords& q,
                                         long long kd)
{
#pragma unroll
    for (int l = 0; l < NL; ++l) {
        Reg<W> t;
        t.v[0] = 0;
        if (l < P.nleaves && ok) {
            const cm_leaf& L = P.leaves[l];
            if (L.sr == -1) {
                load_raw<W>(L, kd + L.sb * q.b, t, 0);
            } else {
                bool z;
                const long long off = leaf_offset(L, q, z);
                load_raw<W>(L, z ? 0 : off, t, 0);
                if (z) t.v[0] = 0;
            }
        }
        lv.k[j][l] = t.v[0];
    }
}

// The same for a program whose leaves are value arrays of this pattern and
// scalars (the common sparse view of same-pattern matrices): no coordinates.
template <bool W, int NL>
__device__ __forceinline__ void sx_fetch_values(const CwProgram& P, Bank<W, NL>& lv, int j, bool ok, long long kd,
                                                long long b)
{
#pragma unroll
    for (int l = 0; l < NL; ++l) {
        Reg<W> t;
        t.v[0] = 0;
        if (l < P.nleaves && ok) {
            const cm_leaf& L = P.leaves[l];
            load_raw<W>(L, (L.sr == -1 ? kd : 0) + L.sb * b, t, 0);
        }
        lv.k[j][l] = t.v[0];
    }
}

// Row of stored entry k within a chunk whose nr + 1 row starts are in rs
// (the largest i with rs[i] <= k; empty rows share their start with the
// next row and are skipped by taking the largest).
__device__ __forceinline__ int sx_row_of(const int* rs, int nr, int k)
{
    int lo = 0, hi = nr;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (rs[mid] <= k) lo = mid;
        else hi = mid;
    }
    return lo;
}

// NEED_ROW: some leaf depends on the row; VALUES: every leaf is a value array
// or a scalar (no coordinates at all)
template <typename T, bool W, int NL, bool NEED_ROW, bool VALUES>
__global__ __launch_bounds__(SX_BS) void csr_expr_spmv_kernel(const CwProgram P, const int* __restrict__ JA,
                                                              const int* __restrict__ IA, int rows, int cols,
                                                              const T* __restrict__ x, int x_batched,
                                                              T* __restrict__ y, int mode)
{
    constexpr int EPT = CwVec<W>::v;  // entries per lane = value-register lanes
    constexpr int WPB = SX_BS / 64;
    __shared__ T prod[WPB][EPT * 64];
    __shared__ int starts[WPB][65];
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    T* slice = prod[wave];
    int* rs = starts[wave];
    const int b = blockIdx.y;
    const T* xb = x + (x_batched ? (long long)b * cols : 0);
    T* yb = y + (long long)b * rows;
    for (ChunkWalk w(div_up_dev(rows, 64), WPB); w.valid(); w.next()) {
        const int r0 = w.cur * 64;
        const int nr = min(64, rows - r0);
        const int s = JA[r0], e = JA[r0 + nr];
        const int rb = lane < nr ? JA[r0 + lane] : 0;
        const int re = lane < nr ? JA[r0 + lane + 1] : 0;
        T acc = T(0);
        if (e - s <= EPT * 64) {
            if (NEED_ROW) {
                if (lane <= nr) rs[lane] = JA[r0 + lane];
                // LDS operations of one wavefront complete in order; the
                // barrier keeps the compiler from moving code across it
                __builtin_amdgcn_wave_barrier();
            }
            int col[EPT];
            Bank<W, NL> lv;
#pragma unroll
            for (int j = 0; j < EPT; ++j) {
                const int k = s + lane + 64 * j;
                const bool ok = k < e;
                col[j] = ok ? IA[k] : 0;
                if constexpr (VALUES) {
                    sx_fetch_values<W, NL>(P, lv, j, ok, k, b);
                } else {
                    Coords q;
                    q.r = (NEED_ROW && ok) ? r0 + sx_row_of(rs, nr, k) : 0;
                    q.c = col[j];
                    q.b = b;
                    sx_fetch<W, NL>(P, lv, j, ok, q, k);
                }
            }
            Reg<W> val;
            run_program<W, NL>(P, lv, val);
            T xv[EPT];
#pragma unroll
            for (int j = 0; j < EPT; ++j) xv[j] = (s + lane + 64 * j < e) ? xb[col[j]] : T(0);
#pragma unroll
            for (int j = 0; j < EPT; ++j) {
                const int k = s + lane + 64 * j;
                if (k < e) slice[k - s] = reg_get<T>(val, j) * xv[j];
            }
            __builtin_amdgcn_wave_barrier();
            const int a = rb - s, bnd = re - s;
            if (lane < nr && a < bnd) {
                acc = slice[a];
                for (int k = a + 1; k < bnd; ++k) acc += slice[k];
            }
            __builtin_amdgcn_wave_barrier();
        } else if (lane < nr) {
            // long rows: each lane walks its own row in CSR order, one entry
            // per pass of the program (value-register lane 0)
            for (int k = rb; k < re; ++k) {
                Bank<W, NL> lv;
                Coords q;
                q.r = r0 + lane;
                q.c = IA[k];
                q.b = b;
                sx_fetch<W, NL>(P, lv, 0, true, q, k);
#pragma unroll
                for (int j = 1; j < EPT; ++j) sx_fetch<W, NL>(P, lv, j, false, q, k);
                Reg<W> val;
                run_program<W, NL>(P, lv, val);
                const T p = reg_get<T>(val, 0) * xb[q.c];
                acc = (k == rb) ? p : acc + p;
            }
        }
        if (lane < nr) {
            T* yy = yb + r0 + lane;
            *yy = (mode == CM_ASSIGN) ? acc : assign_mode<T>(mode, *yy, acc);
        }
    }
}

template <typename T, bool W, int NL>
__global__ __launch_bounds__(SX_BS) void ell_expr_spmv_kernel(const CwProgram P, const int* __restrict__ idx,
                                                              int rows, int cols, int npr,
                                                              const T* __restrict__ x, int x_batched,
                                                              T* __restrict__ y, int mode)
{
    constexpr int EPT = CwVec<W>::v;
    const int r = blockIdx.x * SX_BS + threadIdx.x;
    if (r >= rows) return;  // no barrier below
    const int b = blockIdx.y;
    const T* xb = x + (x_batched ? (long long)b * cols : 0);
    T acc = T(0);
    for (int c0 = 0; c0 < npr; c0 += EPT) {
        int col[EPT];
        Bank<W, NL> lv;
#pragma unroll
        for (int j = 0; j < EPT; ++j) {
            const int ci = c0 + j;
            const long long kd = r + (long long)rows * ci;  // slot (r, ci) of the column-major data
            col[j] = ci < npr ? idx[kd] : -1;
            Coords q;
            q.r = r;
            q.c = col[j] < 0 ? 0 : col[j];
            q.b = b;
            sx_fetch<W, NL>(P, lv, j, col[j] >= 0, q, kd);
        }
        Reg<W> val;
        run_program<W, NL>(P, lv, val);
#pragma unroll
        for (int j = 0; j < EPT; ++j)
            if (col[j] >= 0) acc += reg_get<T>(val, j) * xb[col[j]];  // one expression: fused, as in the plain ELLPACK product
    }
    T* yy = y + r + (long long)rows * b;
    *yy = (mode == CM_ASSIGN) ? acc : assign_mode<T>(mode, *yy, acc);
}

// ------------------------------------------------------------- host side
// Checks shared by both products; returns the program and its width.
static int sx_prepare(const cm_leaf* leaves, int nleaves, const cm_instr* code, int ncode, int dtype, int mode,
                      CwProgram& P, bool& wide)
{
    CM_REQUIRE(dtype == CM_F32 || dtype == CM_F64 || dtype == CM_I32, CM_E_UNSUPPORTED,
               "sparse-expression product: dtype %d (F32, F64 or I32)", dtype);
    CM_REQUIRE(mode == CM_ASSIGN || mode == CM_ADD || mode == CM_SUB, CM_E_ARG,
               "sparse-expression product: mode %d (=, += or -=)", mode);
    CM_REQUIRE(leaves != nullptr || nleaves == 0, CM_E_ARG, "sparse-expression product: NULL leaves");
    CM_REQUIRE(code != nullptr, CM_E_ARG, "sparse-expression product: NULL program");
    CM_TRY(build_program(leaves, nleaves, code, ncode, P, wide));
    CM_REQUIRE(code[ncode - 1].dtype == dtype, CM_E_ARG,
               "sparse-expression product: the program yields dtype %d, the vectors are %d (no implicit casting)",
               code[ncode - 1].dtype, dtype);
    if (dtype == CM_F64) wide = true;
    CM_REQUIRE(!(wide && dtype == CM_I32), CM_E_UNSUPPORTED,
               "sparse-expression product: int32 result of a program with 64-bit values");
    return CM_OK;
}

template <typename T, bool W, int NL>
static int launch_csr_expr(cm_ctx_t ctx, const CwProgram& P, int shape, int rows, int cols, int batches,
                           const int* JA, const int* IA, const void* x, int xb, void* y, int mode)
{
    dim3 grid(walk_grid(div_up(div_up(rows, 64), SX_BS / 64), 2048), batches);
    if (shape == 2)
        csr_expr_spmv_kernel<T, W, NL, false, true><<<grid, SX_BS, 0, ctx->stream>>>(P, JA, IA, rows, cols,
                                                                                      (const T*)x, xb, (T*)y, mode);
    else if (shape == 1)
        csr_expr_spmv_kernel<T, W, NL, true, false><<<grid, SX_BS, 0, ctx->stream>>>(P, JA, IA, rows, cols,
                                                                                      (const T*)x, xb, (T*)y, mode);
    else
        csr_expr_spmv_kernel<T, W, NL, false, false><<<grid, SX_BS, 0, ctx->stream>>>(P, JA, IA, rows, cols,
                                                                                       (const T*)x, xb, (T*)y, mode);
    CM_LAUNCH_CHECK();
    return CM_OK;
}

template <typename T, bool W, int NL>
static int launch_ell_expr(cm_ctx_t ctx, const CwProgram& P, int rows, int cols, int npr, int batches,
                           const int* idx, const void* x, int xb, void* y, int mode)
{
    dim3 grid(div_up(rows, SX_BS), batches);
    ell_expr_spmv_kernel<T, W, NL><<<grid, SX_BS, 0, ctx->stream>>>(P, idx, rows, cols, npr, (const T*)x, xb, (T*)y,
                                                                     mode);
    CM_LAUNCH_CHECK();
    return CM_OK;
}

// dispatch over (result type, register width, leaf count)
template <template <typename, bool, int> class L, typename... A>
static int sx_dispatch(int dtype, bool wide, int nleaves, A&&... a)
{
    const bool few = nleaves <= 4;
    if (dtype == CM_F64)
        return few ? L<double, true, 4>::run(a...) : L<double, true, CM_CWISE_MAX_LEAVES>::run(a...);
    if (dtype == CM_F32 && wide)
        return few ? L<float, true, 4>::run(a...) : L<float, true, CM_CWISE_MAX_LEAVES>::run(a...);
    if (dtype == CM_F32)
        return few ? L<float, false, 4>::run(a...) : L<float, false, CM_CWISE_MAX_LEAVES>::run(a...);
    return few ? L<int, false, 4>::run(a...) : L<int, false, CM_CWISE_MAX_LEAVES>::run(a...);
}
template <typename T, bool W, int NL> struct CsrExprLaunch {
    template <typename... A> static int run(A&&... a) { return launch_csr_expr<T, W, NL>(a...); }
};
template <typename T, bool W, int NL> struct EllExprLaunch {
    template <typename... A> static int run(A&&... a) { return launch_ell_expr<T, W, NL>(a...); }
};


// ------------------------------------------------- generated kernels
// csr_expr_spmv_kernel<T, false, NL, false, true> (32-bit programs whose
// leaves are value arrays of the product's pattern and broadcast scalars) as
// straight-line code for one program: the wave-chunk walk of spmv_wave.hpp
// with the next chunk's indices and value-leaf entries in flight during the
// current chunk's gathers, the program evaluated per stored entry, products
// summed per row in CSR order. Same semantics and summation order as the
// interpreter kernel (long-row chunks included).
constexpr int SX_JIT_DECLINED = 1;

struct JitSxArgs {
    const void* leaf[CM_CWISE_MAX_LEAVES];
    long long sb[CM_CWISE_MAX_LEAVES];
    const int* JA;
    const int* IA;
    const void* x;
    void* y;
    int rows, cols, x_batched, pad;
};

int jit_csr_expr_source(const CwProgram& P, int dtype, int mode, std::string& key, std::string& body)
{
    if (dtype != CM_F32 && dtype != CM_I32) return SX_JIT_DECLINED;
    char buf[512];
    snprintf(buf, sizeof(buf), "csrx32|t%d|m%d|", dtype, mode);
    key = buf;
    std::string members, loads, tail_loads, ptrs, scalars, args_v, args_s, params_v, params_s;
    std::string nm_chunk[CM_CWISE_MAX_LEAVES], nm_tail[CM_CWISE_MAX_LEAVES];
    const char* pc[CM_CWISE_MAX_LEAVES];
    const char* pt[CM_CWISE_MAX_LEAVES];
    for (int l = 0; l < P.nleaves; ++l) {
        const cm_leaf& L = P.leaves[l];
        if (L.dtype != CM_F32 && L.dtype != CM_I32 && L.dtype != CM_BOOL) return SX_JIT_DECLINED;
        const char* et = L.dtype == CM_BOOL ? "unsigned char" : "uint32_t";
        const bool value = L.sr == -1;
        if (!value && !(L.kind == CM_LEAF_STRIDED && L.sr == 0 && L.sc == 0)) return SX_JIT_DECLINED;
        snprintf(buf, sizeof(buf), "L%d%c|", L.dtype, value ? 'v' : 's');
        key += buf;
        const std::string ls = std::to_string(l);
        if (value) {
            members += "    R w" + ls + "[EPT];\n";
            ptrs += std::string("    const ") + et + "* __restrict__ p" + ls + " = (const " + et + "*)A.leaf[" + ls +
                    "] + A.sb[" + ls + "] * b;\n";
            params_v += std::string(", const ") + et + "* __restrict__ p" + ls;
            args_v += ", p" + ls;
            loads += "            C.w" + ls + "[j] = ok ? (R)__builtin_nontemporal_load(p" + ls + " + k) : 0u;\n";
            tail_loads += "                const R t" + ls + " = (R)p" + ls + "[k];\n";
            nm_chunk[l] = "C.w" + ls + "[j]";
            nm_tail[l] = "t" + ls;
        } else {
            scalars += std::string("    const R c") + ls + " = (R)((const " + et + "*)A.leaf[" + ls + "])[A.sb[" + ls +
                       "] * b];\n";
            params_s += ", const R c" + ls;
            args_s += ", c" + ls;
            nm_chunk[l] = "c" + ls;
            nm_tail[l] = "c" + ls;
        }
        pc[l] = nm_chunk[l].c_str();
        pt[l] = nm_tail[l].c_str();
    }
    for (int i = 0; i < P.ncode; ++i) {
        const cm_instr& I = P.code[i];
        uint64_t fb;
        memcpy(&fb, &I.fimm, 8);
        snprintf(buf, sizeof(buf), "%d,%d,%d,%d,%d,%d,%llx,%llx;", I.opcode, I.dtype, I.fn, I.src, I.arg, I.flags,
                 (unsigned long long)fb, (unsigned long long)I.iimm);
        key += buf;
    }
    std::string prog_chunk, prog_tail;
    if (!jit::program_body(P, pc, prog_chunk) || !jit::program_body(P, pt, prog_tail)) return SX_JIT_DECLINED;
    const char* get = dtype == CM_F32 ? "getf" : "geti";
    const char* tt = dtype == CM_F32 ? "float" : "int";
    snprintf(buf, sizeof(buf), "typedef %s TT;\nconstexpr int MODE = %d;\nconstexpr int EPT = 8;\n", tt, mode);
    body = std::string(buf) +
           "typedef uint32_t R;\n"
           "struct JitSxArgs { const void* leaf[" + std::to_string(CM_CWISE_MAX_LEAVES) + "]; long long sb[" +
           std::to_string(CM_CWISE_MAX_LEAVES) +
           "]; const int* JA; const int* IA; const void* x; void* y; int rows, cols, x_batched, pad; };\n"
           "struct Ch {\n    int r0, nr, s, e, rb, re;\n    int col[EPT];\n" + members + "};\n"
           "__device__ __forceinline__ void ld(Ch& C, const int* __restrict__ JA, const int* __restrict__ IA, int rows, int r0" +
           params_v +
           ")\n{\n"
           "    const int lane = threadIdx.x & 63;\n"
           "    C.r0 = r0;\n    C.nr = min(64, rows - r0);\n"
           "    C.s = JA[r0];\n    C.e = JA[r0 + C.nr];\n"
           "    C.rb = lane < C.nr ? JA[r0 + lane] : 0;\n    C.re = lane < C.nr ? JA[r0 + lane + 1] : 0;\n"
           "    if (C.e - C.s <= EPT * 64) {\n"
           "#pragma unroll\n"
           "        for (int j = 0; j < EPT; ++j) {\n"
           "            const int k = C.s + lane + 64 * j;\n"
           "            const bool ok = k < C.e;\n"
           "            C.col[j] = ok ? __builtin_nontemporal_load(IA + k) : 0;\n" +
           loads +
           "        }\n    }\n}\n"
           "__device__ __forceinline__ TT comp(const Ch& C, const int* __restrict__ IA, const TT* __restrict__ xb, TT* slice" +
           params_v + params_s +
           ")\n{\n"
           "    const int lane = threadIdx.x & 63;\n"
           "    TT acc = TT(0);\n"
           "    if (C.e - C.s <= EPT * 64) {\n"
           "        TT xv[EPT];\n"
           "#pragma unroll\n"
           "        for (int j = 0; j < EPT; ++j) xv[j] = (C.s + lane + 64 * j < C.e) ? xb[C.col[j]] : TT(0);\n"
           "#pragma unroll\n"
           "        for (int j = 0; j < EPT; ++j) {\n"
           "            R acc = 0, t = 0, s0 = 0, s1 = 0, s2 = 0, s3 = 0;\n"
           "            (void)t; (void)s0; (void)s1; (void)s2; (void)s3;\n" +
           prog_chunk +
           "            const int k = C.s + lane + 64 * j;\n"
           "            if (k < C.e) slice[k - C.s] = " + get + "(acc) * xv[j];\n"
           "        }\n"
           "        __builtin_amdgcn_wave_barrier();\n"
           "        const int a = C.rb - C.s, bnd = C.re - C.s;\n"
           "        if (lane < C.nr && a < bnd) {\n"
           "            acc = slice[a];\n"
           "            for (int k = a + 1; k < bnd; ++k) acc += slice[k];\n"
           "        }\n"
           "        __builtin_amdgcn_wave_barrier();\n"
           "    } else if (lane < C.nr) {\n"
           "        // long rows: each lane walks its own row in CSR order\n"
           "        for (int k = C.rb; k < C.re; ++k) {\n"
           "            const int c = IA[k];\n"
           "            TT p;\n"
           "            {\n" +
           tail_loads +
           "                R acc = 0, t = 0, s0 = 0, s1 = 0, s2 = 0, s3 = 0;\n"
           "                (void)t; (void)s0; (void)s1; (void)s2; (void)s3;\n" +
           prog_tail +
           "                p = " + get + "(acc) * xb[c];\n"
           "            }\n"
           "            acc = (k == C.rb) ? p : acc + p;\n"
           "        }\n"
           "    }\n"
           "    return acc;\n"
           "}\n"
           "extern \"C\" __global__ __launch_bounds__(256) void cm_jit_csr_expr(const JitSxArgs A)\n{\n"
           "    __shared__ TT lds[4][EPT * 64];\n"
           "    TT* slice = lds[threadIdx.x >> 6];\n"
           "    const int lane = threadIdx.x & 63;\n"
           "    const long long b = blockIdx.y;\n"
           "    const int rows = A.rows;\n"
           "    const TT* __restrict__ xb = (const TT*)A.x + (A.x_batched ? b * A.cols : 0);\n"
           "    TT* __restrict__ yb = (TT*)A.y + b * rows;\n" +
           ptrs + scalars +
           // chunk walk of spmv_wave.hpp: XCD x owns chunks [x*NC/8, (x+1)*NC/8)
           "    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);\n"
           "    const int NC = (rows + 63) >> 6;\n"
           "    const int xcd = blockIdx.x & 7;\n"
           "    const int hi = (int)(((long long)NC * (xcd + 1)) >> 3);\n"
           "    const int step = (gridDim.x >> 3) * 4;\n"
           "    int cur = (int)(((long long)NC * xcd) >> 3) + (blockIdx.x >> 3) * 4 + wave;\n"
           "    if (cur >= hi) return;\n"
           "    Ch cu, nx;\n"
           "    ld(cu, A.JA, A.IA, rows, cur * 64" + args_v + ");\n"
           "    for (;;) {\n"
           "        const int nc = cur + step;\n"
           "        const bool more = nc < hi;\n"
           "        if (more) ld(nx, A.JA, A.IA, rows, nc * 64" + args_v + ");\n"
           "        const TT acc = comp(cu, A.IA, xb, slice" + args_v + args_s + ");\n"
           "        if (lane < cu.nr) {\n"
           "            TT* yy = yb + cu.r0 + lane;\n"
           "            if (MODE == CM_ASSIGN) __builtin_nontemporal_store(acc, yy);\n"
           "            else *yy = assign1<TT>(MODE, *yy, acc);\n"
           "        }\n"
           "        if (!more) break;\n"
           "        cu = nx;\n"
           "        cur = nc;\n"
           "    }\n"
           "}\n";
    return CM_OK;
}

static int jit_csr_expr(cm_ctx_t ctx, const CwProgram& P, int dtype, int rows, int cols, int batches,
                        const int* JA, const int* IA, const void* x, int xb, void* y, int mode)
{
    std::string key, body;
    if (jit_csr_expr_source(P, dtype, mode, key, body) != CM_OK) return SX_JIT_DECLINED;
    jit::Kernel k;
    if (jit::get(key, body, "cm_jit_csr_expr", &k) != 0) return SX_JIT_DECLINED;
    JitSxArgs A;
    memset(&A, 0, sizeof(A));
    for (int l = 0; l < P.nleaves; ++l) {
        A.leaf[l] = P.leaves[l].ptr;
        A.sb[l] = P.leaves[l].sb;
    }
    A.JA = JA;
    A.IA = IA;
    A.x = x;
    A.y = y;
    A.rows = rows;
    A.cols = cols;
    A.x_batched = xb;
    const unsigned grid = (unsigned)walk_grid(div_up(div_up(rows, 64), SX_BS / 64), 2048);
    if (jit::launch(k, grid, SX_BS, ctx->stream, &A, sizeof(A), (unsigned)batches) != 0)
        return fail(CM_E_HIP, "cm_csr_expr_spmv: launch of the generated kernel failed");
    count(CM_CNT_KERNEL_LAUNCH);
    return CM_OK;
}
}  // namespace cm

extern "C" {

int cm_csr_expr_spmv(cm_ctx_t ctx, const cm_leaf* leaves, int nleaves, const cm_instr* code, int ncode, int dtype,
                     int rows, int cols, int nnz, int batches, const int* JA, const int* IA, const void* x,
                     int x_batched, void* y, int mode)
{
    CM_API_BEGIN
    CM_TRY(cm::resolve(ctx, &ctx));
    CM_REQUIRE(rows >= 0 && cols >= 0 && nnz >= 0 && batches >= 1, CM_E_DIM,
               "cm_csr_expr_spmv: bad sizes rows=%d cols=%d nnz=%d batches=%d", rows, cols, nnz, batches);
    cm::CwProgram P;
    bool wide = false;
    CM_TRY(cm::sx_prepare(leaves, nleaves, code, ncode, dtype, mode, P, wide));
    cm::count(CM_CNT_EVAL_MATMUL_SPARSE);
    cm::count(CM_CNT_EVAL_ANY);
    if (rows == 0) return CM_OK;
    CM_REQUIRE(JA && y && (nnz == 0 || (IA && x)), CM_E_ARG, "cm_csr_expr_spmv: NULL pointer");
    P.rows = rows;
    P.cols = cols;
    P.batches = batches;
    // kernel shape: 2 = value arrays and scalars only (no coordinates),
    // 1 = some leaf depends on the row, 0 = leaves depend on the column and
    // batch only
    bool need_row = false, values_only = true;
    for (int l = 0; l < nleaves; ++l) {
        const cm_leaf& L = leaves[l];
        if (L.sr == -1) continue;
        if (L.kind != CM_LEAF_STRIDED || L.sr != 0) need_row = true;
        if (L.kind != CM_LEAF_STRIDED || L.sr != 0 || L.sc != 0) values_only = false;
    }
    const int shape = values_only ? 2 : (need_row ? 1 : 0);
    if (shape == 2 && !wide && cm::jit::want((long long)nnz * batches)) {
        const int s = cm::jit_csr_expr(ctx, P, dtype, rows, cols, batches, JA, IA, x, x_batched, y, mode);
        if (s != cm::SX_JIT_DECLINED) return s;
    }
    return cm::sx_dispatch<cm::CsrExprLaunch>(dtype, wide, nleaves, ctx, P, shape, rows, cols, batches, JA, IA, x,
                                              x_batched, y, mode);
    CM_API_END
}

int cm_ell_expr_spmv(cm_ctx_t ctx, const cm_leaf* leaves, int nleaves, const cm_instr* code, int ncode, int dtype,
                     int rows, int cols, int nnz_per_row, int batches, const int* indices, const void* x,
                     int x_batched, void* y, int mode)
{
    CM_API_BEGIN
    CM_TRY(cm::resolve(ctx, &ctx));
    CM_REQUIRE(rows >= 0 && cols >= 0 && nnz_per_row >= 0 && batches >= 1, CM_E_DIM,
               "cm_ell_expr_spmv: bad sizes rows=%d cols=%d nnz_per_row=%d batches=%d", rows, cols, nnz_per_row,
               batches);
    cm::CwProgram P;
    bool wide = false;
    CM_TRY(cm::sx_prepare(leaves, nleaves, code, ncode, dtype, mode, P, wide));
    cm::count(CM_CNT_EVAL_MATMUL_SPARSE);
    cm::count(CM_CNT_EVAL_ANY);
    if (rows == 0) return CM_OK;
    CM_REQUIRE(y && (nnz_per_row == 0 || (indices && x)), CM_E_ARG, "cm_ell_expr_spmv: NULL pointer");
    P.rows = rows;
    P.cols = cols;
    P.batches = batches;
    return cm::sx_dispatch<cm::EllExprLaunch>(dtype, wide, nleaves, ctx, P, rows, cols, nnz_per_row, batches,
                                              indices, x, x_batched, y, mode);
    CM_API_END
}

}  // extern "C"
