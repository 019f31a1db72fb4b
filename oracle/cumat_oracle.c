/*
 * cumat_oracle.c -- CPU restatement of the reference algorithms on the cuMat
 * hot path (see cumat_oracle.h for the per-function reference citations).
 *
 * TEST INFRASTRUCTURE ONLY. Built into oracle/liboracle.so and loaded by the
 * tests, by the smoke check and by the CPU-baseline leg of bench.py. The
 * product library never links or calls it.
 *
 * Built with -O2 and without -ffast-math, so every floating operation is the
 * IEEE operation written here.
 */
#include "cumat_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

enum { M_ASSIGN = 0, M_ADD, M_SUB, M_MUL, M_DIV, M_MOD, M_AND, M_OR };
enum { R_SUM = 0, R_PROD, R_MIN, R_MAX, R_ALL, R_ANY, R_BAND, R_BOR };
enum {
    B_ADD = 0, B_SUB, B_MUL, B_DIV, B_MOD, B_POW, B_MIN, B_MAX, B_AND, B_OR, B_XOR,
    B_LAND, B_LOR, B_EQ, B_NE, B_LT, B_GT, B_LE, B_GE
};
enum {
    U_NEG = 0, U_ABS, U_ABS2, U_INVERSE, U_INVERSE_CHECK, U_SQRT, U_RSQRT, U_CBRT,
    U_EXP, U_LOG, U_LOG1P, U_LOG10, U_SIN, U_COS, U_TAN, U_ASIN, U_ACOS, U_ATAN,
    U_SINH, U_COSH, U_TANH, U_FLOOR, U_CEIL, U_ROUND, U_NOT, U_BITNOT, U_SIGN
};

/* ------------------------------------------------------------------ */
/* problem generator                                                   */
/* ------------------------------------------------------------------ */
int64_t orc_poisson2d(int g, double shift, int32_t* rowptr, int32_t* colind, double* vals)
{
    /* benchmarks/conjugate_gradient/Implementation_cuMat.cu:29-41 */
    int64_t k = 0;
    for (int x = 0; x < g; ++x) {
        for (int y = 0; y < g; ++y) {
            int row = y + x * g;
            rowptr[row] = (int32_t)k;
            if (x > 0)     { colind[k] = y + (x - 1) * g; vals[k] = -1.0; ++k; }
            if (y > 0)     { colind[k] = (y - 1) + x * g; vals[k] = -1.0; ++k; }
            colind[k] = row; vals[k] = 4.0 + shift; ++k;
            if (y < g - 1) { colind[k] = (y + 1) + x * g; vals[k] = -1.0; ++k; }
            if (x < g - 1) { colind[k] = y + (x + 1) * g; vals[k] = -1.0; ++k; }
        }
    }
    rowptr[g * g] = (int32_t)k;
    return k;
}

/* ------------------------------------------------------------------ */
/* assignment modes: CwiseAssignmentHandler, cuMat/src/CwiseOp.h:15-62 */
/* ------------------------------------------------------------------ */
static float assign_f32(int mode, float old, float v)
{
    switch (mode) {
    case M_ADD: return old + v;
    case M_SUB: return old - v;
    case M_MUL: return old * v;
    case M_DIV: return old / v;
    default: return v;
    }
}
static double assign_f64(int mode, double old, double v)
{
    switch (mode) {
    case M_ADD: return old + v;
    case M_SUB: return old - v;
    case M_MUL: return old * v;
    case M_DIV: return old / v;
    default: return v;
    }
}
static int32_t assign_i32(int mode, int32_t old, int32_t v)
{
    uint32_t uo = (uint32_t)old, uv = (uint32_t)v;
    switch (mode) {
    case M_ADD: return (int32_t)(uo + uv);
    case M_SUB: return (int32_t)(uo - uv);
    case M_MUL: return (int32_t)(uo * uv);
    case M_DIV: return old / v;
    case M_MOD: return old % v;
    case M_AND: return old & v;
    case M_OR:  return old | v;
    default: return v;
    }
}

void orc_assign(int dtype, int mode, int64_t n, void* dst, const void* src)
{
    for (int64_t i = 0; i < n; ++i) {
        if (dtype == ORC_F32) ((float*)dst)[i] = assign_f32(mode, ((float*)dst)[i], ((const float*)src)[i]);
        else if (dtype == ORC_F64) ((double*)dst)[i] = assign_f64(mode, ((double*)dst)[i], ((const double*)src)[i]);
        else if (dtype == ORC_I32) ((int32_t*)dst)[i] = assign_i32(mode, ((int32_t*)dst)[i], ((const int32_t*)src)[i]);
        else {
            uint8_t o = ((uint8_t*)dst)[i], v = ((const uint8_t*)src)[i];
            uint8_t r = v;
            if (mode == M_AND) r = o & v; else if (mode == M_OR) r = o | v;
            ((uint8_t*)dst)[i] = r;
        }
    }
}

/* ------------------------------------------------------------------ */
/* CSR SpMV, cuMat/src/SparseProductEvaluation.h:43-88                 */
/* ------------------------------------------------------------------ */
void orc_csr_spmv_f32(int rows, int cols, int nnz, int batches,
                      const int32_t* JA, const int32_t* IA,
                      const float* vals, int vals_batched,
                      const float* x, int x_batched, float* y, int mode)
{
    for (int b = 0; b < batches; ++b) {
        const float* v = vals + (vals_batched ? (int64_t)b * nnz : 0);
        const float* xb = x + (x_batched ? (int64_t)b * cols : 0);
        for (int r = 0; r < rows; ++r) {
            float acc = 0.0f;
            int s = JA[r], e = JA[r + 1];
            if (s < e) {
                acc = v[s] * xb[IA[s]];
                for (int k = s + 1; k < e; ++k) acc += v[k] * xb[IA[k]];
            }
            float* yy = &y[r + (int64_t)rows * b];
            *yy = assign_f32(mode, *yy, acc);
        }
    }
}

void orc_csr_spmv_f64(int rows, int cols, int nnz, int batches,
                      const int32_t* JA, const int32_t* IA,
                      const double* vals, int vals_batched,
                      const double* x, int x_batched, double* y, int mode)
{
    for (int b = 0; b < batches; ++b) {
        const double* v = vals + (vals_batched ? (int64_t)b * nnz : 0);
        const double* xb = x + (x_batched ? (int64_t)b * cols : 0);
        for (int r = 0; r < rows; ++r) {
            double acc = 0.0;
            int s = JA[r], e = JA[r + 1];
            if (s < e) {
                acc = v[s] * xb[IA[s]];
                for (int k = s + 1; k < e; ++k) acc += v[k] * xb[IA[k]];
            }
            double* yy = &y[r + (int64_t)rows * b];
            *yy = assign_f64(mode, *yy, acc);
        }
    }
}

void orc_csr_spmv_i32(int rows, int cols, int nnz, int batches,
                      const int32_t* JA, const int32_t* IA,
                      const int32_t* vals, int vals_batched,
                      const int32_t* x, int x_batched, int32_t* y, int mode)
{
    for (int b = 0; b < batches; ++b) {
        const int32_t* v = vals + (vals_batched ? (int64_t)b * nnz : 0);
        const int32_t* xb = x + (x_batched ? (int64_t)b * cols : 0);
        for (int r = 0; r < rows; ++r) {
            uint32_t acc = 0;
            for (int k = JA[r]; k < JA[r + 1]; ++k)
                acc += (uint32_t)v[k] * (uint32_t)xb[IA[k]];
            int32_t* yy = &y[r + (int64_t)rows * b];
            *yy = assign_i32(mode, *yy, (int32_t)acc);
        }
    }
}

void orc_bcsr_spmv_f32(int D, int rows, int cols, int nnz, int batches,
                       const int32_t* JA, const int32_t* IA,
                       const float* vals, int vals_batched,
                       const float* x, int x_batched, float* y, int mode)
{
    /* cuMat/src/SparseProductEvaluation.h:43-88 over block scalars: the
     * first block's product of a row, then += in CSR order, each product the
     * block ProductElementFunctor (tests/TestBlockedConjugateGradient.cu:
     * 144-157: component q = dot(row q of the block, x_j)) with dot of
     * tests/TestBlockedConjugateGradient.cu:17-20. A block is D float4 rows
     * (4*D floats), a vector entry one float4; components >= D of a product
     * are 0 (make_real3 sets w = 0). An empty block row assigns 0. */
    for (int b = 0; b < batches; ++b) {
        const float* v = vals + (vals_batched ? (int64_t)b * nnz * 4 * D : 0);
        const float* xb = x + (x_batched ? (int64_t)b * cols * 4 : 0);
        for (int r = 0; r < rows; ++r) {
            float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            for (int k = JA[r]; k < JA[r + 1]; ++k) {
                const float* xc = xb + (int64_t)IA[k] * 4;
                for (int q = 0; q < 4; ++q) {
                    float t = 0.0f;
                    if (q < D) {
                        const float* a = v + ((int64_t)k * D + q) * 4;
                        t = a[0] * xc[0] + a[1] * xc[1] + a[2] * xc[2] + a[3] * xc[3];
                    }
                    acc[q] = (k == JA[r]) ? t : acc[q] + t;
                }
            }
            for (int q = 0; q < 4; ++q) {
                float* yy = &y[((int64_t)rows * b + r) * 4 + q];
                *yy = assign_f32(mode, *yy, acc[q]);
            }
        }
    }
}

void orc_bcsr_inverse_diagonal_f32(int D, int rows, int nnz, int batches, const int32_t* JA, const int32_t* IA,
                                   const float* vals, int vals_batched, float* out)
{
    /* A.diagonal().cwiseInverseCheck() over block scalars
     * (cuMat/src/ConjugateGradient.h:218-246 with the ExtractDiagonalFunctor
     * and cwiseInverseCheck of tests/TestBlockedConjugateGradient.cu:174-210):
     * the diagonal block's diagonal (a zero block where none is stored),
     * each component 1/d or 1 for d == 0; components >= D are 0. */
    for (int b = 0; b < (vals_batched ? batches : 1); ++b) {
        const float* v = vals + (vals_batched ? (int64_t)b * nnz * 4 * D : 0);
        for (int r = 0; r < rows; ++r) {
            int kd = -1;
            for (int k = JA[r]; k < JA[r + 1]; ++k) if (IA[k] == r) { kd = k; break; }
            for (int q = 0; q < 4; ++q) {
                float o = 0.0f;
                if (q < D) {
                    float d = (kd < 0) ? 0.0f : v[((int64_t)kd * D + q) * 4 + q];
                    o = (d == 0.0f) ? 1.0f : 1.0f / d;
                }
                out[((int64_t)rows * b + r) * 4 + q] = o;
            }
        }
    }
}

void orc_ell_spmv_f32(int rows, int cols, int nnz_per_row, int batches,
                      const int32_t* indices, const float* vals, int vals_batched,
                      const float* x, int x_batched, float* y, int mode)
{
    /* cuMat/src/SparseProductEvaluation.h:194-226: column-major index matrix,
     * -1 marks padding, which is skipped; the sum starts from 0. */
    for (int b = 0; b < batches; ++b) {
        const float* v = vals + (vals_batched ? (int64_t)b * rows * nnz_per_row : 0);
        const float* xb = x + (x_batched ? (int64_t)b * cols : 0);
        for (int r = 0; r < rows; ++r) {
            float acc = 0.0f;
            for (int k = 0; k < nnz_per_row; ++k) {
                int c = indices[r + (int64_t)rows * k];
                if (c < 0) continue;
                acc += v[r + (int64_t)rows * k] * xb[c];
            }
            float* yy = &y[r + (int64_t)rows * b];
            *yy = assign_f32(mode, *yy, acc);
        }
    }
}

void orc_ell_spmv_f64(int rows, int cols, int nnz_per_row, int batches,
                      const int32_t* indices, const double* vals, int vals_batched,
                      const double* x, int x_batched, double* y, int mode)
{
    /* cuMat/src/SparseProductEvaluation.h:194-226, as orc_ell_spmv_f32 */
    for (int b = 0; b < batches; ++b) {
        const double* v = vals + (vals_batched ? (int64_t)b * rows * nnz_per_row : 0);
        const double* xb = x + (x_batched ? (int64_t)b * cols : 0);
        for (int r = 0; r < rows; ++r) {
            double acc = 0.0;
            for (int k = 0; k < nnz_per_row; ++k) {
                int c = indices[r + (int64_t)rows * k];
                if (c < 0) continue;
                acc += v[r + (int64_t)rows * k] * xb[c];
            }
            double* yy = &y[r + (int64_t)rows * b];
            *yy = assign_f64(mode, *yy, acc);
        }
    }
}

void orc_ell_spmv_i32(int rows, int cols, int nnz_per_row, int batches,
                      const int32_t* indices, const int32_t* vals, int vals_batched,
                      const int32_t* x, int x_batched, int32_t* y, int mode)
{
    /* cuMat/src/SparseProductEvaluation.h:194-226, int32 arithmetic wraps */
    for (int b = 0; b < batches; ++b) {
        const int32_t* v = vals + (vals_batched ? (int64_t)b * rows * nnz_per_row : 0);
        const int32_t* xb = x + (x_batched ? (int64_t)b * cols : 0);
        for (int r = 0; r < rows; ++r) {
            uint32_t acc = 0;
            for (int k = 0; k < nnz_per_row; ++k) {
                int c = indices[r + (int64_t)rows * k];
                if (c < 0) continue;
                acc += (uint32_t)v[r + (int64_t)rows * k] * (uint32_t)xb[c];
            }
            int32_t* yy = &y[r + (int64_t)rows * b];
            *yy = assign_i32(mode, *yy, (int32_t)acc);
        }
    }
}

/* ------------------------------------------------------------------ */
/* reductions                                                          */
/* ------------------------------------------------------------------ */
static int64_t lin_index(int row_major, int R, int C, int r, int c, int b)
{
    /* cuMat/src/Matrix.h:613-627 */
    return row_major ? (c + (int64_t)C * (r + (int64_t)R * b))
                     : (r + (int64_t)R * (c + (int64_t)C * b));
}

void orc_reduce(int dtype, int op, int axis, int rows, int cols, int batches,
                int in_row_major, const void* in, int out_row_major, void* out)
{
    int oR = (axis & 1) ? 1 : rows;
    int oC = (axis & 2) ? 1 : cols;
    int oB = (axis & 4) ? 1 : batches;
    int64_t on = (int64_t)oR * oC * oB;
    long double* accf = (long double*)calloc((size_t)on, sizeof(long double));
    int64_t* acci = (int64_t*)calloc((size_t)on, sizeof(int64_t));
    uint8_t* seen = (uint8_t*)calloc((size_t)on, 1);
    /* initial values: cuMat/src/ReductionOpsPlugin.inl:9-213 */
    for (int64_t o = 0; o < on; ++o) {
        switch (op) {
        case R_SUM: accf[o] = 0; acci[o] = 0; break;
        case R_PROD: accf[o] = 1; acci[o] = 1; break;
        case R_ALL: acci[o] = 1; break;
        case R_ANY: acci[o] = 0; break;
        case R_BAND: acci[o] = (int64_t)(uint32_t)0xffffffffu; break;
        case R_BOR: acci[o] = 0; break;
        default: break;
        }
    }
    for (int b = 0; b < batches; ++b)
        for (int c = 0; c < cols; ++c)
            for (int r = 0; r < rows; ++r) {
                int64_t ii = lin_index(in_row_major, rows, cols, r, c, b);
                int64_t oo = lin_index(out_row_major, oR, oC, (axis & 1) ? 0 : r,
                                       (axis & 2) ? 0 : c, (axis & 4) ? 0 : b);
                if (dtype == ORC_F32 || dtype == ORC_F64) {
                    long double v = (dtype == ORC_F32) ? (long double)((const float*)in)[ii]
                                                       : (long double)((const double*)in)[ii];
                    switch (op) {
                    case R_SUM: accf[oo] += v; break;
                    case R_PROD: accf[oo] *= v; break;
                    case R_MIN:
                        /* CUB_MIN(a,b) = (a<b)?a:b with init max() */
                        if (!seen[oo]) { accf[oo] = v; seen[oo] = 1; }
                        else accf[oo] = (accf[oo] < v) ? accf[oo] : v;
                        break;
                    case R_MAX:
                        if (!seen[oo]) { accf[oo] = v; seen[oo] = 1; }
                        else accf[oo] = (accf[oo] > v) ? accf[oo] : v;
                        break;
                    case R_ALL: acci[oo] = acci[oo] && (v != 0); break;
                    case R_ANY: acci[oo] = acci[oo] || (v != 0); break;
                    default: break;
                    }
                } else {
                    int32_t v = (dtype == ORC_I32) ? ((const int32_t*)in)[ii]
                                                   : (int32_t)((const uint8_t*)in)[ii];
                    uint32_t a = (uint32_t)acci[oo];
                    switch (op) {
                    case R_SUM: acci[oo] = (int64_t)(uint32_t)(a + (uint32_t)v); break;
                    case R_PROD: acci[oo] = (int64_t)(uint32_t)(a * (uint32_t)v); break;
                    case R_MIN:
                        if (!seen[oo]) { acci[oo] = v; seen[oo] = 1; }
                        else acci[oo] = (acci[oo] < v) ? acci[oo] : v;
                        break;
                    case R_MAX:
                        if (!seen[oo]) { acci[oo] = v; seen[oo] = 1; }
                        else acci[oo] = (acci[oo] > v) ? acci[oo] : v;
                        break;
                    case R_ALL: acci[oo] = acci[oo] && (v != 0); break;
                    case R_ANY: acci[oo] = acci[oo] || (v != 0); break;
                    case R_BAND: acci[oo] = (int64_t)(uint32_t)(a & (uint32_t)v); break;
                    case R_BOR: acci[oo] = (int64_t)(uint32_t)(a | (uint32_t)v); break;
                    default: break;
                    }
                }
            }
    for (int64_t o = 0; o < on; ++o) {
        int isint_result = (op == R_ALL || op == R_ANY || op == R_BAND || op == R_BOR);
        if (dtype == ORC_F32) {
            float v;
            if (isint_result) v = (float)acci[o];
            else if ((op == R_MIN || op == R_MAX) && !seen[o]) v = (op == R_MIN) ? INFINITY : -INFINITY;
            else v = (float)accf[o];
            ((float*)out)[o] = v;
        } else if (dtype == ORC_F64) {
            double v;
            if (isint_result) v = (double)acci[o];
            else if ((op == R_MIN || op == R_MAX) && !seen[o]) v = (op == R_MIN) ? INFINITY : -INFINITY;
            else v = (double)accf[o];
            ((double*)out)[o] = v;
        } else if (dtype == ORC_I32) {
            int32_t v = (int32_t)(uint32_t)acci[o];
            if ((op == R_MIN || op == R_MAX) && !seen[o]) v = (op == R_MIN) ? INT32_MAX : INT32_MIN;
            ((int32_t*)out)[o] = v;
        } else {
            ((uint8_t*)out)[o] = (uint8_t)(acci[o] != 0);
        }
    }
    free(accf); free(acci); free(seen);
}

void orc_dot(int dtype, int64_t n, int batches, const void* a, int a_batched,
             const void* b, int b_batched, void* out)
{
    for (int bb = 0; bb < batches; ++bb) {
        int64_t oa = a_batched ? (int64_t)bb * n : 0, ob = b_batched ? (int64_t)bb * n : 0;
        if (dtype == ORC_F32) {
            long double s = 0;
            for (int64_t i = 0; i < n; ++i)
                s += (long double)((const float*)a)[oa + i] * (long double)((const float*)b)[ob + i];
            ((float*)out)[bb] = (float)s;
        } else if (dtype == ORC_F64) {
            long double s = 0;
            for (int64_t i = 0; i < n; ++i)
                s += (long double)((const double*)a)[oa + i] * (long double)((const double*)b)[ob + i];
            ((double*)out)[bb] = (double)s;
        } else {
            uint32_t s = 0;
            for (int64_t i = 0; i < n; ++i)
                s += (uint32_t)((const int32_t*)a)[oa + i] * (uint32_t)((const int32_t*)b)[ob + i];
            ((int32_t*)out)[bb] = (int32_t)s;
        }
    }
}

/* ------------------------------------------------------------------ */
/* component-wise functors                                             */
/* ------------------------------------------------------------------ */
static double bin_f(int fn, double x, double y, int* isbool)
{
    *isbool = 0;
    switch (fn) {
    case B_ADD: return x + y;
    case B_SUB: return x - y;
    case B_MUL: return x * y;
    case B_DIV: return x / y;
    case B_POW: return pow(x, y);
    case B_MIN: return (x < y) ? x : y;
    case B_MAX: return (x > y) ? x : y;
    case B_EQ: *isbool = 1; return x == y;
    case B_NE: *isbool = 1; return x != y;
    case B_LT: *isbool = 1; return x < y;
    case B_GT: *isbool = 1; return x > y;
    case B_LE: *isbool = 1; return x <= y;
    case B_GE: *isbool = 1; return x >= y;
    default: return NAN;
    }
}
static float bin_f32(int fn, float x, float y, int* isbool)
{
    *isbool = 0;
    switch (fn) {
    case B_ADD: return x + y;
    case B_SUB: return x - y;
    case B_MUL: return x * y;
    case B_DIV: return x / y;
    case B_POW: return powf(x, y);
    case B_MIN: return (x < y) ? x : y;
    case B_MAX: return (x > y) ? x : y;
    case B_EQ: *isbool = 1; return x == y;
    case B_NE: *isbool = 1; return x != y;
    case B_LT: *isbool = 1; return x < y;
    case B_GT: *isbool = 1; return x > y;
    case B_LE: *isbool = 1; return x <= y;
    case B_GE: *isbool = 1; return x >= y;
    default: return NAN;
    }
}
static int32_t bin_i32(int fn, int32_t x, int32_t y, int* isbool)
{
    *isbool = 0;
    switch (fn) {
    case B_ADD: return (int32_t)((uint32_t)x + (uint32_t)y);
    case B_SUB: return (int32_t)((uint32_t)x - (uint32_t)y);
    case B_MUL: return (int32_t)((uint32_t)x * (uint32_t)y);
    case B_DIV: return x / y;
    case B_MOD: return x % y;
    case B_MIN: return (x < y) ? x : y;
    case B_MAX: return (x > y) ? x : y;
    case B_AND: return x & y;
    case B_OR: return x | y;
    case B_XOR: return x ^ y;
    case B_LAND: *isbool = 1; return x && y;
    case B_LOR: *isbool = 1; return x || y;
    case B_EQ: *isbool = 1; return x == y;
    case B_NE: *isbool = 1; return x != y;
    case B_LT: *isbool = 1; return x < y;
    case B_GT: *isbool = 1; return x > y;
    case B_LE: *isbool = 1; return x <= y;
    case B_GE: *isbool = 1; return x >= y;
    default: return 0;
    }
}

/* out has the operand dtype, except comparisons and logical ops whose
 * result is bool (one byte). */
void orc_cwise_binary(int dtype, int fn, int64_t n, const void* a, const void* b, void* out)
{
    for (int64_t i = 0; i < n; ++i) {
        int isbool = 0;
        if (dtype == ORC_F32) {
            float r = bin_f32(fn, ((const float*)a)[i], ((const float*)b)[i], &isbool);
            if (isbool) ((uint8_t*)out)[i] = (uint8_t)(r != 0); else ((float*)out)[i] = r;
        } else if (dtype == ORC_F64) {
            double r = bin_f(fn, ((const double*)a)[i], ((const double*)b)[i], &isbool);
            if (isbool) ((uint8_t*)out)[i] = (uint8_t)(r != 0); else ((double*)out)[i] = r;
        } else if (dtype == ORC_I32) {
            int32_t r = bin_i32(fn, ((const int32_t*)a)[i], ((const int32_t*)b)[i], &isbool);
            if (isbool) ((uint8_t*)out)[i] = (uint8_t)(r != 0); else ((int32_t*)out)[i] = r;
        } else {
            int32_t r = bin_i32(fn, ((const uint8_t*)a)[i], ((const uint8_t*)b)[i], &isbool);
            ((uint8_t*)out)[i] = (uint8_t)(r != 0);
        }
    }
}

static double un_f64(int fn, double x)
{
    switch (fn) {
    case U_NEG: return -x;
    case U_ABS: return fabs(x);
    case U_ABS2: return x * x;
    case U_INVERSE: return 1.0 / x;
    case U_INVERSE_CHECK: return x == 0.0 ? 1.0 : 1.0 / x;
    case U_SQRT: return sqrt(x);
    case U_RSQRT: return 1.0 / sqrt(x);
    case U_CBRT: return cbrt(x);
    case U_EXP: return exp(x);
    case U_LOG: return log(x);
    case U_LOG1P: return log1p(x);
    case U_LOG10: return log10(x);
    case U_SIN: return sin(x);
    case U_COS: return cos(x);
    case U_TAN: return tan(x);
    case U_ASIN: return asin(x);
    case U_ACOS: return acos(x);
    case U_ATAN: return atan(x);
    case U_SINH: return sinh(x);
    case U_COSH: return cosh(x);
    case U_TANH: return tanh(x);
    case U_FLOOR: return floor(x);
    case U_CEIL: return ceil(x);
    case U_ROUND: return round(x);
    case U_SIGN: return (x > 0) - (x < 0);
    default: return NAN;
    }
}
static float un_f32(int fn, float x)
{
    switch (fn) {
    case U_NEG: return -x;
    case U_ABS: return fabsf(x);
    case U_ABS2: return x * x;
    case U_INVERSE: return 1.0f / x;
    case U_INVERSE_CHECK: return x == 0.0f ? 1.0f : 1.0f / x;
    case U_SQRT: return sqrtf(x);
    case U_RSQRT: return 1.0f / sqrtf(x);
    case U_CBRT: return cbrtf(x);
    case U_EXP: return expf(x);
    case U_LOG: return logf(x);
    case U_LOG1P: return log1pf(x);
    case U_LOG10: return log10f(x);
    case U_SIN: return sinf(x);
    case U_COS: return cosf(x);
    case U_TAN: return tanf(x);
    case U_ASIN: return asinf(x);
    case U_ACOS: return acosf(x);
    case U_ATAN: return atanf(x);
    case U_SINH: return sinhf(x);
    case U_COSH: return coshf(x);
    case U_TANH: return tanhf(x);
    case U_FLOOR: return floorf(x);
    case U_CEIL: return ceilf(x);
    case U_ROUND: return roundf(x);
    case U_SIGN: return (float)((x > 0) - (x < 0));
    default: return NAN;
    }
}
static int32_t un_i32(int fn, int32_t x)
{
    switch (fn) {
    case U_NEG: return (int32_t)(0u - (uint32_t)x);
    case U_ABS: return x < 0 ? (int32_t)(0u - (uint32_t)x) : x;
    case U_ABS2: return (int32_t)((uint32_t)x * (uint32_t)x);
    case U_INVERSE: return 1 / x;
    case U_INVERSE_CHECK: return x == 0 ? 1 : 1 / x;
    case U_FLOOR: case U_CEIL: case U_ROUND: return x;
    case U_NOT: return !x;
    case U_BITNOT: return ~x;
    case U_SIGN: return (x > 0) - (x < 0);
    default: return 0;
    }
}

void orc_cwise_unary(int dtype, int fn, int64_t n, const void* a, void* out)
{
    for (int64_t i = 0; i < n; ++i) {
        if (dtype == ORC_F32) ((float*)out)[i] = un_f32(fn, ((const float*)a)[i]);
        else if (dtype == ORC_F64) ((double*)out)[i] = un_f64(fn, ((const double*)a)[i]);
        else if (dtype == ORC_I32) ((int32_t*)out)[i] = un_i32(fn, ((const int32_t*)a)[i]);
        else ((uint8_t*)out)[i] = (uint8_t)(fn == U_NOT ? !((const uint8_t*)a)[i] : ((const uint8_t*)a)[i]);
    }
}

/* ------------------------------------------------------------------ */
/* Jacobi + PCG                                                        */
/* ------------------------------------------------------------------ */
void orc_csr_inverse_diagonal_f64(int rows, const int32_t* JA, const int32_t* IA,
                                  const double* vals, double* invdiag)
{
    /* SparseMatrix::coeff row search (cuMat/src/SparseMatrixBase.h:258-271),
     * then cwiseInverseCheck (cuMat/src/UnaryOps.h:177) */
    for (int r = 0; r < rows; ++r) {
        double d = 0.0;
        for (int k = JA[r]; k < JA[r + 1]; ++k)
            if (IA[k] == r) { d = vals[k]; break; }
        invdiag[r] = (d == 0.0) ? 1.0 : 1.0 / d;
    }
}

/* The f32 solver keeps all vectors and scalars in float; a dot is summed in
 * double and rounded to float once. */
static float dotf(int n, const float* a, const float* b)
{
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += (double)a[i] * (double)b[i];
    return (float)s;
}
static double dotd(int n, const double* a, const double* b)
{
    long double s = 0.0;
    for (int i = 0; i < n; ++i) s += (long double)a[i] * (long double)b[i];
    return (double)s;
}

int orc_pcg_f32(int n, int nnz, int batches, const int32_t* JA, const int32_t* IA,
                const float* vals, const float* b, float* x, int use_guess,
                double tol_d, int max_iter, double* err)
{
    /* cuMat/src/ConjugateGradient.h:123-206 */
    const float tol = (float)tol_d;
    float* invd = (float*)malloc(sizeof(float) * (size_t)n);
    for (int r = 0; r < n; ++r) {
        float d = 0.0f;
        for (int k = JA[r]; k < JA[r + 1]; ++k) if (IA[k] == r) { d = vals[k]; break; }
        invd[r] = (d == 0.0f) ? 1.0f : 1.0f / d;
    }
    size_t N = (size_t)n * batches;
    float* r = (float*)malloc(sizeof(float) * N);
    float* p = (float*)malloc(sizeof(float) * N);
    float* z = (float*)malloc(sizeof(float) * N);
    float* t = (float*)malloc(sizeof(float) * N);
    float* rhs2 = (float*)malloc(sizeof(float) * batches);
    float* thr = (float*)malloc(sizeof(float) * batches);
    float* res2 = (float*)malloc(sizeof(float) * batches);
    float* absNew = (float*)malloc(sizeof(float) * batches);
    if (!use_guess) memset(x, 0, sizeof(float) * N);
    /* residual = rhs - A*x */
    orc_csr_spmv_f32(n, n, nnz, batches, JA, IA, vals, 0, x, 1, t, M_ASSIGN);
    for (size_t i = 0; i < N; ++i) r[i] = b[i] - t[i];
    int all = 1;
    for (int bb = 0; bb < batches; ++bb) {
        rhs2[bb] = dotf(n, b + (size_t)bb * n, b + (size_t)bb * n);
        all = all && (rhs2[bb] < tol * tol);
    }
    int iters = 0;
    *err = 0.0;
    if (all) { memset(x, 0, sizeof(float) * N); goto done; }
    all = 1;
    for (int bb = 0; bb < batches; ++bb) {
        thr[bb] = tol * tol * rhs2[bb];
        res2[bb] = dotf(n, r + (size_t)bb * n, r + (size_t)bb * n);
        all = all && (res2[bb] < thr[bb]);
    }
    if (all) {
        double m = 0;
        for (int bb = 0; bb < batches; ++bb) { double q = (double)(res2[bb] / rhs2[bb]); if (q > m) m = q; }
        *err = sqrt(m);
        goto done;
    }
    for (size_t i = 0; i < N; ++i) p[i] = invd[i % n] * r[i];
    for (int bb = 0; bb < batches; ++bb) absNew[bb] = dotf(n, r + (size_t)bb * n, p + (size_t)bb * n);
    {
        int i = 0;
        while (i < max_iter) {
            orc_csr_spmv_f32(n, n, nnz, batches, JA, IA, vals, 0, p, 1, t, M_ASSIGN);
            for (int bb = 0; bb < batches; ++bb) {
                float alpha = absNew[bb] / dotf(n, p + (size_t)bb * n, t + (size_t)bb * n);
                for (int k = 0; k < n; ++k) {
                    size_t q = (size_t)bb * n + k;
                    x[q] += alpha * p[q];
                    r[q] -= alpha * t[q];
                }
            }
            all = 1;
            for (int bb = 0; bb < batches; ++bb) {
                res2[bb] = dotf(n, r + (size_t)bb * n, r + (size_t)bb * n);
                all = all && (res2[bb] < thr[bb]);
            }
            if (all) break;
            for (size_t q = 0; q < N; ++q) z[q] = invd[q % n] * r[q];
            for (int bb = 0; bb < batches; ++bb) {
                float absOld = absNew[bb];
                absNew[bb] = dotf(n, r + (size_t)bb * n, z + (size_t)bb * n);
                float beta = absNew[bb] / absOld;
                for (int k = 0; k < n; ++k) {
                    size_t q = (size_t)bb * n + k;
                    p[q] = z[q] + beta * p[q];
                }
            }
            i++;
        }
        double m = 0;
        for (int bb = 0; bb < batches; ++bb) { double q = (double)(res2[bb] / rhs2[bb]); if (q > m) m = q; }
        *err = sqrt(m);
        iters = i;
    }
done:
    free(invd); free(r); free(p); free(z); free(t);
    free(rhs2); free(thr); free(res2); free(absNew);
    return iters;
}

int orc_pcg_f64(int n, int nnz, int batches, const int32_t* JA, const int32_t* IA,
                const double* vals, const double* b, double* x, int use_guess,
                double tol, int max_iter, double* err)
{
    /* cuMat/src/ConjugateGradient.h:123-206 in double */
    double* invd = (double*)malloc(sizeof(double) * (size_t)n);
    orc_csr_inverse_diagonal_f64(n, JA, IA, vals, invd);
    size_t N = (size_t)n * batches;
    double* r = (double*)malloc(sizeof(double) * N);
    double* p = (double*)malloc(sizeof(double) * N);
    double* z = (double*)malloc(sizeof(double) * N);
    double* t = (double*)malloc(sizeof(double) * N);
    double* rhs2 = (double*)malloc(sizeof(double) * batches);
    double* thr = (double*)malloc(sizeof(double) * batches);
    double* res2 = (double*)malloc(sizeof(double) * batches);
    double* absNew = (double*)malloc(sizeof(double) * batches);
    if (!use_guess) memset(x, 0, sizeof(double) * N);
    orc_csr_spmv_f64(n, n, nnz, batches, JA, IA, vals, 0, x, 1, t, M_ASSIGN);
    for (size_t i = 0; i < N; ++i) r[i] = b[i] - t[i];
    int all = 1;
    for (int bb = 0; bb < batches; ++bb) {
        rhs2[bb] = dotd(n, b + (size_t)bb * n, b + (size_t)bb * n);
        all = all && (rhs2[bb] < tol * tol);
    }
    int iters = 0;
    *err = 0.0;
    if (all) { memset(x, 0, sizeof(double) * N); goto done; }
    all = 1;
    for (int bb = 0; bb < batches; ++bb) {
        thr[bb] = tol * tol * rhs2[bb];
        res2[bb] = dotd(n, r + (size_t)bb * n, r + (size_t)bb * n);
        all = all && (res2[bb] < thr[bb]);
    }
    if (all) {
        double m = 0;
        for (int bb = 0; bb < batches; ++bb) { double q = res2[bb] / rhs2[bb]; if (q > m) m = q; }
        *err = sqrt(m);
        goto done;
    }
    for (size_t i = 0; i < N; ++i) p[i] = invd[i % n] * r[i];
    for (int bb = 0; bb < batches; ++bb) absNew[bb] = dotd(n, r + (size_t)bb * n, p + (size_t)bb * n);
    {
        int i = 0;
        while (i < max_iter) {
            orc_csr_spmv_f64(n, n, nnz, batches, JA, IA, vals, 0, p, 1, t, M_ASSIGN);
            for (int bb = 0; bb < batches; ++bb) {
                double alpha = absNew[bb] / dotd(n, p + (size_t)bb * n, t + (size_t)bb * n);
                for (int k = 0; k < n; ++k) {
                    size_t q = (size_t)bb * n + k;
                    x[q] += alpha * p[q];
                    r[q] -= alpha * t[q];
                }
            }
            all = 1;
            for (int bb = 0; bb < batches; ++bb) {
                res2[bb] = dotd(n, r + (size_t)bb * n, r + (size_t)bb * n);
                all = all && (res2[bb] < thr[bb]);
            }
            if (all) break;
            for (size_t q = 0; q < N; ++q) z[q] = invd[q % n] * r[q];
            for (int bb = 0; bb < batches; ++bb) {
                double absOld = absNew[bb];
                absNew[bb] = dotd(n, r + (size_t)bb * n, z + (size_t)bb * n);
                double beta = absNew[bb] / absOld;
                for (int k = 0; k < n; ++k) {
                    size_t q = (size_t)bb * n + k;
                    p[q] = z[q] + beta * p[q];
                }
            }
            i++;
        }
        double m = 0;
        for (int bb = 0; bb < batches; ++bb) { double q = res2[bb] / rhs2[bb]; if (q > m) m = q; }
        *err = sqrt(m);
        iters = i;
    }
done:
    free(invd); free(r); free(p); free(z); free(t);
    free(rhs2); free(thr); free(res2); free(absNew);
    return iters;
}

void orc_edge_grad_f64(int nedges, const int32_t* ei, const int32_t* ej,
                       const double* lam, const double* x, double* grad)
{
    /* dA/dtheta_e = (u_i - u_j)(u_i - u_j)^T (boundary: u_i u_i^T), so
     * dL/dtheta_e = -lam^T (dA/dtheta_e) x. */
    for (int e = 0; e < nedges; ++e) {
        int i = ei[e], j = ej[e];
        if (j >= 0) grad[e] = -(lam[i] - lam[j]) * (x[i] - x[j]);
        else grad[e] = -lam[i] * x[i];
    }
}
