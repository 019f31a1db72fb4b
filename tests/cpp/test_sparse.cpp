// C++ host-API tests of the sparse hot path: CSR and ELLPACK products with
// dense vectors (batched, assignment modes) and component-wise evaluation
// into CSR values. Fixtures follow tests/TestSparseMultOp.cu:70-276 and
// tests/TestSparseMatrix.cu:385-510 of the reference.
#include <cuMat/Sparse>

#include <random>
#include <sstream>

#include "harness.h"

using namespace cuMat;

typedef SparseMatrix<float, 1, SparseFlags::CSR> SMatrix_t;
typedef SparsityPattern<SparseFlags::CSR> SPattern;

// {1, 4, 0, 0, 0},
// {0, 2, 3, 0, 0},
// {5, 0, 0, 7, 8},
// {0, 0, 9, 0, 6}
static SPattern fixturePattern()
{
    SPattern pattern;
    pattern.rows = 4;
    pattern.cols = 5;
    pattern.nnz = 9;
    pattern.IA = SPattern::IndexVector::fromHost(std::vector<int>{0, 1, 1, 2, 0, 3, 4, 2, 4});
    pattern.JA = SPattern::IndexVector::fromHost(std::vector<int>{0, 2, 4, 7, 9});
    return pattern;
}

template <typename V>
static void requireVector(const std::vector<float>& expected, const V& actual)
{
    auto h = actual.toHostRowMajor();
    REQUIRE(h.size() == expected.size());
    for (size_t i = 0; i < h.size(); ++i) REQUIRE_MSG(h[i] == expected[i], i << ": " << h[i] << " vs " << expected[i]);
}

TEST_CASE("CSR Matrix-Vector Product")
{
    SPattern pattern = fixturePattern();
    REQUIRE_NOTHROW(pattern.assertValid());
    SMatrix_t A(pattern);
    A.getData().slice(0) = VectorXf::fromHost(std::vector<float>{1, 4, 2, 3, 5, 7, 8, 9, 6});
    VectorXf b = VectorXf::fromHost(std::vector<float>{2, 5, 3, -4, 1});
    Profiling::instance().resetAll();
    VectorXf xActual = A * b;
    REQUIRE(Profiling::instance().get(Profiling::EvalAny) == 1);
    REQUIRE(Profiling::instance().get(Profiling::EvalMatmulSparse) == 1);
    requireVector({22, 19, -10, 33}, xActual);
    // += and -= modes
    xActual += A * b;
    requireVector({44, 38, -20, 66}, xActual);
    xActual -= A * b;
    requireVector({22, 19, -10, 33}, xActual);
    // inplace target
    VectorXf y(4);
    y.inplace() = A * b;
    requireVector({22, 19, -10, 33}, y);
}

TEST_CASE("CSR Matrix-Vector Product, batched matrix and vector")
{
    SPattern pattern = fixturePattern();
    typedef SparseMatrix<float, Dynamic, SparseFlags::CSR> BSMatrix_t;
    BSMatrix_t A(pattern, 3);
    std::vector<float> v{1, 4, 2, 3, 5, 7, 8, 9, 6};
    const float ms[3] = {1, 2, -1}, vs[3] = {1, 2, 3};
    for (int k = 0; k < 3; ++k) {
        std::vector<float> vk(v);
        for (auto& x : vk) x *= ms[k];
        A.getData().slice(k) = VectorXf::fromHost(vk);
    }
    std::vector<float> b{2, 5, 3, -4, 1}, bb;
    for (int k = 0; k < 3; ++k)
        for (float x : b) bb.push_back(x * vs[k]);
    BVectorXf B = BVectorXf::fromHost(bb.data(), 5, 1, 3);
    BVectorXf y = A * B;
    std::vector<float> e;
    const float y0[4] = {22, 19, -10, 33};
    for (int k = 0; k < 3; ++k)
        for (float x : y0) e.push_back(x * ms[k] * vs[k]);
    requireVector(e, y);
    // batched vector, shared matrix
    SMatrix_t A1(pattern);
    A1.getData().slice(0) = VectorXf::fromHost(v);
    BVectorXf y2 = A1 * B;
    std::vector<float> e2;
    for (int k = 0; k < 3; ++k)
        for (float x : y0) e2.push_back(x * vs[k]);
    requireVector(e2, y2);
}

TEST_CASE("CSR product on 2D Poisson with integer x is exact")
{
    const int g = 96, n = g * g;
    std::vector<int> ja{0}, ia;
    std::vector<float> va;
    for (int x = 0; x < g; ++x)
        for (int yy = 0; yy < g; ++yy) {
            int row = yy + x * g;
            if (x > 0) { ia.push_back(row - g); va.push_back(-1); }
            if (yy > 0) { ia.push_back(row - 1); va.push_back(-1); }
            ia.push_back(row); va.push_back(4);
            if (yy < g - 1) { ia.push_back(row + 1); va.push_back(-1); }
            if (x < g - 1) { ia.push_back(row + g); va.push_back(-1); }
            ja.push_back((int)ia.size());
        }
    SPattern P;
    P.rows = P.cols = n;
    P.nnz = (Index)ia.size();
    P.JA = SPattern::IndexVector::fromHost(ja);
    P.IA = SPattern::IndexVector::fromHost(ia);
    SMatrix_t A(P);
    A.getData().slice(0) = VectorXf::fromHost(va);
    std::mt19937 gen(3);
    std::uniform_int_distribution<int> u(-50, 50);
    std::vector<float> x(n);
    for (auto& t : x) t = (float)u(gen);
    VectorXf y = A * VectorXf::fromHost(x);
    auto h = y.toHostRowMajor();
    for (int r = 0; r < n; ++r) {
        float e = 0;
        for (int k = ja[r]; k < ja[r + 1]; ++k) e += va[k] * x[ia[k]];
        REQUIRE(h[r] == e);
    }
    // A * 1 = number of missing neighbours (Dirichlet boundary)
    VectorXf ones = VectorXf::Constant(n, 1.0f);
    auto h1 = VectorXf(A * ones).toHostRowMajor();
    for (int r = 0; r < n; ++r) REQUIRE(h1[r] == 4.0f - (float)(ja[r + 1] - ja[r] - 1));
}

TEST_CASE("ELLPACK Matrix-Vector Product")
{
    typedef SparsityPattern<SparseFlags::ELLPACK> EPattern;
    EPattern P;
    P.rows = 4;
    P.cols = 5;
    P.nnzPerRow = 3;
    // column-major rows x 3 index matrix, -1 = padding
    std::vector<int> ind{0, 1, 0, 2, 1, 2, 3, 4, -1, -1, 4, -1};
    std::vector<float> val{1, 2, 5, 9, 4, 3, 7, 6, 0, 0, 8, 0};
    P.indices = EPattern::IndexMatrix::fromHost(ind.data(), 4, 3, 1);
    REQUIRE_NOTHROW(P.assertValid());
    SMatrixXf_ELLPACK A(P);
    A.getData().slice(0) = MatrixXf::fromHost(val.data(), 4, 3, 1);
    VectorXf b = VectorXf::fromHost(std::vector<float>{2, 5, 3, -4, 1});
    VectorXf y = A * b;
    requireVector({22, 19, -10, 33}, y);
    // ELLPACK MatrixExpression-Vector Product (tests/TestSparseMultOp.cu:329-375)
    Profiling::instance().resetAll();
    VectorXf y2 = (2 * A.direct()).sparseView<SparseFlags::ELLPACK>(A.getSparsityPattern()) * b;
    REQUIRE(Profiling::instance().get(Profiling::EvalAny) == 1);
    REQUIRE(Profiling::instance().get(Profiling::EvalMatmulSparse) == 1);
    requireVector({44, 38, -20, 66}, y2);
}

TEST_CASE("CSR MatrixExpression-Vector Product")
{
    // tests/TestSparseMultOp.cu:104-153: sections "sparse matmul with search"
    // and "sparse matmul with direct access", one evaluation each
    SPattern pattern = fixturePattern();
    SMatrix_t A(pattern);
    A.getData().slice(0) = VectorXf::fromHost(std::vector<float>{1, 4, 2, 3, 5, 7, 8, 9, 6});
    VectorXf b = VectorXf::fromHost(std::vector<float>{2, 5, 3, -4, 1});
    Profiling::instance().resetAll();
    VectorXf x2 = (2 * A).sparseView<SparseFlags::CSR>(A.getSparsityPattern()) * b;
    REQUIRE(Profiling::instance().get(Profiling::EvalAny) == 1);
    REQUIRE(Profiling::instance().get(Profiling::EvalMatmulSparse) == 1);
    requireVector({44, 38, -20, 66}, x2);
    Profiling::instance().resetAll();
    VectorXf x3 = (2 * A.direct()).sparseView<SparseFlags::CSR>(A.getSparsityPattern()) * b;
    REQUIRE(Profiling::instance().get(Profiling::EvalAny) == 1);
    REQUIRE(Profiling::instance().get(Profiling::EvalMatmulSparse) == 1);
    requireVector({44, 38, -20, 66}, x3);
    // an expression over two matrices of the pattern, a row-dependent and a
    // column-dependent vector: v = A + B - r(row) + c(col), x3 += view * b
    SMatrix_t B(pattern);
    B.getData().slice(0) = VectorXf::fromHost(std::vector<float>{1, 1, 1, 1, 1, 1, 1, 1, 1});
    VectorXf r = VectorXf::fromHost(std::vector<float>{1, 2, 3, 4});
    RowVectorXf c = RowVectorXf::fromHost(std::vector<float>{0, 1, 0, 1, 0});
    x3 += (A.direct() + B.direct() - r + c).sparseView<SparseFlags::CSR>(pattern) * b;
    // row 0: (1+1-1+0)*2 + (4+1-1+1)*5 = 27; row 1: (2+1-2+1)*5 + (3+1-2+0)*3 = 16
    // row 2: (5+1-3+0)*2 + (7+1-3+1)*-4 + (8+1-3+0)*1 = -12; row 3: (9+1-4+0)*3 + (6+1-4+0)*1 = 21
    requireVector({44 + 27, 38 + 16, -20 - 12, 66 + 21}, x3);
}

TEST_CASE("Sparse-Cwise and compound operations on CSR values")
{
    SPattern pattern = fixturePattern();
    SMatrix_t s(pattern);
    s.getData().slice(0) = VectorXf::fromHost(std::vector<float>{1, 4, 2, 3, 5, 7, 8, 9, 6});
    SMatrix_t t(pattern);
    Profiling::instance().resetAll();
    t = SMatrix_t::Identity(4, 5) + s.direct().cwiseAbs2();
    REQUIRE(Profiling::instance().get(Profiling::EvalCwiseSparse) == 1);
    // {2,16,0,0,0},{0,5,9,0,0},{25,0,0,49,64},{0,0,81,0,36}: the identity only
    // lands on stored entries
    requireVector({2, 16, 0, 0, 0, 0, 5, 9, 0, 0, 25, 0, 0, 49, 64, 0, 0, 81, 0, 36}, [&] {
        struct W {
            std::vector<float> v;
            std::vector<float> toHostRowMajor() const { return v; }
        };
        return W{t.toDenseHost()};
    }());
    // dense expression evaluated at the stored entries
    int dd[1][4][5] = {{{1, 2, 3, 4, 5}, {6, 7, 8, 9, 10}, {11, 12, 13, 14, 15}, {16, 17, 18, 19, 20}}};
    MatrixXiR D = MatrixXiR::fromArray(dd);
    typedef SparseMatrix<int, 1, SparseFlags::CSR> SMatrixI;
    SMatrixI si(pattern);
    si = D * 2;
    auto hv = si.getData().toHostRowMajor();
    const int expect[9] = {2, 4, 14, 16, 22, 28, 30, 36, 40};
    for (int k = 0; k < 9; ++k) REQUIRE(hv[k] == expect[k]);
    // compound with a matrix of the same pattern, and with a scalar
    SMatrixI s2 = si.deepClone();
    s2 += si;
    s2 *= 3;
    auto h2 = s2.getData().toHostRowMajor();
    for (int k = 0; k < 9; ++k) REQUIRE(h2[k] == 6 * expect[k]);
    s2 -= si;
    h2 = s2.getData().toHostRowMajor();
    for (int k = 0; k < 9; ++k) REQUIRE(h2[k] == 5 * expect[k]);
}

TEST_CASE("malformed sparsity patterns are rejected on the host")
{
    SPattern p = fixturePattern();
    p.IA = SPattern::IndexVector::fromHost(std::vector<int>{0, 1, 1, 2, 0, 3, 5, 2, 4});  // column 5 of 5
    REQUIRE_THROWS_AS(p.assertValid(), std::invalid_argument);
    SPattern q = fixturePattern();
    q.JA = SPattern::IndexVector::fromHost(std::vector<int>{0, 2, 1, 7, 9});  // decreasing
    REQUIRE_THROWS_AS(SMatrix_t(q), std::invalid_argument);
    SPattern r = fixturePattern();
    r.nnz = 8;
    REQUIRE_THROWS_AS(r.assertValid(), std::invalid_argument);
    SMatrix_t A(fixturePattern());
    VectorXf wrong(4);
    REQUIRE_THROWS_AS(VectorXf(A * wrong), std::invalid_argument);
}

TEST_CASE("operator<< of a CSR matrix prints pattern and values")
{
    SMatrix_t A(fixturePattern());
    A.getData().slice(0) = VectorXf::fromHost(std::vector<float>{1, 4, 2, 3, 5, 7, 8, 9, 6});
    std::ostringstream ss;
    ss << A;
    REQUIRE_MSG(ss.str() == "SparseMatrix: \n rows=4, cols=5, batches=1 (compile-time), storage=CSR\n"
                            " Outer Indices (row): 0 2 4 7 9\n Inner Indices (column): 0 1 1 2 0 3 4 2 4\n"
                            " Data:\n   1 4 2 3 5 7 8 9 6\n",
                ss.str());
}

HARNESS_MAIN()
