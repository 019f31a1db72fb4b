// C++ host-API tests of the Eigen interop (CUMAT_EIGEN_SUPPORT=1), built
// only where Eigen's headers are available. Dense conversions follow
// tests/TestEigenInterop.cu:1-120 of the reference (coefficient (x, y) holds
// x + 100 y, every storage order and compile-time / dynamic size there);
// the sparse part follows the reference's CSR benchmark, which builds the
// cuMat pattern from an Eigen::SparseMatrix by hand
// (benchmarks/csrmv/Implementation_cuMat_CSR.cu:30-70), and checks the
// product and the Jacobi-CG solve against Eigen's own on the same matrix:
// SpMV bit-exact (both sum each row in CSR order), CG with the same
// iteration count in FP64 and the iterate within 1e-10 (FP64) / 1e-5 (FP32).
#define CUMAT_EIGEN_SUPPORT 1
#include <cuMat/IterativeLinearSolvers>

#include <Eigen/Core>
#include <Eigen/IterativeLinearSolvers>
#include <Eigen/SparseCore>

#include <random>

#include "harness.h"

using namespace cuMat;

template <typename M>
static void testMatrixToEigen(Index rows, Index cols)
{
    // host data in the matrix's own storage order
    std::vector<typename M::Scalar> h((size_t)(rows * cols));
    for (Index y = 0; y < cols; ++y)
        for (Index x = 0; x < rows; ++x) {
            const Index k = (M::Flags & RowMajor) ? y + cols * x : x + rows * y;
            h[(size_t)k] = (typename M::Scalar)(x + y * 100);
        }
    M m = M::fromHost(h.data(), rows, cols, 1);
    auto e = m.toEigen();
    REQUIRE(e.rows() == rows && e.cols() == cols);
    for (Index y = 0; y < cols; ++y)
        for (Index x = 0; x < rows; ++x) REQUIRE(e(x, y) == (typename M::Scalar)(x + y * 100));
}

TEST_CASE("matrix_to_eigen")
{
    testMatrixToEigen<Matrix<float, 1, 1, 1, ColumnMajor>>(1, 1);
    testMatrixToEigen<Matrix<float, 4, 8, 1, ColumnMajor>>(4, 8);
    testMatrixToEigen<Matrix<int, 16, 8, 1, ColumnMajor>>(16, 8);
    testMatrixToEigen<Matrix<float, Dynamic, Dynamic, 1, ColumnMajor>>(32, 6);
    testMatrixToEigen<Matrix<float, 1, 1, 1, RowMajor>>(1, 1);
    testMatrixToEigen<Matrix<float, 4, 8, 1, RowMajor>>(4, 8);
    testMatrixToEigen<Matrix<int, 16, 8, 1, RowMajor>>(16, 8);
    testMatrixToEigen<Matrix<float, Dynamic, Dynamic, 1, RowMajor>>(32, 6);
    testMatrixToEigen<Matrix<double, Dynamic, Dynamic, Dynamic, ColumnMajor>>(7, 5);
    testMatrixToEigen<Matrix<float, Dynamic, 1, 1, ColumnMajor>>(9, 1);
    testMatrixToEigen<Matrix<float, 1, Dynamic, 1, ColumnMajor>>(1, 9);
    testMatrixToEigen<Matrix<float, 1, Dynamic, 1, RowMajor>>(1, 9);
}

TEST_CASE("matrix_to_eigen rejects batches")
{
    Matrix<float, Dynamic, Dynamic, Dynamic, ColumnMajor> m(3, 2, 2);
    REQUIRE_THROWS_AS(m.toEigen(), std::invalid_argument);
}

template <typename E>
static void testMatrixFromEigen(E host)
{
    const Index sx = host.rows(), sy = host.cols();
    for (Index y = 0; y < sy; ++y)
        for (Index x = 0; x < sx; ++x) host(x, y) = (typename E::Scalar)(x + y * 100);
    typedef typename eigen::MatrixEigenToCuMat<E>::type matrix_t;
    static_assert(int(matrix_t::Flags) == (E::IsRowMajor ? int(RowMajor) : int(ColumnMajor)), "storage carries over");
    matrix_t mat = matrix_t::fromEigen(host);
    REQUIRE(mat.rows() == sx && mat.cols() == sy && mat.batches() == 1);
    const auto v = mat.toHostRowMajor();  // logical [row][col] order
    for (Index y = 0; y < sy; ++y)
        for (Index x = 0; x < sx; ++x) REQUIRE(v[(size_t)(x * sy + y)] == (typename E::Scalar)(x + y * 100));
    // and back
    auto back = mat.toEigen();
    REQUIRE(back.rows() == sx && back.cols() == sy);
    for (Index y = 0; y < sy; ++y)
        for (Index x = 0; x < sx; ++x) REQUIRE(back(x, y) == host(x, y));
}

TEST_CASE("matrix_from_eigen")
{
    testMatrixFromEigen(Eigen::Matrix<float, 8, 6, Eigen::RowMajor>());
    {
        Eigen::Matrix<float, Eigen::Dynamic, 6, Eigen::RowMajor> m;
        m.resize(12, 6);
        testMatrixFromEigen(m);
    }
    {
        Eigen::Matrix<float, Eigen::Dynamic, Eigen::Dynamic, Eigen::RowMajor> m;
        m.resize(12, 24);
        testMatrixFromEigen(m);
    }
    testMatrixFromEigen(Eigen::Matrix<float, 8, 6, Eigen::ColMajor>());
    {
        Eigen::Matrix<float, 16, Eigen::Dynamic, Eigen::ColMajor> m;
        m.resize(16, 8);
        testMatrixFromEigen(m);
    }
    {
        Eigen::Matrix<float, Eigen::Dynamic, Eigen::Dynamic, Eigen::ColMajor> m;
        m.resize(12, 24);
        testMatrixFromEigen(m);
    }
    {
        Eigen::Matrix<int, Eigen::Dynamic, Eigen::Dynamic, Eigen::ColMajor> m;
        m.resize(5, 3);
        testMatrixFromEigen(m);
    }
    {
        Eigen::VectorXd v(17);
        testMatrixFromEigen(v);
        Eigen::RowVectorXf w(13);
        testMatrixFromEigen(w);
    }
}

// 2D 5-point Poisson (diag 4 + shift, off-diagonals -1), rows y + x*g
template <typename S, int Options = Eigen::RowMajor>
static Eigen::SparseMatrix<S, Options, int> eigenPoisson(int g, double shift = 0.0)
{
    std::vector<Eigen::Triplet<S>> t;
    for (int x = 0; x < g; ++x)
        for (int y = 0; y < g; ++y) {
            const int row = y + x * g;
            if (x > 0) t.emplace_back(row, row - g, S(-1));
            if (y > 0) t.emplace_back(row, row - 1, S(-1));
            t.emplace_back(row, row, S(4 + shift));
            if (y < g - 1) t.emplace_back(row, row + 1, S(-1));
            if (x < g - 1) t.emplace_back(row, row + g, S(-1));
        }
    Eigen::SparseMatrix<S, Options, int> m(g * g, g * g);
    m.setFromTriplets(t.begin(), t.end());
    m.makeCompressed();
    return m;
}

template <typename S>
static Eigen::Matrix<S, Eigen::Dynamic, 1> uniformVector(Index n, unsigned seed, bool integers = false)
{
    std::mt19937 gen(seed);
    std::uniform_real_distribution<double> u(-1, 1);
    Eigen::Matrix<S, Eigen::Dynamic, 1> v(n);
    for (Index i = 0; i < n; ++i) v[i] = integers ? (S)std::floor(u(gen) * 9) : (S)u(gen);
    return v;
}

TEST_CASE("csr pattern built by hand from Eigen (the reference's CSR benchmark)")
{
    // benchmarks/csrmv/Implementation_cuMat_CSR.cu:47-62, same steps
    const int matrixSize = 40 * 40;
    Eigen::SparseMatrix<float, Eigen::RowMajor> matrix = eigenPoisson<float>(40);
    Eigen::VectorXf ex = uniformVector<float>(matrixSize, 3);

    typedef SparseMatrix<float, 1, CSR> SMatrix;
    typedef SparsityPattern<CSR> SPattern;
    SPattern pattern;
    pattern.rows = matrixSize;
    pattern.cols = matrixSize;
    pattern.nnz = matrix.nonZeros();
    pattern.JA = SPattern::IndexVector(matrixSize + 1);
    pattern.JA.copyFromHost(matrix.outerIndexPtr());
    pattern.IA = SPattern::IndexVector(pattern.nnz);
    pattern.IA.copyFromHost(matrix.innerIndexPtr());
    pattern.assertValid();
    SMatrix mat(pattern);
    mat.getData().copyFromHost(matrix.valuePtr());

    VectorXf x = VectorXf::fromEigen(ex);
    VectorXf r(matrixSize);
    r = mat * x;
    Eigen::VectorXf er = matrix * ex;
    Eigen::VectorXf gr = r.toEigen();
    for (int i = 0; i < matrixSize; ++i) REQUIRE(gr[i] == er[i]);  // same order per row: bit-exact
}

TEST_CASE("csr pattern and sparse matrix from / to Eigen")
{
    // column-major and uncompressed Eigen input are converted to CSR
    Eigen::SparseMatrix<double, Eigen::ColMajor> cm = eigenPoisson<double, Eigen::ColMajor>(9, 0.5);
    cm.coeffRef(3, 77) = 2.5;  // one extra entry: uncompressed, not symmetric
    REQUIRE(!cm.isCompressed());
    const Eigen::SparseMatrix<double, Eigen::RowMajor> rm(cm);

    auto P = SparsityPattern<CSR>::fromEigen(cm);
    REQUIRE(P.rows == 81 && P.cols == 81 && P.nnz == rm.nonZeros());
    const auto ja = P.JA.toHostRowMajor(), ia = P.IA.toHostRowMajor();
    for (Index r = 0; r <= 81; ++r) REQUIRE(ja[(size_t)r] == rm.outerIndexPtr()[r]);
    for (Index k = 0; k < P.nnz; ++k) REQUIRE(ia[(size_t)k] == rm.innerIndexPtr()[k]);

    auto A = SparseMatrix<double, 1, CSR>::fromEigen(cm);
    REQUIRE(A.rows() == 81 && A.cols() == 81 && A.nnz() == rm.nonZeros());
    const Eigen::SparseMatrix<double, Eigen::RowMajor, int> back = A.toEigen();
    REQUIRE(back.isCompressed() && back.nonZeros() == rm.nonZeros());
    for (Index r = 0; r <= 81; ++r) REQUIRE(back.outerIndexPtr()[r] == rm.outerIndexPtr()[r]);
    for (Index k = 0; k < rm.nonZeros(); ++k) {
        REQUIRE(back.innerIndexPtr()[k] == rm.innerIndexPtr()[k]);
        REQUIRE(back.valuePtr()[k] == rm.valuePtr()[k]);
    }
    REQUIRE(back.coeff(3, 77) == 2.5);

    // product against Eigen's, FP64 bit-exact
    Eigen::VectorXd ex = uniformVector<double>(81, 9);
    VectorXd y = A * VectorXd::fromEigen(ex);
    Eigen::VectorXd ey = rm * ex;
    Eigen::VectorXd gy = y.toEigen();
    for (int i = 0; i < 81; ++i) REQUIRE(gy[i] == ey[i]);

    // batched cuMat matrices have no Eigen form
    SparseMatrix<double, Dynamic, CSR> B(P, 2);
    REQUIRE_THROWS_AS(B.toEigen(), std::invalid_argument);
}

template <typename S>
static void compareCgWithEigen(int g, double shift, S tol, int iterSlack, double xTol)
{
    typedef Eigen::SparseMatrix<S, Eigen::RowMajor, int> ESparse;
    typedef Eigen::Matrix<S, Eigen::Dynamic, 1> EVec;
    const ESparse A = eigenPoisson<S>(g, shift);
    const EVec b = uniformVector<S>(A.rows(), 20261202u + g);

    Eigen::ConjugateGradient<ESparse, Eigen::Lower | Eigen::Upper, Eigen::DiagonalPreconditioner<S>> ecg;
    ecg.setTolerance(tol);
    ecg.setMaxIterations(1000);
    ecg.compute(A);
    const EVec ex = ecg.solve(b);
    REQUIRE(ecg.info() == Eigen::Success);

    typedef SparseMatrix<S, 1, CSR> SMatrix;
    typedef Matrix<S, Dynamic, 1, 1, ColumnMajor> Vec;
    const SMatrix Ad = SMatrix::fromEigen(A);
    ConjugateGradient<SMatrix> cg;
    cg.setTolerance(tol).setMaxIterations(1000);
    cg.compute(Ad);
    const Vec x = cg.solve(Vec::fromEigen(b));
    const EVec gx = x.toEigen();

    REQUIRE(std::abs((long)cg.iterations() - (long)ecg.iterations()) <= iterSlack);
    REQUIRE(cg.error() <= tol);
    REQUIRE((gx - ex).norm() / ex.norm() <= xTol);
}

TEST_CASE("ConjugateGradient matches Eigen's ConjugateGradient + DiagonalPreconditioner")
{
    compareCgWithEigen<double>(32, 0.0, 1e-10, 0, 1e-10);
    compareCgWithEigen<double>(24, 0.3, 1e-12, 0, 1e-11);
    compareCgWithEigen<float>(32, 0.0, 1e-6f, 3, 1e-5);
    compareCgWithEigen<float>(64, 0.0, 1e-6f, 3, 1e-5);
}

HARNESS_MAIN()
