// C++ host-API tests of the Jacobi-preconditioned Conjugate Gradient, in the
// structure of tests/TestConjugateGradient.cu:11-99 of the reference (early
// outs, convergence, batched right-hand sides), on CSR matrices, plus the
// batched and mixed-precision solvers. Expected values come from closed forms
// and from residuals evaluated on the host in double precision.
#include <cuMat/IterativeLinearSolvers>

#include <random>

#include "harness.h"

using namespace cuMat;

struct HostCsr {
    int n = 0;
    std::vector<int> ja{0}, ia;
    std::vector<double> va;
};

// 2D 5-point Poisson (+ shift on the diagonal), rows in y + x*g order
static HostCsr poisson(int g, double shift = 0.0)
{
    HostCsr h;
    h.n = g * g;
    for (int x = 0; x < g; ++x)
        for (int y = 0; y < g; ++y) {
            int row = y + x * g;
            if (x > 0) { h.ia.push_back(row - g); h.va.push_back(-1); }
            if (y > 0) { h.ia.push_back(row - 1); h.va.push_back(-1); }
            h.ia.push_back(row); h.va.push_back(4 + shift);
            if (y < g - 1) { h.ia.push_back(row + 1); h.va.push_back(-1); }
            if (x < g - 1) { h.ia.push_back(row + g); h.va.push_back(-1); }
            h.ja.push_back((int)h.ia.size());
        }
    return h;
}

template <typename S>
static SparseMatrix<S, 1, SparseFlags::CSR> toDevice(const HostCsr& h)
{
    SparsityPattern<SparseFlags::CSR> P;
    P.rows = P.cols = h.n;
    P.nnz = (Index)h.ia.size();
    P.JA = SparsityPattern<SparseFlags::CSR>::IndexVector::fromHost(h.ja);
    P.IA = SparsityPattern<SparseFlags::CSR>::IndexVector::fromHost(h.ia);
    SparseMatrix<S, 1, SparseFlags::CSR> A(P);
    std::vector<S> v(h.va.begin(), h.va.end());
    A.getData().slice(0) = Matrix<S, Dynamic, 1, 1, ColumnMajor>::fromHost(v);
    return A;
}

// |b - A x| / |b| on the host in double precision
template <typename S>
static double relResidual(const HostCsr& h, const std::vector<S>& b, const std::vector<S>& x, int batch = 0)
{
    double rr = 0, bb = 0;
    for (int r = 0; r < h.n; ++r) {
        double ax = 0;
        for (int k = h.ja[r]; k < h.ja[r + 1]; ++k) ax += h.va[k] * (double)x[h.ia[k] + batch * h.n];
        double d = (double)b[r + batch * h.n] - ax;
        rr += d * d;
        bb += (double)b[r + batch * h.n] * b[r + batch * h.n];
    }
    return std::sqrt(rr / bb);
}

template <typename S>
static std::vector<S> randomVector(int n, unsigned seed)
{
    std::mt19937 gen(seed);
    std::uniform_real_distribution<double> u(-1, 1);
    std::vector<S> v(n);
    for (auto& t : v) t = (S)u(gen);
    return v;
}

TEST_CASE("Conjugate Gradient - Early Out Rhs")
{
    for (int g : {5, 10, 50}) {
        auto A = toDevice<float>(poisson(g));
        VectorXf b = VectorXf::Constant(g * g, 0.0f);
        ConjugateGradient<SMatrixXf> cg(A);
        VectorXf x = cg.solve(b);
        REQUIRE(cg.iterations() == 0);
        REQUIRE(cg.error() == 0);
        REQUIRE(x.squaredNorm().getScalar() == 0);
    }
}

TEST_CASE("Conjugate Gradient - Early Out X")
{
    for (int g : {5, 10, 50}) {
        HostCsr h = poisson(g, 0.5);
        auto A = toDevice<double>(h);
        VectorXd xTruth = VectorXd::fromHost(randomVector<double>(h.n, g));
        VectorXd b = A * xTruth;
        ConjugateGradient<SMatrixXd> cg(A);
        VectorXd x = cg.solveWithGuess(b, xTruth);
        REQUIRE(cg.iterations() == 0);
        REQUIRE(cg.error() <= 1e-10);
        auto hx = x.toHostRowMajor(), ht = xTruth.toHostRowMajor();
        for (int i = 0; i < h.n; ++i) REQUIRE(hx[i] == ht[i]);
    }
}

TEST_CASE("Conjugate Gradient - Solve Poisson (FP64 and FP32, both launch paths)")
{
    for (int g : {8, 32, 100}) {
        HostCsr h = poisson(g);
        auto bh = randomVector<double>(h.n, 100 + g);
        {
            auto A = toDevice<double>(h);
            for (int path : {CM_CG_MULTIKERNEL, CM_CG_RESIDENT}) {
                ConjugateGradient<SMatrixXd> cg(A);
                cg.setTolerance(1e-10).setMaxIterations(10 * h.n).setPath(path);
                VectorXd x = cg.solve(VectorXd::fromHost(bh));
                REQUIRE(cg.lastStats().path_used == path);
                REQUIRE(cg.iterations() > 0 && cg.iterations() < 10 * h.n);
                REQUIRE(cg.error() <= 1e-10);
                REQUIRE_MSG(relResidual<double>(h, bh, x.toHostRowMajor()) <= 2e-10,
                            "g=" << g << " path=" << path);
            }
        }
        {
            auto A = toDevice<float>(h);
            std::vector<float> bf(bh.begin(), bh.end());
            ConjugateGradient<SMatrixXf> cg(A);
            cg.setTolerance(1e-6f);
            VectorXf x = cg.solve(VectorXf::fromHost(bf));
            REQUIRE(cg.error() <= 1e-6f);
            // the FP32 recurrence residual drifts from the true residual by
            // about eps_f32 * cond(A) (cond ~ 4e3 at g = 100)
            REQUIRE(relResidual<float>(h, bf, x.toHostRowMajor()) <= 1e-4);
        }
    }
}

TEST_CASE("Conjugate Gradient - batched right-hand sides")
{
    const int g = 40, B = 3;
    HostCsr h = poisson(g, 0.1);
    auto A = toDevice<double>(h);
    auto bh = randomVector<double>(h.n * B, 7);
    for (int i = 0; i < h.n; ++i) bh[i + h.n] = 2 * bh[i];  // batch 1 = 2 x batch 0
    BVectorXd b = BVectorXd::fromHost(bh.data(), h.n, 1, B);
    ConjugateGradient<SMatrixXd> cg(A);
    cg.setTolerance(1e-11);
    BVectorXd x = cg.solve(b);
    REQUIRE(x.batches() == B);
    auto hx = x.toHostRowMajor();
    for (int k = 0; k < B; ++k) REQUIRE(relResidual<double>(h, bh, hx, k) <= 2e-11);
    for (int i = 0; i < h.n; ++i) REQUIRE(harness::close(hx[i + h.n], 2 * hx[i], 1e-9));
}

TEST_CASE("Conjugate Gradient - closed forms: diagonal matrix, compute() re-targets")
{
    // D^-1-preconditioned CG on a diagonal matrix converges in one iteration
    const int n = 1000;
    HostCsr d;
    d.n = n;
    for (int i = 0; i < n; ++i) {
        d.ia.push_back(i);
        d.va.push_back(1.0 + i % 17);
        d.ja.push_back(i + 1);
    }
    auto D = toDevice<double>(d);
    auto bh = randomVector<double>(n, 9);
    ConjugateGradient<SMatrixXd> cg(D);
    cg.setTolerance(1e-12);
    VectorXd x = cg.solve(VectorXd::fromHost(bh));
    REQUIRE(cg.iterations() <= 1);
    auto hx = x.toHostRowMajor();
    for (int i = 0; i < n; ++i) REQUIRE(harness::close(hx[i], bh[i] / d.va[i], 1e-14));
    // the same solver object on another matrix
    HostCsr p = poisson(20, 1.0);
    cg.compute(toDevice<double>(p));
    auto bp = randomVector<double>(p.n, 10);
    VectorXd xp = cg.solve(VectorXd::fromHost(bp));
    REQUIRE(cg.iterations() > 1);
    REQUIRE(relResidual<double>(p, bp, xp.toHostRowMajor()) <= 2e-12);
    // default settings: tolerance epsilon, 2*cols iterations
    ConjugateGradient<SMatrixXd> cg2(D);
    REQUIRE(cg2.tolerance() == std::numeric_limits<double>::epsilon());
    REQUIRE(cg2.maxIterations() == 2 * n);
}

TEST_CASE("Conjugate Gradient - identity preconditioner")
{
    HostCsr h = poisson(24, 0.3);
    auto A = toDevice<double>(h);
    auto bh = randomVector<double>(h.n, 12);
    ConjugateGradient<SMatrixXd, IdentityPreconditioner<SMatrixXd>> cg(A);
    cg.setTolerance(1e-10);
    VectorXd x = cg.solve(VectorXd::fromHost(bh));
    REQUIRE(relResidual<double>(h, bh, x.toHostRowMajor()) <= 2e-10);
}

TEST_CASE("Batched CG over uneven systems, one launch")
{
    const int nsys = 50;
    std::mt19937 gen(4);
    std::uniform_int_distribution<int> side(8, 30);
    HostCsr all;
    std::vector<int> row_off{0};
    std::vector<HostCsr> parts;
    for (int s = 0; s < nsys; ++s) {
        HostCsr p = poisson(side(gen), 0.2 * (s % 5));
        int base = row_off.back(), nbase = (int)all.ia.size();
        for (int r = 0; r < p.n; ++r) all.ja.push_back(nbase + p.ja[r + 1]);
        for (size_t k = 0; k < p.ia.size(); ++k) {
            all.ia.push_back(p.ia[k] + base);
            all.va.push_back(p.va[k]);
        }
        row_off.push_back(base + p.n);
        parts.push_back(p);
    }
    all.n = row_off.back();
    auto A = toDevice<float>(all);
    auto bh = randomVector<float>(all.n, 13);
    VectorXf b = VectorXf::fromHost(bh), x;
    VectorXi its;
    VectorXf errs;
    BatchedConjugateGradient bcg(VectorXi::fromHost(row_off), A);
    bcg.setTolerance(1e-6f).setMaxIterations(2000);
    int bs = bcg.autotune(b);
    REQUIRE(bs == 256 || bs == 512 || bs == 1024);
    bcg.solve(b, x, its, errs);
    auto hx = x.toHostRowMajor();
    auto hi = its.toHostRowMajor();
    auto he = errs.toHostRowMajor();
    for (int s = 0; s < nsys; ++s) {
        REQUIRE(he[s] <= 1e-6f);
        REQUIRE(hi[s] > 0 && hi[s] < 2000);
        std::vector<float> bs_(bh.begin() + row_off[s], bh.begin() + row_off[s + 1]);
        std::vector<float> xs(hx.begin() + row_off[s], hx.begin() + row_off[s + 1]);
        REQUIRE(relResidual<float>(parts[s], bs_, xs) <= 5e-6);
    }
}

TEST_CASE("Mixed-precision CG with FP64 refinement reaches 1e-12")
{
    HostCsr h = poisson(60, 0.0);
    auto A = toDevice<double>(h);
    auto bh = randomVector<double>(h.n, 14);
    for (int storage : {CM_F32, CM_BF16}) {
        MixedPrecisionConjugateGradient ir(A, storage);
        ir.setTolerance(1e-12).setInnerTolerance(storage == CM_BF16 ? 1e-3 : 1e-5);
        VectorXd x = ir.solve(VectorXd::fromHost(bh));
        REQUIRE(ir.error() <= 1e-12);
        REQUIRE(ir.outerIterations() >= 2);
        REQUIRE(relResidual<double>(h, bh, x.toHostRowMajor()) <= 1.5e-12);
    }
}

// ---- dense matrices: ConjugateGradient<MatrixXf> / <MatrixXd> and A * x,
// the four cases of tests/TestConjugateGradient.cu:11-99. The reference fills
// A = U[0,1) + size*I, which is not symmetric; CG reaches the tolerance on it
// only when its skew part happens to be small, so the solve cases take the
// symmetric part (U + U^T)/2 + size*I.
template <typename S>
static std::vector<S> hostDense(int n, unsigned seed, bool spd)
{
    std::mt19937 gen(seed);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    std::vector<double> a((size_t)n * n);
    for (auto& v : a) v = u(gen);
    std::vector<S> out((size_t)n * n);  // column-major
    for (int c = 0; c < n; ++c)
        for (int r = 0; r < n; ++r)
            out[r + (size_t)n * c] = (S)(spd ? 0.5 * (a[r + (size_t)n * c] + a[c + (size_t)n * r]) + (r == c ? n : 0)
                                             : a[r + (size_t)n * c]);
    return out;
}

TEST_CASE("CG dense: early out on a zero right-hand side (TestConjugateGradient.cu:11-25)")
{
    for (int size : {5, 10, 50, 100}) {
        auto ah = hostDense<float>(size, 10 + size, false);
        MatrixXf A = MatrixXf::fromHost(ah.data(), size, size, 1);
        VectorXf b = VectorXf::Zero(size);
        ConjugateGradient<MatrixXf> cg(A);
        VectorXf x = cg.solve(b);
        REQUIRE(cg.iterations() == 0);
        REQUIRE(cg.error() == 0);
        REQUIRE(x.squaredNorm().getScalar() == 0);
    }
}

TEST_CASE("CG dense: early out on an exact guess (TestConjugateGradient.cu:27-44)")
{
    for (int size : {5, 10, 50, 100}) {
        auto ah = hostDense<float>(size, 20 + size, false);
        MatrixXf A = MatrixXf::fromHost(ah.data(), size, size, 1);
        VectorXf xTruth = VectorXf::fromHost(randomVector<float>(size, 30 + size));
        VectorXf b = A * xTruth;
        ConjugateGradient<MatrixXf> cg(A);
        VectorXf x = cg.solveWithGuess(b, xTruth);
        REQUIRE(cg.iterations() == 0);
        REQUIRE(std::fabs(cg.error()) <= 1e-10);
        auto hx = x.toHostRowMajor(), ht = xTruth.toHostRowMajor();
        for (int i = 0; i < size; ++i) REQUIRE(hx[i] == ht[i]);
    }
}

TEST_CASE("CG dense: solve (TestConjugateGradient.cu:46-71)")
{
    for (int size : {5, 10, 50, 100}) {
        auto ah = hostDense<double>(size, 40 + size, true);
        MatrixXd A = MatrixXd::fromHost(ah.data(), size, size, 1);
        VectorXd xTruth = VectorXd::fromHost(randomVector<double>(size, 50 + size));
        VectorXd b = A * xTruth;
        ConjugateGradient<MatrixXd> cg(A);
        cg.setMaxIterations(10 * size);
        VectorXd x = cg.solve(b);
        REQUIRE(cg.iterations() > 0);
        REQUIRE(cg.error() > 0);
        REQUIRE(cg.iterations() < cg.maxIterations());
        REQUIRE(cg.error() <= cg.tolerance());
        auto hx = x.toHostRowMajor(), ht = xTruth.toHostRowMajor();
        for (int i = 0; i < size; ++i) REQUIRE(std::fabs(hx[i] - ht[i]) <= 1e-4 * std::fabs(ht[i]) + 1e-12);
    }
}

TEST_CASE("CG dense: batched right-hand sides (TestConjugateGradient.cu:73-99)")
{
    constexpr int size = 10, batches = 3;
    auto ah = hostDense<double>(size, 60, true);
    MatrixXd A = MatrixXd::fromHost(ah.data(), size, size, 1);
    typedef Matrix<double, Dynamic, 1, batches, ColumnMajor> Vec;
    auto xh = randomVector<double>(size * batches, 61);
    Vec xTruth = Vec::fromHost(xh.data(), size, 1, batches);
    Vec b = A * xTruth;
    ConjugateGradient<MatrixXd> cg(A);
    cg.setMaxIterations(10 * size);
    Vec x = cg.solve(b);
    REQUIRE(cg.iterations() > 0);
    REQUIRE(cg.error() > 0);
    REQUIRE(cg.iterations() < cg.maxIterations());
    REQUIRE(cg.error() <= cg.tolerance());
    auto hx = x.toHostRowMajor();
    for (int i = 0; i < size * batches; ++i) REQUIRE(std::fabs(hx[i] - xh[i]) <= 1e-4 * std::fabs(xh[i]) + 1e-12);
}

TEST_CASE("dense A * x: row-major and column-major against the host, += and -=")
{
    const int rows = 70, cols = 45;
    std::vector<int> h((size_t)rows * cols), xv(cols);
    for (int i = 0; i < rows * cols; ++i) h[i] = (i * 37) % 23 - 11;
    for (int j = 0; j < cols; ++j) xv[j] = (j * 13) % 7 - 3;
    MatrixXi Ac = MatrixXi::fromHost(h.data(), rows, cols, 1);    // column-major storage of h
    MatrixXiR Ar = MatrixXiR::fromHost(h.data(), rows, cols, 1);  // row-major storage of h
    VectorXi x = VectorXi::fromHost(xv);
    VectorXi yc = Ac * x;
    VectorXi yr = Ar * x;
    auto hc = yc.toHostRowMajor(), hr = yr.toHostRowMajor();
    for (int r = 0; r < rows; ++r) {
        long long ec = 0, er = 0;
        for (int c = 0; c < cols; ++c) {
            ec += (long long)h[r + (size_t)rows * c] * xv[c];
            er += (long long)h[(size_t)r * cols + c] * xv[c];
        }
        REQUIRE(hc[r] == ec);
        REQUIRE(hr[r] == er);
    }
    yc += Ac * x;
    yc -= Ac * x;
    yc -= Ac * x;
    auto h2 = yc.toHostRowMajor();
    for (int r = 0; r < rows; ++r) REQUIRE(h2[r] == 0);
    REQUIRE_THROWS_AS(VectorXi(Ac * VectorXi::fromHost(std::vector<int>(cols + 1, 1))), std::invalid_argument);
    // an expression as the left factor is evaluated first: (2 A) x and A^T x
    VectorXi y2 = (Ac * 2) * x;
    auto h3 = y2.toHostRowMajor();
    for (int r = 0; r < rows; ++r) REQUIRE(h3[r] == 2 * hc[r]);
    VectorXi xr = VectorXi::fromHost(std::vector<int>(rows, 1));
    VectorXi yt = Ac.transpose() * xr;
    auto h4 = yt.toHostRowMajor();
    for (int c = 0; c < cols; ++c) {
        long long e = 0;
        for (int r = 0; r < rows; ++r) e += h[r + (size_t)rows * c];
        REQUIRE(h4[c] == e);
    }
}

HARNESS_MAIN()
