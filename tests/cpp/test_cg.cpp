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

HARNESS_MAIN()
