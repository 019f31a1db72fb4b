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

// ---------------------------------------------------------------- block scalars
// The reference's blocked CG (tests/TestBlockedConjugateGradient.cu:244-275)
// with the built-in block types: a 3x3 block is three float4 rows, a vector
// entry one float4 (w = 0), as the reference test's real3x3 / real3.
typedef BlockRows<3> Block3;
typedef SparseMatrix<Block3, 1, SparseFlags::CSR> SMatrix3x3;
typedef Matrix<Float4, Dynamic, 1, 1, ColumnMajor> Vector3X;
typedef Matrix<Float4, Dynamic, 1, Dynamic, ColumnMajor> Vector3XB;

static Float4 r3(float x, float y, float z) { return Float4{x, y, z, 0.0f}; }
static Block3 blk(Float4 a, Float4 b, Float4 c)
{
    Block3 m;
    m.r[0] = a;
    m.r[1] = b;
    m.r[2] = c;
    return m;
}

struct HostBsr {
    int n = 0;
    std::vector<int> ja{0}, ia;
    std::vector<Block3> va;
};

static SMatrix3x3 toDeviceBlocks(const HostBsr& h)
{
    SparsityPattern<SparseFlags::CSR> P;
    P.rows = P.cols = h.n;
    P.nnz = (Index)h.ia.size();
    P.JA = SparsityPattern<SparseFlags::CSR>::IndexVector::fromHost(h.ja);
    P.IA = SparsityPattern<SparseFlags::CSR>::IndexVector::fromHost(h.ia);
    SMatrix3x3 A(P);
    A.getData().copyFromHost(h.va.data());
    return A;
}

// the block product on the host in float, in the device's order: per block
// row the first block's product, then += in CSR order; component q the
// left-to-right four-term dot of row q with x_j
static std::vector<Float4> hostBlockProduct(const HostBsr& h, const std::vector<Float4>& x)
{
    std::vector<Float4> y(h.n, Float4{0, 0, 0, 0});
    for (int i = 0; i < h.n; ++i)
        for (int k = h.ja[i]; k < h.ja[i + 1]; ++k) {
            const Float4& xv = x[h.ia[k]];
            float p[4] = {0, 0, 0, 0};
            for (int q = 0; q < 3; ++q) {
                const Float4& a = h.va[k].r[q];
                p[q] = a.x * xv.x + a.y * xv.y + a.z * xv.z + a.w * xv.w;
            }
            Float4& o = y[i];
            if (k == h.ja[i]) {
                o = Float4{p[0], p[1], p[2], p[3]};
            } else {
                o.x += p[0];
                o.y += p[1];
                o.z += p[2];
                o.w += p[3];
            }
        }
    return y;
}

// |b - A x| / |b| over the first three components, in double
static double blockRelResidual(const HostBsr& h, const std::vector<Float4>& b, const std::vector<Float4>& x)
{
    double rr = 0, bb = 0;
    for (int i = 0; i < h.n; ++i) {
        double ax[3] = {0, 0, 0};
        for (int k = h.ja[i]; k < h.ja[i + 1]; ++k) {
            const Float4& xv = x[h.ia[k]];
            for (int q = 0; q < 3; ++q) {
                const Float4& a = h.va[k].r[q];
                ax[q] += (double)a.x * xv.x + (double)a.y * xv.y + (double)a.z * xv.z + (double)a.w * xv.w;
            }
        }
        const double bv[3] = {b[i].x, b[i].y, b[i].z};
        for (int q = 0; q < 3; ++q) {
            rr += (bv[q] - ax[q]) * (bv[q] - ax[q]);
            bb += bv[q] * bv[q];
        }
    }
    return std::sqrt(rr / bb);
}

// 2D 5-point block Poisson: block (i, j) = L_ij * K with K = tridiag(-0.25, 1.5, -0.25),
// diagonal blocks + 0.05 I
static HostBsr blockPoisson(int g)
{
    HostBsr h;
    h.n = g * g;
    auto K = [](float s, float d) {
        return blk(r3(s * 1.5f + d, -0.25f * s, 0), r3(-0.25f * s, s * 1.5f + d, -0.25f * s),
                   r3(0, -0.25f * s, s * 1.5f + d));
    };
    for (int x = 0; x < g; ++x)
        for (int y = 0; y < g; ++y) {
            int row = y + x * g;
            if (x > 0) { h.ia.push_back(row - g); h.va.push_back(K(-1, 0)); }
            if (y > 0) { h.ia.push_back(row - 1); h.va.push_back(K(-1, 0)); }
            h.ia.push_back(row); h.va.push_back(K(4, 0.05f));
            if (y < g - 1) { h.ia.push_back(row + 1); h.va.push_back(K(-1, 0)); }
            if (x < g - 1) { h.ia.push_back(row + g); h.va.push_back(K(-1, 0)); }
            h.ja.push_back((int)h.ia.size());
        }
    return h;
}

static std::vector<Float4> randomBlockVector(int n, unsigned seed)
{
    std::mt19937 gen(seed);
    std::uniform_real_distribution<float> u(-1, 1);
    std::vector<Float4> v(n);
    for (auto& t : v) {
        const float a = u(gen), b = u(gen), c = u(gen);
        t = r3(a, b, c);
    }
    return v;
}

TEST_CASE("Blocked CG (TestBlockedConjugateGradient.cu:244-275)")
{
    SparsityPattern<SparseFlags::CSR> pattern;
    pattern.rows = 3;
    pattern.cols = 3;
    pattern.nnz = 7;
    pattern.JA = SparsityPattern<SparseFlags::CSR>::IndexVector::fromHost(std::vector<int>{0, 2, 5, 7});
    pattern.IA = SparsityPattern<SparseFlags::CSR>::IndexVector::fromHost(std::vector<int>{0, 1, 0, 1, 2, 1, 2});
    SMatrix3x3 mat(pattern);
    std::vector<Block3> data1 = {
        blk(r3(2, -1, 0), r3(-1, 2, -1), r3(0, -1, 2)), blk(r3(0, 0, 0), r3(0, 0, 0), r3(-1, 0, 0)),
        blk(r3(0, 0, -1), r3(0, 0, 0), r3(0, 0, 0)),   blk(r3(2, -1, 0), r3(-1, 2, -1), r3(0, -1, 2)),
        blk(r3(0, 0, 0), r3(0, 0, 0), r3(-1, 0, 0)),   blk(r3(0, 0, -1), r3(0, 0, 0), r3(0, 0, 0)),
        blk(r3(2, -1, 0), r3(-1, 2, -1), r3(0, -1, 1))};
    mat.getData().copyFromHost(data1.data());

    Vector3X b(3);
    std::vector<Float4> data2 = {r3(0.1f, 0, 0), r3(0, 0, 0), r3(0, 0, -0.1f)};
    b.copyFromHost(data2.data());

    ConjugateGradient<SMatrix3x3> cg(mat);
    cg.setMaxIterations(30);
    float tol = 1e-5f;
    cg.setTolerance(tol);
    Vector3X x = cg.solve(b);
    REQUIRE(cg.error() < cg.tolerance());  // converged
    REQUIRE(cg.iterations() < cg.maxIterations());
    REQUIRE(cg.iterations() > 0);  // but something was done

    HostBsr h;
    h.n = 3;
    h.ja = {0, 2, 5, 7};
    h.ia = {0, 1, 0, 1, 2, 1, 2};
    h.va = data1;
    std::vector<Float4> xh(3);
    x.copyToHost(xh.data());
    REQUIRE(blockRelResidual(h, data2, xh) < 1e-4);
    for (const auto& v : xh) REQUIRE(v.w == 0.0f);
}

TEST_CASE("Block SpMV on a 2D block Poisson matrix: the host's bits, y += A x")
{
    const HostBsr h = blockPoisson(40);
    const SMatrix3x3 A = toDeviceBlocks(h);
    const auto xh = randomBlockVector(h.n, 3);
    Vector3X x(h.n);
    x.copyFromHost(xh.data());
    Vector3X y = A * x;
    std::vector<Float4> yh(h.n);
    y.copyToHost(yh.data());
    const auto ref = hostBlockProduct(h, xh);
    REQUIRE(std::memcmp(yh.data(), ref.data(), sizeof(Float4) * h.n) == 0);
    y += A * x;  // the compound mode of the product
    y.copyToHost(yh.data());
    bool doubled = true;
    for (int i = 0; i < h.n; ++i)
        doubled = doubled && yh[i].x == ref[i].x + ref[i].x && yh[i].y == ref[i].y + ref[i].y &&
                  yh[i].z == ref[i].z + ref[i].z && yh[i].w == 0.0f;
    REQUIRE(doubled);
}

TEST_CASE("Blocked CG on a 2D block Poisson matrix: solve, guess, batched right-hand sides")
{
    const HostBsr h = blockPoisson(40);
    const SMatrix3x3 A = toDeviceBlocks(h);
    const auto bh = randomBlockVector(h.n, 5);
    Vector3X b(h.n);
    b.copyFromHost(bh.data());
    ConjugateGradient<SMatrix3x3> cg(A);
    cg.setTolerance(1e-6f);
    Vector3X x = cg.solve(b);
    REQUIRE(cg.error() < 1e-6f && cg.iterations() > 0);
    std::vector<Float4> xh(h.n);
    x.copyToHost(xh.data());
    REQUIRE(blockRelResidual(h, bh, xh) < 1e-5);
    // the solution as the guess: no more than a couple of iterations
    Vector3X x2 = cg.solveWithGuess(b, x);
    REQUIRE(cg.iterations() <= 2 && cg.error() < 1e-6f);
    // two right-hand sides: both reach the tolerance
    const auto b2 = randomBlockVector(h.n, 6);
    std::vector<Float4> both(bh);
    both.insert(both.end(), b2.begin(), b2.end());
    Vector3XB bb(h.n, 1, 2);
    bb.copyFromHost(both.data());
    auto xb = cg.solve(bb);
    std::vector<Float4> xbh(2 * (size_t)h.n);
    xb.copyToHost(xbh.data());
    REQUIRE(blockRelResidual(h, bh, std::vector<Float4>(xbh.begin(), xbh.begin() + h.n)) < 1e-5);
    REQUIRE(blockRelResidual(h, b2, std::vector<Float4>(xbh.begin() + h.n, xbh.end())) < 1e-5);
}
