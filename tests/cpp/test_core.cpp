// C++ host-API tests of the dense hot path: fused component-wise evaluation
// with assignment modes and broadcasting, casting, views, and reductions.
// Fixtures are the reference's own (tests/TestCompoundAssignment.cu:9-43,
// tests/TestBinaryOps2.cu:13-80, tests/TestReductionOps1a.cu:17-37); the
// expected values of the rest are computed on the host in double precision.
#include <cuMat/Core>

#include <random>
#include <sstream>

#include "harness.h"

using namespace cuMat;

template <typename M, typename T, int B, int R, int C>
static void requireEqual(const T (&expected)[B][R][C], const M& actual)
{
    REQUIRE(actual.rows() == R && actual.cols() == C && actual.batches() == B);
    auto h = actual.toHostRowMajor();
    for (int b = 0; b < B; ++b)
        for (int r = 0; r < R; ++r)
            for (int c = 0; c < C; ++c)
                REQUIRE_MSG(h[(b * R + r) * C + c] == expected[b][r][c],
                            "at (" << r << "," << c << "," << b << "): " << h[(b * R + r) * C + c] << " vs "
                                   << expected[b][r][c]);
}

TEST_CASE("Compound-Matrix")
{
    int data[1][2][2] = {{{1, 2}, {3, 4}}};
    Matrix2iR mat1 = Matrix2iR::fromArray(data);
    Matrix2iR mat2;

    mat2 = mat1.deepClone();
    mat2 += mat1;
    int expected1[1][2][2] = {{{2, 4}, {6, 8}}};
    requireEqual(expected1, mat2);

    mat2 = Matrix2iR::Zero();
    mat2 -= mat1;
    int expected2[1][2][2] = {{{-1, -2}, {-3, -4}}};
    requireEqual(expected2, mat2);

    mat2 = mat1 + 3;
    mat2 %= mat1;
    int expected3[1][2][2] = {{{0, 1}, {0, 3}}};
    requireEqual(expected3, mat2);

    mat2 = mat1.deepClone();
    mat2 *= 2;
    int expected4[1][2][2] = {{{2, 4}, {6, 8}}};
    requireEqual(expected4, mat2);
}

TEST_CASE("broadcasting-multiply")
{
    int data[2][3][4] = {{{5, 2, 4, 1}, {8, 5, 6, 2}, {-5, 0, 3, -9}}, {{11, 0, 0, 3}, {-7, -2, -4, 1}, {9, 4, 7, -7}}};
    BMatrixXiR m1 = BMatrixXiR::fromArray(data);
    BMatrixXiC m2 = m1;  // different storage order: one cwise evaluation
    int expected[2][3][4] = {{{10, 4, 8, 2}, {16, 10, 12, 4}, {-10, 0, 6, -18}},
                             {{22, 0, 0, 6}, {-14, -4, -8, 2}, {18, 8, 14, -14}}};
    requireEqual(expected, BMatrixXiR(2 * m1));
    requireEqual(expected, BMatrixXiR(2 * m2));
    requireEqual(expected, BMatrixXiC(m1 * 2));
    requireEqual(expected, BMatrixXiC(m2 * 2));
    // device scalar (1x1x1) broadcast over all three dimensions
    int scalar[1][1][1] = {{{2}}};
    ScalariR s1 = ScalariR::fromArray(scalar);
    requireEqual(expected, BMatrixXiR(s1.cwiseMul(m1)));
    requireEqual(expected, BMatrixXiR(m2.cwiseMul(s1)));
    // column vector broadcast over columns and batches
    int col[1][3][1] = {{{1}, {2}, {3}}};
    BMatrixXiR c1 = BMatrixXiR::fromArray(col);
    BMatrixXiR prod = m1.cwiseMul(c1);
    auto h = prod.toHostRowMajor();
    for (int b = 0; b < 2; ++b)
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 4; ++c) REQUIRE(h[(b * 3 + r) * 4 + c] == data[b][r][c] * (r + 1));
}

TEST_CASE("fused linear combination: one kernel per assignment")
{
    const int n = 1 << 16;
    std::mt19937 gen(5);
    std::uniform_int_distribution<int> u(-1000, 1000);
    std::vector<std::vector<float>> hv(8, std::vector<float>(n));
    float a[8];
    for (int k = 0; k < 8; ++k) {
        for (auto& x : hv[k]) x = u(gen) / 1000.0f;
        a[k] = u(gen) / 1000.0f;
    }
    std::vector<VectorXf> v;
    for (int k = 0; k < 8; ++k) v.push_back(VectorXf::fromHost(hv[k]));
    VectorXf out(n);
    Profiling::instance().resetAll();
    out.inplace() = v[0] * a[0] + v[1] * a[1] + v[2] * a[2] + v[3] * a[3] + v[4] * a[4] + v[5] * a[5] +
                    v[6] * a[6] + v[7] * a[7];
    REQUIRE(Profiling::instance().get(Profiling::EvalCwise) == 1);
    REQUIRE(Profiling::instance().get(Profiling::EvalAny) == 1);
    REQUIRE(Profiling::instance().get(Profiling::DeviceMemAlloc) == 0);
    auto h = out.toHostRowMajor();
    for (int i = 0; i < n; i += 97) {
        double e = 0;
        for (int k = 0; k < 8; ++k) e += (double)a[k] * hv[k][i];
        REQUIRE_MSG(harness::close(h[i], e, 1e-5), i << ": " << h[i] << " vs " << e);
    }
}

TEST_CASE("integer cwise ops are bit-exact (C semantics of / and %)")
{
    int da[1][1][4] = {{{-7, 7, -7, 7}}};
    int db[1][1][4] = {{{2, -2, -2, 2}}};
    RowVectorXiR a = RowVectorXiR::fromArray(da), b = RowVectorXiR::fromArray(db);
    int q[1][1][4] = {{{-3, -3, 3, 3}}}, m[1][1][4] = {{{-1, 1, -1, 1}}};
    requireEqual(q, RowVectorXiR(a.cwiseDiv(b)));
    requireEqual(m, RowVectorXiR(a % b));
    int x[1][1][4] = {{{-7 ^ 2, 7 ^ -2, -7 ^ -2, 7 ^ 2}}};
    requireEqual(x, RowVectorXiR(a ^ b));
    int big[1][1][1] = {{{2147483647}}};
    ScalariR s = ScalariR::fromArray(big);
    REQUIRE(ScalariR(s + 1).getScalar() == (int)0x80000000u);  // wrap-around
}

TEST_CASE("casting and comparisons")
{
    int data[1][2][3] = {{{1, -2, 3}, {4, 0, -6}}};
    MatrixXiR m = MatrixXiR::fromArray(data);
    MatrixXfR f = m.cast<float>() * 0.5f;
    float ef[1][2][3] = {{{0.5f, -1.f, 1.5f}, {2.f, 0.f, -3.f}}};
    requireEqual(ef, f);
    MatrixXbR pos = m > 0;
    bool eb[1][2][3] = {{{true, false, true}, {true, false, false}}};
    requireEqual(eb, pos);
    MatrixXiR sel = pos.cast<int>().cwiseMul(m);
    int es[1][2][3] = {{{1, 0, 3}, {4, 0, 0}}};
    requireEqual(es, sel);
}

TEST_CASE("views: transpose, diagonal, asDiagonal, inplace, slice")
{
    int data[1][2][3] = {{{1, 2, 3}, {4, 5, 6}}};
    MatrixXiR m = MatrixXiR::fromArray(data);
    MatrixXiR t = m.transpose();
    int et[1][3][2] = {{{1, 4}, {2, 5}, {3, 6}}};
    requireEqual(et, t);
    int sq[1][3][3] = {{{1, 2, 3}, {4, 5, 6}, {7, 8, 9}}};
    MatrixXiR s = MatrixXiR::fromArray(sq);
    VectorXi d = s.diagonal();
    int ed[1][3][1] = {{{1}, {5}, {9}}};
    requireEqual(ed, d);
    MatrixXi D = d.asDiagonal();
    int eD[1][3][3] = {{{1, 0, 0}, {0, 5, 0}, {0, 0, 9}}};
    requireEqual(eD, D);
    MatrixXi I = MatrixXi::Identity(3);
    int eI[1][3][3] = {{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}};
    requireEqual(eI, I);
    // sharing semantics: a copy shares memory, inplace() writes through it
    MatrixXiR shared = m;
    REQUIRE(!m.isExclusiveUse());
    shared.inplace() = shared * 10;
    REQUIRE(m.getCoeff(1, 2, 0) == 60);
    // operator= of an expression on a shared matrix re-allocates
    shared = shared + 1;
    REQUIRE(m.getCoeff(1, 2, 0) == 60);
    REQUIRE(shared.getCoeff(1, 2, 0) == 61);
    // slices of a batched matrix are written in place
    BMatrixXi bm(2, 2, 3);
    bm.setZero();
    bm.slice(1) = Matrix2i::Constant(2, 2, 7);
    auto h = bm.toHostRowMajor();
    for (int k = 0; k < 12; ++k) REQUIRE(h[k] == (k >= 4 && k < 8 ? 7 : 0));
}

static BMatrixXiR createTestMatrix()
{
    int data[2][4][3] = {{{1, 2, 3}, {4, 5, 6}, {7, 8, 9}, {10, 11, 12}},
                         {{13, 14, 15}, {16, 17, 18}, {19, 20, 21}, {22, 23, 24}}};
    return BMatrixXiR::fromArray(data);
}

TEST_CASE("reductions over every axis combination (int32, bit-exact)")
{
    BMatrixXiR m = createTestMatrix();
    BMatrixXiC mc = m;
    auto h = m.toHostRowMajor();
    for (int axis = 0; axis < 8; ++axis) {
        for (int colmajor = 0; colmajor < 2; ++colmajor) {
            auto got = colmajor ? mc.sum(axis).toHostRowMajor() : m.sum(axis).toHostRowMajor();
            int R = (axis & Axis::Row) ? 1 : 4, C = (axis & Axis::Column) ? 1 : 3, B = (axis & Axis::Batch) ? 1 : 2;
            REQUIRE((int)got.size() == R * C * B);
            for (int b = 0; b < B; ++b)
                for (int r = 0; r < R; ++r)
                    for (int c = 0; c < C; ++c) {
                        long long e = 0;
                        for (int bb = 0; bb < 2; ++bb)
                            for (int rr = 0; rr < 4; ++rr)
                                for (int cc = 0; cc < 3; ++cc)
                                    if ((B == 1 || bb == b) && (R == 1 || rr == r) && (C == 1 || cc == c))
                                        e += h[(bb * 4 + rr) * 3 + cc];
                        REQUIRE_MSG(got[(b * R + r) * C + c] == e, "axis " << axis << " at " << r << "," << c << ","
                                                                         << b);
                    }
        }
    }
    int e_row[2][1][3] = {{{22, 26, 30}}, {{70, 74, 78}}};
    requireEqual(e_row, m.sum<Axis::Row>());
    int e_all[1][1][1] = {{{300}}};
    requireEqual(e_all, m.sum<Axis::All>());
    int e_min[2][4][1] = {{{1}, {4}, {7}, {10}}, {{13}, {16}, {19}, {22}}};
    requireEqual(e_min, m.minCoeff<Axis::Column>());
    int e_max[1][4][3] = {{{13, 14, 15}, {16, 17, 18}, {19, 20, 21}, {22, 23, 24}}};
    requireEqual(e_max, m.maxCoeff<Axis::Batch>());
    REQUIRE(m.prod<Axis::Column>().getCoeff(0, 0, 0) == 6);
    REQUIRE(BMatrixXiR(m % 2).bitwiseOr<Axis::All>().getScalar() == 1);
    REQUIRE(BMatrixXiR(m | 1).bitwiseAnd<Axis::All>().getScalar() == 1);
    REQUIRE(BMatrixXbR(m > 0).all<Axis::All>().getScalar() == true);
    REQUIRE(BMatrixXbR(m > 23).any<Axis::All>().getScalar() == true);
    REQUIRE(BMatrixXbR(m > 24).any<Axis::All>().getScalar() == false);
    // reduction of an expression
    REQUIRE((m * 2).sum().getScalar() == 600);
}

TEST_CASE("float reductions, dot, squaredNorm, norm")
{
    const int R = 300, C = 70, B = 3;
    std::mt19937 gen(11);
    std::normal_distribution<float> nd(0.f, 1.f);
    std::vector<float> hv(R * C * B);
    for (auto& x : hv) x = nd(gen);
    BMatrixXf m = BMatrixXf::fromHost(hv.data(), R, C, B);  // column major
    auto cs = m.sum<Axis::Row>().toHostRowMajor();
    for (int b = 0; b < B; ++b)
        for (int c = 0; c < C; ++c) {
            double e = 0, s = 0;
            for (int r = 0; r < R; ++r) {
                e += hv[r + R * (c + C * b)];
                s += std::fabs(hv[r + R * (c + C * b)]);
            }
            REQUIRE(std::fabs(cs[b * C + c] - e) <= 1e-5 * s);
        }
    BVectorXf v = BVectorXf::fromHost(hv.data(), R * C, 1, B);
    auto d = v.dot(v).toHostRowMajor();
    auto sn = v.squaredNorm().toHostRowMajor();
    auto nr = v.norm().toHostRowMajor();
    for (int b = 0; b < B; ++b) {
        double e = 0;
        for (int i = 0; i < R * C; ++i) e += (double)hv[i + R * C * b] * hv[i + R * C * b];
        REQUIRE(harness::close(d[b], e, 1e-5));
        REQUIRE(d[b] == sn[b]);
        REQUIRE(harness::close(nr[b], std::sqrt(e), 1e-5));
    }
}

TEST_CASE("reductions, dot and squaredNorm of expressions run inside the reduction")
{
    // cuMat reads a reduction's input through the expression tree
    // (cuMat/src/Iterator.h:99-105); dot is cwiseDot(rhs).sum<Row|Column>()
    // (cuMat/src/ReductionOpsPlugin.inl:290-297)
    BMatrixXiR m = createTestMatrix();  // 4 x 3 x 2, values 1..24
    BMatrixXiC mc = m;
    auto h = m.toHostRowMajor();
    Profiling::instance().resetAll();
    auto s = (m * 2 - 1).sum<Axis::Row>();
    REQUIRE(Profiling::instance().get(Profiling::EvalCwise) == 0);
    REQUIRE(Profiling::instance().get(Profiling::EvalReduction) == 1);
    auto hs = s.toHostRowMajor();
    for (int b = 0; b < 2; ++b)
        for (int c = 0; c < 3; ++c) {
            int e = 0;
            for (int r = 0; r < 4; ++r) e += 2 * h[(b * 4 + r) * 3 + c] - 1;
            REQUIRE(hs[b * 3 + c] == e);
        }
    for (int axis = 0; axis < 8; ++axis) {
        auto got = (mc + mc).maxCoeff(axis).toHostRowMajor();
        auto ref = BMatrixXiC(mc + mc).maxCoeff(axis).toHostRowMajor();
        REQUIRE_MSG(got == ref, "axis " << axis);
        auto gs = (m - 3).sum(axis).toHostRowMajor();
        auto rs = BMatrixXiR(m - 3).sum(axis).toHostRowMajor();
        REQUIRE_MSG(gs == rs, "axis " << axis);
    }
    const int n = 5000;
    std::vector<float> ha(n), hb(n);
    for (int i = 0; i < n; ++i) {
        ha[i] = std::sin(0.01f * i);
        hb[i] = std::cos(0.02f * i);
    }
    VectorXf a = VectorXf::fromHost(ha.data(), n, 1, 1), b = VectorXf::fromHost(hb.data(), n, 1, 1);
    Profiling::instance().resetAll();
    const float sn = (a - b).squaredNorm().getScalar();
    const float dt = (a * 2.f).dot(b).getScalar();
    REQUIRE(Profiling::instance().get(Profiling::EvalCwise) == 0);
    double esn = 0, edt = 0, sdt = 0;
    for (int i = 0; i < n; ++i) {
        esn += (double)(ha[i] - hb[i]) * (ha[i] - hb[i]);
        edt += 2.0 * ha[i] * hb[i];
        sdt += std::fabs(2.0 * ha[i] * hb[i]);
    }
    REQUIRE(harness::close(sn, esn, 1e-5));
    REQUIRE(std::fabs(dt - edt) <= 1e-5 * sdt);
}

TEST_CASE("dimension and type errors throw")
{
    VectorXf a(5), b(6);
    REQUIRE_THROWS_AS(VectorXf(a + b), std::invalid_argument);
    Matrix2i m;
    REQUIRE_THROWS_AS(m.slice(1), std::out_of_range);
}

TEST_CASE("operator<< prints shape, storage and the batches")
{
    int d[2][2][3] = {{{1, 2, 3}, {4, 5, 6}}, {{7, 8, 9}, {10, 11, 12}}};
    auto m = BMatrixXiR::fromArray(d);
    std::ostringstream ss;
    ss << m;
    REQUIRE_MSG(ss.str() == "Matrix: rows=2 (dynamic), cols=3 (dynamic), batches=2 (dynamic), storage=Row-Major\n"
                            "[\n 1  2  3\n 4  5  6\n],[\n 7  8  9\n10 11 12\n]",
                ss.str());
    std::ostringstream s1;
    s1 << VectorXf::fromHost(std::vector<float>{1.5f, -2.0f});
    REQUIRE_MSG(s1.str() == "Matrix: rows=2 (dynamic), cols=1 (compile-time), batches=1 (compile-time), "
                            "storage=Column-Major\n1.5\n -2",
                s1.str());
}

HARNESS_MAIN()
