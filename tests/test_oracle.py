"""CPU tests that pin the oracle (no GPU): the reference's own golden vectors,
the Eigen build of the reference's vendored Eigen (oracle/_ref, when built),
the committed Eigen fixtures, and closed forms chosen so that a plausible
mistake (dropped term, wrong index, transposed operand) fails."""
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(__file__)
GOLD = json.load(open(os.path.join(HERE, "golden", "reference_tests.json")))
EIG = json.load(open(os.path.join(HERE, "golden", "eigen_cg.json")))
OPS = {"sum": 0, "prod": 1, "min": 2, "max": 3, "all": 4, "any": 5}


def _flat(data3, row_major):
    a = np.array(data3)
    return (a if row_major else np.transpose(a, (0, 2, 1))).reshape(-1)


def test_spmv_golden(orc):
    g = GOLD["csr_4x5"]
    JA, IA = np.array(g["JA"], np.int32), np.array(g["IA"], np.int32)
    V = np.array(g["vals"], np.float32)
    b = np.array(g["b"], np.float32)
    assert np.array_equal(orc.csr_spmv(JA, IA, V, b, 4, 5), np.array(g["y"], np.float32))
    # batched matrix x batched vector with scales (1,2,-1) x (1,2,3)
    V3 = np.concatenate([V * s for s in g["batched_both_matrix_scales"]])
    b3 = np.concatenate([b * s for s in g["batched_both_vector_scales"]])
    y3 = orc.csr_spmv(JA, IA, V3, b3, 4, 5, batches=3, vals_batched=True, x_batched=True)
    exp = np.concatenate([np.array(g["y"], np.float32) * s for s in g["batched_both_expected_scales"]])
    assert np.array_equal(y3, exp)
    e = GOLD["ell_4x5"]
    ind = np.asfortranarray(np.array(e["indices_rows"], np.int32)).reshape(-1, order="F")
    val = np.asfortranarray(np.array(e["values_rows"], np.float32)).reshape(-1, order="F")
    assert np.array_equal(orc.ell_spmv_f32(np.ascontiguousarray(ind), np.ascontiguousarray(val), b, 4, 5, 3),
                          np.array(e["y"], np.float32))


def test_spmv_plants_fail(orc):
    """A transposed operand gives a different answer on the fixture (the
    check is not vacuous)."""
    g = GOLD["csr_4x5"]
    dense = np.zeros((4, 5), np.float32)
    for r in range(4):
        for k in range(g["JA"][r], g["JA"][r + 1]):
            dense[r, g["IA"][k]] = g["vals"][k]
    assert np.array_equal(dense @ np.array(g["b"], np.float32), np.array(g["y"], np.float32))
    assert not np.array_equal(dense[:, :4].T @ np.array(g["b"][:4], np.float32), np.array(g["y"], np.float32))


@pytest.mark.parametrize("fixture", ["reduce_2x4x3_rowmajor", "reduce_2x3x4_axis", "reduce_prod_2x3x2"])
@pytest.mark.parametrize("row_major", [True, False])
def test_reduce_golden(orc, fixture, row_major):
    g = GOLD[fixture]
    B, R, C = g["batches_rows_cols"]
    flat = _flat(g["data"], row_major).astype(np.int32)
    for op in ("sum", "prod", "min", "max"):
        for axis, exp in g.get(op, {}).items():
            got = orc.reduce(flat, OPS[op], int(axis), R, C, B, row_major, True)
            assert np.array_equal(got, np.array(exp, np.int32)), (fixture, op, axis)


def test_reduce_all_any_dot_norm_golden(orc):
    g = GOLD["reduce_all_any_2x3x4"]
    for key, op in (("all", 4), ("any", 5)):
        flat = _flat(g[key + "_data"], True).astype(np.uint8)
        assert np.array_equal(orc.reduce(flat, op, 3, 3, 4, 2, True, True), np.array(g[key + "_rc"], np.uint8))
    d = GOLD["dot_batched"]
    a = np.array(d["a"], np.int32).reshape(-1)
    b = np.array(d["b"], np.int32)
    assert np.array_equal(orc.dot(a, b, 3, batches=2, b_batched=False), np.array(d["dot"], np.int32))
    nn = GOLD["norm_2x1x4"]
    v = np.array(nn["data"], np.float64).reshape(-1)
    assert np.array_equal(orc.dot(v, v, 4, batches=2), np.array(nn["squared_norm"], np.float64))
    t = GOLD["trace_2x4x4"]
    m = np.array(t["data"])
    assert [int(np.trace(m[i])) for i in range(2)] == t["trace"]


def test_reduce_closed_forms(orc):
    """sum 1..N, constant inputs, and axis sums of an arange tensor against
    numpy for every axis mask and layout."""
    R, C, B = 6, 5, 4
    x = np.arange(1, R * C * B + 1, dtype=np.int32)
    assert orc.reduce(x, 0, 7, R, C, B, False, False)[0] == (R * C * B) * (R * C * B + 1) // 2
    for rm in (False, True):
        logical = (x.reshape(B, R, C) if rm else np.transpose(x.reshape(B, C, R), (0, 2, 1))).astype(np.int64)
        for axis in range(8):
            ax = tuple(a for bit, a in ((1, 1), (2, 2), (4, 0)) if axis & bit)
            exp = logical.sum(axis=ax, keepdims=True) if ax else logical
            for orm in (False, True):
                got = orc.reduce(x, 0, axis, R, C, B, rm, orm)
                e = exp if orm else np.transpose(exp, (0, 2, 1))
                assert np.array_equal(got.astype(np.int64), e.reshape(-1)), (rm, axis, orm)


def test_cwise_golden_and_semantics(orc):
    g = GOLD["compound_2x2"]
    m = np.array(g["mat1"], np.int32).reshape(-1)
    assert np.array_equal(orc.assign(1, m, m), np.array(g["add_self"]).reshape(-1))
    assert np.array_equal(orc.assign(2, np.zeros(4, np.int32), m), np.array(g["zero_minus"]).reshape(-1))
    assert np.array_equal(orc.assign(5, m + 3, m), np.array(g["plus3_mod"]).reshape(-1))
    # C semantics of / and % (truncation toward zero), not numpy's floor
    a = np.array([-7, 7, -7, 7], np.int32)
    b = np.array([2, -2, -2, 2], np.int32)
    assert list(orc.cwise_binary(3, a, b)) == [-3, -3, 3, 3]
    assert list(orc.cwise_binary(4, a, b)) == [-1, 1, -1, 1]
    # wrap-around of int32 arithmetic
    big = np.array([2**31 - 1], np.int32)
    assert orc.cwise_binary(0, big, np.array([1], np.int32))[0] == -2**31
    # inverseCheck guards zero (cuMat/src/UnaryOps.h:177)
    assert list(orc.cwise_unary(4, np.array([0.0, 2.0], np.float32))) == [1.0, 0.5]


def test_poisson_generator_matches(orc):
    from cumat_amd.problems import poisson2d
    for g in (1, 2, 5, 128):
        JA, IA, V = poisson2d(g, 0.25, np.float64)
        JA2, IA2, V2 = orc.poisson2d(g, 0.25)
        assert np.array_equal(JA, JA2) and np.array_equal(IA, IA2) and np.array_equal(V, V2)
    JA, IA, V = poisson2d(128)
    assert IA.size == 81408  # the 128^2 system: n=16,384, nnz=81,408
    JA, IA, V = poisson2d(2048)
    assert IA.size == 20963328


def test_pcg_against_eigen_fixtures(orc):
    """The restated PCG against the committed Eigen runs: FP64 iteration
    counts equal and iterates agree to rounding."""
    from cumat_amd.problems import poisson2d
    for c in EIG["cases"]:
        JA, IA, V = poisson2d(c["g"], c["shift"], np.float64)
        n = c["g"] ** 2
        b = np.array(c["b"])
        x, it, err = orc.pcg(JA, IA, V, b, n, tol=c["tol"], max_iter=c["max_iter"])
        assert it == c["iterations_f64"]
        np.testing.assert_allclose(err, c["error_f64"], rtol=1e-6)
        np.testing.assert_allclose(x, np.array(c["x_f64"]), rtol=0, atol=1e-12 * np.abs(x).max())
        x32, it32, err32 = orc.pcg(JA, IA, V.astype(np.float32), b.astype(np.float32), n, tol=c["tol_f32"],
                                   max_iter=c["max_iter"])
        assert abs(it32 - c["iterations_f32"]) <= 2 and err32 <= c["tol_f32"]


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(HERE), "oracle", "_ref", "libeigen_ref.so")),
                    reason="oracle/_ref (Eigen build) not built here")
def test_pcg_against_eigen_build(orc):
    """The config-1 system (128^2 Poisson, FP32, tol 1e-6) against Eigen
    itself: same iteration count, iterates equal to FP32 rounding."""
    from cumat_amd.problems import poisson2d
    JA, IA, V = poisson2d(128)
    n = 16384
    b = np.random.default_rng(20261203).uniform(-1, 1, n).astype(np.float32)
    x, it, err = orc.pcg(JA, IA, V, b, n, tol=1e-6, max_iter=1000)
    xe, ite, erre = orc.eigen_pcg(JA, IA, V, b, n, tol=1e-6, max_iter=1000)
    assert abs(it - ite) <= 1 and abs(err - erre) <= 1e-8
    assert np.linalg.norm(x - xe) / np.linalg.norm(xe) < 1e-5


def test_pcg_closed_forms(orc):
    # diagonal matrix: one step, not counted; x = D^-1 b
    n = 50
    d = np.linspace(1, 5, n)
    JA = np.arange(n + 1, dtype=np.int32)
    IA = np.arange(n, dtype=np.int32)
    b = np.cos(np.arange(n))
    x, it, err = orc.pcg(JA, IA, d, b, n, tol=1e-14)
    assert it == 0 and np.allclose(x, b / d, rtol=1e-15)
    # zero rhs early out
    x, it, err = orc.pcg(JA, IA, d, np.zeros(n), n, x0=np.ones(n))
    assert it == 0 and err == 0 and not x.any()
    # max-iteration exhaustion keeps the count and the last residual
    from cumat_amd.problems import poisson2d
    JA, IA, V = poisson2d(20, 0.0, np.float64)
    bb = np.ones(400)
    x, it, err = orc.pcg(JA, IA, V, bb, 400, tol=1e-14, max_iter=5)
    assert it == 5 and err > 1e-3


def test_edge_grad_closed_form(orc):
    from cumat_amd.problems import grid_laplacian
    rng = np.random.default_rng(0)
    P = grid_laplacian((4, 5), lambda s: np.exp(0.5 * rng.standard_normal(s)))
    n = P["n"]
    lam, x = rng.standard_normal(n), rng.standard_normal(n)
    g = orc.edge_grad(P["ei"], P["ej"], lam, x)
    # central differences of lam^T A(theta) x in theta (bilinear: exact)
    A = np.zeros((n, n))
    rows = np.repeat(np.arange(n), np.diff(P["JA"]))
    A[rows, P["IA"]] = P["vals"]
    for e in range(0, P["nedges"], 3):
        th = P["theta"].copy()
        h = 1e-3
        th[e] += h
        Ap = grid_laplacian((4, 5), _fixed(th))
        th[e] -= 2 * h
        Am = grid_laplacian((4, 5), _fixed(th))
        fd = (_bil(Ap, lam, x) - _bil(Am, lam, x)) / (2 * h)
        assert abs(-fd - g[e]) < 1e-9


def _fixed(theta):
    state = {"o": 0}

    def f(shape):
        k = int(np.prod(shape))
        out = theta[state["o"]:state["o"] + k].reshape(shape)
        state["o"] += k
        return out
    return f


def _bil(P, lam, x):
    rows = np.repeat(np.arange(P["n"]), np.diff(P["JA"]))
    return float(lam @ np.bincount(rows, P["vals"] * x[P["IA"]], minlength=P["n"]))


def test_spmv_i32_golden_and_closed_form(orc):
    """orc_csr_spmv_i32: the reference's 4x5 fixture (tests/TestSparseMultOp.cu:
    70-276) in int32, batched both ways, and the integer closed form of the
    Dirichlet Poisson matrix: A * 1 = number of missing neighbours per row."""
    g = GOLD["csr_4x5"]
    JA, IA = np.array(g["JA"], np.int32), np.array(g["IA"], np.int32)
    V = np.array(g["vals"], np.int32)
    b = np.array(g["b"], np.int32)
    assert list(orc.csr_spmv(JA, IA, V, b, 4, 5)) == [22, 19, -10, 33]
    V3 = np.concatenate([V * s for s in g["batched_both_matrix_scales"]]).astype(np.int32)
    b3 = np.concatenate([b * s for s in g["batched_both_vector_scales"]]).astype(np.int32)
    y3 = orc.csr_spmv(JA, IA, V3, b3, 4, 5, batches=3, vals_batched=True, x_batched=True)
    exp = np.concatenate([np.array(g["y"]) * s for s in g["batched_both_expected_scales"]])
    assert np.array_equal(y3, exp.astype(np.int32))
    # compound modes (+=, -=) on the fixture
    y0 = np.array([1, 2, 3, 4], np.int32)
    assert list(orc.csr_spmv(JA, IA, V, b, 4, 5, y=y0.copy(), mode=1)) == [23, 21, -7, 37]
    assert list(orc.csr_spmv(JA, IA, V, b, 4, 5, y=y0.copy(), mode=2)) == [-21, -17, 13, -29]
    # a transposed operand (columns read as rows) must give another answer
    assert not np.array_equal(orc.csr_spmv(JA, IA, V, b[::-1].copy(), 4, 5), np.array(g["y"], np.int32))
    for gg in (1, 3, 17):
        from cumat_amd.problems import poisson2d
        JAp, IAp, Vp = poisson2d(gg, 0.0, np.float64)
        n = gg * gg
        y1 = orc.csr_spmv(JAp, IAp, Vp.astype(np.int32), np.ones(n, np.int32), n, n)
        assert np.array_equal(y1, (5 - np.diff(JAp)).astype(np.int32))
    # int32 arithmetic wraps like the device's
    big = orc.csr_spmv(np.array([0, 2], np.int32), np.array([0, 1], np.int32), np.array([2**30, 2**30], np.int32),
                       np.array([2, 1], np.int32), 1, 2)
    assert big[0] == np.int32(-2**30)


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.int32])
def test_ell_golden_all_dtypes(orc, dtype):
    """orc_ell_spmv_{f32,f64,i32}: the reference's ELLPACK fixture
    (tests/TestSparseMultOp.cu:278-519, padding value 666 never read), a
    batched matrix with scales 1,2,3, and the += mode."""
    e = GOLD["ell_4x5"]
    ind = np.asfortranarray(np.array(e["indices_rows"], np.int32)).reshape(-1, order="F")
    val = np.asfortranarray(np.array(e["values_rows"], dtype)).reshape(-1, order="F")
    b = np.array(e["b"], dtype)
    assert np.array_equal(orc.ell_spmv(ind, val, b, 4, 5, 3), np.array(e["y"], dtype))
    v3 = np.concatenate([val * s for s in (1, 2, 3)]).astype(dtype)
    y3 = orc.ell_spmv(ind, v3, b, 4, 5, 3, batches=3, vals_batched=True)
    assert np.array_equal(y3, np.concatenate([np.array(e["y"], dtype) * s for s in (1, 2, 3)]))
    y0 = np.array([1, 1, 1, 1], dtype)
    assert np.array_equal(orc.ell_spmv(ind, val, b, 4, 5, 3, y=y0, mode=1), np.array(e["y"], dtype) + 1)


def test_ell_matches_csr_on_poisson(orc):
    """The ELLPACK restatement equals the CSR restatement on a 3D 7-point
    Laplacian stored both ways (entries in the same order, so the same sums)."""
    from cumat_amd.problems import grid_laplacian
    rng = np.random.default_rng(3)
    P = grid_laplacian((6, 7, 8), lambda s: np.exp(rng.standard_normal(s)), edges=False)
    n = P["n"]
    cnt = np.diff(P["JA"])
    ind = -np.ones((n, 7), np.int32)
    val = np.zeros((n, 7))
    slot = np.arange(P["nnz"]) - np.repeat(P["JA"][:-1], cnt)
    rows = np.repeat(np.arange(n), cnt)
    ind[rows, slot] = P["IA"]
    val[rows, slot] = P["vals"]
    x = rng.standard_normal(n)
    ye = orc.ell_spmv(np.asfortranarray(ind).reshape(-1, order="F"), np.asfortranarray(val).reshape(-1, order="F"),
                      x, n, n, 7)
    yc = orc.csr_spmv(P["JA"], P["IA"], P["vals"], x, n, n)
    assert np.array_equal(ye, yc)


# ------------------------------------------------------------------ block CSR
def _bcsr_dense(D, JA, IA, V, rows, cols):
    """Dense (4 rows) x (4 cols) float64 matrix a block matrix acts as."""
    A = np.zeros((4 * rows, 4 * cols))
    V = np.asarray(V, np.float64).reshape(len(IA), D, 4)
    for i in range(rows):
        for k in range(JA[i], JA[i + 1]):
            A[4 * i:4 * i + D, 4 * IA[k]:4 * IA[k] + 4] += V[k]
    return A


def _rand_bcsr(rng, rows, cols, D, max_per_row, integer=True):
    JA = [0]
    IA = []
    for _ in range(rows):
        k = int(rng.integers(0, max_per_row + 1))
        IA.extend(sorted(rng.choice(cols, size=min(k, cols), replace=False)))
        JA.append(len(IA))
    nnz = len(IA)
    V = (rng.integers(-4, 5, (nnz, D, 4)) if integer else rng.standard_normal((nnz, D, 4))).astype(np.float32)
    return np.array(JA, np.int32), np.array(IA, np.int32), V


@pytest.mark.parametrize("D", [2, 3, 4])
def test_bcsr_spmv_dense_expansion(orc, D):
    rng = np.random.default_rng(10 + D)
    rows, cols = 37, 29
    JA, IA, V = _rand_bcsr(rng, rows, cols, D, 9)
    x = rng.integers(-5, 6, (3, cols * 4)).astype(np.float32)
    A = _bcsr_dense(D, JA, IA, V, rows, cols)
    y = orc.bcsr_spmv(D, JA, IA, V, x[0], rows, cols)
    assert np.array_equal(y, (A @ x[0]).astype(np.float32))
    assert np.all(y.reshape(rows, 4)[:, D:] == 0)
    # batched values and vectors (scales 1, 2, -1 on the values)
    V3 = np.concatenate([V * s for s in (1, 2, -1)])
    y3 = orc.bcsr_spmv(D, JA, IA, V3, x.reshape(-1), rows, cols, batches=3, vals_batched=True, x_batched=True)
    exp = np.concatenate([s * (A @ x[b]) for b, s in enumerate((1, 2, -1))]).astype(np.float32)
    assert np.array_equal(y3, exp)
    # compound modes
    y0 = rng.integers(-5, 6, rows * 4).astype(np.float32)
    for mode, f in ((1, lambda a, b: a + b), (2, lambda a, b: a - b)):
        yy = orc.bcsr_spmv(D, JA, IA, V, x[0], rows, cols, y=y0.copy(), mode=mode)
        assert np.array_equal(yy, f(y0, (A @ x[0]).astype(np.float32)))


def test_bcsr_spmv_equals_scalar_csr_when_blocks_hold_one_scalar(orc):
    # blocks whose only entry is (0, 0) and vectors whose only component is
    # x: component x of the block product is the scalar CSR product, bit for
    # bit with random floats (same products, same CSR-order sums)
    rng = np.random.default_rng(7)
    rows, cols = 300, 280
    JA, IA, _ = _rand_bcsr(rng, rows, cols, 3, 12)
    s = rng.standard_normal(len(IA)).astype(np.float32)
    V = np.zeros((len(IA), 3, 4), np.float32)
    V[:, 0, 0] = s
    xs = rng.standard_normal(cols).astype(np.float32)
    x4 = np.zeros((cols, 4), np.float32)
    x4[:, 0] = xs
    y = orc.bcsr_spmv(3, JA, IA, V, x4.reshape(-1), rows, cols).reshape(rows, 4)
    assert np.array_equal(y[:, 0], orc.csr_spmv(JA, IA, s, xs, rows, cols))
    assert np.all(y[:, 1:] == 0)


def test_bcsr_inverse_diagonal(orc):
    from cumat_amd.problems import block_poisson
    JA, IA, V = block_poisson((5, 6), 3, seed=3)
    rows = 30
    V = V.copy()
    # row 4: zero the (1, 1) entry of its diagonal block; row 7: no diagonal block
    k4 = [k for k in range(JA[4], JA[5]) if IA[k] == 4][0]
    V[k4, 1, 1] = 0
    k7 = [k for k in range(JA[7], JA[8]) if IA[k] == 7][0]
    IA = IA.copy()
    IA[k7] = IA[k7 - 1] if k7 > JA[7] else IA[k7 + 1]
    out = orc.bcsr_inverse_diagonal(3, JA, IA, V, rows).reshape(rows, 4)
    for i in range(rows):
        ks = [k for k in range(JA[i], JA[i + 1]) if IA[k] == i]
        d = np.diag(V[ks[0], :, :3]) if ks else np.zeros(3, np.float32)
        exp = np.where(d == 0, np.float32(1), np.float32(1) / np.where(d == 0, np.float32(1), d))
        assert np.array_equal(out[i, :3], exp.astype(np.float32)), i
        assert out[i, 3] == 0
    assert out[4, 1] == 1 and np.all(out[7, :3] == 1)


def test_bcsr_pcg_reference_fixture(orc):
    g = GOLD["blocked_cg_3x3"]
    D, rows = g["block_dim"], g["rows"]
    V = np.array(g["blocks"], np.float32)
    b = np.array(g["b"], np.float32).reshape(-1)
    x, it, err = orc.bcsr_pcg(D, g["JA"], g["IA"], V, b, rows, tol=g["tolerance"],
                              max_iter=g["max_iterations"], dtype=np.float32)
    assert err < g["tolerance"] and 0 < it < g["max_iterations"]
    A = _bcsr_dense(D, np.array(g["JA"]), np.array(g["IA"]), V, rows, rows)
    sel = [4 * i + q for i in range(rows) for q in range(D)]
    xe = np.linalg.solve(A[np.ix_(sel, sel)], b[sel].astype(np.float64))
    assert np.allclose(x[sel], xe, rtol=0, atol=1e-4 * np.abs(xe).max())
    assert np.all(x.reshape(rows, 4)[:, 3] == 0)


def test_bcsr_pcg_closed_forms(orc):
    from cumat_amd.problems import block_poisson
    rng = np.random.default_rng(5)
    # block-diagonal matrix of diagonal blocks: Jacobi-PCG is exact at once
    rows = 40
    JA = np.arange(rows + 1, dtype=np.int32)
    IA = np.arange(rows, dtype=np.int32)
    V = np.zeros((rows, 3, 4), np.float64)
    d = rng.uniform(1, 3, (rows, 3))
    for q in range(3):
        V[:, q, q] = d[:, q]
    b = rng.standard_normal(rows * 4)
    x, it, err = orc.bcsr_pcg(3, JA, IA, V, b, rows, tol=1e-12)
    assert it == 0 and err < 1e-14
    assert np.allclose(x.reshape(rows, 4)[:, :3], b.reshape(rows, 4)[:, :3] / d, rtol=1e-14)
    # a general SPD block system: A x = b to the FP64 tolerance, and the w
    # component of b does not enter (the reference's squaredNorm and
    # preconditioner ignore it)
    JA, IA, V = block_poisson((6, 7), 3, seed=9)
    rows = 42
    b = rng.standard_normal(rows * 4)
    x, it, err = orc.bcsr_pcg(3, JA, IA, V, b, rows, tol=1e-12, max_iter=500)
    A = _bcsr_dense(3, JA, IA, V, rows, rows)
    sel = [4 * i + q for i in range(rows) for q in range(3)]
    r = b[sel] - A[np.ix_(sel, sel)] @ x[sel]
    assert np.linalg.norm(r) <= 1e-11 * np.linalg.norm(b[sel]) and 0 < it <= 3 * rows
    b2 = b.reshape(rows, 4).copy()
    b2[:, 3] = 0
    x2, it2, _ = orc.bcsr_pcg(3, JA, IA, V, b2.reshape(-1), rows, tol=1e-12, max_iter=500)
    assert it2 == it and np.array_equal(x2, x)
