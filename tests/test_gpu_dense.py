"""GPU tests of the dense matrix-vector product (the matrix-vector case of
cuMat's ProductOp, cuMat/src/ProductOp.h:318-420) and of ConjugateGradient
with a dense matrix, restating tests/TestConjugateGradient.cu:11-99 of the
reference (early out on a zero right-hand side, early out on an exact guess,
diagonally dominant dense solve, batched right-hand sides), plus parity of
the dense CG with the oracle's CG on the same matrix in CSR form.

The matrices of those reference tests are uniform random (SimpleRandom
fillUniform); here they come from numpy with fixed seeds. Products are
checked against numpy in FP64: int32 bit-exact, FP64 to 1e-12 and FP32 to
1e-5 relative (the GPU sums the columns in four interleaved partial sums)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dense(cm, a, row_major=False):
    return cm.DeviceArray.from_numpy(a, row_major=row_major)


@pytest.mark.parametrize("row_major", [False, True])
@pytest.mark.parametrize("shape", [(1, 1), (5, 7), (64, 64), (65, 31), (130, 257), (1000, 3), (3, 1000),
                                   (777, 1031)])
def test_matvec_float_shapes(gpu, row_major, shape):
    cm = gpu
    rng = np.random.default_rng(shape[0] * 7919 + shape[1])
    rows, cols = shape
    for dt, tol in ((np.float64, 1e-12), (np.float32, 1e-5)):
        a = rng.uniform(-1, 1, (rows, cols)).astype(dt)
        x = rng.uniform(-1, 1, cols).astype(dt)
        y = cm.dense.matvec(_dense(cm, a, row_major), cm.DeviceArray.from_numpy(x)).numpy()
        ref = a.astype(np.float64) @ x.astype(np.float64)
        scale = np.abs(a).astype(np.float64) @ np.abs(x).astype(np.float64) + 1e-300
        assert np.all(np.abs(y - ref) <= tol * scale), (dt, shape, np.max(np.abs(y - ref) / scale))


@pytest.mark.parametrize("row_major", [False, True])
@pytest.mark.parametrize("shape", [(97, 203), (130, 2600)])  # one column range / five ranges + combine
def test_matvec_int32_exact_and_modes(gpu, row_major, shape):
    cm = gpu
    rng = np.random.default_rng(5)
    rows, cols = shape
    a = rng.integers(-50, 50, (rows, cols)).astype(np.int32)
    x = rng.integers(-50, 50, cols).astype(np.int32)
    A = _dense(cm, a, row_major)
    y = cm.dense.matvec(A, cm.DeviceArray.from_numpy(x))
    np.testing.assert_array_equal(y.numpy(), a.astype(np.int64) @ x)
    y0 = rng.integers(-9, 9, rows).astype(np.int32)
    out = cm.DeviceArray.from_numpy(y0)
    cm.dense.matvec(A, cm.DeviceArray.from_numpy(x), out=out, mode=cm.capi.ADD)
    np.testing.assert_array_equal(out.numpy(), y0 + a @ x)
    cm.dense.matvec(A, cm.DeviceArray.from_numpy(x), out=out, mode=cm.capi.SUB)
    np.testing.assert_array_equal(out.numpy(), y0)


@pytest.mark.parametrize("shape", [(4, 9), (260, 3000), (1024, 700), (2, 513)])
def test_matvec_16byte_column_major(gpu, shape):
    """Column-major A with rows a multiple of the vector width (4 for 4-byte
    types, 2 for FP64) takes the 16-byte path: int32 exact (also += / -= and
    a batched A), FP32 / FP64 within the product tolerances; ragged last row
    tile, one and several column ranges."""
    cm = gpu
    rng = np.random.default_rng(shape[0] + 7 * shape[1])
    rows, cols = shape
    B = 2
    a = rng.integers(-30, 30, (B, rows, cols)).astype(np.int32)
    x = rng.integers(-30, 30, (B, cols, 1)).astype(np.int32)
    A = cm.DeviceArray.from_numpy(a)
    X = cm.DeviceArray.from_numpy(x)
    y = cm.dense.matvec(A, X).numpy().reshape(B, rows)
    exp = np.einsum("brc,bc->br", a.astype(np.int64), x[:, :, 0])
    np.testing.assert_array_equal(y, exp)
    y0 = rng.integers(-9, 9, (B, rows, 1)).astype(np.int32)
    out = cm.DeviceArray.from_numpy(y0)
    cm.dense.matvec(A, X, out=out, mode=cm.capi.ADD)
    np.testing.assert_array_equal(out.numpy().reshape(B, rows), y0[:, :, 0] + exp)
    cm.dense.matvec(A, X, out=out, mode=cm.capi.SUB)
    np.testing.assert_array_equal(out.numpy().reshape(B, rows), y0[:, :, 0])
    for dt, tol in ((np.float64, 1e-12), (np.float32, 1e-5)):
        af = rng.uniform(-1, 1, (rows, cols)).astype(dt)
        xf = rng.uniform(-1, 1, cols).astype(dt)
        yf = cm.dense.matvec(cm.DeviceArray.from_numpy(af), cm.DeviceArray.from_numpy(xf)).numpy()
        ref = af.astype(np.float64) @ xf.astype(np.float64)
        scale = np.abs(af).astype(np.float64) @ np.abs(xf).astype(np.float64) + 1e-300
        assert np.all(np.abs(yf - ref) <= tol * scale), (dt, shape)


def test_matvec_batched_and_broadcast(gpu):
    """Batched A with batched x, batch-1 A broadcast over batched x, batched A
    over a batch-1 x (cuMat's batched products)."""
    cm = gpu
    rng = np.random.default_rng(11)
    B, rows, cols = 3, 70, 45
    a = rng.integers(-20, 20, (B, rows, cols)).astype(np.int32)
    x = rng.integers(-20, 20, (B, cols, 1)).astype(np.int32)
    y = cm.dense.matvec(cm.DeviceArray.from_numpy(a), cm.DeviceArray.from_numpy(x)).numpy()
    np.testing.assert_array_equal(y.reshape(B, rows), np.einsum("brc,bc->br", a, x[:, :, 0]))
    y = cm.dense.matvec(cm.DeviceArray.from_numpy(a[0]), cm.DeviceArray.from_numpy(x)).numpy()
    np.testing.assert_array_equal(y.reshape(B, rows), np.einsum("rc,bc->br", a[0], x[:, :, 0]))
    y = cm.dense.matvec(cm.DeviceArray.from_numpy(a), cm.DeviceArray.from_numpy(x[0])).numpy()
    np.testing.assert_array_equal(y.reshape(B, rows), np.einsum("brc,c->br", a, x[0, :, 0]))


def test_matvec_rejects_bad_shapes(gpu):
    cm = gpu
    A = cm.DeviceArray.from_numpy(np.ones((4, 5), np.float32))
    with pytest.raises(cm.CumatArgError):
        cm.dense.matvec(A, cm.DeviceArray.from_numpy(np.ones(4, np.float32)))
    with pytest.raises(TypeError):
        cm.dense.matvec(A, cm.DeviceArray.from_numpy(np.ones(5, np.float64)))


def test_dense_inverse_diagonal(gpu):
    cm = gpu
    a = np.arange(1, 10, dtype=np.float64).reshape(3, 3)
    a[1, 1] = 0.0
    inv = cm.dense.inverse_diagonal(cm.DeviceArray.from_numpy(a)).numpy()
    np.testing.assert_array_equal(inv, [1.0, 1.0, 1.0 / 9.0])


@pytest.mark.parametrize("size", [5, 10, 50, 100])
def test_cg_dense_early_out_rhs(gpu, size):
    """TestConjugateGradient.cu:11-25: b = 0 -> 0 iterations, error 0, x = 0."""
    cm = gpu
    a = np.random.default_rng(size).uniform(0, 1, (size, size)).astype(np.float32)
    cg = cm.ConjugateGradient(cm.DeviceArray.from_numpy(a))
    x = cg.solve(cm.DeviceArray.from_numpy(np.zeros(size, np.float32))).numpy()
    assert cg.iterations() == 0
    assert cg.error() == 0
    assert np.sum(x.astype(np.float64) ** 2) == 0


@pytest.mark.parametrize("size", [5, 10, 50, 100])
def test_cg_dense_early_out_guess(gpu, size):
    """TestConjugateGradient.cu:27-44: b = A xTruth, solveWithGuess(b, xTruth)
    -> 0 iterations, error ~ 0, x = xTruth."""
    cm = gpu
    rng = np.random.default_rng(100 + size)
    a = rng.uniform(0, 1, (size, size)).astype(np.float32)
    xt = rng.uniform(0, 1, size).astype(np.float32)
    A = cm.DeviceArray.from_numpy(a)
    b = cm.dense.matvec(A, cm.DeviceArray.from_numpy(xt))
    cg = cm.ConjugateGradient(A)
    x = cg.solveWithGuess(b, cm.DeviceArray.from_numpy(xt)).numpy()
    assert cg.iterations() == 0
    assert abs(cg.error()) <= 1e-10
    np.testing.assert_array_equal(x, xt)


@pytest.mark.parametrize("size", [5, 10, 50, 100])
@pytest.mark.parametrize("row_major", [False, True])
def test_cg_dense_solve(gpu, size, row_major):
    """TestConjugateGradient.cu:46-71: FP64, b = A xTruth, maxIterations
    10*size: converges (error <= tolerance = epsilon) to xTruth within 1e-4
    relative. The reference fills A = U[0,1) + size*I, which is not
    symmetric; CG reaches epsilon on it only when its skew part happens to
    be small (the oracle's CG on numpy draws of that matrix stops at 50 and
    100 iterations for sizes 5 and 10 with errors 1e-12 and 1e-7), so the
    test takes the symmetric part (U + U^T)/2 + size*I."""
    cm = gpu
    rng = np.random.default_rng(200 + size)
    u = rng.uniform(0, 1, (size, size))
    a = (u + u.T) / 2 + size * np.eye(size)
    xt = rng.uniform(0, 1, size)
    A = _dense(cm, a, row_major)
    b = cm.dense.matvec(A, cm.DeviceArray.from_numpy(xt))
    cg = cm.ConjugateGradient(A).setMaxIterations(10 * size)
    x = cg.solve(b).numpy()
    assert cg.iterations() > 0
    assert cg.error() > 0
    assert cg.iterations() < cg.maxIterations()
    assert cg.error() <= cg.tolerance()
    np.testing.assert_allclose(x, xt, rtol=1e-4)


def test_cg_dense_solve_batched(gpu):
    """TestConjugateGradient.cu:73-99: size 10, three right-hand sides
    (symmetric part of the reference's matrix, as above)."""
    cm = gpu
    size, batches = 10, 3
    rng = np.random.default_rng(300)
    u = rng.uniform(0, 1, (size, size))
    a = (u + u.T) / 2 + size * np.eye(size)
    xt = rng.uniform(0, 1, (batches, size, 1))
    A = cm.DeviceArray.from_numpy(a)
    b = cm.dense.matvec(A, cm.DeviceArray.from_numpy(xt))
    assert b.batches == batches
    cg = cm.ConjugateGradient(A).setMaxIterations(10 * size)
    x = cg.solve(b).numpy().reshape(batches, size)
    assert cg.iterations() > 0
    assert cg.error() > 0
    assert cg.iterations() < cg.maxIterations()
    assert cg.error() <= cg.tolerance()
    np.testing.assert_allclose(x, xt[:, :, 0], rtol=1e-4)


@pytest.mark.parametrize("n", [256, 301])
def test_cg_dense_float32_both_layouts(gpu, n):
    """FP32 dense CG with row-major A (16-byte row loads when n % 4 == 0)
    and column-major A: both converge to tolerance, to the FP64 solution
    within 1e-4, and within a few iterations of each other."""
    cm = gpu
    rng = np.random.default_rng(n)
    m = rng.standard_normal((n, n))
    a = (m @ m.T / n + 0.5 * np.eye(n)).astype(np.float32)
    b = rng.uniform(-1, 1, n).astype(np.float32)
    x64 = np.linalg.solve(a.astype(np.float64), b.astype(np.float64))
    its = []
    for rm in (True, False):
        cg = cm.ConjugateGradient(_dense(cm, a, rm)).setTolerance(1e-5).setMaxIterations(2000)
        x = cg.solve(cm.DeviceArray.from_numpy(b)).numpy()
        assert cg.error() <= 1e-5
        assert np.linalg.norm(x - x64) / np.linalg.norm(x64) <= 1e-4
        its.append(cg.iterations())
    assert abs(its[0] - its[1]) <= 4, its


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_cg_dense_matches_oracle_on_csr_form(gpu, orc, dt):
    """Dense SPD matrix: the dense CG against the oracle's CG on the same
    matrix stored as a full CSR pattern (identical method; the products sum
    in different orders, so FP32 iteration counts may differ by a few)."""
    cm = gpu
    n = 300
    rng = np.random.default_rng(17)
    m = rng.standard_normal((n, n))
    a = (m @ m.T / n + 0.5 * np.eye(n)).astype(dt)
    b = rng.uniform(-1, 1, n).astype(dt)
    JA = (np.arange(n + 1) * n).astype(np.int32)
    IA = np.tile(np.arange(n, dtype=np.int32), n)
    V = a.reshape(-1).copy()  # row-major = CSR value order
    tol = 1e-10 if dt == np.float64 else 1e-5
    cg = cm.ConjugateGradient(cm.DeviceArray.from_numpy(a)).setTolerance(tol).setMaxIterations(2000)
    x = cg.solve(cm.DeviceArray.from_numpy(b)).numpy()
    x_o, it_o, err_o = orc.pcg(JA, IA, V, b, n, tol=tol, max_iter=2000)
    assert cg.error() <= tol
    slack = 1 if dt == np.float64 else 4
    assert abs(cg.iterations() - it_o) <= slack, (cg.iterations(), it_o)
    x64 = np.linalg.solve(a.astype(np.float64), b.astype(np.float64))
    assert np.linalg.norm(x - x64) / np.linalg.norm(x64) <= (1e-8 if dt == np.float64 else 1e-4)
    assert np.linalg.norm(x - x_o) / np.linalg.norm(x_o) <= (1e-9 if dt == np.float64 else 1e-4)
