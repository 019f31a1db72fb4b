"""GPU tests of the product whose sparse matrix is a component-wise expression
(expr.sparseView<CSR|ELLPACK>(pattern) * x): the reference's fixtures
(tests/TestSparseMultOp.cu:104-153 for CSR, the ELLPACK counterparts in
:329-375) with its one-launch counter assertions, and larger matrices
against the oracle's CSR / ELLPACK products of the same values computed on
the host. Integer-valued data keep every FP product and sum exact, so the
comparisons are bit for bit."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_tests.json")))


def _fixture(cm):
    g = GOLD["csr_4x5"]
    pat = cm.SparsityPattern(4, 5, g["JA"], g["IA"])
    A = cm.SparseMatrix(pat, np.array(g["vals"], np.float32))
    b = cm.DeviceArray.from_numpy(np.array(g["b"], np.float32))
    return A, pat, b, np.array(g["y_times_2A"], np.float32)


def test_csr_expression_product_direct_golden(gpu):
    """(2 * A.direct()).sparseView<CSR>(A.getSparsityPattern()) * b ->
    [44, 38, -20, 66] in ONE evaluation (EvalAny == 1, EvalMatmulSparse == 1)."""
    cm = gpu
    A, pat, b, exp = _fixture(cm)
    cm.reset_counters()
    y = cm.sparse_view(2 * A.direct(), pat) * b
    c = cm.counters()
    assert c["EvalAny"] == 1 and c["EvalMatmulSparse"] == 1 and c["KernelLaunch"] == 1, c
    assert np.array_equal(y.numpy(), exp)


def test_csr_expression_product_search_golden(gpu):
    """(2 * A).sparseView<CSR>(A.getSparsityPattern()) * b: the matrix used
    without direct() on its own pattern."""
    cm = gpu
    A, pat, b, exp = _fixture(cm)
    cm.reset_counters()
    y = cm.sparse_view(2 * A, pat) * b
    c = cm.counters()
    assert c["EvalAny"] == 1 and c["EvalMatmulSparse"] == 1, c
    assert np.array_equal(y.numpy(), exp)


def test_csr_expression_rejects_other_pattern(gpu):
    cm = gpu
    A, pat, b, _ = _fixture(cm)
    g = GOLD["csr_4x5"]
    other = cm.SparsityPattern(4, 5, [0, 1, 2, 3, 4], [0, 1, 2, 3])
    with pytest.raises(cm.capi.CumatArgError):
        cm.sparse_view(2 * A.direct(), other)
    # a dense assignment cannot read stored values directly
    with pytest.raises(TypeError):
        cm.assign(cm.DeviceArray(4, 5, dtype=np.float32), A.direct() * 2.0)
    del g


def _poisson_with_long_rows(g, long_rows, width, rng):
    """2D 5-point pattern plus a few rows with 'width' extra entries each
    (chunks above the lane budget take the per-lane row walk)."""
    from cumat_amd.problems import poisson2d
    JA, IA, _ = poisson2d(g, 0.0, np.float64)
    n = g * g
    rows = []
    for r in range(n):
        cols = list(IA[JA[r]:JA[r + 1]])
        if r in long_rows:
            extra = rng.choice(n, width, replace=False)
            cols = sorted(set(cols) | set(int(c) for c in extra))
        rows.append(cols)
    JA2 = np.zeros(n + 1, np.int32)
    JA2[1:] = np.cumsum([len(c) for c in rows])
    IA2 = np.concatenate([np.array(c, np.int32) for c in rows])
    return JA2, IA2, n


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.int32])
@pytest.mark.parametrize("long_rows", [False, True])
def test_csr_expression_product_against_oracle(gpu, orc, dtype, long_rows):
    """v = 3*A + B - r(row) * c(col) + 1 over two same-pattern batched
    matrices, a column vector r (depends on the row) and a row vector c
    (depends on the column), batched x; modes =, += and -=."""
    cm = gpu
    rng = np.random.default_rng(11)
    g = 150
    lr = set(rng.choice(g * g, 40, replace=False).tolist()) if long_rows else set()
    JA, IA, n = _poisson_with_long_rows(g, lr, 700, rng)
    nnz = IA.size
    B = 2
    va = rng.integers(-5, 6, nnz * B).astype(dtype)
    vb = rng.integers(-5, 6, nnz * B).astype(dtype)
    r = rng.integers(-3, 4, n).astype(dtype)
    c = rng.integers(-3, 4, n).astype(dtype)
    x = rng.integers(-4, 5, n * B).astype(dtype)
    pat = cm.SparsityPattern(n, n, JA, IA)
    A = cm.SparseMatrix(pat, va, batches=B)
    Bm = cm.SparseMatrix(pat, vb, batches=B)
    rd = cm.DeviceArray.from_numpy(r, rows=n, cols=1, batches=1)
    cd = cm.DeviceArray.from_numpy(c, rows=1, cols=n, batches=1, row_major=True)
    xd = cm.DeviceArray.from_numpy(x, rows=n, cols=1, batches=B)
    one = 1.0 if dtype != np.int32 else 1
    three = 3.0 if dtype != np.int32 else 3
    e = A.direct() * three + Bm.direct() - rd.expr() * cd.expr() + one
    rows = np.repeat(np.arange(n), np.diff(JA))
    vals = np.concatenate([(3 * va[k * nnz:(k + 1) * nnz] + vb[k * nnz:(k + 1) * nnz] - r[rows] * c[IA] + 1)
                           for k in range(B)]).astype(dtype)
    ref = orc.csr_spmv(JA, IA, vals, x, n, n, batches=B, vals_batched=True, x_batched=True)
    cm.reset_counters()
    y = cm.sparse_view(e, pat).matvec(xd)
    assert cm.counters()["KernelLaunch"] == 1
    assert np.array_equal(y.numpy(), ref)
    y0 = rng.integers(-100, 101, n * B).astype(dtype)
    for mode in (cm.capi.ADD, cm.capi.SUB):
        yd = cm.DeviceArray.from_numpy(y0.copy(), rows=n, cols=1, batches=B)
        cm.sparse_view(e, pat).matvec(xd, out=yd, mode=mode)
        refm = orc.csr_spmv(JA, IA, vals, x, n, n, batches=B, vals_batched=True, x_batched=True, y=y0.copy(),
                            mode=mode)
        assert np.array_equal(yd.numpy(), refm), mode


def test_csr_expression_product_full_size_random_x(gpu, orc):
    """2048^2 Poisson, v = 2*A: random FP32 x. The products are formed and
    summed in CSR order like the oracle's, so the result is bit-identical."""
    cm = gpu
    from cumat_amd.problems import poisson2d
    g = 2048
    JA, IA, V = poisson2d(g)
    n = g * g
    x = np.random.default_rng(2).uniform(-1, 1, n).astype(np.float32)
    pat = cm.SparsityPattern(n, n, JA, IA)
    A = cm.SparseMatrix(pat, V)
    y = cm.sparse_view(2 * A.direct(), pat) * cm.DeviceArray.from_numpy(x)
    ref = orc.csr_spmv(JA, IA, (2 * V).astype(np.float32), x, n, n)
    assert np.array_equal(y.numpy(), ref)


def test_ell_expression_product_golden_and_batched(gpu, orc):
    """(2 * A.direct()).sparseView<ELLPACK>(pattern) * b on the reference's
    ELLPACK fixture -> [44, 38, -20, 66] in one evaluation; a batched matrix
    (scales 1, 2, 3) with an expression adding a row-dependent leaf."""
    cm = gpu
    g = GOLD["ell_4x5"]
    E = cm.ELLMatrix(4, 5, g["indices_rows"], np.array(g["values_rows"], np.float32))
    b = cm.DeviceArray.from_numpy(np.array(g["b"], np.float32))
    cm.reset_counters()
    y = cm.sparse_view(2 * E.direct(), E) * b
    c = cm.counters()
    assert c["EvalAny"] == 1 and c["EvalMatmulSparse"] == 1 and c["KernelLaunch"] == 1, c
    assert np.array_equal(y.numpy(), np.array(g["y_times_2A"], np.float32))
    vals = np.stack([np.array(g["values_rows"], np.float32) * s for s in (1, 2, 3)])
    E3 = cm.ELLMatrix(4, 5, g["indices_rows"], vals)
    r = np.array([1, -2, 3, 0], np.float32)
    rd = cm.DeviceArray.from_numpy(r)
    y3 = cm.sparse_view(E3.direct() + rd.expr(), E3) * b
    ind = np.array(g["indices_rows"])
    exp = []
    for k, s in enumerate((1, 2, 3)):
        vv = vals[k] + r[:, None]
        exp.append([sum(vv[i, j] * g["b"][ind[i, j]] for j in range(3) if ind[i, j] >= 0) for i in range(4)])
    assert np.array_equal(y3.numpy(), np.array(exp, np.float32).reshape(-1))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_ell_expression_product_poisson(gpu, orc, dtype):
    """Width-7 3D Laplacian in ELLPACK, v = A*2 - 1 + c(col) with integer data:
    bit-exact against the oracle's ELLPACK product of the host-computed
    values (padding slots never read)."""
    cm = gpu
    from cumat_amd.problems import grid_laplacian
    rng = np.random.default_rng(5)
    P = grid_laplacian((30, 30, 30), lambda s: rng.integers(1, 4, s).astype(np.float64), edges=False)
    n = P["n"]
    cnt = np.diff(P["JA"])
    ind = -np.ones((n, 7), np.int32)
    val = np.zeros((n, 7), dtype)
    slot = np.arange(P["nnz"]) - np.repeat(P["JA"][:-1], cnt)
    rows = np.repeat(np.arange(n), cnt)
    ind[rows, slot] = P["IA"]
    val[rows, slot] = P["vals"]
    E = cm.ELLMatrix(n, n, ind, val)
    c = rng.integers(-2, 3, n).astype(dtype)
    cd = cm.DeviceArray.from_numpy(c, rows=1, cols=n, batches=1, row_major=True)
    x = rng.integers(-6, 7, n).astype(dtype)
    y = cm.sparse_view(E.direct() * 2.0 - 1.0 + cd.expr(), E) * cm.DeviceArray.from_numpy(x)
    v2 = np.where(ind >= 0, val * 2 - 1 + c[np.maximum(ind, 0)], 0).astype(dtype)
    ref = orc.ell_spmv(np.asfortranarray(ind).reshape(-1, order="F"), np.asfortranarray(v2).reshape(-1, order="F"),
                       x, n, n, 7)
    assert np.array_equal(y.numpy(), ref)
