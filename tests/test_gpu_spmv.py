"""GPU parity of the CSR / ELLPACK SpMV against the reference's fixtures and the
CPU oracle (restating cuMat/src/SparseProductEvaluation.h:43-88,194-226)."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_tests.json")))


def _csr4x5(cm, batches=1, vscales=(1,)):
    g = GOLD["csr_4x5"]
    vals = np.concatenate([np.array(g["vals"], np.float32) * s for s in vscales])
    pat = cm.SparsityPattern(4, 5, g["JA"], g["IA"])
    return cm.SparseMatrix(pat, vals, batches=len(vscales)), g


def test_csr_golden(gpu):
    cm = gpu
    A, g = _csr4x5(cm)
    b = cm.DeviceArray.from_numpy(np.array(g["b"], np.float32))
    cm.reset_counters()
    y = A.matvec(b)
    c = cm.counters()
    assert c["EvalAny"] == 1 and c["EvalMatmulSparse"] == 1  # one kernel per product
    assert np.array_equal(y.numpy(), np.array(g["y"], np.float32))


@pytest.mark.parametrize("case", ["matrix", "vector", "both"])
def test_csr_golden_batched(gpu, case):
    cm = gpu
    g = GOLD["csr_4x5"]
    if case == "matrix":
        ms, vs, es = g["batched_matrix_scales"], [1], g["batched_matrix_expected_scales"]
    elif case == "vector":
        ms, vs, es = [1], g["batched_vector_scales"], g["batched_vector_expected_scales"]
    else:
        ms, vs, es = g["batched_both_matrix_scales"], g["batched_both_vector_scales"], g["batched_both_expected_scales"]
    A, _ = _csr4x5(cm, vscales=ms)
    b = np.concatenate([np.array(g["b"], np.float32) * s for s in vs])
    bd = cm.DeviceArray.from_numpy(b, rows=5, cols=1, batches=len(vs))
    y = A.matvec(bd).numpy()
    exp = np.concatenate([np.array(g["y"], np.float32) * s for s in es])
    assert np.array_equal(y, exp)


def test_ell_golden(gpu):
    cm = gpu
    g = GOLD["ell_4x5"]
    E = cm.ELLMatrix(4, 5, g["indices_rows"], np.array(g["values_rows"], np.float32))
    y = E.matvec(cm.DeviceArray.from_numpy(np.array(g["b"], np.float32))).numpy()
    assert np.array_equal(y, np.array(g["y"], np.float32))


@pytest.mark.parametrize("g", [1, 7, 64, 333])
def test_poisson_integer_x_bitexact(gpu, orc, g):
    """Integer-valued x makes every FP32 product and partial sum exact: the
    GPU result must equal the oracle bit for bit."""
    cm = gpu
    from cumat_amd.problems import poisson2d
    JA, IA, V = poisson2d(g)
    n = g * g
    x = np.random.default_rng(g).integers(-50, 51, n).astype(np.float32)
    A = cm.SparseMatrix.from_csr(n, n, JA, IA, V)
    y = A.matvec(cm.DeviceArray.from_numpy(x)).numpy()
    assert np.array_equal(y, orc.csr_spmv(JA, IA, V, x, n, n))
    # closed form: A * 1 = number of missing neighbours (Dirichlet rows)
    y1 = A.matvec(cm.DeviceArray.from_numpy(np.ones(n, np.float32))).numpy()
    JA2, IA2, V2 = orc.poisson2d(g)
    assert np.array_equal(JA, JA2) and np.array_equal(IA, IA2)
    assert np.array_equal(y1, 5.0 - np.diff(JA).astype(np.float32))


def test_poisson_2048_full_size(gpu, orc):
    """The benchmark matrix (n = 4,194,304, nnz = 20,963,328) at full size:
    bit-exact against the oracle for integer x, and the oracle order is
    reproduced exactly for random x (products are summed in CSR order)."""
    cm = gpu
    from cumat_amd.problems import poisson2d
    g = 2048
    JA, IA, V = poisson2d(g)
    n = g * g
    assert IA.size == 20963328
    A = cm.SparseMatrix.from_csr(n, n, JA, IA, V)
    rng = np.random.default_rng(5)
    xi = rng.integers(-1000, 1001, n).astype(np.float32)
    assert np.array_equal(A.matvec(cm.DeviceArray.from_numpy(xi)).numpy(), orc.csr_spmv(JA, IA, V, xi, n, n))
    xr = rng.uniform(-1, 1, n).astype(np.float32)
    y = A.matvec(cm.DeviceArray.from_numpy(xr)).numpy()
    assert np.array_equal(y, orc.csr_spmv(JA, IA, V, xr, n, n))


@pytest.mark.parametrize("mode", [1, 2])
def test_spmv_modes(gpu, orc, mode):
    cm = gpu
    from cumat_amd.problems import poisson2d
    JA, IA, V = poisson2d(20)
    n = 400
    rng = np.random.default_rng(3)
    x = rng.uniform(-1, 1, n).astype(np.float32)
    y0 = rng.uniform(-1, 1, n).astype(np.float32)
    A = cm.SparseMatrix.from_csr(n, n, JA, IA, V)
    yd = cm.DeviceArray.from_numpy(y0)
    A.matvec(cm.DeviceArray.from_numpy(x), out=yd, mode=mode)
    exp = orc.csr_spmv(JA, IA, V, x, n, n, y=y0.copy(), mode=mode)
    assert np.array_equal(yd.numpy(), exp)


def test_spmv_ragged_and_empty_rows(gpu, orc):
    """Empty rows, one very long row (the wavefront-per-row fallback) and a
    rectangular matrix."""
    cm = gpu
    rng = np.random.default_rng(11)
    rows, cols = 3000, 9000
    lens = rng.integers(0, 6, rows)
    lens[::7] = 0
    lens[1500] = 8000  # forces the long-row path for its block
    JA = np.zeros(rows + 1, np.int32)
    JA[1:] = np.cumsum(lens)
    IA = np.concatenate([np.sort(rng.choice(cols, l, replace=False)) for l in lens]).astype(np.int32)
    V = rng.uniform(-1, 1, IA.size).astype(np.float32)
    x = rng.uniform(-1, 1, cols).astype(np.float32)
    A = cm.SparseMatrix.from_csr(rows, cols, JA, IA, V)
    yd = cm.DeviceArray.from_numpy(np.full(rows, 7.0, np.float32))  # garbage to be overwritten
    y = A.matvec(cm.DeviceArray.from_numpy(x), out=yd).numpy()
    exp = orc.csr_spmv(JA, IA, V, x, rows, cols)
    assert np.all(y[lens == 0] == 0.0)
    np.testing.assert_allclose(y, exp, rtol=1e-5, atol=1e-5 * np.abs(exp).max())
    short = np.ones(rows, bool)
    short[1024:1536] = False  # rows of the fallback block
    assert np.array_equal(y[short], exp[short])


@pytest.mark.parametrize("seed", [0, 1])
def test_spmv_staging_boundaries_batched(gpu, orc, seed):
    """Rows of 0..16 entries put the 64-row chunks on both sides of the
    staging capacity (512 entries) with every start offset mod 4 (16-byte
    and element-wise staging, and the wavefront-per-row path); three value
    batches with nnz % 4 != 0 make every batch but the first misaligned for
    16-byte loads. Integer-valued operands: exact in FP32, so bit-exact."""
    cm = gpu
    rng = np.random.default_rng(100 + seed)
    rows, cols, B = 64 * 150 + 17, 5000, 3
    lens = rng.integers(0, 17, rows)
    JA = np.zeros(rows + 1, np.int32)
    JA[1:] = np.cumsum(lens)
    if JA[-1] % 4 == 0:
        lens[-1] += 1
        JA[1:] = np.cumsum(lens)
    IA = np.concatenate([np.sort(rng.choice(cols, l, replace=False)) for l in lens]).astype(np.int32)
    nnz = IA.size
    assert nnz % 4 != 0
    V = rng.integers(-8, 9, nnz * B).astype(np.float32)
    x = rng.integers(-8, 9, cols * B).astype(np.float32)
    A = cm.SparseMatrix(cm.SparsityPattern(rows, cols, JA, IA), V, batches=B)
    y = A.matvec(cm.DeviceArray.from_numpy(x, rows=cols, cols=1, batches=B)).numpy()
    exp = orc.csr_spmv(JA, IA, V, x, rows, cols, batches=B, vals_batched=True, x_batched=True)
    assert np.array_equal(y.reshape(-1), exp.reshape(-1))


def test_spmv_int32_and_f64(gpu, orc):
    cm = gpu
    from cumat_amd.problems import poisson2d
    JA, IA, V = poisson2d(50)
    n = 2500
    rng = np.random.default_rng(2)
    Vi = rng.integers(-1000, 1000, IA.size).astype(np.int32)
    xi = rng.integers(-1000, 1000, n).astype(np.int32)
    Ai = cm.SparseMatrix.from_csr(n, n, JA, IA, Vi)
    assert np.array_equal(Ai.matvec(cm.DeviceArray.from_numpy(xi)).numpy(), orc.csr_spmv(JA, IA, Vi, xi, n, n))
    Vd = rng.uniform(-1, 1, IA.size)
    xd = rng.uniform(-1, 1, n)
    Ad = cm.SparseMatrix.from_csr(n, n, JA, IA, Vd)
    assert np.array_equal(Ad.matvec(cm.DeviceArray.from_numpy(xd)).numpy(), orc.csr_spmv(JA, IA, Vd, xd, n, n))


def test_spmv_bf16_values(gpu, orc):
    """BF16 matrix storage with FP32 vectors: exact against the oracle on the
    bf16-rounded values."""
    cm = gpu
    from cumat_amd.problems import poisson2d
    JA, IA, V = poisson2d(40)
    n = 1600
    rng = np.random.default_rng(9)
    V = (V * rng.uniform(0.5, 2, V.size)).astype(np.float32)
    bits = V.view(np.uint32)
    bf = ((bits + 0x7FFF + ((bits >> 16) & 1)) >> 16).astype(np.uint16)
    Vr = (bf.astype(np.uint32) << 16).view(np.float32)
    x = rng.uniform(-1, 1, n).astype(np.float32)
    pat = cm.SparsityPattern(n, n, JA, IA)
    A = cm.SparseMatrix(pat, cm.DeviceArray.from_numpy(bf))
    y = A.matvec(cm.DeviceArray.from_numpy(x)).numpy()
    assert np.array_equal(y, orc.csr_spmv(JA, IA, Vr, x, n, n))


def test_spmv_dimension_errors(gpu):
    cm = gpu
    A, g = _csr4x5(cm)
    with pytest.raises(cm.CumatArgError):
        A.matvec(cm.DeviceArray.from_numpy(np.ones(4, np.float32)))


@pytest.mark.parametrize("dtype", [np.int32, np.float32])
def test_csr_cwise_eval_both_kernels(gpu, dtype):
    """Component-wise programs evaluated at the stored entries of a 2D
    Poisson pattern (cuMat/src/SparseEvaluation.h:14-34): programs whose
    leaves need no row coordinate (values and scalars; a row vector that
    depends on the column only) run one thread per entry, a column vector
    (depends on the row) runs one thread per row. int32 exact against numpy,
    FP32 to 1 ulp-level tolerance; plus a compound mode."""
    cm = gpu
    from cumat_amd.problems import poisson2d
    g = 300
    JA, IA, V = poisson2d(g)
    n = g * g
    rng = np.random.default_rng(41)
    vals = (rng.integers(-50, 50, IA.size) if dtype == np.int32 else rng.uniform(-1, 1, IA.size)).astype(dtype)
    A = cm.SparseMatrix.from_csr(n, n, JA, IA, vals)
    rows = np.repeat(np.arange(n), np.diff(JA))
    rv = (rng.integers(-9, 9, n) if dtype == np.int32 else rng.uniform(-1, 1, n)).astype(dtype)
    cv = (rng.integers(-9, 9, n) if dtype == np.int32 else rng.uniform(-1, 1, n)).astype(dtype)
    R = cm.DeviceArray.from_numpy(rv, rows=1, cols=n, batches=1)
    Cv = cm.DeviceArray.from_numpy(cv, rows=n, cols=1, batches=1)
    two, one = (2, 1) if dtype == np.int32 else (2.0, 1.0)

    def check(got, exp):
        if dtype == np.int32:
            np.testing.assert_array_equal(got, exp)
        else:
            np.testing.assert_allclose(got, exp, rtol=1e-6, atol=1e-6)

    A.assign(A.direct() * two + one)                  # values and scalars only
    v1 = (vals * dtype(two) + dtype(one)).astype(dtype)
    check(A.values.numpy(), v1)
    A.assign(A.direct() * R.expr())                   # depends on the column
    v2 = (v1 * rv[IA]).astype(dtype)
    check(A.values.numpy(), v2)
    A.assign(A.direct() * Cv.expr())                  # depends on the row
    v3 = (v2 * cv[rows]).astype(dtype)
    check(A.values.numpy(), v3)
    A.assign(R.expr() - A.direct(), mode=cm.capi.ADD)  # compound, column leaf
    v4 = (v3 + (rv[IA] - v3)).astype(dtype)
    check(A.values.numpy(), v4)


def test_csr_cwise_eval_broadcast_single_batch_values(gpu):
    """A one-batch sparse matrix used in a program assigned to a two-batch
    matrix of the same pattern is broadcast over the batches (its value leaf
    has batch stride 0): values-only program (dense route) and a
    row-dependent program (row kernel), int32 exact."""
    cm = gpu
    from cumat_amd.problems import poisson2d
    g = 40
    JA, IA, V = poisson2d(g)
    n = g * g
    rng = np.random.default_rng(43)
    v1 = rng.integers(-50, 50, IA.size).astype(np.int32)
    pat = cm.SparsityPattern(n, n, JA, IA)
    S1 = cm.SparseMatrix(pat, v1, batches=1)
    S2 = cm.SparseMatrix(pat, np.zeros(2 * IA.size, np.int32), batches=2)
    S2.assign(S1.direct() * 3)
    np.testing.assert_array_equal(S2.values.numpy().reshape(2, -1), np.stack([v1 * 3, v1 * 3]))
    cv = rng.integers(-9, 9, n).astype(np.int32)
    Cv = cm.DeviceArray.from_numpy(cv, rows=n, cols=1, batches=1)
    S2.assign(S1.direct() * Cv.expr(), mode=cm.capi.ADD)
    rows = np.repeat(np.arange(n), np.diff(JA))
    exp = v1 * 3 + v1 * cv[rows]
    np.testing.assert_array_equal(S2.values.numpy().reshape(2, -1), np.stack([exp, exp]))


def test_csr_cwise_eval_batched_column_leaf(gpu):
    """A column-dependent leaf with its own batch stride (a batched row
    vector) over a two-batch CSR matrix: evaluated per stored entry; and a
    one-batch row vector over the same matrix (gathered along IA, then a
    dense program over the values). int32 exact."""
    cm = gpu
    from cumat_amd.problems import poisson2d
    g = 50
    JA, IA, V = poisson2d(g)
    n = g * g
    rng = np.random.default_rng(47)
    v = rng.integers(-50, 50, (2, IA.size)).astype(np.int32)
    pat = cm.SparsityPattern(n, n, JA, IA)
    S = cm.SparseMatrix(pat, v.reshape(-1), batches=2)
    r2 = rng.integers(-9, 9, (2, 1, n)).astype(np.int32)
    R2 = cm.DeviceArray.from_numpy(r2)  # 1 x n x 2
    S.assign(S.direct() * R2.expr())
    exp = v * r2[:, 0, :][:, IA]
    np.testing.assert_array_equal(S.values.numpy().reshape(2, -1), exp)
    r1 = rng.integers(-9, 9, n).astype(np.int32)
    R1 = cm.DeviceArray.from_numpy(r1, rows=1, cols=n, batches=1)
    S.assign(S.direct() - R1.expr(), mode=cm.capi.ADD)
    exp = exp + (exp - r1[IA][None, :])
    np.testing.assert_array_equal(S.values.numpy().reshape(2, -1), exp)


# ------------------------------------------------------------ ELLPACK
def _ell_from_csr(JA, IA, V, width):
    """rows x width ELLPACK index / value matrices of a CSR matrix (entries in
    CSR order, -1 padding with value 666 as in the reference's fixture)."""
    n = JA.size - 1
    ind = -np.ones((n, width), np.int32)
    val = np.full((n, width), 666, V.dtype)
    cnt = np.diff(JA)
    slot = np.arange(IA.size) - np.repeat(JA[:-1], cnt)
    rows = np.repeat(np.arange(n), cnt)
    ind[rows, slot] = IA
    val[rows, slot] = V
    return ind, val


@pytest.mark.parametrize("case", ["matrix", "vector", "both"])
def test_ell_golden_batched(gpu, case):
    """tests/TestSparseMultOp.cu:377-519: the 4x5 fixture with a batched
    matrix (scales 1,2,3), a batched vector (1,2,3) or both (1,2,-1 x 1,2,3)."""
    cm = gpu
    g = GOLD["ell_4x5"]
    c = GOLD["csr_4x5"]
    if case == "matrix":
        ms, vs, es = c["batched_matrix_scales"], [1], c["batched_matrix_expected_scales"]
    elif case == "vector":
        ms, vs, es = [1], c["batched_vector_scales"], c["batched_vector_expected_scales"]
    else:
        ms, vs, es = c["batched_both_matrix_scales"], c["batched_both_vector_scales"], c["batched_both_expected_scales"]
    vals = np.stack([np.array(g["values_rows"], np.float32) * s for s in ms])
    E = cm.ELLMatrix(4, 5, g["indices_rows"], vals)
    b = np.concatenate([np.array(g["b"], np.float32) * s for s in vs])
    cm.reset_counters()
    y = E.matvec(cm.DeviceArray.from_numpy(b, rows=5, cols=1, batches=len(vs))).numpy()
    cnt = cm.counters()
    assert cnt["EvalAny"] == 1 and cnt["EvalMatmulSparse"] == 1
    exp = np.concatenate([np.array(g["y"], np.float32) * s for s in es])
    assert np.array_equal(y, exp)


@pytest.mark.parametrize("dims,width", [((300, 300), 5), ((40, 40, 40), 7)])
@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.int32])
def test_ell_poisson_width5_width7_bitexact(gpu, orc, dims, width, dtype):
    """2D 5-point and 3D 7-point Laplacians in ELLPACK (the four-entry main
    loop and the tail both run): integer-valued x and integer coefficients
    make every product and sum exact, so the GPU equals the oracle bit for
    bit; batched matrix and vector, and the += / -= modes."""
    cm = gpu
    from cumat_amd.problems import grid_laplacian
    rng = np.random.default_rng(width)
    P = grid_laplacian(dims, lambda s: rng.integers(1, 5, s).astype(np.float64), edges=False)
    n = P["n"]
    ind, v0 = _ell_from_csr(P["JA"], P["IA"], P["vals"], width)
    B = 3
    vals = np.stack([(v0 * (k + 1)).astype(dtype) for k in range(B)])
    E = cm.ELLMatrix(n, n, ind, vals)
    x = rng.integers(-20, 21, n * B).astype(dtype)
    xd = cm.DeviceArray.from_numpy(x, rows=n, cols=1, batches=B)
    ind_cm = np.asfortranarray(ind).reshape(-1, order="F")
    vals_cm = np.concatenate([np.asfortranarray(vals[k]).reshape(-1, order="F") for k in range(B)])
    y = E.matvec(xd).numpy()
    ref = orc.ell_spmv(ind_cm, vals_cm, x, n, n, width, batches=B, vals_batched=True, x_batched=True)
    assert np.array_equal(y, ref)
    # agrees with the CSR product of the same matrix
    Acsr = cm.SparseMatrix.from_csr(n, n, P["JA"], P["IA"], P["vals"].astype(dtype))
    x0 = cm.DeviceArray.from_numpy(x[:n].copy())
    assert np.array_equal(Acsr.matvec(x0).numpy(), y[:n])
    # compound modes, batched vector against a one-batch matrix
    E1 = cm.ELLMatrix(n, n, ind, vals[0])
    y0 = rng.integers(-100, 101, n * B).astype(dtype)
    for mode in (cm.capi.ADD, cm.capi.SUB):
        yd = cm.DeviceArray.from_numpy(y0.copy(), rows=n, cols=1, batches=B)
        E1.matvec(xd, out=yd, mode=mode)
        ref = orc.ell_spmv(ind_cm, vals_cm[:n * width], x, n, n, width, batches=B, vals_batched=False,
                           x_batched=True, y=y0.copy(), mode=mode)
        assert np.array_equal(yd.numpy(), ref), mode


def test_ell_argument_checks(gpu):
    cm = gpu
    g = GOLD["ell_4x5"]
    E = cm.ELLMatrix(4, 5, g["indices_rows"], np.array(g["values_rows"], np.float32))
    with pytest.raises(cm.capi.CumatArgError):
        E.matvec(cm.DeviceArray.from_numpy(np.array(g["b"], np.float64)))  # dtype mismatch
    with pytest.raises(cm.capi.CumatArgError):
        E.matvec(cm.DeviceArray.from_numpy(np.ones(4, np.float32)))  # x too short
    with pytest.raises(cm.capi.CumatArgError):
        E.matvec(cm.DeviceArray.from_numpy(np.array(g["b"], np.float32)), out=cm.DeviceArray(3, dtype=np.float32))


# ------------------------------------------------------------ Jacobi setup guard
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_csr_inverse_diagonal_missing_and_zero(gpu, dtype):
    """invdiag = A_ii == 0 ? 1 : 1/A_ii (cwiseInverseCheck,
    cuMat/src/UnaryOps.h:177): a row with no stored diagonal and a row whose
    stored diagonal is 0 both give 1; the others give 1/A_ii exactly."""
    cm = gpu
    # row 0: diag 4; row 1: no diagonal entry; row 2: stored 0 on the
    # diagonal; row 3: diag -2; row 4: empty
    JA = np.array([0, 2, 4, 6, 8, 8], np.int32)
    IA = np.array([0, 1, 0, 2, 1, 2, 0, 3], np.int32)
    V = np.array([4, -1, -1, 5, 7, 0, 3, -2], dtype)
    A = cm.SparseMatrix.from_csr(5, 5, JA, IA, V)
    inv = A.inverse_diagonal().numpy()
    exp = np.array([0.25, 1.0, 1.0, -0.5, 1.0], dtype)
    assert np.array_equal(inv, exp)
    # batched values: batch 1 has the diagonal scaled by 2
    V2 = np.concatenate([V, V * 2])
    A2 = cm.SparseMatrix(cm.SparsityPattern(5, 5, JA, IA), V2, batches=2)
    inv2 = A2.inverse_diagonal().numpy()
    assert np.array_equal(inv2, np.concatenate([exp, np.array([0.125, 1.0, 1.0, -0.25, 1.0], dtype)]))
