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
