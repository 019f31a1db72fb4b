"""GPU parity of the block-scalar path (block-CSR product, block Jacobi
entries, block Jacobi-PCG) against the CPU oracle and the reference's own
blocked CG test (tests/TestBlockedConjugateGradient.cu:244-275). Blocks are
D float4 rows, vector entries float4 (the reference's real3x3 / real3)."""
import ctypes as C
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_tests.json")))


def _rand_bcsr(rng, rows, cols, D, max_per_row, long_rows=()):
    JA = [0]
    IA = []
    for i in range(rows):
        k = int(rng.integers(0, max_per_row + 1))
        if i in long_rows:
            k = cols
        IA.extend(sorted(rng.choice(cols, size=min(k, cols), replace=False)))
        JA.append(len(IA))
    V = rng.standard_normal((len(IA), D, 4)).astype(np.float32)
    return np.array(JA, np.int32), np.array(IA, np.int32), V


def _dev(cm, a, batches=1):
    a = np.ascontiguousarray(a, np.float32).reshape(-1)
    return cm.DeviceArray.from_numpy(a, rows=a.size // batches, cols=1, batches=batches)


@pytest.mark.parametrize("D", [2, 3, 4])
def test_bcsr_spmv_bit_exact(gpu, orc, D):
    """Random float blocks and vectors: every product and row sum has the
    oracle's bits (dots not fused, rows summed in CSR order). Rows of up to
    20 blocks and one row of 1100 blocks make chunks longer than one LDS
    window (256 blocks); empty rows assign 0."""
    cm = gpu
    rng = np.random.default_rng(100 + D)
    rows, cols = 3000, 1100
    JA, IA, V = _rand_bcsr(rng, rows, cols, D, 20, long_rows=(5, 2000))
    A = cm.BlockSparseMatrix.from_bcsr(rows, cols, JA, IA, V)
    x = rng.standard_normal(cols * 4).astype(np.float32)
    cm.reset_counters()
    y = A.matvec(_dev(cm, x)).numpy()
    assert cm.counters()["EvalAny"] == 1
    exp = orc.bcsr_spmv(D, JA, IA, V, x, rows, cols)
    assert np.array_equal(y, exp)
    assert np.any(np.diff(JA) == 0) and np.all(y.reshape(rows, 4)[np.diff(JA) == 0] == 0)
    assert np.all(y.reshape(rows, 4)[:, D:] == 0)


@pytest.mark.parametrize("case", ["matrix", "vector", "both"])
def test_bcsr_spmv_batched(gpu, orc, case):
    cm = gpu
    rng = np.random.default_rng(7)
    D, rows, cols, B = 3, 700, 650, 3
    JA, IA, V = _rand_bcsr(rng, rows, cols, D, 9)
    Vb = np.stack([V * s for s in (1, 2, -1)]) if case != "vector" else V
    xs = rng.standard_normal((B if case != "matrix" else 1, cols * 4)).astype(np.float32)
    pat = cm.SparsityPattern(rows, cols, JA, IA)
    A = cm.BlockSparseMatrix(pat, D, Vb)
    y = A.matvec(_dev(cm, xs, batches=xs.shape[0])).numpy()
    exp = orc.bcsr_spmv(D, JA, IA, Vb, xs.reshape(-1), rows, cols, batches=B,
                        vals_batched=case != "vector", x_batched=case != "matrix")
    assert np.array_equal(y, exp)


@pytest.mark.parametrize("mode", [1, 2, 3, 4])
def test_bcsr_spmv_compound_modes(gpu, orc, mode):
    cm = gpu
    rng = np.random.default_rng(11)
    D, rows, cols = 4, 500, 480
    JA, IA, V = _rand_bcsr(rng, rows, cols, D, 8)
    A = cm.BlockSparseMatrix.from_bcsr(rows, cols, JA, IA, V)
    x = rng.standard_normal(cols * 4).astype(np.float32)
    y0 = rng.uniform(1, 2, rows * 4).astype(np.float32)
    yd = _dev(cm, y0)
    A.matvec(_dev(cm, x), out=yd, mode=mode)
    exp = orc.bcsr_spmv(D, JA, IA, V, x, rows, cols, y=y0.copy(), mode=mode)
    assert np.array_equal(yd.numpy(), exp)


def test_bcsr_spmv_rejects(gpu):
    cm = gpu
    from cumat_amd import _capi
    A = cm.BlockSparseMatrix.from_bcsr(2, 2, [0, 1, 2], [0, 1], np.ones((2, 3, 4), np.float32))
    with pytest.raises(_capi.CumatArgError):
        A.matvec(_dev(cm, np.ones(7)))  # not one float4 per block column
    with pytest.raises(_capi.CumatArgError):
        A.matvec(cm.DeviceArray.from_numpy(np.ones(8, np.float64)))
    with pytest.raises(_capi.CumatArgError):
        cm.BlockSparseMatrix.from_bcsr(2, 2, [0, 1, 2], [0, 1], np.ones((2, 5, 4), np.float32))
    L = _capi.lib()
    x = _dev(cm, np.ones(8))
    rc = L.cm_bcsr_spmv(None, 5, 2, 2, 2, 1, A.pattern.JA.ptr, A.pattern.IA.ptr, A.values.ptr, 0, x.ptr, 0,
                        x.ptr, 0)
    assert rc == _capi.CM_E_UNSUPPORTED


def test_bcsr_inverse_diagonal(gpu, orc):
    cm = gpu
    from cumat_amd.problems import block_poisson
    JA, IA, V = block_poisson((30, 40), 3, seed=2)
    rows = 1200
    V = V.copy()
    IA = IA.copy()
    k4 = [k for k in range(JA[4], JA[5]) if IA[k] == 4][0]
    V[k4, 2, 2] = 0  # zero diagonal entry -> 1
    k9 = [k for k in range(JA[9], JA[10]) if IA[k] == 9][0]
    IA[k9] = IA[k9 - 1]  # no diagonal block -> 1 1 1
    pat = cm.SparsityPattern(rows, rows, JA, IA)
    A = cm.BlockSparseMatrix(pat, 3, V)
    got = A.inverse_diagonal().numpy()
    exp = orc.bcsr_inverse_diagonal(3, JA, IA, V, rows)
    assert np.array_equal(got, exp)
    assert got[4 * 4 + 2] == 1 and np.all(got[4 * 9:4 * 9 + 3] == 1)


def test_blocked_cg_reference_fixture(gpu):
    """tests/TestBlockedConjugateGradient.cu:244-275: converges within 30
    iterations at tolerance 1e-5, with at least one iteration."""
    cm = gpu
    g = GOLD["blocked_cg_3x3"]
    A = cm.BlockSparseMatrix.from_bcsr(g["rows"], g["cols"], g["JA"], g["IA"], np.array(g["blocks"]))
    b = np.array(g["b"], np.float32).reshape(-1)
    cg = cm.ConjugateGradient(A)
    cg.setMaxIterations(g["max_iterations"])
    cg.setTolerance(g["tolerance"])
    x = cg.solve(_dev(cm, b)).numpy()
    assert cg.error() < cg.tolerance()
    assert 0 < cg.iterations() < cg.maxIterations()
    Ad = np.zeros((9, 9))
    blocks = np.array(g["blocks"], np.float64)
    for i in range(3):
        for k in range(g["JA"][i], g["JA"][i + 1]):
            j = g["IA"][k]
            Ad[3 * i:3 * i + 3, 3 * j:3 * j + 3] = blocks[k, :, :3]
    xe = np.linalg.solve(Ad, b.reshape(3, 4)[:, :3].reshape(-1).astype(np.float64))
    xr = x.reshape(3, 4)
    assert np.allclose(xr[:, :3].reshape(-1), xe, atol=1e-4 * np.abs(xe).max())
    assert np.all(xr[:, 3] == 0)


@pytest.mark.parametrize("D", [2, 3, 4])
def test_block_cg_against_oracle(gpu, orc, D):
    """2D block Poisson (60 x 70 blocks): the FP32 solve tracks the oracle's
    FP32 block PCG (iterations within 4) and reaches the FP64 solution to
    the tolerance the FP32 run allows."""
    cm = gpu
    from cumat_amd.problems import block_poisson
    JA, IA, V = block_poisson((60, 70), D, seed=D)
    rows = 4200
    rng = np.random.default_rng(3)
    b = rng.uniform(-1, 1, rows * 4).astype(np.float32)
    A = cm.BlockSparseMatrix.from_bcsr(rows, rows, JA, IA, V)
    cg = cm.ConjugateGradient(A).setTolerance(1e-6).setMaxIterations(2000)
    x = cg.solve(_dev(cm, b)).numpy()
    xo, ito, erro = orc.bcsr_pcg(D, JA, IA, V, b, rows, tol=1e-6, max_iter=2000, dtype=np.float32)
    assert abs(cg.iterations() - ito) <= 4 and cg.error() < 1e-6
    x64, _, _ = orc.bcsr_pcg(D, JA, IA, V, b, rows, tol=1e-13, max_iter=5000)
    sel = np.arange(rows * 4).reshape(rows, 4)[:, :D].reshape(-1)
    rel = np.linalg.norm(x[sel] - x64[sel]) / np.linalg.norm(x64[sel])
    rel_o = np.linalg.norm(xo[sel] - x64[sel]) / np.linalg.norm(x64[sel])
    assert rel < max(5 * rel_o, 1e-5), (rel, rel_o)
    assert np.all(x.reshape(rows, 4)[:, D:] == 0)


def test_block_cg_batched_guess_and_ignored_components(gpu, orc):
    """Three right-hand sides (all iterate until all converge), a guess, and
    a nonzero w in b that the solve ignores (the reference's squaredNorm and
    Jacobi entries never see it); the guess's w is kept."""
    cm = gpu
    from cumat_amd.problems import block_poisson
    JA, IA, V = block_poisson((40, 45), 3, seed=4)
    rows, B = 1800, 3
    rng = np.random.default_rng(8)
    b = rng.uniform(-1, 1, (B, rows, 4)).astype(np.float32)
    g0 = rng.uniform(-0.1, 0.1, (B, rows, 4)).astype(np.float32)
    A = cm.BlockSparseMatrix.from_bcsr(rows, rows, JA, IA, V)
    cg = cm.ConjugateGradient(A).setTolerance(1e-6).setMaxIterations(2000)
    x = cg.solveWithGuess(_dev(cm, b, B), _dev(cm, g0, B)).numpy().reshape(B, rows, 4)
    bz = b.copy()
    bz[:, :, 3] = 0
    gz = g0.copy()
    gz[:, :, 3] = 0
    xo, ito, _ = orc.bcsr_pcg(3, JA, IA, V, bz.reshape(-1), rows, batches=B, x0=gz.reshape(-1), tol=1e-6,
                              max_iter=2000, dtype=np.float32)
    assert abs(cg.iterations() - ito) <= 4 and cg.error() < 1e-6
    x64, _, _ = orc.bcsr_pcg(3, JA, IA, V, bz.reshape(-1), rows, batches=B, tol=1e-13, max_iter=5000)
    x64 = x64.reshape(B, rows, 4)
    for k in range(B):
        rel = np.linalg.norm(x[k, :, :3] - x64[k, :, :3]) / np.linalg.norm(x64[k, :, :3])
        assert rel < 1e-4, (k, rel)
    assert np.array_equal(x[:, :, 3], g0[:, :, 3])


def test_block_cg_early_outs(gpu):
    cm = gpu
    from cumat_amd.problems import block_poisson
    JA, IA, V = block_poisson((10, 12), 3, seed=1)
    rows = 120
    A = cm.BlockSparseMatrix.from_bcsr(rows, rows, JA, IA, V)
    cg = cm.ConjugateGradient(A).setTolerance(1e-5)
    x = cg.solve(_dev(cm, np.zeros(rows * 4))).numpy()
    assert cg.iterations() == 0 and cg.error() == 0 and np.all(x == 0)
    # only w components set: the right-hand side counts as zero
    bw = np.zeros((rows, 4), np.float32)
    bw[:, 3] = 1
    x = cg.solve(_dev(cm, bw)).numpy()
    assert cg.iterations() == 0 and np.all(x == 0)


def test_block_cg_direct_abi_with_batched_values(gpu, orc):
    """cm_pcg_bcsr_solve with per-batch matrix values (vals_batched) and no
    Jacobi entries given: each batch solves its own system."""
    cm = gpu
    from cumat_amd import _capi
    from cumat_amd.problems import block_poisson
    JA, IA, V0 = block_poisson((20, 25), 4, seed=6)
    rows, B = 500, 2
    V = np.stack([V0, V0 * 2])
    pat = cm.SparsityPattern(rows, rows, JA, IA)
    A = cm.BlockSparseMatrix(pat, 4, V)
    rng = np.random.default_rng(2)
    b = rng.uniform(-1, 1, (B, rows * 4)).astype(np.float32)
    bd = _dev(cm, b, B)
    xd = cm.DeviceArray(rows * 4, 1, B, dtype=np.float32)
    it, err, st = C.c_int(0), C.c_double(0), _capi.CgStats()
    _capi.call("cm_pcg_bcsr_solve", None, 4, rows, pat.nnz, B, pat.JA.ptr, pat.IA.ptr, A.values.ptr, 1, None,
               bd.ptr, xd.ptr, 0, 1e-6, 3000, C.byref(it), C.byref(err), C.byref(st))
    x = xd.numpy().reshape(B, rows * 4)
    assert err.value < 1e-6 and it.value > 0
    for k in range(B):
        x64, _, _ = orc.bcsr_pcg(4, JA, IA, V[k], b[k], rows, tol=1e-13, max_iter=5000)
        rel = np.linalg.norm(x[k] - x64) / np.linalg.norm(x64)
        assert rel < 1e-4, (k, rel)
