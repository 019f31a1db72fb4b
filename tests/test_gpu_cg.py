"""GPU parity of the Jacobi-preconditioned CG (cuMat/src/ConjugateGradient.h:
123-206; Eigen's conjugate_gradient, third-party/Eigen/src/
IterativeLinearSolvers/ConjugateGradient.h:28-91) against the CPU oracle, the
Eigen golden runs and closed forms.

FP32 parity: the reported error is <= tol, the iteration count is within a
few of the oracle's (reduction order differs), and the iterate is within
1e-5 (relative, 2-norm) of the FP64 solution."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PATHS = [1, 2, 3]  # multikernel, resident, fused multikernel


def _poisson(cm, g, shift=0.0, dtype=np.float32):
    from cumat_amd.problems import poisson2d
    JA, IA, V = poisson2d(g, shift, dtype)
    n = g * g
    return cm.SparseMatrix.from_csr(n, n, JA, IA, V), JA, IA, V, n


@pytest.mark.parametrize("path", PATHS)
def test_config1_poisson128(gpu, orc, path):
    """The plain-CG workload: 128x128 Poisson, b ~ U[-1,1], tol 1e-6, max 1000."""
    cm = gpu
    A, JA, IA, V, n = _poisson(cm, 128)
    b = np.random.default_rng(20261203).uniform(-1, 1, n).astype(np.float32)
    cg = cm.ConjugateGradient(A, path=path).setTolerance(1e-6).setMaxIterations(1000)
    x = cg.solve(cm.DeviceArray.from_numpy(b)).numpy()
    x_o, it_o, err_o = orc.pcg(JA, IA, V, b, n, tol=1e-6, max_iter=1000)
    x64, it64, _ = orc.pcg(JA, IA, V.astype(np.float64), b.astype(np.float64), n, tol=1e-12, max_iter=5000)
    assert cg.error() <= 1e-6
    assert abs(cg.iterations() - it_o) <= 4, (cg.iterations(), it_o)
    assert np.linalg.norm(x - x64) / np.linalg.norm(x64) <= 1e-5
    assert np.linalg.norm(x - x_o) / np.linalg.norm(x_o) <= 1e-5
    assert cg.last_stats["path"] == path


def test_eigen_golden_runs(gpu):
    """Small Poisson systems against the committed Eigen runs
    (tests/golden/eigen_cg.json, written by tests/golden/make_eigen_vectors.py)."""
    cm = gpu
    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "eigen_cg.json")))
    for case in gold["cases"]:
        g = case["g"]
        A, JA, IA, V, n = _poisson(cm, g, case["shift"], np.float64)
        b = np.array(case["b"], np.float64)
        for path in PATHS:
            cg = cm.ConjugateGradient(A, path=path).setTolerance(case["tol"]).setMaxIterations(case["max_iter"])
            x = cg.solve(cm.DeviceArray.from_numpy(b)).numpy()
            assert abs(cg.iterations() - case["iterations_f64"]) <= 1, (g, path)
            np.testing.assert_allclose(x, np.array(case["x_f64"]), rtol=0, atol=1e-9 * np.abs(x).max())


@pytest.mark.parametrize("path", PATHS)
def test_early_out_zero_rhs(gpu, path):
    """tests/TestConjugateGradient.cu:11-25: zero rhs -> 0 iterations, error 0, x = 0."""
    cm = gpu
    A, *_ , n = _poisson(cm, 10)
    cg = cm.ConjugateGradient(A, path=path)
    x = cg.solveWithGuess(cm.DeviceArray.from_numpy(np.zeros(n, np.float32)),
                          cm.DeviceArray.from_numpy(np.ones(n, np.float32))).numpy()
    assert cg.iterations() == 0 and cg.error() == 0 and not x.any()


@pytest.mark.parametrize("path", PATHS)
def test_early_out_exact_guess(gpu, orc, path):
    """tests/TestConjugateGradient.cu:27-45: an exact guess gives 0 iterations."""
    cm = gpu
    A, JA, IA, V, n = _poisson(cm, 12)
    xt = np.random.default_rng(1).integers(-3, 4, n).astype(np.float32)
    b = orc.csr_spmv(JA, IA, V, xt, n, n)  # exact: integer data
    cg = cm.ConjugateGradient(A, path=path)
    x = cg.solveWithGuess(cm.DeviceArray.from_numpy(b), cm.DeviceArray.from_numpy(xt)).numpy()
    assert cg.iterations() == 0 and cg.error() == 0.0
    assert np.array_equal(x, xt)


@pytest.mark.parametrize("path", PATHS)
def test_diagonal_matrix_one_iteration(gpu, path):
    """Closed form: Jacobi-PCG on a diagonal matrix converges in one step
    (the converging step is not counted) with x = D^-1 b."""
    cm = gpu
    n = 500
    d = np.random.default_rng(5).uniform(1, 10, n)
    A = cm.SparseMatrix.from_csr(n, n, np.arange(n + 1, dtype=np.int32), np.arange(n, dtype=np.int32), d)
    b = np.random.default_rng(6).uniform(-1, 1, n)
    cg = cm.ConjugateGradient(A, path=path).setTolerance(1e-12)
    x = cg.solve(cm.DeviceArray.from_numpy(b)).numpy()
    assert cg.iterations() == 0
    np.testing.assert_allclose(x, b / d, rtol=1e-14)


def test_small_spd_exact_arithmetic_bound(gpu, orc):
    """CG on an n x n SPD matrix reaches a tiny residual within about n
    iterations in FP64 (exact arithmetic: at most n)."""
    cm = gpu
    rng = np.random.default_rng(12)
    n = 24
    M = rng.standard_normal((n, n))
    S = M @ M.T + n * np.eye(n)
    JA = (np.arange(n + 1) * n).astype(np.int32)
    IA = np.tile(np.arange(n), n).astype(np.int32)
    A = cm.SparseMatrix.from_csr(n, n, JA, IA, S.reshape(-1))
    b = rng.standard_normal(n)
    for path in PATHS:
        cg = cm.ConjugateGradient(A, path=path).setTolerance(1e-13).setMaxIterations(200)
        x = cg.solve(cm.DeviceArray.from_numpy(b)).numpy()
        assert cg.iterations() <= n + 2
        np.testing.assert_allclose(S @ x, b, atol=1e-10)


def test_batched_rhs_and_max_iterations(gpu, orc):
    """Broadcast matrix, 3 right-hand sides (tests/TestConjugateGradient.cu:74-99);
    all batches iterate until all converge; maxIterations exhaustion."""
    cm = gpu
    A, JA, IA, V, n = _poisson(cm, 40, 0.0, np.float64)
    rng = np.random.default_rng(3)
    b = rng.uniform(-1, 1, 3 * n)
    cg = cm.ConjugateGradient(A).setTolerance(1e-10).setMaxIterations(1000)
    x = cg.solve(cm.DeviceArray.from_numpy(b, rows=n, cols=1, batches=3)).numpy()
    xo, ito, erro = orc.pcg(JA, IA, V, b, n, batches=3, tol=1e-10, max_iter=1000)
    assert abs(cg.iterations() - ito) <= 1 and cg.error() <= 1e-10
    np.testing.assert_allclose(x, xo, atol=1e-8 * np.abs(xo).max())
    cg.setMaxIterations(7)
    x = cg.solve(cm.DeviceArray.from_numpy(b, rows=n, cols=1, batches=3)).numpy()
    xo, ito, erro = orc.pcg(JA, IA, V, b, n, batches=3, tol=1e-10, max_iter=7)
    assert cg.iterations() == 7 == ito
    np.testing.assert_allclose(cg.error(), erro, rtol=1e-9)
    np.testing.assert_allclose(x, xo, atol=1e-10)


def test_paths_agree_and_default_tolerance(gpu):
    cm = gpu
    A, JA, IA, V, n = _poisson(cm, 60, 0.0, np.float64)
    b = np.random.default_rng(9).uniform(-1, 1, n)
    xs = []
    for path in PATHS:
        cg = cm.ConjugateGradient(A, path=path)
        assert cg.tolerance() == np.finfo(np.float64).eps and cg.maxIterations() == 2 * n
        cg.setTolerance(1e-13)
        xs.append(cg.solve(cm.DeviceArray.from_numpy(b)).numpy())
    np.testing.assert_allclose(xs[0], xs[1], atol=1e-11)


def test_large_poisson_2048_residual(gpu):
    """Full-size 2048x2048 Poisson with the multi-kernel path: the reported
    recurrence residual meets the tolerance and the true residual agrees."""
    cm = gpu
    A, JA, IA, V, n = _poisson(cm, 2048)
    b = np.random.default_rng(2048).uniform(-1, 1, n).astype(np.float32)
    cg = cm.ConjugateGradient(A, path=1).setTolerance(1e-3).setMaxIterations(20000)
    xd = cg.solve(cm.DeviceArray.from_numpy(b))
    assert cg.error() <= 1e-3 and 0 < cg.iterations() < 20000
    r = A.matvec(xd).numpy().astype(np.float64) - b
    assert np.linalg.norm(r) / np.linalg.norm(b) < 3e-3


def test_auto_path_choice(gpu):
    """AUTO: one workgroup for a small single system, the two-launch form up
    to 2^17 unknowns (also for several right-hand sides), the three-launch
    form beyond."""
    cm = gpu
    for g, expect in ((32, 2), (64, 2), (96, 3), (200, 3), (362, 3), (363, 1), (600, 1)):
        A, JA, IA, V, n = _poisson(cm, g)
        cg = cm.ConjugateGradient(A).setTolerance(1e-4)
        cg.solve(cm.DeviceArray.from_numpy(np.ones(n, np.float32)))
        assert cg.last_stats["path"] == expect, (g, cg.last_stats)
    A, JA, IA, V, n = _poisson(cm, 16)
    cg = cm.ConjugateGradient(A).setTolerance(1e-4)
    cg.solve(cm.DeviceArray.from_numpy(np.ones((2, n, 1), np.float32)))
    assert cg.last_stats["path"] == 3
    A, JA, IA, V, n = _poisson(cm, 300)
    cg = cm.ConjugateGradient(A).setTolerance(1e-4)
    cg.solve(cm.DeviceArray.from_numpy(np.ones((2, n, 1), np.float32)))
    assert cg.last_stats["path"] == 1  # 2 x 90,000 unknowns


def test_fused_and_three_launch_agree(gpu):
    """The two-launch iteration computes p with an explicit multiply-add and
    recomputes gathered p values, so iterates agree with the three-launch
    form to rounding, with the same iteration count within 1 (FP64)."""
    cm = gpu
    A, JA, IA, V, n = _poisson(cm, 150, 0.01, np.float64)
    b = np.random.default_rng(15).uniform(-1, 1, (3, n, 1))
    out = {}
    for path in (1, 3):
        cg = cm.ConjugateGradient(A, path=path).setTolerance(1e-11).setMaxIterations(5000)
        x = cg.solve(cm.DeviceArray.from_numpy(b)).numpy()
        out[path] = (x, cg.iterations(), cg.error())
        assert cg.last_stats["path"] == path
    assert abs(out[1][1] - out[3][1]) <= 1
    np.testing.assert_allclose(out[1][0], out[3][0], rtol=0, atol=1e-8)



def test_solves_repeat_exactly_across_contexts(gpu):
    """Matrix on one context, solves on another, the allocator's cache
    refilled with NaN before every solve: each solve gives the same iteration
    count and a bit-identical, finite x (the Jacobi entries are computed on
    the matrix's context and complete before the constructor returns, so a
    solve on any stream reads them finished)."""
    cm = gpu
    from cumat_amd.problems import grid_laplacian, lognormal_sampler
    ctx1, ctx2 = cm.Context(0), cm.Context(0)
    P = grid_laplacian((100, 100, 100), lognormal_sampler(np.random.default_rng(41), 0.0, 1.0), edges=False)
    n = P["n"]
    A = cm.SparseMatrix(cm.SparsityPattern(n, n, P["JA"], P["IA"], ctx=ctx1),
                        cm.DeviceArray.from_numpy(P["vals"], ctx=ctx1))
    b = cm.DeviceArray.from_numpy(np.random.default_rng(42).uniform(-1, 1, n), ctx=ctx2)
    runs = []
    for rep in range(3):
        junk = []
        for nbytes in [8 * n] * 4 + [16 * n] * 2 + [65536, 16384, 4096] * 2:
            a = cm.DeviceArray(nbytes // 8, dtype=np.float64, ctx=ctx2)
            cm.assign(a, cm.as_expr(np.float64("nan")), ctx=ctx2)
            junk.append(a)
        ctx2.sync()
        for a in junk:
            a.free()
        x = cm.DeviceArray(n, dtype=np.float64, ctx=ctx2)
        cg = cm.ConjugateGradient(A, path=1).setTolerance(1e-8).setMaxIterations(5000)
        cg.solve(b, x, ctx=ctx2)
        runs.append((cg.iterations(), x.numpy()))
    assert np.isfinite(runs[0][1]).all()
    assert runs[0][0] < 5000
    for it, xs in runs[1:]:
        assert it == runs[0][0]
        np.testing.assert_array_equal(xs, runs[0][1])
