"""GPU tests of the three solver additions (no reference counterpart). Their
oracle is the plain method restated in oracle/ (FP64 CG per system), closed
forms, and central finite differences for the gradient."""
import ctypes as C
import gc

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _csr_dense(JA, IA, V, n):
    A = np.zeros((n, n))
    rows = np.repeat(np.arange(n), np.diff(JA))
    A[rows, IA] = V
    return A


# ------------------------------------------------------------ batched values
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_solve_ex_batched_values_per_batch(gpu, orc, dtype):
    """Each batch has its own matrix (same pattern) and stops at its own
    convergence: per batch identical to the plain solver on that system."""
    cm = gpu
    from cumat_amd.problems import poisson2d
    g, B = 30, 5
    n = g * g
    shifts = [0.0, 0.05, 0.3, 1.0, 4.0]
    JA, IA, V0 = poisson2d(g, 0.0, np.float64)
    vals = np.concatenate([poisson2d(g, s, np.float64)[2] for s in shifts]).astype(dtype)
    rng = np.random.default_rng(17)
    b = rng.uniform(-1, 1, n * B).astype(dtype)
    b[2 * n:3 * n] = 0.0  # a zero right-hand side: early out for that batch only
    tol = 1e-10 if dtype == np.float64 else 1e-5
    dJA = cm.DeviceArray.from_numpy(JA)
    dIA = cm.DeviceArray.from_numpy(IA)
    dV = cm.DeviceArray.from_numpy(vals)
    db = cm.DeviceArray.from_numpy(b)
    dx = cm.DeviceArray(n * B, dtype=dtype)
    its = (C.c_int * B)()
    errs = (C.c_double * B)()
    st = cm.capi.CgStats()
    cm.capi.call("cm_pcg_solve_ex", None, cm.capi.F64 if dtype == np.float64 else cm.capi.F32, n, IA.size, B,
                 dJA.ptr, dIA.ptr, dV.ptr, 1, None, db.ptr, dx.ptr, 0, tol, 5000, 1, its, errs, C.byref(st))
    x = dx.numpy()
    for k in range(B):
        xk, itk, errk = orc.pcg(JA, IA, vals[k * IA.size:(k + 1) * IA.size], b[k * n:(k + 1) * n], n, tol=tol,
                                max_iter=5000)
        if k == 2:
            assert its[k] == 0 and errs[k] == 0 and not x[k * n:(k + 1) * n].any()
            continue
        assert abs(its[k] - itk) <= (1 if dtype == np.float64 else 3), (k, its[k], itk)
        assert errs[k] <= tol
        rel = np.linalg.norm(x[k * n:(k + 1) * n] - xk) / np.linalg.norm(xk)
        assert rel <= (1e-8 if dtype == np.float64 else 1e-4), (k, rel)
    # iteration counts differ per system (shift 4 converges much faster)
    assert its[4] < its[0]


# ------------------------------------------------------------ persistent batched CG
def _batched_problem(nsys, lo, hi, seed):
    from cumat_amd.problems import batched_poisson
    return batched_poisson(nsys, lo, hi, seed)


@pytest.mark.parametrize("use_graph", [0, 1])
def test_batched_persistent_cg(gpu, orc, use_graph):
    cm = gpu
    P = _batched_problem(300, 16, 40, seed=5)
    nsys = 300
    ro = P["row_off"]
    d = {k: cm.DeviceArray.from_numpy(P[k]) for k in ("row_off", "JA", "IA", "vals", "b")}
    x = cm.DeviceArray(P["b"].size, dtype=np.float32)
    its = cm.DeviceArray(nsys, dtype=np.int32)
    errs = cm.DeviceArray(nsys, dtype=np.float32)
    for rep in range(2 if use_graph else 1):  # a graph replays the same result
        x.fill_zero()
        cm.pcg_batched(d["row_off"], d["JA"], d["IA"], d["vals"], d["b"], x, 1e-6, 2000, its, errs,
                       use_graph=bool(use_graph))
        xh, ih, eh = x.numpy(), its.numpy(), errs.numpy()
        for s in range(0, nsys, 7):
            lo, hi = ro[s], ro[s + 1]
            n = hi - lo
            JA = P["JA"][lo:hi + 1] - P["JA"][lo]
            IA = P["IA"][P["JA"][lo]:P["JA"][hi]] - lo
            V = P["vals"][P["JA"][lo]:P["JA"][hi]]
            xo, ito, erro = orc.pcg(JA, IA, V, P["b"][lo:hi], n, tol=1e-6, max_iter=2000)
            x64, _, _ = orc.pcg(JA, IA, V.astype(np.float64), P["b"][lo:hi].astype(np.float64), n, tol=1e-13,
                                max_iter=5000)
            assert abs(int(ih[s]) - ito) <= 3, (s, ih[s], ito)
            assert eh[s] <= 1e-6
            assert np.linalg.norm(xh[lo:hi] - x64) / np.linalg.norm(x64) <= 1e-5


def test_batched_autotune_and_block_sizes(gpu, orc):
    cm = gpu
    P = _batched_problem(64, 30, 70, seed=9)
    d = {k: cm.DeviceArray.from_numpy(P[k]) for k in ("row_off", "JA", "IA", "vals", "b")}
    from cumat_amd.solvers import pcg_batched_autotune
    bs, gpc, ms = pcg_batched_autotune(d["row_off"], d["JA"], d["IA"], d["vals"], d["b"], 1e-6, 2000,
                                       with_groups=True)
    assert bs in (256, 512, 1024) and gpc in (1, 2, 4) and ms > 0
    # block_size=0 solves with the tuned shape (threads per system and
    # workgroups per CU): every system converges, iterate close to the others
    xt = cm.DeviceArray(P["b"].size, dtype=np.float32)
    itst = cm.DeviceArray(64, dtype=np.int32)
    errst = cm.DeviceArray(64, dtype=np.float32)
    cm.pcg_batched(d["row_off"], d["JA"], d["IA"], d["vals"], d["b"], xt, 1e-6, 2000, itst, errst, block_size=0)
    assert (errst.numpy() <= 1e-6).all()
    res = []
    for b_s in (256, 512, 1024):
        x = cm.DeviceArray(P["b"].size, dtype=np.float32)
        its = cm.DeviceArray(64, dtype=np.int32)
        errs = cm.DeviceArray(64, dtype=np.float32)
        cm.pcg_batched(d["row_off"], d["JA"], d["IA"], d["vals"], d["b"], x, 1e-6, 2000, its, errs, block_size=b_s)
        res.append((x.numpy(), its.numpy()))
    # block size changes the reduction tree only: same iterations within 2
    assert np.abs(res[0][1] - res[2][1]).max() <= 2
    np.testing.assert_allclose(res[0][0], res[2][0], atol=1e-4)
    np.testing.assert_allclose(xt.numpy(), res[2][0], atol=1e-4)
    assert np.abs(itst.numpy() - res[2][1]).max() <= 2


# ------------------------------------------------------------ mixed precision IR
@pytest.mark.parametrize("storage", ["f32", "bf16"])
def test_mixed_precision_ir_reaches_1e12(gpu, orc, storage):
    """3D 7-point variable-coefficient Poisson (log-normal(0,1) edges): FP64
    iterative refinement around FP32/BF16 inner CG reaches a true FP64
    relative residual <= 1e-12 and the FP64 CG oracle's solution."""
    cm = gpu
    from cumat_amd.problems import grid_laplacian, lognormal_sampler
    g = 24
    P = grid_laplacian((g, g, g), lognormal_sampler(np.random.default_rng(160), 0.0, 1.0))
    n = P["n"]
    b = np.random.default_rng(161).uniform(-1, 1, n)
    dJA, dIA = cm.DeviceArray.from_numpy(P["JA"]), cm.DeviceArray.from_numpy(P["IA"])
    dV = cm.DeviceArray.from_numpy(P["vals"])
    db = cm.DeviceArray.from_numpy(b)
    dx = cm.DeviceArray(n, dtype=np.float64)
    st = cm.pcg_ir(dJA, dIA, dV, db, dx, storage=storage, tol=1e-12, inner_tol=1e-3 if storage == "bf16" else 1e-5)
    x = dx.numpy()
    assert st["rel_residual"] <= 1e-12, st
    A = _csr_dense(P["JA"], P["IA"], P["vals"], n) if n <= 20000 else None
    r = b - orc.csr_spmv(P["JA"], P["IA"], P["vals"], x, n, n)
    assert np.linalg.norm(r) / np.linalg.norm(b) <= 1.5e-12
    x64, it64, err64 = orc.pcg(P["JA"], P["IA"], P["vals"], b, n, tol=1e-13, max_iter=20000)
    assert np.linalg.norm(x - x64) / np.linalg.norm(x64) <= 1e-9
    assert st["outer"] >= 2 and st["inner"] > 0
    del A


# ------------------------------------------------------------ edge kernels
def test_edge_assemble_and_grad(gpu, orc):
    cm = gpu
    from cumat_amd.problems import grid_laplacian
    rng = np.random.default_rng(21)
    P = grid_laplacian((17, 23), lambda s: np.exp(0.5 * rng.standard_normal(s)))
    B = 3
    thetas = np.stack([P["theta"] * (1 + 0.1 * k) for k in range(B)])
    d = {k: cm.DeviceArray.from_numpy(P[k]) for k in ("eptr", "eidx", "ei", "ej")}
    dth = cm.DeviceArray.from_numpy(thetas.reshape(-1))
    dv = cm.DeviceArray(P["nnz"] * B, dtype=np.float64)
    cm.capi.call("cm_edge_assemble", None, cm.capi.F64, P["nnz"], B, P["nedges"], d["eptr"].ptr, d["eidx"].ptr,
                 dth.ptr, dv.ptr)
    got = dv.numpy().reshape(B, -1)
    for k in range(B):
        np.testing.assert_allclose(got[k], P["vals"] * (1 + 0.1 * k), rtol=1e-14, atol=1e-14)
    lam = rng.standard_normal((B, P["n"]))
    x = rng.standard_normal((B, P["n"]))
    dl, dx = cm.DeviceArray.from_numpy(lam.reshape(-1)), cm.DeviceArray.from_numpy(x.reshape(-1))
    dg = cm.DeviceArray(P["nedges"] * B, dtype=np.float64)
    cm.capi.call("cm_edge_grad", None, cm.capi.F64, P["nedges"], B, P["n"], d["ei"].ptr, d["ej"].ptr, dl.ptr,
                 dx.ptr, dg.ptr)
    g = dg.numpy().reshape(B, -1)
    for k in range(B):
        np.testing.assert_array_equal(g[k], orc.edge_grad(P["ei"], P["ej"], lam[k], x[k]))


@pytest.mark.parametrize("per_batch", [1, 0])
def test_large_batched_solve_x_update_in_p_kernel(gpu, orc, per_batch):
    """From 2^20 unknowns on, x += alpha p moves into the p-update launch and a
    batch's last update is applied at the end: per-batch and all-batch
    stopping both match the oracle per system (3 x 600^2 = 1.08 M unknowns)."""
    cm = gpu
    from cumat_amd.problems import poisson2d
    g, B = 600, 3
    n = g * g
    JA, IA, _ = poisson2d(g, 0.0, np.float64)
    shifts = [0.02, 0.2, 1.0]
    vals = np.concatenate([poisson2d(g, s, np.float64)[2] for s in shifts])
    b = np.random.default_rng(31).uniform(-1, 1, n * B)
    dJA, dIA, dV = cm.DeviceArray.from_numpy(JA), cm.DeviceArray.from_numpy(IA), cm.DeviceArray.from_numpy(vals)
    db = cm.DeviceArray.from_numpy(b)
    dx = cm.DeviceArray(n * B, dtype=np.float64)
    its = (C.c_int * B)()
    errs = (C.c_double * B)()
    st = cm.capi.CgStats()
    tol = 1e-8
    cm.capi.call("cm_pcg_solve_ex", None, cm.capi.F64, n, IA.size, B, dJA.ptr, dIA.ptr, dV.ptr, 1, None, db.ptr,
                 dx.ptr, 0, tol, 4000, per_batch, its, errs, C.byref(st))
    x = dx.numpy()
    ref = [orc.pcg(JA, IA, vals[k * IA.size:(k + 1) * IA.size], b[k * n:(k + 1) * n], n, tol=tol, max_iter=4000)
           for k in range(B)]
    # the reduction order differs from the oracle's sequential sums; near the
    # threshold that can move a 360 K-row FP64 solve by an iteration or two
    if per_batch:
        for k in range(B):
            xk, itk, _ = ref[k]
            assert abs(its[k] - itk) <= 3, (k, its[k], itk)
            assert errs[k] <= tol
            assert np.linalg.norm(x[k * n:(k + 1) * n] - xk) / np.linalg.norm(xk) <= 1e-6
        assert its[2] < its[0]
    else:
        # all batches iterate until the slowest converges: same count for all
        assert len(set(its)) == 1
        assert abs(its[0] - max(r[1] for r in ref)) <= 3
        for k in range(B):
            r = orc.csr_spmv(JA, IA, vals[k * IA.size:(k + 1) * IA.size], x[k * n:(k + 1) * n], n, n)
            assert np.linalg.norm(b[k * n:(k + 1) * n] - r) / np.linalg.norm(b[k * n:(k + 1) * n]) <= 2 * tol


def _check_batched(orc, P, xh, ih, eh, tol=1e-6):
    ro = P["row_off"]
    for s in range(ro.size - 1):
        lo, hi = ro[s], ro[s + 1]
        n = hi - lo
        JA = P["JA"][lo:hi + 1] - P["JA"][lo]
        IA = P["IA"][P["JA"][lo]:P["JA"][hi]] - lo
        V = P["vals"][P["JA"][lo]:P["JA"][hi]]
        x64, _, _ = orc.pcg(JA, IA, V.astype(np.float64), P["b"][lo:hi].astype(np.float64), n, tol=1e-13,
                            max_iter=5000)
        assert eh[s] <= tol and ih[s] > 0, (s, ih[s], eh[s])
        assert np.linalg.norm(xh[lo:hi] - x64) / np.linalg.norm(x64) <= 1e-5, s


def test_batched_graph_cache_same_buffers_larger_systems(gpu, orc):
    """Two graph-path calls on the same device buffers: small systems first,
    then larger ones in the same block-size bucket (default block size 1024
    for both). The second call must not replay the first call's captured
    launch (its register tile and LDS size fit only the small systems)."""
    cm = gpu
    from cumat_amd.problems import batched_poisson
    nsys = 40
    small = batched_poisson(nsys, 91, 92, seed=71)   # 8281..8464 rows
    large = batched_poisson(nsys, 118, 120, seed=72)  # 13924..14400 rows
    cap_rows = int(large["row_off"][-1])
    cap_nnz = int(large["IA"].size)
    bufs = {"row_off": cm.DeviceArray(nsys + 1, dtype=np.int32), "JA": cm.DeviceArray(cap_rows + 1, dtype=np.int32),
            "IA": cm.DeviceArray(cap_nnz, dtype=np.int32), "vals": cm.DeviceArray(cap_nnz, dtype=np.float32),
            "b": cm.DeviceArray(cap_rows, dtype=np.float32)}
    x = cm.DeviceArray(cap_rows, dtype=np.float32)
    its = cm.DeviceArray(nsys, dtype=np.int32)
    errs = cm.DeviceArray(nsys, dtype=np.float32)
    for P in (small, large):
        nrows, nnz = int(P["row_off"][-1]), int(P["IA"].size)
        for k in bufs:
            a = np.ascontiguousarray(P[k])
            cm.capi.call("cm_memcpy_h2d", None, bufs[k].ptr, a.ctypes.data, a.nbytes)
        # views of the leading parts of the same buffers (same addresses)
        view = {k: cm.DeviceArray(np.asarray(P[k]).size, dtype=np.asarray(P[k]).dtype, ptr=bufs[k].ptr)
                for k in bufs}
        xv = cm.DeviceArray(nrows, dtype=np.float32, ptr=x.ptr)
        xv.fill_zero()
        cm.pcg_batched(view["row_off"], view["JA"], view["IA"], view["vals"], view["b"], xv, 1e-6, 3000, its, errs,
                       use_graph=True)
        assert nnz == view["IA"].size
        _check_batched(orc, P, xv.numpy(), its.numpy(), errs.numpy())


def test_batched_rejects_columns_outside_own_system(gpu):
    """A system whose column index points into another system's rows is not
    solved (the device reports iterations -1) and the call fails with an
    argument error; the other systems are still solved."""
    cm = gpu
    from cumat_amd.problems import batched_poisson
    P = batched_poisson(6, 10, 14, seed=3)
    IA = P["IA"].copy()
    ro, JA = P["row_off"], P["JA"]
    IA[JA[ro[2]] + 1] = ro[3] + 1  # system 2 reaches into system 3
    d = {k: cm.DeviceArray.from_numpy(P[k]) for k in ("row_off", "JA", "vals", "b")}
    dIA = cm.DeviceArray.from_numpy(IA)
    x = cm.DeviceArray(P["b"].size, dtype=np.float32)
    its = cm.DeviceArray(6, dtype=np.int32)
    errs = cm.DeviceArray(6, dtype=np.float32)
    for g in (False, True):
        with pytest.raises(cm.capi.CumatArgError):
            cm.pcg_batched(d["row_off"], d["JA"], dIA, d["vals"], d["b"], x, 1e-6, 500, its, errs, use_graph=g)
        ih = its.numpy()
        assert ih[2] == -1 and (np.delete(ih, 2) > 0).all()


def test_stream_ordered_frees_issue_no_host_sync(gpu):
    """Temporaries of an expression chain return to the cache without a host
    wait (stream-ordered frees); reuse of a freed block on the same stream is
    ordered by the stream, so the results stay right."""
    cm = gpu
    n = 1 << 20
    rng = np.random.default_rng(5)
    a_h = rng.standard_normal(n).astype(np.float32)
    b_h = rng.standard_normal(n).astype(np.float32)
    a, b = cm.DeviceArray.from_numpy(a_h), cm.DeviceArray.from_numpy(b_h)
    out = cm.DeviceArray(n, dtype=np.float32)
    acc = cm.DeviceArray(n, dtype=np.float32)
    acc.fill_zero()
    gc.collect()  # arrays of earlier tests must not be freed inside the counted loop
    warm = cm.DeviceArray(1, dtype=np.float32)
    cm.dot(a, b, out=warm)  # sizes the reduction scratch before the counted loop
    cm.synchronize()
    cm.reset_counters()
    for k in range(20):
        t = cm.DeviceArray(n, dtype=np.float32)       # cached block, reused every pass
        cm.assign(t, a.expr() * float(k) + b.expr())
        cm.assign(acc, acc.expr() + t.expr())
        d = cm.DeviceArray(1, dtype=np.float32)
        cm.dot(t, a, out=d)
        t.free()                                      # stream-ordered: no wait
        d.free()
    cm.assign(out, acc.expr() * 1.0)
    c = cm.counters()
    assert c["HostSync"] == 0, c
    assert c["DeviceMemFree"] == 40
    exp = sum(a_h.astype(np.float64) * k + b_h for k in range(20))
    got = out.numpy()
    assert np.abs(got - exp).max() <= 1e-4 * np.abs(exp).max()
