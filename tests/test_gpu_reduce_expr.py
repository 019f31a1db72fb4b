"""GPU parity of reductions whose input is an unevaluated expression
(cm_reduce_expr). cuMat's reductions read their input through the expression
tree (cuMat/src/Iterator.h:99-105) and dot / squaredNorm reduce cwiseDot /
cwiseAbs2 (cuMat/src/ReductionOpsPlugin.inl:290-297,395-431): the result must
equal reducing the evaluated expression, in one reduction launch and without
an evaluation pass.

int32 results are checked bit-exact against the oracle's reduction
(orc_reduce, cuMat/src/ReductionOps.h:57-160) of the expression's values;
min / max of floating expressions are exact too; floating sums and products
norm-wise against float64 sums: 1e-5 for float32, 1e-12 for float64."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

OPS = {"sum": 0, "prod": 1, "min": 2, "max": 3, "all": 4, "any": 5, "bitwise_and": 6, "bitwise_or": 7}


def _pack_cm(a3):
    # (batches, rows, cols) logical -> column-major storage order
    return np.ascontiguousarray(np.transpose(a3, (0, 2, 1))).reshape(-1)


def _close(got, exp, tol):
    got = np.asarray(got, np.float64).reshape(-1)
    exp = np.asarray(exp, np.float64).reshape(-1)
    err = np.linalg.norm(got - exp) / max(np.linalg.norm(exp), 1e-30)
    assert err <= tol, err


def _int_inputs(rng, shape):
    B, R, Cc = shape
    return [rng.integers(-300, 301, (B, R, Cc)).astype(np.int32) for _ in range(3)]


@pytest.mark.parametrize("shape", [(1, 1, 1), (2, 5, 3), (7, 64, 33), (5, 1000, 3), (5, 3, 1000), (2000, 7, 9),
                                   (1, 2000, 700)])
@pytest.mark.parametrize("row_major", [False, True])
def test_int_expression_all_axes_and_ops(gpu, orc, shape, row_major):
    """(a*3 - b) ^ (c + 1) reduced over every axis subset with every integer
    operator, both output layouts: bit-exact against the oracle."""
    cm = gpu
    B, R, Cc = shape
    rng = np.random.default_rng(R * 7 + Cc * 3 + B)
    a, b, c = _int_inputs(rng, shape)
    da, db, dc = (cm.DeviceArray.from_numpy(x, row_major=row_major) for x in (a, b, c))
    e = (da.expr() * 3 - db.expr()) ^ (dc.expr() + 1)
    val = ((a * np.int32(3) - b) ^ (c + np.int32(1))).astype(np.int32)
    flat = _pack_cm(val)
    # int32 products wrap; wrapping multiplication is associative, so the
    # product is order-independent and exact too
    for op in ("sum", "min", "max", "bitwise_and", "bitwise_or", "prod"):
        for axis in range(8):
            for orm in (False, True):
                got = cm.reduce(e, op, axis, out_row_major=orm).numpy()
                exp = orc.reduce(flat, OPS[op], axis, R, Cc, B, False, orm)
                assert np.array_equal(got, exp), (op, axis, orm)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape", [(3, 17, 5), (4, 1000, 31), (1, 300000, 1), (600, 30, 20)])
def test_float_expression_axes(gpu, orc, dtype, shape):
    """a*b + c (the FMA-shaped expression) reduced over every axis subset:
    sums within tolerance of the float64 oracle, min / max exact."""
    cm = gpu
    B, R, Cc = shape
    rng = np.random.default_rng(R + Cc + B)
    a, b, c = (rng.standard_normal(shape).astype(dtype) for _ in range(3))
    da, db, dc = (cm.DeviceArray.from_numpy(x) for x in (a, b, c))
    e = da.expr() * db.expr() + dc.expr()
    # the GPU contracts a*b + c into one fused multiply-add; form it in float64
    # and round once, as fma does
    val = (a.astype(np.float64) * b + c).astype(dtype)
    flat = _pack_cm(val)
    tol = 1e-5 if dtype == np.float32 else 1e-12
    for axis in range(8):
        got = cm.reduce(e, "sum", axis).numpy()
        exp = orc.reduce(flat.astype(np.float64), OPS["sum"], axis, R, Cc, B, False, False)
        _close(got, exp, tol)
        for op in ("min", "max"):
            got = cm.reduce(e, op, axis).numpy()
            evaluated = cm.evaluate(e)
            exp2 = cm.reduce(evaluated, op, axis).numpy()
            assert np.array_equal(got, exp2), (op, axis)


@pytest.mark.parametrize("src", [np.int32, np.float32, np.float64])
def test_bool_expressions_every_operator(gpu, src):
    """Comparisons reduced in place ((m > 0).all() and friends; a float64
    comparison runs the 64-bit program): every operator and axis equals the
    stored-bool reduction of the evaluated comparison, bit for bit."""
    cm = gpu
    B, R, Cc = 3, 200, 41
    rng = np.random.default_rng(9)
    a = (rng.standard_normal((B, R, Cc)) + 1.8).astype(src)
    da = cm.DeviceArray.from_numpy(a)
    e = da.expr() > 0
    ev = cm.evaluate(e)
    for op in ("all", "any", "sum", "prod", "min", "max", "bitwise_and", "bitwise_or"):
        for axis in range(8):
            got = cm.reduce(e, op, axis).numpy()
            exp = cm.reduce(ev, op, axis).numpy()
            assert np.array_equal(got, exp), (op, axis)
    assert not cm.reduce(e, "all", 7).numpy()[0] and cm.reduce(e, "any", 7).numpy()[0]


def test_squared_norm_of_difference_is_one_launch(gpu, orc):
    """(a - b).squaredNorm(): one reduction, no evaluation pass, no
    temporary; equal to the oracle's dot of the difference."""
    cm = gpu
    rng = np.random.default_rng(5)
    n = 1 << 16
    a = rng.standard_normal(n).astype(np.float32)
    b = rng.standard_normal(n).astype(np.float32)
    da, db = cm.DeviceArray.from_numpy(a), cm.DeviceArray.from_numpy(b)
    cm.squared_norm(da.expr() - db.expr()).numpy()  # scratch allocated once
    cm.reset_counters()
    out = cm.squared_norm(da.expr() - db.expr())
    cnt = cm.counters()
    assert cnt["EvalReduction"] == 1 and cnt["EvalCwise"] == 0
    assert cnt["KernelLaunch"] <= 2 and cnt["DeviceMemAlloc"] == 0
    d = (a - b).astype(np.float64)
    assert abs(float(out.numpy()[0]) - float(d @ d)) <= 1e-5 * float(d @ d)


@pytest.mark.parametrize("dtype", [np.int32, np.float32, np.float64])
def test_dot_of_expressions_batched(gpu, orc, dtype):
    """dot(2*a + b, c) per batch, and a batch-1 expression broadcast against
    a batched operand (TestReductionOps4.cu:61-62 does this for stored
    matrices)."""
    cm = gpu
    B, R, Cc = 6, 300, 7
    rng = np.random.default_rng(11)
    if dtype == np.int32:
        mk = lambda s: rng.integers(-1000, 1001, s).astype(np.int32)  # noqa: E731
    else:
        mk = lambda s: rng.standard_normal(s).astype(dtype)  # noqa: E731
    a, b, c = mk((B, R, Cc)), mk((B, R, Cc)), mk((B, R, Cc))
    v = mk((1, R, Cc))
    da, db, dc, dv = (cm.DeviceArray.from_numpy(x) for x in (a, b, c, v))
    got = cm.dot(da.expr() * 2 + db.expr(), dc).numpy()
    val = (a * dtype(2) + b).astype(dtype)
    exp = orc.dot(_pack_cm(val), _pack_cm(c), R * Cc, B)
    got_b = cm.dot(dv.expr() + 1, dc.expr()).numpy()
    vb = np.broadcast_to(v + dtype(1), (B, R, Cc)).astype(dtype)
    exp_b = orc.dot(_pack_cm(vb), _pack_cm(c), R * Cc, B)
    if dtype == np.int32:
        assert np.array_equal(got, exp) and np.array_equal(got_b, exp_b)
    else:
        tol = 1e-5 if dtype == np.float32 else 1e-12
        scale = np.sqrt((val.astype(np.float64) ** 2).sum() * (c.astype(np.float64) ** 2).sum())
        assert np.max(np.abs(got.astype(np.float64) - exp)) <= tol * scale
        assert np.max(np.abs(got_b.astype(np.float64) - exp_b)) <= tol * scale


def test_transposed_and_broadcast_leaves(gpu):
    """Leaves read through a transpose and a batch broadcast: the virtual
    reduction order must map every coordinate to the right element (int32,
    exact against numpy)."""
    cm = gpu
    B, R, Cc = 4, 37, 53
    rng = np.random.default_rng(21)
    at = rng.integers(-50, 51, (B, Cc, R)).astype(np.int32)  # stored transposed
    v = rng.integers(-50, 51, (1, R, Cc)).astype(np.int32)
    dat = cm.DeviceArray.from_numpy(at, row_major=True)
    dv = cm.DeviceArray.from_numpy(v)
    e = dat.expr().T * 2 - dv.expr()
    val = (np.transpose(at, (0, 2, 1)) * 2 - v).astype(np.int64)
    for axis in range(8):
        ax = tuple(a for a, bit in ((1, 1), (2, 2), (0, 4)) if axis & bit)
        exp = val.sum(axis=ax, keepdims=True) if ax else val
        got = cm.reduce(e, "sum", axis).numpy(logical=True)
        assert np.array_equal(got.astype(np.int64), exp.reshape(got.shape)), axis


def test_full_size_1024x1024x64_properties(gpu):
    """The benchmark tensor at full size: an int32 expression's column sums
    re-reduced equal its total, which equals the total of the evaluated
    expression; a float32 squared norm of a difference matches the stored
    path within 1e-5 relative."""
    cm = gpu
    R, Cc, B = 1024, 1024, 64
    rng = np.random.default_rng(77)
    a = rng.integers(-1000, 1001, R * Cc * B, dtype=np.int32)
    b = rng.integers(-1000, 1001, R * Cc * B, dtype=np.int32)
    da = cm.DeviceArray.from_numpy(a, rows=R, cols=Cc, batches=B)
    db = cm.DeviceArray.from_numpy(b, rows=R, cols=Cc, batches=B)
    e = da.expr() - db.expr() * 3
    total = np.int64((a.astype(np.int64) - 3 * b.astype(np.int64)).sum()).astype(np.int32)
    assert cm.reduce(e, "sum", 7).numpy()[0] == total
    for axis in (1, 2, 4, 3, 5, 6):
        part = cm.reduce(e, "sum", axis)
        assert cm.reduce(part, "sum", 7).numpy()[0] == total, axis
    ev = cm.evaluate(e)
    for axis in (1, 2, 4):
        for op in ("min", "max", "sum"):
            assert np.array_equal(cm.reduce(e, op, axis).numpy(), cm.reduce(ev, op, axis).numpy()), (op, axis)
    f = rng.standard_normal(R * Cc * B).astype(np.float32)
    g = rng.standard_normal(R * Cc * B).astype(np.float32)
    df = cm.DeviceArray.from_numpy(f, rows=R, cols=Cc, batches=B)
    dg = cm.DeviceArray.from_numpy(g, rows=R, cols=Cc, batches=B)
    got = cm.squared_norm(df.expr() - dg.expr()).numpy().astype(np.float64)
    d = (f - g).reshape(B, -1).astype(np.float64)
    exp = (d * d).sum(axis=1)
    assert np.max(np.abs(got - exp) / exp) <= 1e-5


def test_deterministic_and_rejections(gpu):
    cm = gpu
    f = np.random.default_rng(3).standard_normal(3_000_000).astype(np.float32)
    a = cm.DeviceArray.from_numpy(f)
    e = a.expr() * 1.5 - 0.25
    r1 = cm.reduce(e, "sum", 7).numpy()
    r2 = cm.reduce(e, "sum", 7).numpy()
    assert r1.tobytes() == r2.tobytes()
    with pytest.raises(cm.CumatError):
        cm.reduce(e, "bitwise_and", 7)  # integer-only operator on a float expression
    assert cm.reduce(e, "any", 7).numpy()[0] == 1.0
    bad_out = cm.DeviceArray(2, 1, 1, dtype=np.float32)
    with pytest.raises(ValueError):
        cm.reduce(e, "sum", 7, out=bad_out)
