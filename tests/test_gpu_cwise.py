"""GPU parity of the fused component-wise evaluator (one kernel per
assignment) against the reference's fixtures and the CPU oracle's functors
(cuMat/src/CwiseOp.h:15-85, cuMat/src/BinaryOps.h:230-243,
cuMat/src/UnaryOps.h:128-225)."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_tests.json")))
B = {"add": 0, "sub": 1, "mul": 2, "div": 3, "mod": 4, "pow": 5, "min": 6, "max": 7, "and": 8, "or": 9,
     "xor": 10, "land": 11, "lor": 12, "eq": 13, "ne": 14, "lt": 15, "gt": 16, "le": 17, "ge": 18}


def test_compound_golden(gpu):
    """tests/TestCompoundAssignment.cu:9-43 through assignment modes."""
    cm = gpu
    g = GOLD["compound_2x2"]
    m1 = cm.DeviceArray.from_numpy(np.array(g["mat1"], np.int32), row_major=True)
    m2 = cm.DeviceArray(2, 2, 1, np.int32, row_major=True)
    cm.assign(m2, m1)
    cm.reset_counters()
    cm.assign(m2, m1, mode=cm.capi.ADD)
    assert cm.counters()["EvalCwise"] == 1
    assert np.array_equal(m2.numpy(logical=True)[0], np.array(g["add_self"]))
    m2.fill_zero()
    cm.assign(m2, m1, mode=cm.capi.SUB)
    assert np.array_equal(m2.numpy(logical=True)[0], np.array(g["zero_minus"]))
    cm.assign(m2, m1.expr() + 3)
    cm.assign(m2, m1, mode=cm.capi.MOD)
    assert np.array_equal(m2.numpy(logical=True)[0], np.array(g["plus3_mod"]))
    cm.assign(m2, m1)
    cm.assign(m2, cm.as_expr(2, m1.expr()), mode=cm.capi.MUL)
    assert np.array_equal(m2.numpy(logical=True)[0], np.array(g["times2"]))


def test_broadcast_golden(gpu):
    """tests/TestBinaryOps2.cu:13-60: host scalar and device scalar (1x1x1)."""
    cm = gpu
    g = GOLD["broadcast_mul_2x3x4"]
    for rm in (True, False):
        m = cm.DeviceArray.from_numpy(np.array(g["data"], np.int32), row_major=rm)
        out = cm.evaluate(2 * m.expr(), row_major=rm)
        assert np.array_equal(out.numpy(logical=True), np.array(g["times2"]))
        two = cm.DeviceArray.from_numpy(np.array([2], np.int32))
        out = cm.evaluate(m.expr() * two.expr(), row_major=not rm)
        assert np.array_equal(out.numpy(logical=True), np.array(g["times2"]))


def test_diagonal_golden(gpu):
    cm = gpu
    g = GOLD["diagonal_3x4"]
    m = cm.DeviceArray.from_numpy(np.array(g["data"], np.int32), row_major=True)
    d = cm.evaluate(m.expr().diagonal())
    assert np.array_equal(d.numpy(), np.array(g["diagonal"], np.int32))
    # asDiagonal round trip
    D = cm.evaluate(d.expr().asDiagonal())
    assert np.array_equal(D.numpy(logical=True)[0], np.diag(g["diagonal"]))


@pytest.mark.parametrize("fn", ["add", "sub", "mul", "div", "mod", "min", "max", "and", "or", "xor", "eq", "lt",
                                "ge", "land", "lor"])
def test_int32_binary_bitexact(gpu, orc, fn):
    cm = gpu
    rng = np.random.default_rng(B[fn])
    n = 100003
    a = rng.integers(-2**31, 2**31 - 1, n, dtype=np.int64).astype(np.int32)
    b = rng.integers(-2**31, 2**31 - 1, n, dtype=np.int64).astype(np.int32)
    if fn in ("div", "mod"):
        b[b == 0] = 7
        b[(a == -2**31) & (b == -1)] = 3
    if fn in ("land", "lor"):
        a[::3] = 0
        b[::5] = 0
    da, db = cm.DeviceArray.from_numpy(a), cm.DeviceArray.from_numpy(b)
    e = getattr(cm.expr, "Binary")(B[fn], da.expr(), db.expr())
    got = cm.evaluate(e).numpy()
    exp = orc.cwise_binary(B[fn], a, b)
    assert np.array_equal(got.astype(exp.dtype), exp)


@pytest.mark.parametrize("dtype,modes", [(np.int32, ("ADD", "SUB", "MUL", "DIV", "MOD", "AND", "OR")),
                                          (np.float32, ("ADD", "SUB", "MUL", "DIV")),
                                          (np.float64, ("ADD", "SUB", "MUL", "DIV"))])
def test_compound_modes_large_bitexact(gpu, orc, dtype, modes):
    """Every compound assignment mode (CwiseAssignmentHandler,
    cuMat/src/CwiseOp.h:15-62) on 100003 elements -- full 16-byte groups and
    a ragged tail -- bit-exact against the oracle, for a leaf source and for
    an expression source."""
    cm = gpu
    rng = np.random.default_rng(77)
    n = 100003
    for mode_name in modes:
        mode = getattr(cm.capi, mode_name)
        if dtype == np.int32:
            d0 = rng.integers(-2**20, 2**20, n).astype(np.int32)
            s0 = rng.integers(-1000, 1000, n).astype(np.int32)
            s0[s0 == 0] = 5
        else:
            d0 = rng.standard_normal(n).astype(dtype)
            s0 = rng.standard_normal(n).astype(dtype)
        dst = cm.DeviceArray.from_numpy(d0)
        src = cm.DeviceArray.from_numpy(s0)
        cm.assign(dst, src.expr(), mode=mode)
        exp = orc.assign(mode, d0, s0)
        assert np.array_equal(dst.numpy(), exp), mode_name
        # expression source: (s + s) evaluated in the same pass
        dst2 = cm.DeviceArray.from_numpy(d0)
        cm.assign(dst2, src.expr() + src.expr(), mode=mode)
        exp2 = orc.assign(mode, d0, (s0 + s0).astype(dtype))
        assert np.array_equal(dst2.numpy(), exp2), mode_name + " (expression)"


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("fn", ["add", "sub", "mul", "div", "min", "max", "lt", "eq"])
def test_float_binary_bitexact(gpu, orc, dtype, fn):
    """Single IEEE operations are correctly rounded on both sides."""
    cm = gpu
    rng = np.random.default_rng(B[fn] + 100)
    n = 65537
    a = rng.standard_normal(n).astype(dtype)
    b = rng.standard_normal(n).astype(dtype)
    b[::11] = a[::11]
    da, db = cm.DeviceArray.from_numpy(a), cm.DeviceArray.from_numpy(b)
    got = cm.evaluate(cm.expr.Binary(B[fn], da.expr(), db.expr())).numpy()
    exp = orc.cwise_binary(B[fn], a, b)
    assert np.array_equal(got.astype(exp.dtype), exp)


@pytest.mark.parametrize("name,fn,rtol", [("neg", 0, 0), ("abs", 1, 0), ("abs2", 2, 0), ("inverse", 3, 0),
                                          ("inverseCheck", 4, 0), ("sqrt", 5, 0), ("exp", 8, 4e-7),
                                          ("log", 9, 4e-7), ("sin", 12, 4e-7), ("floor", 21, 0),
                                          ("tanh", 20, 4e-7)])
def test_float_unary(gpu, orc, name, fn, rtol):
    cm = gpu
    rng = np.random.default_rng(fn)
    a = rng.uniform(0.01, 4.0, 50001).astype(np.float32)
    a[::7] *= -1
    if name in ("sqrt", "log"):
        a = np.abs(a)
    if name == "inverseCheck":
        a[::13] = 0
    da = cm.DeviceArray.from_numpy(a)
    got = cm.evaluate(cm.expr.Unary(fn, da.expr())).numpy()
    exp = orc.cwise_unary(fn, a)
    if rtol == 0:
        assert np.array_equal(got, exp)
    else:
        np.testing.assert_allclose(got, exp, rtol=rtol, atol=1e-7)


def test_fused_expression_tree(gpu, orc):
    """A deep mixed tree (stack slots, reversed operands, casts, broadcasts)
    in one launch, checked against the oracle's functors composed on the
    host."""
    cm = gpu
    rng = np.random.default_rng(42)
    R, Cc, Bt = 37, 23, 5
    a = rng.integers(-50, 50, R * Cc * Bt).astype(np.int32)
    b = rng.integers(1, 20, R * Cc * Bt).astype(np.int32)
    col = rng.integers(-9, 9, R).astype(np.int32)  # R x 1 x 1 broadcast
    da = cm.DeviceArray.from_numpy(a, rows=R, cols=Cc, batches=Bt)
    db = cm.DeviceArray.from_numpy(b, rows=R, cols=Cc, batches=Bt)
    dc = cm.DeviceArray.from_numpy(col, rows=R, cols=1, batches=1)
    e = (da.expr() * 3 - db.expr()) % (db.expr() + 2) + (dc.expr() - da.expr()) * (db.expr() ^ 5)
    cm.reset_counters()
    got = cm.evaluate(e).numpy()
    assert cm.counters()["EvalCwise"] == 1
    cb = np.tile(col, Cc * Bt)
    t1 = orc.cwise_binary(B["sub"], orc.cwise_binary(B["mul"], a, np.full_like(a, 3)), b)
    t2 = orc.cwise_binary(B["add"], b, np.full_like(b, 2))
    t3 = orc.cwise_binary(B["mod"], t1, t2)
    t4 = orc.cwise_binary(B["sub"], cb, a)
    t5 = orc.cwise_binary(B["xor"], b, np.full_like(b, 5))
    exp = orc.cwise_binary(B["add"], t3, orc.cwise_binary(B["mul"], t4, t5))
    assert np.array_equal(got, exp)
    # cast + float math + row-major destination
    f = cm.evaluate(cm.cast(da.expr(), np.float32) * 0.5 + cm.cast(db.expr(), np.float32), row_major=True)
    expf = (a.astype(np.float32) * np.float32(0.5) + b.astype(np.float32)).reshape(Bt, Cc, R).transpose(0, 2, 1)
    assert np.array_equal(f.numpy(logical=True), expf)


def test_lincomb8_fused(gpu, orc):
    """out = sum_i a_i v_i over 8 vectors (the linear-combination benchmark)
    by the expression path and by the pre-instantiated kernel."""
    cm = gpu
    import ctypes as C
    rng = np.random.default_rng(8)
    n = (1 << 20) + 3
    vs = [rng.uniform(-1, 1, n).astype(np.float32) for _ in range(8)]
    cs = rng.uniform(-1, 1, 8).astype(np.float32)
    dv = [cm.DeviceArray.from_numpy(v) for v in vs]
    e = dv[0].expr() * float(cs[0])
    for i in range(1, 8):
        e = e + dv[i].expr() * float(cs[i])
    cm.reset_counters()
    got = cm.evaluate(e).numpy()
    assert cm.counters()["EvalCwise"] == 1
    exp = sum(v.astype(np.float64) * float(c) for v, c in zip(vs, cs))
    err = np.abs(got - exp).max() / np.abs(exp).max()
    assert err < 1e-6
    # '+=' into zeros runs the general program interpreter (the '=' form is
    # dispatched to the streaming linear-combination kernel): same values
    acc = cm.DeviceArray(n, dtype=np.float32)
    acc.fill_zero()
    cm.assign(acc, e, mode=cm.capi.ADD)
    assert np.abs(acc.numpy() - exp).max() / np.abs(exp).max() < 1e-6
    out = cm.DeviceArray(n, dtype=np.float32)
    ptrs = (C.c_void_p * 8)(*[d.ptr for d in dv])
    coef = (C.c_float * 8)(*cs)
    cm.capi.call("cm_lincomb_f32", None, 8, ptrs, coef, out.ptr, n)
    got2 = out.numpy()
    assert np.abs(got2 - exp).max() / np.abs(exp).max() < 1e-6


def test_transpose_and_blocks(gpu):
    cm = gpu
    rng = np.random.default_rng(4)
    m = rng.integers(-100, 100, (3, 40, 30)).astype(np.int32)
    dm = cm.DeviceArray.from_numpy(m)
    t = cm.evaluate(dm.expr().T)
    assert np.array_equal(t.numpy(logical=True), np.transpose(m, (0, 2, 1)))


def test_dimension_mismatch_raises(gpu):
    cm = gpu
    a = cm.DeviceArray(4, 3, 1, np.float32)
    b = cm.DeviceArray(3, 4, 1, np.float32)
    with pytest.raises(cm.CumatArgError):
        cm.evaluate(a.expr() + b.expr())
    with pytest.raises(TypeError):
        cm.evaluate(a.expr() + cm.DeviceArray(4, 3, 1, np.int32).expr())
