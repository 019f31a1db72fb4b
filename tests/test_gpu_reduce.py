"""GPU parity of the batched reductions and dot/squaredNorm against the
reference's integer fixtures (bit-exact) and the CPU oracle (restating
cuMat/src/ReductionOps.h:57-160,859-968 and
cuMat/src/ReductionOpsPlugin.inl:124-450)."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_tests.json")))
OPS = {"sum": 0, "prod": 1, "min": 2, "max": 3, "all": 4, "any": 5, "bitwise_and": 6, "bitwise_or": 7}


def _dev(cm, data, dtype, row_major):
    a = np.array(data, dtype)
    return cm.DeviceArray.from_numpy(a, row_major=row_major)


@pytest.mark.parametrize("row_major", [True, False])
@pytest.mark.parametrize("fixture,op", [("reduce_2x4x3_rowmajor", "sum"), ("reduce_2x4x3_rowmajor", "prod"),
                                        ("reduce_2x3x4_axis", "sum"), ("reduce_2x3x4_axis", "min"),
                                        ("reduce_2x3x4_axis", "max"), ("reduce_prod_2x3x2", "prod")])
def test_golden_integer(gpu, row_major, fixture, op):
    cm = gpu
    g = GOLD[fixture]
    a = _dev(cm, g["data"], np.int32, row_major)
    for axis, exp in g[op].items():
        out = cm.reduce(a, op, int(axis), out_row_major=True)
        got = out.numpy(logical=True)
        # expected values are listed in row-major storage order of the output
        assert np.array_equal(got.reshape(-1), np.array(exp, np.int32)), (fixture, op, axis)


def test_golden_all_any(gpu):
    cm = gpu
    g = GOLD["reduce_all_any_2x3x4"]
    for rm in (True, False):
        a = _dev(cm, np.array(g["all_data"], bool), bool, rm)
        assert np.array_equal(cm.reduce(a, "all", 3).numpy(), np.array(g["all_rc"], bool))
        a = _dev(cm, np.array(g["any_data"], bool), bool, rm)
        assert np.array_equal(cm.reduce(a, "any", 3).numpy(), np.array(g["any_rc"], bool))


def test_golden_dot_norm(gpu):
    cm = gpu
    g = GOLD["dot_batched"]
    a = cm.DeviceArray.from_numpy(np.array(g["a"], np.int32).reshape(-1), rows=3, cols=1, batches=2)
    b = cm.DeviceArray.from_numpy(np.array(g["b"], np.int32))
    assert np.array_equal(cm.dot(a, b).numpy(), np.array(g["dot"], np.int32))
    assert np.array_equal(cm.dot(b, a).numpy(), np.array(g["dot"], np.int32))
    g = GOLD["norm_2x1x4"]
    v = cm.DeviceArray.from_numpy(np.array(g["data"], np.float64).reshape(-1), rows=4, cols=1, batches=2)
    assert np.array_equal(cm.squared_norm(v).numpy(), np.array(g["squared_norm"], np.float64))
    np.testing.assert_allclose(cm.norm(v), np.array(g["norm"]), rtol=1e-15)


def _rand(rng, dtype, n):
    if dtype == np.int32:
        return rng.integers(-1000, 1001, n).astype(np.int32)
    if dtype == np.bool_:
        return rng.random(n) < 0.7
    return rng.standard_normal(n).astype(dtype)


def _check(got, exp, dtype, op):
    if dtype in (np.int32, np.bool_) or op in ("min", "max"):
        assert np.array_equal(got, exp)
    else:
        tol = 1e-5 if dtype == np.float32 else 1e-12
        err = np.linalg.norm(got.astype(np.float64) - exp) / max(np.linalg.norm(exp.astype(np.float64)), 1e-30)
        assert err <= tol, err


@pytest.mark.parametrize("dtype", [np.int32, np.float32, np.float64])
@pytest.mark.parametrize("shape", [(1, 1, 1), (5, 3, 2), (64, 33, 7), (1000, 3, 5), (3, 1000, 5), (7, 9, 2000),
                                   (2000, 700, 1)])
@pytest.mark.parametrize("row_major", [False, True])
def test_random_all_axes(gpu, orc, dtype, shape, row_major):
    cm = gpu
    R, Cc, B = shape
    rng = np.random.default_rng(hash((R, Cc, B, row_major)) % 2**32)
    flat = _rand(rng, dtype, R * Cc * B)
    a = cm.DeviceArray.from_numpy(flat, rows=R, cols=Cc, batches=B, row_major=row_major)
    ops = ["sum", "min", "max"] + (["prod", "bitwise_and", "bitwise_or"] if dtype == np.int32 else [])
    for op in ops:
        for axis in range(8):
            for orm in (row_major, not row_major):
                got = cm.reduce(a, op, axis, out_row_major=orm).numpy()
                exp = orc.reduce(flat, OPS[op], axis, R, Cc, B, row_major, orm)
                _check(got, exp, dtype, op)


def test_bool_all_any_random(gpu, orc):
    cm = gpu
    rng = np.random.default_rng(1)
    for shape in [(13, 17, 3), (300, 2, 40)]:
        R, Cc, B = shape
        flat = rng.random(R * Cc * B) < 0.97
        a = cm.DeviceArray.from_numpy(flat, rows=R, cols=Cc, batches=B)
        for op in ("all", "any"):
            for axis in range(8):
                got = cm.reduce(a, op, axis).numpy()
                exp = orc.reduce(flat.astype(np.uint8), OPS[op], axis, R, Cc, B, False, False).astype(bool)
                assert np.array_equal(got, exp), (shape, op, axis)


@pytest.mark.parametrize("shape,row_major,axis", [
    ((512, 512, 7), False, 4),     # batch fold, 256 tiles of 1024 columns
    ((1000, 300, 9), False, 4),    # ragged last tile, fold not a multiple of 8
    ((1024, 256, 64), False, 4),   # 64 batches
    ((9, 2 ** 18, 2), True, 1),    # row fold of a row-major matrix, two batches
    ((40, 5000, 2), True, 1),      # too few columns for the streaming kernel
    ((1000, 300, 64), False, 2),   # long fold, one 1024-thread block per column tile, ragged tile
    ((400, 1000, 64), True, 1),    # the same for a row-major matrix
])
def test_short_fold_many_columns(gpu, orc, shape, row_major, axis):
    """Strided folds over many columns (the batch-sum shape, and long folds
    with a column tile per CU): int32
    exact against the oracle for sum / min / max / prod / and / or; FP32 sum
    norm-wise within 1e-5 of the FP64 sums."""
    cm = gpu
    R, Cc, B = shape
    rng = np.random.default_rng(R * 31 + Cc + B)
    flat = rng.integers(-1000, 1001, R * Cc * B).astype(np.int32)
    a = cm.DeviceArray.from_numpy(flat, rows=R, cols=Cc, batches=B, row_major=row_major)
    for op in ("sum", "min", "max", "prod", "bitwise_and", "bitwise_or"):
        got = cm.reduce(a, op, axis, out_row_major=row_major).numpy()
        exp = orc.reduce(flat, OPS[op], axis, R, Cc, B, row_major, row_major)
        assert np.array_equal(got, exp), op
    f = rng.standard_normal(R * Cc * B).astype(np.float32)
    af = cm.DeviceArray.from_numpy(f, rows=R, cols=Cc, batches=B, row_major=row_major)
    got = cm.reduce(af, "sum", axis, out_row_major=row_major).numpy()
    exp = orc.reduce(f.astype(np.float64), OPS["sum"], axis, R, Cc, B, row_major, row_major)
    _check(got, exp, np.float32, "sum")


def test_full_size_1024x1024x64(gpu):
    """The benchmark tensor at full size through size-independent properties:
    int32 sums are exact, so every axis order must give the same totals, and
    the row / column / batch sums must sum to the full sum."""
    cm = gpu
    R, Cc, B = 1024, 1024, 64
    rng = np.random.default_rng(20261202)
    flat = rng.integers(-1000, 1001, R * Cc * B, dtype=np.int32)
    a = cm.DeviceArray.from_numpy(flat, rows=R, cols=Cc, batches=B)
    total = int(flat.astype(np.int64).sum())
    wrap = np.int32(np.int64(total).astype(np.int32))
    assert cm.reduce(a, "sum", 7).numpy()[0] == wrap
    for axis in (1, 2, 4, 3, 5, 6):
        part = cm.reduce(a, "sum", axis)
        assert cm.reduce(part, "sum", 7).numpy()[0] == wrap, axis
    # row sums against numpy (exact in int64, no overflow at this size)
    rs = cm.reduce(a, "sum", 1).numpy()
    exp = flat.reshape(B, Cc, R).astype(np.int64).sum(axis=2).reshape(-1)
    assert np.array_equal(rs.astype(np.int64), exp)
    bs = cm.reduce(a, "sum", 4).numpy()
    assert np.array_equal(bs.astype(np.int64), flat.reshape(B, -1).astype(np.int64).sum(0))
    # float32 N(0,1): norm-wise relative error vs float64 sums
    f = rng.standard_normal(R * Cc * B).astype(np.float32)
    af = cm.DeviceArray.from_numpy(f, rows=R, cols=Cc, batches=B)
    f3 = f.reshape(B, Cc, R).astype(np.float64)
    for axis, exp in ((1, f3.sum(2)), (2, f3.sum(1)), (4, f3.sum(0))):
        got = cm.reduce(af, "sum", axis).numpy().astype(np.float64)
        err = np.linalg.norm(got - exp.reshape(-1)) / np.linalg.norm(exp)
        assert err <= 1e-5, (axis, err)


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.int32])
@pytest.mark.parametrize("n", [1, 10, 100, 1000, 10000, 100000, 1000000])
def test_dot_sizes(gpu, orc, dtype, n):
    """squaredNorm for n=10..1e6 (tests/TestReductionOps5.cu:13-26) and dot."""
    cm = gpu
    rng = np.random.default_rng(n)
    a = _rand(rng, dtype, n)
    b = _rand(rng, dtype, n)
    da, db = cm.DeviceArray.from_numpy(a), cm.DeviceArray.from_numpy(b)
    got = cm.dot(da, db).numpy()
    exp = orc.dot(a, b, n)
    sq = cm.squared_norm(da).numpy()
    exps = orc.dot(a, a, n)
    if dtype == np.int32:
        assert np.array_equal(got, exp) and np.array_equal(sq, exps)
    else:
        tol = 1e-5 if dtype == np.float32 else 1e-12
        scale = np.sqrt(orc.dot(a, a, n)[0] * orc.dot(b, b, n)[0])
        assert abs(float(got[0]) - float(exp[0])) <= tol * scale
        assert abs(float(sq[0]) - float(exps[0])) <= tol * float(exps[0])


def test_reductions_deterministic(gpu):
    cm = gpu
    f = np.random.default_rng(3).standard_normal(5_000_000).astype(np.float32)
    a = cm.DeviceArray.from_numpy(f)
    r1 = cm.reduce(a, "sum", 7).numpy()
    r2 = cm.reduce(a, "sum", 7).numpy()
    d1 = cm.dot(a, a).numpy()
    d2 = cm.dot(a, a).numpy()
    assert r1.tobytes() == r2.tobytes() and d1.tobytes() == d2.tobytes()
