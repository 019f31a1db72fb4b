"""Run-time generated component-wise kernels (cm_jit_mode) against the
program interpreter, bit for bit: every binary and unary functor in FP32,
FP64 and int32, casts between all dtypes, select, the FMA term, every
assignment mode, bool destinations, broadcast scalars, unaligned leaves and
lengths that end inside a thread's group of elements. Both paths evaluate
the instruction semantics of cwise_ops.hpp; the oracle pins the interpreter
(test_gpu_cwise.py), this file pins the generated kernels to it."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = (1 << 20) + 2051  # not a multiple of a thread's 2048-element span


def _instr(op, dt, fn=0, src=0, arg=0, flags=0, fimm=0.0, iimm=0):
    from cumat_amd import _capi as capi
    i = capi.Instr()
    i.opcode, i.dtype, i.fn, i.src, i.arg, i.flags, i.fimm, i.iimm = op, dt, fn, src, arg, flags, fimm, iimm
    return i


class _Leaf:
    def __init__(self, cm, arr, offset=0, scalar=False):
        self.arr = arr
        self.scalar = scalar
        pad = np.zeros(offset, arr.dtype)
        self.dev = cm.DeviceArray.from_numpy(np.concatenate([pad, arr]))
        self.ptr = self.dev.ptr + offset * arr.dtype.itemsize


def _eval(cm, leaves, code, dst_np_dtype, mode, init, jit):
    from cumat_amd import _capi as capi
    from cumat_amd.device import np_to_dtype
    n = init.size
    L = (capi.Leaf * max(1, len(leaves)))()
    for i, lf in enumerate(leaves):
        L[i].ptr = lf.ptr
        L[i].dtype = np_to_dtype(lf.arr.dtype)
        L[i].kind = capi.LEAF_STRIDED
        L[i].sr, L[i].sc, L[i].sb = (0, 0, 0) if lf.scalar else (1, n, n)
    prog = (capi.Instr * len(code))(*code)
    dst = cm.DeviceArray.from_numpy(init.astype(dst_np_dtype))
    old = cm.jit_mode(jit)
    try:
        capi.call("cm_cwise_eval", None, L, len(leaves), prog, len(code), dst.ptr, np_to_dtype(np.dtype(dst_np_dtype)),
                  n, 1, 1, 1, n, n, mode)
        out = dst.numpy()
    finally:
        cm.jit_mode(old)
    return out


def _same(cm, leaves, code, dst_dtype, mode=0, init=None, n=N):
    if init is None:
        init = np.zeros(n, dst_dtype)
    cm.reset_counters()
    ref = _eval(cm, leaves, code, dst_dtype, mode, init, jit=0)
    assert cm.counters()["JitLaunch"] == 0
    got = _eval(cm, leaves, code, dst_dtype, mode, init, jit=2)
    assert cm.counters()["JitLaunch"] == 1, "the generated kernel did not run"
    assert ref.tobytes() == got.tobytes()
    return got


@pytest.fixture(scope="module")
def data(gpu):
    cm = gpu
    rng = np.random.default_rng(2027)
    f = rng.standard_normal(N).astype(np.float32) * 3
    g = rng.uniform(0.1, 4, N).astype(np.float32)
    g[::97] = 0.0  # zeros for division, inverse and comparisons
    i = rng.integers(-1000, 1000, N).astype(np.int32)
    j = rng.integers(-20, 20, N).astype(np.int32)  # includes zeros (division by zero gives 0)
    b = rng.integers(0, 2, N).astype(np.bool_)
    return {k: _Leaf(cm, v) for k, v in dict(f=f, g=g, i=i, j=j, b=b).items()}


@pytest.mark.parametrize("fn", list(range(19)))
def test_binary_f32(gpu, data, fn):
    from cumat_amd import _capi as capi
    if fn in (capi.BIN_MOD, capi.BIN_AND, capi.BIN_OR, capi.BIN_XOR, capi.BIN_LAND, capi.BIN_LOR):
        pytest.skip("not an FP32 functor")
    cmp = fn >= capi.BIN_EQ
    code = [_instr(capi.OP_LOAD, capi.F32, src=capi.SRC_LEAF, arg=0),
            _instr(capi.OP_BINARY, capi.F32, fn, capi.SRC_LEAF, 1)]
    _same(gpu, [data["f"], data["g"]], code, np.bool_ if cmp else np.float32)
    # reversed operands against an immediate
    code = [_instr(capi.OP_LOAD, capi.F32, src=capi.SRC_LEAF, arg=0),
            _instr(capi.OP_BINARY, capi.F32, fn, capi.SRC_IMM, 0, flags=1, fimm=1.75)]
    _same(gpu, [data["g"]], code, np.bool_ if cmp else np.float32)


@pytest.mark.parametrize("fn", list(range(19)))
def test_binary_i32(gpu, data, fn):
    from cumat_amd import _capi as capi
    if fn == capi.BIN_POW:
        pytest.skip("not an int32 functor")
    code = [_instr(capi.OP_LOAD, capi.I32, src=capi.SRC_LEAF, arg=0),
            _instr(capi.OP_BINARY, capi.I32, fn, capi.SRC_LEAF, 1)]
    _same(gpu, [data["i"], data["j"]], code, np.int32)


@pytest.mark.parametrize("fn", list(range(27)))
def test_unary(gpu, data, fn):
    from cumat_amd import _capi as capi
    if fn != capi.UN_BITNOT:
        code = [_instr(capi.OP_LOAD, capi.F32, src=capi.SRC_LEAF, arg=0), _instr(capi.OP_UNARY, capi.F32, fn)]
        _same(gpu, [data["g"] if fn in (capi.UN_ACOS, capi.UN_ASIN) else data["f"]], code, np.float32)
    code = [_instr(capi.OP_LOAD, capi.I32, src=capi.SRC_LEAF, arg=0), _instr(capi.OP_UNARY, capi.I32, fn)]
    _same(gpu, [data["j"]], code, np.int32)


def test_casts_select_fma_and_scalars(gpu, data):
    """int -> float with a broadcast scalar, exp, an FMA term, a select on a
    bool leaf, a cast to bool, a logical or; bool destination with |=."""
    from cumat_amd import _capi as capi
    cm = gpu
    s = _Leaf(cm, np.array([0.37], np.float32), scalar=True)
    code = [_instr(capi.OP_LOAD, capi.F32, src=capi.SRC_LEAF, arg=0),       # float(i)
            _instr(capi.OP_BINARY, capi.F32, capi.BIN_MUL, capi.SRC_LEAF, 3),  # * scalar
            _instr(capi.OP_UNARY, capi.F32, capi.UN_TANH),
            _instr(capi.OP_FMA, capi.F32, arg=2, fimm=0.5),                    # + f * 0.5
            _instr(capi.OP_PUSH, capi.F32, arg=1),
            _instr(capi.OP_LOAD, capi.BOOL, src=capi.SRC_LEAF, arg=1),
            _instr(capi.OP_SELECT, capi.F32, src=capi.SRC_IMM, arg=1, fimm=2.5),
            _instr(capi.OP_CAST, capi.I32, arg=capi.F32)]
    leaves = [data["i"], data["b"], data["f"], s]
    _same(cm, leaves, code, np.int32)
    code_b = code[:-1] + [_instr(capi.OP_CAST, capi.BOOL, arg=capi.F32),
                          _instr(capi.OP_BINARY, capi.BOOL, capi.BIN_LOR, capi.SRC_LEAF, 1)]
    init = np.zeros(N, np.bool_)
    init[::3] = True
    _same(cm, leaves, code_b, np.bool_, mode=capi.OR, init=init)
    _same(cm, leaves, code_b, np.bool_, mode=capi.AND, init=init)


@pytest.mark.parametrize("mode", [0, 1, 2, 3, 4])
def test_compound_modes_f32(gpu, data, mode):
    from cumat_amd import _capi as capi
    code = [_instr(capi.OP_LOAD, capi.F32, src=capi.SRC_LEAF, arg=0),
            _instr(capi.OP_BINARY, capi.F32, capi.BIN_MUL, capi.SRC_LEAF, 1),
            _instr(capi.OP_BINARY, capi.F32, capi.BIN_ADD, capi.SRC_LEAF, 2)]
    init = np.random.default_rng(mode).uniform(0.5, 2, N).astype(np.float32)
    _same(gpu, [data["f"], data["g"], data["f"]], code, np.float32, mode=mode, init=init)


@pytest.mark.parametrize("mode", [0, 1, 2, 3, 4, 5, 6, 7])
def test_compound_modes_i32(gpu, data, mode):
    from cumat_amd import _capi as capi
    code = [_instr(capi.OP_LOAD, capi.I32, src=capi.SRC_LEAF, arg=0),
            _instr(capi.OP_BINARY, capi.I32, capi.BIN_ADD, capi.SRC_IMM, 0, iimm=25)]
    init = np.random.default_rng(mode).integers(-500, 500, N).astype(np.int32)
    # j + 25 lies in 5..44: no zero divisor for /= and %=
    _same(gpu, [data["j"]], code, np.int32, mode=mode, init=init)


@pytest.mark.parametrize("offset", [1, 2, 3])
def test_unaligned_leaves_and_destination(gpu, data, offset):
    """Leaves and destination off the 16-byte grid take element loads and
    stores in both paths; the stack keeps a partial result."""
    from cumat_amd import _capi as capi
    cm = gpu
    a = _Leaf(cm, data["f"].arr, offset=offset)
    b = _Leaf(cm, data["g"].arr, offset=0)
    code = [_instr(capi.OP_LOAD, capi.F32, src=capi.SRC_LEAF, arg=0),
            _instr(capi.OP_BINARY, capi.F32, capi.BIN_ADD, capi.SRC_LEAF, 1),
            _instr(capi.OP_PUSH, capi.F32, arg=0),
            _instr(capi.OP_LOAD, capi.F32, src=capi.SRC_LEAF, arg=1),
            _instr(capi.OP_BINARY, capi.F32, capi.BIN_SUB, capi.SRC_LEAF, 0),
            _instr(capi.OP_BINARY, capi.F32, capi.BIN_MUL, capi.SRC_STACK, 0, flags=1)]
    _same(cm, [a, b], code, np.float32)
    # unaligned destination: evaluate into a view one element in
    from cumat_amd import _capi as capi2
    from cumat_amd.device import np_to_dtype
    n = N - 7
    L = (capi2.Leaf * 2)()
    for i, lf in enumerate([a, b]):
        L[i].ptr, L[i].dtype, L[i].kind = lf.ptr, capi2.F32, capi2.LEAF_STRIDED
        L[i].sr, L[i].sc, L[i].sb = 1, n, n
    prog = (capi2.Instr * len(code))(*code)
    outs = []
    for jit in (0, 2):
        dst = cm.DeviceArray.from_numpy(np.zeros(n + offset, np.float32))
        old = cm.jit_mode(jit)
        try:
            capi2.call("cm_cwise_eval", None, L, 2, prog, len(code), dst.ptr + 4 * offset, np_to_dtype(np.float32),
                       n, 1, 1, 1, n, n, 0)
        finally:
            cm.jit_mode(old)
        outs.append(dst.numpy())
    assert outs[0].tobytes() == outs[1].tobytes()


@pytest.mark.parametrize("n", [1, 5, 257, 2048, 2049, 4099])
def test_short_lengths(gpu, data, n):
    from cumat_amd import _capi as capi
    cm = gpu
    a = _Leaf(cm, data["f"].arr[:n])
    code = [_instr(capi.OP_LOAD, capi.F32, src=capi.SRC_LEAF, arg=0), _instr(capi.OP_UNARY, capi.F32, capi.UN_EXP)]
    _same(cm, [a], code, np.float32, n=n)


# ---------------------------------------------------------------- 64-bit programs
@pytest.fixture(scope="module")
def data64(gpu):
    cm = gpu
    rng = np.random.default_rng(2028)
    d = rng.standard_normal(N) * 3
    e = rng.uniform(0.1, 4, N)
    e[::89] = 0.0
    return {k: _Leaf(cm, v) for k, v in dict(d=d, e=e).items()}


@pytest.mark.parametrize("fn", list(range(19)))
def test_binary_f64(gpu, data64, fn):
    """FP64 programs run in 64-bit registers in both paths; the generated
    kernel moves two doubles per 16-byte access."""
    from cumat_amd import _capi as capi
    if fn in (capi.BIN_MOD, capi.BIN_AND, capi.BIN_OR, capi.BIN_XOR, capi.BIN_LAND, capi.BIN_LOR):
        pytest.skip("not an FP64 functor")
    cmp = fn >= capi.BIN_EQ
    code = [_instr(capi.OP_LOAD, capi.F64, src=capi.SRC_LEAF, arg=0),
            _instr(capi.OP_BINARY, capi.F64, fn, capi.SRC_LEAF, 1)]
    _same(gpu, [data64["d"], data64["e"]], code, np.bool_ if cmp else np.float64)
    code = [_instr(capi.OP_LOAD, capi.F64, src=capi.SRC_LEAF, arg=0),
            _instr(capi.OP_BINARY, capi.F64, fn, capi.SRC_IMM, 0, flags=1, fimm=1.75)]
    _same(gpu, [data64["e"]], code, np.bool_ if cmp else np.float64)


@pytest.mark.parametrize("fn", list(range(27)))
def test_unary_f64(gpu, data64, fn):
    from cumat_amd import _capi as capi
    if fn == capi.UN_BITNOT:
        pytest.skip("not an FP64 functor")
    code = [_instr(capi.OP_LOAD, capi.F64, src=capi.SRC_LEAF, arg=0), _instr(capi.OP_UNARY, capi.F64, fn)]
    _same(gpu, [data64["e"] if fn in (capi.UN_ACOS, capi.UN_ASIN) else data64["d"]], code, np.float64)


def test_mixed_width_programs(gpu, data, data64):
    """32-bit leaves widened into an FP64 program (casts, FMA, select on a
    bool leaf, a broadcast FP64 scalar) and FP64 programs stored to FP32,
    int32 and bool destinations."""
    from cumat_amd import _capi as capi
    cm = gpu
    s = _Leaf(cm, np.array([0.37]), scalar=True)
    code = [_instr(capi.OP_LOAD, capi.F64, src=capi.SRC_LEAF, arg=0),       # double(f32 f)
            _instr(capi.OP_BINARY, capi.F64, capi.BIN_MUL, capi.SRC_LEAF, 3),  # * scalar
            _instr(capi.OP_UNARY, capi.F64, capi.UN_EXP),
            _instr(capi.OP_FMA, capi.F64, arg=2, fimm=0.5),                    # + d * 0.5
            _instr(capi.OP_PUSH, capi.F64, arg=1),
            _instr(capi.OP_LOAD, capi.BOOL, src=capi.SRC_LEAF, arg=1),
            _instr(capi.OP_SELECT, capi.F64, src=capi.SRC_IMM, arg=1, fimm=2.5)]
    leaves = [data["f"], data["b"], data64["d"], s]
    _same(cm, leaves, code, np.float64)
    _same(cm, leaves, code + [_instr(capi.OP_CAST, capi.F32, arg=capi.F64)], np.float32, mode=capi.ADD,
          init=np.ones(N, np.float32))
    _same(cm, leaves, code + [_instr(capi.OP_CAST, capi.I32, arg=capi.F64)], np.int32)
    code_i = [_instr(capi.OP_LOAD, capi.F64, src=capi.SRC_LEAF, arg=0),        # double(int32 i)
              _instr(capi.OP_BINARY, capi.F64, capi.BIN_DIV, capi.SRC_LEAF, 1),
              _instr(capi.OP_BINARY, capi.F64, capi.BIN_GT, capi.SRC_IMM, 0, fimm=-3.0)]
    _same(cm, [data["i"], data64["e"]], code_i, np.bool_)


@pytest.mark.parametrize("mode", [0, 1, 2, 3, 4])
def test_compound_modes_f64(gpu, data64, mode):
    from cumat_amd import _capi as capi
    code = [_instr(capi.OP_LOAD, capi.F64, src=capi.SRC_LEAF, arg=0),
            _instr(capi.OP_BINARY, capi.F64, capi.BIN_MUL, capi.SRC_LEAF, 1),
            _instr(capi.OP_BINARY, capi.F64, capi.BIN_ADD, capi.SRC_LEAF, 0)]
    init = np.random.default_rng(mode).uniform(0.5, 2, N)
    _same(gpu, [data64["d"], data64["e"]], code, np.float64, mode=mode, init=init)


@pytest.mark.parametrize("offset", [1, 3])
def test_unaligned_f64_and_short_lengths(gpu, data64, offset):
    from cumat_amd import _capi as capi
    from cumat_amd.device import np_to_dtype
    cm = gpu
    a = _Leaf(cm, data64["d"].arr, offset=offset)
    code = [_instr(capi.OP_LOAD, capi.F64, src=capi.SRC_LEAF, arg=0),
            _instr(capi.OP_BINARY, capi.F64, capi.BIN_SUB, capi.SRC_LEAF, 1),
            _instr(capi.OP_UNARY, capi.F64, capi.UN_TANH)]
    _same(cm, [a, data64["e"]], code, np.float64)
    for n in (1, 5, 2049):
        _same(cm, [_Leaf(cm, data64["d"].arr[:n])], code[:1] + code[2:], np.float64, n=n)
    # a destination off the 16-byte grid
    n = N - 5
    L = (capi.Leaf * 1)()
    L[0].ptr, L[0].dtype, L[0].kind = data64["d"].ptr, capi.F64, capi.LEAF_STRIDED
    L[0].sr, L[0].sc, L[0].sb = 1, n, n
    prog = (capi.Instr * 2)(code[0], code[2])
    outs = []
    for jit in (0, 2):
        dst = cm.DeviceArray.from_numpy(np.zeros(n + offset))
        old = cm.jit_mode(jit)
        try:
            capi.call("cm_cwise_eval", None, L, 1, prog, 2, dst.ptr + 8 * offset, np_to_dtype(np.float64),
                      n, 1, 1, 1, n, n, 0)
        finally:
            cm.jit_mode(old)
        outs.append(dst.numpy())
    assert outs[0].tobytes() == outs[1].tobytes()


def test_auto_mode_threshold_and_cache(gpu, data):
    """Mode 1 compiles a program once and reuses it; small evaluations stay
    on the interpreter."""
    from cumat_amd import _capi as capi
    cm = gpu
    code = [_instr(capi.OP_LOAD, capi.F32, src=capi.SRC_LEAF, arg=0),
            _instr(capi.OP_BINARY, capi.F32, capi.BIN_MAX, capi.SRC_IMM, 0, fimm=0.125)]
    init = np.zeros(N, np.float32)
    cm.reset_counters()
    r1 = _eval(cm, [data["f"]], code, np.float32, 0, init, jit=1)
    r2 = _eval(cm, [data["f"]], code, np.float32, 0, init, jit=1)
    c = cm.counters()
    assert c["JitLaunch"] == 2 and c["JitCompile"] <= 1
    assert r1.tobytes() == r2.tobytes()
    small = np.zeros(1000, np.float32)
    a = _Leaf(cm, data["f"].arr[:1000])
    cm.reset_counters()
    _eval(cm, [a], code, np.float32, 0, small, jit=1)
    assert cm.counters()["JitLaunch"] == 0
    assert np.array_equal(r1, np.maximum(data["f"].arr, np.float32(0.125)))


def _long_row_pattern(g, nlong, width, rng):
    """2D 5-point pattern plus 'nlong' rows with 'width' extra entries (their
    chunks exceed the lane budget and take the per-lane row walk)."""
    from cumat_amd.problems import poisson2d
    JA, IA, _ = poisson2d(g, 0.0, np.float64)
    n = g * g
    long_rows = set(rng.choice(n, nlong, replace=False).tolist())
    rows = []
    for r in range(n):
        cols = list(IA[JA[r]:JA[r + 1]])
        if r in long_rows:
            cols = sorted(set(cols) | set(int(c) for c in rng.choice(n, width, replace=False)))
        rows.append(cols)
    JA2 = np.zeros(n + 1, np.int32)
    JA2[1:] = np.cumsum([len(c) for c in rows])
    return JA2, np.concatenate([np.array(c, np.int32) for c in rows]), n


@pytest.mark.parametrize("dtype", [np.float32, np.int32])
@pytest.mark.parametrize("nlong", [0, 30])
def test_csr_expression_product_generated(gpu, dtype, nlong):
    """(a*A + B - s).sparseView(pattern) * x with batched same-pattern value
    arrays, a broadcast scalar and a bool value array in a select; modes =,
    += and -=: the generated product equals the interpreter's bit for bit
    (random FP32 data, so rounding order matters)."""
    from cumat_amd import _capi as capi
    cm = gpu
    rng = np.random.default_rng(31 + nlong)
    JA, IA, n = _long_row_pattern(180, nlong, 900, rng)
    nnz, B = IA.size, 2
    if dtype == np.float32:
        va = rng.standard_normal(nnz * B).astype(dtype)
        vb = rng.standard_normal(nnz * B).astype(dtype)
        x = rng.standard_normal(n * B).astype(dtype)
        three, one = 3.0, 1.0
    else:
        va = rng.integers(-9, 10, nnz * B).astype(dtype)
        vb = rng.integers(-9, 10, nnz * B).astype(dtype)
        x = rng.integers(-9, 10, n * B).astype(dtype)
        three, one = 3, 1
    pat = cm.SparsityPattern(n, n, JA, IA)
    A = cm.SparseMatrix(pat, va, batches=B)
    Bm = cm.SparseMatrix(pat, vb, batches=B)
    mask = cm.SparseMatrix(pat, rng.integers(0, 2, nnz).astype(np.bool_))
    xd = cm.DeviceArray.from_numpy(x, rows=n, cols=1, batches=B)
    e = cm.select(mask.direct(), A.direct() * three + Bm.direct() - one, A.direct())
    y0 = (rng.standard_normal(n * B) * 10).astype(dtype)
    outs = {}
    for jit in (0, 2):
        old = cm.jit_mode(jit)
        try:
            cm.reset_counters()
            y = cm.sparse_view(e, pat).matvec(xd)
            assert cm.counters()["JitLaunch"] == (1 if jit else 0)
            res = [y.numpy()]
            for mode in (capi.ADD, capi.SUB):
                yd = cm.DeviceArray.from_numpy(y0.copy(), rows=n, cols=1, batches=B)
                cm.sparse_view(e, pat).matvec(xd, out=yd, mode=mode)
                res.append(yd.numpy())
        finally:
            cm.jit_mode(old)
        outs[jit] = res
    for a, b in zip(outs[0], outs[2]):
        assert a.tobytes() == b.tobytes()


@pytest.mark.parametrize("dtype", [np.float32, np.int32])
def test_csr_cwise_column_leaves_generated(gpu, dtype):
    """A = A * r(col) + s(col) (row vectors indexed by each stored entry's
    column) and a compound mode, by the generated kernel (one launch, the
    column leaves gathered through IA) and by the interpreter route (gather
    pass + dense program): the same bits."""
    from cumat_amd import _capi as capi
    cm = gpu
    rng = np.random.default_rng(77)
    JA, IA, n = _long_row_pattern(400, 0, 0, rng)
    nnz = IA.size
    if dtype == np.float32:
        v0 = rng.standard_normal(nnz).astype(dtype)
        r = rng.standard_normal(n).astype(dtype)
    else:
        v0 = rng.integers(-50, 50, nnz).astype(dtype)
        r = rng.integers(-7, 8, n).astype(dtype)
    flag = rng.integers(0, 2, n).astype(np.bool_)
    pat = cm.SparsityPattern(n, n, JA, IA)
    rd = cm.DeviceArray.from_numpy(r, rows=1, cols=n, batches=1, row_major=True)
    fd = cm.DeviceArray.from_numpy(flag, rows=1, cols=n, batches=1, row_major=True)
    outs = {}
    for jit in (0, 2):
        S = cm.SparseMatrix(pat, v0.copy())
        old = cm.jit_mode(jit)
        try:
            cm.reset_counters()
            S.assign(cm.select(fd.expr(), S.direct() * rd.expr(), S.direct()))
            c = cm.counters()
            assert c["JitLaunch"] == (1 if jit else 0)
            if jit:
                assert c["KernelLaunch"] == 1  # no gather pass
            S.assign(S.direct() + rd.expr(), mode=capi.ADD)
            outs[jit] = S.values.numpy()
        finally:
            cm.jit_mode(old)
    assert outs[0].tobytes() == outs[2].tobytes()
    rows_c = IA
    exp = np.where(flag[rows_c], v0 * r[rows_c], v0).astype(dtype)
    exp = (exp + (exp + r[rows_c])).astype(dtype)
    if dtype == np.int32:
        assert np.array_equal(outs[2], exp)


def test_csr_cwise_rejects_null_pointers_and_bad_programs_before_launch(gpu):
    from cumat_amd import _capi as capi
    cm = gpu
    from cumat_amd.problems import poisson2d
    JA, IA, V = poisson2d(16)
    pat = cm.SparsityPattern(256, 256, JA, IA)
    S = cm.SparseMatrix(pat, V)
    r = cm.DeviceArray.from_numpy(np.ones(256, np.float32), rows=1, cols=256, batches=1, row_major=True)
    L = (capi.Leaf * 2)()
    L[0].ptr, L[0].dtype, L[0].kind, L[0].sr, L[0].sc, L[0].sb = S.values.ptr, capi.F32, 0, -1, 0, 0
    L[1].ptr, L[1].dtype, L[1].kind, L[1].sr, L[1].sc, L[1].sb = r.ptr, capi.F32, 0, 0, 1, 0
    good = (capi.Instr * 2)(_instr(capi.OP_LOAD, capi.F32, src=capi.SRC_LEAF, arg=0),
                            _instr(capi.OP_BINARY, capi.F32, capi.BIN_MUL, capi.SRC_LEAF, 1))
    bad = (capi.Instr * 2)(_instr(capi.OP_LOAD, capi.F32, src=capi.SRC_LEAF, arg=0), _instr(77, capi.F32))
    Lb = capi.lib()
    cm.reset_counters()
    rc = Lb.cm_csr_cwise_eval(None, L, 2, good, 2, 256, 256, IA.size, 1, pat.JA.ptr, None, S.values.ptr, capi.F32, 0)
    assert rc == capi.CM_E_ARG
    rc = Lb.cm_csr_cwise_eval(None, L, 2, bad, 2, 256, 256, IA.size, 1, pat.JA.ptr, pat.IA.ptr, S.values.ptr,
                              capi.F32, 0)
    assert rc == capi.CM_E_ARG
    assert cm.counters()["KernelLaunch"] == 0


def _rx(cm, leaves, code, dtype, op, axis, shape, out_np, jit):
    """cm_reduce_expr of a raw program over a rows x cols x batches shape;
    leaves are (array, (sr, sc, sb)) with element strides."""
    from cumat_amd import _capi as capi
    from cumat_amd.device import np_to_dtype
    R, Cc, B = shape
    L = (capi.Leaf * max(1, len(leaves)))()
    for i, (lf, st) in enumerate(leaves):
        L[i].ptr, L[i].dtype, L[i].kind = lf.ptr, np_to_dtype(lf.arr.dtype), capi.LEAF_STRIDED
        L[i].sr, L[i].sc, L[i].sb = st
    prog = (capi.Instr * len(code))(*code)
    oR = 1 if axis & 1 else R
    oC = 1 if axis & 2 else Cc
    oB = 1 if axis & 4 else B
    out = cm.DeviceArray(oR * oC * oB, dtype=out_np)
    old = cm.jit_mode(jit)
    try:
        capi.call("cm_reduce_expr", None, L, len(leaves), prog, len(code), dtype, op, axis, R, Cc, B, 0, out.ptr)
        return out.numpy()
    finally:
        cm.jit_mode(old)


@pytest.mark.parametrize("op", ["sum", "prod", "min", "max"])
@pytest.mark.parametrize("axis", [7, 3])
def test_reduce_expression_generated_f32(gpu, data, op, axis):
    """(a*b + c) reduced over all elements or per batch (the virtual order is
    the memory order: the generated reduction) -- the same bits as the
    interpreter; one launch of the generated kernel (plus the combine)."""
    from cumat_amd import _capi as capi
    cm = gpu
    ops = {"sum": capi.RED_SUM, "prod": capi.RED_PROD, "min": capi.RED_MIN, "max": capi.RED_MAX}
    B = 4
    R = (N - 3) // B
    shape = (R, 1, B)
    st = (1, R, R)
    g = data["g"] if op != "prod" else _Leaf(gpu, (1.0 + data["f"].arr * 1e-4).astype(np.float32))
    code = [_instr(capi.OP_LOAD, capi.F32, src=capi.SRC_LEAF, arg=0),
            _instr(capi.OP_BINARY, capi.F32, capi.BIN_MUL, capi.SRC_LEAF, 1),
            _instr(capi.OP_BINARY, capi.F32, capi.BIN_ADD, capi.SRC_LEAF, 2)]
    leaves = [(data["f"], st), (g, st), (data["g"], st)]
    cm.reset_counters()
    ref = _rx(cm, leaves, code, capi.F32, ops[op], axis, shape, np.float32, jit=0)
    assert cm.counters()["JitLaunch"] == 0
    got = _rx(cm, leaves, code, capi.F32, ops[op], axis, shape, np.float32, jit=2)
    assert cm.counters()["JitLaunch"] == 1, "the generated reduction did not run"
    assert ref.tobytes() == got.tobytes()


@pytest.mark.parametrize("op", ["sum", "min", "max", "all", "any", "band", "bor"])
def test_reduce_expression_generated_int_and_bool(gpu, data, op):
    """int32 (i % j + scalar) and bool (i > j) reductions, an unaligned leaf
    and a broadcast scalar leaf."""
    from cumat_amd import _capi as capi
    cm = gpu
    ops = {"sum": capi.RED_SUM, "min": capi.RED_MIN, "max": capi.RED_MAX, "all": capi.RED_ALL,
           "any": capi.RED_ANY, "band": capi.RED_BAND, "bor": capi.RED_BOR}
    n = N - 5
    shape = (n, 1, 1)
    st = (1, n, n)
    iu = _Leaf(cm, data["i"].arr[:n], offset=1)
    sc = _Leaf(cm, np.array([7], np.int32), scalar=True)
    code = [_instr(capi.OP_LOAD, capi.I32, src=capi.SRC_LEAF, arg=0),
            _instr(capi.OP_BINARY, capi.I32, capi.BIN_MOD, capi.SRC_LEAF, 1),
            _instr(capi.OP_BINARY, capi.I32, capi.BIN_ADD, capi.SRC_LEAF, 2)]
    leaves = [(iu, st), (data["j"], st), (sc, (0, 0, 0))]
    ref = _rx(cm, leaves, code, capi.I32, ops[op], 7, shape, np.int32, jit=0)
    cm.reset_counters()
    got = _rx(cm, leaves, code, capi.I32, ops[op], 7, shape, np.int32, jit=2)
    assert cm.counters()["JitLaunch"] == 1
    assert ref.tobytes() == got.tobytes()
    if op in ("sum", "all", "any"):
        codeb = [_instr(capi.OP_LOAD, capi.I32, src=capi.SRC_LEAF, arg=0),
                 _instr(capi.OP_BINARY, capi.I32, capi.BIN_GT, capi.SRC_LEAF, 1)]
        refb = _rx(cm, leaves[:2], codeb, capi.BOOL, ops[op], 7, shape, np.bool_, jit=0)
        gotb = _rx(cm, leaves[:2], codeb, capi.BOOL, ops[op], 7, shape, np.bool_, jit=2)
        assert refb.tobytes() == gotb.tobytes()
    if op == "sum":
        i, j = data["i"].arr[:n].astype(np.int64), data["j"].arr[:n].astype(np.int64)
        m = np.where(j != 0, np.fmod(i, np.where(j == 0, 1, j)), 0) + 7
        exp = np.array([int(m.sum()) & 0xffffffff], np.uint32).view(np.int32)[0]  # int32 sums wrap
        assert int(got[0]) == int(exp)


def test_reduce_expression_non_flat_generated(gpu, data):
    """A row-major leaf reduced along rows walks a virtual order that is not
    its memory order: the generated reduction reads it at the element's
    coordinates (shape and strides compiled in), with the interpreter's
    elements per thread and fold order."""
    from cumat_amd import _capi as capi
    cm = gpu
    R, Cc = 1024, 1000
    code = [_instr(capi.OP_LOAD, capi.F32, src=capi.SRC_LEAF, arg=0), _instr(capi.OP_UNARY, capi.F32, capi.UN_ABS)]
    leaves = [(data["f"], (Cc, 1, R * Cc))]
    ref = _rx(cm, leaves, code, capi.F32, capi.RED_SUM, 1, (R, Cc, 1), np.float32, jit=0)
    cm.reset_counters()
    got = _rx(cm, leaves, code, capi.F32, capi.RED_SUM, 1, (R, Cc, 1), np.float32, jit=2)
    assert cm.counters()["JitLaunch"] == 1
    assert ref.tobytes() == got.tobytes()
    a = np.abs(data["f"].arr[:R * Cc].reshape(R, Cc)).astype(np.float64).sum(axis=0)
    np.testing.assert_allclose(got, a, rtol=1e-5)


@pytest.mark.parametrize("axis", [1, 2, 3, 4, 5, 6, 7])
@pytest.mark.parametrize("op", ["sum", "max"])
def test_reduce_broadcast_expressions_generated(gpu, data, axis, op):
    """(A * colvec - rowvec) over every axis subset of a 300 x 257 x 7
    tensor, and an int32 program with a batch-broadcast leaf: generated and
    interpreted reductions give the same bits."""
    from cumat_amd import _capi as capi
    cm = gpu
    ops = {"sum": capi.RED_SUM, "max": capi.RED_MAX}
    R, Cc, B = 300, 257, 7
    st = (1, R, R * Cc)
    code = [_instr(capi.OP_LOAD, capi.F32, src=capi.SRC_LEAF, arg=0),
            _instr(capi.OP_BINARY, capi.F32, capi.BIN_MUL, capi.SRC_LEAF, 1),
            _instr(capi.OP_BINARY, capi.F32, capi.BIN_SUB, capi.SRC_LEAF, 2)]
    leaves = [(data["f"], st), (data["g"], (1, 0, 0)), (data["g"], (0, 1, 0))]
    ref = _rx(cm, leaves, code, capi.F32, ops[op], axis, (R, Cc, B), np.float32, jit=0)
    cm.reset_counters()
    got = _rx(cm, leaves, code, capi.F32, ops[op], axis, (R, Cc, B), np.float32, jit=2)
    assert cm.counters()["JitLaunch"] == 1
    assert ref.tobytes() == got.tobytes()
    codei = [_instr(capi.OP_LOAD, capi.I32, src=capi.SRC_LEAF, arg=0),
             _instr(capi.OP_BINARY, capi.I32, capi.BIN_ADD, capi.SRC_LEAF, 1)]
    leavesi = [(data["i"], (Cc, 1, R * Cc)), (data["j"], (1, R, 0))]
    refi = _rx(cm, leavesi, codei, capi.I32, ops[op], axis, (R, Cc, B), np.int32, jit=0)
    goti = _rx(cm, leavesi, codei, capi.I32, ops[op], axis, (R, Cc, B), np.int32, jit=2)
    assert refi.tobytes() == goti.tobytes()
    if op == "sum":
        i3 = data["i"].arr[:R * Cc * B].astype(np.int64).reshape(B, R, Cc)
        j2 = data["j"].arr[:R * Cc].astype(np.int64).reshape(Cc, R).T
        tot = i3 + j2[None]
        red = tuple(a for a, bit in ((1, 1), (2, 2), (0, 4)) if axis & bit)
        exp = tot.sum(axis=red)
        assert np.array_equal(np.sort(goti.astype(np.int64)), np.sort((exp.reshape(-1) + 2**31) % 2**32 - 2**31))


# ------------------------------------------- programs with strided leaves
def _st_same(cm, dst_factory, e):
    """dst = e by the interpreter and by the generated kernel (coordinates
    compiled in), bit for bit; returns the logical (batches, rows, cols)
    result."""
    outs = []
    for jit in (0, 2):
        dst = dst_factory()
        old = cm.jit_mode(jit)
        try:
            cm.reset_counters()
            cm.assign(dst, e)
            launched = cm.counters()["JitLaunch"]
        finally:
            cm.jit_mode(old)
        assert launched == (1 if jit else 0), (jit, launched)
        outs.append(dst.numpy(logical=True))
    assert outs[0].tobytes() == outs[1].tobytes()
    return outs[1]


@pytest.mark.parametrize("shape", [(512, 300, 3), (3, 2000, 5), (1000, 3, 7), (1, 5000, 2), (777, 1, 9)])
def test_strided_broadcast_programs(gpu, shape):
    """Column, row and batch broadcasts, a row-major leaf in a column-major
    destination and a row-major destination, odd leading dimensions: the
    generated kernel reads each leaf at the element's coordinates."""
    cm = gpu
    R, C, B = shape
    rng = np.random.default_rng(R * 7 + C)
    a = rng.standard_normal((B, R, C)).astype(np.float32)
    ar = rng.standard_normal((B, R, C)).astype(np.float32)
    col = rng.standard_normal((1, R, 1)).astype(np.float32)
    row = rng.standard_normal((1, 1, C)).astype(np.float32)
    mat = rng.standard_normal((1, R, C)).astype(np.float32)
    A = cm.DeviceArray.from_numpy(a)
    Ar = cm.DeviceArray.from_numpy(ar, row_major=True)
    V = cm.DeviceArray.from_numpy(col)
    W = cm.DeviceArray.from_numpy(row)
    M = cm.DeviceArray.from_numpy(mat)
    colmajor = lambda: cm.DeviceArray(R, C, B, dtype=np.float32)
    rowmajor = lambda: cm.DeviceArray(R, C, B, dtype=np.float32, row_major=True)
    got = _st_same(cm, colmajor, A.expr() * V.expr())
    assert np.array_equal(got, a * col)
    got = _st_same(cm, colmajor, A.expr() * W.expr() + V.expr())
    assert np.array_equal(got, a * row + col)
    got = _st_same(cm, colmajor, A.expr() - M.expr())
    assert np.array_equal(got, a - mat)
    got = _st_same(cm, colmajor, A.expr() + Ar.expr())
    assert np.array_equal(got, a + ar)
    got = _st_same(cm, rowmajor, A.expr() * V.expr() + Ar.expr())
    assert np.array_equal(got, a * col + ar)


def test_strided_programs_f64_and_int(gpu):
    cm = gpu
    R, C, B = 300, 257, 4
    rng = np.random.default_rng(9)
    a = rng.standard_normal((B, R, C))
    row = rng.standard_normal((1, 1, C))
    A = cm.DeviceArray.from_numpy(a)
    W = cm.DeviceArray.from_numpy(row)
    got = _st_same(cm, lambda: cm.DeviceArray(R, C, B, dtype=np.float64), A.expr() * W.expr() - 0.5)
    assert np.array_equal(got, a * row - 0.5)
    ai = rng.integers(-1000, 1000, (B, R, C)).astype(np.int32)
    ci = rng.integers(1, 50, (1, R, 1)).astype(np.int32)
    Ai = cm.DeviceArray.from_numpy(ai, row_major=True)
    Ci = cm.DeviceArray.from_numpy(ci)
    got = _st_same(cm, lambda: cm.DeviceArray(R, C, B, dtype=np.int32), Ai.expr() % Ci.expr())
    assert np.array_equal(got, np.fmod(ai, ci))


@pytest.mark.parametrize("op", ["sum", "prod", "min", "max"])
@pytest.mark.parametrize("axis", [7, 3, 1])
def test_reduce_expression_generated_f64(gpu, data64, op, axis):
    """64-bit programs (4 consecutive elements per thread in 64-bit
    registers): FP64 a*b + c over every element, per batch, per column, with
    a broadcast column vector -- the same bits as the interpreter."""
    from cumat_amd import _capi as capi
    cm = gpu
    ops = {"sum": capi.RED_SUM, "prod": capi.RED_PROD, "min": capi.RED_MIN, "max": capi.RED_MAX}
    R, Cc, B = 1000, 250, 4
    st = (1, R, R * Cc)
    e = data64["e"] if op != "prod" else _Leaf(gpu, 1.0 + data64["d"].arr * 1e-6)
    code = [_instr(capi.OP_LOAD, capi.F64, src=capi.SRC_LEAF, arg=0),
            _instr(capi.OP_BINARY, capi.F64, capi.BIN_MUL, capi.SRC_LEAF, 1),
            _instr(capi.OP_BINARY, capi.F64, capi.BIN_ADD, capi.SRC_LEAF, 2)]
    for leaves in ([(data64["d"], st), (e, st), (data64["e"], st)],
                   [(data64["d"], st), (e, (1, 0, 0)), (data64["e"], (0, 1, 0))]):
        ref = _rx(cm, leaves, code, capi.F64, ops[op], axis, (R, Cc, B), np.float64, jit=0)
        cm.reset_counters()
        got = _rx(cm, leaves, code, capi.F64, ops[op], axis, (R, Cc, B), np.float64, jit=2)
        assert cm.counters()["JitLaunch"] == 1
        assert ref.tobytes() == got.tobytes()


def test_reduce_expression_generated_mixed_width(gpu, data, data64):
    """An FP32 result of a program with an FP64 leaf (64-bit registers, FP32
    accumulation) and a bool result of an FP64 comparison."""
    from cumat_amd import _capi as capi
    cm = gpu
    R, Cc, B = 700, 300, 3
    st = (1, R, R * Cc)
    code = [_instr(capi.OP_LOAD, capi.F64, src=capi.SRC_LEAF, arg=0),
            _instr(capi.OP_BINARY, capi.F64, capi.BIN_SUB, capi.SRC_LEAF, 1),
            _instr(capi.OP_CAST, capi.F32, arg=capi.F64)]
    leaves = [(data["f"], st), (data64["d"], (Cc, 1, R * Cc))]
    for op in (capi.RED_SUM, capi.RED_MAX):
        ref = _rx(cm, leaves, code, capi.F32, op, 6, (R, Cc, B), np.float32, jit=0)
        got = _rx(cm, leaves, code, capi.F32, op, 6, (R, Cc, B), np.float32, jit=2)
        assert ref.tobytes() == got.tobytes()
    codeb = [_instr(capi.OP_LOAD, capi.F64, src=capi.SRC_LEAF, arg=0),
             _instr(capi.OP_BINARY, capi.F64, capi.BIN_LT, capi.SRC_LEAF, 1)]
    for op in (capi.RED_SUM, capi.RED_ALL, capi.RED_ANY):
        ref = _rx(cm, leaves, codeb, capi.BOOL, op, 7, (R, Cc, B), np.bool_, jit=0)
        cm.reset_counters()
        got = _rx(cm, leaves, codeb, capi.BOOL, op, 7, (R, Cc, B), np.bool_, jit=2)
        assert cm.counters()["JitLaunch"] == 1
        assert ref.tobytes() == got.tobytes()
