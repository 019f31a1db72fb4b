"""GPU test of cm_dot_noalloc, the dot product that allocates nothing (its
partial sums live in a fixed device buffer of the library): through the C
ABI, against cm_dot on the same inputs, two calls in a row with the second
one smaller, every dtype and the batched / broadcast forms. The results must
be bit-identical (same kernels, same summation order) and the caching
allocator must not be touched by the no-alloc calls."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dev(cm, a):
    return cm.DeviceArray.from_numpy(np.ascontiguousarray(a))


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.int32])
def test_dot_noalloc_matches_dot_two_calls(gpu, dtype):
    cm = gpu
    L = cm.capi.lib()
    ctx = cm.Context(0)
    dt = {np.float32: cm.capi.F32, np.float64: cm.capi.F64, np.int32: cm.capi.I32}[dtype]
    rng = np.random.default_rng(42)
    # first call: 3 batches of 2^22 + 4 (split into many chunks, 16-byte
    # path for FP32); second call: smaller, 2 batches of 100003 with a
    # broadcast second operand (scalar path)
    cases = [(3, (1 << 22) + 4, 1, 1), (2, 100003, 1, 0)]
    for batches, n, ab, bb in cases:
        if dtype == np.int32:
            a = rng.integers(-100, 101, n * (batches if ab else 1)).astype(dtype)
            b = rng.integers(-100, 101, n * (batches if bb else 1)).astype(dtype)
        else:
            a = rng.uniform(-1, 1, n * (batches if ab else 1)).astype(dtype)
            b = rng.uniform(-1, 1, n * (batches if bb else 1)).astype(dtype)
        da, db = _dev(cm, a), _dev(cm, b)
        o1 = cm.DeviceArray(batches, dtype=dtype, ctx=ctx)
        o2 = cm.DeviceArray(batches, dtype=dtype, ctx=ctx)
        cm.capi.call("cm_dot", ctx.handle, dt, n, batches, da.ptr, ab, db.ptr, bb, o1.ptr)
        ctx.sync()
        before = cm.counters()["DeviceMemAlloc"]
        cm.capi.check(L.cm_dot_noalloc(C.c_void_p(ctx.handle), dt, n, batches, C.c_void_p(da.ptr), ab,
                                       C.c_void_p(db.ptr), bb, C.c_void_p(o2.ptr)))
        ctx.sync()
        assert cm.counters()["DeviceMemAlloc"] == before
        r1, r2 = o1.numpy(), o2.numpy()
        assert np.array_equal(r1, r2), (n, batches, r1, r2)
        # and both are right (FP64 sums of the same inputs)
        for k in range(batches):
            ak = a[k * n:(k + 1) * n] if ab else a
            bk = b[k * n:(k + 1) * n] if bb else b
            if dtype == np.int32:
                exact = np.dot(ak.astype(np.int64), bk.astype(np.int64))
                assert r2[k] == np.array([exact], np.int64).astype(np.int32)[0]  # int32 sums wrap
            else:
                ref = np.dot(ak.astype(np.float64), bk.astype(np.float64))
                mag = np.dot(np.abs(ak.astype(np.float64)), np.abs(bk.astype(np.float64)))
                assert abs(float(r2[k]) - ref) <= (1e-5 if dtype == np.float32 else 1e-13) * mag


def test_dot_noalloc_rejects_null_context_and_pointers(gpu):
    cm = gpu
    L = cm.capi.lib()
    out = cm.DeviceArray(1, dtype=np.float32)
    a = _dev(cm, np.ones(16, np.float32))
    st = L.cm_dot_noalloc(None, cm.capi.F32, 16, 1, C.c_void_p(a.ptr), 1, C.c_void_p(a.ptr), 1, C.c_void_p(out.ptr))
    assert st == cm.capi.CM_E_ARG and "context" in cm.capi.last_error()
    ctx = cm.Context(0)
    st = L.cm_dot_noalloc(C.c_void_p(ctx.handle), cm.capi.F32, 16, 1, None, 1, C.c_void_p(a.ptr), 1,
                          C.c_void_p(out.ptr))
    assert st == cm.capi.CM_E_ARG
