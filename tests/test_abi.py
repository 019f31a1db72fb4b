"""CPU checks of the drop-in boundary: the HIP library loads without a GPU and
exports every entry point include/cumat_amd.h declares; host-side argument
checking; the problem generators."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "cumat_amd.h")).read()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(cm_[a-z0-9_]+)\s*\(", src, re.M)))


def test_library_exports_every_declared_symbol():
    from cumat_amd import _capi
    L = _capi.lib()
    names = _declared()
    assert len(names) >= 30
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_capi.exported_symbols())
    assert L.cm_abi_version() == 1


def test_error_path_without_gpu():
    """Calls that fail return a status and set the thread-local message
    (no exception crosses the ABI)."""
    from cumat_amd import _capi
    L = _capi.lib()
    v = C.c_longlong()
    assert L.cm_counters_get(999, C.byref(v)) == _capi.CM_E_ARG
    assert "out of range" in _capi.last_error()
    assert L.cm_counters_reset() == 0
    n = C.c_int(-1)
    assert L.cm_device_count(C.byref(n)) == 0 and n.value >= 0


def test_expression_lowering_host_only():
    """The accumulator program for a lincomb uses one LOAD/MUL then FMA terms
    (no device memory needed: leaves are fake device arrays)."""
    from cumat_amd import _capi, expr

    class Fake:
        def __init__(self, p):
            self.ptr, self.dtype, self.rows, self.cols, self.batches = p, _capi.F32, 10, 1, 1

        def strides(self):
            return (1, 10, 10)
    leaves = [expr.LeafNode(Fake(1000 + 16 * i)) for i in range(8)]
    e = leaves[0] * 0.5
    for i in range(1, 8):
        e = e + leaves[i] * float(i)
    L, nl, code, nc = expr.lower(e, (10, 1, 1))
    ops = [code[i].opcode for i in range(nc)]
    assert nl == 8 and ops == [_capi.OP_LOAD, _capi.OP_BINARY] + [_capi.OP_FMA] * 7
    # a tree needing the stack: (a+b)*(c+d)
    e2 = (leaves[0] + leaves[1]) * (leaves[2] + leaves[3])
    L, nl, code, nc = expr.lower(e2, (10, 1, 1))
    assert [code[i].opcode for i in range(nc)] == [_capi.OP_LOAD, _capi.OP_BINARY, _capi.OP_PUSH,
                                                   _capi.OP_LOAD, _capi.OP_BINARY, _capi.OP_BINARY]
    assert code[nc - 1].src == _capi.SRC_STACK and code[nc - 1].flags == 1


def test_grid_laplacian_structure():
    from cumat_amd.problems import grid_laplacian, lognormal_sampler
    rng = np.random.default_rng(1)
    for dims in [(3, 4), (4, 3, 5)]:
        P = grid_laplacian(dims, lognormal_sampler(rng))
        n = P["n"]
        A = np.zeros((n, n))
        rows = np.repeat(np.arange(n), np.diff(P["JA"]))
        A[rows, P["IA"]] = P["vals"]
        assert np.allclose(A, A.T)
        assert np.all(np.linalg.eigvalsh(A) > 0)
        for r in range(n):  # ascending columns per row
            cols = P["IA"][P["JA"][r]:P["JA"][r + 1]]
            assert np.all(np.diff(cols) > 0)
        # the gather map reproduces the values from theta
        th = P["theta"]
        v = np.array([sum(np.sign(c) * th[abs(c) - 1] for c in P["eidx"][P["eptr"][k]:P["eptr"][k + 1]])
                      for k in range(P["nnz"])])
        assert np.allclose(v, P["vals"])
        d = len(dims)
        assert P["nedges"] == sum(np.prod(dims) // dims[a] * (dims[a] + 1) for a in range(d))


def test_grid_laplacian_160_counts():
    """The 3D 160^3 mixed-precision workload: n = 4,096,000, nnz = 28,518,400
    (checked on the formula, the full build is done on the GPU side)."""
    g = 160
    n = g ** 3
    interior = 3 * g * g * (g - 1)
    assert n == 4096000 and n + 2 * interior == 28518400
    assert interior == 12211200 and 6 * g * g == 153600


def test_batched_poisson_structure():
    from cumat_amd.problems import batched_poisson
    P = batched_poisson(20, 16, 24, seed=3)
    ro = P["row_off"]
    assert ro[0] == 0 and ro[-1] == P["b"].size
    for s in range(20):
        lo, hi = P["JA"][ro[s]], P["JA"][ro[s + 1]]
        cols = P["IA"][lo:hi]
        assert cols.min() >= ro[s] and cols.max() < ro[s + 1]
