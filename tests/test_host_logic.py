"""CPU tests of the host layers (no device call is made): lowering of
expression trees to the cm_cwise_eval program, broadcasting strides, the
sparsity-pattern checks that run before any upload, and the synthetic
problem generators (symmetry, row sums, the edge gather map)."""
import numpy as np
import pytest

from cumat_amd import _capi as capi
from cumat_amd.device import DeviceArray
from cumat_amd.expr import lower, cast


def _arr(rows, cols=1, batches=1, dtype=np.float32, row_major=False, ptr=0x1000):
    # a non-owning descriptor with a fake address: never handed to the library
    return DeviceArray(rows, cols, batches, dtype=dtype, row_major=row_major, ptr=ptr)


def _ops(code, n):
    return [(code[i].opcode, code[i].src, code[i].arg) for i in range(n)]


def test_lowering_linear_combination_uses_fma_terms():
    a, b, c = _arr(100, ptr=0x1000), _arr(100, ptr=0x2000), _arr(100, ptr=0x3000)
    e = a.expr() * 2.0 + b.expr() * 3.0 + c.expr() * -1.5
    leaves, nl, code, nc = lower(e, (100, 1, 1))
    assert nl == 3
    assert _ops(code, nc) == [(capi.OP_LOAD, capi.SRC_LEAF, 0), (capi.OP_BINARY, capi.SRC_IMM, 0),
                              (capi.OP_FMA, 0, 1), (capi.OP_FMA, 0, 2)]
    assert code[1].fn == capi.BIN_MUL and code[1].fimm == 2.0
    assert (code[2].fimm, code[3].fimm) == (3.0, -1.5)
    assert [leaves[i].ptr for i in range(3)] == [0x1000, 0x2000, 0x3000]


def test_lowering_deduplicates_leaves_and_uses_stack_for_two_subtrees():
    a, b = _arr(10, 10, ptr=0x1000), _arr(10, 10, ptr=0x2000)
    e = (a.expr() + b.expr()) * (a.expr() - b.expr())
    leaves, nl, code, nc = lower(e, (10, 10, 1))
    assert nl == 2  # a and b once each
    ops = [code[i].opcode for i in range(nc)]
    assert ops == [capi.OP_LOAD, capi.OP_BINARY, capi.OP_PUSH, capi.OP_LOAD, capi.OP_BINARY, capi.OP_BINARY]
    last = code[nc - 1]
    assert last.fn == capi.BIN_MUL and last.src == capi.SRC_STACK and last.flags == 1


def test_lowering_broadcast_strides_are_zero():
    m = _arr(4, 5, 3, ptr=0x1000)               # 4x5x3 column-major
    col = _arr(4, 1, 1, ptr=0x2000)             # column vector
    row = _arr(1, 5, 1, row_major=True, ptr=0x3000)
    leaves, nl, code, nc = lower(m.expr() + col.expr() + row.expr(), (4, 5, 3))
    L = {leaves[i].ptr: (leaves[i].sr, leaves[i].sc, leaves[i].sb) for i in range(nl)}
    assert L[0x1000] == (1, 4, 20)
    assert L[0x2000] == (1, 0, 0)
    assert L[0x3000] == (0, 1, 0)


def test_lowering_transpose_diagonal_and_cast():
    m = _arr(3, 4, ptr=0x1000)
    t = m.expr().T
    assert t.shape == (4, 3, 1) and t.strides == (3, 1, 12)
    d = m.expr().diagonal()
    assert d.shape == (3, 1, 1) and d.strides == (4, 0, 12)
    i = _arr(3, 4, dtype=np.int32, ptr=0x2000)
    leaves, nl, code, nc = lower(cast(i.expr(), np.float32) * 0.5, (3, 4, 1))
    assert code[1].opcode == capi.OP_CAST and code[1].arg == capi.I32 and code[1].dtype == capi.F32


def test_lowering_rejects_mixed_types_and_deep_trees():
    a = _arr(8, ptr=0x1000)
    b = _arr(8, dtype=np.int32, ptr=0x2000)
    with pytest.raises(TypeError):
        a.expr() + b.expr()
    x = [_arr(8, ptr=0x1000 * (k + 1)) for k in range(6)]
    e = x[0].expr()
    for k in range(1, 6):  # both operands compound: each level needs one more stack slot
        e = (x[k].expr() + x[k].expr()) * e if k > 1 else (x[1].expr() + x[1].expr()) * (e + e)
    with pytest.raises(capi.CumatArgError):
        lower(e, (8, 1, 1))


def test_sparsity_pattern_checks_run_before_upload():
    from cumat_amd.sparse import SparsityPattern
    JA = np.array([0, 2, 4, 7, 9], np.int32)
    IA = np.array([0, 1, 1, 2, 0, 3, 4, 2, 4], np.int32)
    bad_col = IA.copy()
    bad_col[6] = 5
    with pytest.raises(capi.CumatArgError):
        SparsityPattern(4, 5, JA, bad_col)
    bad_ptr = JA.copy()
    bad_ptr[2] = 1
    with pytest.raises(capi.CumatArgError):
        SparsityPattern(4, 5, bad_ptr, IA)
    with pytest.raises(capi.CumatArgError):
        SparsityPattern(4, 5, JA[:-1], IA)


def test_grid_laplacian_symmetric_with_boundary_row_sums():
    from cumat_amd.problems import grid_laplacian
    rng = np.random.default_rng(3)
    P = grid_laplacian((6, 7, 5), lambda s: np.exp(rng.standard_normal(s)))
    n = P["n"]
    A = np.zeros((n, n))
    rows = np.repeat(np.arange(n), np.diff(P["JA"]))
    A[rows, P["IA"]] = P["vals"]
    assert np.allclose(A, A.T)
    assert np.all(np.diff(P["IA"])[np.diff(rows) == 0] > 0)  # ascending columns per row
    # row sum = coefficients of the row's Dirichlet boundary edges
    bsum = np.zeros(n)
    boundary = P["ej"] < 0
    np.add.at(bsum, P["ei"][boundary], P["theta"][boundary])
    assert np.allclose(A.sum(1), bsum)
    assert np.all(np.linalg.eigvalsh(A) > 0)


def test_edge_gather_map_reassembles_values():
    from cumat_amd.problems import grid_laplacian
    rng = np.random.default_rng(4)
    P = grid_laplacian((9, 11), lambda s: np.exp(0.5 * rng.standard_normal(s)))
    th = P["theta"]
    eidx = P["eidx"]
    sign = np.sign(eidx)
    contrib = sign * th[np.abs(eidx) - 1]
    got = np.add.reduceat(contrib, P["eptr"][:-1])
    assert np.allclose(got, P["vals"], rtol=1e-14, atol=0)
    assert P["eptr"][-1] == eidx.size


def test_batched_poisson_block_structure():
    from cumat_amd.problems import batched_poisson
    P = batched_poisson(12, 4, 9, seed=1)
    ro, JA, IA = P["row_off"], P["JA"], P["IA"]
    assert ro[0] == 0 and ro[-1] == P["b"].size and JA[-1] == IA.size
    for s in range(12):
        lo, hi = ro[s], ro[s + 1]
        cols = IA[JA[lo]:JA[hi]]
        assert cols.min() >= lo and cols.max() < hi  # block diagonal
        assert hi - lo == int(P["sides"][s]) ** 2


def test_dense_matvec_argument_checks_run_before_the_call():
    from cumat_amd import dense
    A = _arr(4, 5, ptr=0x1000)
    with pytest.raises(capi.CumatArgError):
        dense.matvec(A, _arr(4, ptr=0x2000))                      # cols != rows of x
    with pytest.raises(capi.CumatArgError):
        dense.matvec(A, _arr(5, 2, ptr=0x2000))                   # x is not a column
    with pytest.raises(TypeError):
        dense.matvec(A, _arr(5, dtype=np.float64, ptr=0x2000))    # no implicit casting
    with pytest.raises(capi.CumatArgError):
        dense.matvec(_arr(4, 5, 2, ptr=0x1000), _arr(5, 1, 3, ptr=0x2000))  # batches 2 vs 3
    with pytest.raises(capi.CumatArgError):
        dense.matvec(A, _arr(5, ptr=0x2000), out=_arr(3, ptr=0x3000))      # wrong output shape
    with pytest.raises(capi.CumatArgError):
        dense.inverse_diagonal(A)                                  # not square


def test_cg_rejects_non_square_and_batched_dense_matrices():
    from cumat_amd.solvers import ConjugateGradient
    with pytest.raises(capi.CumatArgError):
        ConjugateGradient(_arr(4, 5, ptr=0x1000))
    with pytest.raises(capi.CumatArgError):
        ConjugateGradient(_arr(4, 4, 2, ptr=0x1000))


def test_ell_pattern_indices_checked_before_upload():
    """An ELLPACK index outside [-1, cols) is rejected on the host (the
    kernel gathers x at every index >= 0)."""
    from cumat_amd.sparse import ELLMatrix
    for bad in ([[0, 2], [1, -1]], [[0, -2], [1, 0]]):
        with pytest.raises(capi.CumatArgError):
            ELLMatrix(2, 2, bad, [[1.0, 1.0], [1.0, 1.0]])


def test_torch_cg_rejects_mismatched_arguments():
    """CGSolve / EdgeLaplacian check dtype, batch count, shape and device
    before any pointer reaches the kernels (here on CPU tensors, so no call
    is made at all)."""
    torch = pytest.importorskip("torch")
    from cumat_amd.problems import grid_laplacian
    from cumat_amd.torch_cg import EdgeLaplacian, cg_solve
    P = grid_laplacian((5, 6), lambda s: np.ones(s))
    lap = EdgeLaplacian(P, "cpu")
    n, E = P["n"], P["nedges"]
    th = torch.ones(2, E, dtype=torch.float64)
    b = torch.ones(2, n, dtype=torch.float64)
    cases = [
        (th.float(), b),                                        # theta float32, b float64
        (th[:1], b),                                            # 1 batch of theta, 2 of b
        (th, torch.ones(2, n + 1, dtype=torch.float64)),        # rhs of the wrong length
        (th[:, :-1], b),                                        # theta of the wrong length
        (th.to(torch.int32), b.to(torch.int32)),                # unsupported dtype
        (th, b),                                                # CPU tensors (not a GPU device)
    ]
    for t, bb in cases:
        with pytest.raises(capi.CumatArgError):
            cg_solve(t, bb, lap)
    with pytest.raises(capi.CumatArgError):
        lap.edge_grad(torch.ones(2, n, dtype=torch.float32), torch.ones(2, n, dtype=torch.float64))


def test_ell_host_checks():
    """ELLMatrix rejects index / value shapes and out-of-range indices before
    any upload."""
    from cumat_amd.sparse import ELLMatrix
    with pytest.raises(capi.CumatArgError):
        ELLMatrix(3, 4, [[0, 1], [1, 5], [2, -1]], np.ones((3, 2), np.float32))   # index 5 >= cols
    with pytest.raises(capi.CumatArgError):
        ELLMatrix(3, 4, [[0, 1], [1, 2], [2, -2]], np.ones((3, 2), np.float32))   # index < -1
    with pytest.raises(capi.CumatArgError):
        ELLMatrix(3, 4, [[0, 1], [1, 2], [2, -1]], np.ones((3, 3), np.float32))   # values of another shape


def test_reduction_routes_expressions_and_checks_them_first():
    """A whole stored array (or its plain leaf) takes the stored-array
    reductions; a transposed, broadcast or computed input is an expression
    reduced by cm_reduce_expr. Type and size errors are raised before any
    call into the library."""
    from cumat_amd import reductions as red
    a = _arr(6, 4, 2, ptr=0x1000)
    b = _arr(5, 4, 2, ptr=0x2000)
    assert red._is_stored(a) and red._is_stored(a.expr())
    assert not red._is_stored(a.expr().T)
    assert not red._is_stored(a.expr() * 2)
    assert not red._is_stored(a.expr().diagonal())
    half = _arr(6, 4, 2, dtype=np.uint16, ptr=0x3000)  # bfloat16 storage
    with pytest.raises(TypeError):
        red.reduce(half.expr() * 2, "sum")  # bfloat16 expressions are not reduced in place
    with pytest.raises(capi.CumatArgError):
        red.dot(a.expr() * 2, b.expr())  # 6x4 against 5x4
