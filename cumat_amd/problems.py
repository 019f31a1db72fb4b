"""Seeded synthetic workloads (host side, numpy): grid Laplacians in CSR.

* poisson2d: the 5-point matrix of the reference's CG / CSRMV benchmarks,
  rows in IDX(x, y) = y + x*g order with ascending columns, diagonal 4,
  off-diagonal -1, Dirichlet boundary
  (benchmarks/conjugate_gradient/Implementation_cuMat.cu:29-41).
* grid_laplacian: variable-coefficient 2D/3D edge Laplacians. Every grid edge
  (including the Dirichlet edges to the boundary) carries a coefficient c_e;
  A_ii = sum of incident c_e, A_ij = -c_e. The same routine returns the edge
  endpoints and the CSR-entry -> edge gather map used by the adjoint kernels.
* batched_poisson: block-diagonal concatenation of uneven shifted Poisson
  systems (persistent batched CG workload).
* block_poisson: a block-CSR matrix of D x D blocks (D float4 rows each), the
  Kronecker product of a 2D or 3D Poisson pattern with an SPD coupling block
  plus a random SPD perturbation of the diagonal blocks (block-scalar CG
  workload).

All random inputs come from numpy's PCG64 generator seeded explicitly.
"""
from __future__ import annotations

import numpy as np


def poisson2d(g: int, shift: float = 0.0, dtype=np.float32):
    """Returns (JA, IA, vals) of the g*g 5-point Poisson matrix (+ shift*I)."""
    n = g * g
    row = np.arange(n, dtype=np.int64)
    x, y = row // g, row % g
    cols = np.stack([row - g, row - 1, row, row + 1, row + g], axis=1)
    valid = np.stack([x > 0, y > 0, np.ones(n, bool), y < g - 1, x < g - 1], axis=1)
    vals = np.broadcast_to(np.array([-1.0, -1.0, 4.0 + shift, -1.0, -1.0]), (n, 5))
    JA = np.zeros(n + 1, np.int32)
    JA[1:] = np.cumsum(valid.sum(1))
    return JA, cols[valid].astype(np.int32), vals[valid].astype(dtype)


def grid_laplacian(dims, coef_sampler, dtype=np.float64, edges=True):
    """Variable-coefficient edge Laplacian on a grid of shape 'dims' (C order:
    the last axis is the fastest). coef_sampler(shape) returns the edge
    coefficients of one axis, an array of 'dims' with dims[a]+1 on axis a
    (entries 0 and dims[a] along that axis are the Dirichlet boundary edges).

    Returns dict with JA, IA, vals, n, nedges, theta (edge coefficients) and,
    if edges, ei, ej (edge endpoints, ej = -1 for boundary edges), eptr, eidx
    (gather map of cm_edge_assemble)."""
    dims = tuple(int(d) for d in dims)
    d = len(dims)
    n = int(np.prod(dims))
    strides = [int(np.prod(dims[a + 1:])) for a in range(d)]
    coords = np.indices(dims).reshape(d, -1)
    node = np.arange(n, dtype=np.int64)
    E, offs, off = [], [], 0
    for a in range(d):
        shp = list(dims)
        shp[a] += 1
        e = np.asarray(coef_sampler(tuple(shp)), dtype=np.float64)
        E.append(e)
        offs.append(off)
        off += e.size
    nedges = off
    theta = np.concatenate([e.ravel() for e in E])

    def edge_id(a, shift):
        shp = list(dims)
        shp[a] += 1
        c = coords.copy()
        c[a] = c[a] + shift
        return offs[a] + np.ravel_multi_index(tuple(c), tuple(shp))

    minus_e = [edge_id(a, 0) for a in range(d)]
    plus_e = [edge_id(a, 1) for a in range(d)]
    # slots in ascending column order: minus neighbours (largest stride
    # first), diagonal, plus neighbours (smallest stride first)
    slot_cols, slot_valid, slot_edges, slot_sign = [], [], [], []
    for a in range(d):
        slot_cols.append(node - strides[a])
        slot_valid.append(coords[a] > 0)
        slot_edges.append(minus_e[a])
    slot_cols.append(node)
    slot_valid.append(np.ones(n, bool))
    slot_edges.append(None)
    for a in reversed(range(d)):
        slot_cols.append(node + strides[a])
        slot_valid.append(coords[a] < dims[a] - 1)
        slot_edges.append(plus_e[a])
    ns = len(slot_cols)
    cols = np.stack(slot_cols, 1)
    valid = np.stack(slot_valid, 1)
    diag = np.zeros(n)
    for a in range(d):
        diag += theta[minus_e[a]] + theta[plus_e[a]]
    vals = np.zeros((n, ns))
    for s in range(ns):
        vals[:, s] = diag if slot_edges[s] is None else -theta[slot_edges[s]]
    JA = np.zeros(n + 1, np.int32)
    JA[1:] = np.cumsum(valid.sum(1))
    IA = cols[valid].astype(np.int32)
    V = vals[valid].astype(dtype)
    del cols, vals
    if not edges:
        return {"JA": JA, "IA": IA, "vals": V, "n": n, "nnz": int(IA.size), "nedges": nedges, "theta": theta}
    # gather map: off-diagonal entry -> -(edge+1); diagonal -> +(edge+1) x 2d
    cnt = np.where(np.arange(ns) == d, 2 * d, 1)
    per_entry = np.broadcast_to(cnt, (n, ns))[valid]
    eptr = np.zeros(per_entry.size + 1, np.int32)
    eptr[1:] = np.cumsum(per_entry)
    lists = np.zeros((n, ns, 2 * d), np.int64)
    for s in range(ns):
        if slot_edges[s] is None:
            inc = [minus_e[a] for a in range(d)] + [plus_e[a] for a in range(d)]
            lists[:, s, :] = np.stack(inc, 1) + 1
        else:
            lists[:, s, 0] = -(slot_edges[s] + 1)
    mask = np.zeros((n, ns, 2 * d), bool)
    mask[:, :, 0] = True
    mask[:, d, :] = True
    eidx = lists[valid][mask[valid]].astype(np.int32)
    # edge endpoints
    ei = np.empty(nedges, np.int32)
    ej = np.empty(nedges, np.int32)
    for a in range(d):
        shp = list(dims)
        shp[a] += 1
        q = np.indices(shp).reshape(d, -1)
        qa = q[a]
        lo = q.copy()
        lo[a] = np.clip(qa - 1, 0, dims[a] - 1)
        hi = q.copy()
        hi[a] = np.clip(qa, 0, dims[a] - 1)
        nlo = np.ravel_multi_index(tuple(lo), dims)
        nhi = np.ravel_multi_index(tuple(hi), dims)
        first = qa == 0
        last = qa == dims[a]
        sl = slice(offs[a], offs[a] + E[a].size)
        ei[sl] = np.where(first, nhi, nlo)
        ej[sl] = np.where(first | last, -1, nhi)
    return {"JA": JA, "IA": IA, "vals": V, "n": n, "nnz": int(IA.size), "nedges": nedges,
            "theta": theta, "ei": ei, "ej": ej, "eptr": eptr, "eidx": eidx}


def lognormal_sampler(rng, mu=0.0, sigma=1.0):
    return lambda shape: np.exp(mu + sigma * rng.standard_normal(shape))


def batched_poisson(nsys: int, side_lo: int, side_hi: int, seed: int, dtype=np.float32):
    """nsys independent 2D Poisson systems, side ~ U{side_lo..side_hi},
    A_k = P_side + sigma_k I with sigma_k ~ U[0,1], b ~ U[-1,1]; returned as
    one block-diagonal CSR with global column indices and row offsets."""
    rng = np.random.default_rng(seed)
    sides = rng.integers(side_lo, side_hi + 1, nsys)
    sig = rng.uniform(0.0, 1.0, nsys)
    JAs, IAs, Vs = [], [], []
    row_off = np.zeros(nsys + 1, np.int64)
    nnz_off = 0
    for s in range(nsys):
        JA, IA, V = poisson2d(int(sides[s]), float(sig[s]), dtype)
        n = JA.size - 1
        JAs.append(JA[:-1].astype(np.int64) + nnz_off)
        IAs.append(IA.astype(np.int64) + row_off[s])
        Vs.append(V)
        row_off[s + 1] = row_off[s] + n
        nnz_off += IA.size
    JA = np.concatenate(JAs + [np.array([nnz_off])]).astype(np.int32)
    IA = np.concatenate(IAs).astype(np.int32)
    V = np.concatenate(Vs).astype(dtype)
    b = rng.uniform(-1.0, 1.0, int(row_off[-1])).astype(dtype)
    return {"JA": JA, "IA": IA, "vals": V, "row_off": row_off.astype(np.int32), "b": b,
            "sides": sides, "sigma": sig}


def block_poisson(dims, D: int = 3, seed: int = 0, coupling: float = 0.5):
    """Block-CSR (JA, IA, vals) over a 2D or 3D grid (C order, Dirichlet):
    block (i, j) = L_ij * K for the 5- / 7-point Laplacian L and the SPD
    D x D block K = I + coupling * (tridiagonal -1/2, 1, -1/2), plus on each
    diagonal block a seeded positive semi-definite term (0.1/D) R R^T with R
    uniform in [-1, 1], so the matrix stays SPD. vals is float32 (nnz, D, 4):
    D float4 rows, w = 0."""
    dims = tuple(int(d) for d in dims)
    n = int(np.prod(dims))
    idx = np.arange(n, dtype=np.int64).reshape(dims)
    coords = np.unravel_index(np.arange(n), dims)
    strides = [int(np.prod(dims[a + 1:])) for a in range(len(dims))]
    offs, valid = [], []
    for a in reversed(range(len(dims))):  # most distant neighbour first: ascending columns
        offs.append(-strides[a])
        valid.append(coords[a] > 0)
    offs.append(0)
    valid.append(np.ones(n, bool))
    for a in range(len(dims)):
        offs.append(strides[a])
        valid.append(coords[a] < dims[a] - 1)
    order = np.argsort(offs, kind="stable")
    offs = [offs[k] for k in order]
    valid = np.stack([valid[k] for k in order], axis=1)
    row = idx.reshape(-1)
    cols = row[:, None] + np.array(offs)[None, :]
    JA = np.zeros(n + 1, np.int32)
    JA[1:] = np.cumsum(valid.sum(1))
    IA = cols[valid].astype(np.int32)
    K = np.eye(D) + coupling * (np.eye(D) - 0.5 * np.eye(D, k=1) - 0.5 * np.eye(D, k=-1))
    lap = np.where(np.array(offs) == 0, 2.0 * len(dims), -1.0)
    scal = np.broadcast_to(lap, valid.shape)[valid]
    vals = np.zeros((IA.size, D, 4), np.float64)
    vals[:, :, :D] = scal[:, None, None] * K[None, :, :]
    rng = np.random.default_rng(seed)
    diag_pos = np.nonzero(IA == np.repeat(row, valid.sum(1)))[0]
    R = rng.uniform(-1.0, 1.0, (n, D, D))
    vals[diag_pos, :, :D] += (0.1 / D) * (R @ np.transpose(R, (0, 2, 1)))  # PSD: the sum stays SPD
    return JA, IA, vals.astype(np.float32)

