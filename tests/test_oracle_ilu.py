"""The oracle's block ILU(0) (block_ilu.cpp:3-84, block_graph.cpp:7-44) pinned
without retyping it:

* a dense block graph has no fill, so ILU(0) is the exact block LU and the
  apply solves exactly, in any row order (check_1d.cpp:110-142);
* the natural-order path graph has no fill either (same check);
* on the 2D periodic element graph, ILU(0) is characterised by (L U)_ij = A_ij
  on every pattern block with L unit block-lower and U block-upper (Saad,
  Iterative Methods, Prop. 10.4): an independent dense factorisation that
  satisfies that identity gives M = L U, and the oracle's apply must be M^-1;
* in red-black order the factors have the closed form L = [I 0; A_br D_r^-1 I],
  U = [D_r A_rb; 0 S_b] with S_b the block diagonal of
  D_b - A_br D_r^-1 A_rb, because no black-black coupling exists.
"""
import numpy as np
import pytest

import oracle_py as O


def grid_adjacency(nx, ny):
    adj = []
    for y in range(ny):
        for x in range(nx):
            c = x + nx * y
            adj.append(sorted({c, (x + nx - 1) % nx + nx * y, (x + 1) % nx + nx * y,
                               x + nx * ((y + ny - 1) % ny), x + nx * ((y + 1) % ny)}))
    return adj


def random_blocks(adj, width, rng, strength=3.0):
    blocks = {}
    for r, row in enumerate(adj):
        for c in row:
            b = 0.3 * rng.standard_normal((width, width))
            if r == c:
                b += strength * np.eye(width) + np.diag(rng.random(width))
            blocks[(r, c)] = b
    return blocks


def dense(adj, blocks, width, order=None):
    rows = len(adj)
    order = list(range(rows)) if order is None else list(order)
    inv = np.empty(rows, dtype=int)
    inv[order] = np.arange(rows)
    A = np.zeros((rows * width, rows * width))
    for (r, c), b in blocks.items():
        i, j = inv[r], inv[c]
        A[i * width:(i + 1) * width, j * width:(j + 1) * width] = b
    return A


def permute(v, order, width):
    return np.concatenate([v[o * width:(o + 1) * width] for o in order])


def unpermute(v, order, width):
    out = np.empty_like(v)
    for i, o in enumerate(order):
        out[o * width:(o + 1) * width] = v[i * width:(i + 1) * width]
    return out


def ilu0_dense(A, pattern, width):
    """Block ILU(0) by Gaussian elimination restricted to the pattern
    (Saad, Alg. 10.4, IKJ form, one block at a time), returning L and U."""
    n = A.shape[0] // width
    W = A.copy()
    blk = lambda i, j: (slice(i * width, (i + 1) * width), slice(j * width, (j + 1) * width))
    for i in range(1, n):
        for k in range(i):
            if (i, k) not in pattern:
                continue
            W[blk(i, k)] = W[blk(i, k)] @ np.linalg.inv(W[blk(k, k)])
            for j in range(k + 1, n):
                if (i, j) in pattern:
                    W[blk(i, j)] -= W[blk(i, k)] @ W[blk(k, j)]
    L = np.eye(A.shape[0])
    U = np.zeros_like(A)
    for (i, j) in pattern:
        if i > j:
            L[blk(i, j)] = W[blk(i, j)]
        else:
            U[blk(i, j)] = W[blk(i, j)]
    return L, U


@pytest.mark.parametrize("order_kind", ["natural", "reversed"])
def test_full_graph_is_exact(order_kind):
    rng = np.random.default_rng(1)
    rows, width = 7, 4
    adj = [list(range(rows)) for _ in range(rows)]
    blocks = random_blocks(adj, width, rng)
    order = list(range(rows)) if order_kind == "natural" else list(range(rows))[::-1]
    exact = rng.standard_normal(rows * width)
    A = dense(adj, blocks, width)
    x = O.block_ilu_solve(adj, blocks, order, A @ exact)
    assert np.linalg.norm(x - exact) < 1e-12


def test_path_graph_natural_is_exact():
    rng = np.random.default_rng(2)
    rows, width = 7, 4
    adj = [[j for j in range(rows) if abs(i - j) <= 1] for i in range(rows)]
    blocks = random_blocks(adj, width, rng)
    exact = rng.standard_normal(rows * width)
    x = O.block_ilu_solve(adj, blocks, list(range(rows)), dense(adj, blocks, width) @ exact)
    assert np.linalg.norm(x - exact) < 1e-12


@pytest.mark.parametrize("nx,ny,order_kind", [(4, 4, "natural"), (5, 4, "natural"), (3, 4, "natural"),
                                              (4, 4, "redblack"), (6, 4, "redblack")])
def test_pattern_identity_on_periodic_grid(nx, ny, order_kind):
    rng = np.random.default_rng(nx * 10 + ny)
    width = 4
    adj = grid_adjacency(nx, ny)
    blocks = random_blocks(adj, width, rng)
    rows = nx * ny
    order = list(range(rows))
    if order_kind == "redblack":
        order = [c for colour in (0, 1) for c in range(rows) if (c % nx + c // nx) % 2 == colour]
    inv = np.empty(rows, dtype=int)
    inv[order] = np.arange(rows)
    pattern = {(int(inv[r]), int(inv[c])) for r, row in enumerate(adj) for c in row}
    A = dense(adj, blocks, width, order)
    L, U = ilu0_dense(A, pattern, width)
    LU = L @ U
    for (i, j) in pattern:  # the defining identity of ILU(0)
        sl = (slice(i * width, (i + 1) * width), slice(j * width, (j + 1) * width))
        assert np.abs(LU[sl] - A[sl]).max() < 1e-12
    r = rng.standard_normal(rows * width)
    want = unpermute(np.linalg.solve(LU, permute(r, order, width)), order, width)
    got = O.block_ilu_solve(adj, blocks, order, r)
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-12


def test_red_black_closed_form():
    rng = np.random.default_rng(9)
    nx, ny, width = 4, 6, 3
    adj = grid_adjacency(nx, ny)
    blocks = random_blocks(adj, width, rng)
    rows = nx * ny
    red = [c for c in range(rows) if (c % nx + c // nx) % 2 == 0]
    black = [c for c in range(rows) if (c % nx + c // nx) % 2 == 1]
    order = red + black
    A = dense(adj, blocks, width, order)
    h = len(red) * width
    Dr, Arb, Abr, Db = A[:h, :h], A[:h, h:], A[h:, :h], A[h:, h:]
    G = Abr @ np.linalg.solve(Dr, Arb)
    Sb = Db.copy()
    for k in range(len(black)):
        sl = slice(k * width, (k + 1) * width)
        Sb[sl, sl] -= G[sl, sl]
    L = np.block([[np.eye(h), np.zeros_like(Arb)], [Abr @ np.linalg.inv(Dr), np.eye(A.shape[0] - h)]])
    U = np.block([[Dr, Arb], [np.zeros_like(Abr), Sb]])
    r = rng.standard_normal(rows * width)
    want = unpermute(np.linalg.solve(L @ U, permute(r, order, width)), order, width)
    got = O.block_ilu_solve(adj, blocks, order, r)
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-12


@pytest.mark.parametrize("ordering", ["natural", "redblack"])
def test_dg_operator_ilu_matches_dense(ordering):
    """orc_ilu_refresh on the DG element blocks (DGOperator::block) against the
    dense ILU(0) of the same blocks, 4x4, k=1."""
    args = ["--nx", "4", "--ny", "4", "--k", "1", "--testcase", "diagonal", "--cfl", "5"]
    o = O.Oracle(args)
    u = o.initial()
    rate = o.freeze(u)
    o.set_step(5.0 / rate, u)
    o.dg_refresh(u)
    o.ilu_refresh(ordering)
    adj = grid_adjacency(4, 4)
    blocks = {(r, c): o.dg_block(r, c) for r, row in enumerate(adj) for c in row}
    order = list(range(16))
    if ordering == "redblack":
        order = [c for colour in (0, 1) for c in range(16) if (c % 4 + c // 4) % 2 == colour]
    r = np.random.default_rng(3).standard_normal(o.size)
    assert np.linalg.norm(o.ilu_apply(r) - O.block_ilu_solve(adj, blocks, order, r)) <= 1e-13 * np.linalg.norm(
        o.ilu_apply(r))
