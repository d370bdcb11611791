"""Pins of the selected-inverse oracle (no GPU): the reference's own checks
(test/test_selected_inverse.jl) against a brute-force no-pivot LU and
numpy's dense inverse."""
import numpy as np
import pytest

from conftest import relfro
from oracle.selinv import (set_sparse_Block, symbolic_fill, Block_factorization, Block_inverse,
                           full, dense_nopivot_lu)

RIND = [1, 1, 1, 2, 3, 3, 3, 4, 5]          # test/test_selected_inverse.jl:8-9
CIND = [1, 3, 5, 2, 3, 4, 5, 4, 5]


@pytest.mark.parametrize("b", [1, 20])
def test_block_factorization_vs_nopivot_lu(b):
    npl = 5
    sizes = [b] * npl
    T = set_sparse_Block(sizes, RIND, CIND, True)
    F = Block_factorization(T, npl)
    assert set(F) == symbolic_fill(T.keys(), npl)
    lu = dense_nopivot_lu(full(T, sizes))
    FM = full(F, sizes)
    assert relfro(FM, lu) < 1e-14
    L = np.tril(FM, -1) + np.eye(FM.shape[0])
    U = np.triu(FM)
    assert relfro(L @ U, full(T, sizes)) < 1e-14


@pytest.mark.parametrize("b", [1, 12])
def test_block_inverse_vs_dense(b):
    npl = 5
    sizes = [b] * npl
    T = set_sparse_Block(sizes, RIND, CIND, True, rng=np.random.default_rng(3))
    X = Block_inverse(Block_factorization(T, npl), npl)
    inv = np.linalg.inv(full(T, sizes))
    for (i, j), blk in X.items():
        ref = inv[i * b:(i + 1) * b, j * b:(j + 1) * b]
        assert np.abs(blk - ref).max() < 1e-13 * max(1.0, np.abs(ref).max()), (i, j)
    # diag(A X) = 1 (the reference checks it on the b = 1 case)
    if b == 1:
        d = np.diag(full(T, sizes) @ full(X, sizes))
        assert np.abs(1 - d).max() < 1e-11


def test_block_inverse_tridiagonal_matches_rgf_band():
    from oracle.rgf import bndiag_of_inv_rgf_local
    from libnegf_amd.synthetic import BlockTridiag
    npl, b = 6, 7
    rng = np.random.default_rng(8)
    T = {}
    for i in range(npl):
        T[(i, i)] = np.asfortranarray(rng.random((b, b)) + 1j * rng.random((b, b)) + 3 * b * np.eye(b))
    for i in range(npl - 1):
        T[(i, i + 1)] = np.asfortranarray(rng.random((b, b)) + 1j * rng.random((b, b)))
        T[(i + 1, i)] = np.asfortranarray(rng.random((b, b)) + 1j * rng.random((b, b)))
    X = Block_inverse(Block_factorization(T, npl), npl)
    A = BlockTridiag([T[(i, i)] for i in range(npl)], [T[(i, i + 1)] for i in range(npl - 1)],
                     [T[(i + 1, i)] for i in range(npl - 1)])
    G = bndiag_of_inv_rgf_local(A)
    for k in G:
        assert relfro(X[k], G[k]) < 1e-13
