"""Pins of the general n-diagonal RGF oracle and the fused route (no GPU)."""
import numpy as np
import pytest

from conftest import max_block_err
from oracle.general_rgf import (random_ndiag, ndiag_to_dense, band_of_dense,
                                bndiag_of_inv_general_rgf, bm_fuse_to_tridiagonal)
from oracle.rgf import bndiag_of_inv_rgf_local
from libnegf_amd.synthetic import BlockTridiag


@pytest.mark.parametrize("N,w", [(10, 2), (9, 3), (7, 1), (6, 5)])
def test_general_rgf_vs_dense(N, w):
    sizes = [8 + (i % 3) for i in range(N)]
    A = random_ndiag(sizes, w, seed=N * w)
    G = bndiag_of_inv_general_rgf(A, sizes, w)
    ref = band_of_dense(np.linalg.inv(ndiag_to_dense(A, sizes)), sizes, w)
    assert set(G) == set(ref)
    assert max_block_err(G, ref) < 1e-12


def test_general_equals_fused():
    # test/test_matinvertndiag_general_rgf_fused.jl:39-90 (n = 5, i.e. w = 2)
    N, w = 10, 2
    sizes = [6] * N
    A = random_ndiag(sizes, w, seed=1)
    G = bndiag_of_inv_general_rgf(A, sizes, w)
    fs, d, u, l = bm_fuse_to_tridiagonal(A, sizes, w)
    F = bndiag_of_inv_rgf_local(BlockTridiag(d, u, l))
    # cut the original w-band out of the fused tridiagonal band
    off = np.concatenate([[0], np.cumsum(sizes)])
    foff = np.concatenate([[0], np.cumsum(fs)])
    for (i, j), blk in G.items():
        I, J = i // w, j // w
        sub = F[(I, J)][off[i] - foff[I]:off[i + 1] - foff[I], off[j] - foff[J]:off[j + 1] - foff[J]]
        assert np.linalg.norm(sub - blk) <= 1e-12 * np.linalg.norm(blk)


def test_general_w1_is_plain_rgf():
    N = 8
    sizes = [5] * N
    A = random_ndiag(sizes, 1, seed=4)
    G = bndiag_of_inv_general_rgf(A, sizes, 1)
    B = BlockTridiag([A[(i, i)] for i in range(N)], [A[(i, i + 1)] for i in range(N - 1)],
                     [A[(i + 1, i)] for i in range(N - 1)])
    assert max_block_err(G, bndiag_of_inv_rgf_local(B)) < 1e-13


def test_general_rgf_does_not_modify_input():
    sizes = [4] * 6
    A = random_ndiag(sizes, 2, seed=2)
    before = {k: v.copy() for k, v in A.items()}
    bndiag_of_inv_general_rgf(A, sizes, 2)
    assert all(np.array_equal(A[k], before[k]) for k in A)
