"""Pins of the DD-RGF oracle (no GPU), the reference's own checks
(test/test_matinvertndiag_ddrgf.jl): DD-RGF equals sequential RGF on every
band block, the Schur complement equals T22 - T21 inv(T11) T12 from a dense
inverse, and the task split / permutation match a hand count."""
import numpy as np
import pytest

from conftest import relfro, max_block_err
from libnegf_amd.synthetic import bm_create_synthetic_random, device_hamiltonian, assemble_A
from oracle.ddrgf import (get_nr_tasks, create_permutation_vector, chunks, bndiag_of_inv_ddrgf,
                          schur_complement)
from oracle.rgf import bndiag_of_inv_rgf_local


def test_task_split_and_permutation():
    # the shapes of the uneven-device workload: exactly 8 tasks each
    assert get_nr_tasks(128, 15) == (8, 0)
    assert get_nr_tasks(64, 7) == (8, 0)
    assert get_nr_tasks(32, 3) == (8, 0)
    assert get_nr_tasks(10, 3) == (3, 1)      # [s,3][s,3][s,1]
    assert get_nr_tasks(9, 3) == (-1, 0)      # leftover chunk without interior
    perm, sizes = create_permutation_vector(10, 3)
    assert perm == [0, 4, 8, 1, 2, 3, 5, 6, 7, 9]
    assert sizes == [1, 1, 1, 3, 3, 1]
    assert sorted(perm) == list(range(10))
    assert chunks(10, 3) == [(0, 1, 3), (4, 5, 3), (8, 9, 1)]


@pytest.mark.parametrize("npl,bs", [(16, 3), (10, 3), (14, 2), (12, 1)])
def test_ddrgf_equals_rgf(npl, bs):
    A = bm_create_synthetic_random(npl, 12, seed=npl * 3 + bs)
    G = bndiag_of_inv_ddrgf(A, bs)
    R = bndiag_of_inv_rgf_local(A)
    assert set(G) == set(R)
    assert max_block_err(G, R) < 2.5e-10     # the reference's tolerance (:17)


def test_ddrgf_two_levels_equals_rgf():
    A = bm_create_synthetic_random(36, 8, seed=5)
    G = bndiag_of_inv_ddrgf(A, [3, 2])       # 9 separators, then 3 tasks of [s, 2]
    assert max_block_err(G, bndiag_of_inv_rgf_local(A)) < 1e-11


def test_schur_complement_exact():
    dev = device_hamiltonian(16, 10, seed=3)
    A = assemble_A(dev, 0.1, 1e-2)
    bs = 3
    S, g, Y, Z = schur_complement(A, bs)
    ch = chunks(16, bs)
    sizes = A.block_sizes
    off = np.concatenate([[0], np.cumsum(sizes)])
    M = A.to_dense()
    seps = [c[0] for c in ch]
    inter = [l for (s, f, L) in ch for l in range(f, f + L)]
    idx = lambda ls: np.concatenate([np.arange(off[l], off[l + 1]) for l in ls])
    i1, i2 = idx(inter), idx(seps)
    Sx = M[np.ix_(i2, i2)] - M[np.ix_(i2, i1)] @ np.linalg.solve(M[np.ix_(i1, i1)], M[np.ix_(i1, i2)])
    so = np.concatenate([[0], np.cumsum([sizes[s] for s in seps])])
    for c in range(len(seps)):
        assert relfro(S.diag[c], Sx[so[c]:so[c + 1], so[c]:so[c + 1]]) < 1e-12
        if c + 1 < len(seps):
            assert relfro(S.upper[c], Sx[so[c]:so[c + 1], so[c + 1]:so[c + 2]]) < 1e-12
            assert relfro(S.low(c), Sx[so[c + 1]:so[c + 2], so[c]:so[c + 1]]) < 1e-12
    # separators two chunks apart do not couple: S is block tridiagonal
    assert np.abs(Sx[so[0]:so[1], so[2]:so[3]]).max() < 1e-12
