"""The fused / general n-diagonal route and the CSC entry point on the GPU."""
import numpy as np
import pytest
import scipy.sparse as sp

from conftest import max_block_err

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from libnegf_amd import general, csc  # noqa: E402
from libnegf_amd.synthetic import device_hamiltonian, assemble_A, bm_create_synthetic_random  # noqa
from oracle.general_rgf import random_ndiag, bndiag_of_inv_general_rgf, ndiag_to_dense, band_of_dense  # noqa
from oracle.rgf import bndiag_of_inv_rgf_local  # noqa: E402


@pytest.mark.parametrize("N,w", [(10, 2), (9, 3), (7, 1)])
def test_general_rgf_fused_route_vs_oracle(N, w):
    print(f"general rgf N={N} w={w}", flush=True)
    sizes = [16 + (i % 3) for i in range(N)]
    A = random_ndiag(sizes, w, seed=7 * N + w)
    G = general.bndiag_of_inv_general_rgf(A, sizes, w)
    ref = bndiag_of_inv_general_rgf(A, sizes, w)
    assert set(G) == set(ref)
    assert max_block_err(G, ref) < 1e-12


def _csc_of(A):
    return sp.csc_matrix(A.to_dense())


@pytest.mark.parametrize("method", [0, 1, 2])
def test_run_bndiag_wrapper_uniform(method):
    print(f"csc wrapper method={method}", flush=True)
    dev = device_hamiltonian(16, 24, seed=3)
    A = assemble_A(dev, 0.2, 1e-2)
    out = csc.run_bndiag_wrapper(_csc_of(A), A.block_sizes, method=method)
    ref = bndiag_of_inv_rgf_local(A)
    off = np.concatenate([[0], np.cumsum(A.block_sizes)])
    D = out.toarray()
    got = {k: D[off[k[0]]:off[k[0] + 1], off[k[1]]:off[k[1] + 1]] for k in ref}
    assert max_block_err(got, ref) < 2.5e-10
    # nothing outside the band
    mask = np.ones_like(D, dtype=bool)
    for (i, j) in ref:
        mask[off[i]:off[i + 1], off[j]:off[j + 1]] = False
    assert np.abs(D[mask]).max() == 0


def test_run_bndiag_wrapper_uneven_and_hermitian_storage():
    print("csc wrapper uneven / hermitian", flush=True)
    A = bm_create_synthetic_random(9, 20, seed=8)
    out = csc.run_bndiag_wrapper(_csc_of(A), A.block_sizes)
    ref = bndiag_of_inv_rgf_local(A)
    off = np.concatenate([[0], np.cumsum(A.block_sizes)])
    D = out.toarray()
    got = {k: D[off[k[0]]:off[k[0] + 1], off[k[1]]:off[k[1] + 1]] for k in ref}
    assert max_block_err(got, ref) < 1e-12
    # Hermitian input given with only its upper band: the lower band is read as the adjoint
    dev = device_hamiltonian(8, 16, seed=2)
    full = _herm(dev)
    upper_only = full.copy()
    for i in range(7):      # Hermitian storage drops the lower off-diagonal blocks only
        upper_only[(i + 1) * 16:(i + 2) * 16, i * 16:(i + 1) * 16] = 0
    out = csc.run_bndiag_wrapper(sp.csc_matrix(upper_only), [16] * 8, is_hermitian=True,
                                 method=1)
    ref = band_of_dense(np.linalg.inv(full), [16] * 8, 1)
    Dh = out.toarray()
    o = np.arange(9) * 16
    got = {k: Dh[o[k[0]]:o[k[0] + 1], o[k[1]]:o[k[1] + 1]] for k in ref}
    assert max_block_err(got, ref) < 1e-11


def _herm(dev):
    """A Hermitian, well-conditioned block-tridiagonal matrix: 5 I - H."""
    from libnegf_amd.synthetic import BlockTridiag
    N, b = dev.nblocks, dev.b
    d = [np.asfortranarray(5 * np.eye(b) - h) for h in dev.H_diag]
    u = [np.asfortranarray(-x) for x in dev.H_upper]
    return BlockTridiag(d, u, None).to_dense()
