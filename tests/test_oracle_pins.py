"""Pins of the CPU oracle against what the mathematics fixes (no GPU).

Each check is chosen so that a plausible slip in the restatement (a dropped
term, a wrong sign or index, a transposed operand) fails:
  * RGF band vs numpy's dense inverse (bndiag_of_inv_direct semantics), on
    the reference's own uneven synthetic matrices and on the NEGF workload;
  * the b = 1 uniform chain against its closed-form Green's function;
  * block-diagonal A (G_ii = A_ii^{-1}, zero off-diagonal), N = 1, N = 2;
  * RKD vs dense G Sigma G^H (test/test_keldyshndiag.jl:86-95);
  * the Ward identity i (G - G^H) = G (2 eta I + Gamma_L + Gamma_R) G^H,
    tying RKD to RGF without a dense inverse;
  * transmission vs the dense G_{1N}; T = 1 inside the band of an ideal
    chain with exact lead self-energies; T >= 0; T_LR = T_RL.
"""
import numpy as np
import pytest

from conftest import relfro, max_block_err
from libnegf_amd.synthetic import (BlockTridiag, bm_create_synthetic_random, device_hamiltonian,
                                   assemble_A, random_hermitian_btd)
from oracle.rgf import (bndiag_of_inv_rgf_local, bndiag_of_inv_direct, band_blocks, rgf_upward,
                        first_row_chain)
from oracle.keldysh import keldyshndiag, dense_keldysh_band, transmission_dense
from oracle.backend import be_lu, be_mldivide, be_gemm, be_inv, LapackError


def test_rgf_vs_dense_reference_generator():
    A = bm_create_synthetic_random(10, 32, seed=11)
    G = bndiag_of_inv_rgf_local(A)
    D = bndiag_of_inv_direct(A)
    assert max_block_err(G, D) < 1e-12


def test_rgf_full_inverse_vs_dense():
    A = bm_create_synthetic_random(7, 12, seed=3)
    G = bndiag_of_inv_rgf_local(A, full_inverse=True)
    D = band_blocks(np.linalg.inv(A.to_dense()), A.block_sizes, ndiag=2 * 7 + 1)
    assert set(G) == set(D)
    assert max_block_err(G, D) < 1e-12


def test_rgf_negf_workload_small():
    dev = device_hamiltonian(16, 64, seed=1234)
    A = assemble_A(dev, 0.3, 1e-3)
    G = bndiag_of_inv_rgf_local(A)
    assert max_block_err(G, bndiag_of_inv_direct(A)) < 1e-12


def _chain_closed_form(N, z, eps, t):
    k = np.arange(1, N + 1)
    i = np.arange(1, N + 1)
    S = np.sqrt(2.0 / (N + 1)) * np.sin(np.outer(i, k) * np.pi / (N + 1))
    lam = eps + 2 * t * np.cos(k * np.pi / (N + 1))
    return (S / (z - lam)) @ S.T


def test_rgf_b1_chain_closed_form():
    N, eps, t, z = 12, 0.1, -0.7, 0.35 + 0.02j
    diag = [np.array([[z - eps]], dtype=complex, order="F") for _ in range(N)]
    up = [np.array([[-t]], dtype=complex, order="F") for _ in range(N - 1)]
    A = BlockTridiag(diag, up, [u.copy() for u in up])
    G = bndiag_of_inv_rgf_local(A, full_inverse=True)
    ref = _chain_closed_form(N, z, eps, t)
    for (i, j), blk in G.items():
        assert abs(blk[0, 0] - ref[i, j]) < 1e-12 * abs(ref).max()


def test_rgf_block_diagonal_and_small_N():
    rng = np.random.default_rng(0)
    sizes = [5, 3, 4]
    diag = [np.asfortranarray(rng.standard_normal((s, s)) + 1j * rng.standard_normal((s, s))
                              + 4 * np.eye(s)) for s in sizes]
    up = [np.zeros((sizes[i], sizes[i + 1]), complex, order="F") for i in range(2)]
    lo = [np.zeros((sizes[i + 1], sizes[i]), complex, order="F") for i in range(2)]
    G = bndiag_of_inv_rgf_local(BlockTridiag(diag, up, lo))
    for i in range(3):
        assert relfro(G[(i, i)], np.linalg.inv(diag[i])) < 1e-14
    for i in range(2):
        assert np.abs(G[(i, i + 1)]).max() == 0 and np.abs(G[(i + 1, i)]).max() == 0
    one = BlockTridiag([diag[0]], [], [])
    assert relfro(bndiag_of_inv_rgf_local(one)[(0, 0)], np.linalg.inv(diag[0])) < 1e-14
    A2 = bm_create_synthetic_random(2, 6, seed=5)
    assert max_block_err(bndiag_of_inv_rgf_local(A2), bndiag_of_inv_direct(A2)) < 1e-13


def test_backend_ops_vs_numpy():
    rng = np.random.default_rng(1)
    n = 9
    S = np.asfortranarray(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    B = np.asfortranarray(rng.standard_normal((n, 4)) + 1j * rng.standard_normal((n, 4)))
    lu = be_lu(S)
    assert relfro(be_mldivide("N", B, lu), np.linalg.solve(S, B)) < 1e-13
    assert relfro(be_mldivide("C", B, lu), np.linalg.solve(S.conj().T, B)) < 1e-13
    assert relfro(be_inv(S), np.linalg.inv(S)) < 1e-13
    C = be_gemm("C", "N", -1.0, S, B, 0.0, None)
    assert relfro(C, -S.conj().T @ B) < 1e-14
    with pytest.raises(LapackError):
        be_lu(np.zeros((3, 3), complex))


def test_keldysh_vs_dense_reference_shape():
    # the reference's test shape: N = 10, b ~ 32 +- 4, Sigma random Hermitian band
    A = bm_create_synthetic_random(10, 32, seed=21)
    S = random_hermitian_btd(A.block_sizes, seed=22)
    Sd, Su = keldyshndiag(A, S.diag, S.upper)
    Bd, Bu = dense_keldysh_band(A, S.diag, S.upper)
    num = sum(np.linalg.norm(a - b) ** 2 for a, b in zip(Sd + Su, Bd + Bu))
    den = sum(np.linalg.norm(b) ** 2 for b in Bd + Bu)
    assert np.sqrt(num / den) < 1e-12          # test/test_keldyshndiag.jl:94-95


def test_keldysh_ward_identity():
    N, b, eta = 8, 16, 1e-2
    dev = device_hamiltonian(N, b, seed=5)
    A = assemble_A(dev, 0.2, eta)
    G = bndiag_of_inv_rgf_local(A)
    Sd = [2 * eta * np.eye(b, dtype=complex) for _ in range(N)]
    Sd[0] = Sd[0] + dev.gamma_L
    Sd[-1] = Sd[-1] + dev.gamma_R
    Su = [np.zeros((b, b), complex, order="F") for _ in range(N - 1)]
    Gd, Gu = keldyshndiag(A, [np.asfortranarray(x) for x in Sd], Su)
    for i in range(N):
        spec = 1j * (G[(i, i)] - G[(i, i)].conj().T)
        assert relfro(Gd[i], spec) < 1e-12
    for i in range(N - 1):
        spec = 1j * (G[(i, i + 1)] - G[(i + 1, i)].conj().T)
        assert relfro(Gu[i], spec) < 1e-12


def test_transmission_dense_and_symmetry():
    dev = device_hamiltonian(6, 10, seed=9)
    A = assemble_A(dev, -0.1, 1e-3)
    buf = rgf_upward(A)
    G = bndiag_of_inv_rgf_local(A, upward=buf)
    row = first_row_chain(A, G, buf)
    g1n = row[-1]
    T = np.real(np.trace(dev.gamma_L @ g1n @ dev.gamma_R @ g1n.conj().T))
    assert abs(T - transmission_dense(A, dev.gamma_L, dev.gamma_R)) < 1e-12 * max(1.0, abs(T))
    assert T >= 0
    # T_LR = T_RL (two terminals, no absorption: take eta -> 0)
    Gd = np.linalg.inv(assemble_A(dev, -0.1, 1e-12).to_dense())
    g1n = Gd[:10, -10:]
    T = np.real(np.trace(dev.gamma_L @ g1n @ dev.gamma_R @ g1n.conj().T))
    b = 10
    gn1 = Gd[-b:, :b]
    T_rl = np.real(np.trace(dev.gamma_R @ gn1 @ dev.gamma_L @ gn1.conj().T))
    assert abs(T - T_rl) < 1e-10 * max(1.0, abs(T))


def test_transmission_ideal_chain_is_one():
    # uniform chain, eps = 0, hopping t = -1, exact semi-infinite lead self-energies
    N, t, E = 10, -1.0, 0.4
    x = E / (2 * abs(t))
    g_surf = (E - 1j * np.sqrt(4 * t * t - E * E)) / (2 * t * t)  # retarded surface GF
    sig = t * t * g_surf
    diag = [np.array([[E + 0j]], order="F") for _ in range(N)]
    diag[0] = diag[0] - sig
    diag[-1] = diag[-1] - sig
    up = [np.array([[-t + 0j]], order="F") for _ in range(N - 1)]
    A = BlockTridiag(diag, up, [u.copy() for u in up])
    gam = np.array([[-2 * sig.imag + 0j]], order="F")
    buf = rgf_upward(A)
    G = bndiag_of_inv_rgf_local(A, upward=buf)
    g1n = first_row_chain(A, G, buf)[-1]
    T = np.real(gam[0, 0] * g1n[0, 0] * gam[0, 0] * np.conj(g1n[0, 0]))
    assert abs(T - 1.0) < 1e-12 and abs(x) < 1


def test_direct_oracle_is_lapack_inverse():
    A = bm_create_synthetic_random(4, 6, seed=2)
    D = bndiag_of_inv_direct(A)
    ref = band_blocks(np.linalg.inv(A.to_dense()), A.block_sizes)
    assert max_block_err(D, ref) < 1e-13
