"""Writes the golden fixtures of tests/golden/*.npz from the CPU oracle.

The reference ships no golden vectors (its tests compare against dense
inverses inside one Julia run), so these fixtures are the oracle's outputs on
small seeded inputs; the oracle itself is pinned by tests/test_oracle_pins.py
(dense inverse, closed forms, Ward identity).  Inputs and outputs are stored
as packed column-major block arrays (see include/negf_hip.h).

Run from the repository root:  python tests/golden/make_golden.py
"""
import pathlib
import sys

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "libnegf.jl_amd"))

from libnegf_amd.synthetic import (device_hamiltonian, assemble_A, bm_create_synthetic_random,  # noqa
                                   random_hermitian_btd, fermi)
from oracle.rgf import bndiag_of_inv_rgf_local, rgf_upward, first_row_chain  # noqa: E402
from oracle.keldysh import keldyshndiag  # noqa: E402

OUT = pathlib.Path(__file__).resolve().parent


def pack(blocks):
    return np.concatenate([np.asarray(b).ravel(order="F") for b in blocks]) if blocks else \
        np.zeros(0, complex)


def pack_btd(A, prefix):
    d = {f"{prefix}_diag": pack(A.diag), f"{prefix}_upper": pack(A.upper)}
    if A.lower is not None:
        d[f"{prefix}_lower"] = pack(A.lower)
    return d


def band(G, N):
    return (pack([G[(i, i)] for i in range(N)]), pack([G[(i, i + 1)] for i in range(N - 1)]),
            pack([G[(i + 1, i)] for i in range(N - 1)]))


def main():
    # 1. RGF on the NEGF workload (uniform blocks)
    dev = device_hamiltonian(4, 8, seed=7)
    A = assemble_A(dev, 0.3, 1e-3)
    gd, gu, gl = band(bndiag_of_inv_rgf_local(A), 4)
    np.savez(OUT / "rgf_workload_N4_b8.npz", sizes=np.array(A.block_sizes), **pack_btd(A, "A"),
             G_diag=gd, G_upper=gu, G_lower=gl)
    # 2. RGF on the reference's uneven synthetic generator
    A = bm_create_synthetic_random(5, 8, seed=3)
    gd, gu, gl = band(bndiag_of_inv_rgf_local(A), 5)
    np.savez(OUT / "rgf_uneven_N5.npz", sizes=np.array(A.block_sizes), **pack_btd(A, "A"),
             G_diag=gd, G_upper=gu, G_lower=gl)
    # 3. Keldysh with a random Hermitian central operator (the reference's test shape, smaller)
    A = bm_create_synthetic_random(5, 6, seed=13)
    S = random_hermitian_btd(A.block_sizes, seed=14)
    nd, nu = keldyshndiag(A, S.diag, S.upper)
    np.savez(OUT / "keldysh_N5.npz", sizes=np.array(A.block_sizes), **pack_btd(A, "A"),
             **pack_btd(S, "S"), Gn_diag=pack(nd), Gn_upper=pack(nu))
    # 4. energy-batched workload: G^R, G^< diagonals and T at 3 energies
    N, b, eta, kT, muL, muR = 3, 6, 1e-3, 0.025, 0.1, -0.1
    dev = device_hamiltonian(N, b, seed=21)
    Es = np.array([-0.3, 0.05, 0.6])
    GR, GL, T = [], [], []
    for E in Es:
        A = assemble_A(dev, E, eta)
        buf = rgf_upward(A)
        G = bndiag_of_inv_rgf_local(A, upward=buf)
        Sd = [np.zeros((b, b), complex, order="F") for _ in range(N)]
        Sd[0] = np.asfortranarray(fermi(E, muL, kT) * dev.gamma_L)
        Sd[-1] = np.asfortranarray(Sd[-1] + fermi(E, muR, kT) * dev.gamma_R)
        Su = [np.zeros((b, b), complex, order="F") for _ in range(N - 1)]
        gd_, _ = keldyshndiag(A, Sd, Su, upward=buf)
        g1n = first_row_chain(A, G, buf)[-1]
        GR.append(pack([G[(i, i)] for i in range(N)]))
        GL.append(pack([1j * x for x in gd_]))
        T.append(np.real(np.trace(dev.gamma_L @ g1n @ dev.gamma_R @ g1n.conj().T)))
    np.savez(OUT / "energies_N3_b6.npz", N=N, b=b, eta=eta, kT=kT, mu_L=muL, mu_R=muR,
             energies=Es, H_diag=pack(dev.H_diag), H_upper=pack(dev.H_upper),
             gamma_L=pack([dev.gamma_L]), gamma_R=pack([dev.gamma_R]), GR_diag=np.array(GR),
             Gless_diag=np.array(GL), T=np.array(T))
    print("wrote", sorted(p.name for p in OUT.glob("*.npz")))


if __name__ == "__main__":
    main()
