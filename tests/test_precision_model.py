"""Reduced-precision products on the RGF + Keldysh energy path: what error
they leave at the benchmark's conditioning (eta = 1e-4).

The model (oracle/precision.py) is first pinned to the oracle: with FP64
products it must reproduce keldyshndiag / RGF / T.  Then two facts that
decide the mixed-precision form are checked on the device generator:

* FP32 products (the error class of 4M FP32 and 3-way-split BF16 MFMA
  products with FP32 accumulation) leave relative errors far above 1e-8 --
  the recursion amplifies the 2^-24 rounding by 1e3..1e5;
* an integer-emulated FP64 product that keeps 44 bits per row/column scale
  (Ozaki scheme II with 13 int8 moduli at k <= 512) stays below 1e-8 with a
  wide margin.
"""
import numpy as np
import pytest

from conftest import relfro
from libnegf_amd.synthetic import assemble_A, device_hamiltonian, fermi, band_edge_energies
from oracle import precision as P
from oracle.keldysh import keldyshndiag
from oracle.rgf import bndiag_of_inv_rgf_local, first_row_chain, rgf_upward

ETA = 1e-4


def _energies(dev):
    lo, hi = band_edge_energies(dev)
    return [0.3, -0.7, lo + 3e-4]


def test_model_matches_oracle():
    dev = device_hamiltonian(6, 16, seed=5)
    E = 0.3
    fL, fR = fermi(E, 0.1, 0.025), fermi(E, -0.1, 0.025)
    G, Gn, T = P.energy_sweep(dev, E, ETA, fL, fR, P.exact_product)
    A = assemble_A(dev, E, ETA)
    up = rgf_upward(A)
    Gref = bndiag_of_inv_rgf_local(A, upward=up)
    N = dev.nblocks
    for i in range(N):
        assert relfro(G[i], Gref[(i, i)]) < 1e-12
    Sd = [np.zeros((16, 16), complex) for _ in range(N)]
    Su = [np.zeros((16, 16), complex) for _ in range(N - 1)]
    Sd[0] = fL * dev.gamma_L
    Sd[-1] = Sd[-1] + fR * dev.gamma_R
    Kd, _ = keldyshndiag(A, Sd, Su, upward=up)
    for i in range(N):
        assert relfro(Gn[i], Kd[i]) < 1e-12
    row = first_row_chain(A, Gref, up)
    G1N = row[-1]
    Tref = np.real(np.trace(dev.gamma_L @ G1N @ dev.gamma_R @ G1N.conj().T))
    assert abs(T - Tref) <= 1e-12 * abs(Tref)


def test_truncated_product_is_exact_on_short_integers():
    # inputs that already fit in beta bits are not changed by the model
    rng = np.random.default_rng(0)
    X = rng.integers(-100, 100, (8, 8)) + 1j * rng.integers(-100, 100, (8, 8))
    Y = rng.integers(-100, 100, (8, 8)) + 1j * rng.integers(-100, 100, (8, 8))
    assert np.array_equal(P.truncated_product(20)(X, Y), X @ Y)
    # and beta bits of the row maximum are kept: error <= 2^-beta relative
    Z = rng.standard_normal((8, 8)) + 1j * rng.standard_normal((8, 8))
    err = np.abs(P.truncated_product(30)(Z, np.eye(8)) - Z).max()
    assert 0 < err <= 2.0 ** -30 * np.abs(Z).max()


@pytest.mark.slow
def test_fp32_misses_and_44_bits_meet_1e8():
    dev = device_hamiltonian(32, 64, seed=1234)
    worst32, worst44 = 0.0, 0.0
    for E in _energies(dev):
        fL, fR = fermi(E, 0.1, 0.025), fermi(E, -0.1, 0.025)
        G, Gn, T = P.energy_sweep(dev, E, ETA, fL, fR, P.exact_product)
        for prod, name in ((P.fp32_product, "fp32"), (P.truncated_product(44), "b44")):
            g, gn, t = P.energy_sweep(dev, E, ETA, fL, fR, prod)
            e = max(P.max_relerr(g, G), P.max_relerr(gn, Gn))
            print(f"E={E:+.4f} {name}: max block relerr {e:.2e}")
            if name == "fp32":
                worst32 = max(worst32, e)
            else:
                worst44 = max(worst44, e)
    assert worst32 > 1e-6          # FP32 products cannot reach 1e-8 here
    assert worst44 < 1e-9          # 44 kept bits do, with margin
