"""negf_rgf_rkd_mixed: the band sweep (RGF, and RGF + RKD in one pass) with
its products in FP32 or FP64 emulated on the int8 matrix cores, the
per-instance error estimate and the FP64 re-run of instances above the
tolerance, against the oracle."""
import numpy as np
import pytest

from conftest import max_block_err, relfro

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from libnegf_amd import api  # noqa: E402
from libnegf_amd.btd import DeviceBlockTridiag  # noqa: E402
from libnegf_amd.synthetic import (device_hamiltonian, assemble_A,  # noqa: E402
                                   bm_create_synthetic_random, random_hermitian_btd)
from oracle.rgf import bndiag_of_inv_rgf_local  # noqa: E402
from oracle.keldysh import keldyshndiag  # noqa: E402


def _band(h):
    G = {(i, i): h.diag[i] for i in range(h.nblocks)}
    for i in range(h.nblocks - 1):
        G[(i, i + 1)] = h.upper[i]
        G[(i + 1, i)] = h.lower[i]
    return G


@pytest.mark.parametrize("precision", ["emulated", "fp32"])
def test_rgf_mixed_band_vs_oracle(precision):
    """N = 8, b = 256, eta = 1e-4 at three energies.  Emulated products keep
    every instance under 1e-8 with no re-run, and the estimate bounds the
    error; FP32 products estimate far above 1e-8, so every instance is
    re-run in FP64 and matches the oracle at the FP64 level."""
    N, b, eta = 8, 256, 1e-4
    print(f"rgf mixed {precision} N={N} b={b}", flush=True)
    dev = device_hamiltonian(N, b, seed=77)
    mats = [assemble_A(dev, E, eta) for E in (-0.4, 0.3, 0.9)]
    Min = DeviceBlockTridiag.from_host(mats)
    Mout = Min.similar()
    est, fb = api.bndiag_of_inv_rgf_local_mixed(Mout, Min, precision, tol=1e-8)
    print(f"  est {est}, fallback {fb}")
    for k, A in enumerate(mats):
        err = max_block_err(_band(Mout.to_host(k)), bndiag_of_inv_rgf_local(A))
        print(f"  instance {k}: relerr {err:.2e}")
        if fb[k]:
            assert err < 1e-11
        else:
            assert err < 1e-8 and err <= est[k]
    if precision == "fp32":
        assert fb.all()
    else:
        assert not fb.any()


def test_rgf_rkd_mixed_keldysh_band_and_uneven_blocks():
    """RGF + RKD with a random Hermitian central operator on the reference's
    uneven generator, emulated products: G and G^n within 1e-8 of the
    oracle; with tol = 0 every instance is re-run and matches at the FP64
    level."""
    print("rgf+rkd mixed uneven", flush=True)
    A = bm_create_synthetic_random(6, 40, seed=19)
    S = random_hermitian_btd(A.block_sizes, seed=20)
    Min = DeviceBlockTridiag.from_host(A)
    Sd = DeviceBlockTridiag.from_host(S)
    nd_ref, nu_ref = keldyshndiag(A, S.diag, S.upper)
    G_ref = bndiag_of_inv_rgf_local(A)
    for tol, bar in ((1e-8, 1e-8), (0.0, 1e-12)):
        Mout, Gn = Min.similar(), Sd.similar()
        est, fb = api.bndiag_of_inv_rgf_local_mixed(Mout, Min, "emulated", tol=tol, S=Sd, Gn=Gn)
        assert bool(fb[0]) == (tol == 0.0)
        g = Gn.to_host()
        eG = max_block_err(_band(Mout.to_host()), G_ref)
        eN = max(max(relfro(g.diag[i], nd_ref[i]) for i in range(6)),
                 max(relfro(g.upper[i], nu_ref[i]) for i in range(5)))
        print(f"  tol={tol}: est {est[0]:.2e}, G {eG:.2e}, G^n {eN:.2e}")
        assert eG < bar and eN < bar
