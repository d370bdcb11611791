"""The reference's RGF sweep issued op by op through the backend binding
(libnegf_amd.perop: the be_lu / be_mldivide / be_gemm / be_ctranspose
sequence of src/matinvertndiag_rgf.jl:114-200 that backend_amd.jl executes
under the unchanged sweep), against the oracle's RGF, which runs the same
sequence on the CPU."""
import numpy as np
import pytest

from conftest import max_block_err

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from libnegf_amd import perop  # noqa: E402
from libnegf_amd.btd import DeviceBlockTridiag  # noqa: E402
from libnegf_amd.synthetic import device_hamiltonian, assemble_A, bm_create_synthetic_random  # noqa
from oracle.rgf import bndiag_of_inv_rgf_local  # noqa: E402


def _band(h):
    G = {(i, i): h.diag[i] for i in range(h.nblocks)}
    for i in range(h.nblocks - 1):
        G[(i, i + 1)] = h.upper[i]
        G[(i + 1, i)] = h.lower[i]
    return G


def _perop(A):
    Min = DeviceBlockTridiag.from_host(A)
    Mout = Min.similar()
    perop.bndiag_of_inv_rgf_local_perop(Mout, Min)
    return _band(Mout.to_host())


@pytest.mark.parametrize("N,b", [(1, 48), (2, 32), (6, 64), (5, 130)])
def test_perop_sweep_vs_oracle(N, b):
    print(f"perop sweep N={N} b={b}", flush=True)
    A = assemble_A(device_hamiltonian(N, b, seed=N * b), 0.2, 1e-3)
    assert max_block_err(_perop(A), bndiag_of_inv_rgf_local(A)) < 1e-12


def test_perop_sweep_uneven_blocks():
    print("perop sweep uneven", flush=True)
    A = bm_create_synthetic_random(7, 20, seed=71)
    assert max_block_err(_perop(A), bndiag_of_inv_rgf_local(A)) < 1e-12


def test_perop_sweep_production_block_size():
    """b = 512 at eta = 1e-4, E = 0.3 (the single-energy drop-in's block size
    and broadening, 6 layers): getrf, getrs (N and C) with 512 right-hand
    sides and the GEMMs as one sequence.  The bar is 5x the FP64 spread of
    two correct formulations on this input (tests/_floors.py)."""
    from _floors import rgf_floor
    print("perop sweep N=6 b=512 eta=1e-4", flush=True)
    A = assemble_A(device_hamiltonian(6, 512, seed=1234), 0.3, 1e-4)
    floor, ref = rgf_floor(A)
    err = max_block_err(_perop(A), ref)
    print(f"  perop vs oracle {err:.2e} (floor {floor:.2e})")
    assert err < 5 * max(floor, 1e-13)
