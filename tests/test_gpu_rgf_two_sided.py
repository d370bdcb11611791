"""The two-sided band sweep (negf_rgf_schedule: the upward recursion from
both ends of a uniform chain at once, meeting at layer N/2) against the
oracle's RGF and the one-sided sweep, through the same C entry negf_rgf.

Shapes cover both parities of N (an even N pads the right part by an
identity layer), N = 4 (two layers per part), Hermitian coupling storage
(A_{i+1,i} read as A_{i,i+1}^H), outputs without the off-diagonal blocks, a
batch above the automatic limit of 16 (forced), and a singular block whose
original layer the info array must name.
"""
import numpy as np
import pytest

from conftest import max_block_err

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from libnegf_amd import api  # noqa: E402
from libnegf_amd._lib import call  # noqa: E402
from libnegf_amd.btd import DeviceBlockTridiag  # noqa: E402
from libnegf_amd.synthetic import (BlockTridiag, bm_create_synthetic_random,  # noqa: E402
                                   device_hamiltonian, assemble_A)
from oracle.rgf import bndiag_of_inv_rgf_local as rgf_ref  # noqa: E402


def _band(host, upper=True, lower=True):
    G = {(i, i): host.diag[i] for i in range(host.nblocks)}
    for i in range(host.nblocks - 1):
        if upper:
            G[(i, i + 1)] = host.upper[i]
        if lower:
            G[(i + 1, i)] = host.lower[i]
    return G


def _sweep(mats, schedule):
    Min = DeviceBlockTridiag.from_host(mats)
    Mout = Min.similar(hermitian=False)
    call("negf_rgf_schedule", schedule)
    try:
        api.bndiag_of_inv_rgf_local(Mout, Min, api.allocate_aux_data_RGF(Min))
        torch.cuda.synchronize()
    finally:
        call("negf_rgf_schedule", 0)
    return Mout


@pytest.mark.parametrize("N,b,nE,herm", [(4, 48, 1, False), (5, 64, 2, False), (16, 64, 3, False),
                                         (17, 40, 1, True), (12, 96, 20, False)])
def test_two_sided_vs_oracle_and_one_sided(N, b, nE, herm):
    print(f"two-sided rgf N={N} b={b} batch={nE} hermitian={herm}", flush=True)
    dev = device_hamiltonian(N, b, seed=N + b)
    Es = np.linspace(-0.7, 0.6, nE)
    mats = [assemble_A(dev, E, 1e-3) for E in Es]
    if herm:   # a Hermitian-coupling matrix: A_{i+1,i} = A_{i,i+1}^H, stored as upper only
        mats = [BlockTridiag(A.diag, A.upper, None) for A in mats]
    two = _sweep(mats, 2)
    one = _sweep(mats, 1)
    for k, A in enumerate(mats):
        if herm:
            A = BlockTridiag(A.diag, A.upper, [u.conj().T.copy(order="F") for u in A.upper])
        ref = rgf_ref(A)
        e2 = max_block_err(_band(two.to_host(k)), ref)
        e1 = max_block_err(_band(one.to_host(k)), ref)
        print(f"  instance {k}: two-sided {e2:.1e}, one-sided {e1:.1e}", flush=True)
        assert e2 < 1e-12 and e1 < 1e-12


def test_two_sided_is_the_automatic_schedule_at_small_batch():
    """negf_rgf at one instance of a uniform chain takes the two-sided form
    (its result differs from the one-sided sweep in rounding only)."""
    dev = device_hamiltonian(10, 64, seed=3)
    mats = [assemble_A(dev, 0.2, 1e-3)]
    auto = _sweep(mats, 0).to_host(0)
    two = _sweep(mats, 2).to_host(0)
    one = _sweep(mats, 1).to_host(0)
    assert all(np.array_equal(a, t) for a, t in zip(auto.diag, two.diag))
    assert not all(np.array_equal(a, o) for a, o in zip(auto.diag, one.diag))


def test_two_sided_without_offdiagonal_outputs():
    dev = device_hamiltonian(9, 32, seed=5)
    A = assemble_A(dev, -0.3, 1e-3)
    Min = DeviceBlockTridiag.from_host(A)
    Gd = Min.similar()
    call("negf_rgf_schedule", 2)
    try:
        L = Min.layout
        info = torch.zeros(1, dtype=torch.int32, device="cuda")
        import ctypes as C
        vp = lambda t: C.c_void_p(t.data_ptr())
        s = torch.cuda.current_stream()
        call("negf_rgf", L.N, L.sizes_c(), vp(Min.diag), L.nd, vp(Min.upper), L.no, vp(Min.lower),
             L.no, vp(Gd.diag), None, None, 1, vp(info), C.c_void_p(s.cuda_stream))
        torch.cuda.synchronize()
    finally:
        call("negf_rgf_schedule", 0)
    assert int(info.cpu()[0]) == 0
    ref = rgf_ref(A)
    host = Gd.to_host(0)
    assert max(np.linalg.norm(host.diag[i] - ref[(i, i)]) / np.linalg.norm(ref[(i, i)])
               for i in range(9)) < 1e-12


def test_two_sided_singular_block_names_its_layer():
    """Layer 4 of 6 zero and uncoupled: its Schur complement is zero.  The
    right part (layers 3, 4, 5) meets it at its position 1: info = 4 + 1."""
    A = bm_create_synthetic_random(6, 8, seed=2, uneven=False)
    A.diag[4][:] = 0.0
    for i in (3, 4):
        A.upper[i][:] = 0.0
        A.lower[i][:] = 0.0
    Min = DeviceBlockTridiag.from_host(A)
    call("negf_rgf_schedule", 2)
    try:
        with pytest.raises(api.SingularError) as e:
            api.bndiag_of_inv_rgf_local(Min.similar(), Min)
    finally:
        call("negf_rgf_schedule", 0)
    assert e.value.info == [5]
