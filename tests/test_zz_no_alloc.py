"""negf_rgf_noalloc (the allocation-free variant of negf_rgf) against
negf_rgf itself and the oracle, through the C ABI.

Two calls in a row on the same device array, the second on a smaller input
(fewer, uneven blocks and a smaller batch), so stale contents of the first
call cannot leak into the second.  Both entry points run the same kernels in
the same order at these shapes (the in-place inverse schedule below 32
instances), so the outputs must agree bit for bit; each is also held to the
oracle at 1e-12.  (The refusal of a request beyond the array's size is a
host-side check, tested without a GPU in test_golden_and_host.py.)
"""
import ctypes as C

import numpy as np
import pytest

from conftest import max_block_err

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from libnegf_amd import api  # noqa: E402
from libnegf_amd._lib import call  # noqa: E402
from libnegf_amd.btd import DeviceBlockTridiag  # noqa: E402
from libnegf_amd.synthetic import (bm_create_synthetic_random, device_hamiltonian,  # noqa: E402
                                   assemble_A)
from oracle.rgf import bndiag_of_inv_rgf_local as rgf_ref  # noqa: E402


def _band(host):
    G = {}
    for i in range(host.nblocks):
        G[(i, i)] = host.diag[i]
    for i in range(host.nblocks - 1):
        G[(i, i + 1)] = host.upper[i]
        G[(i + 1, i)] = host.lower[i]
    return G


def _run(name, M):
    """One call of `name` (negf_rgf or negf_rgf_noalloc) through the C ABI on
    the packed device matrix M; returns the output and the info array."""
    out = M.similar()
    info = torch.full((M.batch,), -7, dtype=torch.int32, device="cuda")
    sizes = (C.c_int32 * M.nblocks)(*M.block_sizes)
    vp = lambda t: C.c_void_p(t.data_ptr()) if t is not None else None
    s = torch.cuda.current_stream()
    call(name, M.nblocks, sizes, vp(M.diag), M.diag.shape[-1], vp(M.upper), M.upper.shape[-1],
         vp(M.lower), M.lower.shape[-1], vp(out.diag), vp(out.upper), vp(out.lower), M.batch,
         vp(info), C.c_void_p(s.cuda_stream))
    torch.cuda.synchronize()
    return out, info


def _cases():
    dev = device_hamiltonian(12, 64, seed=21)
    first = [assemble_A(dev, E, 1e-3) for E in (-0.6, 0.05, 0.3)]
    second = [bm_create_synthetic_random(5, 40, seed=22)]
    second.append(second[0].copy())
    for d in second[1].diag:
        d += 0.5 * np.eye(d.shape[0])
    return first, second


def test_noalloc_matches_negf_rgf_two_calls():
    first, second = _cases()
    assert len(set(second[0].block_sizes)) > 1
    for k, mats in enumerate((first, second)):
        M = DeviceBlockTridiag.from_host(mats)
        print(f"call {k}: N={M.nblocks} sizes={M.block_sizes[:3]}... batch={M.batch}", flush=True)
        ref, ref_info = _run("negf_rgf", M)
        got, got_info = _run("negf_rgf_noalloc", M)
        assert torch.equal(got_info, ref_info) and int(got_info.abs().sum()) == 0
        for name in ("diag", "upper", "lower"):
            assert torch.equal(getattr(got, name), getattr(ref, name)), name
        for e, A in enumerate(mats):
            err = max_block_err(_band(got.to_host(e)), rgf_ref(A))
            print(f"  instance {e}: max block relerr vs oracle {err:.2e}")
            assert err < 1e-12


def test_noalloc_without_info_and_repeat():
    """info = NULL (the accumulator then lives in the array), then a repeat
    of the same call: identical results."""
    _, second = _cases()
    M = DeviceBlockTridiag.from_host(second)
    outs = []
    for _ in range(2):
        out = M.similar()
        sizes = (C.c_int32 * M.nblocks)(*M.block_sizes)
        s = torch.cuda.current_stream()
        call("negf_rgf_noalloc", M.nblocks, sizes, C.c_void_p(M.diag.data_ptr()),
             M.diag.shape[-1], C.c_void_p(M.upper.data_ptr()), M.upper.shape[-1],
             C.c_void_p(M.lower.data_ptr()), M.lower.shape[-1], C.c_void_p(out.diag.data_ptr()),
             None, None, M.batch, None, C.c_void_p(s.cuda_stream))
        torch.cuda.synchronize()
        outs.append(out.diag.clone())
    assert torch.equal(outs[0], outs[1])
    for e, A in enumerate(second):
        ref = rgf_ref(A)
        host = M.similar()
        host.diag.copy_(outs[0])
        hd = host.to_host(e).diag
        for i in range(M.nblocks):
            d = np.linalg.norm(hd[i] - ref[(i, i)]) / np.linalg.norm(ref[(i, i)])
            assert d < 1e-12

