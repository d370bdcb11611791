"""The reference's RGF sweep issued op by op through the backend binding.

``bndiag_of_inv_rgf_local_perop`` is what LibNEGF.jl's unchanged
``bndiag_of_inv_rgf_local!`` (``src/matinvertndiag_rgf.jl:114-200``)
executes when ``backend_amd.jl`` is the selected backend: one be_lu /
be_mldivide / be_gemm / be_ctranspose / be_copy_in_hw call per block
operation, in the reference's order and with its buffers (LU(S_i), t_i,
l_{i+1}), each call a launch sequence in libnegf_hip.so, and the binding's
info check (a host read of the LU's info, where the reference exits on a zero
pivot) after every LU.  It is the per-op path a Julia user gets without the
sweep-level entry points; it is timed beside ``negf_rgf`` (one call per
sweep) and tested as a sequence at production block size.
"""
from __future__ import annotations

import torch

from . import api
from .btd import DeviceBlockTridiag

__all__ = ["AuxDataRGFPerOp", "bndiag_of_inv_rgf_local_perop"]


class AuxDataRGFPerOp:
    """The reference's ``AuxDataRGF`` buffers (``src/matinvertndiag_rgf.jl:
    7-57``) for the per-op sweep: ``buffM1``'s LU factors per layer, t_i and
    l_{i+1}, plus the identity and adjoint scratch the calls use."""

    def __init__(self, M: DeviceBlockTridiag):
        if M.batch != 1:
            raise ValueError("the per-op sweep takes one instance (the reference's BlockMatrix)")
        b = M.block_sizes
        N = len(b)
        dev = M.diag.device
        self.lu = [api.GpuLU(api.be_zero_array((b[i], b[i]), dev),
                             torch.zeros(b[i], dtype=torch.int32, device=dev)) for i in range(N)]
        self.t = [api.be_zero_array((b[i], b[i + 1]), dev) for i in range(N - 1)]
        self.l = [None] + [api.be_zero_array((b[i + 1], b[i]), dev) for i in range(N - 1)]
        self.tmp = [api.be_zero_array((b[i + 1], b[i]), dev) for i in range(N - 1)]
        self.tmp2 = [api.be_zero_array((b[i + 1], b[i]), dev) for i in range(N - 1)]
        self.S = [api.be_zero_array((b[i], b[i]), dev) for i in range(N)]
        self.eye = {n: api.be_identity(n, dev) for n in set(b)}


def bndiag_of_inv_rgf_local_perop(Mout: DeviceBlockTridiag, Min: DeviceBlockTridiag,
                                  aux: AuxDataRGFPerOp | None = None) -> DeviceBlockTridiag:
    """band(Min^{-1}) into Mout by the reference's call sequence, op by op."""
    aux = AuxDataRGFPerOp(Min) if aux is None else aux
    N = Min.nblocks
    A = lambda kind, i: Min.block(kind, i)
    G = lambda kind, i: Mout.block(kind, i)
    # upward (UDL) pass, :143-162
    api.be_lu(aux.lu[N - 1], A("diag", N - 1))                                     # :144
    for i in range(N - 2, -1, -1):                                                 # :147
        api.be_ctranspose(aux.tmp[i], A("upper", i))                               # :152
        api.be_mldivide("C", aux.tmp2[i], aux.tmp[i], aux.lu[i + 1])               # :153
        api.be_ctranspose(aux.t[i], aux.tmp2[i])                                   # :154
        api.be_mldivide("N", aux.l[i + 1], A("lower", i), aux.lu[i + 1])           # :157
        api.be_copy_in_hw(aux.S[i], A("diag", i))                                  # :159
        api.be_gemm("N", "N", -1.0, A("upper", i), aux.l[i + 1], 1.0, aux.S[i])    # :160
        api.be_lu(aux.lu[i], aux.S[i])                                             # :161
    # downward pass, :165-182
    b = Min.block_sizes
    api.be_mldivide("N", G("diag", 0), aux.eye[b[0]], aux.lu[0])                   # :169
    for i in range(1, N):                                                          # :172
        api.be_gemm("N", "N", -1.0, G("diag", i - 1), aux.t[i - 1], 0.0, G("upper", i - 1))   # :174
        api.be_gemm("N", "N", -1.0, aux.l[i], G("diag", i - 1), 0.0, G("lower", i - 1))       # :177
        api.be_mldivide("N", G("diag", i), aux.eye[b[i]], aux.lu[i])                         # :180
        api.be_gemm("N", "N", -1.0, aux.l[i], G("upper", i - 1), 1.0, G("diag", i))          # :181
    return Mout
