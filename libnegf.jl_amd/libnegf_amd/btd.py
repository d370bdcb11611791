"""Device-resident block-tridiagonal matrices in the packed layout of the C ABI.

``DeviceBlockTridiag`` is the GPU counterpart of the reference's
``BlockMatrix`` with ``ndiag = 3`` (``src/blockmatrix.jl:10-17``): the N
diagonal blocks, the N-1 upper blocks A_{i,i+1} and (unless Hermitian
storage) the N-1 lower blocks A_{i+1,i}, each column-major, packed into three
complex128 device arrays of shape (batch, nd) / (batch, no).  Torch only
provides the device memory; all arithmetic goes through libnegf_hip.so.
"""
from __future__ import annotations

import numpy as np
import torch

from .synthetic import BlockTridiag

__all__ = ["Layout", "DeviceBlockTridiag"]


class Layout:
    """Offsets of the packed layout (element units)."""

    def __init__(self, block_sizes):
        self.b = [int(x) for x in block_sizes]
        self.N = len(self.b)
        self.od = np.concatenate([[0], np.cumsum([x * x for x in self.b])]).astype(np.int64)
        self.oo = np.concatenate([[0], np.cumsum([self.b[i] * self.b[i + 1]
                                                  for i in range(self.N - 1)])]).astype(np.int64)
        self.nd = int(self.od[-1])
        self.no = int(self.oo[-1])

    def sizes_c(self):
        import ctypes
        return (ctypes.c_int32 * self.N)(*self.b)


def _seg_to_block(seg: torch.Tensor, m: int, n: int) -> torch.Tensor:
    """A flat column-major segment as an (m, n) tensor view (strides (1, m))."""
    return seg.view(n, m).transpose(0, 1)


class DeviceBlockTridiag:
    def __init__(self, block_sizes, batch: int = 1, hermitian: bool = False, device="cuda",
                 zero: bool = True):
        self.layout = Layout(block_sizes)
        self.batch = int(batch)
        self.hermitian = bool(hermitian)
        alloc = torch.zeros if zero else torch.empty
        L = self.layout
        self.diag = alloc((self.batch, L.nd), dtype=torch.complex128, device=device)
        self.upper = alloc((self.batch, max(L.no, 1)), dtype=torch.complex128, device=device)
        self.lower = None if hermitian else alloc((self.batch, max(L.no, 1)),
                                                  dtype=torch.complex128, device=device)

    # ---- shape helpers
    @property
    def block_sizes(self):
        return self.layout.b

    @property
    def nblocks(self):
        return self.layout.N

    def similar(self, hermitian=None, zero=True) -> "DeviceBlockTridiag":
        """``bm_similar`` (``src/blockmatrix.jl:307-319``) on the device."""
        h = self.hermitian if hermitian is None else hermitian
        return DeviceBlockTridiag(self.layout.b, self.batch, h, self.diag.device, zero)

    def block(self, kind: str, i: int, inst: int = 0) -> torch.Tensor:
        """View of block ``kind`` in {'diag', 'upper', 'lower'} number i."""
        L = self.layout
        if kind == "diag":
            m = n = L.b[i]
            return _seg_to_block(self.diag[inst, L.od[i]:L.od[i + 1]], m, n)
        arr = self.upper if kind == "upper" else self.lower
        if arr is None:
            raise ValueError("Hermitian storage holds no lower blocks")
        m, n = (L.b[i], L.b[i + 1]) if kind == "upper" else (L.b[i + 1], L.b[i])
        return _seg_to_block(arr[inst, L.oo[i]:L.oo[i + 1]], m, n)

    # ---- host <-> device (be_copy_to_hw / be_copy_from_hw, src/backend_cpu.jl:26-40)
    @staticmethod
    def from_host(mats, device="cuda") -> "DeviceBlockTridiag":
        if isinstance(mats, BlockTridiag):
            mats = [mats]
        m0 = mats[0]
        out = DeviceBlockTridiag(m0.block_sizes, len(mats), m0.is_hermitian, device, zero=False)
        L = out.layout
        hd = np.empty((len(mats), L.nd), dtype=np.complex128)
        hu = np.zeros((len(mats), max(L.no, 1)), dtype=np.complex128)
        hl = None if m0.is_hermitian else np.zeros((len(mats), max(L.no, 1)), dtype=np.complex128)
        for k, m in enumerate(mats):
            if m.block_sizes != m0.block_sizes or m.is_hermitian != m0.is_hermitian:
                raise ValueError("all instances of a batch need the same block structure")
            for i in range(L.N):
                hd[k, L.od[i]:L.od[i + 1]] = np.asarray(m.diag[i]).ravel(order="F")
            for i in range(L.N - 1):
                hu[k, L.oo[i]:L.oo[i + 1]] = np.asarray(m.upper[i]).ravel(order="F")
                if hl is not None:
                    hl[k, L.oo[i]:L.oo[i + 1]] = np.asarray(m.lower[i]).ravel(order="F")
        out.diag.copy_(torch.from_numpy(hd))
        out.upper.copy_(torch.from_numpy(hu))
        if hl is not None:
            out.lower.copy_(torch.from_numpy(hl))
        return out

    def to_host(self, inst: int = 0) -> BlockTridiag:
        L = self.layout
        hd = self.diag[inst].cpu().numpy()
        hu = self.upper[inst].cpu().numpy()
        hl = None if self.lower is None else self.lower[inst].cpu().numpy()

        def blk(arr, o0, o1, m, n):
            return np.asfortranarray(arr[o0:o1].reshape(n, m).T)

        diag = [blk(hd, L.od[i], L.od[i + 1], L.b[i], L.b[i]) for i in range(L.N)]
        upper = [blk(hu, L.oo[i], L.oo[i + 1], L.b[i], L.b[i + 1]) for i in range(L.N - 1)]
        lower = None if hl is None else [blk(hl, L.oo[i], L.oo[i + 1], L.b[i + 1], L.b[i])
                                         for i in range(L.N - 1)]
        return BlockTridiag(diag, upper, lower)
