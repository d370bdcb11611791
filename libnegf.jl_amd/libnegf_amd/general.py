"""Block n-diagonal RGF on the GPU: natively (``negf_general_rgf``, the
recursion of ``src/matinvertndiag_general_rgf.jl:31-146``) or through the
fused route.

``bm_fuse_to_tridiagonal`` mirrors ``src/blockmatrix.jl:1075-1106``: w
adjacent principal layers become one block, which turns a block
(2w+1)-diagonal matrix into a block-tridiagonal one.  Every original pair
|i - j| <= w then lies inside the fused band, so
``bndiag_of_inv_general_rgf`` (the result of
``src/matinvertndiag_general_rgf.jl:31-146``) is plain RGF on the fused
matrix (negf_rgf) with the original w-band cut out of it.  Blocks are given
as a dict {(i, j): array}, 0-based, |i - j| <= w.
"""
from __future__ import annotations

import numpy as np

from .synthetic import BlockTridiag

__all__ = ["bm_fuse_to_tridiagonal", "bndiag_of_inv_general_rgf"]


def bm_fuse_to_tridiagonal(A: dict, sizes, w: int) -> tuple:
    """(fused block sizes, fused BlockTridiag)."""
    if w <= 1:
        N = len(sizes)
        return list(sizes), BlockTridiag([A[(i, i)] for i in range(N)],
                                         [A[(i, i + 1)] for i in range(N - 1)],
                                         [A[(i + 1, i)] for i in range(N - 1)])
    N = len(sizes)
    newN = -(-N // w)
    fsizes = [int(sum(sizes[I * w:min(N, (I + 1) * w)])) for I in range(newN)]
    off = np.concatenate([[0], np.cumsum(sizes)])
    foff = np.concatenate([[0], np.cumsum(fsizes)])

    def fused(I, J):
        out = np.zeros((fsizes[I], fsizes[J]), dtype=np.complex128, order="F")
        for i in range(I * w, min(N, (I + 1) * w)):
            for j in range(J * w, min(N, (J + 1) * w)):
                if (i, j) in A:
                    out[off[i] - foff[I]:off[i + 1] - foff[I],
                        off[j] - foff[J]:off[j + 1] - foff[J]] = A[(i, j)]
        return out

    diag = [fused(I, I) for I in range(newN)]
    upper = [fused(I, I + 1) for I in range(newN - 1)]
    lower = [fused(I + 1, I) for I in range(newN - 1)]
    return fsizes, BlockTridiag(diag, upper, lower)


def band_layout(sizes, w):
    """Offsets of the band layout of negf_general_rgf: blocks (i, j),
    |i - j| <= w, in row-major order of (i, j)."""
    N = len(sizes)
    off, o = {}, 0
    for i in range(N):
        for j in range(max(0, i - w), min(N, i + w + 1)):
            off[(i, j)] = o
            o += sizes[i] * sizes[j]
    return off, o


def bndiag_of_inv_general_rgf(A, sizes, w: int, device="cuda", method: str = "native"):
    """The w-band of A^{-1} computed on the GPU.  ``A``: a dict {(i, j):
    block} (one instance; the result is a host dict) or a list of such dicts
    (a batch: a list of dicts).  ``method`` "native" runs the general
    recursion (negf_general_rgf), "fused" plain RGF on w-fused layers."""
    if method == "fused":
        if isinstance(A, list):
            return [_fused(a, sizes, w, device) for a in A]
        return _fused(A, sizes, w, device)
    import ctypes as C
    import torch
    from . import api
    from ._lib import call
    mats = A if isinstance(A, list) else [A]
    off, total = band_layout(sizes, w)
    host = np.zeros((len(mats), total), dtype=np.complex128)
    for e, a in enumerate(mats):
        for (i, j), o in off.items():
            host[e, o:o + sizes[i] * sizes[j]] = np.asarray(a[(i, j)]).ravel(order="F")
    dA = torch.from_numpy(host).to(device)
    dG = torch.empty_like(dA)
    info = torch.zeros(len(mats), dtype=torch.int32, device=device)
    sz = (C.c_int32 * len(sizes))(*[int(x) for x in sizes])
    call("negf_general_rgf", len(sizes), sz, int(w), C.c_void_p(dA.data_ptr()), total,
         C.c_void_p(dG.data_ptr()), len(mats), C.c_void_p(info.data_ptr()),
         C.c_void_p(torch.cuda.current_stream().cuda_stream))
    api._check_info("bndiag_of_inv_general_rgf", info)
    hG = dG.cpu().numpy()
    out = []
    for e in range(len(mats)):
        out.append({k: np.asfortranarray(hG[e, o:o + sizes[k[0]] * sizes[k[1]]]
                                         .reshape(sizes[k[1]], sizes[k[0]]).T)
                    for k, o in off.items()})
    return out if isinstance(A, list) else out[0]


def _fused(A: dict, sizes, w: int, device="cuda") -> dict:
    """The w-band of A^{-1} (host dict) through the fused route."""
    from . import api
    from .btd import DeviceBlockTridiag
    fsizes, F = bm_fuse_to_tridiagonal(A, sizes, w)
    Min = DeviceBlockTridiag.from_host(F, device=device)
    Mout = Min.similar()
    api.bndiag_of_inv_rgf_local(Mout, Min)
    H = Mout.to_host()
    N = len(sizes)
    ww = max(w, 1)
    off = np.concatenate([[0], np.cumsum(sizes)])
    foff = np.concatenate([[0], np.cumsum(fsizes)])

    def fblock(I, J):
        if I == J:
            return H.diag[I]
        return H.upper[I] if J == I + 1 else H.lower[J]

    G = {}
    for i in range(N):
        for j in range(max(0, i - w), min(N, i + w + 1)):
            I, J = i // ww, j // ww
            G[(i, j)] = np.asfortranarray(
                fblock(I, J)[off[i] - foff[I]:off[i + 1] - foff[I],
                             off[j] - foff[J]:off[j + 1] - foff[J]])
    return G
