"""Selected inverse on a block-sparse pattern, on the GPU.

Mirror of ``src/selected_inverse.jl`` and the ``Block`` container of
``src/block.jl``: ``BlockSparse`` holds the assigned b x b blocks of a
block-sparse matrix as column-major complex128 device tensors (``T[i, j]``,
0-based, like ``isassigned(A, i, j)``); ``Block_factorization`` and
``Block_inverse`` return new containers (the reference's copying wrappers,
``:124-128, :222-226``); the ``!`` forms work in place.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from ._lib import call, SingularError

__all__ = ["BlockSparse", "set_sparse_Block", "Block_factorization", "Block_factorization_",
           "Block_inverse", "Block_inverse_", "full", "fill_pattern"]


def fill_pattern(npl: int, pattern) -> list:
    """The pattern plus the factorization's fill-in (negf_block_fill_pattern)."""
    pat = list(pattern)
    br = (C.c_int32 * len(pat))(*[p[0] for p in pat])
    bc = (C.c_int32 * len(pat))(*[p[1] for p in pat])
    n = call("negf_block_fill_pattern", npl, len(pat), br, bc, 0, None, None)
    orow = (C.c_int32 * n)()
    ocol = (C.c_int32 * n)()
    call("negf_block_fill_pattern", npl, len(pat), br, bc, n, orow, ocol)
    return [(orow[k], ocol[k]) for k in range(n)]


class BlockSparse:
    """Assigned blocks of an npl x npl block matrix with uniform block size b,
    all instances of a batch in one (batch, nnzb, b, b) device tensor."""

    def __init__(self, npl: int, b: int, pattern, batch: int = 1, device="cuda"):
        self.npl, self.b, self.batch = int(npl), int(b), int(batch)
        self.pattern = sorted(set(pattern))
        self.index = {p: k for k, p in enumerate(self.pattern)}
        # (batch, nnzb, b*b): block k of instance m is column-major at [m, k]
        self.data = torch.zeros((self.batch, len(self.pattern), b * b), dtype=torch.complex128,
                                device=device)
        # pivots of a "partial" factorization ([batch, npl * b], 1-based rows)
        self.ipiv = None

    def isassigned(self, i, j) -> bool:
        return (i, j) in self.index

    def block(self, i, j, inst=0) -> torch.Tensor:
        return self.data[inst, self.index[(i, j)]].view(self.b, self.b).transpose(0, 1)

    def copy(self) -> "BlockSparse":
        out = BlockSparse(self.npl, self.b, self.pattern, self.batch, self.data.device)
        out.data.copy_(self.data)
        ipiv = getattr(self, "ipiv", None)
        out.ipiv = None if ipiv is None else ipiv.clone()
        return out

    def with_pattern(self, pattern) -> "BlockSparse":
        out = BlockSparse(self.npl, self.b, pattern, self.batch, self.data.device)
        for p, k in self.index.items():
            out.data[:, out.index[p]] = self.data[:, k]
        return out

    def _args(self):
        n = len(self.pattern)
        br = (C.c_int32 * n)(*[p[0] for p in self.pattern])
        bc = (C.c_int32 * n)(*[p[1] for p in self.pattern])
        base = self.data.data_ptr()
        step = self.b * self.b * 16
        ptrs = (C.c_void_p * n)(*[base + k * step for k in range(n)])
        return n, br, bc, ptrs, self.b, len(self.pattern) * self.b * self.b

    @staticmethod
    def from_host(blocks: dict, npl: int, device="cuda") -> "BlockSparse":
        b = next(iter(blocks.values())).shape[0]
        out = BlockSparse(npl, b, blocks.keys(), 1, device)
        host = np.zeros((1, len(out.pattern), b * b), dtype=np.complex128)
        for p, k in out.index.items():
            host[0, k] = np.asarray(blocks[p]).ravel(order="F")
        out.data.copy_(torch.from_numpy(host))
        return out

    def to_host(self, inst=0) -> dict:
        h = self.data[inst].cpu().numpy()
        return {p: np.asfortranarray(h[k].reshape(self.b, self.b).T) for p, k in self.index.items()}


def set_sparse_Block(bind, rind, cind, symmetric=True, seed=0, dominant=True, device="cuda"):
    """Random blocks on a pattern given by 1-based block row / column indices
    (``set_sparse_Block``, ``src/block.jl:582-687``); the diagonal gets + n I
    as in ``test/test_selected_inverse.jl:13-15``."""
    rng = np.random.default_rng(seed)
    npl = len(bind)
    b = int(bind[0])
    if any(int(x) != b for x in bind):
        raise ValueError("the selected inverse assumes uniform block sizes (:150-153)")
    pat = set()
    for r, c in zip(rind, cind):
        pat.add((r - 1, c - 1))
        if symmetric:
            pat.add((c - 1, r - 1))
    n = b * npl
    blocks = {}
    for (i, j) in sorted(pat):
        blk = rng.random((b, b)) + 1j * rng.random((b, b))
        if i == j and dominant:
            blk = blk + n * np.eye(b)
        blocks[(i, j)] = np.asfortranarray(blk)
    return BlockSparse.from_host(blocks, npl, device)


PIVOTING = {"none": 0, "reference": 1, "partial": 2}


def Block_factorization_(T: BlockSparse, ref_pivoting: bool = False,
                         pivoting: str | None = None) -> BlockSparse:
    """In place.  The pattern must already hold the fill-in (use
    ``Block_factorization``, which adds it).

    ``pivoting``: "partial" (default) pivots inside every diagonal block and
    applies the permutation to its U row blocks, keeping the pivots on the
    container for ``Block_inverse``; "none" factors without row swaps;
    "reference" reproduces the reference (partial pivoting whose pivots are
    discarded, ``src/backend_cpu.jl:255-265``: right only without swaps).
    ``ref_pivoting = True`` is "reference"."""
    if pivoting is None:
        pivoting = "reference" if ref_pivoting else "partial"
    mode = PIVOTING[pivoting]
    n, br, bc, ptrs, b, stride = T._args()
    info = torch.zeros(T.batch, dtype=torch.int32, device=T.data.device)
    T.ipiv = torch.zeros((T.batch, T.npl * b), dtype=torch.int32,
                         device=T.data.device) if mode == 2 else None
    call("negf_block_factorization_pivoted", T.npl, b, n, br, bc, ptrs, b, stride, T.batch, mode,
         C.c_void_p(T.ipiv.data_ptr()) if T.ipiv is not None else None,
         C.c_void_p(info.data_ptr()), C.c_void_p(torch.cuda.current_stream().cuda_stream))
    bad = info.cpu()
    if bool((bad != 0).any()):
        raise SingularError("Block_factorization", bad.tolist())
    return T


def Block_factorization(T: BlockSparse, ref_pivoting: bool = False,
                        pivoting: str | None = None) -> BlockSparse:
    """Copy with the fill-in blocks added, then factor
    (``src/selected_inverse.jl:124-128``)."""
    F = T.with_pattern(fill_pattern(T.npl, T.pattern))
    return Block_factorization_(F, ref_pivoting, pivoting)


def Block_inverse_(T: BlockSparse) -> BlockSparse:
    """In place; uses the pivots a "partial" factorization kept on T."""
    n, br, bc, ptrs, b, stride = T._args()
    ipiv = getattr(T, "ipiv", None)
    call("negf_block_inverse_pivoted", T.npl, b, n, br, bc, ptrs, b, stride, T.batch,
         C.c_void_p(ipiv.data_ptr()) if ipiv is not None else None,
         C.c_void_p(torch.cuda.current_stream().cuda_stream))
    return T


def Block_inverse(T: BlockSparse) -> BlockSparse:
    """``src/selected_inverse.jl:222-226``."""
    return Block_inverse_(T.copy())


def full(T: BlockSparse, inst=0) -> np.ndarray:
    """Dense host matrix of the assigned blocks (``src/block.jl:519-569``)."""
    b = T.b
    M = np.zeros((T.npl * b, T.npl * b), dtype=np.complex128)
    for (i, j), blk in T.to_host(inst).items():
        M[i * b:(i + 1) * b, j * b:(j + 1) * b] = blk
    return M
