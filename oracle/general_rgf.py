"""CPU oracle: RGF for a block n-diagonal matrix and the fused route.

TEST INFRASTRUCTURE ONLY (see oracle/backend.py).

``bndiag_of_inv_general_rgf`` restates ``bndiag_of_inv_general_rgf!``
(``src/matinvertndiag_general_rgf.jl:31-146``) with bandwidth w =
(ndiag - 1) / 2: the upward pass eliminates layer i+1 from the w layers
above it (multipliers ``:58-69``, w x w Schur update ``:72-89``), the
downward pass fills the w-band (``:100-128``).  Two defects of the reference
are not reproduced: it reads a ``nrsType`` field that ``BlockMatrix`` does
not have (``:13, 17, 33-35``; ``src/blockmatrix.jl:10-17``), so it cannot
run as written, and it writes the Schur updates back into ``Min``
(``:86``); here the input is copied first.

``bm_fuse_to_tridiagonal`` restates ``src/blockmatrix.jl:1075-1106``: w
adjacent layers fused into one block, through the dense matrix.

Blocks: dict ``{(i, j): array}`` for |i - j| <= w, 0-based.

Pinning (tests/test_oracle_general.py): the w-band of numpy's dense
inverse; general == fused (the reference's own check,
``test/test_matinvertndiag_general_rgf_fused.jl:39-90``); w = 1 equals
plain RGF.
"""
from __future__ import annotations

import numpy as np

from .backend import be_lu, be_mldivide, be_gemm, be_ctranspose, be_identity

__all__ = ["random_ndiag", "ndiag_to_dense", "bndiag_of_inv_general_rgf",
           "bm_fuse_to_tridiagonal", "band_of_dense"]


def random_ndiag(sizes, w, seed=0, dominant=True):
    rng = np.random.default_rng(seed)
    N = len(sizes)
    A = {}
    for i in range(N):
        for j in range(max(0, i - w), min(N, i + w + 1)):
            blk = rng.random((sizes[i], sizes[j])) + 1j * rng.random((sizes[i], sizes[j]))
            if i == j and dominant:
                blk = blk + 2 * w * max(sizes) * np.eye(sizes[i])
            A[(i, j)] = np.asfortranarray(blk if i == j else 0.3 * blk)
    return A


def ndiag_to_dense(A, sizes):
    off = np.concatenate([[0], np.cumsum(sizes)])
    M = np.zeros((off[-1], off[-1]), complex)
    for (i, j), blk in A.items():
        M[off[i]:off[i + 1], off[j]:off[j + 1]] = blk
    return M


def band_of_dense(M, sizes, w):
    off = np.concatenate([[0], np.cumsum(sizes)])
    N = len(sizes)
    return {(i, j): np.asfortranarray(M[off[i]:off[i + 1], off[j]:off[j + 1]])
            for i in range(N) for j in range(max(0, i - w), min(N, i + w + 1))}


def bndiag_of_inv_general_rgf(Ain: dict, sizes, w: int) -> dict:
    N = len(sizes)
    Min = {k: np.array(v, order="F") for k, v in Ain.items()}
    buf = {}                                  # buffM1: LU on the diagonal, multipliers off it
    buf[(N - 1, N - 1)] = be_lu(Min[(N - 1, N - 1)])                                  # :51
    for ix in range(N - 2, -1, -1):                                                  # :54
        kmax = min(w, ix + 1)
        for k in range(1, kmax + 1):                                                  # :58
            t = ix - k + 1
            tmp = be_ctranspose(Min[(t, ix + 1)])                                     # :63
            tmp2 = be_mldivide("C", tmp, buf[(ix + 1, ix + 1)])                       # :64
            buf[(t, ix + 1)] = be_ctranspose(tmp2)                                    # :65
            buf[(ix + 1, t)] = be_mldivide("N", Min[(ix + 1, t)], buf[(ix + 1, ix + 1)])  # :68
        for i in range(1, kmax + 1):                                                  # :72
            for j in range(1, kmax + 1):
                r, c = ix - i + 1, ix - j + 1
                upd = be_gemm("N", "N", -1.0, Min[(r, ix + 1)], buf[(ix + 1, c)], 1.0,
                              Min[(r, c)])                                            # :77-80
                if r == c == ix:
                    buf[(ix, ix)] = be_lu(upd)                                        # :84
                else:
                    Min[(r, c)] = upd                                                 # :86
    G = {}
    G[(0, 0)] = be_mldivide("N", be_identity(sizes[0]), buf[(0, 0)])                  # :97
    for ix in range(1, N):                                                            # :100
        kmax = min(w, ix)
        for k in range(1, kmax + 1):                                                  # :103
            tp = ix - k
            up = np.zeros((sizes[tp], sizes[ix]), complex, order="F")
            lo = np.zeros((sizes[ix], sizes[tp]), complex, order="F")
            for k2 in range(1, kmax + 1):                                             # :111
                m = ix - k2
                up = be_gemm("N", "N", -1.0, G[(tp, m)], buf[(m, ix)], 1.0, up)       # :115
                lo = be_gemm("N", "N", -1.0, buf[(ix, m)], G[(m, tp)], 1.0, lo)       # :118
            G[(tp, ix)] = up
            G[(ix, tp)] = lo
        gii = be_mldivide("N", be_identity(sizes[ix]), buf[(ix, ix)])                 # :123
        for k2 in range(1, kmax + 1):
            m = ix - k2
            gii = be_gemm("N", "N", -1.0, buf[(ix, m)], G[(m, ix)], 1.0, gii)         # :126
        G[(ix, ix)] = gii
    return G


def bm_fuse_to_tridiagonal(A: dict, sizes, w: int):
    """Fused sizes and the fused block-tridiagonal blocks (diag, upper, lower)."""
    N = len(sizes)
    newN = -(-N // w)
    fsizes = [sum(sizes[I * w:min(N, (I + 1) * w)]) for I in range(newN)]
    M = ndiag_to_dense(A, sizes)
    off = np.concatenate([[0], np.cumsum(fsizes)])
    diag = [np.asfortranarray(M[off[I]:off[I + 1], off[I]:off[I + 1]]) for I in range(newN)]
    upper = [np.asfortranarray(M[off[I]:off[I + 1], off[I + 1]:off[I + 2]]) for I in range(newN - 1)]
    lower = [np.asfortranarray(M[off[I + 1]:off[I + 2], off[I]:off[I + 1]]) for I in range(newN - 1)]
    return fsizes, diag, upper, lower
