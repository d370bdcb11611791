"""CPU oracle: plain block-tridiagonal RGF, its full-inverse extension and the
dense ground truth.

TEST INFRASTRUCTURE ONLY (see oracle/backend.py).

``bndiag_of_inv_rgf_local`` restates ``src/matinvertndiag_rgf.jl:114-200``
call for call: the same be_lu / be_mldivide / be_gemm / be_ctranspose
sequence, the same buffers (``buffM1`` holds LU(S_i) on the diagonal, t_i at
[i, i+1] and l_{i+1} at [i+1, i]; ``Mout`` doubles as scratch).

Input: any object with ``diag[i]``, ``upper[i]`` and ``low(i)`` (A_{i+1,i}),
e.g. ``BlockTridiag``.  Output: a dict-of-blocks ``{(i, j): block}`` holding
the band of A^{-1} (and, with ``full_inverse``, every block), 0-based.

Pinning (tests/test_oracle_pins.py): band of a dense numpy inverse; the b = 1
uniform chain against its closed-form Green's function; block-diagonal A
(G_ii = A_ii^{-1}, zero off-diagonals); N = 1 (G = A^{-1}).
"""
from __future__ import annotations

import numpy as np

from .backend import (be_lu, be_mldivide, be_gemm, be_ctranspose, be_identity,
                      be_inv)

__all__ = ["rgf_upward", "bndiag_of_inv_rgf_local", "bndiag_of_inv_direct",
           "band_blocks", "first_row_chain"]


def rgf_upward(A):
    """The upward (UDL) pass, ``src/matinvertndiag_rgf.jl:143-162``.

    Returns ``buffM1`` as a dict: ('lu', i) -> CpuLU(S_i), ('t', i) -> t_i =
    A_{i,i+1} S_{i+1}^{-1}, ('l', i+1) -> l_{i+1} = S_{i+1}^{-1} A_{i+1,i}.
    """
    N = len(A.diag)
    buf = {}
    buf[("lu", N - 1)] = be_lu(A.diag[N - 1])                                  # :144
    for i in range(N - 2, -1, -1):                                            # :147
        tmp = be_ctranspose(A.upper[i])                                       # :152
        tmp2 = be_mldivide("C", tmp, buf[("lu", i + 1)])                      # :153
        buf[("t", i)] = be_ctranspose(tmp2)                                   # :154
        buf[("l", i + 1)] = be_mldivide("N", A.low(i), buf[("lu", i + 1)])    # :157
        S = np.array(A.diag[i], order="F")                                    # :159
        S = be_gemm("N", "N", -1.0, A.upper[i], buf[("l", i + 1)], 1.0, S)    # :160
        buf[("lu", i)] = be_lu(S)                                             # :161
    return buf


def bndiag_of_inv_rgf_local(A, full_inverse: bool = False, upward=None):
    """band(A^{-1}) by RGF (``src/matinvertndiag_rgf.jl:114-200``).

    ``full_inverse`` mirrors ``auxData.buildFullInv`` (``:185-199``).
    """
    N = len(A.diag)
    buf = rgf_upward(A) if upward is None else upward
    G = {}
    G[(0, 0)] = be_mldivide("N", be_identity(A.diag[0].shape[0]), buf[("lu", 0)])  # :169
    for i in range(1, N):                                                            # :172
        G[(i - 1, i)] = be_gemm("N", "N", -1.0, G[(i - 1, i - 1)], buf[("t", i - 1)], 0.0, None)   # :174
        G[(i, i - 1)] = be_gemm("N", "N", -1.0, buf[("l", i)], G[(i - 1, i - 1)], 0.0, None)       # :177
        Gii = be_mldivide("N", be_identity(A.diag[i].shape[0]), buf[("lu", i)])                   # :180
        G[(i, i)] = be_gemm("N", "N", -1.0, buf[("l", i)], G[(i - 1, i)], 1.0, Gii)                 # :181
    if full_inverse and N > 2:                                                        # :185-199
        for i in range(N):
            for j in range(i + 2, N):
                G[(i, j)] = be_gemm("N", "N", -1.0, G[(i, j - 1)], buf[("t", j - 1)], 0.0, None)
                G[(j, i)] = be_gemm("N", "N", -1.0, buf[("l", j)], G[(j - 1, i)], 0.0, None)
    return G


def first_row_chain(A, G, buf):
    """G_{0,j} for j = 0..N-1 by the full-inverse row recursion
    G_{0,j} = -G_{0,j-1} t_{j-1} (``src/matinvertndiag_rgf.jl:191``), which
    the transmission T = Tr(Gamma_L G_{1N} Gamma_R G_{1N}^H) needs."""
    N = len(A.diag)
    row = [G[(0, 0)]]
    for j in range(1, N):
        if (0, j) in G:
            row.append(G[(0, j)])
        else:
            row.append(be_gemm("N", "N", -1.0, row[-1], buf[("t", j - 1)], 0.0, None))
    return row


def band_blocks(Ginv: np.ndarray, sizes, ndiag: int = 3):
    """Cut the block band out of a dense matrix, like ``bndiag_of_inv_direct!``
    (``src/matinvertndiag_direct.jl:17-61``)."""
    off = np.concatenate([[0], np.cumsum(sizes)])
    w = (ndiag - 1) // 2
    N = len(sizes)
    out = {}
    for i in range(N):
        for j in range(max(0, i - w), min(N, i + w + 1)):
            out[(i, j)] = np.asfortranarray(Ginv[off[i]:off[i + 1], off[j]:off[j + 1]])
    return out


def bndiag_of_inv_direct(A, ndiag: int = 3):
    """Dense ground truth: be_inv of the assembled matrix, then the band
    (``src/matinvertndiag_direct.jl:17-61``, ``:79-86``)."""
    return band_blocks(be_inv(A.to_dense()), A.block_sizes, ndiag)
