"""CPU oracle: block LU without inter-block pivoting and the Takahashi-style
selected inverse on a block-sparse pattern.

TEST INFRASTRUCTURE ONLY (see oracle/backend.py).

``Block_factorization`` restates ``Block_factorization!``
(``src/selected_inverse.jl:9-55``) and ``Block_inverse`` restates
``Block_inverse!`` (``:139-206``), op for op: be_getrf (pivots discarded),
be_trsm, be_gemm with fill-in allocation, be_getri (identity pivots), on a
dict ``{(i, j): block}`` of the assigned blocks (0-based), as the
reference's ``Matrix{Block}`` with ``isassigned``.

``set_sparse_Block`` builds the pattern of ``src/block.jl:582-687`` from row /
column block indices (1-based as in the reference's test), symmetric when
asked.  ``dense_nopivot_lu`` is the brute-force check of the factorization.

Pinning (tests/test_oracle_selinv.py): the factors against a hand-written
no-pivot dense LU and their product against A (test/test_selected_inverse.jl:
6-22); the selected inverse against numpy's dense inverse on every assigned
block, and diag(A X) = 1 (:23-55), at b = 1 and b > 1 on diagonally dominant
inputs (where partial pivoting makes no swaps and the reference is exact).
"""
from __future__ import annotations

import numpy as np

from .backend import be_getrf, be_getri, be_trsm, be_gemm

__all__ = ["set_sparse_Block", "symbolic_fill", "Block_factorization", "Block_inverse",
           "full", "dense_nopivot_lu"]


def set_sparse_Block(sizes, rind, cind, symmetric=True, rng=None, dominant=True):
    """Random blocks on the pattern (1-based rind/cind), symmetric pattern if
    asked; diagonal blocks get + n I (the reference's test, :13-15)."""
    rng = np.random.default_rng(0) if rng is None else rng
    pat = set()
    for r, c in zip(rind, cind):
        pat.add((r - 1, c - 1))
        if symmetric:
            pat.add((c - 1, r - 1))
    n = sum(sizes)
    A = {}
    for (i, j) in sorted(pat):
        blk = rng.random((sizes[i], sizes[j])) + 1j * rng.random((sizes[i], sizes[j]))
        if i == j and dominant:
            blk = blk + n * np.eye(sizes[i])
        A[(i, j)] = np.asfortranarray(blk)
    return A


def symbolic_fill(pattern, npl):
    """The blocks Block_factorization! creates by fill-in (:41-46)."""
    pat = set(pattern)
    for i in range(npl):
        for k in range(i + 1, npl):
            if (i, k) in pat:
                for l in range(k + 1, npl):
                    if (i, l) in pat and (k, l) not in pat:
                        pat.add((k, l))
                        pat.add((l, k))
    return pat


def Block_factorization(A: dict, npl: int) -> dict:
    A = {k: np.array(v, order="F") for k, v in A.items()}
    for i in range(npl):                                                   # :16
        A[(i, i)] = be_getrf(A[(i, i)])                                    # :18
        for j in range(i + 1, npl):
            if (j, i) in A:
                A[(j, i)] = be_trsm("R", "U", "N", "N", 1.0, A[(i, i)], A[(j, i)])   # :23
            if (i, j) in A:
                A[(i, j)] = be_trsm("L", "L", "N", "U", 1.0, A[(i, i)], A[(i, j)])   # :27
        for k in range(i + 1, npl):                                        # :32
            if (i, k) in A:
                A[(k, k)] = be_gemm("N", "N", -1.0, A[(k, i)], A[(i, k)], 1.0, A[(k, k)])  # :34
                for l in range(k + 1, npl):
                    if (i, l) in A:
                        if (k, l) in A:                                     # :37-40
                            A[(k, l)] = be_gemm("N", "N", -1.0, A[(k, i)], A[(i, l)], 1.0, A[(k, l)])
                            A[(l, k)] = be_gemm("N", "N", -1.0, A[(l, i)], A[(i, k)], 1.0, A[(l, k)])
                        else:                                               # :41-46 fill-in
                            A[(k, l)] = be_gemm("N", "N", -1.0, A[(k, i)], A[(i, l)], 0.0, None)
                            A[(l, k)] = be_gemm("N", "N", -1.0, A[(l, i)], A[(i, k)], 0.0, None)
    return A


def Block_inverse(A: dict, npl: int) -> dict:
    A = {k: np.array(v, order="F") for k, v in A.items()}
    b = A[(0, 0)].shape[0]
    A[(npl - 1, npl - 1)] = be_getri(A[(npl - 1, npl - 1)])               # :158
    for i in range(npl - 2, -1, -1):                                        # :160
        U, Lt, vidx = [], [], []
        for j in range(i + 1, npl):
            if (i, j) in A:
                u = np.zeros((b, b), complex, order="F")
                lt = np.zeros((b, b), complex, order="F")
                for k in range(i + 1, npl):
                    if (i, k) in A and (k, j) in A:
                        u = be_gemm("N", "N", 1.0, A[(i, k)], A[(k, j)], 1.0, u)     # :165
                        lt = be_gemm("N", "N", 1.0, A[(j, k)], A[(k, i)], 1.0, lt)   # :167
                u = be_trsm("L", "U", "N", "N", -1.0, A[(i, i)], u)                 # :172
                lt = be_trsm("R", "L", "N", "U", -1.0, A[(i, i)], lt)               # :174
                U.append(u)
                Lt.append(lt)
                vidx.append(j)
        Aii = np.zeros((b, b), complex, order="F")
        for vj, j in enumerate(vidx):                                        # :180
            if (j, i) in A:
                Aii = be_gemm("N", "N", 1.0, U[vj], A[(j, i)], 1.0, Aii)     # :183
                A[(j, i)] = Lt[vj]
                A[(i, j)] = U[vj]
        Aii = be_trsm("R", "L", "N", "U", 1.0, A[(i, i)], Aii)               # :195
        A[(i, i)] = be_getri(A[(i, i)]) - Aii                                # :197-199
    return A


def full(A: dict, sizes) -> np.ndarray:
    """Dense matrix of the assigned blocks (``src/block.jl:519-569``)."""
    off = np.concatenate([[0], np.cumsum(sizes)])
    M = np.zeros((off[-1], off[-1]), complex)
    for (i, j), blk in A.items():
        M[off[i]:off[i + 1], off[j]:off[j + 1]] = blk
    return M


def dense_nopivot_lu(M: np.ndarray) -> np.ndarray:
    """Brute-force Doolittle LU without pivoting (packed: unit L below, U on
    and above the diagonal), the ``lu(full(T), NoPivot())`` of the test."""
    A = np.array(M, dtype=complex)
    n = A.shape[0]
    for k in range(n):
        A[k + 1:, k] /= A[k, k]
        A[k + 1:, k + 1:] -= np.outer(A[k + 1:, k], A[k, k + 1:])
    return A
