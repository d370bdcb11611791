"""CPU oracle, backend op set.

TEST INFRASTRUCTURE ONLY.  This package restates LibNEGF.jl's CPU path so the
GPU library can be checked against it; nothing in the product imports it.

Each ``be_*`` op issues the same LAPACK/BLAS routine as the reference's CPU
backend (``src/backend_cpu.jl``), through scipy's wrappers of OpenBLAS 0.3.29
(the version the reference pins, ``Manifest_cpu.toml:366-369``).  Block
buffers are Fortran-ordered complex128 numpy arrays; an LU holder is a
``CpuLU`` with ``A`` in getrf format and 0-based ``piv``.

Pinning: the op set is pinned by tests/test_oracle_pins.py against numpy's
dense solve/inverse on random inputs and, for the pivot-discarding selected
inverse ops, against a hand-written no-pivot LU on diagonally dominant inputs.
"""
from __future__ import annotations

import numpy as np
from scipy.linalg import lapack, blas

__all__ = [
    "CpuLU", "LapackError", "be_lu", "be_mldivide", "be_gemm", "be_ctranspose",
    "be_identity", "be_zero_array", "be_zero_lu", "be_getrf", "be_getri",
    "be_trsm", "be_inv",
]

_TRANS = {"N": 0, "T": 1, "C": 2}


class LapackError(RuntimeError):
    """getrf reported info != 0 (the reference prints and calls exit(),
    ``src/backend_cpu.jl:151-155``)."""

    def __init__(self, info: int):
        super().__init__(f"LAPACK getrf returned info = {info}")
        self.info = info


class CpuLU:
    """``CpuLU`` (``src/backend_cpu.jl:7-10``): factors in getrf layout plus
    the pivot vector (0-based here, 1-based in the reference)."""

    __slots__ = ("A", "piv")

    def __init__(self, A: np.ndarray, piv: np.ndarray):
        self.A = A
        self.piv = piv


def be_zero_array(shape) -> np.ndarray:
    """``src/backend_cpu.jl:50-52``."""
    return np.zeros(shape, dtype=np.complex128, order="F")


def be_identity(n: int) -> np.ndarray:
    """``src/backend_cpu.jl:54-56``."""
    return np.asfortranarray(np.eye(n, dtype=np.complex128))


def be_zero_lu(n: int) -> CpuLU:
    """``src/backend_cpu.jl:130-136``."""
    return CpuLU(be_zero_array((n, n)), np.zeros(n, dtype=np.int32))


def be_ctranspose(Min: np.ndarray) -> np.ndarray:
    """adjoint! (``src/backend_cpu.jl:58-68``)."""
    return np.asfortranarray(Min.conj().T)


def be_lu(Min: np.ndarray) -> CpuLU:
    """copy + zgetrf with partial pivoting (``src/backend_cpu.jl:147-169``)."""
    lu, piv, info = lapack.zgetrf(np.array(Min, dtype=np.complex128, order="F"))
    if info != 0:
        raise LapackError(int(info))
    return CpuLU(lu, piv)


def be_mldivide(trans: str, Min: np.ndarray, Mlu: CpuLU) -> np.ndarray:
    """copy + zgetrs: returns op(LU)^{-1} Min with op = N or C
    (``src/backend_cpu.jl:193-206``)."""
    x, info = lapack.zgetrs(Mlu.A, Mlu.piv, np.array(Min, dtype=np.complex128, order="F"),
                            trans=_TRANS[trans])
    assert info == 0
    return x


def be_gemm(tA: str, tB: str, alpha, A, B, beta, C) -> np.ndarray:
    """zgemm: alpha op(A) op(B) + beta C (``src/backend_cpu.jl:208-219``).
    Returns the new C (the reference updates C in place)."""
    if C is None:
        m = A.shape[1] if tA != "N" else A.shape[0]
        n = B.shape[0] if tB != "N" else B.shape[1]
        C = be_zero_array((m, n))
        beta = 0.0
    return blas.zgemm(alpha, A, B, beta=beta, c=np.asfortranarray(C), trans_a=_TRANS[tA],
                      trans_b=_TRANS[tB], overwrite_c=0)


def be_getrf(A: np.ndarray) -> np.ndarray:
    """zgetrf whose pivots are *discarded* (``src/backend_cpu.jl:255-265``):
    the returned factors are only meaningful when no row swap happened."""
    lu, piv, info = lapack.zgetrf(np.array(A, dtype=np.complex128, order="F"))
    return lu


def be_getri(A: np.ndarray) -> np.ndarray:
    """zgetri with identity pivots (``src/backend_cpu.jl:229-239``): inverts
    the factors L U stored in A as if no row swap had been made."""
    n = A.shape[0]
    inv, info = lapack.zgetri(np.array(A, dtype=np.complex128, order="F"),
                              np.arange(n, dtype=np.int32))
    return inv


def be_trsm(side: str, ul: str, tA: str, dA: str, alpha, A, B) -> np.ndarray:
    """ztrsm (``src/backend_cpu.jl:241-253``); returns the new B."""
    return blas.ztrsm(alpha, A, np.asfortranarray(B), side=0 if side == "L" else 1,
                      lower=1 if ul == "L" else 0, trans_a=_TRANS[tA],
                      diag=1 if dA == "U" else 0)


def be_inv(M: np.ndarray) -> np.ndarray:
    """Dense inverse, getrf + getri (``src/backend_cpu.jl:143-145``)."""
    lu, piv, info = lapack.zgetrf(np.array(M, dtype=np.complex128, order="F"))
    if info != 0:
        raise LapackError(int(info))
    inv, info = lapack.zgetri(lu, piv)
    return inv
