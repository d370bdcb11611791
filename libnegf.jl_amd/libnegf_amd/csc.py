"""The reference's C entry point on CSC matrices, from Python.

``run_bndiag_wrapper`` mirrors ``src/c_interface/libnegf_c_interface.jl:
85-142``: a scipy CSC block-tridiagonal matrix in, the CSC matrix of the
block band of its inverse out (DD-RGF on the GPU for uniform blocks, RGF
otherwise).  ``method`` selects explicitly (0 auto, 1 RGF, 2 DD-RGF).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import scipy.sparse as sp

from . import _lib
from ._lib import CCscMatrix, SingularError, NegfError

__all__ = ["run_bndiag_wrapper"]


def run_bndiag_wrapper(A: sp.spmatrix, block_sizes, ndiag=None, is_hermitian=False,
                       method: int = 0) -> sp.csc_matrix:
    lib = _lib.load()
    A = sp.csc_matrix(A, dtype=np.complex128)
    A.sort_indices()
    n = A.shape[0]
    colptr = np.ascontiguousarray(A.indptr, dtype=np.int32)
    rowval = np.ascontiguousarray(A.indices, dtype=np.int32)
    nz = np.ascontiguousarray(A.data.view(np.float64))
    bs = np.ascontiguousarray(block_sizes, dtype=np.int32)
    ndiag = ndiag or {"in": 3, "out": 3}
    keys = (C.c_char_p * len(ndiag))(*[k.encode() for k in ndiag])
    vals = (C.c_int32 * len(ndiag))(*ndiag.values())
    out = CCscMatrix()
    p32 = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
    rc = lib.negf_run_bndiag_csc(n, n, A.nnz, p32(colptr), p32(rowval),
                                 nz.ctypes.data_as(C.POINTER(C.c_double)), len(bs), p32(bs),
                                 len(ndiag), keys, vals, int(bool(is_hermitian)), int(method),
                                 C.byref(out), None)
    if rc < 0:
        raise NegfError("negf_run_bndiag_csc", rc, lib.negf_status_string(rc).decode())
    if rc > 0:
        raise SingularError("run_bndiag_wrapper", rc)
    libc = C.CDLL(None)
    try:
        nnz = out.nnz
        cp = np.ctypeslib.as_array(out.colptr, shape=(out.n + 1,)).copy()
        rv = np.ctypeslib.as_array(out.rowval, shape=(max(nnz, 1),))[:nnz].copy()
        dv = np.ctypeslib.as_array(out.nzval, shape=(max(2 * nnz, 2),))[:2 * nnz].copy()
    finally:
        for ptr in (out.colptr, out.rowval, out.nzval):
            libc.free(C.cast(ptr, C.c_void_p))
    return sp.csc_matrix((dv.view(np.complex128), rv, cp), shape=(out.m, out.n))
