"""ctypes binding of libnegf_hip.so (the C ABI in include/negf_hip.h).

The library is the only compute path: when it is missing, or no GPU is
visible, every entry point raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os
import pathlib

_HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = pathlib.Path(os.environ.get("LIBNEGF_HIP_LIB", _HERE / "libnegf_hip.so"))

c_int32_p = C.POINTER(C.c_int32)
c_double_p = C.POINTER(C.c_double)
vp = C.c_void_p
i32 = C.c_int32
i64 = C.c_int64
ch = C.c_char
d = C.c_double

class CCscMatrix(C.Structure):
    """Mirror of CCscMatrix (include/negf_hip.h, the reference's
    src/c_interface/libnegf_c_interface.h:12-19)."""

    _fields_ = [("m", C.c_int32), ("n", C.c_int32), ("nnz", C.c_int32),
                ("colptr", C.POINTER(C.c_int32)), ("rowval", C.POINTER(C.c_int32)),
                ("nzval", C.POINTER(C.c_double))]


# name -> (restype, argtypes); mirrors include/negf_hip.h
class NegfBandJob(C.Structure):
    """include/negf_hip.h NegfBandJob (one device of negf_bndiag_devices)."""
    _fields_ = [("nblocks", C.c_int32), ("block_sizes", C.POINTER(C.c_int32)),
                ("diag", C.c_void_p), ("diag_bstride", C.c_int64),
                ("upper", C.c_void_p), ("upper_bstride", C.c_int64),
                ("lower", C.c_void_p), ("lower_bstride", C.c_int64),
                ("g_diag", C.c_void_p), ("g_upper", C.c_void_p), ("g_lower", C.c_void_p),
                ("batch", C.c_int32), ("block_sizes_d1", C.POINTER(C.c_int32)),
                ("nlevels", C.c_int32), ("info", C.c_void_p)]


SIGNATURES = {
    "run_bndiag_wrapper": (CCscMatrix, [i32, i32, i32, c_int32_p, c_int32_p, c_double_p, i32,
                                        c_int32_p, i32, C.POINTER(C.c_char_p), c_int32_p, i32]),
    "negf_run_bndiag_csc": (C.c_int, [i32, i32, i32, c_int32_p, c_int32_p, c_double_p, i32,
                                      c_int32_p, i32, C.POINTER(C.c_char_p), c_int32_p, i32, i32,
                                      C.POINTER(CCscMatrix), vp]),
    "negf_version": (C.c_int, []),
    "negf_status_string": (C.c_char_p, [C.c_int]),
    "negf_workspace_release": (C.c_int, []),
    "negf_gemm_variant": (C.c_int, [i32]),
    "negf_inv_variant": (C.c_int, [i32]),
    "negf_rgf_schedule": (C.c_int, [i32]),
    "negf_zinv_schedule": (C.c_int, [i32, i32, i32, c_int32_p]),
    "negf_zinv_workspace": (C.c_int64, [i32, i32]),
    "negf_zgemm": (C.c_int, [ch, ch, i32, i32, i32, c_double_p, vp, i64, vp, i64, c_double_p, vp,
                             i64, vp]),
    "negf_zgemm_strided_batched": (C.c_int, [ch, ch, i32, i32, i32, c_double_p, vp, i64, i64, vp,
                                             i64, i64, c_double_p, vp, i64, i64, i32, vp]),
    "negf_zgemm_prec_strided_batched": (C.c_int, [ch, ch, i32, i32, i32, c_double_p, vp, i64, i64,
                                                  vp, i64, i64, c_double_p, vp, i64, i64, i32, i32,
                                                  i32, vp]),
    "negf_zgetrf": (C.c_int, [i32, vp, i64, vp, vp, vp]),
    "negf_zgetrf_strided_batched": (C.c_int, [i32, vp, i64, i64, vp, i64, vp, i32, vp]),
    "negf_zgetrs": (C.c_int, [ch, i32, i32, vp, i64, vp, vp, i64, vp]),
    "negf_zgetrs_strided_batched": (C.c_int, [ch, i32, i32, vp, i64, i64, vp, i64, vp, i64, i64,
                                              i32, vp]),
    "negf_ztrsm": (C.c_int, [ch, ch, ch, ch, i32, i32, c_double_p, vp, i64, vp, i64, vp]),
    "negf_ztrsm_strided_batched": (C.c_int, [ch, ch, ch, ch, i32, i32, c_double_p, vp, i64, i64,
                                             vp, i64, i64, i32, vp]),
    "negf_zgetri": (C.c_int, [i32, vp, i64, vp, vp, vp]),
    "negf_zgetri_strided_batched": (C.c_int, [i32, vp, i64, i64, vp, i64, vp, i32, vp]),
    "negf_zinv_strided_batched": (C.c_int, [i32, vp, i64, i64, vp, i64, i64, i32, vp, vp]),
    "negf_zadjoint_strided_batched": (C.c_int, [i32, i32, vp, i64, i64, vp, i64, i64, i32, vp]),
    "negf_zcopy_strided_batched": (C.c_int, [i32, i32, vp, i64, i64, vp, i64, i64, i32, vp]),
    "negf_zfill_strided_batched": (C.c_int, [i32, i32, c_double_p, c_double_p, vp, i64, i64, i32,
                                             vp]),
    "negf_rgf": (C.c_int, [i32, c_int32_p, vp, i64, vp, i64, vp, i64, vp, vp, vp, i32, vp, vp]),
    "negf_general_rgf": (C.c_int, [i32, c_int32_p, i32, vp, i64, vp, i32, vp, vp]),
    "negf_general_rgf_band_elements": (C.c_int64, [i32, c_int32_p, i32]),
    "negf_rgf_full": (C.c_int, [i32, c_int32_p, vp, i64, vp, i64, vp, i64, vp, i32, vp, vp]),
    "negf_rgf_noalloc": (C.c_int, [i32, c_int32_p, vp, i64, vp, i64, vp, i64, vp, vp, vp, i32, vp,
                                   vp]),
    "negf_rgf_rkd": (C.c_int, [i32, c_int32_p, vp, i64, vp, i64, vp, i64, vp, i64, vp, i64, vp, vp,
                               vp, vp, vp, vp, i64, vp, i64, vp, i32, vp, vp]),
    "negf_rgf_keldysh_energies": (C.c_int, [i32, i32, vp, vp, vp, vp, vp, i32, d, d, d, d, vp, vp,
                                            vp, vp, i32, i32, vp]),
    "negf_rgf_keldysh_energies_workspace": (C.c_int64, [i32, i32, i32, i32, i32]),
    "negf_rgf_keldysh_energies_mixed": (C.c_int, [i32, i32, vp, vp, vp, vp, vp, i32, d, d, d, d, vp,
                                                  vp, vp, vp, i32, i32, i32, i32, d, vp, vp, vp,
                                                  vp]),
    "negf_rgf_keldysh_energies_mixed_workspace": (C.c_int64, [i32, i32, i32, i32, i32, i32, i32]),
    "negf_rgf_keldysh_energies_graph_create": (C.c_int, [i32, i32, vp, vp, vp, vp, vp, i32, d, d, d,
                                                         d, vp, vp, vp, vp, i32, i32,
                                                         C.POINTER(vp)]),
    "negf_rgf_keldysh_energies_graph_create_prec": (C.c_int, [i32, i32, vp, vp, vp, vp, vp, i32, d,
                                                              d, d, d, vp, vp, vp, vp, i32, i32,
                                                              i32, C.POINTER(vp)]),
    "negf_energies_graph_launch": (C.c_int, [vp, vp]),
    "negf_energies_graph_estimates": (C.c_int, [vp, vp, vp]),
    "negf_energies_graph_destroy": (C.c_int, [vp]),
    "negf_ddrgf": (C.c_int, [i32, i32, vp, i64, vp, i64, vp, i64, vp, vp, vp, i32, c_int32_p, i32,
                             vp, vp]),
    "negf_ddrgf_workspace": (C.c_int64, [i32, i32, i32, c_int32_p, i32]),
    "negf_ddrgf_blocksizes": (C.c_int, [i32, c_int32_p, vp, i64, vp, i64, vp, i64, vp, vp, vp, i32,
                                        c_int32_p, i32, vp, vp]),
    "negf_ddrgf_blocksizes_workspace": (C.c_int64, [i32, c_int32_p, i32, c_int32_p, i32]),
    "negf_bndiag_devices": (C.c_int, [i32, vp, vp]),
    "negf_bndiag_devices_prec": (C.c_int, [i32, vp, i32, vp]),
    "negf_rgf_rkd_mixed": (C.c_int, [i32, c_int32_p, vp, i64, vp, i64, vp, i64, vp, i64, vp, i64,
                                     vp, vp, vp, vp, vp, vp, i64, vp, i64, vp, i32, vp, i32, i32,
                                     C.c_double, vp, vp, vp, vp]),
    "negf_tune_set": (C.c_int, [i32, c_int32_p, i32, c_int32_p, i32]),
    "negf_tune_lookup": (C.c_int, [i32, c_int32_p, i32, c_int32_p, i32]),
    "negf_tune_clear": (C.c_int, []),
    "negf_tune_count": (C.c_int, []),
    "negf_tune_load": (C.c_int, [C.c_char_p]),
    "negf_block_fill_pattern": (C.c_int, [i32, i32, c_int32_p, c_int32_p, i32, c_int32_p,
                                          c_int32_p]),
    "negf_block_factorization": (C.c_int, [i32, i32, i32, c_int32_p, c_int32_p,
                                           C.POINTER(vp), i64, i64, i32, i32, vp, vp]),
    "negf_block_inverse": (C.c_int, [i32, i32, i32, c_int32_p, c_int32_p, C.POINTER(vp), i64,
                                     i64, i32, vp]),
    "negf_block_factorization_pivoted": (C.c_int, [i32, i32, i32, c_int32_p, c_int32_p,
                                                   C.POINTER(vp), i64, i64, i32, i32, vp, vp, vp]),
    "negf_block_inverse_pivoted": (C.c_int, [i32, i32, i32, c_int32_p, c_int32_p, C.POINTER(vp),
                                             i64, i64, i32, vp, vp]),
    "negf_build_M_from_HS": (C.c_int, [i32, c_int32_p, vp, i32] + [vp, i64] * 9 + [vp, vp, vp, vp]),
}

_lib = None


class NegfError(RuntimeError):
    """A negative status from the library (bad argument, HIP error, ...)."""

    def __init__(self, fn: str, status: int, msg: str):
        super().__init__(f"{fn} failed with status {status}: {msg}")
        self.status = status


class SingularError(RuntimeError):
    """A zero pivot / singular Schur complement (LAPACK info > 0).  The
    reference prints and exits there (src/backend_cpu.jl:151-155)."""

    def __init__(self, where: str, info):
        super().__init__(f"{where}: singular pivot, info = {info}")
        self.info = info


def load(path: pathlib.Path | None = None) -> C.CDLL:
    """Load the shared library (once) and declare every signature."""
    global _lib
    if _lib is not None:
        return _lib
    p = pathlib.Path(path) if path else LIB_PATH
    if not p.exists():
        raise FileNotFoundError(
            f"{p} is not built; run `python libnegf.jl_amd/build_lib.py` (no fallback path exists)")
    lib = C.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    # the measured choices of the tuner (libnegf_amd/tune.py), when a table exists
    from . import tune
    tune.load_table()
    return lib


def call(name: str, *args) -> int:
    lib = load()
    rc = getattr(lib, name)(*args)
    if name.endswith("_workspace") or name == "run_bndiag_wrapper":
        return rc
    if rc < 0:
        raise NegfError(name, rc, lib.negf_status_string(rc).decode())
    return rc


def cplx(z) -> C.Array:
    z = complex(z)
    return (C.c_double * 2)(z.real, z.imag)
