"""Host-side mirror of LibNEGF.jl's interface for the NEGF hot path.

Names, argument meaning and error behaviour follow the reference so the
parity tests read like its own tests; every call lands in libnegf_hip.so.

Sweeps (``src/matinvertndiag_rgf.jl``, ``src/keldyshndiag.jl``):
  allocate_aux_data_RGF, bndiag_of_inv_rgf_local, allocate_aux_data_Keldysh,
  keldyshndiag, rgf_keldysh_transmission, rgf_keldysh_energies.
Backend ops (``src/backend_cpu.jl``) on column-major complex128 device
tensors: be_gemm, be_lu, be_mldivide, be_getrf, be_getri, be_trsm,
be_ctranspose, be_copy_in_hw, be_fill, be_identity, be_zero_array, be_inv.
"""
from __future__ import annotations

import ctypes as C
import dataclasses

import numpy as np
import torch

from . import _lib
from ._lib import call, cplx, SingularError
from .btd import DeviceBlockTridiag, Layout

__all__ = [
    "AuxDataRGF", "AuxDataKeldysh", "allocate_aux_data_RGF", "allocate_aux_data_Keldysh",
    "bndiag_of_inv_rgf_local", "keldyshndiag", "rgf_keldysh_transmission",
    "rgf_keldysh_energies", "GpuLU", "be_gemm", "be_lu", "be_mldivide", "be_getrf", "be_getri",
    "be_trsm", "be_ctranspose", "be_copy_in_hw", "be_fill", "be_identity", "be_zero_array",
    "be_inv", "col_major", "release_workspace", "build_M_from_HS", "EnergiesGraph",
    "bndiag_of_inv_direct", "bndiag_of_inv_rgf_local_mixed",
]


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("libnegf_amd needs a visible GPU; there is no CPU fallback")
    _lib.load()


def release_workspace():
    call("negf_workspace_release")


def _check_info(where, info: torch.Tensor):
    bad = info.cpu()
    if bool((bad < 0).any()):
        # the energy job's code for a contact broadening that is not
        # Hermitian positive semidefinite (its G^< uses Gamma = F F^H)
        raise ValueError(f"{where}: a contact broadening Gamma is not Hermitian positive "
                         f"semidefinite (info = {bad.tolist()})")
    if bool((bad != 0).any()):
        raise SingularError(where, bad.tolist())


# ----------------------------------------------------------------------------- sweeps
@dataclasses.dataclass
class AuxDataRGF:
    """``AuxDataRGF`` (``src/matinvertndiag_rgf.jl:10-14``).  The device
    sweep keeps its buffers in the library's cached workspace; the holder
    carries the per-instance info array, the ``buildFullInv`` flag and, when
    set, ``fullInv``: every block of the inverse as a dense column-major
    matrix per instance, shape (batch, n, n) indexed [instance, column, row]."""

    info: torch.Tensor
    buildFullInv: bool = False
    fullInv: torch.Tensor | None = None


def allocate_aux_data_RGF(M: DeviceBlockTridiag) -> AuxDataRGF:
    """``src/matinvertndiag_rgf.jl:80-97``."""
    return AuxDataRGF(torch.zeros(M.batch, dtype=torch.int32, device=M.diag.device))


def bndiag_of_inv_rgf_local(Mout: DeviceBlockTridiag, Min: DeviceBlockTridiag,
                            auxData: AuxDataRGF | None = None):
    """band(Min^{-1}) into Mout (``src/matinvertndiag_rgf.jl:114-200``).

    Min may hold a batch of instances of one block structure; Mout must match.
    Raises SingularError where the reference would print and exit()."""
    _require_gpu()
    if Mout.layout.b != Min.layout.b or Mout.batch != Min.batch or Mout.lower is None:
        raise ValueError("Mout must be a non-Hermitian BlockMatrix shaped like Min")
    L = Min.layout
    info = auxData.info if auxData is not None else torch.zeros(Min.batch, dtype=torch.int32,
                                                                 device=Min.diag.device)
    if auxData is not None and auxData.buildFullInv:
        # every block (:185-199) into auxData.fullInv, the band copied into Mout
        n = sum(L.b)
        full = torch.empty((Min.batch, n, n), dtype=torch.complex128, device=Min.diag.device)
        call("negf_rgf_full", L.N, L.sizes_c(), _ptr(Min.diag), L.nd, _ptr(Min.upper),
             max(L.no, 1), _ptr(Min.lower) if Min.lower is not None else None, max(L.no, 1),
             _ptr(full), Min.batch, _ptr(info), _stream())
        _check_info("bndiag_of_inv_rgf_local (buildFullInv)", info)
        off = [0]
        for x in L.b:
            off.append(off[-1] + x)

        def blk(i, j):   # column-major flat block (i, j) of every instance
            return full[:, off[j]:off[j + 1], off[i]:off[i + 1]].reshape(Min.batch, -1)

        for i in range(L.N):
            Mout.diag[:, L.od[i]:L.od[i + 1]] = blk(i, i)
        for i in range(L.N - 1):
            Mout.upper[:, L.oo[i]:L.oo[i + 1]] = blk(i, i + 1)
            Mout.lower[:, L.oo[i]:L.oo[i + 1]] = blk(i + 1, i)
        auxData.fullInv = full
        return Mout
    call("negf_rgf", L.N, L.sizes_c(), _ptr(Min.diag), L.nd, _ptr(Min.upper), max(L.no, 1),
         _ptr(Min.lower) if Min.lower is not None else None, max(L.no, 1),
         _ptr(Mout.diag), _ptr(Mout.upper), _ptr(Mout.lower), Min.batch, _ptr(info), _stream())
    _check_info("bndiag_of_inv_rgf_local", info)
    return Mout


def bndiag_of_inv_rgf_local_mixed(Mout: DeviceBlockTridiag, Min: DeviceBlockTridiag,
                                  precision: str = "emulated", tol: float = 1e-8,
                                  n_moduli: int = 0, S: DeviceBlockTridiag | None = None,
                                  Gn: DeviceBlockTridiag | None = None):
    """``bndiag_of_inv_rgf_local`` (and, with S and Gn, ``keldyshndiag``'s
    band of Min^{-1} S Min^{-H} in one pass) with the block products in FP32
    ("fp32") or FP64 emulated on the int8 matrix cores ("emulated"); the
    diagonal inverses stay FP64.  Instances whose error estimate exceeds
    ``tol`` are computed again in FP64.  Returns (est, fallback) per
    instance (numpy)."""
    _require_gpu()
    if Mout.layout.b != Min.layout.b or Mout.batch != Min.batch or Mout.lower is None:
        raise ValueError("Mout must be a non-Hermitian BlockMatrix shaped like Min")
    L = Min.layout
    info = torch.zeros(Min.batch, dtype=torch.int32, device=Min.diag.device)
    est = np.zeros(Min.batch, dtype=np.float64)
    fb = np.zeros(Min.batch, dtype=np.int32)
    nfb = C.c_int32(0)
    if (S is None) != (Gn is None):
        raise ValueError("S and Gn go together")
    if S is not None and not S.hermitian:
        raise ValueError("the central operator S must use Hermitian storage")
    call("negf_rgf_rkd_mixed", L.N, L.sizes_c(), _ptr(Min.diag), L.nd, _ptr(Min.upper),
         max(L.no, 1), _ptr(Min.lower) if Min.lower is not None else None, max(L.no, 1),
         _ptr(S.diag) if S is not None else None, (L.nd if S is not None and S.batch > 1 else 0),
         _ptr(S.upper) if S is not None else None,
         (max(L.no, 1) if S is not None and S.batch > 1 else 0),
         _ptr(Mout.diag), _ptr(Mout.upper), _ptr(Mout.lower),
         _ptr(Gn.diag) if Gn is not None else None, _ptr(Gn.upper) if Gn is not None else None,
         None, 0, None, 0, None, Min.batch, _ptr(info), PRECISIONS[precision], int(n_moduli),
         float(tol), est.ctypes.data_as(C.c_void_p), fb.ctypes.data_as(C.c_void_p),
         C.byref(nfb), _stream())
    _check_info("bndiag_of_inv_rgf_local_mixed", info)
    return est, fb


def _full_storage(M: DeviceBlockTridiag) -> DeviceBlockTridiag:
    """Non-Hermitian storage of M (lower blocks materialised as adjoints of the
    upper ones when M is in Hermitian storage)."""
    if M.lower is not None:
        return M
    out = DeviceBlockTridiag(M.layout.b, M.batch, hermitian=False, device=M.diag.device, zero=False)
    out.diag.copy_(M.diag)
    out.upper.copy_(M.upper)
    L = M.layout
    for i in range(L.N - 1):
        seg_u, seg_l = slice(int(L.oo[i]), int(L.oo[i + 1])), slice(int(L.oo[i]), int(L.oo[i + 1]))
        call("negf_zadjoint_strided_batched", L.b[i], L.b[i + 1], _ptr(M.upper[:, seg_u]), L.b[i],
             L.no, _ptr(out.lower[:, seg_l]), L.b[i + 1], L.no, M.batch, _stream())
    return out


def build_M_from_HS(H: DeviceBlockTridiag, S: DeviceBlockTridiag, Se: DeviceBlockTridiag | None,
                    energies) -> DeviceBlockTridiag:
    """M(E) = E S - H - Se for every energy (``build_M_from_HS``,
    ``src/matloader.jl:118-122``), on the device: a batch of len(energies)
    instances.  H, S, Se hold one instance (shared by all energies) or one per
    energy; energies are real (E) or complex (E + i eta)."""
    _require_gpu()
    z = torch.as_tensor(energies, dtype=torch.complex128, device=H.diag.device).reshape(-1)
    nE = int(z.numel())
    L = H.layout
    for X in (S, Se):
        if X is not None and X.layout.b != L.b:
            raise ValueError("H, S and Se must share the block structure")
    ins = []
    for X in (H, S, Se):
        if X is None:
            ins.append(None)
            continue
        if X.batch not in (1, nE):
            raise ValueError("inputs hold one instance or one per energy")
        ins.append(_full_storage(X))
    Hf, Sf, Ef = ins
    bs = lambda X, per: 0 if X.batch == 1 else per
    M = DeviceBlockTridiag(L.b, nE, hermitian=False, device=H.diag.device, zero=False)
    no = max(L.no, 1)
    args = []
    for X in (Hf, Sf, Ef):
        if X is None:
            args += [None, 0, None, 0, None, 0]
        else:
            args += [_ptr(X.diag), bs(X, L.nd), _ptr(X.upper), bs(X, no), _ptr(X.lower), bs(X, no)]
    call("negf_build_M_from_HS", L.N, L.sizes_c(), _ptr(z), nE, *args, _ptr(M.diag),
         _ptr(M.upper), _ptr(M.lower), _stream())
    return M


@dataclasses.dataclass
class AuxDataKeldysh:
    """``AuxDataKeldysh`` (``src/keldyshndiag.jl:7-12``): ``buffS`` receives
    the result (diagonal and upper blocks of G^n, Hermitian storage)."""

    auxDataRGF: AuxDataRGF
    buffS: DeviceBlockTridiag


def allocate_aux_data_Keldysh(M: DeviceBlockTridiag, S: DeviceBlockTridiag) -> AuxDataKeldysh:
    """``src/keldyshndiag.jl:79-107``."""
    return AuxDataKeldysh(allocate_aux_data_RGF(M), S.similar(hermitian=True))


def keldyshndiag(M: DeviceBlockTridiag, S: DeviceBlockTridiag, auxData: AuxDataKeldysh,
                 Gr: DeviceBlockTridiag | None = None):
    """band(M^{-1} S M^{-H}) into ``auxData.buffS`` (``src/keldyshndiag.jl:120-131``,
    ``keldyshndiag_v3!`` ``:166-318``).  S is Hermitian block tridiagonal in
    Hermitian storage.  With ``Gr`` given, the band of M^{-1} comes out of the
    same shared upward pass (what the reference gets from a separate
    ``bndiag_of_inv_rgf_local!`` call)."""
    _require_gpu()
    if not S.hermitian:
        raise ValueError("the central operator S must use Hermitian storage")
    L = M.layout
    out = auxData.buffS
    info = auxData.auxDataRGF.info
    call("negf_rgf_rkd", L.N, L.sizes_c(), _ptr(M.diag), L.nd, _ptr(M.upper), max(L.no, 1),
         _ptr(M.lower) if M.lower is not None else None, max(L.no, 1),
         _ptr(S.diag), L.nd if S.batch > 1 else 0, _ptr(S.upper), max(L.no, 1) if S.batch > 1 else 0,
         _ptr(Gr.diag) if Gr is not None else None, _ptr(Gr.upper) if Gr is not None else None,
         _ptr(Gr.lower) if Gr is not None else None,
         _ptr(out.diag), _ptr(out.upper), None, 0, None, 0, None, M.batch, _ptr(info), _stream())
    _check_info("keldyshndiag", info)
    return out


def rgf_keldysh_transmission(M: DeviceBlockTridiag, S: DeviceBlockTridiag | None,
                             gamma_L: torch.Tensor, gamma_R: torch.Tensor,
                             Gr: DeviceBlockTridiag | None = None,
                             Gn: DeviceBlockTridiag | None = None):
    """One shared upward pass, then any of: band(G^R) into Gr, band(G^n)
    into Gn, and T = Re Tr(Gamma_L G_{1N} Gamma_R G_{1N}^H) (returned, one per
    instance).  gamma_L / gamma_R: (b_0, b_0) / (b_{N-1}, b_{N-1}) column-major
    device tensors, shared by all instances."""
    _require_gpu()
    L = M.layout
    info = torch.zeros(M.batch, dtype=torch.int32, device=M.diag.device)
    T = torch.zeros(M.batch, dtype=torch.float64, device=M.diag.device)
    gl, gr = col_major(gamma_L), col_major(gamma_R)
    call("negf_rgf_rkd", L.N, L.sizes_c(), _ptr(M.diag), L.nd, _ptr(M.upper), max(L.no, 1),
         _ptr(M.lower) if M.lower is not None else None, max(L.no, 1),
         _ptr(S.diag) if S is not None else None, (L.nd if S is not None and S.batch > 1 else 0),
         _ptr(S.upper) if S is not None else None,
         (max(L.no, 1) if S is not None and S.batch > 1 else 0),
         _ptr(Gr.diag) if Gr is not None else None, _ptr(Gr.upper) if Gr is not None else None,
         _ptr(Gr.lower) if Gr is not None else None,
         _ptr(Gn.diag) if Gn is not None else None, _ptr(Gn.upper) if Gn is not None else None,
         _ptr(gl), 0, _ptr(gr), 0, _ptr(T), M.batch, _ptr(info), _stream())
    _check_info("rgf_keldysh_transmission", info)
    return T


def _pack_blocks(blocks, device):
    flat = np.concatenate([np.asarray(b).ravel(order="F") for b in blocks]) if blocks else \
        np.zeros(1, dtype=np.complex128)
    return torch.from_numpy(flat.astype(np.complex128)).to(device)


class DeviceHamiltonianGPU:
    """Energy-independent inputs of the energy-batched workload, resident in
    HBM: H diagonal blocks, H_{i,i+1}, Gamma_L, Gamma_R (column-major)."""

    def __init__(self, dev, device="cuda"):
        self.N = dev.nblocks
        self.b = dev.b
        self.H_diag = _pack_blocks(dev.H_diag, device)
        self.H_upper = _pack_blocks(dev.H_upper, device)
        self.gamma_L = _pack_blocks([dev.gamma_L], device)
        self.gamma_R = _pack_blocks([dev.gamma_R], device)


def _check_energy_job(where, H, energies, GR_diag, Gless_diag, transmission):
    """The energy job's tensors as the C ABI reads them: energies a contiguous
    float64 vector, G^R / G^< complex128 and T float64, contiguous, with room
    for every energy, all on a GPU (checked before any device call: a float32
    energy vector would otherwise be read as half as many doubles, past its
    end)."""
    def need(name, t, dtype, count):
        if t is None:
            return
        if not isinstance(t, torch.Tensor) or t.dtype != dtype:
            raise ValueError(f"{where}: {name} must be a {dtype} tensor")
        if not t.is_contiguous() or t.numel() < count:
            raise ValueError(f"{where}: {name} must be contiguous with at least {count} elements")
    if energies is None:
        raise ValueError(f"{where}: energies must be a torch.float64 tensor")
    need("energies", energies, torch.float64, 0)
    nE = int(energies.numel())
    per = H.N * H.b * H.b
    need("GR_diag", GR_diag, torch.complex128, nE * per)
    need("Gless_diag", Gless_diag, torch.complex128, nE * per)
    need("transmission", transmission, torch.float64, nE)
    for name, t in (("energies", energies), ("GR_diag", GR_diag), ("Gless_diag", Gless_diag),
                    ("transmission", transmission)):
        if t is not None and t.device.type != "cuda":
            raise ValueError(f"{where}: {name} must be on the GPU")


def rgf_keldysh_energies(H: DeviceHamiltonianGPU, energies: torch.Tensor, eta: float,
                         mu_L: float, mu_R: float, kT: float, GR_diag: torch.Tensor | None = None,
                         Gless_diag: torch.Tensor | None = None,
                         transmission: torch.Tensor | None = None, wave: int = 0,
                         streams: int = 0, check: bool = True):
    """G^R and G^< diagonal blocks and T(E) for every energy (device float64
    tensor).  GR_diag / Gless_diag: complex128 device tensors of shape
    (nE, N*b*b) (packed diag layout per energy) or None."""
    _check_energy_job("rgf_keldysh_energies", H, energies, GR_diag, Gless_diag, transmission)
    _require_gpu()
    nE = int(energies.numel())
    info = torch.zeros(nE, dtype=torch.int32, device=energies.device)
    call("negf_rgf_keldysh_energies", H.N, H.b, _ptr(H.H_diag), _ptr(H.H_upper),
         _ptr(H.gamma_L), _ptr(H.gamma_R), _ptr(energies), nE, float(eta), float(mu_L),
         float(mu_R), float(kT), _ptr(GR_diag), _ptr(Gless_diag), _ptr(transmission), _ptr(info),
         int(wave), int(streams), _stream())
    if check:
        _check_info("rgf_keldysh_energies", info)
    return info


PRECISIONS = {"fp64": 0, "fp32": 1, "emulated": 2, "fp64_3m": 3}


def rgf_keldysh_energies_mixed(H: DeviceHamiltonianGPU, energies: torch.Tensor, eta: float,
                               mu_L: float, mu_R: float, kT: float, precision: str = "emulated",
                               tol: float = 1e-8, n_moduli: int = 0,
                               GR_diag: torch.Tensor | None = None,
                               Gless_diag: torch.Tensor | None = None,
                               transmission: torch.Tensor | None = None, wave: int = 0,
                               streams: int = 0, check: bool = True):
    """``rgf_keldysh_energies`` with the complex products in FP32 ("fp32"),
    FP64 emulated on the int8 matrix cores ("emulated") or FP64 with three
    real products per complex product ("fp64_3m"); the diagonal-block
    inverses stay FP64.  Energies whose error estimate exceeds ``tol`` are
    computed again in FP64.  Returns (info, est, fallback): the device info
    array, the per-energy estimates and re-run flags (numpy)."""
    if precision not in PRECISIONS:
        raise ValueError(f"rgf_keldysh_energies_mixed: precision {precision!r}")
    _check_energy_job("rgf_keldysh_energies_mixed", H, energies, GR_diag, Gless_diag,
                      transmission)
    _require_gpu()
    nE = int(energies.numel())
    info = torch.zeros(nE, dtype=torch.int32, device=energies.device)
    est = np.zeros(nE, dtype=np.float64)
    fb = np.zeros(nE, dtype=np.int32)
    nfb = C.c_int32(0)
    call("negf_rgf_keldysh_energies_mixed", H.N, H.b, _ptr(H.H_diag), _ptr(H.H_upper),
         _ptr(H.gamma_L), _ptr(H.gamma_R), _ptr(energies), nE, float(eta), float(mu_L),
         float(mu_R), float(kT), _ptr(GR_diag), _ptr(Gless_diag), _ptr(transmission), _ptr(info),
         int(wave), int(streams), PRECISIONS[precision], int(n_moduli), float(tol),
         est.ctypes.data_as(C.c_void_p), fb.ctypes.data_as(C.c_void_p), C.byref(nfb), _stream())
    if check:
        _check_info("rgf_keldysh_energies_mixed", info)
    return info, est, fb


class EnergiesGraph:
    """The energy job of ``rgf_keldysh_energies`` captured once into a HIP
    graph (negf_rgf_keldysh_energies_graph_create[_prec]) and replayed by
    ``launch()`` as one submission.  The tensors passed here are bound for the
    graph's lifetime: change energies / H / Gamma by writing into them in
    place between launches; eta, mu and kT are fixed.  ``precision`` "fp64"
    (four real products per complex product) or "fp64_3m" (three, Gauss; each
    launch also writes the per-energy estimates, read by ``estimates()``)."""

    def __init__(self, H: DeviceHamiltonianGPU, energies: torch.Tensor, eta: float, mu_L: float,
                 mu_R: float, kT: float, GR_diag: torch.Tensor | None = None,
                 Gless_diag: torch.Tensor | None = None,
                 transmission: torch.Tensor | None = None, wave: int = 0, streams: int = 0,
                 precision: str = "fp64"):
        if precision not in ("fp64", "fp64_3m"):
            raise ValueError(f"EnergiesGraph: precision {precision!r} (fp64 or fp64_3m)")
        _check_energy_job("EnergiesGraph", H, energies, GR_diag, Gless_diag, transmission)
        _require_gpu()
        nE = int(energies.numel())
        self.precision = precision
        self.info = torch.zeros(nE, dtype=torch.int32, device=energies.device)
        # keep every bound tensor alive as long as the graph
        self._bound = (H, energies, GR_diag, Gless_diag, transmission)
        h = C.c_void_p()
        call("negf_rgf_keldysh_energies_graph_create_prec", H.N, H.b, _ptr(H.H_diag),
             _ptr(H.H_upper), _ptr(H.gamma_L), _ptr(H.gamma_R), _ptr(energies), nE, float(eta),
             float(mu_L), float(mu_R), float(kT), _ptr(GR_diag), _ptr(Gless_diag),
             _ptr(transmission), _ptr(self.info), int(wave), int(streams),
             PRECISIONS[precision], C.byref(h))
        self._h = h
        self._nE = nE

    def estimates(self) -> np.ndarray:
        """The last launch's per-energy error estimates (three-product form)."""
        if self._h is None:
            raise RuntimeError("graph destroyed")
        est = np.zeros(self._nE, dtype=np.float64)
        call("negf_energies_graph_estimates", self._h, est.ctypes.data_as(C.c_void_p), _stream())
        return est

    def launch(self, check: bool = True):
        if self._h is None:
            raise RuntimeError("graph destroyed")
        call("negf_energies_graph_launch", self._h, _stream())
        if check:
            _check_info("rgf_keldysh_energies (graph)", self.info)
        return self.info

    def close(self):
        if self._h is not None:
            call("negf_energies_graph_destroy", self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def energies_workspace_bytes(N: int, b: int, nE: int, wave: int, streams: int = 1) -> int:
    return int(call("negf_rgf_keldysh_energies_workspace", N, b, nE, wave, streams))


def energies_mixed_workspace_bytes(N: int, b: int, nE: int, wave: int, streams: int = 1,
                                   precision: str = "fp64_3m", n_moduli: int = 0) -> int:
    """Device workspace of ``rgf_keldysh_energies_mixed`` for a wave."""
    return int(call("negf_rgf_keldysh_energies_mixed_workspace", N, b, nE, wave, streams,
                    PRECISIONS[precision], int(n_moduli)))


# ----------------------------------------------------------------------------- per-op
def col_major(x: torch.Tensor) -> torch.Tensor:
    """Column-major (Fortran) copy of a 2-D tensor unless it already is one."""
    if x.dim() != 2:
        raise ValueError("expected a matrix")
    if x.stride(0) == 1 and x.stride(1) >= max(1, x.shape[0]):
        return x
    return x.transpose(0, 1).contiguous().transpose(0, 1)


def _ld(x: torch.Tensor) -> int:
    if x.stride(0) != 1 and x.shape[0] > 1:
        raise ValueError("matrices must be column-major (use col_major())")
    if x.shape[1] <= 1:
        # one column: torch keeps any stride for it (often 1); every ld >= m
        # addresses the same elements
        return max(int(x.shape[0]), 1)
    return max(int(x.stride(1)), 1)


def be_zero_array(shape, device="cuda") -> torch.Tensor:
    """``src/backend_cpu.jl:50-52``."""
    return col_major(torch.zeros(shape, dtype=torch.complex128, device=device))


def be_identity(n: int, device="cuda") -> torch.Tensor:
    """``src/backend_cpu.jl:54-56``."""
    out = be_zero_array((n, n), device)
    call("negf_zfill_strided_batched", n, n, cplx(0), cplx(1), _ptr(out), _ld(out), 0, 1,
         _stream())
    return out


def be_fill(M: torch.Tensor, x) -> torch.Tensor:
    """``src/backend_cpu.jl:86-88``."""
    call("negf_zfill_strided_batched", M.shape[0], M.shape[1], cplx(x), cplx(x), _ptr(M), _ld(M),
         0, 1, _stream())
    return M


def be_copy_in_hw(Mout: torch.Tensor, Min: torch.Tensor) -> torch.Tensor:
    """``src/backend_cpu.jl:42-44``."""
    call("negf_zcopy_strided_batched", Min.shape[0], Min.shape[1], _ptr(Min), _ld(Min), 0,
         _ptr(Mout), _ld(Mout), 0, 1, _stream())
    return Mout


def be_ctranspose(Mout: torch.Tensor, Min: torch.Tensor | None = None) -> torch.Tensor:
    """``be_ctranspose!(Mout, Min)`` = ``adjoint!`` (``src/backend_cpu.jl:58-68``);
    with one argument the non-mutating ``be_ctranspose(Min)`` (``:70-80``): a
    new device matrix holding Min^H."""
    if Min is None:
        Min = Mout
        Mout = be_zero_array((Min.shape[1], Min.shape[0]), device=Min.device)
    call("negf_zadjoint_strided_batched", Min.shape[0], Min.shape[1], _ptr(Min), _ld(Min), 0,
         _ptr(Mout), _ld(Mout), 0, 1, _stream())
    return Mout


def be_gemm(tA: str, tB: str, alpha, A: torch.Tensor, B: torch.Tensor, beta,
            C: torch.Tensor) -> torch.Tensor:
    """zgemm (``src/backend_cpu.jl:208-219``); C is updated in place."""
    m, n = C.shape
    k = A.shape[1] if tA == "N" else A.shape[0]
    call("negf_zgemm", tA.encode(), tB.encode(), m, n, k, cplx(alpha), _ptr(A), _ld(A), _ptr(B),
         _ld(B), cplx(beta), _ptr(C), _ld(C), _stream())
    return C


class GpuLU:
    """LU holder of the AMD backend (the counterpart of ``CpuLU``,
    ``src/backend_cpu.jl:7-10``): LAPACK-layout factors and 1-based pivots on
    the device."""

    def __init__(self, A: torch.Tensor, piv: torch.Tensor):
        self.A = A
        self.piv = piv


def be_lu(Mout: GpuLU | None, Min: torch.Tensor) -> GpuLU:
    """copy + zgetrf (``src/backend_cpu.jl:147-169``); SingularError where the
    reference exits."""
    n = Min.shape[0]
    if Mout is None:
        Mout = GpuLU(be_zero_array((n, n), Min.device),
                     torch.zeros(n, dtype=torch.int32, device=Min.device))
    be_copy_in_hw(Mout.A, Min)
    info = torch.zeros(1, dtype=torch.int32, device=Min.device)
    call("negf_zgetrf", n, _ptr(Mout.A), _ld(Mout.A), _ptr(Mout.piv), _ptr(info), _stream())
    _check_info("be_lu", info)
    return Mout


def be_mldivide(trans: str, Mout: torch.Tensor, Min: torch.Tensor, Mlu: GpuLU) -> torch.Tensor:
    """Mout = op(LU)^{-1} Min via copy + zgetrs (``src/backend_cpu.jl:193-206``)."""
    be_copy_in_hw(Mout, Min)
    n, nrhs = Mout.shape
    call("negf_zgetrs", trans.encode(), n, nrhs, _ptr(Mlu.A), _ld(Mlu.A), _ptr(Mlu.piv),
         _ptr(Mout), _ld(Mout), _stream())
    return Mout


def be_getrf(A: torch.Tensor) -> torch.Tensor:
    """zgetrf in place with the pivots discarded (``src/backend_cpu.jl:255-265``)."""
    n = A.shape[0]
    piv = torch.zeros(n, dtype=torch.int32, device=A.device)
    info = torch.zeros(1, dtype=torch.int32, device=A.device)
    call("negf_zgetrf", n, _ptr(A), _ld(A), _ptr(piv), _ptr(info), _stream())
    return A


def be_getri(A: torch.Tensor) -> torch.Tensor:
    """zgetri with identity pivots (``src/backend_cpu.jl:229-239``)."""
    n = A.shape[0]
    work = torch.empty(n * n, dtype=torch.complex128, device=A.device)
    call("negf_zgetri", n, _ptr(A), _ld(A), None, _ptr(work), _stream())
    return A


def be_trsm(side: str, ul: str, tA: str, dA: str, alpha, A: torch.Tensor,
            B: torch.Tensor) -> torch.Tensor:
    """ztrsm in place on B (``src/backend_cpu.jl:241-253``)."""
    m, n = B.shape
    call("negf_ztrsm", side.encode(), ul.encode(), tA.encode(), dA.encode(), m, n, cplx(alpha),
         _ptr(A), _ld(A), _ptr(B), _ld(B), _stream())
    return B


def bndiag_of_inv_direct(M, block_sizes, ndiag: dict | None = None):
    """``bndiag_of_inv_direct`` (``src/matinvertndiag_direct.jl:17-86``): the
    dense inverse of M by ``be_inv`` on the device, then its block
    n-diagonal (``ndiag["out"]``) as a scipy CSC matrix holding exactly the
    band blocks.  M: a scipy sparse or numpy matrix (copied to the device, as
    the reference's be_copy_to_hw) or a column-major device tensor."""
    import scipy.sparse as sp
    _require_gpu()
    ndiag = ndiag or {"in": 3, "out": 3}
    if isinstance(M, torch.Tensor):
        Md = col_major(M.to(torch.complex128))
    else:
        dense = M.toarray() if sp.issparse(M) else np.asarray(M)
        Md = col_major(torch.from_numpy(np.asarray(dense, dtype=np.complex128)).to("cuda"))
    Minv = be_inv(Md).cpu().numpy()
    sizes = [int(x) for x in block_sizes]
    off = np.concatenate([[0], np.cumsum(sizes)])
    w = (int(ndiag["out"]) - 1) // 2
    N = len(sizes)
    out = sp.lil_matrix(Minv.shape, dtype=np.complex128)
    for i in range(N):
        for j in range(max(0, i - w), min(N, i + w + 1)):
            out[off[i]:off[i + 1], off[j]:off[j + 1]] = Minv[off[i]:off[i + 1], off[j]:off[j + 1]]
    return out.tocsc()


def be_inv(M: torch.Tensor) -> torch.Tensor:
    """Explicit inverse (``src/backend_cpu.jl:143-145``): Gauss-Jordan up to
    n = 2048, above that LU + triangular inverses + their product (up to
    n = 16384)."""
    n = M.shape[0]
    out = be_zero_array((n, n), M.device)
    info = torch.zeros(1, dtype=torch.int32, device=M.device)
    call("negf_zinv_strided_batched", n, _ptr(M), _ld(M), 0, _ptr(out), _ld(out), 0, 1,
         _ptr(info), _stream())
    _check_info("be_inv", info)
    return out
