"""Domain-decomposed RGF on the GPU: host mirror of
``src/matinvertndiag_ddrgf_{core,setup,utils}.jl``.

``bndiag_of_inv_ddrgf_get_nr_tasks`` is the reference's task split
(utils.jl:278-312); ``allocate_aux_data_DDRGF`` picks the levels the way the
reference's ``opt_params`` (utils.jl:357-453) does, by timing: the tune
table's entry for the shape, else the candidates (plain RGF among them) timed
on the device by ``tune.tune_bndiag``; ``bndiag_of_inv_ddrgf`` runs
negf_ddrgf_blocksizes (any block sizes; the partitions of every instance are
solved as one batch) or, when the tuned choice is RGF, negf_rgf.
"""
from __future__ import annotations

import ctypes as C
import dataclasses

import torch

from ._lib import call, SingularError
from .btd import DeviceBlockTridiag

__all__ = ["bndiag_of_inv_ddrgf_get_nr_tasks", "levels_for_tasks", "AuxDataDDRGF",
           "allocate_aux_data_DDRGF", "bndiag_of_inv_ddrgf", "bndiag_of_inv_ddrgf_error_inv_of_T11",
           "bndiag_devices"]


def bndiag_of_inv_ddrgf_get_nr_tasks(npl: int, blockSizeD1: int, blockSizeD2: int = 1):
    """[nrTasks, leftover interior size of the last task] (0 = full), or
    [-1, 0] when the last task would have no interior (utils.jl:278-312)."""
    if blockSizeD2 != 1:
        raise ValueError("The value of blockSizeD2 should be set to 1")
    per = blockSizeD1 + blockSizeD2
    nr = npl // per + 1
    left = npl - (nr - 1) * per
    if left == 0:
        return [nr - 1, 0]
    if left - 1 == 0:
        return [-1, 0]
    return [nr, left - 1]


def levels_for_tasks(npl: int, tasks: int) -> list:
    """blockSizeD1 giving `tasks` tasks at one level (e.g. 15 / 7 / 3 for
    npl = 128 / 64 / 32 and 8 tasks)."""
    for bs in range(1, npl + 1):
        nr, _ = bndiag_of_inv_ddrgf_get_nr_tasks(npl, bs)
        if nr == tasks:
            return [bs]
    raise ValueError(f"no blockSizeD1 splits {npl} layers into {tasks} tasks")


@dataclasses.dataclass
class AuxDataDDRGF:
    """Levels (blockSizeD1 per level; [0] = plain RGF, the tuner's choice
    where it is faster) and the per-instance info array; the library keeps
    the buffers (AuxDataDDRGF, core:1-30)."""

    levels: list
    info: torch.Tensor
    times: dict | None = None   # the tuner's measurements when it chose the levels

    @property
    def rgf(self) -> bool:
        return not self.levels or self.levels[0] == 0


def allocate_aux_data_DDRGF(M: DeviceBlockTridiag, levels=None, tasks=None, max_levels: int = 2,
                            persist: bool = False) -> AuxDataDDRGF:
    """``allocate_aux_data_DDRGF`` (setup.jl:4-51).  `levels` fixes the
    blockSizeD1 per level; `tasks` asks for a task count at one level; else
    the tune table's choice for (N, bmax, batch), else the fastest candidate
    timed now on a matrix of M's shape (recorded in the table, and in the
    table file with `persist`)."""
    from . import tune
    info = torch.zeros(M.batch, dtype=torch.int32, device=M.diag.device)
    times = None
    if levels is None:
        if tasks is not None:
            levels = levels_for_tasks(M.nblocks, tasks)
        else:
            levels = tune.lookup("bndiag", (M.nblocks, max(M.block_sizes), M.batch))
            if levels is None:
                levels, times = tune.tune_bndiag(M.block_sizes, M.batch, max_levels=max_levels,
                                                 persist=persist)
    return AuxDataDDRGF(list(levels), info, times)


def bndiag_of_inv_ddrgf(Mout: DeviceBlockTridiag, Min: DeviceBlockTridiag, aux: AuxDataDDRGF,
                        check: bool = True):
    """band(Min^{-1}) into Mout (core:719-758), any block sizes."""
    L = Min.layout
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    p = lambda t: C.c_void_p(t.data_ptr()) if t is not None else None
    if aux.rgf:
        call("negf_rgf", L.N, L.sizes_c(), p(Min.diag), L.nd, p(Min.upper), max(L.no, 1),
             p(Min.lower), max(L.no, 1), p(Mout.diag), p(Mout.upper), p(Mout.lower), Min.batch,
             p(aux.info), s)
    else:
        lv = (C.c_int32 * len(aux.levels))(*aux.levels)
        call("negf_ddrgf_blocksizes", L.N, L.sizes_c(), p(Min.diag), L.nd, p(Min.upper),
             max(L.no, 1), p(Min.lower), max(L.no, 1), p(Mout.diag), p(Mout.upper),
             p(Mout.lower), Min.batch, lv, len(aux.levels), p(aux.info), s)
    if check:
        bad = aux.info.cpu()
        if bool((bad != 0).any()):
            raise SingularError("bndiag_of_inv_ddrgf", bad.tolist())
    return Mout


def bndiag_of_inv_ddrgf_error_inv_of_T11(Min: DeviceBlockTridiag, aux: AuxDataDDRGF) -> list:
    """``bndiag_of_inv_ddrgf_error_inv_of_T11`` (utils.jl:776-824): how well
    the interiors' inverses T11^{-1} that DD-RGF eliminates with solve T11,
    per instance

        sqrt( sum over tasks, block rows i, block columns j of
              ||(T11 T11^{-1})_ij - delta_ij I||_F^2  /  sum of the interior sizes ),

    with T11^{-1} every block of each interior's inverse (negf_rgf_full, the
    recursion DD-RGF runs on the interiors) and T11 the interior's
    block-tridiagonal part of Min, at the first level's task split (any block
    sizes; the reference assumes equal ones).  Returns one value per
    instance."""
    import ctypes as Cc
    from ._lib import cplx
    if aux.rgf:
        raise ValueError("the tuned method is plain RGF: there is no T11")
    L = Min.layout
    bs = aux.levels[0]
    nr, left = bndiag_of_inv_ddrgf_get_nr_tasks(L.N, bs)
    if nr < 2:
        raise ValueError("the first level's block size does not split the chain")
    dev = Min.diag.device
    s = Cc.c_void_p(torch.cuda.current_stream().cuda_stream)
    num = torch.zeros(Min.batch, dtype=torch.float64, device=dev)
    den = 0
    for c in range(nr):
        f = c * (bs + 1) + 1
        Lc = left if (c == nr - 1 and left) else bs
        Lc = min(Lc, L.N - f)
        if Lc <= 0:
            continue
        sizes = [L.b[f + i] for i in range(Lc)]
        chunk = DeviceBlockTridiag(sizes, Min.batch, False, dev, zero=False)
        for i in range(Lc):
            chunk.diag[:, chunk.layout.od[i]:chunk.layout.od[i + 1]] = \
                Min.diag[:, L.od[f + i]:L.od[f + i + 1]]
        for i in range(Lc - 1):
            chunk.upper[:, chunk.layout.oo[i]:chunk.layout.oo[i + 1]] = \
                Min.upper[:, L.oo[f + i]:L.oo[f + i + 1]]
            if Min.lower is not None:
                chunk.lower[:, chunk.layout.oo[i]:chunk.layout.oo[i + 1]] = \
                    Min.lower[:, L.oo[f + i]:L.oo[f + i + 1]]
            else:
                for e in range(Min.batch):
                    chunk.block("lower", i, e).copy_(Min.block("upper", f + i, e).conj().T)
        n = sum(sizes)
        X = torch.empty((Min.batch, n, n), dtype=torch.complex128, device=dev)
        info = torch.zeros(Min.batch, dtype=torch.int32, device=dev)
        CL = chunk.layout
        call("negf_rgf_full", CL.N, CL.sizes_c(), Cc.c_void_p(chunk.diag.data_ptr()), CL.nd,
             Cc.c_void_p(chunk.upper.data_ptr()), max(CL.no, 1),
             Cc.c_void_p(chunk.lower.data_ptr()), max(CL.no, 1), Cc.c_void_p(X.data_ptr()),
             Min.batch, Cc.c_void_p(info.data_ptr()), s)
        # the interior as a dense column-major matrix (stored transposed: [inst, col, row])
        T = torch.zeros((Min.batch, n, n), dtype=torch.complex128, device=dev)
        off = [0]
        for x in sizes:
            off.append(off[-1] + x)
        for e in range(Min.batch):
            for i in range(Lc):
                T[e, off[i]:off[i + 1], off[i]:off[i + 1]] = chunk.block("diag", i, e).T
                if i + 1 < Lc:
                    T[e, off[i + 1]:off[i + 2], off[i]:off[i + 1]] = chunk.block("upper", i, e).T
                    T[e, off[i]:off[i + 1], off[i + 1]:off[i + 2]] = chunk.block("lower", i, e).T
        R = torch.eye(n, dtype=torch.complex128, device=dev).expand(Min.batch, n, n).clone()
        call("negf_zgemm_strided_batched", b"N", b"N", n, n, n, cplx(1.0),
             Cc.c_void_p(T.data_ptr()), n, n * n, Cc.c_void_p(X.data_ptr()), n, n * n, cplx(-1.0),
             Cc.c_void_p(R.data_ptr()), n, n * n, Min.batch, s)
        num += torch.linalg.vector_norm(R.reshape(Min.batch, -1), dim=1) ** 2
        den += n
    if bool((num != num).any()):
        raise SingularError("bndiag_of_inv_ddrgf_error_inv_of_T11", [])
    return torch.sqrt(num / den).cpu().tolist()


def bndiag_devices(jobs, check: bool = True, precision: str = "fp64"):
    """Several devices in one call (negf_bndiag_devices): ``jobs`` is a list
    of (Mout, Min, levels) with levels a blockSizeD1 list, [0] for RGF or
    None for the tune table's choice.  The devices run concurrently on the
    library's streams, joined back into the current stream.  ``precision``
    "fp64_3m" computes every block product with three real products
    (negf_bndiag_devices_prec).  Returns the per-device info arrays."""
    if precision not in ("fp64", "fp64_3m"):
        raise ValueError(f"bndiag_devices: precision {precision!r} (fp64 or fp64_3m)")
    from ._lib import NegfBandJob
    if len(jobs) > 8:
        raise ValueError("at most 8 devices per call")
    arr = (NegfBandJob * len(jobs))()
    keep, infos = [], []
    p = lambda t: C.c_void_p(t.data_ptr()) if t is not None else None
    for k, (Mout, Min, levels) in enumerate(jobs):
        L = Min.layout
        info = torch.zeros(Min.batch, dtype=torch.int32, device=Min.diag.device)
        sizes = L.sizes_c()
        lv = None if levels is None else (C.c_int32 * len(levels))(*levels)
        keep += [sizes, lv]
        infos.append(info)
        arr[k] = NegfBandJob(L.N, C.cast(sizes, C.POINTER(C.c_int32)), p(Min.diag), L.nd,
                             p(Min.upper), max(L.no, 1), p(Min.lower), max(L.no, 1),
                             p(Mout.diag), p(Mout.upper), p(Mout.lower), Min.batch,
                             C.cast(lv, C.POINTER(C.c_int32)) if lv is not None else None,
                             len(levels) if levels is not None else 0, p(info))
    call("negf_bndiag_devices_prec", len(jobs), C.cast(arr, C.c_void_p),
         0 if precision == "fp64" else 3, C.c_void_p(torch.cuda.current_stream().cuda_stream))
    if check:
        for info in infos:
            bad = info.cpu()
            if bool((bad != 0).any()):
                raise SingularError("bndiag_devices", bad.tolist())
    return infos
