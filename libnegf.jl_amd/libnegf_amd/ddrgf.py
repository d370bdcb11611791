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
           "allocate_aux_data_DDRGF", "bndiag_of_inv_ddrgf"]


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
