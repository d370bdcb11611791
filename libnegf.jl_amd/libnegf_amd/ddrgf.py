"""Domain-decomposed RGF on the GPU: host mirror of
``src/matinvertndiag_ddrgf_{core,setup,utils}.jl``.

``bndiag_of_inv_ddrgf_get_nr_tasks`` is the reference's task split
(utils.jl:278-312); ``allocate_aux_data_DDRGF`` picks the levels (the role of
the reference's timing-based tuner ``opt_params``, utils.jl:357-453, here a
deterministic choice: the block size that gives the requested task count,
or the one minimising a GEMM-count cost model); ``bndiag_of_inv_ddrgf`` runs
negf_ddrgf, whose partitions are solved as one batch on the GPU.
"""
from __future__ import annotations

import ctypes as C
import dataclasses

import torch

from ._lib import call, SingularError
from .btd import DeviceBlockTridiag

__all__ = ["bndiag_of_inv_ddrgf_get_nr_tasks", "levels_for_tasks", "cost_ddrgf_gpu",
           "AuxDataDDRGF", "allocate_aux_data_DDRGF", "bndiag_of_inv_ddrgf"]


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


def cost_ddrgf_gpu(npl: int, levels: list) -> float:
    """GEMM-equivalents on the critical path of one level-recursion (each
    batched launch counted once, whatever its batch): the chunk chains are
    O(bs) long, the reduced chain O(tasks)."""
    cost, n = 0.0, npl
    for bs in levels:
        nr, last = bndiag_of_inv_ddrgf_get_nr_tasks(n, bs)
        if nr <= 1:
            return cost + 4.0 * n
        cost += 14.0 * bs + 12.0
        n = nr
    return cost + 4.0 * n


@dataclasses.dataclass
class AuxDataDDRGF:
    """Levels (blockSizeD1 per level) and the per-instance info array; the
    library keeps the buffers (AuxDataDDRGF, core:1-30)."""

    levels: list
    info: torch.Tensor


def allocate_aux_data_DDRGF(M: DeviceBlockTridiag, levels=None, tasks=None,
                            max_levels: int = 3) -> AuxDataDDRGF:
    """``allocate_aux_data_DDRGF`` (setup.jl:4-51).  `levels` fixes the
    blockSizeD1 per level; `tasks` asks for a task count at one level; else
    the levels minimising ``cost_ddrgf_gpu`` over blockSizeD1 in 1..16 and up
    to `max_levels` levels."""
    npl = M.nblocks
    if levels is None:
        if tasks is not None:
            levels = levels_for_tasks(npl, tasks)
        else:
            best = (cost_ddrgf_gpu(npl, [npl]), [npl])
            frontier = [([], npl)]
            for _ in range(max_levels):
                nxt = []
                for lv, n in frontier:
                    for bs in range(1, 17):
                        nr, _ = bndiag_of_inv_ddrgf_get_nr_tasks(n, bs)
                        if nr < 2:
                            continue
                        cand = lv + [bs]
                        c = cost_ddrgf_gpu(npl, cand)
                        if c < best[0]:
                            best = (c, cand)
                        nxt.append((cand, nr))
                frontier = nxt
            levels = best[1]
    return AuxDataDDRGF(list(levels), torch.zeros(M.batch, dtype=torch.int32,
                                                  device=M.diag.device))


def bndiag_of_inv_ddrgf(Mout: DeviceBlockTridiag, Min: DeviceBlockTridiag, aux: AuxDataDDRGF):
    """band(Min^{-1}) into Mout (core:719-758).  Uniform block sizes."""
    L = Min.layout
    if len(set(L.b)) != 1:
        raise ValueError("the GPU DD-RGF takes uniform block sizes")
    lv = (C.c_int32 * len(aux.levels))(*aux.levels)
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    p = lambda t: C.c_void_p(t.data_ptr()) if t is not None else None
    call("negf_ddrgf", L.N, L.b[0], p(Min.diag), L.nd, p(Min.upper), max(L.no, 1),
         p(Min.lower), max(L.no, 1), p(Mout.diag), p(Mout.upper), p(Mout.lower), Min.batch, lv,
         len(aux.levels), p(aux.info), s)
    bad = aux.info.cpu()
    if bool((bad != 0).any()):
        raise SingularError("bndiag_of_inv_ddrgf", bad.tolist())
    return Mout
