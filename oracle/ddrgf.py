"""CPU oracle: domain-decomposed RGF (DD-RGF), one or more levels.

TEST INFRASTRUCTURE ONLY (see oracle/backend.py).

Restates ``bndiag_of_inv_ddrgf!`` (``src/matinvertndiag_ddrgf_core.jl:719-758``)
phase by phase, on dense blocks:
  task split        ``bndiag_of_inv_ddrgf_get_nr_tasks`` (utils.jl:278-312)
  permutation       ``..._create_permutation_vector`` (utils.jl:478-575):
                    separators (domain 2) first, then the interior layers
                    (domain 1), chunk c = [separator s_c, L_c interior layers]
  T11^{-1}          ``..._inv_of_T11!`` (core:437-498): per-chunk RGF with the
                    full inverse of the chunk
  T11^{-1} T12      ``..._compute_THat11Inv_x_THat12!`` (core:48-114)
  T21 T11^{-1}      ``..._compute_THat21_x_THat11Inv!`` (core:203-266)
  Schur complement  ``..._build_Schur_compl!`` (core:500-654):
                    S = T22 - T21 T11^{-1} T12, block tridiagonal over separators
  band(S^{-1})      ``..._inv_Schur_compl!`` (core:656-680): sequential RGF at
                    the last level, DD-RGF of S otherwise
  G12 / G21 / G11   ``..._compute_minus_THat11Inv_x_THat12_x_THatSInv!``
                    (core:116-201), ``..._compute_minus_x_THatSInv_THat21_x_THat11Inv!``
                    (core:268-344), ``..._compute_11_part!`` (core:346-435)
Output: the band of A^{-1} in the original layer order, dict {(i, j): block}.

Pinning (tests/test_oracle_ddrgf.py): equality with sequential RGF on every
band block (test/test_matinvertndiag_ddrgf.jl:254-293) and the Schur
complement against S computed from a dense inverse of T11 (:214-241), on the
reference's uneven generator; the permutation against a hand count.
"""
from __future__ import annotations

import numpy as np

from .backend import be_gemm
from .rgf import bndiag_of_inv_rgf_local

__all__ = ["get_nr_tasks", "create_permutation_vector", "chunks", "bndiag_of_inv_ddrgf",
           "schur_complement"]


def get_nr_tasks(npl: int, bs1: int, bs2: int = 1):
    """``bndiag_of_inv_ddrgf_get_nr_tasks`` (utils.jl:278-312): [nrTasks,
    leftover interior size of the last chunk (0 = full)], or [-1, 0]."""
    if bs2 != 1:
        raise ValueError("The value of blockSizeD2 should be set to 1")
    per = bs1 + bs2
    nr = npl // per + 1
    left = npl - (nr - 1) * per
    if left == 0:
        return nr - 1, 0
    if left - 1 == 0:
        return -1, 0
    return nr, left - 1


def create_permutation_vector(npl: int, bs1: int):
    """``..._create_permutation_vector`` (utils.jl:478-575), 0-based:
    permVecInv[new] = old, and the sizes of the domains (separators, then
    chunk interiors)."""
    nr, last = get_nr_tasks(npl, bs1)
    if nr < 0:
        raise ValueError("this blockSizeD1 does not split npl")
    lastL = bs1 if last == 0 else last
    perm, sizes = [], []
    old = 0
    for c in range(nr):
        perm.append(old)
        sizes.append(1)
        old += 1 + (lastL if c == nr - 1 else bs1)
    old = 0
    for c in range(nr):
        L = lastL if c == nr - 1 else bs1
        old += 1
        perm.extend(range(old, old + L))
        sizes.append(L)
        old += L
    return perm, sizes


def chunks(npl: int, bs1: int):
    """[(separator layer, first interior layer, interior length)] per task."""
    nr, last = get_nr_tasks(npl, bs1)
    if nr < 0:
        raise ValueError(f"blockSizeD1 = {bs1} leaves a last chunk without interior layers")
    lastL = bs1 if last == 0 else last
    out, s = [], 0
    for c in range(nr):
        L = lastL if c == nr - 1 else bs1
        out.append((s, s + 1, L))
        s += 1 + L
    return out


class _Sub:
    """A chunk's interior as a block-tridiagonal view (for the oracle RGF)."""

    def __init__(self, A, f, L):
        self.diag = [A.diag[f + j] for j in range(L)]
        self.upper = [A.upper[f + j] for j in range(L - 1)]
        self._low = [A.low(f + j) for j in range(L - 1)]
        self.block_sizes = [d.shape[0] for d in self.diag]

    def low(self, i):
        return self._low[i]


class _Btd:
    def __init__(self, diag, upper, lower):
        self.diag, self.upper, self._low = diag, upper, lower
        self.block_sizes = [d.shape[0] for d in diag]

    def low(self, i):
        return self._low[i]


def schur_complement(A, bs1, g=None):
    """S = T22 - T21 T11^{-1} T12 over the separators (core:500-654), from the
    chunk inverses g[c] (full, dict) -- computed here when not given."""
    ch = chunks(len(A.diag), bs1)
    if g is None:
        g = [bndiag_of_inv_rgf_local(_Sub(A, f, L), full_inverse=True) for (_, f, L) in ch]
    T = len(ch)
    # Y[c][j, side] = T11^{-1} T12 (side 0: separator c, side 1: separator c+1)
    Y, Z = [], []
    for c, (s, f, L) in enumerate(ch):
        yc, zc = {}, {}
        for j in range(L):                                                         # core:84-110
            yc[(j, 0)] = be_gemm("N", "N", 1.0, g[c][(j, 0)], A.low(s), 0.0, None)
            if c + 1 < T:
                yc[(j, 1)] = be_gemm("N", "N", 1.0, g[c][(j, L - 1)], A.upper[f + L - 1], 0.0, None)
            zc[(0, j)] = be_gemm("N", "N", 1.0, A.upper[s], g[c][(0, j)], 0.0, None)   # core:203-266
            if c + 1 < T:
                zc[(1, j)] = be_gemm("N", "N", 1.0, A.low(f + L - 1), g[c][(L - 1, j)], 0.0, None)
        Y.append(yc)
        Z.append(zc)
    Sd, Su, Sl = [], [], []
    for c, (s, f, L) in enumerate(ch):
        d = be_gemm("N", "N", -1.0, A.upper[s], Y[c][(0, 0)], 1.0, A.diag[s])
        if c > 0:
            sp, fp, Lp = ch[c - 1]
            d = be_gemm("N", "N", -1.0, A.low(s - 1), Y[c - 1][(Lp - 1, 1)], 1.0, d)
        Sd.append(d)
        if c + 1 < T:
            Su.append(be_gemm("N", "N", -1.0, A.upper[s], Y[c][(0, 1)], 0.0, None))
            Sl.append(be_gemm("N", "N", -1.0, A.low(f + L - 1), Y[c][(L - 1, 0)], 0.0, None))
    return _Btd(Sd, Su, Sl), g, Y, Z


def bndiag_of_inv_ddrgf(A, bs_levels):
    """band(A^{-1}) by DD-RGF with blockSizeD1 = bs_levels[0] at this level
    and the reduced system solved by DD-RGF with bs_levels[1:] (or RGF when
    none are left), as ``bndiag_of_inv_ddrgf!`` with ``depth``."""
    if isinstance(bs_levels, int):
        bs_levels = [bs_levels]
    npl = len(A.diag)
    bs1 = bs_levels[0]
    ch = chunks(npl, bs1)
    T = len(ch)
    if T == 1:                                                                    # core:724-729
        return bndiag_of_inv_rgf_local(A)
    S, g, Y, Z = schur_complement(A, bs1)
    Sinv = (bndiag_of_inv_ddrgf(S, bs_levels[1:]) if len(bs_levels) > 1
            else bndiag_of_inv_rgf_local(S))                                      # core:656-680
    G = {}
    for c in range(T):
        G[(ch[c][0], ch[c][0])] = Sinv[(c, c)]
    for c, (s, f, L) in enumerate(ch):
        right = c + 1 < T

        def sv(a, b):
            return Sinv[(a, b)]
        # G12 columns of this chunk (core:116-201): -T11^{-1} T12 S^{-1}
        G12 = {}
        for j in range(L):
            for side, cc in ((0, c), (1, c + 1)):
                if cc >= T:
                    continue
                x = be_gemm("N", "N", -1.0, Y[c][(j, 0)], sv(c, cc), 0.0, None)
                if right:
                    x = be_gemm("N", "N", -1.0, Y[c][(j, 1)], sv(c + 1, cc), 1.0, x)
                G12[(j, side)] = x
        # G21 rows (core:268-344): -S^{-1} T21 T11^{-1}
        G21 = {}
        for k in range(L):
            for side, cc in ((0, c), (1, c + 1)):
                if cc >= T:
                    continue
                x = be_gemm("N", "N", -1.0, sv(cc, c), Z[c][(0, k)], 0.0, None)
                if right:
                    x = be_gemm("N", "N", -1.0, sv(cc, c + 1), Z[c][(1, k)], 1.0, x)
                G21[(side, k)] = x
        # band entries between the chunk and its separators
        G[(f, s)] = G12[(0, 0)]
        G[(s, f)] = G21[(0, 0)]
        if right:
            G[(f + L - 1, f + L)] = G12[(L - 1, 1)]
            G[(f + L, f + L - 1)] = G21[(1, L - 1)]
        # 11 part (core:346-435): T11^{-1} + T11^{-1} T12 S^{-1} T21 T11^{-1}
        for j in range(L):
            for k in range(max(0, j - 1), min(L, j + 2)):
                x = be_gemm("N", "N", -1.0, Y[c][(j, 0)], G21[(0, k)], 1.0, g[c][(j, k)])
                if right:
                    x = be_gemm("N", "N", -1.0, Y[c][(j, 1)], G21[(1, k)], 1.0, x)
                G[(f + j, f + k)] = x
    return G
