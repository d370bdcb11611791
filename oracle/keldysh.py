"""CPU oracle: recursive Keldysh (RKD) and the NEGF observables around it.

TEST INFRASTRUCTURE ONLY (see oracle/backend.py).

``keldyshndiag`` restates ``keldyshndiag_v3!`` (``src/keldyshndiag.jl:166-188``)
and its four passes call for call:
  upward RGF  ``:191-219``, upward RKD ``:221-235``, central RKD ``:237-302``,
  downward RKD ``:304-318``.
Input: A (non-Hermitian storage) and a Hermitian block-tridiagonal Sigma in
Hermitian storage (diag + upper).  Output: (diag, upper) of
band(A^{-1} Sigma A^{-H}), Hermitian storage, like ``auxK.buffS``.

``transmission``, ``lesser_sigma`` and friends have no reference
counterpart; they are the NEGF observables of the benchmark workload:
T(E) = Tr(Gamma_L G_{1N} Gamma_R G_{1N}^H), Sigma^n = f_L Gamma_L (+) f_R Gamma_R,
G^< = i G^n.

Pinning (tests/test_oracle_pins.py): RKD against the dense
G Sigma G^H (the reference's own check, ``test/test_keldyshndiag.jl:86-95``);
the Ward identity i (G - G^H) = G (Gamma_L + Gamma_R + 2 eta I) G^H, which ties
RKD to RGF with no dense inverse; T against the dense G_{1N} and, for the
b = 1 chain, the closed form (T = 1 inside an ideal-lead band).
"""
from __future__ import annotations

import numpy as np

from .backend import be_lu, be_mldivide, be_gemm, be_ctranspose
from .rgf import rgf_upward

__all__ = ["keldyshndiag", "dense_keldysh_band", "transmission_dense"]


def keldyshndiag(A, S_diag, S_upper, upward=None):
    N = len(A.diag)
    buf = rgf_upward(A) if upward is None else upward      # keldyshndiag_upward_rgf!, :191-219
    Sd = [np.array(s, order="F") for s in S_diag]           # bm_copy!(auxData.buffS, S), :178
    Su = [np.array(s, order="F") for s in S_upper]
    t = lambda i: buf[("t", i)]
    l = lambda i: buf[("l", i)]
    lu = lambda i: buf[("lu", i)]

    # upward RKD, :230-234
    for i in range(N - 2, -1, -1):
        Sd[i] = be_gemm("N", "C", -1.0, t(i), Su[i], 1.0, Sd[i])
        Su[i] = be_gemm("N", "N", -1.0, t(i), Sd[i + 1], 1.0, Su[i])
        Sd[i] = be_gemm("N", "C", -1.0, Su[i], t(i), 1.0, Sd[i])

    # central RKD, :256-301
    if N > 1:
        m12 = be_mldivide("N", Su[0], lu(0))                              # :257
        m21 = be_ctranspose(m12)                                          # :258
        x21 = be_mldivide("N", m21, lu(1))                                # :259
        Su[0] = be_ctranspose(x21)                                        # :260
    X = np.asfortranarray(np.eye(A.diag[0].shape[0], dtype=np.complex128))  # :263-268
    for i in range(1, N - 1):                                             # :270
        m = be_gemm("N", "N", 1.0, X, t(i - 1), 0.0, None)                 # :273
        xo = be_mldivide("N", m, lu(i - 1))                               # :274
        X = be_gemm("N", "N", 1.0, A.low(i - 1), xo, 0.0, None)            # :275
        X[np.diag_indices(X.shape[0])] += 1.0                             # :280-283
        mii = be_mldivide("N", X, lu(i))                                  # :285
        mup = be_gemm("N", "N", 1.0, mii, Su[i], 0.0, None)                # :286
        mlo = be_ctranspose(mup)                                          # :289
        xlo = be_mldivide("N", mlo, lu(i + 1))                            # :290
        Su[i] = be_ctranspose(xlo)                                        # :291
    for i in range(N):                                                    # :296-301
        m = be_mldivide("N", Sd[i], lu(i))
        xd = be_ctranspose(m)
        m = be_mldivide("N", xd, lu(i))
        Sd[i] = be_ctranspose(m)

    # downward RKD, :313-317
    for i in range(N - 1):
        Sd[i + 1] = be_gemm("C", "C", -1.0, Su[i], l(i + 1), 1.0, Sd[i + 1])
        Su[i] = be_gemm("N", "C", -1.0, Sd[i], l(i + 1), 1.0, Su[i])
        Sd[i + 1] = be_gemm("N", "N", -1.0, l(i + 1), Su[i], 1.0, Sd[i + 1])
    return Sd, Su


def dense_keldysh_band(A, S_diag, S_upper):
    """Brute force Gr (S Gr^H) on the band, as ``test/test_keldyshndiag.jl:86-91``."""
    from .rgf import band_blocks
    from .backend import be_inv
    n = len(S_diag)
    sizes = [d.shape[0] for d in S_diag]
    off = np.concatenate([[0], np.cumsum(sizes)])
    S = np.zeros((off[-1], off[-1]), dtype=np.complex128)
    for i in range(n):
        S[off[i]:off[i + 1], off[i]:off[i + 1]] = S_diag[i]
    for i in range(n - 1):
        S[off[i]:off[i + 1], off[i + 1]:off[i + 2]] = S_upper[i]
        S[off[i + 1]:off[i + 2], off[i]:off[i + 1]] = S_upper[i].conj().T
    Gr = be_inv(A.to_dense())
    Gn = Gr @ (S @ Gr.conj().T)
    band = band_blocks(Gn, sizes)
    return [band[(i, i)] for i in range(n)], [band[(i, i + 1)] for i in range(n - 1)]


def transmission_dense(A, gamma_L, gamma_R):
    """T = Tr(Gamma_L G_{1N} Gamma_R G_{1N}^H) from a dense inverse."""
    from .backend import be_inv
    sizes = A.block_sizes
    off = np.concatenate([[0], np.cumsum(sizes)])
    G = be_inv(A.to_dense())
    g1n = G[off[0]:off[1], off[-2]:off[-1]]
    return float(np.real(np.trace(gamma_L @ g1n @ gamma_R @ g1n.conj().T)))
