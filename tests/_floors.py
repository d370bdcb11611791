"""FP64 noise floors of the production-shape problems, measured on the CPU
on the very inputs the GPU tests use.

At N = 64, b = 512, eta = 1e-4 the problem itself limits how closely two
correct FP64 computations agree: the oracle's RGF (LU solves, the reference's
call sequence) and the same recursion with explicit inverses g_i = S_i^{-1}
(the form the GPU runs) differ by ~3e-12 per block.  The GPU bars are a
stated multiple of these spreads, computed per input rather than assumed.
"""
import numpy as np

from oracle.rgf import bndiag_of_inv_rgf_local, rgf_upward
from oracle.keldysh import keldyshndiag


def relfro(a, b):
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def rgf_explicit(A):
    """band(A^{-1}) by RGF with explicit inverses g_i = S_i^{-1} (numpy
    inv + products), the recursion of src/matinvertndiag_rgf.jl:143-182 in
    the GPU's form.  Also returns g, t, l."""
    N = len(A.diag)
    g, t, l = [None] * N, [None] * N, [None] * N
    g[N - 1] = np.linalg.inv(A.diag[N - 1])
    for i in range(N - 2, -1, -1):
        t[i] = A.upper[i] @ g[i + 1]
        l[i + 1] = g[i + 1] @ A.low(i)
        g[i] = np.linalg.inv(A.diag[i] - A.upper[i] @ l[i + 1])
    G = {(0, 0): g[0]}
    for i in range(1, N):
        G[(i - 1, i)] = -G[(i - 1, i - 1)] @ t[i - 1]
        G[(i, i - 1)] = -l[i] @ G[(i - 1, i - 1)]
        G[(i, i)] = g[i] - l[i] @ G[(i - 1, i)]
    return G, g, t, l


def rgf_floor(A, ref=None):
    """Largest block spread between the oracle's RGF and the explicit-inverse
    FP64 RGF on A; returns (spread, oracle band)."""
    ref = bndiag_of_inv_rgf_local(A) if ref is None else ref
    X, *_ = rgf_explicit(A)
    return max(relfro(X[k], ref[k]) for k in ref), ref


def keldysh_injection_fp64(A, G, gamma_L, gamma_R, fL, fR, t, l):
    """G^n diagonal blocks of a contact-only Sigma^n = fL Gamma_L (+) fR Gamma_R
    in the injection form the GPU energy job uses (FP64 numpy): C_j =
    l_j C_{j-1} l_j^H from C_0 = G_00 fL Gamma_L G_00^H, B_j = t_j B_{j+1}
    t_j^H from B_{N-1} = Gamma_R, G^n_jj = C_j + fR G_jj B_j G_jj^H."""
    N = len(A.diag)
    B = [None] * N
    B[N - 1] = gamma_R
    for i in range(N - 2, -1, -1):
        B[i] = (t[i] @ B[i + 1]) @ t[i].conj().T
    C = (G[(0, 0)] @ (fL * gamma_L)) @ G[(0, 0)].conj().T
    out = []
    for j in range(N):
        if j > 0:
            C = (l[j] @ C) @ l[j].conj().T
        out.append(C + (G[(j, j)] @ (fR * B[j])) @ G[(j, j)].conj().T)
    return out


def energy_floor(dev, A, fL, fR):
    """Spreads between the oracle (RGF + keldyshndiag, the reference's call
    sequence) and the explicit-inverse / injection form, both FP64, for one
    energy: (G^R spread, G^n spread, oracle G^R diag, oracle G^n diag, up)."""
    N, b = dev.nblocks, dev.b
    up = rgf_upward(A)
    ref = bndiag_of_inv_rgf_local(A, upward=up)
    X, g, t, l = rgf_explicit(A)
    Sd = [np.zeros((b, b), complex, order="F") for _ in range(N)]
    Sd[0] = np.asfortranarray(fL * dev.gamma_L)
    Sd[-1] = np.asfortranarray(Sd[-1] + fR * dev.gamma_R)
    Su = [np.zeros((b, b), complex, order="F") for _ in range(N - 1)]
    gd, _ = keldyshndiag(A, Sd, Su, upward=up)
    inj = keldysh_injection_fp64(A, X, dev.gamma_L, dev.gamma_R, fL, fR, t, l)
    sG = max(relfro(X[(i, i)], ref[(i, i)]) for i in range(N))
    sN = max(relfro(inj[i], gd[i]) for i in range(N))
    return sG, sN, [ref[(i, i)] for i in range(N)], gd, up
