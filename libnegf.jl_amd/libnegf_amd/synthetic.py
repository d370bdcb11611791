"""Seeded synthetic inputs for the block-tridiagonal NEGF path (host side, numpy).

Two generators live here:

* ``device_hamiltonian`` / ``assemble_A``: the tight-binding workload the GPU
  benchmarks are quoted on.  A principal-layer Hamiltonian H with Hermitian
  diagonal blocks, Hermitian-conjugate coupling blocks and wide-band contact
  broadenings Gamma_L / Gamma_R on the first and last layer:

      H_ii      = diag(U[-2, 2]) + (X + X^H) / 2,  X ~ CN(0, 1) / sqrt(b)
      H_{i,i+1} = Y ~ CN(0, 1) / sqrt(b),          H_{i+1,i} = H_{i,i+1}^H
      Gamma     = Z Z^H / b,                        Z ~ CN(0, 1)
      A(E)      = (E + i eta) I - H - Sigma_L - Sigma_R,  Sigma = -i Gamma / 2

  CN(0, 1) is the circular complex normal: re and im each N(0, 1/2).  Draw
  order (fixed, so every consumer sees identical inputs for a seed):
  block sizes are all equal; for i = 0..N-1: the onsite diagonal (b uniforms),
  then X (re block, im block); then for i = 0..N-2 the coupling (re, im);
  then Z_L (re, im) and Z_R (re, im).

* ``bm_create_synthetic_random``: the reference's own generator
  (``src/blockmatrix.jl:878-919``): block sizes round(b - 4 + 8 U), diagonal
  blocks with re, im ~ U[0, 1), off-diagonal blocks 0.3 times the same.  The
  reference draws from Julia's Xoshiro stream; this one draws from numpy's
  PCG64, so the *distribution* matches and the exact matrices do not.

Energy grids (``energy_grid``) are linspace(lo, hi, n) with the endpoints
included.  ``band_edge_energies`` defines the lead band edges as the extrema of
eig(H_11 + H_12 e^{ik} + H_12^H e^{-ik}) over a k grid; this is the rule used to
place the near-edge energies of the conditioning-fallback workload.
"""
from __future__ import annotations

import dataclasses
import numpy as np

__all__ = [
    "BlockTridiag",
    "DeviceHamiltonian",
    "device_hamiltonian",
    "assemble_A",
    "bm_create_synthetic_random",
    "random_hermitian_btd",
    "fermi",
    "energy_grid",
    "band_edge_energies",
    "near_edge_grid",
    "CONFIGS",
]


def _cn(rng: np.random.Generator, shape, scale: float = 1.0) -> np.ndarray:
    """Circular complex normal CN(0, scale^2): re, im ~ N(0, scale^2 / 2)."""
    s = scale / np.sqrt(2.0)
    re = rng.standard_normal(shape) * s
    im = rng.standard_normal(shape) * s
    return np.asfortranarray(re + 1j * im)


@dataclasses.dataclass
class BlockTridiag:
    """Host-side block-tridiagonal matrix.

    ``diag[i]`` is A_{i,i} (b_i x b_i), ``upper[i]`` is A_{i,i+1}
    (b_i x b_{i+1}) and ``lower[i]`` is A_{i+1,i} (b_{i+1} x b_i), every block
    a Fortran-ordered complex128 array, as in the reference's ``BlockMatrix``
    with ``ndiag = 3`` (``src/blockmatrix.jl:10-17``).  ``lower`` may be None
    for Hermitian storage (``isHermitian = true``): then A_{i+1,i} reads as
    ``upper[i]^H`` (``src/blockmatrix.jl:105-109``).
    """

    diag: list
    upper: list
    lower: list | None

    @property
    def nblocks(self) -> int:
        return len(self.diag)

    @property
    def block_sizes(self) -> list:
        return [d.shape[0] for d in self.diag]

    @property
    def is_hermitian(self) -> bool:
        return self.lower is None

    def low(self, i: int) -> np.ndarray:
        """A_{i+1,i}, reading Hermitian storage as the adjoint of the upper block."""
        if self.lower is None:
            return np.asfortranarray(self.upper[i].conj().T)
        return self.lower[i]

    def offsets(self):
        """Start row of every block in the dense matrix."""
        return np.concatenate([[0], np.cumsum(self.block_sizes)])

    def to_dense(self) -> np.ndarray:
        off = self.offsets()
        n = off[-1]
        M = np.zeros((n, n), dtype=np.complex128, order="F")
        for i in range(self.nblocks):
            M[off[i]:off[i + 1], off[i]:off[i + 1]] = self.diag[i]
        for i in range(self.nblocks - 1):
            M[off[i]:off[i + 1], off[i + 1]:off[i + 2]] = self.upper[i]
            M[off[i + 1]:off[i + 2], off[i]:off[i + 1]] = self.low(i)
        return M

    @staticmethod
    def from_dense(M: np.ndarray, block_sizes, hermitian: bool = False) -> "BlockTridiag":
        off = np.concatenate([[0], np.cumsum(block_sizes)])
        n = len(block_sizes)
        diag = [np.asfortranarray(M[off[i]:off[i + 1], off[i]:off[i + 1]]) for i in range(n)]
        upper = [np.asfortranarray(M[off[i]:off[i + 1], off[i + 1]:off[i + 2]]) for i in range(n - 1)]
        lower = None if hermitian else [
            np.asfortranarray(M[off[i + 1]:off[i + 2], off[i]:off[i + 1]]) for i in range(n - 1)]
        return BlockTridiag(diag, upper, lower)

    def copy(self) -> "BlockTridiag":
        return BlockTridiag([d.copy(order="F") for d in self.diag],
                            [u.copy(order="F") for u in self.upper],
                            None if self.lower is None else [l.copy(order="F") for l in self.lower])


@dataclasses.dataclass
class DeviceHamiltonian:
    """Energy-independent part of the workload: H (block tridiagonal,
    Hermitian) and the two contact broadenings."""

    H_diag: list
    H_upper: list
    gamma_L: np.ndarray
    gamma_R: np.ndarray
    seed: int

    @property
    def nblocks(self) -> int:
        return len(self.H_diag)

    @property
    def b(self) -> int:
        return self.H_diag[0].shape[0]


def device_hamiltonian(N: int, b: int, seed: int = 1234) -> DeviceHamiltonian:
    """Tight-binding chain of N principal layers of size b (draw order in the
    module docstring)."""
    rng = np.random.default_rng(seed)
    H_diag = []
    for _ in range(N):
        onsite = rng.uniform(-2.0, 2.0, size=b)
        X = _cn(rng, (b, b), 1.0 / np.sqrt(b))
        h = (X + X.conj().T) / 2.0
        h[np.diag_indices(b)] += onsite
        H_diag.append(np.asfortranarray(h))
    H_upper = [_cn(rng, (b, b), 1.0 / np.sqrt(b)) for _ in range(N - 1)]
    ZL = _cn(rng, (b, b))
    ZR = _cn(rng, (b, b))
    gL = np.asfortranarray(ZL @ ZL.conj().T / b)
    gR = np.asfortranarray(ZR @ ZR.conj().T / b)
    return DeviceHamiltonian(H_diag, H_upper, gL, gR, seed)


def assemble_A(dev: DeviceHamiltonian, E: float, eta: float) -> BlockTridiag:
    """A(E) = (E + i eta) I - H - Sigma_L - Sigma_R with Sigma = -i Gamma / 2,
    stored non-Hermitian (both off-diagonal bands), as the RGF input."""
    N, b = dev.nblocks, dev.b
    z = complex(E, eta)
    diag = []
    for i in range(N):
        d = -dev.H_diag[i].copy(order="F")
        d[np.diag_indices(b)] += z
        if i == 0:
            d += 0.5j * dev.gamma_L
        if i == N - 1:
            d += 0.5j * dev.gamma_R
        diag.append(np.asfortranarray(d))
    upper = [np.asfortranarray(-u) for u in dev.H_upper]
    lower = [np.asfortranarray(-u.conj().T) for u in dev.H_upper]
    return BlockTridiag(diag, upper, lower)


def fermi(E, mu: float, kT: float):
    """Fermi-Dirac occupation 1 / (1 + exp((E - mu) / kT)), overflow-safe."""
    x = (np.asarray(E, dtype=np.float64) - mu) / kT
    return 0.5 * (1.0 - np.tanh(0.5 * x))


def energy_grid(n: int, lo: float = -1.0, hi: float = 1.0) -> np.ndarray:
    return np.linspace(lo, hi, n)


def band_edge_energies(dev: DeviceHamiltonian, nk: int = 64) -> tuple:
    """Lowest and highest band energy of the periodic lead built from the
    first layer: extrema of eig(H_11 + H_12 e^{ik} + H_12^H e^{-ik})."""
    lo, hi = np.inf, -np.inf
    h0, h1 = dev.H_diag[0], dev.H_upper[0]
    for k in np.linspace(0.0, np.pi, nk):
        hk = h0 + h1 * np.exp(1j * k) + h1.conj().T * np.exp(-1j * k)
        w = np.linalg.eigvalsh(hk)
        lo, hi = min(lo, w[0]), max(hi, w[-1])
    return lo, hi


def near_edge_grid(dev: DeviceHamiltonian, n: int, n_edge: int, lo=-1.0, hi=1.0,
                   delta: float = 1e-3, nk: int = 64, seed: int = 7) -> np.ndarray:
    """n energies: n - n_edge on linspace(lo, hi) plus n_edge placed uniformly
    within +-delta of the two lead band edges (half at each edge), sorted."""
    e_lo, e_hi = band_edge_energies(dev, nk)
    rng = np.random.default_rng(seed)
    half = n_edge // 2
    edge = np.concatenate([e_lo + rng.uniform(-delta, delta, half),
                           e_hi + rng.uniform(-delta, delta, n_edge - half)])
    return np.sort(np.concatenate([np.linspace(lo, hi, n - n_edge), edge]))


def bm_create_synthetic_random(N: int, b: int, seed: int = 1234, hermitian: bool = False,
                               uneven: bool = True) -> BlockTridiag:
    """The reference's synthetic matrix (``src/blockmatrix.jl:878-919``).

    Block sizes round(8 U + b - 4) (``:884``) unless ``uneven`` is False;
    diagonal blocks re, im ~ U[0, 1) (``be_random_array``,
    ``src/backend_cpu.jl:82``) and off-diagonal blocks 0.3 times the same
    (``:898, :909``).  Draw order follows the reference's loop: block sizes,
    then per row i the left block, the centre block, the right block.
    """
    rng = np.random.default_rng(seed)
    if uneven:
        sizes = [int(np.round(8.0 * u + (b - 4.0))) for u in rng.random(N)]
    else:
        sizes = [b] * N

    def rnd(m, n):
        return np.asfortranarray(rng.random((m, n)) + 1j * rng.random((m, n)))

    diag = [None] * N
    upper = [None] * (N - 1)
    lower = None if hermitian else [None] * (N - 1)
    for i in range(N):
        if not hermitian and i > 0:
            lower[i - 1] = 0.3 * rnd(sizes[i], sizes[i - 1])
        diag[i] = rnd(sizes[i], sizes[i])
        if i < N - 1:
            upper[i] = 0.3 * rnd(sizes[i], sizes[i + 1])
    return BlockTridiag(diag, upper, lower)


def random_hermitian_btd(sizes, seed: int = 99) -> BlockTridiag:
    """Random Hermitian block-tridiagonal matrix in Hermitian storage (the
    central operator of the Keldysh test, ``test/test_keldyshndiag.jl:66-71``:
    a random band symmetrised as (S + S^H) / 2)."""
    rng = np.random.default_rng(seed)
    n = len(sizes)
    diag, upper = [], []
    for i in range(n):
        X = rng.random((sizes[i], sizes[i])) + 1j * rng.random((sizes[i], sizes[i]))
        diag.append(np.asfortranarray((X + X.conj().T) / 2.0))
    for i in range(n - 1):
        upper.append(np.asfortranarray(
            (rng.random((sizes[i], sizes[i + 1])) + 1j * rng.random((sizes[i], sizes[i + 1]))) / 2.0))
    return BlockTridiag(diag, upper, None)


# The benchmark workloads (N layers, block size b, energies, eta, chemical
# potentials, temperature).  Kept here so tests, the bench and the docs agree.
CONFIGS = {
    "rgf_small": dict(N=16, b=64, E=0.3, eta=1e-3, seed=1234),
    "rgf_prod": dict(N=64, b=512, E=0.3, eta=1e-4, seed=1234),
    "rgf_keldysh_batched": dict(N=64, b=512, n_energies=256, lo=-1.0, hi=1.0, eta=1e-4,
                                kT=0.025, mu_L=0.1, mu_R=-0.1, seed=1234),
    "mixed_precision": dict(N=64, b=512, n_energies=512, n_edge=32, eta=1e-4,
                            kT=0.025, mu_L=0.1, mu_R=-0.1, seed=1234),
    "ddrgf_uneven": dict(devices=[(128, 256), (64, 512), (32, 1024)], parts=8,
                         n_energies=128, eta=1e-4, seed=1234),
}
