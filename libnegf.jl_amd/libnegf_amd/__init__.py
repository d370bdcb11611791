"""libnegf_amd: MI355X-native block-tridiagonal NEGF kernels behind a C ABI.

The compute path is libnegf_hip.so (gfx950 HIP kernels, include/negf_hip.h);
this package is the host-side mirror of LibNEGF.jl's interface for that path
plus the seeded synthetic inputs.  Importing it does not touch the GPU.
"""
from .synthetic import (BlockTridiag, DeviceHamiltonian, device_hamiltonian, assemble_A,
                        bm_create_synthetic_random, random_hermitian_btd, fermi, energy_grid,
                        band_edge_energies, near_edge_grid, CONFIGS)
from ._lib import LIB_PATH, NegfError, SingularError, load as load_library
from .flops import (flops_rgf, flops_rkd, flops_upward, flops_transmission,
                    flops_rgf_keldysh_energy, flops_energy_executed)

__version__ = "1.0.0"


def __getattr__(name):
    # the device-side API needs torch; import it lazily
    if name in ("api", "btd"):
        import importlib
        return importlib.import_module(f".{name}", __name__)
    raise AttributeError(name)
