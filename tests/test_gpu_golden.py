"""The GPU path against the committed golden fixtures (tests/golden/*.npz,
written by tests/golden/make_golden.py and re-checked against the oracle by
tests/test_golden_and_host.py): RGF on uniform and uneven blocks, RKD with a
random Hermitian central operator, and the energy-batched G^R / G^< / T job.
The fixtures are packed column-major block arrays in the layout of
include/negf_hip.h, so the device outputs are compared segment by segment."""
import pathlib

import numpy as np
import pytest

from conftest import relfro

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from libnegf_amd import api  # noqa: E402
from libnegf_amd.btd import DeviceBlockTridiag  # noqa: E402
from libnegf_amd.synthetic import BlockTridiag, DeviceHamiltonian  # noqa: E402

GOLD = pathlib.Path(__file__).resolve().parent / "golden"


def _unpack(flat, shapes):
    out, o = [], 0
    for m, n in shapes:
        out.append(np.asfortranarray(flat[o:o + m * n].reshape(n, m).T))
        o += m * n
    return out


def _btd(z, prefix, sizes, hermitian=False):
    N = len(sizes)
    diag = _unpack(z[f"{prefix}_diag"], [(s, s) for s in sizes])
    up = _unpack(z[f"{prefix}_upper"], [(sizes[i], sizes[i + 1]) for i in range(N - 1)])
    lo = None if hermitian else _unpack(z[f"{prefix}_lower"],
                                        [(sizes[i + 1], sizes[i]) for i in range(N - 1)])
    return BlockTridiag(diag, up, lo)


def _segments(dev_flat, gold_flat, shapes):
    """Largest per-block relative error of a packed device array vs the fixture."""
    got = _unpack(np.asarray(dev_flat), shapes)
    want = _unpack(gold_flat, shapes)
    return max((relfro(g, w) for g, w in zip(got, want)), default=0.0)


@pytest.mark.parametrize("name", ["rgf_workload_N4_b8.npz", "rgf_uneven_N5.npz"])
def test_rgf_matches_golden(name):
    print(f"golden rgf {name}", flush=True)
    z = np.load(GOLD / name)
    sizes = [int(x) for x in z["sizes"]]
    N = len(sizes)
    M = DeviceBlockTridiag.from_host(_btd(z, "A", sizes))
    out = M.similar()
    api.bndiag_of_inv_rgf_local(out, M)
    nd = sum(s * s for s in sizes)
    no = sum(sizes[i] * sizes[i + 1] for i in range(N - 1))
    dshape = [(s, s) for s in sizes]
    ushape = [(sizes[i], sizes[i + 1]) for i in range(N - 1)]
    lshape = [(sizes[i + 1], sizes[i]) for i in range(N - 1)]
    assert _segments(out.diag[0, :nd].cpu(), z["G_diag"], dshape) < 1e-12
    assert _segments(out.upper[0, :no].cpu(), z["G_upper"], ushape) < 1e-12
    assert _segments(out.lower[0, :no].cpu(), z["G_lower"], lshape) < 1e-12


def test_keldysh_matches_golden():
    print("golden keldysh N=5", flush=True)
    z = np.load(GOLD / "keldysh_N5.npz")
    sizes = [int(x) for x in z["sizes"]]
    N = len(sizes)
    M = DeviceBlockTridiag.from_host(_btd(z, "A", sizes))
    S = DeviceBlockTridiag.from_host(_btd(z, "S", sizes, hermitian=True))
    out = api.keldyshndiag(M, S, api.allocate_aux_data_Keldysh(M, S))
    nd = sum(s * s for s in sizes)
    no = sum(sizes[i] * sizes[i + 1] for i in range(N - 1))
    assert _segments(out.diag[0, :nd].cpu(), z["Gn_diag"], [(s, s) for s in sizes]) < 1e-12
    assert _segments(out.upper[0, :no].cpu(), z["Gn_upper"],
                     [(sizes[i], sizes[i + 1]) for i in range(N - 1)]) < 1e-12


def test_energies_match_golden():
    print("golden energies N=3 b=6", flush=True)
    z = np.load(GOLD / "energies_N3_b6.npz")
    N, b = int(z["N"]), int(z["b"])
    dev = DeviceHamiltonian(_unpack(z["H_diag"], [(b, b)] * N),
                            _unpack(z["H_upper"], [(b, b)] * (N - 1)),
                            _unpack(z["gamma_L"], [(b, b)])[0],
                            _unpack(z["gamma_R"], [(b, b)])[0], seed=-1)
    H = api.DeviceHamiltonianGPU(dev)
    Es = torch.tensor(z["energies"], dtype=torch.float64, device="cuda")
    nE = Es.numel()
    GR = torch.zeros((nE, N * b * b), dtype=torch.complex128, device="cuda")
    GL = torch.zeros_like(GR)
    T = torch.zeros(nE, dtype=torch.float64, device="cuda")
    api.rgf_keldysh_energies(H, Es, float(z["eta"]), float(z["mu_L"]), float(z["mu_R"]),
                             float(z["kT"]), GR, GL, T)
    torch.cuda.synchronize()
    GR, GL, T = GR.cpu().numpy(), GL.cpu().numpy(), T.cpu().numpy()
    for e in range(nE):
        assert _segments(GR[e], z["GR_diag"][e], [(b, b)] * N) < 1e-12
        assert _segments(GL[e], z["Gless_diag"][e], [(b, b)] * N) < 1e-12
        assert abs(T[e] - z["T"][e]) < 1e-11 * max(1.0, abs(z["T"][e]))
