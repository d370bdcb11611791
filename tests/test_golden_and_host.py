"""CPU-only checks: the oracle against the committed golden fixtures, the
host-side logic (packed layout, generator properties, flop counts), and that
libnegf_hip.so loads and exports every entry point include/negf_hip.h
declares (no compute calls: there may be no GPU here)."""
import ctypes
import pathlib
import re

import numpy as np
import pytest

from conftest import relfro
from libnegf_amd import synthetic, flops_rgf_keldysh_energy, flops_rgf, flops_rkd
from libnegf_amd.synthetic import BlockTridiag, fermi
from oracle.rgf import bndiag_of_inv_rgf_local, rgf_upward, first_row_chain
from oracle.keldysh import keldyshndiag

ROOT = pathlib.Path(__file__).resolve().parents[1]
GOLD = ROOT / "tests" / "golden"


def unpack(flat, shapes):
    out, o = [], 0
    for m, n in shapes:
        out.append(np.asfortranarray(flat[o:o + m * n].reshape(n, m).T))
        o += m * n
    return out


def load_btd(z, prefix, sizes, hermitian=False):
    N = len(sizes)
    diag = unpack(z[f"{prefix}_diag"], [(s, s) for s in sizes])
    up = unpack(z[f"{prefix}_upper"], [(sizes[i], sizes[i + 1]) for i in range(N - 1)])
    lo = None if hermitian else unpack(z[f"{prefix}_lower"],
                                       [(sizes[i + 1], sizes[i]) for i in range(N - 1)])
    return BlockTridiag(diag, up, lo)


@pytest.mark.parametrize("name", ["rgf_workload_N4_b8.npz", "rgf_uneven_N5.npz"])
def test_oracle_rgf_golden(name):
    z = np.load(GOLD / name)
    sizes = [int(x) for x in z["sizes"]]
    A = load_btd(z, "A", sizes)
    G = bndiag_of_inv_rgf_local(A)
    N = len(sizes)
    gd = unpack(z["G_diag"], [(s, s) for s in sizes])
    gu = unpack(z["G_upper"], [(sizes[i], sizes[i + 1]) for i in range(N - 1)])
    gl = unpack(z["G_lower"], [(sizes[i + 1], sizes[i]) for i in range(N - 1)])
    for i in range(N):
        assert relfro(G[(i, i)], gd[i]) < 1e-13
    for i in range(N - 1):
        assert relfro(G[(i, i + 1)], gu[i]) < 1e-13
        assert relfro(G[(i + 1, i)], gl[i]) < 1e-13


def test_oracle_keldysh_golden():
    z = np.load(GOLD / "keldysh_N5.npz")
    sizes = [int(x) for x in z["sizes"]]
    A = load_btd(z, "A", sizes)
    S = load_btd(z, "S", sizes, hermitian=True)
    nd, nu = keldyshndiag(A, S.diag, S.upper)
    N = len(sizes)
    rd = unpack(z["Gn_diag"], [(s, s) for s in sizes])
    ru = unpack(z["Gn_upper"], [(sizes[i], sizes[i + 1]) for i in range(N - 1)])
    for i in range(N):
        assert relfro(nd[i], rd[i]) < 1e-13
    for i in range(N - 1):
        assert relfro(nu[i], ru[i]) < 1e-13


def test_oracle_energies_golden():
    z = np.load(GOLD / "energies_N3_b6.npz")
    N, b = int(z["N"]), int(z["b"])
    dev = synthetic.DeviceHamiltonian(unpack(z["H_diag"], [(b, b)] * N),
                                      unpack(z["H_upper"], [(b, b)] * (N - 1)),
                                      unpack(z["gamma_L"], [(b, b)])[0],
                                      unpack(z["gamma_R"], [(b, b)])[0], seed=-1)
    for e, E in enumerate(z["energies"]):
        A = synthetic.assemble_A(dev, E, float(z["eta"]))
        buf = rgf_upward(A)
        G = bndiag_of_inv_rgf_local(A, upward=buf)
        g1n = first_row_chain(A, G, buf)[-1]
        T = np.real(np.trace(dev.gamma_L @ g1n @ dev.gamma_R @ g1n.conj().T))
        assert abs(T - z["T"][e]) < 1e-12
        gr = unpack(z["GR_diag"][e], [(b, b)] * N)
        for i in range(N):
            assert relfro(G[(i, i)], gr[i]) < 1e-13


def test_generator_properties():
    dev = synthetic.device_hamiltonian(5, 12, seed=3)
    for h in dev.H_diag:
        assert np.allclose(h, h.conj().T)
    for g in (dev.gamma_L, dev.gamma_R):
        assert np.allclose(g, g.conj().T)
        assert np.linalg.eigvalsh(g).min() > -1e-12
    A = synthetic.assemble_A(dev, 0.2, 1e-3)
    M = A.to_dense()
    # anti-Hermitian part of A is eta I + Gamma/2 >= eta I (so A is invertible)
    ah = (M - M.conj().T) / 2j
    assert np.linalg.eigvalsh(ah).min() > 1e-3 - 1e-12
    # same seed, same matrices
    dev2 = synthetic.device_hamiltonian(5, 12, seed=3)
    assert all(np.array_equal(a, b) for a, b in zip(dev.H_diag, dev2.H_diag))
    R = synthetic.bm_create_synthetic_random(50, 64, seed=1)
    assert all(60 <= s <= 68 for s in R.block_sizes)
    assert abs(fermi(0.1, 0.1, 0.025) - 0.5) < 1e-15
    assert fermi(10.0, 0.0, 0.025) == 0.0 and fermi(-10.0, 0.0, 0.025) == 1.0


def test_flop_counts_of_the_path():
    assert abs(flops_rgf(64, 512) - 4.975e11) / 4.975e11 < 1e-3
    assert abs(flops_rkd(64, 512) - 1.171e12) / 1.171e12 < 1e-3
    assert abs(flops_rgf_keldysh_energy(64, 512) - 1.511e12) / 1.511e12 < 1e-3
    from libnegf_amd import flops_energy_executed
    # what the energy driver launches per energy at N = 64: RGF upward 3 x 63
    # (t, l, Schur), downward 2 x 63, the factored Keldysh pass's E_j 63,
    # transmission 2, F_j 64, W_j 64 and G^< 64 at depth 2b (2 units each) =
    # 636 units of a 512^3 GEMM, and 64 inverses
    n_gemm = 3 * 63 + 2 * 63 + 63 + 2 + 64 + 64 + 2 * 64
    assert n_gemm == 636
    # the 64 G^< products compute 36 of their 64 tiles (anti-Hermitian G^<)
    assert flops_energy_executed(64, 512) == (8 * n_gemm + 8 * 64 - 16 * 64 * 28 / 64) * 512 ** 3
    assert flops_energy_executed(64, 512, 96) == flops_energy_executed(64, 512)
    # b = 160: 32 x 32 tiles, 15 of 25 computed; b = 512 at a wave of 4
    # (256 tiles of 64 x 64): 32 x 32 tiles, 136 of 256
    assert flops_energy_executed(4, 160) == (8 * (10 * 3 + 6) + 8 * 4 - 16 * 4 * 10 / 25) * 160 ** 3
    assert flops_energy_executed(64, 512, 4) == (8 * n_gemm + 8 * 64 - 16 * 64 * 120 / 256) * 512 ** 3
    # b = 100: register-staged GEMM, nothing mirrored
    assert flops_energy_executed(4, 100) == (8 * (10 * 3 + 6) + 8 * 4) * 100 ** 3
    # the three-product form: the GEMM part at 3/4, the inverses unchanged
    assert flops_energy_executed(64, 512, 96, 3) == \
        ((8 * n_gemm - 16 * 64 * 28 / 64) * 3 / 4 + 8 * 64) * 512 ** 3


def test_packed_layout_roundtrip_on_cpu_tensors():
    torch = pytest.importorskip("torch")
    from libnegf_amd.btd import DeviceBlockTridiag, Layout
    A = synthetic.bm_create_synthetic_random(4, 6, seed=2)
    L = Layout(A.block_sizes)
    assert L.nd == sum(s * s for s in A.block_sizes)
    assert L.no == sum(A.block_sizes[i] * A.block_sizes[i + 1] for i in range(3))
    D = DeviceBlockTridiag.from_host([A, A.copy()], device="cpu")
    B = D.to_host(1)
    for x, y in zip(A.diag + A.upper + A.lower, B.diag + B.upper + B.lower):
        assert np.array_equal(x, y)
    assert np.array_equal(D.block("upper", 2, 0).numpy(), A.upper[2])


def _declared_symbols():
    text = (ROOT / "include" / "negf_hip.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|const char\*|CCscMatrix)\s+(\w+)\s*\(",
                                 text, re.M)))


def test_library_exports_every_declared_symbol():
    from libnegf_amd._lib import LIB_PATH, SIGNATURES
    if not LIB_PATH.exists():
        pytest.skip("libnegf_hip.so not built (python libnegf.jl_amd/build_lib.py)")
    lib = ctypes.CDLL(str(LIB_PATH))
    syms = _declared_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), s
        assert s in SIGNATURES, f"{s} has no ctypes signature"
    assert lib.negf_version() == 10000


def test_inverse_workspace_serves_every_smaller_block():
    """A sweep carves one inverse workspace for its largest block (bmax) and
    inverts every layer's block in it.  The schedule's strip width falls as n
    grows (16 columns up to 512, 8 up to 1024, 4 beyond), so a smaller layer
    can need wider W panels than bmax: e.g. blocks [600, 512, 600] at 32
    energies send the 512 layer into split steps with 64-column panels.
    Checks, for every schedule variant, that the W buffers of every n <= nmax
    fit what the workspace for nmax holds (host-side queries, no GPU)."""
    from libnegf_amd._lib import LIB_PATH, load
    if not LIB_PATH.exists():
        pytest.skip("libnegf_hip.so not built")
    lib = load()
    out = (ctypes.c_int32 * 3)()
    for nmax in (96, 300, 512, 513, 600, 1024, 1100, 2048):
        for batch in (1, 8, 32, 96):
            ws = lib.negf_zinv_workspace(nmax, batch)
            # X (n^2), piv + used (2n ints) per instance; W and R are the rest
            w_bytes = ws - 1024 - 16 * nmax * nmax * batch - 8 * nmax * batch
            for v in range(7):
                for n in sorted({1, 16, 200, 256, 257, 511, 512, 513, 640, 1024, 1025, nmax}):
                    if n > nmax:
                        continue
                    rc = lib.negf_zinv_schedule(n, batch, v, out)
                    assert rc in (0, 1)
                    strip, panel, nbuf = out[0], out[1], out[2]
                    assert strip * panel <= 64
                    need = 16 * nbuf * n * strip * panel * batch
                    assert need <= w_bytes, (nmax, batch, v, n, need, w_bytes)
    # the case that wrote past the workspace before: a 512 layer beside 600s
    assert lib.negf_zinv_schedule(512, 32, 0, out) == 1 and out[0] * out[1] == 64
    assert lib.negf_zinv_schedule(0, 1, 0, out) < 0
    assert lib.negf_zinv_schedule(512, 32, 6, out) == 1 and out[2] == 4
    assert lib.negf_zinv_schedule(300, 32, 6, out) == 1 and out[2] == 3
    assert lib.negf_zinv_schedule(16, 1, 7, out) < 0


def test_rgf_noalloc_refuses_beyond_its_scratch():
    """negf_rgf_noalloc checks the sweep's workspace against its static
    device array (64 MiB per GPU) before any HIP call: N = 64 blocks of 512
    (about 1 GiB) is refused with NEGF_TOO_LARGE; bad sizes with
    NEGF_BAD_ARG.  The pointers are never dereferenced."""
    from libnegf_amd._lib import LIB_PATH, load
    if not LIB_PATH.exists():
        pytest.skip("libnegf_hip.so not built")
    lib = load()
    dummy = ctypes.c_void_p(16)
    sizes = (ctypes.c_int32 * 64)(*([512] * 64))
    rc = lib.negf_rgf_noalloc(64, sizes, dummy, 0, dummy, 0, dummy, 0, dummy, None, None, 1, None,
                              None)
    assert rc == -2 and b"beyond" in lib.negf_status_string(rc)
    # 64 MiB: 16 * (2 * 48 * 400^2 ...) just beyond at one instance of N = 48, b = 400
    sizes = (ctypes.c_int32 * 48)(*([400] * 48))
    assert lib.negf_rgf_noalloc(48, sizes, dummy, 0, dummy, 0, dummy, 0, dummy, None, None, 1,
                                None, None) == -2
    bad = (ctypes.c_int32 * 2)(64, 0)
    assert lib.negf_rgf_noalloc(2, bad, dummy, 0, dummy, 0, dummy, 0, dummy, None, None, 1, None,
                                None) == -1
    assert lib.negf_rgf_noalloc(2, None, dummy, 0, dummy, 0, dummy, 0, dummy, None, None, 1, None,
                                None) == -1


def test_tune_table_file_and_lookup_rules(tmp_path):
    """The tune table (host-side, no GPU): a file written by tune.record loads
    back into the library; a lookup matches every key but the batch-like last
    one exactly and takes the largest batch key <= the request, else the
    smallest above it; malformed lines are refused."""
    from libnegf_amd._lib import LIB_PATH, load, call
    from libnegf_amd import tune
    if not LIB_PATH.exists():
        pytest.skip("libnegf_hip.so not built")
    load()
    p = tmp_path / "t.txt"
    try:
        call("negf_tune_clear")
        tune.record("gemm", (512, 512, 512, 32), (9,), {"v9": 1e-3, "v10": 2e-3}, path=p)
        tune.record("gemm", (512, 512, 512, 96), (6,), path=p)
        tune.record("bndiag", (128, 256, 8), (15,), path=p)
        tune.record("bndiag", (128, 256, 8), (0,), {"rgf": 0.5}, path=p)   # replaces the line
        tune.record("bndiag", (64, 512, 8), (7, 2), path=p)
        tune.record("inv", (512, 96), (4,), path=p)
        tune.record("getrs", (512, 16), (1,), path=p)
        tune.record("getrs", (512, 512), (2,), {"r1": 2e-3, "r2": 1e-3}, path=p)
        text = p.read_text()
        assert text.count("bndiag 128 256 8") == 1 and "bndiag 128 256 8 0  # rgf:500.000ms" in text
        call("negf_tune_clear")
        assert call("negf_tune_count") == 0 and tune.lookup("gemm", (512, 512, 512, 32)) is None
        assert tune.load_table(p) == 7
        assert tune.lookup("gemm", (512, 512, 512, 64)) == [9]     # largest <= 64
        assert tune.lookup("gemm", (512, 512, 512, 200)) == [6]
        assert tune.lookup("gemm", (512, 512, 512, 4)) == [9]      # none <= 4: smallest above
        assert tune.lookup("gemm", (256, 256, 256, 64)) is None    # other keys must match
        assert tune.lookup("bndiag", (128, 256, 1)) == [0]
        assert tune.lookup("bndiag", (64, 512, 16)) == [7, 2]
        assert tune.lookup("inv", (512, 8)) == [4]
        assert tune.lookup("getrs", (512, 100)) == [1] and tune.lookup("getrs", (512, 600)) == [2]
        assert tune.lookup("getrs", (256, 600)) is None
        bad = tmp_path / "bad.txt"
        for line in ("frob 1 2 3", "gemm 512 512 512", "inv 512 -1 2", "inv 512 8 1 2 3 4 5",
                     "getrs 512 16"):
            bad.write_text(line + "\n")
            with pytest.raises(Exception):
                tune.load_table(bad)
    finally:
        call("negf_tune_clear")
        tune.load_table()


def test_tuner_candidates_always_include_rgf_and_split():
    from libnegf_amd import tune
    from libnegf_amd.ddrgf import bndiag_of_inv_ddrgf_get_nr_tasks as nr_tasks
    for npl in (5, 16, 32, 64, 128, 160):
        c = tune.bndiag_candidates(npl)
        assert c[0] == [0]
        for lv in c[1:]:
            n = npl
            for bs in lv:
                nr, _ = nr_tasks(n, bs)
                assert nr >= 2, (npl, lv)
                n = nr


def test_new_entry_points_refuse_bad_arguments_without_a_gpu():
    """Argument checks of the multi-device, mixed-precision and uneven DD-RGF
    entries run on the host before any device call."""
    from libnegf_amd._lib import LIB_PATH, load, NegfBandJob
    if not LIB_PATH.exists():
        pytest.skip("libnegf_hip.so not built")
    lib = load()
    BAD = -1
    assert lib.negf_bndiag_devices(0, None, None) == 0
    assert lib.negf_bndiag_devices(9, ctypes.cast((NegfBandJob * 9)(), ctypes.c_void_p), None) == BAD
    assert lib.negf_bndiag_devices(-1, None, None) == BAD
    job = (NegfBandJob * 1)()          # no info array, no sizes
    assert lib.negf_bndiag_devices(1, ctypes.cast(job, ctypes.c_void_p), None) == BAD
    sizes = (ctypes.c_int32 * 3)(8, 8, 8)
    for prec, nmod, tol in ((4, 0, 1e-8), (-1, 0, 1e-8), (2, 1, 1e-8), (2, 17, 1e-8),
                            (2, 0, float("nan")), (2, 0, -1.0)):
        assert lib.negf_rgf_rkd_mixed(3, sizes, None, 0, None, 0, None, 0, None, 0, None, 0,
                                      None, None, None, None, None, None, 0, None, 0, None, 1,
                                      None, prec, nmod, tol, None, None, None, None) == BAD
    lv = (ctypes.c_int32 * 1)(3)
    uneven = (ctypes.c_int32 * 12)(*([16, 18] * 6))
    uniform = (ctypes.c_int32 * 12)(*([18] * 12))
    wu = lib.negf_ddrgf_blocksizes_workspace(12, uneven, 2, lv, 1)
    wf = lib.negf_ddrgf_blocksizes_workspace(12, uniform, 2, lv, 1)
    assert wu > wf > 0     # the padded copies of A and G come on top of the same sweep
    # blockSizeD1 = 10 leaves an empty last task (utils.jl:278-312): refused
    assert lib.negf_ddrgf_blocksizes_workspace(12, uneven, 2, (ctypes.c_int32 * 1)(10), 1) == 0
    big = (ctypes.c_int32 * 1025)(*([4, 5] * 512 + [4]))
    assert lib.negf_ddrgf_blocksizes_workspace(1025, big, 1, lv, 1) == 0   # > 1024 uneven blocks
    # the three-product forms: precision 0 or 3 only, checked before any device call
    for prec in (1, 2, 4, -1):
        assert lib.negf_bndiag_devices_prec(0, None, prec, None) == BAD
        g = ctypes.c_void_p()
        assert lib.negf_rgf_keldysh_energies_graph_create_prec(
            4, 64, None, None, None, None, None, 8, 1e-4, 0.1, -0.1, 0.025, None, None, None,
            None, 8, 1, prec, ctypes.byref(g)) == BAD
        assert not g.value
    assert lib.negf_bndiag_devices_prec(0, None, 3, None) == 0
    assert lib.negf_energies_graph_estimates(None, None, None) == BAD


def test_python_precision_arguments_are_checked_first():
    """The Python wrappers refuse an unknown product form before touching a
    device (so the error is the argument's, on any host)."""
    from libnegf_amd import api, ddrgf as dd
    with pytest.raises(ValueError):
        dd.bndiag_devices([], precision="fp32")
    with pytest.raises(ValueError):
        api.EnergiesGraph(None, None, 1e-4, 0.1, -0.1, 0.025, precision="emulated")


def test_energy_job_tensor_checks_before_any_device_call():
    """The energy job's wrappers refuse tensors the C ABI would misread (a
    float32 energy vector read as doubles, outputs too small for every
    energy, host memory) before touching a device."""
    import types
    import torch
    from libnegf_amd import api
    H = types.SimpleNamespace(N=2, b=4)
    e32 = torch.zeros(3, dtype=torch.float32)
    e64 = torch.zeros(3, dtype=torch.float64)
    GR = torch.zeros((3, 2 * 16), dtype=torch.complex128)
    with pytest.raises(ValueError, match="float64"):
        api.rgf_keldysh_energies(H, e32, 1e-4, 0.1, -0.1, 0.025, GR)
    with pytest.raises(ValueError, match="at least 96"):
        api.rgf_keldysh_energies(H, e64, 1e-4, 0.1, -0.1, 0.025, GR[:2])
    with pytest.raises(ValueError, match="complex128"):
        api.rgf_keldysh_energies(H, e64, 1e-4, 0.1, -0.1, 0.025, None, GR.to(torch.complex64))
    with pytest.raises(ValueError, match="at least 3"):
        api.rgf_keldysh_energies_mixed(H, e64, 1e-4, 0.1, -0.1, 0.025, "fp64_3m",
                                       transmission=e64[:2])
    with pytest.raises(ValueError, match="on the GPU"):
        api.EnergiesGraph(H, e64, 1e-4, 0.1, -0.1, 0.025, GR, precision="fp64_3m")
    with pytest.raises(ValueError, match="precision"):
        api.rgf_keldysh_energies_mixed(H, e64, 1e-4, 0.1, -0.1, 0.025, "fp16")


def test_info_codes_map_to_python_errors():
    """A positive info (the layer of a singular Schur complement) raises
    SingularError, the energy job's -1 (a contact broadening that is not
    Hermitian positive semidefinite) ValueError, all zeros nothing."""
    import torch
    from libnegf_amd.api import _check_info
    from libnegf_amd._lib import SingularError
    _check_info("x", torch.zeros(3, dtype=torch.int32))
    with pytest.raises(SingularError):
        _check_info("x", torch.tensor([0, 2, 0], dtype=torch.int32))
    with pytest.raises(ValueError, match="positive semidefinite"):
        _check_info("x", torch.tensor([-1, -1], dtype=torch.int32))
    with pytest.raises(ValueError):
        _check_info("x", torch.tensor([-1, 3], dtype=torch.int32))


def test_leading_dimension_of_single_column_and_single_row_matrices():
    """torch gives a one-column tensor any stride for its column (often 1); the
    binding passes ld = m for it, and ld = the column stride otherwise."""
    import torch
    from libnegf_amd.api import _ld, col_major
    v = col_major(torch.zeros((64, 1), dtype=torch.complex128))
    assert _ld(v) == 64
    assert _ld(col_major(torch.zeros((1, 5), dtype=torch.complex128))) >= 1
    assert _ld(col_major(torch.zeros((7, 3), dtype=torch.complex128))) == 7
    with pytest.raises(ValueError):
        _ld(torch.zeros((7, 3), dtype=torch.complex128))      # row-major


def test_rgf_schedule_knob_checks_its_argument():
    """negf_rgf_schedule takes 0 (automatic), 1 (one-sided) or 2 (two-sided
    wherever the chain allows); anything else is refused on the host."""
    from libnegf_amd._lib import call, NegfError
    with pytest.raises(NegfError):
        call("negf_rgf_schedule", 3)
    with pytest.raises(NegfError):
        call("negf_rgf_schedule", -1)
    call("negf_rgf_schedule", 1)
    call("negf_rgf_schedule", 0)
